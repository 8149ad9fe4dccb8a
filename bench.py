"""Benchmark of the density-peak hot path (kNN -> rho -> delta/nneigh) on one MI355X.

A step is one irc_rho_delta_batch call over the whole drift-10M workload (384 sites,
10M spikes, D=42, 200 drift chunks, 15 links within +/-64, k=30) with features resident in
HBM.  `value` is spikes/s of the headline form (BF16 screen + FP64 re-rank); the plain
FP32-screen form is timed in the same run ("forms").  The CPU baseline is the FP32
restatement of the reference CPU path timed on a bounded sample of sites.

  python bench.py [--steps K] [--warmup W] [--workload drift-10M] [--forms bf16,fp32]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spikes/s via kNN+rho+delta at 10M spikes; BF16-screen vs FP32, exact kNN"
PEAK = {"fp32": 157.3e12, "bf16": 2.5e15, "fp8": 5.0e15, "exact": 78.6e12}
PREC = {"exact": 0, "fp32": 1, "bf16": 2, "fp8": 4}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="drift-10M")
    ap.add_argument("--forms", default="bf16,fp32")
    ap.add_argument("--cpu-sites", type=int, default=0, help="sites in the CPU baseline sample (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="verify size-independent properties of the output")
    ap.add_argument("--autotune", action="store_true",
                    help="pick k_over / grid / list depth / pre-pass / group size per form by timing them first "
                         "(outside the timed steps)")
    ap.add_argument("--graph", action="store_true",
                    help="also time replays of the call's device phase captured in a HIP graph")
    ap.add_argument("--per-site", type=int, default=0,
                    help="also time the per-site call pattern (irc_knn_drift + irc_delta_drift per site, host "
                         "buffers, the reference's rho_paged_site_/delta_paged_site_ loop) on this many sites")
    ap.add_argument("--host", action="store_true",
                    help="also time the host-buffer form of the headline call (features copied in and results "
                         "copied out over PCIe inside the timed call)")
    ap.add_argument("--no-lanes", action="store_true",
                    help="skip the launch-pattern lane on lognormal-10M (persistent queue vs per-site launches vs "
                         "graph replay of the autotuned setting)")
    ap.add_argument("--lane-steps", type=int, default=3)
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(the path would shard by site)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from ironclust_amd import irc2, libirc, synth

    if a.copy_calib > 0:
        n = int(a.copy_calib * (1 << 30)) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n}))
        return
    t0 = time.time()
    ds = synth.make_workload(a.workload, device="cuda")
    lay = irc2.site_layout(ds)
    torch.cuda.synchronize()
    N, k, D = ds["n_spikes"], ds["k"], ds["D"]
    pairs = synth.n_pairs(ds)
    print(f"# data {a.workload}: N={N} D={D} sites={ds['n_sites']} chunks={ds['n_drift']} "
          f"pairs/pass={pairs:.4e} gen {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = dict(F=lay["F"], col_spike=lay["col_spike"], site_ptr=lay["site_ptr"],
               site_n_primary=lay["site_n_primary"], lim_drift=lay["lim_drift"], link=lay["link"],
               n_spikes=N, k=k)
    out = dict(rho=torch.empty(N, dtype=torch.float32, device="cuda"),
               delta=torch.empty(N, dtype=torch.float32, device="cuda"),
               nneigh=torch.empty(N, dtype=torch.int64, device="cuda"),
               knn=torch.empty((N, k), dtype=torch.int64, device="cuda"))
    ctx = libirc.Context(0)
    forms = {}
    results = {}
    tuned = {}
    for form in a.forms.split(","):
        opts = libirc.make_options(precision=PREC[form])
        if a.autotune and form != "exact":
            from ironclust_amd import autotune
            choice, rep = autotune.tune(ctx, dev, out, precision=PREC[form])
            tuned[form] = {"options": choice, **rep}
            opts = libirc.make_options(precision=PREC[form], **choice)
            print(f"# autotune {form}: {choice}", file=sys.stderr, flush=True)
        for _ in range(a.warmup):
            ctx.rho_delta_batch_device(dev, out, opts)
        torch.cuda.synchronize()
        st_acc = {"ms_knn": 0.0, "ms_delta": 0.0, "ms_total": 0.0, "ms_screen": 0.0, "ms_rerank": 0.0}
        t1 = time.perf_counter()
        for _ in range(a.steps):
            ctx.rho_delta_batch_device(dev, out, opts)
            s = ctx.stats()
            for kk in st_acc:
                st_acc[kk] += s.get(kk, 0.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / a.steps
        s = ctx.stats()
        ms_kernel = st_acc["ms_screen"] / a.steps if st_acc["ms_screen"] > 0 else st_acc["ms_knn"] / a.steps
        flops = 2.0 * 2 * D * pairs / 2          # one pass (kNN) of 2*D flops per pair
        forms[form] = {
            "spikes_per_s": N / dt, "ms_per_step": dt * 1e3,
            "ms_knn": st_acc["ms_knn"] / a.steps, "ms_delta": st_acc["ms_delta"] / a.steps,
            "ms_device_total": st_acc["ms_total"] / a.steps,
            "knn_kernel_ms": ms_kernel, "rerank_ms": st_acc["ms_rerank"] / a.steps,
            "knn_tflops": flops / (ms_kernel * 1e-3) / 1e12,
            "roofline_frac": flops / (ms_kernel * 1e-3) / PEAK[form],
            "screen_fallback": int(s["n_screen_fallback"]), "rerank_overflow": int(s["n_rerank_overflow"]),
            "delta_full_scan": int(s["n_delta_full"]),
            "precision_used": int(s["precision_used"]),
        }
        if len(a.forms.split(",")) > 1 or a.graph:
            results[form] = {kk: v.clone() for kk, v in out.items()}
        else:
            results[form] = out
        if a.graph:
            go = libirc.make_options(precision=PREC[form], **dict(tuned.get(form, {}).get("options", {}),
                                                                   flags=libirc.FLAG_GRAPH |
                                                                   tuned.get(form, {}).get("options", {}).get("flags", 0)))
            ctx.rho_delta_batch_device(dev, out, go)          # runs once and records the graph
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for _ in range(a.steps):
                ctx.rho_delta_batch_device(dev, out, go)
            torch.cuda.synchronize()
            forms[form]["graph_replay_ms_per_step"] = (time.perf_counter() - t2) / a.steps * 1e3
            forms[form]["graph_ms_device"] = ctx.stats()["ms_graph"]
            forms[form]["graph_identical"] = all(torch.equal(out[kk], results[form][kk]) for kk in out)
        print(f"# {form}: {forms[form]}", file=sys.stderr, flush=True)
    head = a.forms.split(",")[0]
    host_rate = None
    if a.host:
        import numpy as np
        hl = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in dev.items()}
        hopts = libirc.make_options(precision=PREC[head])
        ctx.rho_delta_batch(hl, k, hopts, want_knn=True)
        t3 = time.perf_counter()
        hr = ctx.rho_delta_batch(hl, k, hopts, want_knn=True)
        host_dt = time.perf_counter() - t3
        assert np.array_equal(hr["rho"], results[head]["rho"].cpu().numpy())
        host_rate = {"spikes_per_s": N / host_dt, "ms_per_call": host_dt * 1e3}
        print(f"# host buffers ({head}): {host_rate}", file=sys.stderr, flush=True)
    # all forms must agree bit for bit
    for form in results:
        for kk in results[head]:
            assert torch.equal(results[form][kk], results[head][kk]), (form, kk)
    if a.check:
        r = results[head]
        ids = torch.arange(1, N + 1, device="cuda")
        has = r["nneigh"] != ids
        assert bool((r["rho"][r["nneigh"][has] - 1] > r["rho"][has]).all())
        assert bool((r["knn"][:, 0] == ids).all())
    hf = forms[head]
    flops = 2.0 * D * pairs
    traffic = screen_traffic(a.workload, head)
    peak = PEAK[head]
    ach = flops / (hf["knn_kernel_ms"] * 1e-3)
    line = {
        "metric": METRIC, "value": hf["spikes_per_s"], "unit": "spikes/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": hf["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"bf16": "bf16", "fp32": "fp32", "fp8": "fp8", "exact": "f64"}[head],
        "data": "synthetic: seeded drift recording generated on the GPU, features resident in HBM",
        "config": {"workload": a.workload, "n_spikes": N, "n_sites": ds["n_sites"], "D": D,
                   "n_drift": ds["n_drift"], "k": k, "pairs_per_pass": pairs,
                   "form": {"bf16": "BF16 hi/lo MFMA screen + FP64 re-rank", "fp32": "FP32 MFMA screen + FP64 re-rank",
                            "fp8": "FP8 e4m3 three-term MFMA screen (per-site scale) + FP64 re-rank",
                            "exact": "FP64 exact"}[head]},
        "roofline": {"bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": ach / peak, "traffic": traffic,
                     "kernel": "kNN pass (screen kernel when screened)",
                     "flops_per_launch": flops,
                     "traffic_source": TRAFFIC_SRC if traffic is not None else None,
                     "frac_by_form": {f: forms[f]["roofline_frac"] for f in forms}},
        "forms": forms,
    }
    if tuned:
        line["autotune"] = tuned
    if host_rate:
        line["host_buffers_pcie_inclusive"] = host_rate
    if a.per_site > 0:
        line["per_site_launches"] = per_site(ds, lay, a, ctx, PREC[head], results[head])
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ds, a)
    if not a.no_lanes and a.workload == "drift-10M":
        del results, out, dev, lay, ds
        line["launch_lane"] = launch_lane(a, ctx, head)
    print(json.dumps(line), flush=True)


TRAFFIC_SRC = ("profiles/traffic_<workload>.json: rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch of the screen "
               "kernel, each scaled by the factor a plain device copy of known size reads")


def screen_traffic(workload, form):
    """HBM bytes of one screen launch (its kernels: the item split, the full-item and the half-item
    screen kernels) from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if form not in ("bf16", "fp32", "fp8") or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    prefix = "k_screen_knn<" + {"fp32": "0", "bf16": "1", "fp8": "2"}[form] + ","
    tot, hit = 0.0, False
    for name, d in t.get("forms", {}).get(form, {}).items():
        if name.startswith(prefix) or name.startswith("k_split_items"):
            tot += d["hbm_bytes"]
            hit = hit or name.startswith(prefix)
    return tot if hit else None


def launch_lane(a, ctx, form):
    """lognormal-10M (lognormal site sizes, Dirichlet chunk shares): the batch call with the
    persistent per-XCD item queue, with one screen launch per site (the per-site launch pattern of
    the reference's site loop, matlab/irc2.m:3537-3561, 3764-3788), and as a graph replay of the
    setting the library's autotuner picks (irc_autotune).  Each is timed over `lane_steps` calls
    after one warm-up; all three must give identical outputs."""
    import torch
    from ironclust_amd import irc2, libirc, synth
    ds = synth.make_workload("lognormal-10M", device="cuda")
    lay = irc2.site_layout(ds)
    N, k = ds["n_spikes"], ds["k"]
    dev = dict(F=lay["F"], col_spike=lay["col_spike"], site_ptr=lay["site_ptr"],
               site_n_primary=lay["site_n_primary"], lim_drift=lay["lim_drift"], link=lay["link"],
               n_spikes=N, k=k)
    out = dict(rho=torch.empty(N, dtype=torch.float32, device="cuda"),
               delta=torch.empty(N, dtype=torch.float32, device="cuda"),
               nneigh=torch.empty(N, dtype=torch.int64, device="cuda"),
               knn=torch.empty((N, k), dtype=torch.int64, device="cuda"))
    prec = PREC[form]

    def timed(opts):
        ctx.rho_delta_batch_device(dev, out, opts)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(a.lane_steps):
            ctx.rho_delta_batch_device(dev, out, opts)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / a.lane_steps
        return {"ms_per_step": dt * 1e3, "spikes_per_s": N / dt, "ms_screen": ctx.stats()["ms_screen"]}

    res = {"workload": "lognormal-10M", "form": form, "n_spikes": N}
    res["persistent_queue"] = timed(libirc.make_options(precision=prec))
    ref = {kk: v.clone() for kk, v in out.items()}
    res["site_launches"] = timed(libirc.make_options(precision=prec, flags=libirc.FLAG_SITE_LAUNCHES))
    same = all(torch.equal(out[kk], ref[kk]) for kk in out)
    choice, rep = ctx.autotune_device(dev, out, libirc.make_options(precision=prec))
    res["autotune"] = {"choice": choice, "ms_base": rep["ms_base"], "ms_best": rep["ms_best"],
                       "trials": len(rep["trials"])}
    res["graph_autotuned"] = timed(libirc.make_options(precision=prec, flags=libirc.FLAG_GRAPH, **choice))
    same = same and all(torch.equal(out[kk], ref[kk]) for kk in out)
    res["identical"] = same
    print(f"# launch lane: {res}", file=sys.stderr, flush=True)
    return res


def per_site(ds, lay, a, ctx, prec, ref):
    """The per-site call pattern on the first `a.per_site` sites: for each site one
    irc_knn_drift over its primary columns, then (after every site's rho is known) one
    irc_delta_drift, with host buffers -- the loop of rho_paged_site_ / delta_paged_site_
    (matlab/irc2.m:3635-3726) on the library's per-site entry points.  Results are checked
    against the batched call."""
    import numpy as np
    import torch
    from ironclust_amd import libirc
    S = min(a.per_site, ds["n_sites"])
    F = lay["F"].cpu().numpy()
    col_spike = lay["col_spike"].cpu().numpy()
    site_ptr, site_np = lay["site_ptr"], lay["site_n_primary"]
    lim = np.asarray(ds["lim"], dtype=np.int64)
    link = np.asarray(ds["link"], dtype=np.uint8)
    k = ds["k"]
    opts = libirc.make_options(precision=prec)
    rho_ref = ref["rho"].cpu().numpy()
    sites = []
    for s in range(S):
        b, e = int(site_ptr[s]), int(site_ptr[s + 1])
        ids = col_spike[b:e].astype(np.int64)
        drift = (np.searchsorted(lim, ids, side="right")).astype(np.int32)
        sites.append((np.ascontiguousarray(F[b:e]), drift, ids, int(site_np[s])))
    n_spk = sum(x[3] for x in sites)
    torch.cuda.synchronize()
    t = time.perf_counter()
    rhos = []
    for Fs, drift, ids, npri in sites:
        dk, _ = ctx.knn_drift(Fs, drift, link, 0, npri, k, id_of_col=ids, opts=opts)
        rhos.append((1.0 / dk.astype(np.float64)).astype(np.float32))
    for (Fs, drift, ids, npri), rho in zip(sites, rhos):
        rho_cols = rho_ref[ids - 1] if len(ids) > npri else rho
        ctx.delta_drift(Fs, drift, link, 0, npri, rho_cols, id_of_col=ids, opts=opts)
    dt = time.perf_counter() - t
    # the per-site entry returns d_k in single precision, so 1/d_k may differ from the batched
    # rho (computed from d_k^2 in FP64) in the last bit
    rel = max(float(np.max(np.abs(r / rho_ref[ids[:npri] - 1] - 1.0))) for (Fs, drift, ids, npri), r in
              zip(sites, rhos))
    return {"value": n_spk / dt, "unit": "spikes/s", "sites": S, "spikes": int(n_spk), "seconds": dt,
            "rho_max_rel_diff_vs_batched": rel}


def cpu_baseline(ds, a):
    """FP32 restatement of knn_cpu_ + delta_cpu_ (OpenMP over sites) on a bounded sample of
    sites of the same workload; spikes/s = sampled spikes / wall time."""
    import numpy as np
    from oracle import oracle as O
    host = dict(ds)
    for kk in ("F1", "site1"):
        host[kk] = ds[kk].cpu().numpy()
    host["F2"] = None if ds.get("F2") is None else ds["F2"].cpu().numpy()
    host["site2"] = None if ds.get("site2") is None else ds["site2"].cpu().numpy()
    lib = O.cpu_port()
    cores = int(lib.cpu_max_threads())
    S = ds["n_sites"]
    # probe one site per core, then size the sample to about 15 s of CPU work
    n_sites = a.cpu_sites or max(1, min(S, cores))
    if not a.cpu_sites:
        sites = np.arange(1, n_sites + 1, dtype=np.int32)
        t = time.perf_counter()
        O.cpu_sort_batch(host, sites=sites, nthreads=cores)
        probe = time.perf_counter() - t
        n_sites = int(min(S, max(n_sites, n_sites * 15.0 / max(probe, 1e-3))))
    sites = np.arange(1, n_sites + 1, dtype=np.int32)
    n_spk = int(np.isin(host["site1"], sites).sum())
    t = time.perf_counter()
    O.cpu_sort_batch(host, sites=sites, nthreads=cores)
    dt = time.perf_counter() - t
    return {"value": n_spk / dt, "unit": "spikes/s", "cores": cores, "kind": "port",
            "sample": f"{n_sites} of {S} sites ({n_spk} spikes) of {a.workload}, {dt:.1f} s"}


if __name__ == "__main__":
    main()
