"""GPU parity: the gfx950 library against the oracle through the C ABI, bit-exact for
indices, rho, delta and nneigh (the canonical values are defined in FP64 and the library
reports FP64-derived values whatever its screen precision)."""
import numpy as np
import pytest

from ironclust_amd import irc2, libirc, synth
from oracle import oracle as O
from tests.helpers import int_dataset

pytestmark = pytest.mark.gpu

PRECS = ["exact", "fp32", "bf16", "fp8"]


def _check_batch(ds, prec, ctxname=""):
    lay = irc2.site_layout({kk: (__import__("torch").as_tensor(v) if kk in ("F1", "F2", "site1", "site2")
                                 and v is not None else v) for kk, v in ds.items()})
    ref = O.sort_batch(ds)
    rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": ds["k"], "precision": prec})
    np.testing.assert_array_equal(knn.T, ref["knn"], err_msg=f"knn {ctxname} {prec}")
    np.testing.assert_array_equal(rho, ref["rho"], err_msg=f"rho {ctxname} {prec}")
    np.testing.assert_array_equal(nn, ref["nneigh"], err_msg=f"nneigh {ctxname} {prec}")
    np.testing.assert_array_equal(delta, ref["delta"], err_msg=f"delta {ctxname} {prec}")
    return ref


@pytest.mark.parametrize("prec", PRECS)
def test_batch_drift_with_secondary_sites(prec):
    ds = synth.make_dataset(12000, 5, 14, 10, 4, k=16, drift_step=0.3, n_links=4, band=3, seed=31, f2=True)
    _check_batch(ds, prec, "drift-f2")


@pytest.mark.parametrize("prec", PRECS)
def test_batch_plain_single_chunk(prec):
    ds = synth.make_dataset(8000, 3, 24, 1, 8, k=30, seed=32)
    _check_batch(ds, prec, "plain")


@pytest.mark.parametrize("prec", PRECS)
def test_batch_integer_ties(prec):
    # integer features: many exactly tied distances and duplicate points (d2 = 0, rho = inf)
    ds = int_dataset(N=1500, S=3, C=6, D=3, k=9, seed=33, f2=True, lo=-2, hi=3)
    _check_batch(ds, prec, "ties")


@pytest.mark.parametrize("prec", PRECS)
def test_batch_d42_k30(prec):
    ds = synth.make_dataset(30000, 4, 42, 20, 10, k=30, drift_step=0.3, n_links=15, band=8, seed=34)
    _check_batch(ds, prec, "d42")


@pytest.mark.parametrize("prec", PRECS)
def test_batch_fewer_candidates_than_k(prec):
    # sparse sites: many (site, chunk) blocks hold fewer than k candidates, some none
    ds = synth.make_dataset(600, 40, 6, 12, 2, k=20, drift_step=0.2, n_links=2, band=1, seed=35)
    _check_batch(ds, prec, "sparse")


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_rerank_second_pass(prec, ctx):
    # integer features in 6 dimensions: runs of tied distances longer than the first re-rank
    # pass's window send thousands of queries through the second (long-window) pass, and the
    # longest runs on to the exact kernel; every route must give the oracle's lists
    ds = int_dataset(N=20000, S=2, C=4, D=6, k=16, seed=62, lo=-3, hi=4)
    _check_batch(ds, prec, "rerank-second")
    lay = irc2.site_layout({kk: (__import__("torch").as_tensor(v) if kk in ("F1", "F2", "site1", "site2")
                                 and v is not None else v) for kk, v in ds.items()})
    host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in lay.items()}
    ctx.rho_delta_batch(host, ds["k"], opts=libirc.make_options(libirc.PREC_BF16_SCREEN if prec == "bf16"
                                                                  else libirc.PREC_FP32))
    st = ctx.stats()
    assert st["n_rerank_second"] > 100, st
    assert st["n_rerank_overflow"] <= st["n_rerank_second"] <= st["n_queries"], st


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batch_large_blocks_screened(prec, ctx):
    # blocks of 15k-22k candidates: stream positions need 15 bits of the screen key, so the
    # truncated keys are coarser; the screen must still certify almost every query
    ds = synth.make_dataset(45000, 2, 16, 3, 6, k=30, drift_step=0.2, n_links=2, band=1, seed=38)
    _check_batch(ds, prec, "large-blocks")
    lay = irc2.site_layout(ds)
    host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in lay.items()}
    ctx.rho_delta_batch(host, 30, opts=libirc.make_options(libirc.PREC_BF16_SCREEN if prec == "bf16"
                                                           else libirc.PREC_FP32))
    st = ctx.stats()
    assert st["precision_used"] == (libirc.PREC_BF16_SCREEN if prec == "bf16" else libirc.PREC_FP32)
    assert st["n_screen_fallback"] < 0.01 * ds["n_spikes"], st


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("large_ncand", [100, 30000])
def test_screen_list_depth_paths(prec, large_ncand):
    # blocks of 1k-22k candidates: large_ncand = 100 sends every block through the 64-deep list
    # body, 30000 keeps the 40/48-deep body past the finalize's byte table (binary-search lookup)
    ds = synth.make_dataset(45000, 2, 16, 6, 6, k=30, drift_step=0.2, n_links=3, band=2, seed=39)
    lay = irc2.site_layout(ds)
    ref = O.sort_batch(ds)
    rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": 30, "precision": prec, "large_ncand": large_ncand})
    np.testing.assert_array_equal(knn.T, ref["knn"])
    np.testing.assert_array_equal(rho, ref["rho"])
    np.testing.assert_array_equal(nn, ref["nneigh"])
    np.testing.assert_array_equal(delta, ref["delta"])


def test_batch_from_detect_files(tmp_path):
    # a real irc2 run's hand-over: _fet_/_fet2_ files of two loads -> site-grouped columns -> one
    # batch call -> per-site _knn_ files; results equal the oracle's on the same spikes
    from ironclust_amd import irc_files as IF
    from tests.helpers import write_detect_files
    ds = synth.make_dataset(4000, 6, 12, 5, 3, k=8, drift_step=0.3, n_links=2, band=2, seed=42, f2=True)
    prefix = str(tmp_path / "rec")
    cc1, cc2, dimm, ml = write_detect_files(ds, prefix, 3, 4)
    lay = IF.layout_from_fet_files(prefix, cc1, cc2, dimm, ml, ds["lim"], ds["link"], ds["n_spikes"], ds["k"])
    ref = O.sort_batch(ds)
    rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": ds["k"], "precision": "bf16"})
    np.testing.assert_array_equal(knn.T, ref["knn"])
    np.testing.assert_array_equal(rho, ref["rho"])
    np.testing.assert_array_equal(nn, ref["nneigh"])
    np.testing.assert_array_equal(delta, ref["delta"])
    IF.save_knn_by_site(prefix, lay, knn)
    for s in range(ds["n_sites"]):
        ids = lay["col_spike"][lay["site_ptr"][s]:lay["site_ptr"][s] + lay["site_n_primary"][s]]
        np.testing.assert_array_equal(IF.load_miKnn_site_(prefix, s + 1, ds["k"]), ref["knn"][ids - 1].T)


def test_assignment_matches_oracle():
    # assignCluster_knn_ on the batch outputs: host merge of overlapping peaks + GPU chain labels,
    # against the oracle's sequential restatement; the device-pointer form agrees with the host one
    import torch
    from oracle import assign as A
    ds = synth.make_dataset(60000, 8, 12, 10, 6, k=20, drift_step=0.3, n_links=3, band=2, seed=52)
    lay = irc2.site_layout(ds)
    rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": 20, "precision": "bf16"})
    for q in (0.99, 0.999):
        cut = float(np.quantile(delta, q))
        w_clu, w_peaks, w_pre = A.assign_cluster_knn(rho, delta, nn, knn, delta_cut=cut)
        g_clu, g_peaks, g_pre = irc2.assignCluster_knn_({"rho": rho, "delta": delta, "nneigh": nn},
                                                        {"delta_cut": cut}, knn)
        assert g_pre == w_pre
        np.testing.assert_array_equal(g_peaks, w_peaks)
        np.testing.assert_array_equal(g_clu, w_clu)
    lab0 = np.zeros(len(rho), np.int32)
    pk = np.flatnonzero(delta >= cut)
    lab0[pk] = np.arange(1, len(pk) + 1, dtype=np.int32)
    ctx = irc2.get_ctx(0)
    host = ctx.assign_chain(nn, lab0)
    dev = ctx.assign_chain(torch.as_tensor(nn, dtype=torch.int64).cuda(), torch.as_tensor(lab0).cuda())
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    assert host.min() >= 1


def test_isolation_and_merge_match_oracle():
    from oracle import assign as A
    ds = synth.make_dataset(40000, 6, 12, 8, 6, k=16, drift_step=0.3, n_links=3, band=2, seed=53)
    lay = irc2.site_layout(ds)
    rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": 16, "precision": "bf16"})
    clu, _, _ = A.assign_cluster_knn(rho, delta, nn, knn, delta_cut=float(np.quantile(delta, 0.995)))
    clu = clu.copy()
    clu[::97] = 0                                   # some spikes outside every cluster
    w_s, w_nn = A.isolation_score(clu, knn)
    g_s, g_nn = irc2.isolation_score_(clu, knn)
    np.testing.assert_array_equal(np.isnan(g_s), np.isnan(w_s))
    np.testing.assert_array_equal(g_s[~np.isnan(w_s)], w_s[~np.isnan(w_s)])
    np.testing.assert_array_equal(np.isnan(g_nn), np.isnan(w_nn))
    np.testing.assert_array_equal(g_nn[~np.isnan(w_nn)], w_nn[~np.isnan(w_nn)])
    for th in (0.5, 0.8, 0.95):
        np.testing.assert_array_equal(irc2.knn_overlap_merge_(clu, knn, th), A.knn_overlap_merge(clu, knn, th))


def test_graph_replay_matches_eager():
    """IRC_FLAG_GRAPH: the first call runs and captures the device phase, later calls replay it.
    Every replay (with unrelated eager work and synchronizations in between) must equal the eager
    call, also after the features change in place."""
    import torch
    ds = synth.make_dataset(24000, 4, 16, 8, 5, k=12, drift_step=0.3, n_links=3, band=2, seed=39, f2=True,
                            device="cuda")
    lay = irc2.site_layout(ds)
    N, k = ds["n_spikes"], ds["k"]
    dev = dict(F=lay["F"], col_spike=lay["col_spike"], site_ptr=lay["site_ptr"],
               site_n_primary=lay["site_n_primary"], lim_drift=lay["lim_drift"], link=lay["link"], n_spikes=N, k=k)
    out = dict(rho=torch.empty(N, dtype=torch.float32, device="cuda"),
               delta=torch.empty(N, dtype=torch.float32, device="cuda"),
               nneigh=torch.empty(N, dtype=torch.int64, device="cuda"),
               knn=torch.empty((N, k), dtype=torch.int64, device="cuda"))
    ctx = libirc.Context(0)
    try:
        for feats_round in range(2):
            torch.cuda.synchronize()
            ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN))
            ref = {kk: v.clone() for kk, v in out.items()}
            for rep in range(4):
                for v in out.values():
                    v.zero_()
                junk = torch.randn(1 << 20, device="cuda") * 2.0          # eager work between replays
                torch.cuda.synchronize()
                ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN,
                                                                         flags=libirc.FLAG_GRAPH))
                st = ctx.stats()
                for kk in ref:
                    assert torch.equal(out[kk], ref[kk]), (feats_round, rep, kk)
                if rep > 0 or feats_round > 0:
                    assert st["ms_graph"] > 0.0, st
                del junk
            if feats_round == 0:
                # the features change in place: the replays must read the new values
                torch.cuda.synchronize()
                dev["F"].mul_(1.25)
        r = O.sort_batch(dict(ds, F1=ds["F1"].cpu() * 1.25, F2=ds["F2"].cpu() * 1.25,
                              site1=ds["site1"].cpu(), site2=ds["site2"].cpu()))
        np.testing.assert_array_equal(out["rho"].cpu().numpy(), r["rho"])
        np.testing.assert_array_equal(out["knn"].cpu().numpy(), r["knn"])
    finally:
        ctx.close()


def test_batch_odd_dims_and_k():
    for D, k in ((1, 1), (5, 3), (45, 33), (62, 64), (64, 7)):
        ds = synth.make_dataset(3000, 2, D, 3, 3, k=k, drift_step=0.1, n_links=2, band=1, seed=36 + D)
        for prec in PRECS:
            _check_batch(ds, prec, f"D{D}k{k}")


def test_site_entry_points(ctx):
    ds = int_dataset(N=900, S=1, C=5, D=4, k=8, seed=41, f2=False, lo=-3, hi=4)
    F = np.asarray(ds["F1"])
    drift = np.searchsorted(ds["lim"], np.arange(1, 901), side="right").astype(np.int32)
    for (a, n) in ((0, 900), (100, 300), (899, 1)):
        dk_r, I_r = O.knn_site(F, drift, ds["link"], a, n, 8)
        dk_g, I_g = ctx.knn_drift(F, drift, ds["link"], a, n, 8, opts=libirc.make_options(libirc.PREC_EXACT))
        np.testing.assert_array_equal(I_g, I_r)
        np.testing.assert_array_equal(dk_g, dk_r)
    rho = np.random.default_rng(0).random(900).astype(np.float32)
    dm_r, nn_r = O.delta_site(F, drift, ds["link"], 50, 700, rho)
    dm_g, nn_g = ctx.delta_drift(F, drift, ds["link"], 50, 700, rho)
    np.testing.assert_array_equal(nn_g, nn_r)
    np.testing.assert_array_equal(dm_g, dm_r)


def test_site_entry_unsorted_drift_and_ids(ctx):
    # columns in arbitrary drift order with explicit spike ids for tie-breaking
    rng = np.random.default_rng(5)
    n = 700
    F = rng.integers(-2, 3, size=(n, 3)).astype(np.float32)
    drift = rng.integers(1, 5, size=n).astype(np.int32)
    ids = rng.permutation(n).astype(np.int64) + 1000
    link = (rng.random((4, 4)) < 0.6).astype(np.uint8)
    np.fill_diagonal(link, 1)
    dk_r, I_r = O.knn_site(F, drift, link, 0, n, 6, id_of_col=ids)
    dk_g, I_g = ctx.knn_drift(F, drift, link, 0, n, 6, id_of_col=ids)
    np.testing.assert_array_equal(I_g, I_r)
    np.testing.assert_array_equal(dk_g, dk_r)


def test_mirror_functions_match_reference_orientation():
    ds = synth.make_dataset(2000, 1, 9, 4, 3, k=10, drift_step=0.3, n_links=2, band=1, seed=43)
    F = ds["F1"].numpy()
    drift = np.searchsorted(ds["lim"], np.arange(1, 2001), side="right").astype(np.int32)
    vl = np.zeros(2000, bool)
    vl[500:1500] = True
    vrKnn, miKnn = irc2.search_knn_drift_(vl, F.T, drift, ds["link"].astype(bool), {"knn": 10})
    dk_r, I_r = O.knn_site(F, drift, ds["link"], 500, 1000, 10)
    assert miKnn.shape == (10, 1000)
    np.testing.assert_array_equal(miKnn.T, I_r)
    np.testing.assert_array_equal(vrKnn, dk_r)
    rho = (1.0 / np.maximum(vrKnn, 1e-30)).astype(np.float32)
    rho_out = np.random.default_rng(1).random(2000).astype(np.float32)
    rho_out[500:1500] = rho
    d_g, n_g = irc2.search_delta_drift_(vl, F.T, rho_out, drift, ds["link"].astype(bool), {"knn": 10})
    d_r, n_r = O.delta_site(F, drift, ds["link"], 500, 1000, rho_out)
    np.testing.assert_array_equal(n_g, n_r)
    np.testing.assert_array_equal(d_g, d_r)
    K, fGpu, I = irc2.cuda_knn_(F[:300].T, F[300:400].T, {"knn": 7})
    K_r, I_r = O.knn_ab(F[:300], F[300:400], 7)
    np.testing.assert_array_equal(I.T, I_r)
    np.testing.assert_array_equal(K, K_r)


def test_ab_and_dc_entry_points(ctx):
    rng = np.random.default_rng(7)
    B = rng.standard_normal((500, 11)).astype(np.float32)
    A = rng.standard_normal((200, 11)).astype(np.float32)
    for k in (1, 5, 30):
        K_g, I_g = ctx.knn_ab(B, A, k)
        K_r, I_r = O.knn_ab(B, A, k)
        np.testing.assert_array_equal(I_g, I_r)
        np.testing.assert_array_equal(K_g, K_r)
    K_g, I_n = ctx.knn_ab(B[:3], A, 5, want_index=False)
    np.testing.assert_array_equal(K_g, O.knn_ab(B[:3], A, 5)[0])
    RB = rng.random(500).astype(np.float32)
    RA = rng.random(200).astype(np.float32)
    np.testing.assert_array_equal(np.stack(ctx.delta_knn_ab(B, RB, A, RA)), np.stack(O.delta_knn_ab(B, RB, A, RA)))
    n12, n1 = 3000, 1200
    site = np.stack([rng.integers(1, 6, n12), rng.integers(1, 6, n12)], 1).astype(np.int32)
    f1 = rng.standard_normal((n12, 9)).astype(np.float32)
    f2 = rng.standard_normal((n12, 9)).astype(np.float32)
    rho_g = ctx.rho_drift_dc(n1, site, f1, f2, 6.0)
    rho_r = O.rho_drift_dc(n1, site, f1, f2, 6.0)
    np.testing.assert_array_equal(rho_g, rho_r)
    rho12 = np.concatenate([rho_r, rng.random(n12 - n1).astype(np.float32)])
    dl_g, nn_g = ctx.delta_drift_dc(n1, site, f1, f2, rho12, 6.0)
    dl_r, nn_r = O.delta_drift_dc(n1, site, f1, f2, rho12, 6.0)
    np.testing.assert_array_equal(nn_g, nn_r)
    np.testing.assert_array_equal(dl_g, dl_r)


def test_golden_fixtures():
    import glob
    import os
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
    assert files
    for f in files:
        z = np.load(f)
        if "site_ptr" not in z:
            continue
        lay = dict(F=z["F"], col_spike=z["col_spike"], site_ptr=z["site_ptr"], site_n_primary=z["site_n_primary"],
                   lim_drift=z["lim"], link=z["link"], n_spikes=int(z["n_spikes"]), k=int(z["k"]))
        for prec in PRECS:
            rho, delta, nn, knn = irc2.sort_page_(lay, {"knn": int(z["k"]), "precision": prec})
            np.testing.assert_array_equal(knn.T, z["knn"], err_msg=f)
            np.testing.assert_array_equal(rho, z["rho"], err_msg=f)
            np.testing.assert_array_equal(delta, z["delta"], err_msg=f)
            np.testing.assert_array_equal(nn, z["nneigh"], err_msg=f)


def test_force_fail_and_errors(ctx):
    ds = synth.make_dataset(500, 2, 4, 2, 2, k=5, seed=50)
    lay = irc2.site_layout(ds)
    host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in lay.items()}
    with pytest.raises(libirc.IrcError) as e:
        ctx.rho_delta_batch(host, 5, opts=libirc.make_options(force_fail=True))
    assert e.value.status == libirc.IRC_E_FORCED
    # a bad descriptor is rejected cleanly and the context stays usable
    bad = dict(host, site_n_primary=host["site_n_primary"] + 1)
    with pytest.raises(libirc.IrcError) as e:
        ctx.rho_delta_batch(bad, 5)
    assert e.value.status == libirc.IRC_E_ARG
    with pytest.raises(libirc.IrcError):
        ctx.rho_delta_batch(host, 65)          # k above the supported maximum
    r = ctx.rho_delta_batch(host, 5)
    ref = O.sort_batch(ds)
    np.testing.assert_array_equal(r["rho"], ref["rho"])


def test_fp64_sqrt_div_rounding_matches_host(ctx):
    # rho = float(1/sqrt(d2)) relies on correctly rounded FP64 sqrt and division on the GPU;
    # the lattice-free A/B form with one candidate reports float(sqrt(d2)) for random d2.
    rng = np.random.default_rng(9)
    A = (rng.standard_normal((4000, 3)) * rng.choice([1e-3, 1.0, 1e3], size=(4000, 1))).astype(np.float32)
    B = np.zeros((1, 3), np.float32)
    K_g, _ = ctx.knn_ab(B, A, 1)
    K_r, _ = O.knn_ab(B, A, 1)
    np.testing.assert_array_equal(K_g, K_r)


def test_mfma_lane_layout(ctx):
    """The screen relies on the 32x32 MFMA output layout (register r of lane l holds row
    8*(r//4) + 4*(l//32) + r%4, column l%32) and on v_permlane32_swap exchanging the upper
    half of its first operand with the lower half of its second."""
    out = ctx.selftest_mfma().reshape(6, 16, 64)
    r = np.arange(16)[:, None]
    l = np.arange(64)[None, :]
    row = 8 * (r // 4) + 4 * (l // 32) + r % 4
    col = l % 32
    want = (row + 1) * 64.0 * (col + 1)
    np.testing.assert_array_equal(out[0], want, err_msg="bf16 32x32x16 layout")
    np.testing.assert_array_equal(out[1], want, err_msg="f32 32x32x2 layout")
    acc0 = want
    acc1 = want + 0.5
    first = np.where(l < 32, acc0, np.roll(acc1, 32, axis=1))
    second = np.where(l < 32, np.roll(acc0, -32, axis=1), acc1)
    np.testing.assert_array_equal(out[2], first, err_msg="permlane32_swap first result")
    np.testing.assert_array_equal(out[3], second, err_msg="permlane32_swap second result")
    # 32x32x64 e4m3 (OCP: bias 7, subnormals m * 2^-9): row m's 32 bytes of one lane half all hold
    # code 4m + 1, column n's code (8 + n) with the sign of odd n; the product sums 32 equal terms
    def e4m3(c):
        s = -1.0 if c & 0x80 else 1.0
        e, mt = (c >> 3) & 15, c & 7
        return s * (mt * 2.0 ** -9 if e == 0 else (1 + mt / 8) * 2.0 ** (e - 7))
    xa = np.array([e4m3(4 * i + 1) for i in range(32)])
    xb = np.array([e4m3((8 + i) | (0x80 if i & 1 else 0)) for i in range(32)])
    want8 = (32.0 * xa[row] * xb[col]).astype(np.float32)
    np.testing.assert_array_equal(out[4], want8, err_msg="fp8 32x32x64 layout and e4m3 decoding (half 0)")
    np.testing.assert_array_equal(out[5], want8, err_msg="fp8 32x32x64 layout and e4m3 decoding (half 1)")


def _bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("prec,K", [(1, 48), (1, 64), (2, 48), (2, 64)])
def test_mfma_accumulation_within_the_screen_bound(ctx, prec, K):
    """The certified screen bound (screen_eta) models every FP32 accumulation inside the matrix core
    with unit roundoff 2^-23, twice FP32's.  Measured here on the screen's own MFMA forms (FP32
    32x32x2 chains; BF16 32x32x16 chains, whose bf16 products are exact): every output of random tiles
    with rows and columns scaled over 2^+-12, of tiles whose products cancel exactly, and of tiles
    with entries scaled over 2^+-20 one by one stays within gamma_{K+2}(2^-23) * sum |a_i b_i| of the
    exact sum."""
    rng = np.random.default_rng(100 + 10 * prec + K)
    T = 384
    A = rng.standard_normal((T, 32, K)) * 2.0 ** rng.integers(-12, 13, size=(T, 32, 1))
    B = rng.standard_normal((T, K, 32)) * 2.0 ** rng.integers(-12, 13, size=(T, 1, 32))
    c = slice(128, 256)   # exact cancellation: the second half of k repeats the first with B negated
    A[c, :, K // 2:] = A[c, :, :K // 2]
    B[c, K // 2:, :] = -B[c, :K // 2, :]
    w = slice(256, T)     # entries scaled one by one
    A[w] *= 2.0 ** rng.integers(-20, 21, size=A[w].shape)
    B[w] *= 2.0 ** rng.integers(-20, 21, size=B[w].shape)
    A = A.astype(np.float32)
    B = B.astype(np.float32)
    if prec == 2:
        A, B = _bf16(A), _bf16(B)
    out = ctx.selftest_mfma_accum(prec, A, B).astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    exact = np.matmul(A64, B64)
    scale = np.matmul(np.abs(A64), np.abs(B64))
    u = 2.0 ** -23
    gamma = (K + 2) * u / (1 - (K + 2) * u)
    err = np.abs(out - exact)
    ratio = err / (u * scale)
    print(f"prec {prec} K {K}: largest error {ratio.max():.3f} u * sum|ab| (bound {gamma / u:.1f}); "
          f"cancelling tiles: largest |result| {err[c].max() / scale[c].max():.2e} of sum|ab|")
    assert (err <= gamma * scale).all(), ratio.max()
    assert np.abs(out[c]).max() <= gamma * scale[c].max()


def test_mfma_accum_selftest_rejects_bad_shapes(ctx):
    A = np.ones((2, 32, 47), np.float32)
    B = np.ones((2, 47, 32), np.float32)
    with pytest.raises(libirc.IrcError):
        ctx.selftest_mfma_accum(1, A, B)                      # odd K for the FP32 form
    A, B = np.ones((2, 32, 40), np.float32), np.ones((2, 40, 32), np.float32)
    with pytest.raises(libirc.IrcError):
        ctx.selftest_mfma_accum(2, A, B)                      # K not a multiple of 16 for BF16
    with pytest.raises(libirc.IrcError):
        ctx.selftest_mfma_accum(3, A, B)                      # no such form
    np.testing.assert_array_equal(ctx.selftest_mfma_accum(1, A, B), np.full((2, 32, 32), 40.0, np.float32))


def _dev_batch(ds, k):
    import torch
    lay = irc2.site_layout(ds)
    N = ds["n_spikes"]
    dev = dict(F=lay["F"], col_spike=lay["col_spike"], site_ptr=lay["site_ptr"],
               site_n_primary=lay["site_n_primary"], lim_drift=lay["lim_drift"], link=lay["link"], n_spikes=N, k=k)
    out = dict(rho=torch.empty(N, dtype=torch.float32, device="cuda"),
               delta=torch.empty(N, dtype=torch.float32, device="cuda"),
               nneigh=torch.empty(N, dtype=torch.int64, device="cuda"),
               knn=torch.empty((N, k), dtype=torch.int64, device="cuda"))
    return dev, out


def test_host_buffers_after_device_calls():
    """A fresh context used in device-pointer form first, then with host buffers: the host form's own
    copies of F and col_spike must be allocated before the set-up kernels read them."""
    ds = synth.make_dataset(5000, 3, 8, 4, 3, k=7, drift_step=0.3, n_links=2, band=1, seed=70, device="cuda")
    dev, out = _dev_batch(ds, 7)
    ctx = libirc.Context(0)
    try:
        ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN))
        host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in irc2.site_layout(ds).items()}
        r = ctx.rho_delta_batch(host, 7, opts=libirc.make_options(libirc.PREC_BF16_SCREEN))
        ref = O.sort_batch(dict(ds, F1=ds["F1"].cpu(), site1=ds["site1"].cpu()))
        np.testing.assert_array_equal(r["knn"], ref["knn"])
        np.testing.assert_array_equal(r["delta"], ref["delta"])
        np.testing.assert_array_equal(out["knn"].cpu().numpy(), ref["knn"])
    finally:
        ctx.close()


def test_graph_replay_after_buffer_growth():
    """A replay must not run on context buffers that another call reallocated: capture with
    IRC_FLAG_GRAPH, grow buffers with other calls (a larger A/B kNN, an FP32 batch on a larger
    recording), replay, and compare with the eager call and the oracle."""
    import torch
    ds = synth.make_dataset(20000, 4, 12, 8, 5, k=10, drift_step=0.3, n_links=3, band=2, seed=71, device="cuda")
    big = synth.make_dataset(60000, 4, 12, 8, 5, k=16, drift_step=0.3, n_links=3, band=2, seed=72, device="cuda")
    dev, out = _dev_batch(ds, 10)
    bdev, bout = _dev_batch(big, 16)
    ctx = libirc.Context(0)
    try:
        g = libirc.make_options(libirc.PREC_BF16_SCREEN, flags=libirc.FLAG_GRAPH)
        ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN))
        ref = {kk: v.clone() for kk, v in out.items()}
        ctx.rho_delta_batch_device(dev, out, g)            # runs and captures
        ctx.rho_delta_batch_device(dev, out, g)            # replays
        assert ctx.stats()["ms_graph"] > 0.0
        rng = np.random.default_rng(3)
        B = rng.standard_normal((30000, 12)).astype(np.float32)
        A = rng.standard_normal((40000, 12)).astype(np.float32)
        for grow in range(2):
            if grow == 0:
                ctx.knn_ab(B, A, 30)                               # grows the kNN and plan buffers
            else:
                ctx.rho_delta_batch_device(bdev, bout, libirc.make_options(libirc.PREC_FP32))
            for v in out.values():
                v.zero_()
            torch.cuda.synchronize()
            ctx.rho_delta_batch_device(dev, out, g)
            for kk in ref:
                assert torch.equal(out[kk], ref[kk]), (grow, kk)
            ctx.rho_delta_batch_device(dev, out, g)            # and the fresh capture replays
            assert ctx.stats()["ms_graph"] > 0.0
            for kk in ref:
                assert torch.equal(out[kk], ref[kk]), (grow, kk)
        r = O.sort_batch(dict(ds, F1=ds["F1"].cpu(), site1=ds["site1"].cpu()))
        np.testing.assert_array_equal(out["knn"].cpu().numpy(), r["knn"])
        np.testing.assert_array_equal(out["delta"].cpu().numpy(), r["delta"])
    finally:
        ctx.close()


def test_site_launches_match_persistent_launch():
    ds = synth.make_dataset(30000, 6, 14, 10, 5, k=12, drift_step=0.3, n_links=4, band=3, seed=73, lognormal=True,
                            device="cuda")
    dev, out = _dev_batch(ds, 12)
    ctx = irc2.get_ctx(0)
    ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN))
    ref = {kk: v.clone() for kk, v in out.items()}
    ctx.rho_delta_batch_device(dev, out, libirc.make_options(libirc.PREC_BF16_SCREEN,
                                                             flags=libirc.FLAG_SITE_LAUNCHES))
    for kk in ref:
        assert bool((out[kk] == ref[kk]).all()), kk
    with pytest.raises(libirc.IrcError):
        ctx.rho_delta_batch_device(dev, out, libirc.make_options(
            libirc.PREC_BF16_SCREEN, flags=libirc.FLAG_SITE_LAUNCHES | libirc.FLAG_GRAPH))


@pytest.mark.parametrize("prec", [libirc.PREC_BF16_SCREEN, libirc.PREC_FP32])
def test_autotune_choice_reproduces_default(prec, tmp_path):
    """irc_autotune: every trial setting gives the base setting's outputs bit for bit, the choice is
    one of them, and batch calls with the choice, with IRC_FLAG_AUTOTUNE and captured in a graph after
    tuning all equal the oracle."""
    import torch
    from ironclust_amd import autotune
    ds = synth.make_dataset(40000, 5, 16, 10, 6, k=20, drift_step=0.3, n_links=4, band=3, seed=74, device="cuda")
    dev, out = _dev_batch(ds, 20)
    ctx = libirc.Context(0)
    try:
        choice, rep = autotune.tune(ctx, dev, out, precision=prec, cache=str(tmp_path / "tune.json"), force=True)
        assert len(rep["trials"]) >= 5
        assert all(t["identical"] for t in rep["trials"]), rep["trials"]
        r = O.sort_batch(dict(ds, F1=ds["F1"].cpu(), site1=ds["site1"].cpu()))
        for opts in (libirc.make_options(prec, **choice), libirc.make_options(prec, flags=libirc.FLAG_AUTOTUNE),
                     libirc.make_options(prec, flags=libirc.FLAG_GRAPH, **choice),
                     libirc.make_options(prec, flags=libirc.FLAG_GRAPH, **choice)):
            for v in out.values():
                v.zero_()
            ctx.rho_delta_batch_device(dev, out, opts)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(out["knn"].cpu().numpy(), r["knn"])
            np.testing.assert_array_equal(out["rho"].cpu().numpy(), r["rho"])
            np.testing.assert_array_equal(out["nneigh"].cpu().numpy(), r["nneigh"])
            np.testing.assert_array_equal(out["delta"].cpu().numpy(), r["delta"])
        again, rep2 = autotune.tune(ctx, dev, out, precision=prec, cache=str(tmp_path / "tune.json"))
        assert rep2["cached"] and again == choice
    finally:
        ctx.close()


def test_ab_full_and_indirect_forms(ctx):
    rng = np.random.default_rng(8)
    B = rng.standard_normal((700, 9)).astype(np.float32)
    A = rng.standard_normal((300, 9)).astype(np.float32)
    for k in (1, 6, 30):
        K_g, I_g = ctx.knn_ab_full(B, A, k)
        Ka, _, I_r = O.knn_ab_list(B, A, k)
        np.testing.assert_array_equal(I_g, I_r)
        np.testing.assert_array_equal(K_g, Ka)
    IB = np.concatenate([rng.permutation(700)[:250] + 1, [17, 17, 3]]).astype(np.int32)
    for k in (1, 9, 30):
        K_g, I_g = ctx.knn_ab_ib(B, IB, A, k)
        _, K_r, I_r = O.knn_ab_list(B, A, k, IB=IB)
        np.testing.assert_array_equal(I_g, I_r)
        np.testing.assert_array_equal(K_g, K_r)
    # fewer listed candidates than k: padded with the last one
    K_g, I_g = ctx.knn_ab_ib(B, np.array([5, 2], np.int32), A, 4)
    _, K_r, I_r = O.knn_ab_list(B, A, 4, IB=np.array([5, 2], np.int32))
    np.testing.assert_array_equal(I_g, I_r)
    np.testing.assert_array_equal(K_g, K_r)
    with pytest.raises(libirc.IrcError):
        ctx.knn_ab_ib(B, np.array([0, 701], np.int32), A, 4)
    mK, mI = irc2.cuda_knn_full_(B.T, A.T, {"knn": 6})
    assert mK.shape == (6, 300) and mI.shape == (6, 300)
    np.testing.assert_array_equal(mI.T, O.knn_ab_list(B, A, 6)[2])
    vK, vI = irc2.cuda_knn_ib_(B.T, IB, A.T, {"knn": 5})
    np.testing.assert_array_equal(vI.T, O.knn_ab_list(B, A, 5, IB=IB)[2])


def test_sort_chain_matches_oracle(tmp_path):
    """sort_: drift linkage on the GPU, the batch, rho / max(rho) and ordrho (the NaN rule of ordrho_
    is pinned on the CPU), with the features in memory and from detect files (two loads, secondary
    sites; the per-site _knn_ files are written on the way); the cache round-trips."""
    import torch
    from ironclust_amd import irc_files as IF
    from tests.helpers import write_detect_files
    rng = np.random.default_rng(5)
    N = 6000
    ds = synth.make_dataset(N, 4, 6, 1, 5, k=8, seed=75, f2=True)
    amp = rng.gamma(2.0, 20.0, size=N).astype(np.float32)
    pos = np.stack([np.zeros(N), rng.uniform(0, 400, size=N)], 1).astype(np.float32)
    P = {"nTime_batch": 3, "nTime_drift": 12, "nQuantile_drift": 6, "nSites": 8, "knn": 8, "precision": "bf16"}
    S0 = dict(vrAmp_spk=amp, mrPos_spk=pos, F1=ds["F1"].numpy(), viSite_spk=ds["site1"].numpy(),
              F2=ds["F2"].numpy(), viSite2_spk=ds["site2"].numpy(), nSites=4)
    S_clu = irc2.sort_(S0, P)
    S_drift = S_clu["S_drift"]
    assert S_drift["mlDrift"].shape == (12, 12)
    r = O.sort_batch(dict(ds, lim=S_drift["viLim_drift"], link=S_drift["mlDrift"].astype(np.uint8), n_drift=12))
    rho_n = (r["rho"] / r["rho"].max()).astype(np.float32)
    np.testing.assert_array_equal(S_clu["rho"], rho_n)
    np.testing.assert_array_equal(S_clu["delta"], r["delta"])
    np.testing.assert_array_equal(S_clu["nneigh"], r["nneigh"])
    assert S_clu["ordrho"].dtype == np.float64
    want = sorted(range(N), key=lambda i: (-float(rho_n[i]), i))
    np.testing.assert_array_equal(S_clu["ordrho"].astype(np.int64) - 1, want)
    prefix = str(tmp_path / "rec")
    cc1, cc2, dimm, ml = write_detect_files(ds, prefix, 2, 3)
    S0f = dict(vrAmp_spk=amp, mrPos_spk=pos, vcFile_prm=prefix, ccviSpk_site_load=cc1, ccviSpk_site2_load=cc2,
               dimm_fet=dimm, mlPc=ml)
    S_f = irc2.sort_(S0f, dict(P, vcFile_prm=prefix))
    for kk in ("rho", "delta", "nneigh", "ordrho"):
        np.testing.assert_array_equal(S_f[kk], S_clu[kk])
    s1 = ds["site1"].numpy()
    for s in range(1, 5):
        np.testing.assert_array_equal(IF.load_miKnn_site_(prefix, s, 8), r["knn"][np.flatnonzero(s1 == s)].T)
    f = str(tmp_path / "sort_irc.mat")
    IF.save_sort_(S_f, f)
    back = IF.load_sort_(f)
    for kk in ("rho", "delta", "nneigh", "ordrho"):
        np.testing.assert_array_equal(np.ravel(back[kk]), S_f[kk])
