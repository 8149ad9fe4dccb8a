"""irc_rho_delta_batch_noalloc through the C ABI: its outputs equal irc_rho_delta_batch's bit for bit
on two calls in a row (the second on a smaller recording, so the scratch array holds the first call's
contents), in device and host memory, and a recording too large for its scratch is rejected with
IRC_E_UNSUPPORTED (decided on the host before the buffer would be used)."""
import numpy as np
import pytest
import torch

from ironclust_amd import irc2, libirc, synth

pytestmark = pytest.mark.gpu


def _dev(ds, k):
    lay = irc2.site_layout(ds)
    N = ds["n_spikes"]
    dev = dict(F=lay["F"], col_spike=lay["col_spike"], site_ptr=lay["site_ptr"],
               site_n_primary=lay["site_n_primary"], lim_drift=lay["lim_drift"], link=lay["link"], n_spikes=N, k=k)
    out = dict(rho=torch.empty(N, dtype=torch.float32, device="cuda"),
               delta=torch.empty(N, dtype=torch.float32, device="cuda"),
               nneigh=torch.empty(N, dtype=torch.int64, device="cuda"),
               knn=torch.empty((N, k), dtype=torch.int64, device="cuda"),
               dk=torch.empty(N, dtype=torch.float32, device="cuda"))
    return lay, dev, out


@pytest.mark.parametrize("prec", [libirc.PREC_BF16_SCREEN, libirc.PREC_FP32, libirc.PREC_EXACT])
def test_noalloc_matches_batch_on_two_calls(prec):
    ctx = libirc.Context(0)
    try:
        for i, (N, S, D, C, k, seed) in enumerate(((30000, 6, 42, 20, 30, 81), (9000, 3, 24, 8, 16, 82))):
            print(f"noalloc call {i}: N={N} S={S} D={D} C={C} k={k} prec={prec}", flush=True)
            ds = synth.make_dataset(N, S, D, C, 6, k=k, drift_step=0.3, n_links=5, band=3, seed=seed, f2=True,
                                    device="cuda")
            lay, dev, out = _dev(ds, k)
            ctx.rho_delta_batch_device(dev, out, libirc.make_options(prec))
            ref = {kk: v.clone() for kk, v in out.items()}
            for v in out.values():
                v.fill_(-7)
            torch.cuda.synchronize()
            ctx.rho_delta_batch_noalloc_device(dev, out, libirc.make_options(prec))
            torch.cuda.synchronize()
            for kk in ref:
                assert torch.equal(out[kk], ref[kk]), (i, kk)
    finally:
        ctx.close()


def test_noalloc_host_memory_and_limit():
    ctx = libirc.Context(0)
    try:
        ds = synth.make_dataset(12000, 4, 16, 6, 5, k=12, drift_step=0.3, n_links=3, band=2, seed=83)
        lay = irc2.site_layout(ds)
        host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in lay.items()}
        ref = ctx.rho_delta_batch(host, 12, opts=libirc.make_options(libirc.PREC_BF16_SCREEN))
        import ctypes
        F = np.ascontiguousarray(host["F"], dtype=np.float32)
        cs = np.ascontiguousarray(host["col_spike"], dtype=np.int32)
        sp = np.ascontiguousarray(host["site_ptr"], dtype=np.int64)
        npri = np.ascontiguousarray(host["site_n_primary"], dtype=np.int64)
        lim = np.ascontiguousarray(host["lim_drift"], dtype=np.int64)
        lk = np.ascontiguousarray(np.asarray(host["link"], np.uint8).reshape(-1, order="F"))
        N = int(host["n_spikes"])
        b = libirc.Batch()
        b.n_spikes, b.D, b.k, b.n_sites, b.n_drift, b.n_cols = N, F.shape[1], 12, len(npri), len(lim) - 1, F.shape[0]
        b.F, b.col_spike = libirc._p(F, libirc._c_f32p), libirc._p(cs, libirc._c_i32p)
        b.site_ptr, b.site_n_primary = libirc._p(sp, libirc._c_i64p), libirc._p(npri, libirc._c_i64p)
        b.lim_drift, b.link = libirc._p(lim, libirc._c_i64p), libirc._p(lk, libirc._c_u8p)
        rho = np.zeros(N, np.float32)
        delta = np.zeros(N, np.float32)
        nn = np.zeros(N, np.int64)
        knn = np.zeros((N, 12), np.int64)
        outs = libirc.Outputs(libirc._p(rho, libirc._c_f32p), libirc._p(delta, libirc._c_f32p),
                              libirc._p(nn, libirc._c_i64p), libirc._p(knn, libirc._c_i64p), None)
        o = libirc.make_options(libirc.PREC_BF16_SCREEN)
        st = ctx.lib.irc_rho_delta_batch_noalloc(ctx.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(outs))
        assert st == libirc.IRC_OK, ctx.lib.irc_last_error(ctx.h)
        np.testing.assert_array_equal(rho, ref["rho"])
        np.testing.assert_array_equal(delta, ref["delta"])
        np.testing.assert_array_equal(nn, ref["nneigh"])
        np.testing.assert_array_equal(knn, ref["knn"])
        # too large for the scratch array: a clear error, outputs untouched, the context still usable
        big = synth.make_dataset(400000, 8, 42, 10, 6, k=30, drift_step=0.3, n_links=3, band=2, seed=84,
                                 device="cuda")
        lay_b, dev_b, out_b = _dev(big, 30)
        for v in out_b.values():
            v.fill_(-3)
        with pytest.raises(libirc.IrcError) as e:
            ctx.rho_delta_batch_noalloc_device(dev_b, out_b, libirc.make_options(libirc.PREC_BF16_SCREEN))
        assert e.value.status == libirc.IRC_E_UNSUPPORTED and "scratch" in str(e.value)
        torch.cuda.synchronize()
        assert bool((out_b["knn"] == -3).all()) and bool((out_b["rho"] == -3).all())
        with pytest.raises(libirc.IrcError) as e:
            ctx.rho_delta_batch_noalloc_device(dev_b, out_b, libirc.make_options(libirc.PREC_BF16_SCREEN,
                                                                                 flags=libirc.FLAG_GRAPH))
        assert e.value.status == libirc.IRC_E_ARG
        st = ctx.lib.irc_rho_delta_batch_noalloc(ctx.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(outs))
        assert st == libirc.IRC_OK
        np.testing.assert_array_equal(knn, ref["knn"])
    finally:
        ctx.close()
