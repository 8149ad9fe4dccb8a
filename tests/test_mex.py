"""The MEX gateway (ironclust_amd/csrc/mex/irc_hip_mex.c) compiled against the mock MEX API of
tests/mex_mock: every array whose length disagrees with the sizes derived from the others is rejected
with irc:hip:arg before the library is called (CPU), and a well-formed rho_delta_page call returns the
oracle's results with int64 nneigh and miKnn (GPU)."""
import numpy as np
import pytest

from tests.mexcall import MexError, mex


def _page_args():
    rng = np.random.default_rng(1)
    D, n = 4, 40
    F = rng.standard_normal((D, n)).astype(np.float32)          # mrFet, D x n_cols
    spk = np.arange(1, n + 1, dtype=np.int32)                    # one site, primary spikes only
    site_ptr = np.array([0, n], np.int64)
    n_prim = np.array([n], np.int64)
    lim = np.array([1, 21, 41], np.int64)
    link = np.ones((2, 2), bool)
    return [F, spk, site_ptr, n_prim, lim, link, np.float64(5)]


def _raises(args, cmd="rho_delta_page", nlhs=4):
    with pytest.raises(MexError) as e:
        mex(cmd, *args, nlhs=nlhs)
    assert e.value.ident == "irc:hip:arg", e.value
    return str(e.value)


def test_rho_delta_page_length_checks():
    a = _page_args()
    bad = list(a); bad[1] = a[1][:-1]
    assert "viSpk_col" in _raises(bad)
    bad = list(a); bad[2] = np.array([0, 20, 40], np.int64)
    assert "viSite_ptr" in _raises(bad)
    bad = list(a); bad[2] = np.array([0, 39], np.int64)
    assert "viSite_ptr(end)" in _raises(bad)
    bad = list(a); bad[5] = np.ones((3, 3), bool)
    assert "mlDrift" in _raises(bad)
    bad = list(a); bad[4] = np.array([1], np.int64)
    assert "viLim_drift" in _raises(bad)
    bad = list(a); bad[3] = np.zeros(0, np.int64)
    _raises(bad)
    assert "options" in _raises(a + [np.zeros(7)])
    assert "options" in _raises(a + [np.zeros(0)], "autotune", 2)
    bad = list(a); bad[1] = a[1][:-1]
    assert "viSpk_col" in _raises(bad, "autotune", 2)


def test_site_call_length_checks():
    rng = np.random.default_rng(2)
    F = rng.standard_normal((3, 30)).astype(np.float32)
    vl = np.zeros(30, bool); vl[5:15] = True
    drift = np.ones(30, np.int32)
    link = np.ones((1, 1), bool)
    assert "vl_in" in _raises([vl[:-1], F, drift, link, np.float64(4)], "knn_drift", 2)
    assert "viDrift_out" in _raises([vl, F, drift[:-2], link, np.float64(4)], "knn_drift", 2)
    assert "square" in _raises([vl, F, drift, np.ones((1, 2), bool), np.float64(4)], "knn_drift", 2)
    gap = vl.copy(); gap[9] = False
    assert "contiguous" in _raises([gap, F, drift, link, np.float64(4)], "knn_drift", 2)
    rho = np.ones(29, np.float32)
    assert "vrRho_out" in _raises([vl, F, rho, drift, link], "delta_drift", 2)


def test_ab_and_unknown_command_checks():
    B = np.ones((4, 10), np.float32)
    A = np.ones((3, 5), np.float32)
    assert "rows" in _raises([B, A, np.float64(3)], "knn_ab", 2)
    assert "rows" in _raises([B, np.array([1, 2], np.int32), A, np.float64(3)], "knn_ab_ib", 2)
    _raises([B, B, np.float64(0)], "knn_ab", 2)
    with pytest.raises(MexError) as e:
        mex("no_such_command", nlhs=1)
    assert e.value.ident == "irc:hip:arg"


@pytest.mark.gpu
def test_rho_delta_page_through_the_gateway():
    from ironclust_amd import irc2, synth
    from oracle import oracle as O
    ds = synth.make_dataset(3000, 3, 8, 4, 3, k=9, drift_step=0.3, n_links=2, band=1, seed=12, f2=True)
    lay = irc2.site_layout(ds)
    F = np.ascontiguousarray(lay["F"].numpy().T)                 # D x n_cols
    rho, delta, nn, knn = mex("rho_delta_page", F, lay["col_spike"].numpy(), lay["site_ptr"],
                              lay["site_n_primary"], lay["lim_drift"], lay["link"].astype(bool),
                              np.float64(9), np.float64(2), nlhs=4)
    ref = O.sort_batch(ds)
    assert nn.dtype == np.int64 and knn.dtype == np.int64 and knn.shape == (9, 3000)
    np.testing.assert_array_equal(rho.ravel(), ref["rho"])
    np.testing.assert_array_equal(delta.ravel(), ref["delta"])
    np.testing.assert_array_equal(nn.ravel(), ref["nneigh"])
    np.testing.assert_array_equal(knn.T, ref["knn"])


@pytest.mark.gpu
def test_autotune_through_the_gateway():
    """'autotune' returns the chosen settings and its trials; rho_delta_page with those settings, and
    with the autotune flag (8) that uses the context's choice, reproduces the oracle."""
    from ironclust_amd import irc2, synth
    from oracle import oracle as O
    ds = synth.make_dataset(6000, 3, 8, 4, 3, k=9, drift_step=0.3, n_links=2, band=1, seed=13, f2=True)
    lay = irc2.site_layout(ds)
    args = [np.ascontiguousarray(lay["F"].numpy().T), lay["col_spike"].numpy(), lay["site_ptr"],
            lay["site_n_primary"], lay["lim_drift"], lay["link"].astype(bool), np.float64(9)]
    mex("reset", nlhs=0)
    opt, trials = mex("autotune", *args, np.array([2.0]), nlhs=2)
    assert opt.shape == (1, 6) and opt[0, 0] == 2 and opt[0, 1] == 0
    assert trials.shape[1] == 7 and trials.shape[0] >= 2
    assert trials[:, 6].sum() == 1 and trials[trials[:, 6] == 1, 5] == 1
    assert (trials[:, 4] > 0).all()
    ref = O.sort_batch(ds)
    for o in (opt.ravel(), np.array([2.0, 8.0])):
        rho, delta, nn, knn = mex("rho_delta_page", *args, o, nlhs=4)
        np.testing.assert_array_equal(rho.ravel(), ref["rho"])
        np.testing.assert_array_equal(delta.ravel(), ref["delta"])
        np.testing.assert_array_equal(nn.ravel(), ref["nneigh"])
        np.testing.assert_array_equal(knn.T, ref["knn"])
