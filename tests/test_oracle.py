"""Pins the CPU oracle (oracle/oracle.c) against checks that do not retype it:
exact brute force on integer features (every summation order is exact), exact rational
distances, a k-d tree, closed forms on a lattice, hand-computed link cases, and the
invariants/properties the reference's structure implies (link mask, paging, layout)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import brute_batch, int_dataset
from ironclust_amd import irc2, synth


def test_distance_is_exact_fma_chain():
    rng = np.random.default_rng(0)
    for D in (1, 3, 24, 42, 45):
        for _ in range(50):
            q = (rng.standard_normal(D) * rng.choice([1e-3, 1, 1e3])).astype(np.float32)
            x = (rng.standard_normal(D) * rng.choice([1e-3, 1, 1e3])).astype(np.float32)
            exact = sum((Fraction(float(a)) - Fraction(float(b))) ** 2 for a, b in zip(q, x))
            got = O.dist2(q, x)
            # an fma chain of D terms is within D+1 roundings of the exact sum
            assert abs(Fraction(got) - exact) <= Fraction((D + 1) * 2.0 ** -52) * exact
    # integer inputs are exact
    q = np.array([1, -2, 3], np.float32)
    x = np.array([4, 2, -1], np.float32)
    assert O.dist2(q, x) == 9 + 16 + 16


def test_rho_and_sentinel_formulas():
    lib = O.oracle()
    assert lib.orc_rho_from_d2(4.0) == np.float32(0.5)
    assert lib.orc_rho_from_d2(0.0) == np.float32(np.inf)
    assert lib.orc_sentinel_delta() == np.float32(math.sqrt(3.402e38))
    assert lib.orc_sentinel_delta() == np.float32(1.8444511e19)


@pytest.mark.parametrize("seed,f2", [(1, False), (2, True), (3, True), (4, False)])
def test_batch_matches_integer_brute_force(seed, f2):
    ds = int_dataset(N=300, S=3, C=5, D=3, k=7, seed=seed, f2=f2)
    got = O.sort_batch(ds)
    ref = brute_batch(ds)
    np.testing.assert_array_equal(got["knn"], ref["knn"])
    np.testing.assert_array_equal(got["rho"], ref["rho"])
    np.testing.assert_array_equal(got["nneigh"], ref["nneigh"])
    np.testing.assert_array_equal(got["delta"], ref["delta"])


def test_lattice_closed_form():
    # 1-D lattice, one site and chunk: kNN of point i is i, then i-1, i+1, i-2, i+2 ... by
    # distance with the lower id first on ties; rho = 1/floor(k/2) away from the edges.
    N, k = 40, 5
    F1 = np.arange(N, dtype=np.float32)[:, None]
    ds = dict(n_spikes=N, n_sites=1, D=1, n_drift=1, k=k, F1=F1, F2=None,
              site1=np.ones(N, np.int32), site2=None, lim=np.array([1, N + 1]), link=np.ones((1, 1), np.uint8))
    got = O.sort_batch(ds)
    for i in range(2, N - 2):
        assert list(got["knn"][i]) == [i + 1, i, i + 2, i - 1, i + 3]
        assert got["rho"][i] == np.float32(0.5)
    # edges: point 0 sees 0,1,2,3,4 -> d_k = 4
    assert list(got["knn"][0]) == [1, 2, 3, 4, 5] and got["rho"][0] == np.float32(0.25)
    # interior points all have rho 0.5; denser points are the edge neighbours 1 and N-2 (rho 1/3)?
    # point 1: 0,1,2,3 and then 4 at distance 3 -> rho 1/3; so interior (rho 1/2) is denser than 1.
    assert got["rho"][1] == np.float32(1 / 3)
    # interior points share rho 0.5: none is denser, so they get the sentinel
    assert got["nneigh"][10] == 11 and got["delta"][10] == np.float32(1.8444511e19)
    # point 1 (rho 1/3): nearest denser is point 2 (id 3) at distance 1 -> delta = 1 * (1/3)
    assert got["nneigh"][1] == 3 and got["delta"][1] == np.float32(1.0 * np.float64(np.float32(1 / 3)))


def test_kdtree_cross_check():
    from scipy.spatial import cKDTree
    ds = synth.make_dataset(4000, 1, 8, 1, 5, k=12, seed=7)
    got = O.sort_batch(ds)
    X = ds["F1"].numpy().astype(np.float64)
    d, idx = cKDTree(X).query(X, k=12)
    assert np.mean(np.all(np.sort(got["knn"] - 1, 1) == np.sort(idx, 1), 1)) > 0.999
    np.testing.assert_allclose(got["rho"], 1.0 / d[:, -1], rtol=1e-6)


def test_asymmetric_link_hand_case():
    # two chunks; chunk 0 spikes are candidates for chunk 1 queries but not vice versa.
    F1 = np.array([[0.0], [1.0], [10.0], [11.5]], np.float32)
    link = np.array([[1, 1], [0, 1]], np.uint8)          # link[r, c]: r=0 -> c=1 only
    ds = dict(n_spikes=4, n_sites=1, D=1, n_drift=2, k=2, F1=F1, F2=None, site1=np.ones(4, np.int32),
              site2=None, lim=np.array([1, 3, 5]), link=link)
    got = O.sort_batch(ds)
    # chunk 0 queries (ids 1,2) see only chunk 0: kNN = self, other at distance 1
    assert list(got["knn"][0]) == [1, 2] and list(got["knn"][1]) == [2, 1]
    # chunk 1 queries see both chunks: id 3 at 10 -> self, then id 2 (dist 9)? no: id 4 (1.5) nearer
    assert list(got["knn"][2]) == [3, 4] and list(got["knn"][3]) == [4, 3]
    assert got["rho"][0] == np.float32(1.0) and got["rho"][2] == np.float32(1 / 1.5)
    # delta: id 3 (rho 2/3) -- denser candidates in chunks 0,1: ids 1,2 (rho 1) -> nearest id 2 at 9
    assert got["nneigh"][2] == 2 and got["delta"][2] == np.float32(9.0 * np.float64(np.float32(1 / 1.5)))
    # transposing the link matrix changes the answer (chunk 0 would then see chunk 1)
    got_t = O.sort_batch(dict(ds, link=link.T.copy()))
    assert list(got_t["knn"][2]) == [3, 4] and got_t["nneigh"][2] == 3


def test_link_identity_equals_independent_chunks():
    ds = synth.make_dataset(3000, 3, 6, 4, 3, k=8, drift_step=0.5, seed=11)
    ds["link"] = np.eye(4, dtype=np.uint8)
    got = O.sort_batch(ds)
    lim = ds["lim"]
    for c in range(4):
        a, b = lim[c] - 1, lim[c + 1] - 1
        sub = dict(ds, n_spikes=b - a, n_drift=1, F1=ds["F1"][a:b], site1=ds["site1"][a:b],
                   lim=np.array([1, b - a + 1]), link=np.ones((1, 1), np.uint8))
        r = O.sort_batch(sub)
        np.testing.assert_array_equal(got["knn"][a:b], r["knn"] + a)
        np.testing.assert_array_equal(got["rho"][a:b], r["rho"])


def test_link_all_true_equals_single_chunk():
    ds = synth.make_dataset(2500, 2, 5, 5, 3, k=6, seed=12)
    ds["link"] = np.ones((5, 5), np.uint8)
    got = O.sort_batch(ds)
    one = O.sort_batch(dict(ds, n_drift=1, lim=np.array([1, ds["n_spikes"] + 1]), link=np.ones((1, 1), np.uint8)))
    for key in ("knn", "rho", "delta", "nneigh"):
        np.testing.assert_array_equal(got[key], one[key])


def test_paging_does_not_change_results():
    # 40 chunks, pages of width 16 (quarter 4): links limited to +/-4 chunks, as calc_drift_ does
    ds = synth.make_dataset(8000, 2, 4, 40, 3, k=6, drift_step=0.3, n_links=5, band=4, seed=13)
    full = O.sort_batch(ds)
    so, si, do, di = irc2.plan_sort_page_(40, ds["lim"], nTime_max_drift=16)
    assert si[0, 0] == 1 and si[-1, 1] == ds["n_spikes"]
    assert np.all(si[1:, 0] == si[:-1, 1] + 1)          # "in" ranges tile the recording
    for p in range(len(so)):
        a, b = so[p, 0] - 1, so[p, 1]
        ca, cb = do[p, 0] - 1, do[p, 1]
        sub = dict(ds, n_spikes=b - a, n_drift=cb - ca, F1=ds["F1"][a:b], site1=ds["site1"][a:b],
                   lim=ds["lim"][ca:cb + 1] - a, link=ds["link"][ca:cb, ca:cb].copy())
        r = O.sort_batch(sub)
        ia, ib = si[p, 0] - 1, si[p, 1]
        np.testing.assert_array_equal(r["knn"][ia - a:ib - a] + a, full["knn"][ia:ib])
        np.testing.assert_array_equal(r["rho"][ia - a:ib - a], full["rho"][ia:ib])


def test_site_form_matches_batch_form_and_layout_permutation():
    ds = int_dataset(N=400, S=2, C=4, D=3, k=5, seed=21, f2=True)
    full = O.sort_batch(ds)
    lay = irc2.site_layout({kk: (__import__("torch").as_tensor(v) if kk in ("F1", "F2", "site1", "site2") else v)
                            for kk, v in ds.items()})
    F = lay["F"].numpy()
    spk = lay["col_spike"].numpy()
    for s in range(2):
        a, b = lay["site_ptr"][s], lay["site_ptr"][s + 1]
        npri = lay["site_n_primary"][s]
        Fs, ids = F[a:b], spk[a:b].astype(np.int64)
        drift = np.searchsorted(ds["lim"], ids, side="right").astype(np.int32)
        dk, I = O.knn_site(Fs, drift, ds["link"], 0, npri, ds["k"], id_of_col=ids)
        np.testing.assert_array_equal(ids[I - 1], full["knn"][ids[:npri] - 1])
        # permuting the columns (ties are broken by id, not position) changes nothing
        perm = np.random.default_rng(s).permutation(b - a)
        perm = np.concatenate([np.sort(perm[perm < npri]), perm[perm >= npri]])  # keep queries first
        dk2, I2 = O.knn_site(Fs[perm], drift[perm], ds["link"], 0, npri, ds["k"], id_of_col=ids[perm])
        np.testing.assert_array_equal(ids[perm][I2 - 1], ids[I - 1])
        rho_cols = full["rho"][ids - 1]
        dm, nn = O.delta_site(Fs, drift, ds["link"], 0, npri, rho_cols, id_of_col=ids)
        has = full["nneigh"][ids[:npri] - 1] != ids[:npri]
        np.testing.assert_array_equal(ids[nn - 1][has], full["nneigh"][ids[:npri] - 1][has])
        assert np.all(np.isinf(dm[~has]))


def test_invariants_on_drift_data():
    ds = synth.make_dataset(6000, 3, 10, 12, 4, k=10, drift_step=0.3, n_links=4, band=3, seed=5, f2=True)
    r = O.sort_batch(ds)
    N = ds["n_spikes"]
    ids = np.arange(1, N + 1)
    assert np.all(r["knn"][:, 0] == ids)                   # self first (no duplicates in float data)
    d = r["dk2"]
    assert np.all(r["rho"] == (1.0 / np.sqrt(d)).astype(np.float32))
    has = r["nneigh"] != ids
    assert np.all(r["rho"][r["nneigh"][has] - 1] > r["rho"][has])
    top = np.argmax(r["rho"])
    assert r["nneigh"][top] == top + 1


def test_ab_list_forms_brute_force():
    # cuda_knn_full (every neighbour's distance) and cuda_knn_ib (candidates listed by IB, 1-based,
    # with a repeated column and columns out of order; indices reported as columns of B)
    rng = np.random.default_rng(4)
    B = rng.integers(-3, 4, size=(70, 5)).astype(np.float32)
    A = rng.integers(-3, 4, size=(30, 5)).astype(np.float32)
    d = ((A[:, None, :].astype(np.float64) - B[None].astype(np.float64)) ** 2).sum(2)
    Ka, K, I = O.knn_ab_list(B, A, 7)
    for i in range(30):
        o = np.lexsort((np.arange(70), d[i]))[:7]
        assert list(I[i] - 1) == list(o)
        assert list(Ka[i]) == [np.float32(np.sqrt(d[i, j])) for j in o]
        assert K[i] == Ka[i, -1]
    IB = np.concatenate([rng.permutation(70)[:25] + 1, [5, 5]]).astype(np.int32)
    Ka, K, I = O.knn_ab_list(B, A, 9, IB=IB)
    for i in range(30):
        cols = IB - 1
        o = np.lexsort((cols, d[i, cols]))[:9]
        assert list(I[i] - 1) == list(cols[o])
        assert K[i] == np.float32(np.sqrt(d[i, cols[o[-1]]]))
    # fewer listed candidates than k: padded with the last one
    Ka, K, I = O.knn_ab_list(B, A, 5, IB=np.array([3, 9], np.int32))
    for i in range(30):
        o = np.lexsort((np.array([2, 8]), d[i, [2, 8]]))
        assert list(I[i] - 1) == [2 if o[0] == 0 else 8] + [2 if o[1] == 0 else 8] * 4


def test_ab_and_dc_forms_brute_force():
    rng = np.random.default_rng(3)
    B = rng.integers(-3, 4, size=(60, 4)).astype(np.float32)
    A = rng.integers(-3, 4, size=(25, 4)).astype(np.float32)
    K, I = O.knn_ab(B, A, 6)
    d = ((A[:, None, :].astype(np.float64) - B[None].astype(np.float64)) ** 2).sum(2)
    for i in range(25):
        o = np.lexsort((np.arange(60), d[i]))[:6]
        assert list(I[i] - 1) == list(o)
        assert K[i] == np.float32(np.sqrt(d[i, o[-1]]))
    RB = rng.random(60).astype(np.float32)
    RA = rng.random(25).astype(np.float32)
    Dd, Nn = O.delta_knn_ab(B, RB, A, RA)
    for i in range(25):
        m = RB > RA[i]
        if not m.any():
            assert Nn[i] == 0 and np.isinf(Dd[i])
            continue
        j = np.flatnonzero(m)[np.lexsort((np.flatnonzero(m), d[i, m]))[0]]
        assert Nn[i] == j + 1 and Dd[i] == np.float32(np.sqrt(d[i, j]))
    # dc density with the site-match rule
    n12, n1 = 80, 30
    site = np.stack([rng.integers(1, 4, n12), rng.integers(1, 4, n12)], 1).astype(np.int32)
    f1 = rng.integers(-2, 3, size=(n12, 3)).astype(np.float32)
    f2 = rng.integers(-2, 3, size=(n12, 3)).astype(np.float32)
    rho = O.rho_drift_dc(n1, site, f1, f2, 4.0)
    dl, nn = O.delta_drift_dc(n1, site, f1, f2, np.concatenate([rho, rng.random(n12 - n1).astype(np.float32)]), 4.0)
    rho12 = np.concatenate([rho, rng.random(n12 - n1).astype(np.float32)])
    dl, nn = O.delta_drift_dc(n1, site, f1, f2, rho12, 4.0)
    for i in range(n1):
        m1 = site[:, 0] == site[i, 0]
        m2 = (~m1) & (site[:, 1] == site[i, 0])
        X = np.where(m1[:, None], f1, f2).astype(np.float64)
        dd = ((X - f1[i].astype(np.float64)) ** 2).sum(1)
        m = m1 | m2
        assert rho[i] == np.float32(np.sum(dd[m] <= 4.0) / np.sum(m))
        den = m & (rho12 > rho12[i])
        if den.any():
            j = np.flatnonzero(den)[np.argmin(dd[den])]
            assert nn[i] == j + 1 and dl[i] == np.float32(np.sqrt(dd[j] / 4.0))
        else:
            assert nn[i] == i + 1


def test_cpu_port_agrees_with_oracle():
    ds = synth.make_dataset(20000, 4, 12, 8, 5, k=15, drift_step=0.3, n_links=3, band=2, seed=9)
    r = O.sort_batch(ds)
    c = O.cpu_sort_batch(ds)
    np.testing.assert_allclose(c["rho"], r["rho"], rtol=2e-5)
    assert np.mean(c["nneigh"] == r["nneigh"]) > 0.995
    ok = c["nneigh"] == r["nneigh"]
    np.testing.assert_allclose(c["delta"][ok], r["delta"][ok], rtol=1e-3)
    # the kNN ids: the same lists up to FP32 near-ties (self first wherever it has no duplicate)
    same = (c["knn"] == r["knn"]).all(axis=1)
    assert np.mean(same) > 0.999
    assert (c["knn"][:, 0] == np.arange(1, ds["n_spikes"] + 1)).mean() > 0.999
    # the k-th id lies at d_k = 1/rho from its query (primary-site features only here)
    assert ds["F2"] is None
    F = np.asarray(ds["F1"], np.float64)
    d = F[c["knn"][:, -1] - 1] - F
    np.testing.assert_allclose(np.sqrt((d * d).sum(1)), 1.0 / c["rho"].astype(np.float64), rtol=1e-4)


def test_golden_fixtures_reproduce():
    import glob
    import os
    from tests.helpers import layout_to_ds
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
    assert len(files) >= 4
    for f in files:
        z = np.load(f)
        ds = layout_to_ds(z)
        r = O.sort_batch(ds)
        for key in ("knn", "rho", "delta", "nneigh"):
            np.testing.assert_array_equal(r[key], z[key], err_msg=f"{f} {key}")
        if "int_ties" in f:
            ref = brute_batch(ds)
            for key in ("knn", "rho", "delta", "nneigh"):
                np.testing.assert_array_equal(ref[key], z[key], err_msg=f"{f} {key} brute")


def test_site_subset_preserves_results():
    from tests.helpers import subset_sites
    ds = synth.make_dataset(6000, 4, 6, 10, 3, k=7, drift_step=0.3, n_links=3, band=2, seed=23)
    full = O.sort_batch(ds)
    sub, orig = subset_sites(ds, [2, 4])
    r = O.sort_batch(sub)
    np.testing.assert_array_equal(orig[r["knn"] - 1], full["knn"][orig - 1])
    np.testing.assert_array_equal(r["rho"], full["rho"][orig - 1])
    np.testing.assert_array_equal(orig[r["nneigh"] - 1], full["nneigh"][orig - 1])
    np.testing.assert_array_equal(r["delta"], full["delta"][orig - 1])


def test_lognormal_generator_total_and_chunks():
    # per-site lognormal counts and per-chunk Dirichlet shares must add up to the requested total
    ds = synth.make_dataset(30000, 16, 4, 12, 3, lognormal=True, seed=4)
    assert ds["site1"].numel() == 30000
    assert ds["lim"][0] == 1 and ds["lim"][-1] == 30001
    assert np.all(np.diff(ds["lim"]) >= 0)
