"""On-disk formats (_fet_/_fet2_, _knn_) of the detect and sort stages: round trips, the PC mask
rule, and the assembled column layout against the in-memory one (CPU)."""
import numpy as np
import pytest
import torch

from ironclust_amd import irc2, synth
from ironclust_amd import irc_files as IF


def test_vi2cell_matches_brute_force():
    rng = np.random.default_rng(3)
    v = rng.integers(0, 6, size=200)                  # 0 = no site (dropped)
    got = IF.vi2cell_(v, 6)
    for s in range(1, 7):
        np.testing.assert_array_equal(got[s - 1], np.flatnonzero(v == s) + 1)


def test_get_mlPc_rules():
    ml, n = IF.get_mlPc_((3, 14), nC_max=42)
    assert n == 42 and ml.all()
    # 42 > 40: the loop trims from the last site backwards until the total fits:
    # [3]*12 + [2, 1] (39 features)
    ml, n = IF.get_mlPc_((3, 14), nC_max=40)
    assert n == 39
    assert list(ml.sum(0)) == [3] * 12 + [2, 1]
    assert ml[:2, 12].all() and not ml[2, 12] and ml[0, 13] and not ml[1:, 13].any()
    ml, n = IF.get_mlPc_((3, 5), nC_max=100, fMode_mlPc=1)       # nPc:-1:1 over the first 3 sites
    assert list(ml.sum(0)) == [3, 2, 1, 0, 0] and n == 6
    ml, n = IF.get_mlPc_((3, 4), nC_max=100, nPcPerChan=2)
    assert list(ml.sum(0)) == [2, 2, 2, 2]
    ml, n = IF.get_mlPc_((3, 4), nC_max=100, nPcPerChan=[3, 1, 0, 2])
    assert list(ml.sum(0)) == [3, 1, 0, 2] and n == 6


def _two_loads(tmp_path, with2=True):
    rng = np.random.default_rng(5)
    nPc, nS, nSites = 3, 4, 5
    loads = []
    for n in (50, 70):
        tr = rng.standard_normal((nPc, nS, n)).astype(np.float32)
        s1 = rng.integers(1, nSites + 1, size=n)
        tr2 = rng.standard_normal((nPc, nS, n)).astype(np.float32) if with2 else None
        s2 = np.where(rng.random(n) < 0.5, rng.integers(1, nSites + 1, size=n), 0) if with2 else None
        loads.append((tr, s1, tr2, s2))
    prefix = str(tmp_path / "rec")
    c1, c2 = [], []
    for i, (tr, s1, tr2, s2) in enumerate(loads, start=1):
        a, b, dimm = IF.detect_paged_save_(prefix, i, tr, s1, nSites, tr2, s2)
        c1.append(a)
        c2.append(b)
    vn = [l[0].shape[2] for l in loads]
    return prefix, loads, IF.merge_site_lists_(c1, vn), IF.merge_site_lists_(c2, vn), (nPc, nS, sum(vn)), nSites


def test_fet_round_trip_every_site(tmp_path):
    prefix, loads, cc1, cc2, dimm, nSites = _two_loads(tmp_path)
    tr_all = np.concatenate([l[0] for l in loads], axis=2)
    s1_all = np.concatenate([l[1] for l in loads])
    tr2_all = np.concatenate([l[2] for l in loads], axis=2)
    s2_all = np.concatenate([l[3] for l in loads])
    ml, nF = IF.get_mlPc_(dimm, nC_max=10)
    m = ml.ravel(order="F")
    for s in range(1, nSites + 1):
        for iFet, tr, sa in ((1, tr_all, s1_all), (2, tr2_all, s2_all)):
            cc = cc1 if iFet == 1 else cc2
            ids = np.flatnonzero(sa == s) + 1
            t, vi = IF.load_fet_site_(prefix, iFet, s, cc, dimm)
            np.testing.assert_array_equal(vi, ids)
            np.testing.assert_array_equal(t, tr[:, :, ids - 1])
            f, vi = IF.load_fet_site_(prefix, iFet, s, cc, dimm, ml)
            assert f.shape == (nF, len(ids))
            np.testing.assert_array_equal(f, tr[:, :, ids - 1].reshape(-1, len(ids), order="F")[m])


def test_layout_from_files_equals_in_memory_layout(tmp_path):
    ds = synth.make_dataset(3000, 6, 12, 5, 3, k=8, drift_step=0.3, n_links=2, band=2, seed=41, f2=True)
    lay = irc2.site_layout(ds)
    nPc, nS = 3, 4                                    # D = 12 = 3 PCs x 4 sites, all selected
    N = ds["n_spikes"]
    F1 = ds["F1"].numpy().reshape(N, nPc, nS, order="F").transpose(1, 2, 0)
    F2 = ds["F2"].numpy().reshape(N, nPc, nS, order="F").transpose(1, 2, 0)
    s1 = ds["site1"].numpy()
    s2 = ds["site2"].numpy()
    prefix = str(tmp_path / "rec")
    cut = [0, 1100, N]                                # two loads
    c1, c2 = [], []
    for i in range(2):
        a, b = cut[i], cut[i + 1]
        x, y, _ = IF.detect_paged_save_(prefix, i + 1, F1[:, :, a:b], s1[a:b], ds["n_sites"], F2[:, :, a:b], s2[a:b])
        c1.append(x)
        c2.append(y)
    vn = [cut[1] - cut[0], cut[2] - cut[1]]
    cc1, cc2 = IF.merge_site_lists_(c1, vn), IF.merge_site_lists_(c2, vn)
    ml, _ = IF.get_mlPc_((nPc, nS, N), nC_max=12)
    got = IF.layout_from_fet_files(prefix, cc1, cc2, (nPc, nS, N), ml, ds["lim"], ds["link"], N, ds["k"])
    np.testing.assert_array_equal(got["F"], lay["F"].numpy())
    np.testing.assert_array_equal(got["col_spike"], lay["col_spike"].numpy())
    np.testing.assert_array_equal(got["site_ptr"], lay["site_ptr"])
    np.testing.assert_array_equal(got["site_n_primary"], lay["site_n_primary"])


def _read_knn_as_reference(path, knn):
    """load_bin_(vcFile_knn_site1, 'int64', [knn, n]) (matlab/irc2.m:3395): raw little-endian int64,
    column-major knn x n, read here without the module under test."""
    raw = open(path, "rb").read()
    assert len(raw) % (8 * knn) == 0
    n = len(raw) // (8 * knn)
    return np.frombuffer(raw, dtype="<i8").reshape((n, knn)).T


def test_knn_files_append_per_page(tmp_path):
    rng = np.random.default_rng(7)
    prefix = str(tmp_path / "rec")
    # ids above 2^31 exist in int64 files (the reference type, matlab/irc2.m:3670)
    p1 = rng.integers(1, 1 << 40, size=(5, 30)).astype(np.int64)
    p2 = rng.integers(1, 1000, size=(5, 12)).astype(np.int32)      # an int32 page is widened on write
    path = IF.save_miKnn_site_(prefix, 3, p1)
    IF.save_miKnn_site_(prefix, 3, p2, fAppend=True)
    want = np.concatenate([p1, p2.astype(np.int64)], axis=1)
    assert (tmp_path / "rec_knn_3.irc").stat().st_size == 8 * want.size
    np.testing.assert_array_equal(_read_knn_as_reference(path, 5), want)
    np.testing.assert_array_equal(IF.load_miKnn_site_(prefix, 3, 5), want)


def test_sort_cache_round_trip(tmp_path):
    """save_sort_ writes rho, delta, nneigh, ordrho back to back in struct_save_bin_'s layout
    (matlab/irc2.m:1733-1753) and S_var beside them; the blob is read here from S_var's classes and
    sizes alone (as struct_load_bin_ does, matlab/irc2.m:1775-1805), and by load_sort_."""
    from scipy.io import loadmat
    from oracle import oracle as O
    ds = synth.make_dataset(2500, 3, 6, 4, 3, k=7, drift_step=0.3, n_links=2, band=1, seed=8)
    r = O.sort_batch(ds)
    rho = (r["rho"] / r["rho"].max()).astype(np.float32)
    S_clu = dict(rho=rho, delta=r["delta"], nneigh=r["nneigh"], ordrho=irc2.ordrho_(rho), nFeatures=6)
    f = str(tmp_path / "sort_irc.mat")
    blob, S_var = IF.save_sort_(S_clu, f)
    assert blob.endswith("sort.irc")
    m = loadmat(f, squeeze_me=True, simplify_cells=True)
    sv = m["S_var"]
    assert list(sv["csName_var"]) == ["rho", "delta", "nneigh", "ordrho"]
    assert list(sv["csType_var"]) == ["single", "single", "int64", "double"]
    raw = open(blob, "rb").read()
    pos = 0
    width = {"single": 4, "int64": 8, "double": 8}
    code = {"single": "<f4", "int64": "<i8", "double": "<f8"}
    for name, typ, dimm in zip(sv["csName_var"], sv["csType_var"], sv["cDimm_var"]):
        n = int(np.prod(dimm))
        assert list(np.ravel(dimm)) == [2500, 1]
        got = np.frombuffer(raw[pos:pos + n * width[typ]], dtype=code[typ])
        pos += n * width[typ]
        np.testing.assert_array_equal(got, np.ravel(S_clu[name]), err_msg=name)
    assert pos == len(raw)
    back = IF.load_sort_(f)
    for name in ("rho", "delta", "nneigh", "ordrho"):
        np.testing.assert_array_equal(np.ravel(back[name]), np.ravel(S_clu[name]))
        assert back[name].dtype == np.asarray(S_clu[name]).dtype
    assert int(back["nFeatures"]) == 6


def test_ordrho_matches_matlab_descending_sort():
    # MATLAB's sort(v, 'descend') is stable and puts NaN first; rho / max(rho) is NaN where rho and
    # max(rho) are both Inf (duplicate points, d_k = 0)
    v = np.array([0.5, np.nan, 1.0, 0.5, 0.0, np.nan, 1.0, -0.0, 0.25], np.float32)
    want = [2, 6, 3, 7, 1, 4, 9, 5, 8]
    assert list(irc2.ordrho_(v).astype(int)) == want
    rng = np.random.default_rng(2)
    v = rng.integers(0, 5, size=300).astype(np.float32)
    v[rng.random(300) < 0.1] = np.nan
    key = lambda i: (0 if np.isnan(v[i]) else 1, 0.0 if np.isnan(v[i]) else -float(v[i]), i)
    assert list(irc2.ordrho_(v).astype(int) - 1) == sorted(range(300), key=key)
