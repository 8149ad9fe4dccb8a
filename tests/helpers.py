"""Small independent restatements used to pin the oracle (numpy brute force on inputs whose
squared distances are exact in any summation order, i.e. small integer features)."""
import numpy as np


def brute_batch(ds):
    """Brute-force canonical outputs for integer-valued features (exact arithmetic)."""
    N, k = ds["n_spikes"], ds["k"]
    F1 = np.asarray(ds["F1"], np.float64)
    F2 = None if ds.get("F2") is None else np.asarray(ds["F2"], np.float64)
    s1 = np.asarray(ds["site1"])
    s2 = None if ds.get("site2") is None else np.asarray(ds["site2"])
    lim = np.asarray(ds["lim"])
    link = np.asarray(ds["link"]).astype(bool)
    chunk = np.searchsorted(lim, np.arange(1, N + 1), side="right") - 1
    ids = np.arange(N)
    cand = []
    for a in range(N):
        rows = link[:, chunk[a]]
        m1 = (s1 == s1[a]) & rows[chunk]
        b1 = ids[m1]
        d1 = ((F1[b1] - F1[a]) ** 2).sum(1)
        if s2 is not None:
            m2 = (s2 == s1[a]) & rows[chunk]
            b2 = ids[m2]
            d2 = ((F2[b2] - F1[a]) ** 2).sum(1)
            b1 = np.concatenate([b1, b2])
            d1 = np.concatenate([d1, d2])
        cand.append((b1, d1))
    knn = np.zeros((N, k), np.int64)
    rho = np.zeros(N, np.float32)
    for a, (b, d) in enumerate(cand):
        o = np.lexsort((b, d))
        kk = min(k, len(o))
        sel = list(o[:kk]) + [o[kk - 1]] * (k - kk)
        knn[a] = b[sel] + 1
        rho[a] = np.float32(1.0 / np.sqrt(d[o[kk - 1]]))
    delta = np.zeros(N, np.float32)
    nn = np.zeros(N, np.int64)
    sent = np.float32(np.sqrt(3.402e38))
    for a, (b, d) in enumerate(cand):
        m = rho[b] > rho[a]
        if not m.any():
            delta[a], nn[a] = sent, a + 1
            continue
        bb, dd = b[m], d[m]
        o = np.lexsort((bb, dd))[0]
        v = np.float32(np.sqrt(dd[o]) * np.float64(rho[a]))
        if not np.isfinite(v):
            delta[a], nn[a] = sent, a + 1
        else:
            delta[a], nn[a] = v, bb[o] + 1
    return dict(rho=rho, delta=delta, nneigh=nn, knn=knn)


def int_dataset(N, S, C, D, k, seed, lo=-3, hi=4, f2=False, link=None, lim=None):
    """Integer-valued features (many exact ties) with random sites and chunk links."""
    rng = np.random.default_rng(seed)
    F1 = rng.integers(lo, hi, size=(N, D)).astype(np.float32)
    s1 = rng.integers(1, S + 1, size=N).astype(np.int32)
    if lim is None:
        c = np.arange(1, C + 1)
        lim = np.concatenate([[0], np.ceil(c / C * N).astype(np.int64)]) + 1
    if link is None:
        link = (rng.random((C, C)) < 0.5).astype(np.uint8)
        np.fill_diagonal(link, 1)
    ds = dict(n_spikes=N, n_sites=S, D=D, n_drift=C, k=k, F1=F1, F2=None, site1=s1, site2=None,
              lim=np.asarray(lim, np.int64), link=np.asarray(link, np.uint8))
    if f2:
        s2 = np.where(rng.random(N) < 0.5, (s1 + rng.integers(1, S, size=N) - 1) % S + 1, 0).astype(np.int32)
        ds["site2"] = s2
        ds["F2"] = rng.integers(lo, hi, size=(N, D)).astype(np.float32)
    return ds


def layout_to_ds(z):
    """Spike-level arrays (F1, site1, F2, site2) back from the site-grouped batch layout."""
    N = int(z["n_spikes"])
    F, spk = np.asarray(z["F"]), np.asarray(z["col_spike"]).astype(np.int64)
    sp, npri = np.asarray(z["site_ptr"]), np.asarray(z["site_n_primary"])
    S = len(npri)
    D = F.shape[1]
    F1 = np.zeros((N, D), np.float32)
    F2 = np.zeros((N, D), np.float32)
    s1 = np.zeros(N, np.int32)
    s2 = np.zeros(N, np.int32)
    for s in range(S):
        a, b, p = sp[s], sp[s + 1], sp[s] + npri[s]
        F1[spk[a:p] - 1] = F[a:p]
        s1[spk[a:p] - 1] = s + 1
        F2[spk[p:b] - 1] = F[p:b]
        s2[spk[p:b] - 1] = s + 1
    has2 = bool((s2 > 0).any())
    lim = np.asarray(z["lim"])
    return dict(n_spikes=N, n_sites=S, D=D, n_drift=len(lim) - 1, k=int(z["k"]), F1=F1,
                F2=F2 if has2 else None, site1=s1, site2=s2 if has2 else None, lim=lim,
                link=np.asarray(z["link"]))


def subset_sites(ds, sites):
    """Spike-level sub-recording holding only the spikes whose primary site is in `sites`
    (primary-only data).  Ids are renumbered in time order and the drift chunks keep their
    membership, so kNN, rho, delta and nneigh of those spikes are unchanged (ties follow id
    order, which the renumbering preserves).  Returns (sub, orig_ids) with orig_ids 1-based."""
    s1 = np.asarray(ds["site1"].cpu() if hasattr(ds["site1"], "cpu") else ds["site1"])
    keep = np.isin(s1, np.asarray(sites))
    orig = np.flatnonzero(keep) + 1
    lim = np.asarray(ds["lim"])
    chunk = np.searchsorted(lim, orig, side="right") - 1
    C = len(lim) - 1
    sub_lim = np.searchsorted(chunk, np.arange(C + 1), side="left") + 1
    F1 = ds["F1"].cpu().numpy() if hasattr(ds["F1"], "cpu") else np.asarray(ds["F1"])
    sub = dict(n_spikes=len(orig), n_sites=ds["n_sites"], D=ds["D"], n_drift=C, k=ds["k"],
               F1=np.ascontiguousarray(F1[orig - 1]), F2=None, site1=s1[orig - 1].astype(np.int32),
               site2=None, lim=sub_lim.astype(np.int64), link=np.asarray(ds["link"]))
    return sub, orig


def write_detect_files(ds, prefix, nPc, nS, n_loads=2):
    """Write a synthetic dataset as irc2 detect files (_fet_/_fet2_ per load, D = nPc x nS with
    every feature selected); returns (ccviSpk_site_load, ccviSpk_site2_load, dimm_fet, mlPc)."""
    from ironclust_amd import irc_files as IF
    N = ds["n_spikes"]
    F1 = ds["F1"].cpu().numpy().reshape(N, nPc, nS, order="F").transpose(1, 2, 0)
    s1 = ds["site1"].cpu().numpy()
    has2 = ds.get("F2") is not None
    F2 = ds["F2"].cpu().numpy().reshape(N, nPc, nS, order="F").transpose(1, 2, 0) if has2 else None
    s2 = ds["site2"].cpu().numpy() if has2 else None
    cut = np.linspace(0, N, n_loads + 1).astype(np.int64)
    c1, c2 = [], []
    for i in range(n_loads):
        a, b = cut[i], cut[i + 1]
        x, y, _ = IF.detect_paged_save_(prefix, i + 1, F1[:, :, a:b], s1[a:b], ds["n_sites"],
                                        F2[:, :, a:b] if has2 else None, s2[a:b] if has2 else None)
        c1.append(x)
        c2.append(y)
    vn = np.diff(cut)
    ml, _ = IF.get_mlPc_((nPc, nS, N), nC_max=nPc * nS)
    return IF.merge_site_lists_(c1, vn), (IF.merge_site_lists_(c2, vn) if has2 else None), (nPc, nS, N), ml
