"""Host-side mirror of irc2's hot-path functions, calling the gfx950 library.

Names, argument meaning and orientation follow the reference (matlab/irc2.m) so a
caller can switch call sites one for one.  Arrays use MATLAB orientation: features
are D x n (one column per spike), indices returned are 1-based.

  search_knn_drift_   matlab/irc2.m:3731-3789
  search_delta_drift_ matlab/irc2.m:3794-3865
  cuda_knn_           matlab/irc2.m:4202-4238
  sort_page_          matlab/irc2.m:3456-3503 (rho_page_, delta_page_, rho_paged_site_,
                      delta_paged_site_ folded into one batched call)
  plan_sort_page_     matlab/irc2.m:3914-3946
  calc_drift_         matlab/irc2.m:6373-6447 (drift linkage that feeds the path)
  sort_               matlab/irc2.m:3251-3274 (calc_drift_ -> sort_page_ -> rho/max(rho) -> ordrho)
  sort_cache_         matlab/irc2.m:359-383 (the save/load of its result: irc_files.save_sort_)
  cuda_knn_full_, cuda_knn_ib_  the cuda_knn_full.cu / cuda_knn_ib.cu forms
  prepare_page_site_  matlab/irc2.m:3869-3910 (site_layout builds the same column grouping
                      for every site at once)

There is no CPU fallback: a library error raises IrcError (the reference's callers
catch errors and fall back to their own CPU code, matlab/irc2.m:3643-3652).
"""
import numpy as np

from . import libirc

_CTX = {}


def get_ctx(device=0):
    """One context per device per process, like the reference's `persistent CK`."""
    ctx = _CTX.get(device)
    if ctx is None:
        ctx = libirc.Context(device)
        _CTX[device] = ctx
    return ctx


def _opts(P):
    P = P or {}
    prec = P.get("precision", libirc.PREC_AUTO)
    if isinstance(prec, str):
        prec = {"exact": libirc.PREC_EXACT, "fp32": libirc.PREC_FP32, "bf16": libirc.PREC_BF16_SCREEN,
                "fp8": libirc.PREC_FP8_SCREEN, "auto": libirc.PREC_AUTO}[prec]
    return libirc.make_options(precision=prec, k_over=int(P.get("k_over", 0)),
                               force_fail=bool(P.get("force_fail", False)), flags=int(P.get("flags", 0)),
                               large_ncand=int(P.get("large_ncand", 0)))


def search_knn_drift_(vl_in, mrFet_out, viDrift_out, mlDrift1, P=None, viSpk_out=None):
    """[vrKnn_in, miKnn_in] = search_knn_drift_(vl_in, mrFet_out, viDrift_out, mlDrift1, P).
    vl_in: logical over columns (a contiguous run); mrFet_out: D x nF single;
    viDrift_out: 1-based chunk per column; mlDrift1: C x C logical.
    Returns vrKnn_in (n_in,) single = distance to the k-th neighbour and miKnn_in (knn x n_in)
    int32 1-based column indices.  viSpk_out (optional) gives spike ids for tie-breaking."""
    P = P or {}
    knn = int(P.get("knn", 30))
    vl_in = np.asarray(vl_in, dtype=bool).ravel()
    idx = np.flatnonzero(vl_in)
    if idx.size == 0:
        return np.zeros(0, np.float32), np.zeros((knn, 0), np.int32)
    if idx[-1] - idx[0] + 1 != idx.size:
        raise ValueError("vl_in must mark a contiguous run of columns")
    F = np.ascontiguousarray(np.asarray(mrFet_out, dtype=np.float32).T)
    dk, I = get_ctx(P.get("device", 0)).knn_drift(F, viDrift_out, np.asarray(mlDrift1, dtype=np.uint8),
                                                  idx[0], idx.size, knn, id_of_col=viSpk_out, opts=_opts(P))
    return dk, np.ascontiguousarray(I.T)


def search_delta_drift_(vl_in, mrFet_out, vrRho_out, viDrift_out, mlDrift1, P=None, viSpk_out=None):
    """[vrDelta_in, viNneigh_in] = search_delta_drift_(vl_in, mrFet_out, vrRho_out, viDrift_out, mlDrift1, P).
    vrDelta_in is the distance (not yet multiplied by rho; +Inf when no denser column exists),
    viNneigh_in the 1-based column (own column when none)."""
    P = P or {}
    vl_in = np.asarray(vl_in, dtype=bool).ravel()
    idx = np.flatnonzero(vl_in)
    if idx.size == 0:
        return np.zeros(0, np.float32), np.zeros(0, np.int32)
    if idx[-1] - idx[0] + 1 != idx.size:
        raise ValueError("vl_in must mark a contiguous run of columns")
    F = np.ascontiguousarray(np.asarray(mrFet_out, dtype=np.float32).T)
    return get_ctx(P.get("device", 0)).delta_drift(F, viDrift_out, np.asarray(mlDrift1, dtype=np.uint8), idx[0],
                                                   idx.size, vrRho_out, id_of_col=viSpk_out, opts=_opts(P))


def cuda_knn_(mrFet2, mrFet1, P=None):
    """[vrKnn, fGpu, miKnn] = cuda_knn_(mrFet2, mrFet1, P): kNN of the columns of mrFet1 among
    the columns of mrFet2 (knn = min(knn, n2) rows, as the reference)."""
    P = P or {}
    knn = int(P.get("knn", 30))
    B = np.ascontiguousarray(np.asarray(mrFet2, dtype=np.float32).T)
    A = np.ascontiguousarray(np.asarray(mrFet1, dtype=np.float32).T)
    knn = min(knn, B.shape[0])
    K, I = get_ctx(P.get("device", 0)).knn_ab(B, A, knn)
    return K, 1, np.ascontiguousarray(I.T)


def cuda_knn_full_(mrFet2, mrFet1, P=None):
    """[mrKnn, miKnn] of the cuda_knn_full kernel (cuda_knn_full.cu:27-100): every one of the knn
    nearest columns of mrFet2 for each column of mrFet1, distances (knn x n1) and 1-based indices."""
    P = P or {}
    B = np.ascontiguousarray(np.asarray(mrFet2, dtype=np.float32).T)
    A = np.ascontiguousarray(np.asarray(mrFet1, dtype=np.float32).T)
    knn = min(int(P.get("knn", 30)), B.shape[0])
    K, I = get_ctx(P.get("device", 0)).knn_ab_full(B, A, knn)
    return np.ascontiguousarray(K.T), np.ascontiguousarray(I.T)


def cuda_knn_ib_(mrFet2, viB, mrFet1, P=None):
    """[vrKnn, miKnn] of the cuda_knn_ib kernel (cuda_knn_ib.cu:27-97): the candidates are the
    columns viB (1-based) of mrFet2; miKnn holds columns of mrFet2."""
    P = P or {}
    B = np.ascontiguousarray(np.asarray(mrFet2, dtype=np.float32).T)
    A = np.ascontiguousarray(np.asarray(mrFet1, dtype=np.float32).T)
    ib = np.asarray(viB).ravel()
    knn = min(int(P.get("knn", 30)), ib.size)
    K, I = get_ctx(P.get("device", 0)).knn_ab_ib(B, ib, A, knn)
    return K, np.ascontiguousarray(I.T)


def calc_drift_(vrAmp_spk, mrPos_spk, P=None):
    """S_drift = calc_drift_(S0, P) (matlab/irc2.m:6373-6447) on the spikes' amplitudes and
    positions: returns dict(nTime_drift, mlDrift (C x C logical), viLim_drift (1-based),
    mrCount_drift).  P keys: nTime_batch, nTime_drift, nQuantile_drift (10), nTime_max_drift (256),
    viSite2Chan (its length is the number of position bins)."""
    P = P or {}
    pos = np.asarray(mrPos_spk, dtype=np.float32)
    if pos.ndim == 2:
        pos = pos[:, 1]
    n_pos = len(P["viSite2Chan"]) if "viSite2Chan" in P else int(P.get("nSites", 1))
    link, lim, counts = get_ctx(P.get("device", 0)).calc_drift(
        np.asarray(vrAmp_spk, dtype=np.float32).ravel(), pos, n_pos, int(P.get("nQuantile_drift", 10)),
        int(P["nTime_drift"]), int(P["nTime_batch"]), int(P.get("nTime_max_drift", 256)), want_counts=True)
    return dict(nTime_drift=link.shape[0], mlDrift=link, viLim_drift=lim, mrCount_drift=counts)


def site_layout(ds, F1=None):
    """Group spike features by site (the _fet_/_fet2_ layout, matlab/irc2.m:4583-4618):
    for each site, the spikes whose primary site it is (primary features, ascending id), then
    the spikes whose secondary site it is (secondary features, ascending id).
    Returns numpy/torch arrays for irc_rho_delta_batch."""
    import torch
    site1 = ds["site1"]
    dev = site1.device
    N = ds["n_spikes"]
    S = ds["n_sites"]
    ids = torch.arange(1, N + 1, device=dev, dtype=torch.int32)
    key1 = (site1.long() - 1) * 2
    F1 = ds["F1"] if F1 is None else F1
    if ds.get("site2") is not None:
        s2 = ds["site2"]
        m = s2 > 0
        key2 = (s2[m].long() - 1) * 2 + 1
        keys = torch.cat([key1, key2])
        allids = torch.cat([ids, ids[m]])
        src = torch.cat([torch.zeros(N, dtype=torch.bool, device=dev), torch.ones(int(m.sum()), dtype=torch.bool, device=dev)])
        rows = torch.cat([torch.arange(N, device=dev), torch.nonzero(m).flatten()])
    else:
        keys = key1
        allids = ids
        src = None
        rows = torch.arange(N, device=dev)
    order = torch.sort(keys, stable=True).indices
    col_spike = allids[order].contiguous()
    rows_o = rows[order]
    if src is None:
        F = F1[rows_o].contiguous()
    else:
        so = src[order]
        F = torch.empty((order.numel(), F1.shape[1]), dtype=torch.float32, device=dev)
        F[~so] = F1[rows_o[~so]]
        F[so] = ds["F2"][rows_o[so]]
    cnt = torch.bincount(keys[order], minlength=2 * S).reshape(S, 2).cpu().numpy().astype(np.int64)
    site_ptr = np.concatenate([[0], np.cumsum(cnt.sum(1))]).astype(np.int64)
    return dict(F=F, col_spike=col_spike, site_ptr=site_ptr, site_n_primary=cnt[:, 0].copy(),
                lim_drift=np.asarray(ds["lim"], dtype=np.int64), link=np.asarray(ds["link"], dtype=np.uint8),
                n_spikes=N, k=ds["k"])


def sort_page_(layout, P=None, want_knn=True):
    """[vrRho, vrDelta, viNneigh] (+ kNN, k x N) for every site and drift chunk in one call.
    rho is unnormalised (sort_ divides by max afterwards, matlab/irc2.m:3268)."""
    P = P or {}
    ctx = get_ctx(P.get("device", 0))
    host = {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v) for kk, v in layout.items()}
    r = ctx.rho_delta_batch(host, int(P.get("knn", layout.get("k", 30))), opts=_opts(P), want_knn=want_knn)
    out = [r["rho"], r["delta"], r["nneigh"]]
    if want_knn:
        out.append(np.ascontiguousarray(r["knn"].T))
    return tuple(out)


def ordrho_(vrRho):
    """[~, ordrho] = sort(vrRho, 'descend') (matlab/irc2.m:3269): 1-based positions (double), equal
    values in ascending position (MATLAB's sort is stable), NaN first (a descending sort places NaN
    before every number)."""
    r = np.asarray(vrRho).ravel()
    idx = np.arange(r.size)
    nan = np.isnan(r)
    rest = idx[~nan]
    o = np.concatenate([idx[nan], rest[np.argsort(-r[~nan], kind="stable")]])
    return (o + 1).astype(np.float64)


def sort_(S0, P=None):
    """S_clu = sort_(S0, P) (matlab/irc2.m:3251-3274): drift linkage, then every site and drift chunk
    in one batch call, rho divided by its maximum (3268) and ordrho (3269).
    S0 holds the drift inputs vrAmp_spk, mrPos_spk and the features, either
      * in memory: F1 (N x D rows), viSite_spk, and optionally F2, viSite2_spk, or
      * as detect files: vcFile_prm (the prefix the _fet_ files were written under),
        ccviSpk_site_load, ccviSpk_site2_load (or None), dimm_fet, and mlPc (get_mlPc_).
    P: calc_drift_'s keys (nTime_batch, nTime_drift, nQuantile_drift, nTime_max_drift, viSite2Chan or
    nSites), knn, precision; P['vcFile_prm'] set: the per-site _knn_<site>.irc files are written as
    rho_paged_site_ does (matlab/irc2.m:3670-3679).  Returns S_clu (rho, delta, ordrho, nneigh, P,
    miKnn (k x N int64 when P['fKeep_knn'], else empty, as the reference), S_drift, nFeatures)."""
    from . import irc_files
    P = dict(P or {})
    knn = int(P.get("knn", 30))
    S_drift = calc_drift_(S0["vrAmp_spk"], S0["mrPos_spk"], P)
    lim, link = S_drift["viLim_drift"], S_drift["mlDrift"]
    N = int(np.asarray(S0["vrAmp_spk"]).size)
    if "ccviSpk_site_load" in S0:
        layout = irc_files.layout_from_fet_files(S0["vcFile_prm"], S0["ccviSpk_site_load"],
                                                 S0.get("ccviSpk_site2_load"), S0["dimm_fet"], S0["mlPc"], lim,
                                                 link, N, knn)
    else:
        import torch
        site1 = torch.as_tensor(np.asarray(S0["viSite_spk"], dtype=np.int32))
        has2 = S0.get("viSite2_spk") is not None
        ds = dict(n_spikes=N, n_sites=int(S0.get("nSites", int(site1.max()))), k=knn,
                  F1=torch.as_tensor(np.asarray(S0["F1"], dtype=np.float32)), site1=site1,
                  F2=torch.as_tensor(np.asarray(S0["F2"], dtype=np.float32)) if has2 else None,
                  site2=torch.as_tensor(np.asarray(S0["viSite2_spk"], dtype=np.int32)) if has2 else None,
                  lim=lim, link=np.asarray(link, dtype=np.uint8))
        layout = site_layout(ds)
    vrRho, vrDelta, viNneigh, miKnn = sort_page_(layout, dict(P, knn=knn))
    if P.get("vcFile_prm"):
        irc_files.save_knn_by_site(P["vcFile_prm"], {kk: (v.cpu().numpy() if hasattr(v, "cpu") else v)
                                                      for kk, v in layout.items()}, miKnn)
    vrRho = (vrRho / vrRho.max()).astype(np.float32)
    return dict(rho=vrRho, delta=vrDelta, ordrho=ordrho_(vrRho), nneigh=viNneigh, P=P,
                miKnn=miKnn if P.get("fKeep_knn") else np.zeros((0, 0), np.int64), S_drift=S_drift,
                nFeatures=int(layout["F"].shape[1]))


def plan_sort_page_(nTime_drift, viLim_drift, nTime_max_drift=256):
    """Pages of drift chunks (matlab/irc2.m:3914-3946): returns (miSpk_lim_out, miSpk_lim_in,
    miDrift_lim_out, miDrift_lim_in), 1-based inclusive limits, one row per page."""
    n = int(nTime_drift)
    width = int(nTime_max_drift)
    lim = np.asarray(viLim_drift, dtype=np.int64)
    quarter, half = width // 4, width // 2
    n_pages = n // half
    if n_pages == 0:
        so = np.array([[lim[0], lim[-1] - 1]], np.int64)
        do = np.array([[1, len(lim) - 1]], np.int32)
        return so, so.copy(), do, do.copy()
    so = np.zeros((n_pages, 2), np.int64)
    si = np.zeros((n_pages, 2), np.int64)
    do = np.zeros((n_pages, 2), np.int32)
    di = np.zeros((n_pages, 2), np.int32)
    ib_prev = 0
    for p in range(1, n_pages + 1):
        iB = min((p - 1) * half + width, n)
        iA = max(1, iB - width + 1)
        ia = 1 if p == 1 else min(iA + quarter, n)
        ia = max(ib_prev + 1, ia)
        ib = n if p == n_pages else max(iB - quarter, 1)
        ib_prev = ib
        do[p - 1] = (iA, iB)
        di[p - 1] = (ia, ib)
        si[p - 1] = (lim[ia - 1], lim[ib] - 1)
        so[p - 1] = (lim[iA - 1], lim[iB] - 1)
    return so, si, do, di


# ---- density-peak assignment (the consumer of rho, delta, nneigh and the kNN) ---------------

def _merge_overlap_groups(groups, n_repeat=10):
    """cell2map_ (matlab/irc2.m:6754-6803) on overlap lists (1-based member sets): up to n_repeat
    rounds in which every list with more than one member is replaced, for all its members, by the
    union of its members' lists (later lists overwrite earlier ones); each entry then maps to its
    smallest member (itself when its list is empty) and the map is compacted.  Returns
    (cluster of each entry, 1-based, and the sorted representatives)."""
    cur = [frozenset(g) for g in groups]
    n = len(cur)
    for _ in range(n_repeat):
        nxt = [frozenset()] * n
        for i in range(n):
            if len(cur[i]) < 2:
                continue
            grown = set(cur[i])
            for m in cur[i]:
                grown |= cur[m - 1]
            g = frozenset(grown)
            for m in g:
                nxt[m - 1] = g
        if all(len(a) == len(b) for a, b in zip(nxt, cur)):
            break
        cur = nxt
    rep = np.array([min(g) if g else i + 1 for i, g in enumerate(cur)], dtype=np.int64)
    reps = np.unique(rep)
    ids = np.searchsorted(reps, rep) + 1
    return ids.astype(np.int32), reps


def assignCluster_knn_(S_clu, P=None, miKnn=None):
    """[viClu_spk, viSpk_peak, nClu_pre] = assignCluster_knn_(S_clu, viSite_spk, P)
    (matlab/irc2.m:6640-6694).  S_clu: dict with rho, delta, nneigh (1-based); miKnn: k x N kNN
    (1-based spike ids; only the peaks' columns are read).  Peaks (delta >= delta_cut, by rho
    descending) whose kNN lists share a spike are merged on the host; the chain loop runs on the
    GPU (irc_assign_chain)."""
    P = P or {}
    rho = np.asarray(S_clu["rho"])
    delta = np.asarray(S_clu["delta"])
    nneigh = np.asarray(S_clu["nneigh"], dtype=np.int64)
    N = rho.size
    ordrho = np.argsort(-rho.astype(np.float64), kind="stable") + 1
    peaks = np.flatnonzero(delta >= float(P.get("delta_cut", 1.0))) + 1
    n_pre = len(peaks)
    peaks = peaks[np.argsort(-rho[peaks - 1].astype(np.float64), kind="stable")]
    if len(peaks):
        K = np.asarray(miKnn)[:, peaks - 1]
        # spike -> peaks whose kNN list holds it; a peak's overlap list is the union over its spikes
        owners = {}
        for j in range(K.shape[1]):
            for s in set(K[:, j].tolist()):
                owners.setdefault(s, []).append(j + 1)
        groups = [set().union(*(owners[s] for s in set(K[:, j].tolist()))) for j in range(K.shape[1])]
        clu_peak, reps = _merge_overlap_groups(groups)
        peaks_kept = peaks[reps - 1]
    else:
        peaks_kept = peaks
    if len(peaks_kept) <= 1:
        return np.ones(N, np.int32), ordrho[:1], n_pre
    lab = np.zeros(N, np.int32)
    lab[peaks - 1] = clu_peak
    lab = get_ctx(P.get("device", 0)).assign_chain(nneigh, lab)
    return lab, peaks_kept, n_pre


def postCluster_(S_clu, P=None, miKnn=None):
    """viClu and icl after assignCluster_knn_ and dpclus_remove_count_ (matlab/irc2.m:6520-6534,
    6735-6749): clusters with fewer than P['min_count'] spikes are dropped and the ids compacted."""
    P = P or {}
    clu, icl, _ = assignCluster_knn_(S_clu, P, miKnn)
    min_count = int(P.get("min_count", 0))
    n = len(icl)
    counts = np.bincount(clu[clu > 0], minlength=n + 1)[1:n + 1]
    keep = counts >= min_count
    if keep.all():
        return clu, icl
    new_id = np.zeros(n + 1, np.int32)
    new_id[1:][keep] = np.arange(1, int(keep.sum()) + 1)
    return np.where(clu > 0, new_id[np.clip(clu, 0, n)], 0).astype(np.int32), icl[keep]


def isolation_score_(viClu, miKnn, nClu=None, P=None):
    """[vrScore_clu, viClu_nn_clu] = isolation_score_(S_auto, miKnn_spk) (matlab/irc2.m:6699-6731)
    from the GPU's neighbour-cluster counts: n11 = own-cluster kNN entries of cluster c, c2 = the
    most frequent other cluster (ties to the smaller id, as MATLAB's mode), score = n11/(n11+n12);
    with no other cluster the score is n11/n11 and c2 NaN (MATLAB's mode of an empty set)."""
    P = P or {}
    clu = np.asarray(viClu, dtype=np.int32)
    n_clu = int(clu.max()) if nClu is None else int(nClu)
    cnt = get_ctx(P.get("device", 0)).isolation_counts(miKnn, clu, n_clu).astype(np.int64)
    c = np.arange(n_clu)
    n11 = cnt[c, c + 1]
    other = cnt[:, 1:].copy()
    other[c, c] = 0
    c2 = np.argmax(other, axis=1)
    n12 = other[c, c2]
    with np.errstate(invalid="ignore", divide="ignore"):
        score = np.where(n12 > 0, n11 / (n11 + n12), n11 / n11.astype(np.float64))
    nn = np.where(n12 > 0, c2 + 1, np.nan)
    return score.astype(np.float64), nn.astype(np.float64)


def knn_overlap_merge_(viClu, miKnn, thresh, nClu=None, P=None):
    """knn_overlap_merge_ (matlab/irc2.m:2156-2174): clusters whose isolation score is below thresh
    are grouped with their most frequent neighbour cluster (cell2map_); returns relabelled viClu."""
    score, nn = isolation_score_(viClu, miKnn, nClu, P)
    groups = [{c, int(nn[c - 1])} if score[c - 1] < thresh else {c} for c in range(1, len(score) + 1)]
    m, _ = _merge_overlap_groups(groups)
    clu = np.asarray(viClu).copy()
    upd = clu > 0
    clu[upd] = m[clu[upd] - 1]
    return clu
