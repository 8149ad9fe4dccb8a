"""Density-peak assignment oracle (test infrastructure only; never the product path).

CPU restatement of the step after the hot path (reference paths relative to the ironclust
repository):

  assignCluster_knn_   matlab/irc2.m:6640-6694  peaks = delta >= delta_cut, ordered by rho
                                                 (descending, stable); peaks whose kNN lists share
                                                 a spike are merged (cell2map_); every other
                                                 spike takes the cluster of its nneigh, visited
                                                 in descending rho order
  cell2map_            matlab/irc2.m:6754-6803  iterated union of the overlap lists (at most 10
                                                 rounds), map to the smallest member, compact ids
  dpclus_remove_count_ matlab/irc2.m:6735-6749  clusters under min_count dropped (ids compacted)
  isolation_score_     matlab/irc2.m:6699-6731  per cluster: own-label share of its spikes' kNN
                                                 against the most frequent other cluster
  knn_overlap_merge_   matlab/irc2.m:2156-2174  clusters under the isolation threshold merged
                                                 with that neighbour through cell2map_

Indices are 1-based as in MATLAB.
"""
import numpy as np


def cell2map(cvi_clu, n_repeat=10):
    """cell2map_ without removals: returns (viMapClu_new, viUniq) -- the compacted cluster of each
    entry and the sorted representatives (1-based)."""
    cvi = [np.unique(np.asarray(x, dtype=np.int64)) for x in cvi_clu]
    n = len(cvi)
    for _ in range(n_repeat):
        cvi1 = [np.zeros(0, np.int64) for _ in range(n)]
        for i in range(n):
            if len(cvi[i]) <= 1:
                continue
            members = cvi[i]
            grown = np.unique(np.concatenate([members] + [cvi[m - 1] for m in members]))
            for m in grown:
                cvi1[m - 1] = grown
        if all(len(a) == len(b) for a, b in zip(cvi1, cvi)):
            break
        cvi = cvi1
    vi_map = np.arange(1, n + 1, dtype=np.int64)
    for i in range(n):
        if len(cvi[i]):
            vi_map[i] = cvi[i].min()
    uniq = np.unique(vi_map)
    uniq = uniq[uniq != 0]
    remap = np.zeros(n + 1, np.int64)
    remap[uniq] = np.arange(1, len(uniq) + 1)
    return remap[vi_map], uniq


def peak_overlap_lists(knn_peak):
    """cvi_peak{i} = peaks whose kNN column shares any spike with peak i's (itself included)."""
    K = np.asarray(knn_peak)
    P = K.shape[1]
    sets = [set(K[:, i].tolist()) for i in range(P)]
    return [np.array([j + 1 for j in range(P) if sets[j] & sets[i]], dtype=np.int64) for i in range(P)]


def assign_cluster_knn(rho, delta, nneigh, knn, delta_cut=1.0, n_repeat=10):
    """assignCluster_knn_: returns (viClu_spk int32, viSpk_peak 1-based, nClu_pre).
    knn: k x N (1-based spike ids) -- the peaks' columns are read from it."""
    rho = np.asarray(rho)
    delta = np.asarray(delta)
    nneigh = np.asarray(nneigh, dtype=np.int64)
    N = rho.size
    peaks = np.flatnonzero(delta >= delta_cut) + 1
    n_pre = len(peaks)
    clu = np.zeros(N, np.int32)
    ord_rho = np.argsort(-rho.astype(np.float64), kind="stable") + 1      # sort(vrRho, 'descend')
    pk_order = np.argsort(-rho[peaks - 1].astype(np.float64), kind="stable")
    peaks = peaks[pk_order]
    knn = np.asarray(knn)
    cvi = peak_overlap_lists(knn[:, peaks - 1]) if len(peaks) else []
    if len(peaks):
        m, uniq = cell2map(cvi)
        clu[peaks - 1] = m
        peaks = peaks[uniq - 1]
    if len(peaks) <= 1:
        return np.ones(N, np.int32), ord_rho[:1], n_pre
    nn1 = nneigh[ord_rho - 1]
    for _ in range(n_repeat):
        todo = np.flatnonzero(clu[ord_rho - 1] <= 0)
        if len(todo) == 0:
            break
        for ii in todo:
            clu[ord_rho[ii] - 1] = clu[nn1[ii] - 1]
        if not (clu <= 0).any():
            break
    return clu, peaks, n_pre


def remove_count(clu, icl, min_count):
    """dpclus_remove_count_: clusters with fewer than min_count spikes removed, ids compacted."""
    n = len(icl)
    counts = np.bincount(np.asarray(clu)[np.asarray(clu) > 0], minlength=n + 1)[1:n + 1]
    keep = counts >= min_count
    if keep.all():
        return np.asarray(clu), np.asarray(icl)
    vmap = np.zeros(n + 1, np.int64)
    vmap[1:][keep] = np.arange(1, keep.sum() + 1)
    c = np.asarray(clu)
    out = np.where(c > 0, vmap[np.clip(c, 0, n)], 0).astype(np.int32)
    return out, np.asarray(icl)[keep]


def isolation_score(clu, knn, n_clu=None):
    """isolation_score_ (matlab/irc2.m:6699-6731): for cluster c, the labels of its spikes' kNN;
    n11 = how many are c, c2 = the most frequent other positive label (ties to the smaller, as
    MATLAB's mode), n12 = its count; score = n11 / (n11 + n12), or n11 / n11 with c2 = NaN when no
    other cluster occurs (MATLAB's mode of an empty set is NaN).  Returns (score, c2) float64."""
    clu = np.asarray(clu, dtype=np.int64)
    knn = np.asarray(knn, dtype=np.int64)
    n_clu = int(clu.max()) if n_clu is None else int(n_clu)
    score = np.full(n_clu, np.nan)
    nn = np.full(n_clu, np.nan)
    for c in range(1, n_clu + 1):
        spikes = np.flatnonzero(clu == c)
        lab = clu[knn[:, spikes] - 1].ravel()
        n11 = int((lab == c).sum())
        other = lab[(lab != c) & (lab > 0)]
        if other.size == 0:
            score[c - 1] = n11 / n11 if n11 else np.nan
            continue
        vals, cnt = np.unique(other, return_counts=True)
        c2 = int(vals[np.argmax(cnt)])            # first maximum = smallest label among ties
        n12 = int((lab == c2).sum())
        score[c - 1] = n11 / (n11 + n12)
        nn[c - 1] = c2
    return score, nn


def knn_overlap_merge(clu, knn, thresh, n_clu=None):
    """knn_overlap_merge_ (matlab/irc2.m:2156-2174): clusters scoring below thresh are grouped with
    their most frequent neighbour cluster through cell2map_; returns the relabelled clusters."""
    score, nn = isolation_score(clu, knn, n_clu)
    n = len(score)
    groups = [[c, int(nn[c - 1])] if score[c - 1] < thresh else [c] for c in range(1, n + 1)]
    m, _ = cell2map(groups)
    clu = np.asarray(clu).copy()
    upd = clu > 0
    clu[upd] = m[clu[upd] - 1]
    return clu
