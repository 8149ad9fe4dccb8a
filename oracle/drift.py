"""Drift linkage oracle (test infrastructure only; never the product path).

CPU restatement of calc_drift_ and its helpers (reference paths relative to the ironclust
repository):

  calc_drift_      matlab/irc2.m:6373-6447  chunk limits, amplitude/position quantiles,
                                             per-chunk 2-D histograms, nearest-chunk links
  dist_mask_ml_    matlab/irc2.m:6452-6462  masked nearest chunks (long recordings)
  quantile_vr_     matlab/irc2.m:6478-6486  quantiles of a subsample
  subsample_vr_    matlab/irc2.m:1557-1572  deterministic subsample
  mi2ml_drift_     matlab/irc2.m:6508-6515  member lists -> link matrix

histcounts2 (MATLAB) bins are [e_i, e_{i+1}) with the last bin closed, values outside the
edges uncounted -- numpy.histogram2d's convention.  Counts are integers, so the squared
distances between count vectors are exact integers and the chunk ranking is by
(distance^2, chunk index): MATLAB's sort is stable and sqrt is monotone.
"""
import numpy as np


def lim_drift(n_spikes, n_drift):
    """viLim_drift = int64([0, ceil((1:C)/C*N)]) + 1 (matlab/irc2.m:6395), 1-based."""
    c = np.arange(1, n_drift + 1, dtype=np.float64)
    return np.concatenate([[0], np.ceil(c / n_drift * n_spikes)]).astype(np.int64) + 1


def _round_half_away(x):
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def subsample_vr(v, n_max=100000):
    """subsample_vr_ (matlab/irc2.m:1557-1572)."""
    v = np.asarray(v)
    if v.size <= n_max:
        return v
    n_skip = v.size // n_max
    if n_skip > 1:
        v = v[::n_skip]
    if v.size > n_max:
        n_remove = v.size - n_max
        rem = _round_half_away(np.linspace(1, v.size, n_remove)).astype(np.int64)
        rem = np.clip(rem, 1, v.size)
        keep = np.ones(v.size, bool)
        keep[rem - 1] = False
        v = v[keep]
    return v


def quantile_vr(v, q, n_max=100000):
    """quantile_vr_ (matlab/irc2.m:6478-6486): sorted subsample at max(min(round(n q), n), 1)."""
    v = subsample_vr(np.asarray(v, dtype=np.float32), n_max)
    n = v.size
    idx = np.maximum(np.minimum(_round_half_away(n * np.asarray(q, dtype=np.float64)), n), 1).astype(np.int64)
    return np.sort(v, kind="stable")[idx - 1]


def hist_counts(amp, pos, lim, amp_edges, pos_edges):
    """mrCount_drift: (nAmp*nPos, C) counts, column-major bins (iAmp + nAmp*iPos) as mn_(:)."""
    C = len(lim) - 1
    nA, nP = len(amp_edges) - 1, len(pos_edges) - 1
    out = np.zeros((nA * nP, C), dtype=np.int64)
    for c in range(C):
        a = np.asarray(amp[lim[c] - 1:lim[c + 1] - 1], dtype=np.float32)
        p = np.asarray(pos[lim[c] - 1:lim[c + 1] - 1], dtype=np.float32)
        h, _, _ = np.histogram2d(a.astype(np.float64), p.astype(np.float64),
                                 bins=[np.asarray(amp_edges, np.float64), np.asarray(pos_edges, np.float64)])
        out[:, c] = h.astype(np.int64).reshape(-1, order="F")
    return out


def drift_mask(C, n_time_max):
    """mlMask of calc_drift_ for C > nTime_max_drift (matlab/irc2.m:6425-6430): |i-j| <= q plus
    the corner blocks, q = nTime_max_drift / 4.  mask[r, c] = row r usable for column c."""
    q = n_time_max // 4
    i = np.arange(C)
    m = np.abs(i[None, :] - i[:, None]) <= q
    m[:2 * q + 1, :q] = True
    m[C - 2 * q - 1:, C - q:] = True
    return m


def links_from_counts(counts, n_batch, n_time_max):
    """mlDrift (C, C) bool: column c marks the n_batch chunks nearest to c by Euclidean distance
    of their count vectors, ties to the lower chunk index (matlab/irc2.m:6425-6437, 6452-6462)."""
    Fm = np.asarray(counts, dtype=np.int64)
    C = Fm.shape[1]
    d2 = ((Fm[:, :, None] - Fm[:, None, :]) ** 2).sum(0)        # exact integers
    mask = drift_mask(C, n_time_max) if C > n_time_max else np.ones((C, C), bool)
    ml = np.zeros((C, C), bool)
    for c in range(C):
        rows = np.flatnonzero(mask[:, c])
        order = rows[np.lexsort((rows, d2[rows, c]))]
        ml[order[:n_batch], c] = True
    return ml


def calc_drift(amp, pos, n_pos, n_amp, n_drift, n_batch, n_time_max=256):
    """Return (mlDrift (C, C) bool, viLim_drift (C+1,) 1-based) as calc_drift_ does."""
    N = len(amp)
    if n_batch >= n_drift:
        return np.ones((1, 1), bool), np.array([1, N + 1], np.int64)
    lim = lim_drift(N, n_drift)
    qa = quantile_vr(amp, np.arange(n_amp + 1) / n_amp)
    qp = quantile_vr(pos, np.arange(n_pos + 1) / n_pos)
    counts = hist_counts(amp, pos, lim, qa, qp)
    return links_from_counts(counts, n_batch, n_time_max), lim
