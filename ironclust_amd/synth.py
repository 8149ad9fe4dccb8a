"""Seeded synthetic spike-feature recordings for the density-peak hot path.

The reference's path consumes per-site PCA features written by its detect stage;
here they are synthesized directly: per site, a mixture of Gaussian clusters whose
centres drift with a rigid random walk over time.  Workloads (sizes as used by the
benchmarks and tests):

  plain-500k   64 sites, 500,000 spikes, D=24, 8 clusters/site, 1 drift chunk, k=30
  drift-10M    384 sites, 10,000,000 spikes, D=42, 10 clusters/site, 200 chunks, drift step
               N(0, 0.3^2), each chunk linked to its 15 nearest chunks within +/-64, k=30
  lognormal-10M  as drift-10M with lognormal per-site counts (median 20k, 500..600k) and
               Dirichlet(2) per-chunk shares
  multiday-100M  100M spikes, 384 sites, D=42, 2000 chunks, +/-64 band, 15 links, k=30

Drift chunks follow viLim_drift = [0, ceil((1:C)/C*N)] + 1 (matlab/irc2.m:6395) for
equal-count chunks; links mirror calc_drift_'s nearest-chunk rule restricted to a band
(matlab/irc2.m:6425-6437).
"""
import math

import numpy as np
import torch

WORKLOADS = {
    "plain-500k": dict(n_spikes=500_000, n_sites=64, D=24, n_drift=1, n_clusters=8, k=30,
                       drift_step=0.0, n_links=1, band=0, seed=1),
    "drift-10M": dict(n_spikes=10_000_000, n_sites=384, D=42, n_drift=200, n_clusters=10, k=30,
                      drift_step=0.3, n_links=15, band=64, seed=2),
    "lognormal-10M": dict(n_spikes=10_000_000, n_sites=384, D=42, n_drift=200, n_clusters=10, k=30,
                          drift_step=0.3, n_links=15, band=64, seed=4, lognormal=True),
    "multiday-100M": dict(n_spikes=100_000_000, n_sites=384, D=42, n_drift=2000, n_clusters=10, k=30,
                          drift_step=0.3, n_links=15, band=64, seed=5),
}


def lim_equal(n_spikes, n_drift):
    """viLim_drift for equal-count chunks: [0, ceil((1:C)/C*N)] + 1 (1-based)."""
    c = np.arange(1, n_drift + 1, dtype=np.float64)
    ends = np.ceil(c / n_drift * n_spikes).astype(np.int64)
    return np.concatenate([[0], ends]).astype(np.int64) + 1


def drift_links(delta, n_links, band):
    """link[r, c] = True for the n_links chunks r within |r-c| <= band whose drift offsets are
    nearest to chunk c's (self included, ties to the lower r).  Column c lists the chunks
    whose spikes are candidates for queries in chunk c, as mlDrift(:, c)."""
    C = len(delta)
    link = np.zeros((C, C), dtype=np.uint8)
    for c in range(C):
        lo, hi = max(0, c - band), min(C, c + band + 1)
        r = np.arange(lo, hi)
        d = np.abs(np.asarray(delta[lo:hi], dtype=np.float64) - float(delta[c]))
        order = np.lexsort((r, d))
        link[r[order[:n_links]], c] = 1
    return link


def lognormal_counts(n_sites, n_spikes, rng, median=20_000, lo=500, hi=600_000, sigma=1.0):
    x = rng.lognormal(mean=math.log(median), sigma=sigma, size=n_sites)
    x = np.clip(x, lo, hi)
    for _ in range(50):
        x = np.clip(x * (n_spikes / x.sum()), lo, hi)
    n = np.floor(x).astype(np.int64)
    n[np.argmax(n)] += n_spikes - n.sum()
    return n


PIECE = 1 << 22


def _draw(fn, n, gen, seed, tag, device, **kw):
    """fn(size, generator=...) over n values in pieces of PIECE: the first piece from the shared
    generator (so recordings of up to PIECE spikes draw exactly as one call), every further piece
    from its own generator seeded by (seed, tag, piece) -- one huge draw from a single device
    generator is avoided, and the pieces are independent streams."""
    if n <= PIECE:
        return fn(n, generator=gen, **kw)
    parts = []
    for i, a in enumerate(range(0, n, PIECE)):
        b = min(n, a + PIECE)
        g = gen if i == 0 else torch.Generator(device=device).manual_seed(
            (seed * 1_000_003 + tag * 10_007 + i) % (1 << 62))
        parts.append(fn(b - a, generator=g, **kw))
    return torch.cat(parts)


def make_dataset(n_spikes, n_sites, D, n_drift, n_clusters, k=30, drift_step=0.0, n_links=1, band=0,
                 seed=1, lognormal=False, f2=False, f2_frac=0.5, center_std=4.0, device="cpu",
                 link=None):
    """Return a dict with spike-level arrays (ids are time-ordered, 1..N):
    F1 (N, D) float32, F2 (N, D) or None, site1 (N,) int32 1-based, site2 (N,) int32 (0 = none)
    or None, lim (C+1,) int64 (viLim_drift), link (C, C) uint8 (link[r, c] = mlDrift(r, c)),
    plus sizes.  Arrays are torch tensors on `device` (link and lim are numpy)."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device).manual_seed(seed)
    S, C, N = n_sites, n_drift, n_spikes
    # drift offsets per chunk: rigid random walk
    steps = rng.normal(0.0, drift_step, size=C) if drift_step > 0 else np.zeros(C)
    steps[0] = 0.0
    delta = np.cumsum(steps)
    if link is None:
        if C == 1:
            link = np.ones((1, 1), np.uint8)
        else:
            link = drift_links(delta, n_links, band)
    if lognormal:
        counts = lognormal_counts(S, N, rng)
        shares = rng.dirichlet(np.full(C, 2.0), size=S)           # (S, C)
        nsc = np.floor(shares * counts[:, None]).astype(np.int64)
        rem = counts - nsc.sum(1)
        for s in range(S):
            if rem[s] > 0:
                np.add.at(nsc[s], rng.integers(0, C, size=rem[s]), 1)
        chunk_tot = nsc.sum(0)
        lim = np.concatenate([[0], np.cumsum(chunk_tot)]).astype(np.int64) + 1
        site_parts = []
        for c in range(C):
            sv = np.repeat(np.arange(1, S + 1, dtype=np.int32), nsc[:, c])
            rng.shuffle(sv)
            site_parts.append(sv)
        site1 = torch.from_numpy(np.concatenate(site_parts)).to(device)
    else:
        lim = lim_equal(N, C)
        site1 = _draw(lambda n, **kw: torch.randint(1, S + 1, (n,), device=device, dtype=torch.int32, **kw),
                      N, gen, seed, 1, device)
    chunk = torch.zeros(N, dtype=torch.int64, device=device)
    bounds = torch.from_numpy(lim[1:-1] - 1).to(device)           # 0-based starts of chunks 1..C-1
    if C > 1:
        chunk = torch.bucketize(torch.arange(N, device=device), bounds, right=True)
    centers = torch.randn((S, n_clusters, D), generator=gen, device=device) * center_std
    u = torch.randn((S, D), generator=gen, device=device)
    u = u / u.norm(dim=1, keepdim=True)
    dl = torch.from_numpy(delta.astype(np.float32)).to(device)
    z = _draw(lambda n, **kw: torch.randint(0, n_clusters, (n,), device=device, **kw), N, gen, seed, 2, device)
    s0 = site1.long() - 1

    def feats(site0, zz, tag):
        out = torch.empty((N, D), dtype=torch.float32, device=device)
        for i, a in enumerate(range(0, N, PIECE)):
            b = min(N, a + PIECE)
            g = gen if i == 0 else torch.Generator(device=device).manual_seed(
                (seed * 1_000_003 + tag * 10_007 + i) % (1 << 62))
            x = torch.randn((b - a, D), generator=g, device=device)
            x += centers[site0[a:b], zz[a:b]]
            x += dl[chunk[a:b]][:, None] * u[site0[a:b]]
            out[a:b] = x
        return out

    F1 = feats(s0, z, 3)
    F2 = site2 = None
    if f2 and S > 1:
        has2 = _draw(lambda n, **kw: torch.rand((n,), device=device, **kw), N, gen, seed, 4, device) < f2_frac
        off = _draw(lambda n, **kw: torch.randint(1, S, (n,), device=device, **kw), N, gen, seed, 5, device)
        s2 = (s0 + off) % S
        site2 = torch.where(has2, (s2 + 1).int(), torch.zeros_like(site1))
        z2 = (z + 1) % n_clusters
        F2 = feats(s2, z2, 6)
        F2[~has2] = 0.0
    return dict(n_spikes=N, n_sites=S, D=D, n_drift=C, k=k, F1=F1, F2=F2, site1=site1, site2=site2,
                lim=lim, link=link, delta=delta)


def make_workload(name, device="cpu", **over):
    cfg = dict(WORKLOADS[name])
    cfg.update(over)
    return make_dataset(device=device, **cfg)


def n_pairs(ds):
    """Sum over queries of |C(a)| (candidates per query, self included): the algorithmic
    size of one pass.  Computed from per-(site, chunk) counts."""
    S, C = ds["n_sites"], ds["n_drift"]
    lim = ds["lim"]
    N = ds["n_spikes"]
    site1 = ds["site1"].long().cpu()
    ch = torch.bucketize(torch.arange(N), torch.from_numpy(lim[1:-1] - 1), right=True)
    cnt = torch.zeros((S, C), dtype=torch.int64)
    cnt.index_put_((site1 - 1, ch), torch.ones(N, dtype=torch.int64), accumulate=True)
    cand = cnt.clone()
    if ds.get("site2") is not None:
        s2 = ds["site2"].long().cpu()
        m = s2 > 0
        cand.index_put_((s2[m] - 1, ch[m]), torch.ones(int(m.sum()), dtype=torch.int64), accumulate=True)
    L = torch.from_numpy(ds["link"].astype(np.int64))            # (C, C): L[r, c]
    cand_c = cand @ L                                              # (S, C): candidates of block (s, c)
    return int((cnt * cand_c).sum())
