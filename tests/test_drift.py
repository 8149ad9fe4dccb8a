"""Drift linkage (calc_drift_): the oracle pinned on CPU, the gfx950 entry point against it."""
import numpy as np
import pytest

from ironclust_amd import synth
from oracle import drift as DR


def test_lim_drift_is_the_equal_count_rule():
    for N, C in ((1000, 7), (10_000_000, 200), (5, 5), (12345, 1)):
        np.testing.assert_array_equal(DR.lim_drift(N, C), synth.lim_equal(N, C))


def test_quantile_rounds_half_away_from_zero():
    # n = 5, q = 0.5: round(2.5) = 3 in MATLAB (numpy's round would give 2)
    q = DR.quantile_vr(np.array([5, 1, 4, 2, 3], np.float32), [0.0, 0.5, 1.0])
    np.testing.assert_array_equal(q, [1, 3, 5])
    q = DR.quantile_vr(np.array([10, 20], np.float32), [0.0, 0.25, 0.75, 1.0])   # round(0.5)=1, round(1.5)=2
    np.testing.assert_array_equal(q, [10, 10, 20, 20])


def test_subsample_sizes_and_positions():
    v = np.arange(1, 250_002)                     # 250001 values, skip 2 -> 125001, remove 25001
    s = DR.subsample_vr(v, 100_000)
    assert s.size == 100_000
    # the removed positions are round(linspace(1, 125001, 25001)) = 1, 6, 11, ... of the skipped list
    kept_skipped = np.arange(1, 250_002, 2)
    removed = kept_skipped[np.arange(0, 125_001, 5)]
    assert not np.isin(removed, s).any()
    assert np.array_equal(s[:4], kept_skipped[[1, 2, 3, 4]])
    assert DR.subsample_vr(np.arange(10), 100).size == 10


def test_histogram_edges_semantics():
    lim = np.array([1, 6], np.int64)
    amp = np.array([0.0, 1.0, 2.0, 3.0, -1.0], np.float32)    # 3.0 = top edge (closed), -1 outside
    pos = np.array([0.0, 0.0, 1.0, 2.0, 0.0], np.float32)
    c = DR.hist_counts(amp, pos, lim, np.array([0, 1, 2, 3], np.float32), np.array([0, 1, 2], np.float32))
    h = np.zeros((3, 2), np.int64)
    h[0, 0] += 1          # (0, 0)
    h[1, 0] += 1          # (1, 0)
    h[2, 1] += 1          # (2, 1)
    h[2, 1] += 1          # (3 on the closed top edge, 2 on the closed top edge)
    np.testing.assert_array_equal(c[:, 0], h.reshape(-1, order="F"))


def test_links_match_an_independent_ranking():
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 20, size=(24, 30))
    ml = DR.links_from_counts(counts, 6, 256)
    d = cdist(counts.T.astype(np.float64), counts.T.astype(np.float64))     # exact for these integers
    for c in range(30):
        order = np.argsort(d[:, c], kind="stable")
        assert set(np.flatnonzero(ml[:, c])) == set(order[:6])
        assert ml[c, c]                                                       # a chunk is nearest itself
    # identical histograms tie: the lower chunk index wins
    tied = np.tile(counts[:, :1], (1, 5))
    ml = DR.links_from_counts(tied, 2, 256)
    assert ml[:, 3].tolist() == [True, True, False, False, False]


def test_band_mask_for_long_recordings():
    m = DR.drift_mask(40, 16)              # q = 4
    assert m[0, 0] and m[8, 0] and not m[9, 0]          # corner block: rows 0..8 for columns 0..3
    assert m[6, 10] and not m[5, 10] and m[14, 10] and not m[15, 10]    # band |r - c| <= 4
    assert m[31, 39] and not m[30, 39]                  # corner block: rows 31..39 for columns 36..39
    rng = np.random.default_rng(6)
    counts = rng.integers(0, 9, size=(6, 40))
    ml = DR.links_from_counts(counts, 5, 16)
    assert not (ml & ~m).any()
    assert ml.sum(0).tolist() == [5] * 40


def test_single_chunk_when_batch_covers_all():
    ml, lim = DR.calc_drift(np.ones(50, np.float32), np.ones(50, np.float32), 4, 3, 5, 5)
    assert ml.shape == (1, 1) and ml[0, 0]
    np.testing.assert_array_equal(lim, [1, 51])


def drift_spikes(N, n_sites, C, seed):
    """Amplitudes and y-positions of a recording whose probe drifts: positions move with a
    random walk over time, so chunks with similar offsets have similar histograms."""
    rng = np.random.default_rng(seed)
    t = np.arange(N)
    chunk = np.minimum(t * C // N, C - 1)
    walk = np.cumsum(rng.normal(0, 8.0, C))
    y = rng.integers(0, n_sites, N) * 20.0 + rng.normal(0, 5.0, N) + walk[chunk]
    amp = rng.lognormal(4.0, 0.5, N)
    return amp.astype(np.float32), y.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("C,n_batch,n_time_max", [(60, 10, 256), (300, 15, 256), (8, 3, 4)])
def test_calc_drift_matches_oracle(ctx, C, n_batch, n_time_max):
    amp, pos = drift_spikes(240_000, 32, C, seed=C)
    link, lim, counts = ctx.calc_drift(amp, pos, 32, 10, C, n_batch, n_time_max, want_counts=True)
    ml_r, lim_r = DR.calc_drift(amp, pos, 32, 10, C, n_batch, n_time_max)
    np.testing.assert_array_equal(lim, lim_r)
    qa = DR.quantile_vr(amp, np.arange(11) / 10)
    qp = DR.quantile_vr(pos, np.arange(33) / 32)
    np.testing.assert_array_equal(counts, DR.hist_counts(amp, pos, lim_r, qa, qp))
    np.testing.assert_array_equal(link, ml_r)


@pytest.mark.gpu
def test_calc_drift_large_input_subsample_and_single_chunk(ctx):
    amp, pos = drift_spikes(1_000_003, 16, 40, seed=9)      # quantiles from a 100000-value subsample
    link, lim, _ = ctx.calc_drift(amp, pos, 16, 10, 40, 7, 256)
    ml_r, lim_r = DR.calc_drift(amp, pos, 16, 10, 40, 7, 256)
    np.testing.assert_array_equal(link, ml_r)
    np.testing.assert_array_equal(lim, lim_r)
    link, lim, _ = ctx.calc_drift(amp[:1000], pos[:1000], 16, 10, 5, 5, 256)
    assert link.shape == (1, 1) and link[0, 0]
    np.testing.assert_array_equal(lim, [1, 1001])
