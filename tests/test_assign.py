"""Density-peak assignment oracle (assignCluster_knn_, cell2map_, dpclus_remove_count_), pinned
by hand-computed cases and independent chain walks (CPU)."""
import numpy as np

from oracle import assign as A
from oracle import oracle as O
from ironclust_amd import synth


def test_cell2map_hand_cases():
    # overlap lists (self-inclusive): {1,2,3} chain together, 4 alone, {5,6}
    m, reps = A.cell2map([[1, 2], [1, 2, 3], [2, 3], [4], [5, 6], [5, 6]])
    assert list(m) == [1, 1, 1, 2, 3, 3] and list(reps) == [1, 4, 5]
    # a path 1-2-3-4-5, traced by hand through the reference's update order (later lists overwrite
    # earlier ones within a round): round 1 leaves {1..5}, {2..5}, {3,4,5} x 3, round 2 changes no
    # size, so the path splits into three clusters -- [1, 2, 3, 3, 3]
    path = [[1, 2], [1, 2, 3], [2, 3, 4], [3, 4, 5], [4, 5]]
    m, reps = A.cell2map(path)
    assert list(m) == [1, 2, 3, 3, 3] and list(reps) == [1, 2, 3]
    m, reps = A.cell2map([[1], [2], [3]])
    assert list(m) == [1, 2, 3] and list(reps) == [1, 2, 3]
    # a mutual pair and a star: {1,3} and {2,3} both reach 3 -> one cluster
    m, reps = A.cell2map([[1, 3], [2, 3], [1, 2, 3]])
    assert list(m) == [1, 1, 1] and list(reps) == [1]


def _chain_walk_labels(lab0, nneigh):
    """Independent restatement: the label of the first labelled spike on a -> nneigh(a) -> ..."""
    out = lab0.copy()
    for a in range(len(lab0)):
        x, seen = a, 0
        while out[x] == 0 and lab0[x] == 0 and nneigh[x] - 1 != x and seen <= len(lab0):
            x = nneigh[x] - 1
            seen += 1
        out[a] = lab0[x] if lab0[x] > 0 else out[x]
    return out


def test_assign_follows_chains_and_merges_overlapping_peaks():
    ds = synth.make_dataset(4000, 3, 6, 4, 3, k=10, drift_step=0.3, n_links=2, band=2, seed=51)
    r = O.sort_batch(ds)
    rho, delta, nn, knn = r["rho"], r["delta"], r["nneigh"], r["knn"].T
    for cut in (np.quantile(delta, 0.99), np.quantile(delta, 0.999)):
        clu, peaks, n_pre = A.assign_cluster_knn(rho, delta, nn, knn, delta_cut=cut)
        assert n_pre == int((delta >= cut).sum())
        assert clu.min() >= 1
        # peaks: every original peak carries the cluster its kNN-overlap group maps to
        pk = np.flatnonzero(delta >= cut) + 1
        lab0 = np.zeros(len(rho), np.int32)
        lab0[pk - 1] = clu[pk - 1]
        np.testing.assert_array_equal(clu, _chain_walk_labels(lab0, nn))
        # two peaks whose kNN lists share a spike are in one cluster
        for i in pk[:40]:
            for j in pk[:40]:
                if set(knn[:, i - 1]) & set(knn[:, j - 1]):
                    assert clu[i - 1] == clu[j - 1]
        assert len(peaks) == clu.max()


def test_remove_count():
    clu = np.array([1, 1, 2, 3, 3, 3, 0], np.int32)
    out, icl = A.remove_count(clu, np.array([10, 20, 30]), 2)
    assert list(out) == [1, 1, 0, 2, 2, 2, 0] and list(icl) == [10, 30]
    out, icl = A.remove_count(clu, np.array([10, 20, 30]), 1)
    assert list(out) == list(clu)


def test_isolation_score_hand_case():
    # 6 spikes, clusters 1 1 2 2 2 0, k = 2
    clu = np.array([1, 1, 2, 2, 2, 0])
    knn = np.array([[1, 2, 3, 4, 5, 6],
                    [3, 6, 4, 1, 1, 5]])
    s, nn = A.isolation_score(clu, knn)
    # cluster 1: labels of knn[:, [1,2]] = {1, 2} and {1, 0} -> n11 = 2, other = [2] -> n12 = 1
    # cluster 2: columns 3,4,5 -> {2,2}, {2,1}, {2,1} -> n11 = 4, other = [1,1] -> n12 = 2
    assert s[0] == 2 / 3 and nn[0] == 2
    assert s[1] == 4 / 6 and nn[1] == 1
    # merging both (threshold above both scores): one cluster
    assert list(A.knn_overlap_merge(clu, knn, 0.9)) == [1, 1, 1, 1, 1, 0]
    assert list(A.knn_overlap_merge(clu, knn, 0.5)) == list(clu)


def test_isolation_mode_ties_and_empty():
    # cluster 1's neighbours: 2 and 3 once each -> MATLAB's mode picks the smaller (2)
    clu = np.array([1, 2, 3, 4])
    knn = np.array([[1, 2, 3, 4], [2, 2, 3, 4], [3, 2, 3, 4]])
    s, nn = A.isolation_score(clu, knn)
    assert nn[0] == 2 and s[0] == 1 / 2
    # clusters 2..4 only see themselves: score 1, no neighbour (NaN)
    assert np.all(s[1:] == 1) and np.all(np.isnan(nn[1:]))


def _chain_jump_interleaved(lab0, nneigh, rng, p_stale):
    """The device's pointer jumping (k_chain_init, then ceil(log2 n) + 2 rounds of k_chain_jump,
    ironclust_amd/csrc/irc_assign.hip) run on the host with the threads of each round in a random
    order and every read of another spike's lab / nxt entry seeing, at random, the value at the
    start of the round or the one already written in this round -- the interleavings a GPU round
    may produce, since lab and nxt are updated in place."""
    n = len(lab0)
    lab = lab0.astype(np.int64).copy()
    j0 = nneigh.astype(np.int64) - 1
    nxt = np.where((lab > 0) | (j0 < 0) | (j0 >= n), np.arange(n), j0)
    rounds = 1
    while (1 << (rounds - 1)) < n:
        rounds += 1
    for _ in range(rounds + 1):
        lab_old, nxt_old = lab.copy(), nxt.copy()
        for i in rng.permutation(n):
            if lab[i] > 0:
                continue
            j = nxt[i]
            lj = lab_old[j] if rng.random() < p_stale else lab[j]
            if lj > 0:
                lab[i] = lj
            else:
                nxt[i] = nxt_old[j] if rng.random() < p_stale else nxt[j]
    return lab


def test_chain_jump_result_does_not_depend_on_interleaving():
    """k_chain_jump reads neighbours' entries while other threads write them; whatever mix of old and
    new values a round sees, every spike must end with the label of the first peak on its chain
    within the launched number of rounds.  Long chains (a line with monotone density) and a random
    density-peak forest from the oracle are both checked under many interleavings."""
    rng = np.random.default_rng(5)
    # a single path n -> n-1 -> ... -> 1 (peak at spike 1) plus a second peak mid-way
    n = 300
    nneigh = np.concatenate([[1], np.arange(1, n)]).astype(np.int64)
    lab0 = np.zeros(n, np.int32)
    lab0[0] = 1
    lab0[150] = 2
    want = _chain_walk_labels(lab0, nneigh)
    for p in (0.0, 0.5, 1.0):
        for _ in range(5):
            assert np.array_equal(_chain_jump_interleaved(lab0, nneigh, rng, p), want)
    # one path of 1024 spikes to a single peak: with every read stale (a synchronous round) its far
    # end needs 10 doublings and one more round to take the label, 11 of the 12 rounds launched
    n = 1024
    nneigh = np.concatenate([[1], np.arange(1, n)]).astype(np.int64)
    lab0 = np.zeros(n, np.int32)
    lab0[0] = 7
    for p in (0.0, 1.0):
        assert np.array_equal(_chain_jump_interleaved(lab0, nneigh, rng, p), np.full(n, 7))
    # the nearest-denser forest of a small recording, peaks labelled by delta
    ds = synth.make_dataset(3000, 2, 5, 3, 3, k=8, drift_step=0.3, n_links=2, band=1, seed=77)
    r = O.sort_batch(ds)
    nn = r["nneigh"].astype(np.int64)
    order = np.argsort(-r["delta"])
    lab0 = np.zeros(len(nn), np.int32)
    lab0[order[:12]] = np.arange(1, 13)
    lab0[nn - 1 == np.arange(len(nn))] = 13    # every chain end (no denser spike) is a peak
    want = _chain_walk_labels(lab0, nn)
    assert (want > 0).all()
    for p in (0.0, 0.3, 0.7, 1.0):
        for _ in range(3):
            assert np.array_equal(_chain_jump_interleaved(lab0, nn, rng, p), want)
