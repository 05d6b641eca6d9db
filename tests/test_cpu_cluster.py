"""CPU pins of the oracle's exact clustering at the threshold layer (reduction, centroid_sel and
coarse_cluster, src/SDGC/SNICIT/kernel.hpp:8-69,138-194; pool and compaction,
src/SDGC/SNICIT/SNICIT.hpp:406-509).

Each stage is checked against what the mathematics fixes, not against a second copy of the
oracle's loops:
  * the pool is the first 32 live rows of Y_t, and Y_t is the plain forward of t layers;
  * the segment sums equal float64 numpy sums within the FP32 summation error bound;
  * Alg. 1's survivors are characterised without running it: a pool row is pruned iff an
    earlier survivor lies within eps on every segment sum (a row can only be pruned by an
    earlier base, and every base that prunes survives), which fixes the set uniquely;
  * the assignment is numpy's first argmin of the L0 distances to the centroids;
  * ne(i) = (L0 distance > 0) for non-centroid rows;
  * a hand-built input fixes the centroid set, the assignment (including ties) and ne by
    construction.
"""
import numpy as np
import pytest

import oracle as orc
from tests import nets

EPS = 0.03


def seg_sums64(Y, segments, starts=None):
    """float64 segment sums of the rows of Y (segment g = N/segments neurons from starts[g])."""
    N = Y.shape[1]
    L = N // segments
    if starts is None:
        starts = [g * L for g in range(segments)]
    return np.stack([Y[:, s:s + L].astype(np.float64).sum(axis=1) for s in starts], axis=1)


def close(Fa, Fb):
    """The reference's pruning test between two rows of segment sums: the count of sums that
    differ by more than 0.03 (compared in double) is below segments * 0.03."""
    cnt = int((np.abs(Fa.astype(np.float64) - Fb.astype(np.float64)) > EPS).sum())
    return cnt < len(Fa) * EPS


def check_alg1_characterisation(F, pool, centroids):
    """Survivors = pool rows with no earlier survivor within eps; pruned rows have one."""
    surv = set(int(c) for c in centroids)
    assert surv <= set(int(p) for p in pool)
    order = {int(s): q for q, s in enumerate(pool)}
    assert sorted(surv, key=order.get) == [int(c) for c in centroids], "centroids in pool order"
    for q, s in enumerate(pool):
        earlier = [r for r in range(q) if int(pool[r]) in surv]
        has_close = any(close(F[q], F[r]) for r in earlier)
        if int(s) in surv:
            assert not has_close, f"survivor {s} is within eps of an earlier survivor"
        else:
            assert has_close, f"pruned row {s} has no earlier survivor within eps"


def sum_error_bound(F64, n_terms):
    """|fl(sum) - sum| <= (n - 1) u sum for n nonnegative terms (u = 2^-24), plus a margin."""
    return (n_terms - 1) * 2.0 ** -24 * np.abs(F64) + 1e-6


@pytest.mark.parametrize("t", [2, 6])
def test_cluster_stages_against_numpy(t):
    N, L, B = 1024, 10, 400
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    y0[200:260] = y0[3]  # duplicates of a pool row: zero distance, pruned in the pool
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, t)
    Yt, _, _ = orc.plain_f32(cols[:t], vals[:t], -0.3, 32.0, y0)
    live = np.flatnonzero(Yt.any(axis=1))
    assert r["live_at_t"] == len(live) > 40
    np.testing.assert_array_equal(r["pool"], live[:32])
    # segment sums
    F64 = seg_sums64(Yt[live[:32]], 16)
    assert np.all(np.abs(r["F"] - F64) <= sum_error_bound(F64, N // 16))
    # Alg. 1
    check_alg1_characterisation(r["F"], r["pool"], r["centroids"])
    # nearest centroid: first minimum of the L0 distance
    C = Yt[r["centroids"]]
    d = (Yt[live][:, None, :] != C[None, :, :]).sum(axis=2)
    first = np.argmin(d, axis=1)
    np.testing.assert_array_equal(r["assign"][live], first)
    dead = np.setdiff1d(np.arange(B), live)
    assert (r["assign"][dead] == -1).all()
    # non-empty delta rows: non-centroids with a nonzero distance
    is_cent = np.isin(live, r["centroids"])
    assert r["nerow_at_t"] == int(((d.min(axis=1) > 0) & ~is_cent).sum())
    # the duplicated pool row is one centroid; its copies map to it with distance 0
    assert 3 in r["centroids"]
    c3 = int(np.flatnonzero(r["centroids"] == 3)[0])
    assert (r["assign"][200:260] == c3).all()


def test_crafted_clustering_case():
    """Centroids, assignment (with ties) and ne fixed by construction."""
    y0, e = nets.crafted_cluster_case()
    N = y0.shape[1]
    cols, vals = nets.sdgc(N, 3, seed=7)
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, 0)
    np.testing.assert_array_equal(r["pool"], e["pool"])
    # segment sums of multiples of 1/64 are exact
    np.testing.assert_array_equal(r["F"].astype(np.float64), seg_sums64(y0[e["pool"]], 16))
    np.testing.assert_array_equal(r["centroids"], e["centroids"])
    check_alg1_characterisation(r["F"], r["pool"], r["centroids"])
    np.testing.assert_array_equal(r["assign"], e["assign"])
    assert r["nerow_at_t"] == int(e["ne"].sum())
    # categories still equal the plain forward's
    _, cat, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    np.testing.assert_array_equal(r["cat"], cat)


def test_pruning_is_sequential():
    """A pruned row prunes nothing: rows 8 (= 7 + 1/64) and 9 (= 7 + 2/64) are within eps
    of each other, but 8 is pruned by 7 first, so 9 survives; likewise 11 and 12."""
    y0, e = nets.crafted_cluster_case()
    F = seg_sums64(y0, 16)
    assert close(F[7], F[8]) and close(F[8], F[9]) and not close(F[7], F[9])
    assert close(F[10], F[11]) and close(F[11], F[12]) and not close(F[10], F[12])
    for s in (7, 9, 10, 12):
        assert s in e["centroids"]
    for s in (8, 11):
        assert s not in e["centroids"]


def test_q2_reference_segments():
    """With the reference's indexing (blockDim (64, 16), kernel.hpp:147-148) segment g of an
    N = 4096 row covers neurons [64 g, 64 g + 256): the segments overlap and the last
    N - 1216 neurons are never summed.  The intended layout sums [256 g, 256 g + 256)."""
    N = 4096
    rng = np.random.default_rng(4)
    y0 = (rng.integers(0, 5, size=(40, N)) / 4.0).astype(np.float32)
    cols, vals = nets.sdgc(N, 2, seed=3)
    ref = orc.snicit_exact(cols, vals, -0.4, 32.0, y0, 0, ref_segments=True)
    own = orc.snicit_exact(cols, vals, -0.4, 32.0, y0, 0)
    pool = own["pool"]
    np.testing.assert_array_equal(ref["pool"], pool)
    np.testing.assert_array_equal(own["F"].astype(np.float64), seg_sums64(y0[pool], 16))
    np.testing.assert_array_equal(ref["F"].astype(np.float64),
                                  seg_sums64(y0[pool], 16, [64 * g for g in range(16)]))
    # a change past neuron 64*15 + 256 = 1216 is invisible to the reference's sums
    y1 = y0.copy()
    y1[:, 2000:] = 0.0
    ref1 = orc.snicit_exact(cols, vals, -0.4, 32.0, y1, 0, ref_segments=True)
    np.testing.assert_array_equal(ref1["F"], ref["F"])
    # at N = 1024 both layouts are the same segments
    cols1, vals1 = nets.sdgc(1024, 2, seed=3)
    a = orc.snicit_exact(cols1, vals1, -0.3, 32.0, y0[:, :1024], 0, ref_segments=True)
    b = orc.snicit_exact(cols1, vals1, -0.3, 32.0, y0[:, :1024], 0)
    np.testing.assert_array_equal(a["F"], b["F"])
    np.testing.assert_array_equal(a["centroids"], b["centroids"])
    # the reference's indexing runs past the row below N = 1024 / 16 * 15 + N / 16
    with pytest.raises(AssertionError):
        colsm, valsm = nets.sdgc(512, 2, seed=3)
        orc.snicit_exact(colsm, valsm, -0.3, 32.0, y0[:, :512], 0, ref_segments=True)


def test_alg1_prunes_identical_rows():
    """Identical pool rows collapse to one centroid per distinct row, every copy maps to its
    own distinct row's centroid with distance 0, and nothing is left in the delta list."""
    N, L = 256, 6
    cols, vals = nets.random_ell(N, L, K=32, seed=1, share=4)
    base = nets.random_inputs(N, 4, seed=5, density=0.3)
    y0 = np.repeat(base, 16, axis=0)  # 4 distinct rows, 16 copies each
    r = orc.snicit_exact(cols, vals, -0.01, 32.0, y0, 0)
    assert r["live_at_t"] == 64
    np.testing.assert_array_equal(r["centroids"], [0, 16])  # the pool holds rows 0..31
    # rows 32.. are copies of rows the pool does not hold: nearest of the two centroids
    assert r["nerow_at_t"] == 32
    np.testing.assert_array_equal(r["assign"][:16], 0)
    np.testing.assert_array_equal(r["assign"][16:32], 1)
    _, cat, _ = orc.plain_f32(cols, vals, -0.01, 32.0, y0)
    np.testing.assert_array_equal(r["cat"], cat)
