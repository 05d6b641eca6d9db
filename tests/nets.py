"""Seeded nets and inputs shared by the tests (numpy only)."""
import numpy as np

import snicit_amd as sa


def sdgc(N, L, seed=7, weight=0.0625):
    cols = sa.gen_sdgc_cols(N, L, seed)
    vals = np.full(cols.shape, weight, np.float32)
    return cols, vals


def random_ell(N, L, K=32, seed=0, scale=None, share=1):
    """Rows with K distinct random inputs and N(0, scale) weights.  ``share`` consecutive rows
    read the same inputs (share=1: no panel sharing)."""
    rng = np.random.default_rng(seed)
    cols = np.empty((L, N, K), np.int32)
    for l in range(L):
        for g in range(0, N, share):
            c = rng.choice(N, size=K, replace=False)
            cols[l, g:g + share] = c
    if scale is None:
        scale = np.sqrt(2.0 / K)
    vals = (rng.standard_normal((L, N, K)) * scale).astype(np.float32)
    return cols, vals


def images(N, B, seed=11, first=0):
    return sa.gen_images(N, seed, first, B)


def random_inputs(N, B, seed=3, density=0.15):
    rng = np.random.default_rng(seed)
    return (rng.random((B, N)) < density).astype(np.float32)


def csr_inputs_to_ell(row_ptr, cols, vals, N):
    """Hidden layers given as CSR over input neurons (row i lists the outputs input i feeds,
    absolute offsets) -> per-output-row lists (L, N, K) in ascending input order, padded with
    value-0 entries that repeat the row's last input."""
    L = row_ptr.shape[0]
    per = [[[] for _ in range(N)] for _ in range(L)]
    for l in range(L):
        for i in range(N):
            for e in range(row_ptr[l, i], row_ptr[l, i + 1]):
                per[l][cols[e]].append((i, vals[e]))
    K = max(1, max(len(r) for layer in per for r in layer))
    ec = np.zeros((L, N, K), np.int32)
    ev = np.zeros((L, N, K), np.float32)
    for l in range(L):
        for o in range(N):
            r = per[l][o]
            for k in range(K):
                if k < len(r):
                    ec[l, o, k], ev[l, o, k] = r[k]
                else:
                    ec[l, o, k] = r[-1][0] if r else 0
    return ec, ev


# The reference's beyond nets (main/beyond.cu:57-100): hidden size, hidden layers, density,
# default threshold, and whether the inputs are CIFAR-10 features fed to the hidden layers.
BEYOND_NETS = {"A": dict(N=128, L=18, density=0.6, t=8, cifar=False),
               "B": dict(N=256, L=18, density=0.6, t=8, cifar=False),
               "C": dict(N=256, L=12, density=0.5, t=6, cifar=False),
               "D": dict(N=256, L=12, density=0.5, t=6, cifar=True)}


def beyond_net(key):
    """The reference's beyond net `key` from the committed fixture (tests/golden/
    beyond_net{key}.npz, written from its TSV files by make_beyond_fixture.py): row_ptr
    (L, N+1) absolute offsets of a CSR over input neurons, cols (int32), vals, bias (L, N),
    w_out (N, 10) input-major, b_out; net A also w_in (784, N) input-major and b_in."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", f"beyond_net{key}.npz"))
    out = {k: d[k] for k in d.files}
    out["cols"] = out["cols"].astype(np.int32)
    return out


def beyond_net_a():
    """The reference's beyond net A (784 -> 128 x 18 -> 10), from the committed fixture."""
    return beyond_net("A")


def prototype_features(Nin, B, seed):
    """Samples from 10 seeded U[0,1) prototypes, half with N(0, 0.1) noise on ~5% of features."""
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


def crafted_cluster_case(N=1024, seed=0):
    """Threshold-layer input (use t = 0, so Y_t is the input itself) whose exact-clustering
    result is fixed by construction (16 segments of N/16 neurons, eps 0.03; every value a
    multiple of 1/64, so every segment sum is exact in FP32).

    Rows 1 and 13 are all zero (dropped before pooling).  In the pool (the first 32 live rows):
    2 = row 0 + 1/64 at one neuron and 4 = row 0 with two values of one segment swapped (both
    pruned by 0); 3 = row 0 + 3/64 (kept: one sum differs by more than eps); 6 = a copy of 5;
    8 = 7 + 1/64 (pruned by 7) and 9 = 7 + 2/64 (kept: 2/64 > eps from 7, and the row within
    eps of it, 8, was itself pruned, so it prunes nothing); 11 = row 10 - 1/64 (pruned by 10),
    12 = row 10 - 2/64 (kept); 14..33 are random.  After the pool: a copy of row 3 (distance
    0), rows equidistant from two or three centroids (the first of them wins), random rows.
    Rows 8 and 11 are also equidistant (distance 1) from the two centroids next to them.
    Returns (y0, expect) with expect = {pool, centroids, assign (per sample: centroid index or
    -1), ne (per sample)}."""
    rng = np.random.default_rng(seed)
    vals = np.array([0.0, 0.5, 1.0, 2.0], np.float32)

    def pattern():
        return rng.choice(vals, size=N, p=[0.55, 0.15, 0.15, 0.15]).astype(np.float32)

    seg = N // 16
    rows = {}
    A, Bp, Cp, D = pattern(), pattern(), pattern(), pattern()
    n1, n2, n3, n4 = 5, seg + 7, 3 * seg + 1, 9 * seg + 2
    a, b = 2 * seg, 2 * seg + 1
    A[a], A[b] = 1.0, 2.0
    rows[0] = A
    rows[1] = np.zeros(N, np.float32)
    rows[2] = A.copy()
    rows[2][n1] += 1 / 64
    rows[3] = A.copy()
    rows[3][n2] += 3 / 64
    rows[4] = A.copy()
    rows[4][a], rows[4][b] = 2.0, 1.0  # two values of one segment swapped: same sums
    rows[5] = Bp
    rows[6] = Bp.copy()
    rows[7] = Cp
    rows[8] = Cp.copy()
    rows[8][n3] += 1 / 64
    rows[9] = Cp.copy()
    rows[9][n3] += 2 / 64
    D[n4] = 1.0
    rows[10] = D
    rows[11] = D.copy()
    rows[11][n4] -= 1 / 64
    rows[12] = D.copy()
    rows[12][n4] -= 2 / 64
    rows[13] = np.zeros(N, np.float32)
    for i in range(14, 34):
        rows[i] = pattern()
    cent = [0, 3, 5, 7, 9, 10, 12] + list(range(14, 34))
    cidx = {s: c for c, s in enumerate(cent)}
    assign = {s: cidx[s] for s in cent}
    assign.update({2: cidx[0], 4: cidx[0], 6: cidx[5], 8: cidx[7], 11: cidx[10]})

    def between(P, Q, third):
        """A row at equal L0 distance from P and Q, equal to both where they agree."""
        x = P.copy()
        d = np.flatnonzero(P != Q)
        h = len(d) // 2
        x[d[:h]] = Q[d[:h]]
        if len(d) % 2:
            x[d[-1]] = third  # differs from both
        return x

    rows[34] = rows[3].copy()
    assign[34] = cidx[3]
    rows[35] = between(A, Bp, 0.25)
    assign[35] = cidx[0]
    rows[36] = between(rows[20], rows[15], 0.25)
    assign[36] = cidx[15]
    rows[37] = np.zeros(N, np.float32)
    rows[38] = between(rows[12], rows[16], 0.25)
    assign[38] = cidx[12]
    for i in range(39, 48):
        rows[i] = pattern()
    y0 = np.stack([rows[i] for i in range(48)]).astype(np.float32)
    live = [i for i in range(48) if y0[i].any()]
    # samples after the pool: nearest centroid by brute force (first minimum); the rows
    # built to be equidistant were placed so that the expected winner is the first one
    C = y0[cent]
    for i in range(39, 48):
        assign[i] = int(np.argmin((y0[i][None, :] != C).sum(axis=1)))
    assign_arr = np.full(48, -1, np.int32)
    ne = np.zeros(48, bool)
    for i, c in assign.items():
        assign_arr[i] = c
        ne[i] = i not in cidx and bool((y0[i] != C[c]).any())
    return y0, {"pool": np.array(live[:32]), "centroids": np.array(cent), "assign": assign_arr,
                "ne": ne}


def post_tolerance_case():
    """One hidden layer (N = 64, one entry per row, clip 1, bias 0) and two samples: sample 0
    is all zero and becomes the only centroid; sample 1 is one-hot at neuron 7, so its delta
    at clustering is 1 at neuron 7.  The layer's delta is then exactly w for the rows that read
    neuron 7: w = 0.05f (just above 0.05, so kept by the reference's |v| > 0.05 test in
    double, src/beyond/SNICIT/kernel.hpp:161) on rows 0 and 2, and the float just below
    0.05f (pruned) on row 1.  Rows 3.. have bias 0.5 (equal for both samples, so their
    deltas are 0), which makes the net one without dead-row skipping, as lossy clustering
    needs.  Returns (cols, vals, bias [1, N], y0, kept_rows)."""
    N = 64
    cols = np.arange(N, dtype=np.int32).reshape(1, N, 1).copy()
    vals = np.zeros((1, N, 1), np.float32)
    for o, w in ((0, np.float32(0.05)), (1, np.nextafter(np.float32(0.05), np.float32(0))),
                 (2, np.float32(0.05))):
        cols[0, o, 0] = 7
        vals[0, o, 0] = w
    y0 = np.zeros((2, N), np.float32)
    y0[1, 7] = 1.0
    bias = np.full((1, N), 0.5, np.float32)
    bias[0, :3] = 0.0
    return cols, vals, bias, y0, [0, 2]
