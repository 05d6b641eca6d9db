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


def beyond_net_a():
    """The reference's beyond net A (784 -> 128 x 18 -> 10), from the committed fixture."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "beyond_netA.npz"))
    return {k: d[k] for k in d.files}


def prototype_features(Nin, B, seed):
    """Samples from 10 seeded U[0,1) prototypes, half with N(0, 0.1) noise on ~5% of features."""
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)
