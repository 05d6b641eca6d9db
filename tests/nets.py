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
