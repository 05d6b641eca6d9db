"""GPU parity of the wide-row (ELL) layer kernels on shapes that reach their edge cases: an
input count that is not a multiple of the 128-neuron LDS chunk, rows of very different
lengths given as CSR (empty rows, rows whose entries all sit in one chunk, rows ending in
zero-valued entries), and batch sizes that select every rows-per-wavefront count and both
tile mappings of the wavefront-per-row kernel.  Every result is compared bit for bit with
the oracle's FP32 forward."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa

pytestmark = pytest.mark.gpu


def _ragged_csr(N, L, seed):
    """Per-layer CSR over output rows with lengths 0..300; row 0 of each layer is empty, row 1
    reads only inputs of one chunk, row 2 ends in zero-valued entries."""
    r = np.random.default_rng(seed)
    row_ptr = np.zeros((L, N + 1), np.int64)
    cols, vals = [], []
    off = 0
    for l in range(L):
        for o in range(N):
            row_ptr[l, o] = off
            n = int(r.integers(0, 301))
            if o == 0:
                n = 0
            if o == 1:
                c = r.choice(np.arange(256, 384), size=60, replace=False)
            else:
                c = r.choice(N, size=n, replace=False)
            v = (r.standard_normal(c.size) * np.sqrt(2.0 / max(c.size, 1))).astype(np.float32)
            if o == 2 and c.size:
                c = np.sort(c)
                v[-min(5, c.size):] = 0.0
            cols.append(c.astype(np.int32))
            vals.append(v)
            off += c.size
        row_ptr[l, N] = off
    return row_ptr, np.concatenate(cols), np.concatenate(vals)


def _to_ell(N, row_ptr, cols, vals):
    """Oracle form: (L, N, K) rows in ascending input order, value-0 padding."""
    L = row_ptr.shape[0]
    K = int(max(1, max(row_ptr[l, o + 1] - row_ptr[l, o] for l in range(L) for o in range(N))))
    ec = np.zeros((L, N, K), np.int32)
    ev = np.zeros((L, N, K), np.float32)
    for l in range(L):
        for o in range(N):
            a, b = row_ptr[l, o], row_ptr[l, o + 1]
            order = np.argsort(cols[a:b], kind="stable")
            c, v = cols[a:b][order], vals[a:b][order]
            ec[l, o, :c.size], ev[l, o, :c.size] = c, v
            if c.size:
                ec[l, o, c.size:] = c[-1]
    return ec, ev


@pytest.fixture(scope="module")
def ragged(ctx):
    N, L = 1008, 3  # 7 full chunks of 128 inputs and one of 112; N % 16 == 0 for SNICIT
    row_ptr, cols, vals = _ragged_csr(N, L, 4)
    net = sa.Net.from_csr(ctx, N, row_ptr, cols, vals, -0.05)
    ec, ev = _to_ell(N, row_ptr, cols, vals)
    return N, L, net, ec, ev


@pytest.mark.parametrize("B", [1, 63, 65, 520, 5000])
def test_wave_kernel_ragged_rows(ctx, ragged, B):
    N, L, net, ec, ev = ragged
    y0 = (np.random.default_rng(B).random((B, N)) < 0.3).astype(np.float32)
    yL, cat, _ = orc.plain_f32(ec, ev, -0.05, 32.0, y0)
    for mode in ("BF", "PLAIN_DENSE"):
        for chunked in (2, 1):
            print(f"B={B} mode={mode} ell_chunked={chunked}", flush=True)
            plan = sa.Plan(net, sa.default_cfg(mode), B)
            plan.set_shape(16384, 64, 80, chunked)
            plan.set_input(y0)
            plan.run()
            np.testing.assert_array_equal(plan.activations(), yL)
            np.testing.assert_array_equal(plan.categories(), cat)


@pytest.mark.parametrize("t", [1, 2])
def test_wave_kernel_snicit_epilogues(ctx, ragged, t):
    """Centroid and delta epilogues of the wavefront-per-row kernel (post-threshold layers
    of exact SNICIT on a wide-row net)."""
    N, L, net, ec, ev = ragged
    B = 900
    base = (np.random.default_rng(31).random((12, N)) < 0.3).astype(np.float32)
    y0 = base[np.random.default_rng(32).integers(0, 12, B)]
    flip = np.random.default_rng(33).random((B, N)) < 0.002
    y0 = np.where(flip, 1.0 - y0, y0).astype(np.float32)
    r = orc.snicit_exact(ec, ev, -0.05, 32.0, y0, t)
    print(f"SNICIT t={t} B={B}", flush=True)
    plan = sa.Plan(net, sa.default_cfg("SNICIT", threshold=t), B)
    plan.set_input(y0)
    plan.run()
    np.testing.assert_array_equal(plan.categories(), r["cat"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
