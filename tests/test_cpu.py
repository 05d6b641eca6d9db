"""CPU tests: the oracle pinned by closed forms, invariants and brute force; the host
generators; the C-ABI library loads and exports every declared symbol.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- the C-ABI boundary -------------------------------------------------------------
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "snicit.h")).read()
    declared = set(re.findall(r"\b(snicit_[a-z_0-9]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(sa.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(sa.SIGNATURES), declared ^ set(sa.SIGNATURES)


@pytest.mark.parametrize("cname,py", [("snicit_run_cfg", "RunCfg"), ("snicit_stats", "Stats"),
                                      ("snicit_tune_info", "TuneInfo"),
                                      ("snicit_segment", "Segment"),
                                      ("snicit_server_cfg", "ServerCfg"),
                                      ("snicit_request_info", "RequestInfo"),
                                      ("snicit_server_stats", "ServerStats")])
def test_ctypes_structs_follow_the_header(cname, py):
    """Every struct of the C-ABI and its ctypes mirror list the same fields in the same order
    with the same element types and counts."""
    hdr = open(os.path.join(ROOT, "include", "snicit.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), hdr, re.S).group(1)
    ctype = {"int32_t": ctypes.c_int32, "int64_t": ctypes.c_int64, "double": ctypes.c_double,
             "float": ctypes.c_float, "uint64_t": ctypes.c_uint64, "char": ctypes.c_char}
    want = []
    for decl in filter(None, (d.strip() for d in body.split(";"))):
        typ, names = decl.split(None, 1)
        for name in (n.strip() for n in names.split(",")):
            m = re.fullmatch(r"(\w+)\[(\d+)\]", name)
            want.append((m.group(1), ctype[typ] * int(m.group(2))) if m else (name, ctype[typ]))
    got = getattr(sa, py)._fields_
    assert [n for n, _ in got] == [n for n, _ in want]
    for (n, a), (_, b) in zip(got, want):
        assert ctypes.sizeof(a) == ctypes.sizeof(b), n
        assert getattr(a, "_type_", a) == getattr(b, "_type_", b), n


def test_error_path_without_device():
    # bad arguments are rejected before any device work
    rc = sa.lib().snicit_gen_sdgc_ell(1000, 2, 1, None)
    assert rc == -1
    assert b"" != sa.lib().snicit_last_error()


# ---- generators -------------------------------------------------------------------------
@pytest.mark.parametrize("N", [1024, 4096])
def test_sdgc_generator_structure(N):
    L = 12
    cols = sa.gen_sdgc_cols(N, L, 7)
    assert cols.shape == (L, N, 32)
    for l in range(L):
        c = cols[l]
        # 32 distinct ascending inputs per row
        assert (np.diff(c, axis=1) > 0).all()
        # in-degree exactly 32
        assert (np.bincount(c.ravel(), minlength=N) == 32).all()
        # 16-row groups share their inputs, made of two aligned 16-chunks
        g = c.reshape(N // 16, 16, 32)
        assert (g == g[:, :1, :]).all()
        assert (g[:, 0, 0] % 16 == 0).all() and (g[:, 0, 16] % 16 == 0).all()
        assert (g[:, 0, 15] - g[:, 0, 0] == 15).all()
    # deterministic in the seed, different across seeds
    assert (sa.gen_sdgc_cols(N, 3, 7) == cols[:3]).all()
    assert not (sa.gen_sdgc_cols(N, 3, 8) == cols[:3]).all()


def test_image_generator():
    N = 1024
    a = sa.gen_images(N, 11, 0, 400)
    assert set(np.unique(a)) <= {0.0, 1.0}
    assert 0.12 < a.mean() < 0.18  # ~15% density
    # sample s depends only on (seed, s), not on the batch it is generated in
    b = sa.gen_images(N, 11, 100, 50)
    np.testing.assert_array_equal(a[100:150], b)
    # spatially clustered: a lit pixel's right neighbour is lit far more often than chance
    img = a.reshape(-1, 32, 32)
    p_right = (img[:, :, 1:] * img[:, :, :-1]).sum() / img[:, :, :-1].sum()
    assert p_right > 0.8


# ---- oracle pins ------------------------------------------------------------------------
def test_layer0_closed_form():
    """Binary inputs and 0.0625 weights: z = 0.0625 * popcount(inputs) + b exactly."""
    N = 1024
    cols, vals = nets.sdgc(N, 1, seed=3)
    y0 = nets.images(N, 64, seed=4)
    yL, _, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    pop = y0[:, cols[0]].sum(axis=2)  # (B, N)
    expect = np.minimum(np.maximum(np.float32(0.0625) * pop.astype(np.float32) + np.float32(-0.3), 0), 32)
    np.testing.assert_array_equal(yL, expect.astype(np.float32))


def test_dead_bias_kills_everything():
    """max z = 32 inputs * 32 * 0.0625 = 64, so b <= -64 zeroes every row."""
    N, L = 1024, 3
    cols, vals = nets.sdgc(N, L, seed=3)
    y0 = np.full((8, N), 32.0, np.float32)
    yL, cat, live = orc.plain_f32(cols, vals, -64.0, 32.0, y0)
    assert not yL.any() and not cat.any() and (live == 0).all()
    yL, cat, _ = orc.plain_f32(cols[:1], vals[:1], -63.9, 32.0, y0)
    assert cat.all() and np.allclose(yL, 0.1, atol=1e-5)


def test_zero_input_stays_zero():
    N, L = 256, 5
    cols, vals = nets.random_ell(N, L, seed=2)
    yL, cat, live = orc.plain_f32(cols, vals, -0.1, 32.0, np.zeros((4, N), np.float32))
    assert not yL.any() and not cat.any()


def test_permutation_net():
    """K=1, w=1, b=0, huge clip: the net is a composition of permutations."""
    N, L, B = 300, 4, 10
    rng = np.random.default_rng(0)
    perms = [rng.permutation(N) for _ in range(L)]
    cols = np.stack([p[:, None] for p in perms]).astype(np.int32)
    vals = np.ones((L, N, 1), np.float32)
    y0 = rng.random((B, N)).astype(np.float32)
    yL, _, _ = orc.plain_f32(cols, vals, 0.0, 1e30, y0)
    expect = y0
    for p in perms:
        expect = expect[:, p]
    np.testing.assert_array_equal(yL, expect)


def test_neuron_renumbering_keeps_categories():
    """The bucket-hash renumbering the SDGC command line applies with --hash-reorder
    (src/SDGC/reorder/hash.h:14-21: v = q * buckets + r -> r * (N / buckets) + q, 64 buckets
    at N = 1024) is a permutation; renumbering every weight index and input feature alike
    leaves the categories unchanged and renumbers the final activations.  Rows sum in
    ascending input-neuron order, which the renumbering changes, so the activations agree to
    FP32 rounding (1e-4 relative), not bit for bit."""
    N, L, B = 1024, 8, 300
    v = np.arange(N)
    to = v // 64 + (v % 64) * (N // 64)
    assert np.array_equal(np.sort(to), v)
    assert to[1] == 16 and to[64] == 1 and to[N - 1] == N - 1
    cols = sa.gen_sdgc_cols(N, L, 5)
    vals = np.full(cols.shape, 0.0625, np.float32)
    y0 = sa.gen_images(N, 6, 0, B)
    pc = np.empty_like(cols)
    pc[:, to] = to[cols]
    py = np.empty_like(y0)
    py[:, to] = y0
    yL, cat, _ = orc.plain_f32(cols, vals, sa.SDGC_BIAS[N], 32.0, y0)
    pyL, pcat, _ = orc.plain_f32(pc, vals, sa.SDGC_BIAS[N], 32.0, py)
    assert 0 < cat.sum() < B
    np.testing.assert_array_equal(pcat, cat)
    np.testing.assert_allclose(pyL[:, to], yL, rtol=1e-4, atol=1e-4)


def test_dense_brute_force():
    """Small net against a dense float64 matmul forward (clip and bias included)."""
    N, L, B = 128, 6, 40
    cols, vals = nets.random_ell(N, L, K=24, seed=6)
    rng = np.random.default_rng(1)
    bias = (rng.standard_normal((L, N)) * 0.1).astype(np.float32)
    y0 = rng.random((B, N)).astype(np.float32)
    yL, cat, _ = orc.plain_f32(cols, vals, bias, 1.0, y0)
    y = y0.astype(np.float64)
    for l in range(L):
        W = np.zeros((N, N))
        for o in range(N):
            np.add.at(W[o], cols[l, o], vals[l, o])
        y = np.clip(y @ W.T + bias[l], 0, 1.0)
    assert np.abs(yL - y).max() < 1e-5
    np.testing.assert_array_equal(cat, (y > 0).any(axis=1))
    yd, catd = orc.plain_f64(cols, vals, bias, 1.0, y0)
    assert np.abs(yd - y).max() < 1e-12


def test_fixed_order_fma_chain():
    """One row, FP32: the oracle's sum is the ascending-column fmaf chain, not the slot order."""
    N = 64
    cols = np.array([[[5, 1, 3]]], np.int32).repeat(N, axis=1)
    vals = np.array([[[1e8, 1.0, -1e8]]], np.float32).repeat(N, axis=1)
    y0 = np.zeros((1, N), np.float32)
    y0[0, [1, 3, 5]] = [1.0, 1.0, 1.0]
    yL, _, _ = orc.plain_f32(cols, vals, 0.0, 1e30, y0)
    # ascending columns 1, 3, 5: fma(1e8, 1, fma(-1e8, 1, fma(1, 1, 0))) = 1e8 - 1e8 (1 lost) ... = 0
    acc = np.float32(0)
    for w, x in ((1.0, 1.0), (-1e8, 1.0), (1e8, 1.0)):
        acc = np.float32(np.float64(w) * np.float64(x) + np.float64(acc))
    assert yL[0, 0] == acc == 0.0


@pytest.mark.parametrize("t", [0, 1, 2, 4, 9])
def test_snicit_exact_equals_plain(t):
    """Exact SNICIT categories equal plain categories for any centroid choice; recovered
    activations equal the plain ones within ulps (delta rounding)."""
    N, L, B = 512, 12, 300
    cols, vals = nets.random_ell(N, L, K=32, seed=3, share=8)
    y0 = nets.random_inputs(N, B, seed=9, density=0.2)
    y0[100:140] = y0[0]
    yL, cat, _ = orc.plain_f32(cols, vals, -0.02, 32.0, y0)
    r = orc.snicit_exact(cols, vals, -0.02, 32.0, y0, t)
    np.testing.assert_array_equal(r["cat"], cat)
    assert np.abs(r["yL"] - yL).max() <= 1e-3 * max(1.0, np.abs(yL).max())
    assert r["k"] >= 1 and len(r["centroids"]) == r["k"]
    # centroids are pool rows, listed in pool order
    assert set(r["centroids"]) <= set(r["pool"])
    assert list(r["centroids"]) == sorted(r["centroids"])
    # the first pool row always survives Alg. 1
    assert r["centroids"][0] == r["pool"][0]


def test_snicit_dead_rows():
    N, L = 1024, 20
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = np.zeros((50, N), np.float32)
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, 5)
    assert r["k"] == 0 and r["live_at_t"] == 0 and not r["cat"].any()
    assert (r["assign"] == -1).all()


def test_fp64_agrees_on_sdgc():
    N, L, B = 1024, 120, 200
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    y32, c32, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    y64, c64 = orc.plain_f64(cols, vals, -0.3, 32.0, y0)
    np.testing.assert_array_equal(c32, c64)
    assert np.abs(y32 - y64).max() <= 1e-4
