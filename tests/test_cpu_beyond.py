"""CPU checks of the oracle's unstructured-model path (dense input/output layers, wide rows)
and of its lossy SNICIT restatement.  Each check is a closed form, a hand-built known answer
or an invariant that a dropped term, a wrong index or a swapped operand breaks.
"""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets


def _identity_net(N, L):
    cols = np.broadcast_to(np.arange(N, dtype=np.int32)[None, :, None], (L, N, 1)).copy()
    vals = np.ones((L, N, 1), np.float32)
    return cols, vals


def test_model_without_io_equals_plain():
    N, L, B = 256, 6, 120
    cols, vals = nets.random_ell(N, L, K=48, seed=2)
    y0 = nets.random_inputs(N, B, seed=5, density=0.3)
    yL, lg, cat = orc.model_plain(cols, vals, -0.05, 32.0, y0)
    ref, ref_cat, _ = orc.plain_f32(cols, vals, -0.05, 32.0, y0)
    assert lg is None
    np.testing.assert_array_equal(yL, ref)
    np.testing.assert_array_equal(cat, ref_cat)


def test_dense_layers_closed_form():
    """Identity hidden layers: yL is the input layer's output and the logits are the output
    layer applied to it; both against float64 numpy, the category against numpy's argmax."""
    Nin, N, Nout, B = 40, 64, 10, 50
    rng = np.random.default_rng(0)
    w_in = rng.standard_normal((N, Nin)).astype(np.float32) * 0.3
    b_in = rng.standard_normal(N).astype(np.float32) * 0.1
    w_out = rng.standard_normal((Nout, N)).astype(np.float32)
    b_out = rng.standard_normal(Nout).astype(np.float32)
    x = rng.random((B, Nin)).astype(np.float32)
    cols, vals = _identity_net(N, 2)
    yL, lg, cat = orc.model_plain(cols, vals, 0.0, 1e30, x, w_in=w_in, b_in=b_in, in_clip=0.7,
                                  w_out=w_out, b_out=b_out)
    h = np.clip(x.astype(np.float64) @ w_in.T.astype(np.float64) + b_in, 0.0, 0.7)
    np.testing.assert_allclose(yL, h, rtol=0, atol=1e-5)
    ref_lg = yL.astype(np.float64) @ w_out.T.astype(np.float64) + b_out
    np.testing.assert_allclose(lg, ref_lg, rtol=0, atol=1e-4)
    np.testing.assert_array_equal(cat, np.argmax(lg, axis=1))
    assert 0.05 < (yL > 0).mean() < 0.95 and (yL == 0.7).any()


def test_unstructured_generator():
    Nin, N, L, K, Nout = 50, 128, 3, 20, 10
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, seed=4)
    assert cols.shape == (L, N, K) and w_in.shape == (N, Nin) and w_out.shape == (Nout, N)
    # K distinct ascending inputs per row
    assert (np.diff(cols, axis=2) > 0).all() and cols.min() >= 0 and cols.max() < N
    assert abs(vals.std() - np.sqrt(2.0 / K)) < 0.1 * np.sqrt(2.0 / K)
    assert abs(w_in.std() - np.sqrt(2.0 / Nin)) < 0.1 * np.sqrt(2.0 / Nin)
    again = sa.gen_unstructured(Nin, N, L, K, Nout, seed=4)
    np.testing.assert_array_equal(again[0], cols)
    np.testing.assert_array_equal(again[1], vals)
    f = sa.gen_features(Nin, 3, 10, 5)
    np.testing.assert_array_equal(f[2:], sa.gen_features(Nin, 3, 12, 3))
    assert f.min() >= 0 and f.max() < 1


def _known_rows(N):
    a = np.zeros(N, np.float32)
    a[: N // 2] = 1.0
    b = np.zeros(N, np.float32)
    b[N // 2:] = 1.0
    r1 = a.copy()
    r1[[3, 7]] = 0.5  # 2 entries off A by 0.5 (negative deltas)
    r3 = a.copy()
    r3[N // 2 + np.arange(5)] = 0.5  # 5 entries off A by +0.5
    return np.stack([a, r1, b, r3])


@pytest.mark.parametrize("positive_only", [False, True])
def test_lossy_known_answer(positive_only):
    """Four hand-built rows of 100 neurons, pool of 3 (rows 0, 1, 2), identity layers.
    Row 1 differs from row 0 in 2 < 100*0.03 entries, so row 0 prunes it; the centroids are
    rows 0 and 2; rows 1 and 3 are assigned to row 0 with deltas of 2 and 5 entries.  Under the
    reference's positive-only test row 1 (all deltas negative) counts as empty and is
    recovered as its centroid."""
    N = 100
    x = _known_rows(N)
    cols, vals = _identity_net(N, 3)
    r = orc.model_snicit_lossy(cols, vals, 0.0, 32.0, x, 1, pool=3, positive_only=positive_only)
    np.testing.assert_array_equal(r["pool"], [0, 1, 2])
    np.testing.assert_array_equal(r["centroids"], [0, 2])
    np.testing.assert_array_equal(r["assign"], [0, 0, 1, 0])
    want = x.copy()
    if positive_only:
        assert r["nerow_at_t"] == 1
        want[1] = x[0]
    else:
        assert r["nerow_at_t"] == 2
    np.testing.assert_array_equal(r["nerow"], [r["nerow_at_t"]] * 2)
    np.testing.assert_array_equal(r["yL"], want)
    np.testing.assert_array_equal(r["cat"], [1, 1, 1, 1])


def test_lossy_post_prune_drops_small_changes():
    """A delta of 0.04 survives clustering with eps_assign 0.03 and is pruned by the first
    post layer with eps_post 0.05: the row recovers as its centroid from there on."""
    N = 100
    x = _known_rows(N)[:3].copy()
    x[1] = x[0]
    x[1, 60] = 0.04
    cols, vals = _identity_net(N, 2)
    r = orc.model_snicit_lossy(cols, vals, 0.0, 32.0, x, 0, pool=2, eps_assign=0.03,
                               eps_post=0.05)
    # pool rows 0 and 1 (y*3/2): row 1 is 1 entry off row 0 -> pruned
    np.testing.assert_array_equal(r["centroids"], [0])
    assert r["nerow_at_t"] == 2  # row 1 (0.04 > 0.03) and row 2 (far from row 0)
    np.testing.assert_array_equal(r["nerow"], [1, 1])
    np.testing.assert_array_equal(r["yL"][1], x[0])
    np.testing.assert_array_equal(r["yL"][2], x[2])


def test_lossy_zero_tolerance_recovers_plain():
    """With eps_assign = eps_post = 0 no delta entry is dropped, so the recovered activations
    equal the plain forward up to rounding (f(zc + z) - f(zc) + f(zc) versus f(z'))."""
    Nin, N, L, K, Nout, B = 64, 256, 6, 24, 10, 300
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, seed=8)
    rng = np.random.default_rng(3)
    b_in = np.zeros(N, np.float32)
    b_out = rng.standard_normal(Nout).astype(np.float32) * 0.1
    x = sa.gen_features(Nin, 5, 0, B)
    kw = dict(w_in=w_in, b_in=b_in, w_out=w_out, b_out=b_out)
    yL, lg, cat = orc.model_plain(cols, vals, 0.0, 1.0, x, **kw)
    for t in (0, 2, 5):
        r = orc.model_snicit_lossy(cols, vals, 0.0, 1.0, x, t, eps_assign=0.0, eps_post=0.0, **kw)
        assert 1 <= r["k"] <= 128
        np.testing.assert_allclose(r["yL"], yL, rtol=0, atol=2e-5)
        np.testing.assert_allclose(r["logits"], lg, rtol=0, atol=2e-4)
        assert (r["cat"] == cat).mean() > 0.98


def test_lossy_clustering_invariants():
    """On a real forward: the pool is strided, pool row 0 is always kept, surviving centroids
    are pairwise farther apart than the pruning limit, every centroid is assigned to itself
    and every other row to its first nearest centroid (L0 distance with eps_assign)."""
    Nin, N, L, K, Nout, B, t = 64, 256, 6, 24, 10, 500, 3
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, seed=8)
    b_in = np.zeros(N, np.float32)
    b_out = np.zeros(Nout, np.float32)
    x = sa.gen_features(Nin, 5, 0, B)
    kw = dict(w_in=w_in, b_in=b_in, w_out=w_out, b_out=b_out)
    # Y_t = the first t hidden layers of the plain forward
    Yt, _, _ = orc.model_plain(cols[:t], vals[:t], 0.0, 1.0, x, w_in=w_in, b_in=b_in)
    r = orc.model_snicit_lossy(cols, vals, 0.0, 1.0, x, t, **kw)
    np.testing.assert_array_equal(r["pool"], np.arange(128) * B // 128)
    cent = r["centroids"]
    assert cent[0] == 0 and np.isin(cent, r["pool"]).all() and len(cent) == r["k"] >= 2
    C = Yt[cent]
    for a in range(len(cent)):
        for b in range(a + 1, len(cent)):
            assert (np.abs(C[a] - C[b]) > 0.03).sum() >= N * 0.03
    dist = (np.abs(Yt[:, None, :] - C[None, :, :]) > 0.04).sum(axis=2)
    want = np.argmin(dist, axis=1)
    want[cent] = np.arange(len(cent))
    np.testing.assert_array_equal(r["assign"], want)
    nonempty = ((np.abs(Yt - C[want]) > 0.04).any(axis=1))
    nonempty[cent] = False
    assert r["nerow_at_t"] == nonempty.sum()


def test_real_net_a_formats_against_float64():
    """The reference's net A files (fixture): CSR over input neurons and input-major dense
    layers.  The oracle's model forward on the transposed lists agrees with a float64 dense
    forward built straight from the file arrays (y[k] = sum_i x[i] W[i, k])."""
    a = nets.beyond_net_a()
    N, L = 128, 18
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], N)
    x = nets.prototype_features(784, 200, 3)
    yL, lg, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, w_in=a["w_in"].T.copy(),
                                  b_in=a["b_in"], in_clip=1.0, w_out=a["w_out"].T.copy(),
                                  b_out=a["b_out"])
    h = np.clip(x.astype(np.float64) @ a["w_in"].astype(np.float64) + a["b_in"], 0, 1)
    for l in range(L):
        W = np.zeros((N, N))
        rp = a["row_ptr"][l]
        for i in range(N):
            for e in range(rp[i], rp[i + 1]):
                W[i, a["cols"][e]] += a["vals"][e]
        h = np.clip(h @ W + a["bias"][l], 0, 1)
    ref_lg = h @ a["w_out"].astype(np.float64) + a["b_out"]
    assert 0.05 < (yL > 0).mean() < 0.95
    np.testing.assert_allclose(yL, h, rtol=0, atol=1e-4)
    np.testing.assert_allclose(lg, ref_lg, rtol=0, atol=1e-3)
    assert (cat == np.argmax(ref_lg, axis=1)).mean() > 0.99


def test_post_tolerance_is_the_decimal_value():
    """A post-layer delta of exactly 0.05f is kept (0.05f > 0.05 in double, as the reference
    compares) and the float below it is pruned."""
    cols, vals, bias, y0, kept = nets.post_tolerance_case()
    r = orc.model_snicit_lossy(cols, vals, bias, 1.0, y0, 0)
    assert r["k"] == 1 and r["nerow_at_t"] == 1
    expect = np.where(np.arange(y0.shape[1]) >= 3, np.float32(0.5), np.float32(0))
    np.testing.assert_array_equal(r["yL"][0], expect)  # the centroid
    expect[kept] = np.float32(0.05)
    np.testing.assert_array_equal(r["yL"][1], expect)


@pytest.mark.parametrize("key", ["A", "B", "C", "D"])
def test_reference_beyond_fixtures(key):
    """The fixtures of the reference's beyond nets have the shapes and densities of its net
    table (main/beyond.cu:57-100), and the oracle's forward over them (hidden layers as CSR
    over inputs, dense output layer, argmax) agrees with a float64 dense numpy forward."""
    spec = nets.BEYOND_NETS[key]
    a = nets.beyond_net(key)
    N, L = spec["N"], spec["L"]
    rp = a["row_ptr"]
    assert rp.shape == (L, N + 1) and a["bias"].shape == (L, N) and a["w_out"].shape == (N, 10)
    assert (np.diff(rp, axis=1) >= 0).all() and rp[0, 0] == 0
    assert (rp[1:, 0] == rp[:-1, -1]).all() and rp[-1, -1] == len(a["vals"])
    dens = (rp[:, -1] - rp[:, 0]) / (N * N)
    assert abs(dens.mean() - spec["density"]) < 1e-3
    assert ("w_in" in a) == (key == "A") and ("b_in" in a) == (not spec["cifar"])
    rng = np.random.default_rng(7)
    x = ((rng.random((40, N)) < 0.3) * rng.random((40, N))).astype(np.float32)
    cols, vals = nets.csr_inputs_to_ell(rp, a["cols"], a["vals"], N)
    yL, lg, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, w_out=a["w_out"].T.copy(),
                                  b_out=a["b_out"])
    y = x.astype(np.float64)
    for h in range(L):
        W = np.zeros((N, N))  # W[i, o]: input i feeds output o
        for i in range(N):
            for e in range(rp[h, i], rp[h, i + 1]):
                W[i, a["cols"][e]] += a["vals"][e]
        y = np.clip(y @ W + a["bias"][h], 0, 1)
    logits = y @ a["w_out"].astype(np.float64) + a["b_out"]
    assert np.abs(yL - y).max() < 1e-4
    assert np.abs(lg - logits).max() < 1e-4
    margin = np.sort(logits, axis=1)[:, -1] - np.sort(logits, axis=1)[:, -2]
    sure = margin > 1e-3
    np.testing.assert_array_equal(cat[sure], logits.argmax(axis=1)[sure])
