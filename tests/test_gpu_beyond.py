"""GPU parity of the unstructured-model path: wide rows (ELL layers), dense input and output
layers with argmax categories, and lossy SNICIT, against the CPU oracle through the C-ABI.
All FP32 results compare bit for bit."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


def _run(net, cfg, x):
    plan = sa.Plan(net, cfg, x.shape[0])
    plan.set_input(x)
    st = plan.run()
    return plan, st


def _model(Nin, N, L, K, Nout, seed):
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, seed)
    rng = np.random.default_rng(seed)
    b_in = (rng.standard_normal(N) * 0.05).astype(np.float32)
    b_out = (rng.standard_normal(Nout) * 0.1).astype(np.float32)
    return dict(cols=cols, vals=vals, w_in=w_in, b_in=b_in, w_out=w_out, b_out=b_out)


def _clustered(Nin, B, seed):
    """Features drawn from 10 prototypes; half the samples get noise on ~5% of features, so
    clustering finds tens of centroids and both empty and non-empty delta rows."""
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


@pytest.mark.parametrize("K,bias", [(64, -0.2), (100, -0.2), (48, "per_neuron")])
def test_wide_rows_plain_bitexact(ctx, K, bias):
    N, L, B = 1024, 6, 300
    cols, vals = nets.random_ell(N, L, K=K, seed=K)
    y0 = nets.random_inputs(N, B, seed=1, density=0.25)
    clip = 32.0
    if bias == "per_neuron":
        bias = (np.random.default_rng(2).standard_normal((L, N)) * 0.05).astype(np.float32)
        clip = 1.0
    net = sa.Net.from_ell(ctx, cols, vals, bias, clip=clip)
    yL, cat, _ = orc.plain_f32(cols, vals, bias, clip, y0)
    assert 0 < cat.sum() < B or K == 48
    for mode in ("BF", "PLAIN_DENSE"):
        plan, _ = _run(net, sa.default_cfg(mode), y0)
        np.testing.assert_array_equal(plan.activations(), yL)
        np.testing.assert_array_equal(plan.categories(), cat)


def test_wide_rows_exact_snicit(ctx):
    N, L, B, t = 512, 10, 400, 3
    cols, vals = nets.random_ell(N, L, K=80, seed=6)
    y0 = nets.random_inputs(N, B, seed=4, density=0.2)
    y0[50:90] = y0[1]
    net = sa.Net.from_ell(ctx, cols, vals, -0.05)
    plan, st = _run(net, sa.default_cfg("SNICIT", threshold=t), y0)
    r = orc.snicit_exact(cols, vals, -0.05, 32.0, y0, t)
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


@pytest.mark.parametrize("mode", ["BF", "SNIG"])
def test_model_io_plain_bitexact(ctx, mode):
    """784 -> 1024 x 6 (103 nonzeros per row) -> 10 with argmax categories."""
    Nin, N, L, K, Nout, B = 784, 1024, 6, 103, 10, 500
    m = _model(Nin, N, L, K, Nout, seed=3)
    x = sa.gen_features(Nin, 9, 0, B)
    net = sa.Net.from_ell(ctx, m["cols"], m["vals"], 0.0, clip=1.0)
    net.set_io(m["w_in"], m["b_in"], 1.0, m["w_out"], m["b_out"])
    plan, _ = _run(net, sa.default_cfg(mode), x)
    yL, lg, cat = orc.model_plain(m["cols"], m["vals"], 0.0, 1.0, x, w_in=m["w_in"],
                                  b_in=m["b_in"], in_clip=1.0, w_out=m["w_out"],
                                  b_out=m["b_out"])
    assert len(np.unique(cat)) > 3
    np.testing.assert_array_equal(plan.activations(), yL)
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, Nout), lg)
    np.testing.assert_array_equal(plan.categories(), cat)


@pytest.mark.parametrize("B,cap,N", [(1, 1, 256), (64, 4096, 256), (65, 65, 256),
                                     (777, 3000, 256), (2303, 2400, 256), (2304, 2304, 256),
                                     (8200, 8200, 256), (777, 3000, 160), (2304, 2304, 160)])
def test_dense_input_layer_forms(ctx, B, cap, N):
    """Dense layers on the matrix cores: sample counts that take the (row, column) chains
    (<= 64 columns), 32 x 64 wavefront tiles (< 2304 columns) and 64 x 128 tiles, each with a
    partial last tile, in plans whose capacity launches one tile form or both; 300 inputs pad
    to 304, not a multiple of 32; 160 neurons pad to 192 rows (not whole 64-row tiles, so
    only the 32 x 64 form runs); the 10-row output layer runs on one zero-padded 32-row
    tile.  The vector path (ell_chunked = 1) gives the same bits."""
    Nin, L, K, Nout = 300, 1, 40, 10
    m = _model(Nin, N, L, K, Nout, seed=11)
    x = sa.gen_features(Nin, 5, 0, B)
    net = sa.Net.from_ell(ctx, m["cols"], m["vals"], 0.0, clip=1.0)
    net.set_io(m["w_in"], m["b_in"], 1.0, m["w_out"], m["b_out"])
    yL, lg, cat = orc.model_plain(m["cols"], m["vals"], 0.0, 1.0, x, w_in=m["w_in"],
                                  b_in=m["b_in"], in_clip=1.0, w_out=m["w_out"],
                                  b_out=m["b_out"])
    for chunked in (2, 1):
        print(f"B={B} capacity={cap} ell_chunked={chunked}", flush=True)
        plan = sa.Plan(net, sa.default_cfg("BF"), cap)
        plan.set_shape(16384, 64, 80, chunked)
        plan.set_input(x)
        plan.run()
        np.testing.assert_array_equal(plan.activations(), yL)
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, Nout), lg)
        np.testing.assert_array_equal(plan.categories(), cat)


@pytest.mark.parametrize("t,positive_only", [(0, 0), (2, 0), (2, 1), (5, 0)])
def test_lossy_snicit_matches_oracle(ctx, t, positive_only):
    Nin, N, L, K, Nout, B = 784, 1024, 6, 103, 10, 700
    m = _model(Nin, N, L, K, Nout, seed=3)
    x = _clustered(Nin, B, 1)
    net = sa.Net.from_ell(ctx, m["cols"], m["vals"], 0.0, clip=1.0)
    net.set_io(m["w_in"], m["b_in"], 1.0, m["w_out"], m["b_out"])
    cfg = sa.default_cfg("SNICIT", threshold=t, cluster=sa.CLUSTER_LOSSY,
                         lossy_positive_only=positive_only)
    plan, st = _run(net, cfg, x)
    r = orc.model_snicit_lossy(m["cols"], m["vals"], 0.0, 1.0, x, t, w_in=m["w_in"],
                               b_in=m["b_in"], in_clip=1.0, w_out=m["w_out"], b_out=m["b_out"],
                               positive_only=bool(positive_only))
    assert 10 < r["k"] < 128 and 0 < r["nerow_at_t"] < B
    assert st.centroids == r["k"] and st.nerow_first == r["nerow_at_t"]
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_POOL), r["pool"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, Nout),
                                  r["logits"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


def test_lossy_without_io_layers(ctx):
    """Lossy clustering on a hidden-only net whose per-neuron biases are partly positive."""
    N, L, B, t = 512, 8, 400, 2
    cols, vals = nets.random_ell(N, L, K=32, seed=12, share=4)
    bias = (np.random.default_rng(5).standard_normal((L, N)) * 0.05).astype(np.float32)
    y0 = nets.random_inputs(N, B, seed=3, density=0.2)
    y0[200:260] = y0[0]
    net = sa.Net.from_ell(ctx, cols, vals, bias, clip=1.0)
    cfg = sa.default_cfg("SNICIT", threshold=t, cluster=sa.CLUSTER_LOSSY)
    plan, st = _run(net, cfg, y0)
    r = orc.model_snicit_lossy(cols, vals, bias, 1.0, y0, t)
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


def test_lossy_graph_replays_and_rerun(ctx):
    Nin, N, L, K, Nout, B = 784, 1024, 6, 103, 10, 300
    m = _model(Nin, N, L, K, Nout, seed=4)
    net = sa.Net.from_ell(ctx, m["cols"], m["vals"], 0.0, clip=1.0)
    net.set_io(m["w_in"], m["b_in"], 1.0, m["w_out"], m["b_out"])
    cfg = sa.default_cfg("SNICIT", threshold=2, cluster=sa.CLUSTER_LOSSY)
    plan = sa.Plan(net, cfg, B)
    for seed in (0, 1, 0):
        x = _clustered(Nin, B, seed)
        plan.set_input(x)
        plan.run()
        r = orc.model_snicit_lossy(m["cols"], m["vals"], 0.0, 1.0, x, 2, w_in=m["w_in"],
                                   b_in=m["b_in"], w_out=m["w_out"], b_out=m["b_out"])
        np.testing.assert_array_equal(plan.categories(), r["cat"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])


def test_infer_rounds_with_input_layer(ctx):
    """snicit_infer steps through host samples of Nin features per round."""
    Nin, N, L, K, Nout, B, batch = 784, 1024, 4, 103, 10, 600, 200
    m = _model(Nin, N, L, K, Nout, seed=6)
    x = _clustered(Nin, B, 4)
    net = sa.Net.from_ell(ctx, m["cols"], m["vals"], 0.0, clip=1.0)
    net.set_io(m["w_in"], m["b_in"], 1.0, m["w_out"], m["b_out"])
    cat = sa.SNICIT(net, threshold=1, cluster=sa.CLUSTER_LOSSY).infer(x, batch=batch)
    for off in range(0, B, batch):
        r = orc.model_snicit_lossy(m["cols"], m["vals"], 0.0, 1.0, x[off:off + batch], 1,
                                   w_in=m["w_in"], b_in=m["b_in"], w_out=m["w_out"],
                                   b_out=m["b_out"])
        np.testing.assert_array_equal(cat[off:off + batch], r["cat"])
    cat_bf = sa.BF(net).infer(x, batch=batch)
    _, _, ref = orc.model_plain(m["cols"], m["vals"], 0.0, 1.0, x, w_in=m["w_in"], b_in=m["b_in"],
                                w_out=m["w_out"], b_out=m["b_out"])
    np.testing.assert_array_equal(cat_bf, ref)


@pytest.mark.parametrize("mode,t", [("BF", 0), ("SNICIT", 2), ("SNICIT", 8)])
def test_real_net_a_from_csr(ctx, mode, t):
    """The reference's beyond net A (fixture of its TSV files): CSR over input neurons through
    snicit_net_create_csr, input-major dense ends transposed; BF and lossy SNICIT (t = 8 is the
    reference's default for this net) equal the oracle bit for bit.  Inputs: 20% of 784 pixels
    set to U[0,1) values, which leaves hundreds of non-empty delta rows at t = 2."""
    a = nets.beyond_net_a()
    N, B = 128, 700
    r0 = np.random.default_rng(1)
    x = ((r0.random((B, 784)) < 0.2) * r0.random((B, 784))).astype(np.float32)
    net = sa.Net.from_csr(ctx, N, a["row_ptr"], a["cols"], a["vals"], a["bias"], clip=1.0,
                          rows_are_inputs=True)
    net.set_io(a["w_in"].T.copy(), a["b_in"], 1.0, a["w_out"].T.copy(), a["b_out"])
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], N)
    kw = dict(w_in=a["w_in"].T.copy(), b_in=a["b_in"], in_clip=1.0, w_out=a["w_out"].T.copy(),
              b_out=a["b_out"])
    if mode == "BF":
        plan, _ = _run(net, sa.default_cfg("BF"), x)
        yL, lg, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, **kw)
        np.testing.assert_array_equal(plan.activations(), yL)
    else:
        plan, st = _run(net, sa.default_cfg("SNICIT", threshold=t, cluster=sa.CLUSTER_LOSSY), x)
        r = orc.model_snicit_lossy(cols, vals, a["bias"], 1.0, x, t, **kw)
        assert st.centroids == r["k"]
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])
        lg, cat = r["logits"], r["cat"]
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, 10), lg)
    np.testing.assert_array_equal(plan.categories(), cat)


def test_post_tolerance_is_the_decimal_value(ctx):
    """A post-layer delta of exactly 0.05f is kept, as the reference's double comparison
    keeps it, and the float below it is pruned; equal to the oracle."""
    cols, vals, bias, y0, kept = nets.post_tolerance_case()
    net = sa.Net.from_ell(ctx, cols, vals, bias, clip=1.0)
    cfg = sa.default_cfg("SNICIT", threshold=0, cluster=sa.CLUSTER_LOSSY)
    plan, st = _run(net, cfg, y0)
    r = orc.model_snicit_lossy(cols, vals, bias, 1.0, y0, 0)
    assert st.centroids == r["k"] == 1
    y = plan.activations()
    np.testing.assert_array_equal(y, r["yL"])
    assert (y[1, kept] == np.float32(0.05)).all()
    np.testing.assert_array_equal(plan.categories(), r["cat"])


@pytest.mark.parametrize("key", ["B", "C", "D"])
@pytest.mark.parametrize("mode", ["BF", "SNICIT"])
def test_real_nets_bcd_from_csr(ctx, key, mode):
    """The reference's beyond nets B, C and D (fixtures of their TSV files) through
    snicit_net_create_csr with their dense output layers; the hidden layers take inputs of
    width N directly (net D's CIFAR-10 features are that by design; the reference tree lacks
    the input-layer weights of B and C).  BF and lossy SNICIT at the reference's default
    threshold equal the oracle bit for bit.  Inputs: clustered features in [0, 1]."""
    spec = nets.BEYOND_NETS[key]
    a = nets.beyond_net(key)
    N, B, t = spec["N"], 900, spec["t"]
    x = nets.prototype_features(N, B, 31)
    net = sa.Net.from_csr(ctx, N, a["row_ptr"], a["cols"], a["vals"], a["bias"], clip=1.0,
                          rows_are_inputs=True)
    net.set_io(None, None, 1.0, a["w_out"].T.copy(), a["b_out"])
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], N)
    kw = dict(w_out=a["w_out"].T.copy(), b_out=a["b_out"])
    if mode == "BF":
        plan, _ = _run(net, sa.default_cfg("BF"), x)
        yL, lg, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, **kw)
        np.testing.assert_array_equal(plan.activations(), yL)
    else:
        plan, st = _run(net, sa.default_cfg("SNICIT", threshold=t, cluster=sa.CLUSTER_LOSSY), x)
        r = orc.model_snicit_lossy(cols, vals, a["bias"], 1.0, x, t, **kw)
        assert st.centroids == r["k"] and r["k"] > 1
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])
        lg, cat = r["logits"], r["cat"]
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, 10), lg)
    np.testing.assert_array_equal(plan.categories(), cat)
