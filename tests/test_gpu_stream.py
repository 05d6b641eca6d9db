"""Streamed weights: a net whose hidden layers stay in pinned host memory and are copied into
a plan's device slots layer by layer must give the same bits as the same net resident in
device memory, in every mode, through graph replays, batches and autotuning."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


def _run(net, cfg, B, set_input):
    plan = sa.Plan(net, cfg, B)
    set_input(plan)
    plan.run()
    return plan


def _clustered(Nin, B, seed):
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


@pytest.mark.parametrize("buffers", [2, 3])
@pytest.mark.parametrize("mode,t", [("BF", 0), ("SNICIT", 8), ("SNICIT", 0)])
def test_sdgc_streamed_equals_resident(ctx, buffers, mode, t):
    """Panel layers (SDGC-like, 40 layers); t = 0 leaves the pre-threshold stage without
    layers.  Two runs with different inputs replay the captured stage graphs."""
    N, L, B = 1024, 40, 700
    res = sa.Net.generate_sdgc(ctx, N, L, 7, bias=-0.3)
    stm = sa.Net.generate_sdgc(ctx, N, L, 7, bias=-0.3)
    stm.stream_weights(buffers)
    cfg = sa.default_cfg(mode, threshold=t)
    p_res = sa.Plan(res, cfg, B)
    p_stm = sa.Plan(stm, cfg, B)
    for first in (0, 5000):
        print(f"buffers={buffers} mode={mode} t={t} first sample {first}", flush=True)
        p_res.set_input_images(11, first, B)
        p_stm.set_input_images(11, first, B)
        p_res.run()
        p_stm.run()
        np.testing.assert_array_equal(p_stm.activations(), p_res.activations())
        np.testing.assert_array_equal(p_stm.categories(), p_res.categories())


def test_wide_rows_streamed_matches_oracle(ctx):
    """Wide-row layers (the wavefront-per-row kernel's entry stream and range table come from
    the device slot), BF and exact SNICIT, against the oracle."""
    N, L, B, t = 512, 6, 300, 2
    cols, vals = nets.random_ell(N, L, K=80, seed=6)
    y0 = nets.random_inputs(N, B, seed=4, density=0.2)
    y0[50:90] = y0[1]
    net = sa.Net.from_ell(ctx, cols, vals, -0.05)
    net.stream_weights(2)
    yL, cat, _ = orc.plain_f32(cols, vals, -0.05, 32.0, y0)
    plan = _run(net, sa.default_cfg("BF"), B, lambda p: p.set_input(y0))
    np.testing.assert_array_equal(plan.activations(), yL)
    np.testing.assert_array_equal(plan.categories(), cat)
    r = orc.snicit_exact(cols, vals, -0.05, 32.0, y0, t)
    plan = _run(net, sa.default_cfg("SNICIT", threshold=t), B, lambda p: p.set_input(y0))
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


def test_model_streamed_lossy_matches_oracle(ctx):
    """784 -> 1024 x 6 -> 10 with dense ends (resident) and streamed hidden layers, lossy
    SNICIT, against the oracle."""
    Nin, N, L, K, Nout, B, t = 784, 1024, 6, 103, 10, 700, 2
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, 3)
    rng = np.random.default_rng(3)
    b_in = (rng.standard_normal(N) * 0.05).astype(np.float32)
    b_out = (rng.standard_normal(Nout) * 0.1).astype(np.float32)
    net = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=1.0)
    net.set_io(w_in, b_in, 1.0, w_out, b_out)
    net.stream_weights(2)
    x = _clustered(Nin, B, 1)
    r = orc.model_snicit_lossy(cols, vals, 0.0, 1.0, x, t, w_in=w_in, b_in=b_in, in_clip=1.0,
                               w_out=w_out, b_out=b_out)
    cfg = sa.default_cfg("SNICIT", threshold=t, cluster=sa.CLUSTER_LOSSY)
    plan = _run(net, cfg, B, lambda p: p.set_input(x))
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LOGITS).reshape(B, Nout), r["logits"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


def test_batch_and_autotune_with_streamed_net(ctx):
    """A batch over a streamed and a resident net (one captured graph, each plan forking its
    copy stream inside it), and the autotuner on a streamed plan."""
    N, L = 1024, 30
    stm = sa.Net.generate_sdgc(ctx, N, L, 7, bias=-0.3)
    stm.stream_weights(2)
    res = sa.Net.generate_sdgc(ctx, N, L, 7, bias=-0.3)
    cfg = sa.default_cfg("SNICIT", threshold=8)
    segments = [(0, 300), (1, 200), (0, 57), (1, 64)]
    batch = sa.Batch(ctx, [stm, res], [cfg, cfg], segments)
    for i, (n, B) in enumerate(segments):
        batch.set_input_images(i, 11, 1000 * i)
    for _ in range(2):
        batch.run()
    for i, (n, B) in enumerate(segments):
        print(f"segment {i}: net {n}, {B} samples", flush=True)
        alone = _run(res, cfg, B, lambda p: p.set_input_images(11, 1000 * i, B))
        np.testing.assert_array_equal(batch.categories(i), alone.categories())
        np.testing.assert_array_equal(batch.activations(i), alone.activations())
    plan = sa.Plan(stm, cfg, 2000)
    plan.set_input_images(3, 0, 2000)
    plan.autotune()
    plan.run()
    alone = _run(res, cfg, 2000, lambda p: p.set_input_images(3, 0, 2000))
    np.testing.assert_array_equal(plan.categories(), alone.categories())


def test_stream_weights_argument_errors(ctx):
    net = sa.Net.generate_sdgc(ctx, 1024, 4, 7, bias=-0.3)
    with pytest.raises(sa.SnicitError):
        net.stream_weights(1)
    plan = sa.Plan(net, sa.default_cfg("BF"), 64)
    with pytest.raises(sa.SnicitError):
        net.stream_weights(2)  # plans exist
    plan.close()
    net.stream_weights(2)
    with pytest.raises(sa.SnicitError):
        net.stream_weights(2)  # already streamed
