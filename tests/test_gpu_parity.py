"""GPU parity: the HIP layer loop against the CPU oracle, through the C-ABI.

FP32 results compare bit for bit (both sides use the same fixed fused-multiply-add order);
categories compare exactly; the FP64 oracle bounds the layer-local FP32 error by 1e-4.
"""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


def _run(ctx, net, cfg, y0, stop=None):
    plan = sa.Plan(net, cfg, y0.shape[0])
    plan.set_input(y0)
    st = plan.run(stop)
    return plan, st


@pytest.mark.parametrize("mode", ["BF", "SNIG", "XY", "PLAIN_DENSE"])
def test_plain_sdgc_bitexact(ctx, mode):
    N, L, B = 1024, 120, 700
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    plan, st = _run(ctx, net, sa.default_cfg(mode), y0)
    yL, cat, live = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    np.testing.assert_array_equal(plan.categories(), cat)
    np.testing.assert_array_equal(plan.activations(), yL)
    assert st.active == cat.sum()
    if mode != "PLAIN_DENSE":
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE), live)


def test_generated_net_equals_ell_net(ctx):
    N, L, B = 4096, 24, 300
    cols, vals = nets.sdgc(N, L, seed=5)
    y0 = nets.images(N, B, seed=2)
    a = sa.Net.from_ell(ctx, cols, vals, -0.35)
    g = sa.Net.generate_sdgc(ctx, N, L, 5, bias=-0.35)
    assert a.panels == g.panels == L * N // 32
    pa, _ = _run(ctx, a, sa.default_cfg("BF"), y0)
    pg, _ = _run(ctx, g, sa.default_cfg("BF"), y0)
    np.testing.assert_array_equal(pa.activations(), pg.activations())


def test_device_images_equal_host(ctx):
    N, B = 16384, 257
    net = sa.Net.generate_sdgc(ctx, N, 2, 1)
    plan = sa.Plan(net, sa.default_cfg("BF"), B)
    plan.set_input_images(99, 1000, B)
    st = plan.run(stop_layer=0)
    host = sa.gen_images(N, 99, 1000, B)
    np.testing.assert_array_equal(plan.activations(), host)


@pytest.mark.parametrize("bias", [-0.2, "per_neuron_mixed"])
def test_plain_random_ell_bitexact(ctx, bias):
    N, L, B = 512, 12, 300
    cols, vals = nets.random_ell(N, L, K=32, seed=1)
    rng = np.random.default_rng(4)
    y0 = rng.random((B, N)).astype(np.float32) * (rng.random((B, N)) < 0.3)
    if bias == "per_neuron_mixed":
        b = (rng.standard_normal((L, N)) * 0.05).astype(np.float32)
        clip = 1.0
    else:
        b = bias
        clip = 32.0
    net = sa.Net.from_ell(ctx, cols, vals, b, clip=clip)
    assert net.panels == L * N  # no shared inputs: one row per panel
    for mode in ("BF", "PLAIN_DENSE"):
        plan, _ = _run(ctx, net, sa.default_cfg(mode), y0)
        yL, cat, _ = orc.plain_f32(cols, vals, b, clip, y0)
        np.testing.assert_array_equal(plan.activations(), yL)
        np.testing.assert_array_equal(plan.categories(), cat)


def test_short_rows_padded(ctx):
    N, L, B = 256, 6, 130
    cols, vals = nets.random_ell(N, L, K=20, seed=8)
    y0 = nets.random_inputs(N, B, seed=1, density=0.4)
    net = sa.Net.from_ell(ctx, cols, vals, -0.05)
    plan, _ = _run(ctx, net, sa.default_cfg("BF"), y0)
    yL, cat, _ = orc.plain_f32(cols, vals, -0.05, 32.0, y0)
    np.testing.assert_array_equal(plan.activations(), yL)


def test_layer_local_fp64(ctx):
    """Per-layer FP32 error against FP64 fed the same FP32 input: <= 1e-4."""
    N, L, B = 1024, 40, 400
    cols, vals = nets.random_ell(N, L, K=32, seed=2, share=16)
    y0 = nets.images(N, B, seed=5)
    net = sa.Net.from_ell(ctx, cols, vals, -0.01)
    plan = sa.Plan(net, sa.default_cfg("BF"), B)
    plan.set_input(y0)
    prev = y0
    for l in (1, 2, 5, 17, 40):
        plan.run(stop_layer=l - 1)
        yin = plan.activations()
        plan.run(stop_layer=l)
        yout = plan.activations()
        ref = orc.layer_f64(cols[l - 1], vals[l - 1], -0.01, 32.0, yin)
        assert np.abs(yout - ref).max() <= 1e-4, l
        prev = yout


@pytest.mark.parametrize("t", [0, 1, 3, 8, 30])
def test_snicit_sdgc_matches_oracle(ctx, t):
    N, L, B = 1024, 60, 800
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    plan, st = _run(ctx, net, sa.default_cfg("SNICIT", threshold=t), y0)
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, t)
    _, cat_plain, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0, want_y=False)
    np.testing.assert_array_equal(r["cat"], cat_plain)
    np.testing.assert_array_equal(plan.categories(), r["cat"])
    assert st.live_at_threshold == r["live_at_t"]
    assert st.centroids == r["k"]
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_POOL), r["pool"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_F).reshape(-1, 16), r["F"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])


@pytest.mark.parametrize("t", [2, 9])
def test_panels_with_repeated_weight_rows(ctx, t):
    """Panels whose rows repeat other rows' weights, not next to each other in row order
    (rows 8g..8g+7 share inputs; rows 8g+5 and 8g+7 copy row 8g+1, row 8g+6 copies 8g+3, so
    the packer groups them and the post kernel runs one chain per distinct row), with a
    per-neuron bias that differs between the copies: exact SNICIT equals the oracle."""
    N, L, B = 512, 24, 600
    cols, vals = nets.random_ell(N, L, K=32, seed=21, share=8)
    v = vals.reshape(L, N // 8, 8, 32)
    v[:, :, 5] = v[:, :, 1]
    v[:, :, 7] = v[:, :, 1]
    v[:, :, 6] = v[:, :, 3]
    bias = (np.random.default_rng(2).standard_normal((L, N)) * 0.02 - 0.03).astype(np.float32)
    bias = np.minimum(bias, 0)
    y0 = nets.random_inputs(N, B, seed=4, density=0.3)
    y0[300:400] = y0[7]
    net = sa.Net.from_ell(ctx, cols, vals, bias)
    plan, st = _run(ctx, net, sa.default_cfg("SNICIT", threshold=t), y0)
    r = orc.snicit_exact(cols, vals, bias, 32.0, y0, t)
    assert r["nerow_at_t"] > 0 and r["nerow"][-1] > 0
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


@pytest.mark.parametrize("mode,t", [("BF", 0), ("SNICIT", 4)])
def test_compaction_over_many_tiles(ctx, mode, t):
    """50000 columns (13 tiles of 4096): the live-column compaction runs as per-tile counts
    and per-tile placement while more than 16384 columns live, and in one workgroup below;
    the live counts per layer, activations and categories equal the oracle's."""
    N, L, B = 1024, 12, 50000
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    plan, st = _run(ctx, net, sa.default_cfg(mode, threshold=t), y0)
    if mode == "BF":
        yL, cat, live = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
        assert live[0] > 16384 > live[-1]
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE), live)
    else:
        r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, t)
        yL, cat = r["yL"], r["cat"]
        assert r["live"][0] > 16384 and st.live_at_threshold == r["live_at_t"]
    np.testing.assert_array_equal(plan.activations(), yL)
    np.testing.assert_array_equal(plan.categories(), cat)


@pytest.mark.parametrize("t", [1, 2, 4])
def test_snicit_random_net_delta_rows(ctx, t):
    """Unconverged nets: many non-empty delta rows through every post layer."""
    N, L, B = 512, 10, 600
    cols, vals = nets.random_ell(N, L, K=32, seed=3, share=8)
    y0 = nets.random_inputs(N, B, seed=9, density=0.2)
    # a few duplicated rows so exact-zero deltas exist too
    y0[100:140] = y0[0]
    net = sa.Net.from_ell(ctx, cols, vals, -0.02)
    plan, st = _run(ctx, net, sa.default_cfg("SNICIT", threshold=t), y0)
    r = orc.snicit_exact(cols, vals, -0.02, 32.0, y0, t)
    assert r["nerow_at_t"] > 100
    np.testing.assert_array_equal(plan.categories(), r["cat"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    _, cat_plain, _ = orc.plain_f32(cols, vals, -0.02, 32.0, y0, want_y=False)
    np.testing.assert_array_equal(plan.categories(), cat_plain)


def test_snicit_positive_bias_no_skipping(ctx):
    N, L, B = 256, 8, 200
    cols, vals = nets.random_ell(N, L, K=32, seed=5, share=4)
    rng = np.random.default_rng(1)
    b = (rng.standard_normal((L, N)) * 0.05).astype(np.float32)
    y0 = nets.random_inputs(N, B, seed=2, density=0.1)
    y0[:50] = 0.0  # all-zero rows are live when f(b) != 0
    net = sa.Net.from_ell(ctx, cols, vals, b, clip=1.0)
    plan, _ = _run(ctx, net, sa.default_cfg("SNICIT", threshold=3), y0)
    r = orc.snicit_exact(cols, vals, b, 1.0, y0, 3)
    np.testing.assert_array_equal(plan.categories(), r["cat"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])


def test_edge_cases(ctx):
    N, L = 1024, 20
    cols, vals = nets.sdgc(N, L, seed=7)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    # all-dead batch: every category 0, k = 0
    y0 = np.zeros((70, N), np.float32)
    plan, st = _run(ctx, net, sa.default_cfg("SNICIT", threshold=5), y0)
    assert plan.categories().sum() == 0 and st.centroids == 0
    # fewer live rows than the pool
    y0 = nets.images(N, 90, seed=11)
    y0[5:] = 0.0
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, 5)
    plan, st = _run(ctx, net, sa.default_cfg("SNICIT", threshold=5), y0)
    np.testing.assert_array_equal(plan.categories(), r["cat"])
    assert st.live_at_threshold == r["live_at_t"] <= 5
    # one sample, ragged (not a multiple of 64) batches
    for B in (1, 63, 65, 129):
        y0 = nets.images(N, B, seed=3)
        plan, _ = _run(ctx, net, sa.default_cfg("BF"), y0)
        yL, cat, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
        np.testing.assert_array_equal(plan.activations(), yL)
    # empty batch
    plan = sa.Plan(net, sa.default_cfg("SNICIT", threshold=3), 64)
    plan.set_input(np.zeros((0, N), np.float32))
    st = plan.run()
    assert st.batch == 0


def test_invalid_arguments(ctx):
    N, L = 1024, 10
    net = sa.Net.generate_sdgc(ctx, N, L, 1)
    with pytest.raises(sa.SnicitError):
        sa.Plan(net, sa.default_cfg("SNICIT", threshold=L), 64)
    with pytest.raises(sa.SnicitError):
        sa.Plan(net, sa.default_cfg("SNICIT", threshold=-1), 64)
    plan = sa.Plan(net, sa.default_cfg("BF"), 64)
    with pytest.raises(sa.SnicitError):
        plan.run()  # no input yet
    with pytest.raises(sa.SnicitError):
        plan.set_input(np.zeros((65, N), np.float32))
    # non-finite inputs are refused, and the plan needs a new input before it runs
    for bad in (np.inf, -np.inf, np.nan):
        y = np.zeros((64, N), np.float32)
        y[37, 1000] = bad
        with pytest.raises(sa.SnicitError):
            plan.set_input(y)
        with pytest.raises(sa.SnicitError):
            plan.run()
    plan.set_input(np.zeros((64, N), np.float32))
    plan.run()


def test_graph_replay_deterministic(ctx):
    """Captured graphs replayed several times (with eager work between) equal eager runs."""
    N, L, B = 4096, 40, 500
    net = sa.Net.generate_sdgc(ctx, N, L, 9, bias=-0.35)
    eager = sa.Plan(net, sa.default_cfg("SNICIT", threshold=6, use_graph=0), B)
    graph = sa.Plan(net, sa.default_cfg("SNICIT", threshold=6, use_graph=1), B)
    eager.set_input_images(4, 0, B)
    graph.set_input_images(4, 0, B)
    eager.run()
    ref_cat, ref_y = eager.categories(), eager.activations()
    for i in range(3):
        graph.run()
        np.testing.assert_array_equal(graph.categories(), ref_cat)
        np.testing.assert_array_equal(graph.activations(), ref_y)
        eager.run()
    # a smaller batch through the same graphs
    graph.set_input_images(4, 0, B // 2)
    eager.set_input_images(4, 0, B // 2)
    graph.run()
    eager.run()
    np.testing.assert_array_equal(graph.categories(), eager.categories())


def test_infer_rounds(ctx):
    """snicit_infer in rounds of `batch` equals one full-batch plan (SNICIT::infer rounds)."""
    N, L, B = 1024, 30, 600
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=12)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    cat_r = sa.SNICIT(net, threshold=10).infer(y0, batch=200)
    _, cat, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0, want_y=False)
    np.testing.assert_array_equal(cat_r, cat)
    for cls in (sa.BF, sa.SNIG, sa.XY):
        np.testing.assert_array_equal(cls(net).infer(y0, batch=300), cat)


@pytest.mark.parametrize("K,share", [(8, 1), (12, 3), (5, 2)])
def test_merged_short_row_panels(ctx, K, share):
    """Rows with fewer than 32 inputs share panels when their input sets fit 32 slots
    (inspector-style grouping): fewer panels, and BF and exact SNICIT equal the oracle bit for
    bit under every grouping, as does the layout without merging.  One row lists an input
    twice; its group keeps its own panel."""
    N, L, B = 1024, 8, 700
    cols, vals = nets.random_ell(N, L, K=K, seed=13, share=share)
    cols[2, 5, 1] = cols[2, 5, 0]
    y0 = nets.random_inputs(N, B, seed=4, density=0.3)
    yL, _, _ = orc.plain_f32(cols, vals, -0.01, 32.0, y0)
    r = orc.snicit_exact(cols, vals, -0.01, 32.0, y0, 3)
    panels = {}
    try:
        for grouping in (ctx.GROUP_SEARCH, ctx.GROUP_CONSECUTIVE, ctx.GROUP_NONE):
            print(f"K={K} share={share} grouping={grouping}", flush=True)
            ctx.set_grouping(grouping)
            net = sa.Net.from_ell(ctx, cols, vals, -0.01)
            panels[grouping] = net.panels
            for mode, t in (("BF", 0), ("SNICIT", 3)):
                plan, _ = _run(ctx, net, sa.default_cfg(mode, threshold=t), y0)
                np.testing.assert_array_equal(plan.activations(), yL if mode == "BF" else r["yL"])
    finally:
        ctx.set_grouping(ctx.GROUP_SEARCH)
    assert panels[ctx.GROUP_SEARCH] <= panels[ctx.GROUP_CONSECUTIVE] < panels[ctx.GROUP_NONE]


@pytest.mark.parametrize("K", [6, 12])
def test_search_grouping_finds_overlaps_at_any_distance(ctx, K):
    """Banded short rows (row o reads inputs o .. o+K-1, so neighbouring rows share most
    inputs) listed in a shuffled order: consecutive grouping finds few merges, the search
    packs up to 33-K rows per panel; both equal the oracle bit for bit."""
    N, L, B = 1024, 6, 500
    rng = np.random.default_rng(21)
    cols = np.empty((L, N, K), np.int32)
    for l in range(L):
        band = (np.arange(N)[:, None] + np.arange(K)[None, :]) % N
        cols[l] = band[rng.permutation(N)]
    vals = (rng.standard_normal((L, N, K)) * 0.3).astype(np.float32)
    y0 = nets.random_inputs(N, B, seed=8, density=0.3)
    yL, _, _ = orc.plain_f32(cols, vals, -0.01, 32.0, y0)
    panels = {}
    try:
        for grouping in (ctx.GROUP_SEARCH, ctx.GROUP_CONSECUTIVE):
            ctx.set_grouping(grouping)
            net = sa.Net.from_ell(ctx, cols, vals, -0.01)
            panels[grouping] = net.panels
            plan, _ = _run(ctx, net, sa.default_cfg("BF"), y0)
            np.testing.assert_array_equal(plan.activations(), yL)
    finally:
        ctx.set_grouping(ctx.GROUP_SEARCH)
    print(f"K={K} panels: search {panels[ctx.GROUP_SEARCH]} consecutive "
          f"{panels[ctx.GROUP_CONSECUTIVE]}", flush=True)
    assert panels[ctx.GROUP_SEARCH] * 3 < panels[ctx.GROUP_CONSECUTIVE]


@pytest.mark.parametrize("share", [1, 13, 16, 32])
def test_matrix_core_and_vector_kernels_agree(ctx, share, monkeypatch):
    """The matrix-core layer kernels (K=1 fused steps: the default plain-layer kernel with
    weights in LDS, and the slot-major one) and the vector-ALU kernel give the same bits, and
    all equal the oracle."""
    N, L, B = 1024, 8, 700
    cols, vals = nets.random_ell(N, L, K=32, seed=11, share=share)
    y0 = nets.random_inputs(N, B, seed=4, density=0.25)
    outs = {}
    for kind in ("auto", "valu", "mfma"):
        monkeypatch.setenv("SNICIT_KERNEL", kind)
        net = sa.Net.from_ell(ctx, cols, vals, -0.01)
        for mode, t in (("BF", 0), ("SNICIT", 3)):
            plan, _ = _run(ctx, net, sa.default_cfg(mode, threshold=t), y0)
            outs[(kind, mode)] = plan.activations()
    yL, _, _ = orc.plain_f32(cols, vals, -0.01, 32.0, y0)
    r = orc.snicit_exact(cols, vals, -0.01, 32.0, y0, 3)
    for kind in ("auto", "valu", "mfma"):
        np.testing.assert_array_equal(outs[(kind, "BF")], yL)
        np.testing.assert_array_equal(outs[(kind, "SNICIT")], r["yL"])
