"""Launch-shape autotuning: any launch shape gives the same bits, and the tuner's choice is
cached and reused."""
import json

import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["BF", "SNICIT"])
def test_launch_shapes_do_not_change_results(ctx, mode):
    N, L, B = 1024, 30, 700
    cols, vals = nets.random_ell(N, L, K=32, seed=3, share=8)
    y0 = nets.random_inputs(N, B, seed=9, density=0.2)
    y0[100:140] = y0[0]
    net = sa.Net.from_ell(ctx, cols, vals, -0.02)
    cfg = sa.default_cfg(mode, threshold=4)
    ref = None
    for grid_target, narrow, post, post_narrow in [
            (16384, 64, 2048, 64), (1, 0, 1, 0), (1, 100000, 1 << 20, 100000),
            (1 << 20, 0, 7, 256), (4096, 256, 16384, 0), (16384, 0, 1, 100000)]:
        print(f"shape grid_target={grid_target} narrow={narrow} post_grid_target={post} "
              f"post_narrow={post_narrow}", flush=True)
        plan = sa.Plan(net, cfg, B)
        plan.set_shape(grid_target, narrow)
        plan.set_post_shape(post, post_narrow)
        plan.set_input(y0)
        plan.run()
        got = (plan.categories(), plan.activations())
        if ref is None:
            ref = got
            if mode == "SNICIT":
                r = orc.snicit_exact(cols, vals, -0.02, 32.0, y0, 4)
                np.testing.assert_array_equal(got[1], r["yL"])
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_lossy_ell_launch_shapes(ctx):
    Nin, N, L, K, Nout, B = 784, 1024, 4, 103, 10, 300
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, 5)
    b_in = np.zeros(N, np.float32)
    b_out = np.zeros(Nout, np.float32)
    net = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=1.0)
    net.set_io(w_in, b_in, 1.0, w_out, b_out)
    x = sa.gen_features(Nin, 2, 0, B)
    cfg = sa.default_cfg("SNICIT", threshold=1, cluster=sa.CLUSTER_LOSSY)
    r = orc.model_snicit_lossy(cols, vals, 0.0, 1.0, x, 1, w_in=w_in, b_in=b_in, w_out=w_out,
                               b_out=b_out)
    shapes = [(1, 0, 0, 0), (1, 100000, 64, 0), (65536, 128, 16, 0), (16384, 64, 32, 0),
              (7, 0, 8, 0), (16384, 64, 80, 1), (1, 0, 5, 1), (3, 64, 160, 1), (65536, 0, 40, 1),
              (16384, 64, 80, 2), (1, 0, 0, 2), (1, 100000, 80, 2), (4096, 128, 80, 2)]
    for grid_target, narrow, lds_kb, chunked in shapes:
        print(f"shape grid_target={grid_target} narrow={narrow} ell_lds_kb={lds_kb} "
              f"chunked={chunked}", flush=True)
        plan = sa.Plan(net, cfg, B)
        plan.set_shape(grid_target, narrow, lds_kb, chunked)
        plan.set_input(x)
        plan.run()
        np.testing.assert_array_equal(plan.categories(), r["cat"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])


def test_autotune_caches_and_keeps_parity(ctx, tmp_path):
    N, L, B = 4096, 12, 3000
    net = sa.Net.generate_sdgc(ctx, N, L, 5, bias=-0.35)
    cfg = sa.default_cfg("SNICIT", threshold=4)
    cache = tmp_path / "tune.json"
    plan = sa.Plan(net, cfg, B)
    plan.set_input_images(3, 0, B)
    info = plan.autotune(cache)
    print(info, flush=True)
    assert not info["from_cache"] and info["ms_tuned"] <= info["ms_default"]
    # the post phase is tuned on whole runs, timed by the post stage
    assert 0 < info["ms_post_tuned"] <= info["ms_post_default"]
    assert info["post_grid_target"] in (512, 1024, 2048, 4096, 8192, 16384)
    assert info["post_form"] in (0, 1, 2) and info["sample_size"] == 32
    # whole runs before and after; the pre-threshold stage is tuned around layer 0's choice
    # and the default grid (separate measurements: a small allowance for noise)
    assert info["ms_run_default"] > 0 and info["ms_run_tuned"] > 0
    assert 0 < info["ms_pre_tuned"] <= 1.1 * info["ms_pre_default"] + 0.05
    assert info["placements"] == 1 and info["placement"] == 0
    key = info["key"]
    assert "/D" in key and "/G-1" in key and not key.endswith("/S"), key
    entry = json.loads(cache.read_text())[key]
    assert entry["post_grid_target"] == info["post_grid_target"]
    assert entry["post_form"] == info["post_form"]
    plan.run()
    cat = plan.categories()
    imgs = sa.gen_images(N, 3, 0, B)
    _, ref, _ = orc.plain_f32(sa.gen_sdgc_cols(N, L, 5), np.full((L, N, 32), 0.0625, np.float32),
                              -0.35, 32.0, imgs, want_y=False)
    np.testing.assert_array_equal(cat, ref)
    plan2 = sa.Plan(net, cfg, B)
    plan2.set_input_images(3, 0, B)
    info2 = plan2.autotune(cache)
    assert info2["from_cache"]
    for k in ("grid_target", "narrow", "post_grid_target", "post_narrow", "post_form"):
        assert info2[k] == info[k], k
    assert abs(info2["ms_post_tuned"] - info["ms_post_tuned"]) < 1e-5
    plan2.run()
    np.testing.assert_array_equal(plan2.categories(), ref)


def test_autotune_sample_size_and_placement(ctx, tmp_path):
    """The tuner's optional phases: the exact clustering sample size (kept only when the
    categories equal those of the configured size) and fresh placements of the activation
    buffers.  The tuned plan, on whichever buffers it kept, still matches the oracle, run
    after run, and a second plan takes the sample size from the cache."""
    N, L, B = 4096, 12, 3000
    net = sa.Net.generate_sdgc(ctx, N, L, 5, bias=-0.35)
    cfg = sa.default_cfg("SNICIT", threshold=4)
    cache = tmp_path / "tune.json"
    imgs = sa.gen_images(N, 3, 0, B)
    _, ref, _ = orc.plain_f32(sa.gen_sdgc_cols(N, L, 5), np.full((L, N, 32), 0.0625, np.float32),
                              -0.35, 32.0, imgs, want_y=False)
    plan = sa.Plan(net, cfg, B)
    plan.set_input_images(3, 0, B)
    info = plan.autotune(cache, sample=True, placement=True)
    print(info, flush=True)
    assert info["key"].endswith("/S")
    assert info["sample_size"] in (8, 16, 24, 32)
    assert 1 <= info["placements"] <= 5 and 0 <= info["placement"] < info["placements"]
    assert info["ms_run_tuned"] > 0
    for _ in range(2):
        st = plan.run()
        np.testing.assert_array_equal(plan.categories(), ref)
    plan2 = sa.Plan(net, cfg, B)
    plan2.set_input_images(3, 0, B)
    info2 = plan2.autotune(cache, sample=True)
    assert info2["from_cache"] and info2["sample_size"] == info["sample_size"]
    assert info2["placements"] == 1
    plan2.run()
    np.testing.assert_array_equal(plan2.categories(), ref)
    # without the option the configured size stays, under its own cache key
    plan3 = sa.Plan(net, cfg, B)
    plan3.set_input_images(3, 0, B)
    info3 = plan3.autotune(cache)
    assert not info3["from_cache"] and info3["sample_size"] == 32
    with pytest.raises(sa.SnicitError, match="unknown autotune flags"):
        sa._check(sa.lib().snicit_plan_autotune_ex(plan3.h, None, 4, None))


def _panel_net(N, L, distinct, seed):
    """32 rows per input set (one panel each); within a panel the rows take `distinct`
    weight vectors, in runs of random length."""
    rng = np.random.default_rng(seed)
    cols = np.empty((L, N, 32), np.int32)
    vals = np.empty((L, N, 32), np.float32)
    for l in range(L):
        for g in range(0, N, 32):
            cols[l, g:g + 32] = np.sort(rng.choice(N, size=32, replace=False))
            cuts = np.sort(rng.choice(np.arange(1, 32), size=distinct - 1, replace=False))
            for d, (a, b) in enumerate(zip(np.r_[0, cuts], np.r_[cuts, 32])):
                vals[l, g + a:g + b] = (rng.standard_normal(32) * 0.25).astype(np.float32)
    return cols, vals


@pytest.mark.parametrize("distinct,per_neuron_bias", [(1, False), (2, False), (2, True),
                                                      (8, False)])
def test_post_forms_give_identical_bits(ctx, distinct, per_neuron_bias):
    """The LDS-staged post kernel, the wavefront-local one and the wavefront-local one fed by
    the centroid stream give the oracle's activations and clustering, bit for bit; 60 post
    layers cover several chain blocks and delta-list compactions."""
    N, L, B, t = 1024, 64, 600, 4
    cols, vals = _panel_net(N, L, distinct, seed=10 + distinct)
    bias = (np.random.default_rng(5).uniform(-0.05, 0.0, (L, N)).astype(np.float32)
            if per_neuron_bias else np.float32(-0.02))
    y0 = nets.random_inputs(N, B, seed=9, density=0.2)
    y0[100:160] = y0[0]
    y0[300:320] = y0[7]
    net = sa.Net.from_ell(ctx, cols, vals, bias)
    r = orc.snicit_exact(cols, vals, bias, 32.0, y0, t)
    assert max(r["nerow"]) > 64  # wide delta launches, not only the narrow path
    for form in (sa.Plan.POST_LDS, sa.Plan.POST_WAVE, sa.Plan.POST_WAVE_CHAINS):
        print(f"distinct rows {distinct} per-neuron bias {per_neuron_bias} post form {form}",
              flush=True)
        plan = sa.Plan(net, sa.default_cfg("SNICIT", threshold=t), B)
        plan.set_post_form(form)
        plan.set_input(y0)
        for _ in range(2):  # captured graphs replayed
            plan.run()
            np.testing.assert_array_equal(plan.activations(), r["yL"])
            np.testing.assert_array_equal(plan.categories(), r["cat"])
            np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])


@pytest.mark.parametrize("B", [3000, 20000])
@pytest.mark.parametrize("slack", [0, 32, 1023])
def test_compaction_slack_keeps_results(ctx, slack, B):
    """Compaction slack: a plain layer's compaction keeps the live list as it is (dead columns
    carried as zeros) when at most slack/1024 of it died; the lists entering clustering and
    after the last layer stay exact.  BF and SNICIT (t = 3, 20) give the oracle's categories,
    activations, live counts and clustering traces at every slack, 1023 keeping lists almost
    always; B = 20000 takes the multi-workgroup compaction."""
    from tests import nets

    N, L = 1024, 40
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=11)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    yL, cat, live = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    print(f"slack {slack} B {B}: BF", flush=True)
    plan = sa.Plan(net, sa.default_cfg("BF"), B)
    plan.set_compact_slack(slack)
    plan.set_input(y0)
    st = plan.run()
    np.testing.assert_array_equal(plan.categories(), cat)
    np.testing.assert_array_equal(plan.activations(), yL)
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE), live)
    # work: the lists carried at least the live columns
    assert st.rows_processed >= B + int(np.sum(live[:-1]))
    for t in (3, 20):
        print(f"slack {slack} B {B}: SNICIT t={t}", flush=True)
        r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, t)
        plan = sa.Plan(net, sa.default_cfg("SNICIT", threshold=t), B)
        plan.set_compact_slack(slack)
        plan.set_input(y0)
        st = plan.run()
        np.testing.assert_array_equal(plan.categories(), r["cat"])
        assert st.live_at_threshold == r["live_at_t"] and st.centroids == r["k"]
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE)[:t], live[:t])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_POOL), r["pool"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_F).reshape(-1, 16), r["F"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])
    with pytest.raises(sa.SnicitError):
        plan.set_compact_slack(1024)


@pytest.mark.parametrize("B", [3001, 20000])
@pytest.mark.parametrize("tiles", [1, 2, 4])
def test_plain_tiles_keep_results(ctx, tiles, B):
    """The plain vector kernel with 1, 2 or 4 neighbouring 64-column tiles per wavefront
    (runs of 256, 512 or 1024 columns per workgroup, a partial last unit at B = 3001) gives
    the oracle's categories, activations and live counts in BF, and SNICIT's at t = 3, with
    exact compaction and with lists kept almost always (slack 1023)."""
    from tests import nets

    N, L = 1024, 30
    cols, vals = nets.sdgc(N, L, seed=7)
    y0 = nets.images(N, B, seed=13)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    yL, cat, live = orc.plain_f32(cols, vals, -0.3, 32.0, y0)
    r = orc.snicit_exact(cols, vals, -0.3, 32.0, y0, 3)
    for slack in (0, 1023):
        print(f"tiles {tiles} B {B} slack {slack}", flush=True)
        plan = sa.Plan(net, sa.default_cfg("BF"), B)
        plan.set_plain_tiles(tiles)
        plan.set_compact_slack(slack)
        plan.set_input(y0)
        plan.run()
        np.testing.assert_array_equal(plan.categories(), cat)
        np.testing.assert_array_equal(plan.activations(), yL)
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE), live)
        plan = sa.Plan(net, sa.default_cfg("SNICIT", threshold=3), B)
        plan.set_plain_tiles(tiles)
        plan.set_compact_slack(slack)
        plan.set_input(y0)
        plan.run()
        np.testing.assert_array_equal(plan.categories(), r["cat"])
        np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
        np.testing.assert_array_equal(plan.activations(), r["yL"])
    with pytest.raises(sa.SnicitError):
        plan.set_plain_tiles(3)
