"""Batches: uneven segments over several nets in one captured graph.  Every segment's
results must equal a run of that batch alone (bit for bit), because clustering only sees the
segment's own rows, and the categories must equal the oracle's."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


def _alone(net, cfg, set_input, B):
    plan = sa.Plan(net, cfg, B)
    set_input(plan)
    st = plan.run()
    return plan.categories(), plan.activations(), st


def _clustered(Nin, B, seed):
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


def test_batch_segments_equal_alone(ctx):
    # net 0: SDGC-like, exact SNICIT; net 1: unstructured model, lossy SNICIT; net 2: BF
    N0, L0 = 1024, 40
    net0 = sa.Net.generate_sdgc(ctx, N0, L0, 7, bias=-0.3)
    cfg0 = sa.default_cfg("SNICIT", threshold=8)
    Nin, N1, L1, K1, Nout = 784, 1024, 6, 103, 10
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N1, L1, K1, Nout, 3)
    rng = np.random.default_rng(3)
    b_in = (rng.standard_normal(N1) * 0.05).astype(np.float32)
    b_out = (rng.standard_normal(Nout) * 0.1).astype(np.float32)
    net1 = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=1.0)
    net1.set_io(w_in, b_in, 1.0, w_out, b_out)
    cfg1 = sa.default_cfg("SNICIT", threshold=2, cluster=sa.CLUSTER_LOSSY)
    cols2, vals2 = nets.random_ell(512, 10, K=32, seed=3, share=8)
    net2 = sa.Net.from_ell(ctx, cols2, vals2, -0.02)
    cfg2 = sa.default_cfg("BF")
    segments = [(0, 300), (1, 200), (0, 57), (2, 400), (0, 1000), (1, 33), (0, 5), (1, 700),
                (0, 640), (2, 64)]
    batch = sa.Batch(ctx, [net0, net1, net2], [cfg0, cfg1, cfg2], segments)
    inputs = {}
    for i, (n, B) in enumerate(segments):
        if n == 0:
            batch.set_input_images(i, 11, 1000 * i)
        elif n == 1:
            inputs[i] = _clustered(Nin, B, i)
            batch.set_input(i, inputs[i])
        else:
            inputs[i] = nets.random_inputs(512, B, seed=i, density=0.2)
            batch.set_input(i, inputs[i])
    ms, stats = batch.run()
    assert ms > 0
    k_alone = {0: 0, 1: 0}
    for i, (n, B) in enumerate(segments):
        print(f"segment {i}: net {n}, {B} samples", flush=True)
        if n == 0:
            cat, y, st = _alone(net0, cfg0, lambda p: p.set_input_images(11, 1000 * i, B), B)
            imgs = sa.gen_images(N0, 11, 1000 * i, B)
            _, ref_cat, _ = orc.plain_f32(sa.gen_sdgc_cols(N0, L0, 7),
                                          np.full((L0, N0, 32), 0.0625, np.float32), -0.3, 32.0,
                                          imgs, want_y=False)
            np.testing.assert_array_equal(cat, ref_cat)
        elif n == 1:
            cat, y, st = _alone(net1, cfg1, lambda p: p.set_input(inputs[i]), B)
            r = orc.model_snicit_lossy(cols, vals, 0.0, 1.0, inputs[i], 2, w_in=w_in, b_in=b_in,
                                       w_out=w_out, b_out=b_out)
            np.testing.assert_array_equal(cat, r["cat"])
        else:
            cat, y, st = _alone(net2, cfg2, lambda p: p.set_input(inputs[i]), B)
        if n in k_alone:
            k_alone[n] += st.centroids
        np.testing.assert_array_equal(batch.categories(i), cat)
        np.testing.assert_array_equal(batch.activations(i), y)
    assert stats[0].centroids == k_alone[0] and stats[1].centroids == k_alone[1]
    assert stats[0].batch == 2002 and stats[1].batch == 933 and stats[2].batch == 464


def test_batch_replay_with_new_inputs(ctx):
    N, L = 1024, 30
    net = sa.Net.generate_sdgc(ctx, N, L, 5, bias=-0.3)
    cfg = sa.default_cfg("SNICIT", threshold=6)
    segments = [(0, 130), (0, 2000), (0, 77)]
    batch = sa.Batch(ctx, [net], [cfg], segments)
    for rnd in range(3):
        for i, (_, B) in enumerate(segments):
            batch.set_input_images(i, 4, 10000 * rnd + 3000 * i)
        batch.run()
        for i, (_, B) in enumerate(segments):
            cat, y, _ = _alone(net, cfg, lambda p: p.set_input_images(4, 10000 * rnd + 3000 * i, B),
                               B)
            np.testing.assert_array_equal(batch.categories(i), cat)
            np.testing.assert_array_equal(batch.activations(i), y)


def test_batch_rejects_bad_layouts(ctx):
    net = sa.Net.generate_sdgc(ctx, 1024, 4, 5)
    cfg = sa.default_cfg("BF")
    with pytest.raises(sa.SnicitError):
        sa.Batch(ctx, [net], [cfg], [(1, 10)])
    with pytest.raises(sa.SnicitError):
        sa.Batch(ctx, [net], [cfg], [(0, 0)])
    b = sa.Batch(ctx, [net], [cfg], [(0, 10), (0, 20)])
    b.set_input_images(0, 1, 0)
    with pytest.raises(sa.SnicitError):
        b.run()  # segment 1 has no input


def test_batch_autotune_refines_on_whole_batches(ctx, tmp_path):
    """The batch tuner: each plan's layer-0 shape, then pre and post shapes refined on
    whole-batch launches; results unchanged (bit for bit, against the untuned batch), the
    refined shapes in the cache, and a second batch that takes them from there."""
    import json

    N, L = 1024, 30
    net = sa.Net.generate_sdgc(ctx, N, L, 5, bias=-0.3)
    cfg = sa.default_cfg("SNICIT", threshold=6)
    cols2, vals2 = nets.random_ell(512, 10, K=32, seed=3, share=8)
    net2 = sa.Net.from_ell(ctx, cols2, vals2, -0.02)
    segments = [(0, 700), (1, 300), (0, 2000), (1, 64)]
    inputs = {i: nets.random_inputs(512, B, seed=i, density=0.2)
              for i, (n, B) in enumerate(segments) if n == 1}

    def make():
        b = sa.Batch(ctx, [net, net2], [cfg, sa.default_cfg("BF")], segments)
        for i, (n, B) in enumerate(segments):
            if n == 0:
                b.set_input_images(i, 4, 3000 * i)
            else:
                b.set_input(i, inputs[i])
        return b

    batch = make()
    batch.run()
    want = [(batch.categories(i), batch.activations(i)) for i in range(len(segments))]
    cache = tmp_path / "tune.json"
    info = batch.autotune(cache)
    print(info, flush=True)
    assert not any(t["from_cache"] for t in info)
    assert all(t["ms_run_default"] > 0 and t["ms_run_tuned"] > 0 for t in info)
    # whole-batch times: every plan reports the batch's
    assert info[0]["ms_run_tuned"] == info[1]["ms_run_tuned"]
    entries = json.loads(cache.read_text())
    for t in info:
        e = entries[t["key"]]
        for k in ("grid_target", "post_grid_target", "post_narrow", "post_form"):
            assert e[k] == t[k], k
    for _ in range(2):
        batch.run()
        for i in range(len(segments)):
            np.testing.assert_array_equal(batch.categories(i), want[i][0])
            np.testing.assert_array_equal(batch.activations(i), want[i][1])
    batch2 = make()
    info2 = batch2.autotune(cache)
    assert all(t["from_cache"] for t in info2)
    for a, b in zip(info, info2):
        for k in ("grid_target", "narrow", "post_grid_target", "post_narrow", "post_form"):
            assert a[k] == b[k], k
    batch2.run()
    for i in range(len(segments)):
        np.testing.assert_array_equal(batch2.categories(i), want[i][0])
