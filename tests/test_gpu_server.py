"""Serving loop: requests of mixed sizes over several resident nets, grouped and overlapped
on the device, must give every request the categories it gets when run alone."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


def _clustered(Nin, B, seed):
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)].copy()
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(m.sum()).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


def _alone(net, cfg, set_input, B):
    plan = sa.Plan(net, cfg, B)
    set_input(plan)
    plan.run()
    return plan.categories()


@pytest.mark.parametrize("round_ms", [0.0, 4.0])
def test_server_results_equal_alone(ctx, round_ms):
    """Greedy launch rule (round_ms 0) and fixed periods alike."""
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
    srv = sa.Server(ctx, [net0, net1, net2], [cfg0, cfg1, cfg2], slot_samples=2048,
                    slot_segments=16, max_slots_per_net=2, group_window_ms=5.0,
                    hbm_budget_bytes=8 << 30, round_ms=round_ms)
    st = srv.stats()
    assert st["slots"] == 6
    r = np.random.default_rng(11)
    reqs = []
    for i in range(60):
        net = int(r.integers(0, 3))
        B = int(np.exp(r.uniform(0, np.log(600))))
        if net == 0:
            rid = srv.submit_images(0, 11, 10000 * i, B)
            reqs.append((rid, net, B, ("img", 10000 * i)))
        elif net == 1:
            x = _clustered(Nin, B, i)
            reqs.append((srv.submit(1, x), net, B, x))
        else:
            x = nets.random_inputs(512, B, seed=i, density=0.2)
            reqs.append((srv.submit(2, x), net, B, x))
    results = {}
    for rid, net, B, inp in reqs:
        cat, info = srv.wait(rid, 60000.0)
        assert cat.shape == (B,) and info["ms_total"] >= info["ms_queue"] >= 0
        results[rid] = (cat, info)
    assert max(info["group_requests"] for _, info in results.values()) > 1
    st = srv.stats()
    assert st["requests_done"] == 60 and st["queued"] == 0
    srv.close()
    for rid, net, B, inp in reqs:
        print(f"request {rid}: net {net}, {B} samples", flush=True)
        if net == 0:
            want = _alone(net0, cfg0, lambda p: p.set_input_images(11, inp[1], B), B)
        elif net == 1:
            want = _alone(net1, cfg1, lambda p: p.set_input(inp), B)
        else:
            want = _alone(net2, cfg2, lambda p: p.set_input(inp), B)
        np.testing.assert_array_equal(results[rid][0], want)
    # the exact net's categories also equal the plain oracle's
    rid, net, B, inp = next(q for q in reqs if q[1] == 0)
    imgs = sa.gen_images(N0, 11, inp[1], B)
    _, ref, _ = orc.plain_f32(sa.gen_sdgc_cols(N0, L0, 7), np.full((L0, N0, 32), 0.0625, np.float32),
                              -0.3, 32.0, imgs, want_y=False)
    np.testing.assert_array_equal(results[rid][0], ref)


def test_server_argument_errors_and_timeout(ctx):
    net = sa.Net.generate_sdgc(ctx, 1024, 8, 5)
    srv = sa.Server(ctx, [net], [sa.default_cfg("BF")], slot_samples=256, slot_segments=4,
                    max_slots_per_net=1, hbm_budget_bytes=1 << 30)
    with pytest.raises(sa.SnicitError):
        srv.submit_images(0, 1, 0, 257)  # larger than a slot
    with pytest.raises(sa.SnicitError):
        srv.submit_images(1, 1, 0, 10)  # no such net
    with pytest.raises(sa.SnicitError):
        srv.wait(12345, 0.0)  # unknown id
    rid = srv.submit_images(0, 1, 0, 200)
    cat, info = srv.wait(rid, 60000.0)
    assert cat.shape == (200,)
    srv.close()
    with pytest.raises(sa.SnicitError):
        sa.Server(ctx, [net], [sa.default_cfg("BF")], slot_samples=256, hbm_budget_bytes=1)


@pytest.mark.parametrize("pinned_mb", [0, 2, 256])
def test_server_pinned_input_arena(ctx, pinned_mb):
    """Request inputs held in the server's pinned arena (large ones copied to the device from
    it directly, small ones packed with the others), in pageable memory when the arena is
    off or full, and arena blocks reused as groups finish: every request's categories are
    those of its run alone."""
    Nin, N1, L1, K1, Nout = 784, 1024, 6, 103, 10
    cols, vals, w_in, w_out = sa.gen_unstructured(Nin, N1, L1, K1, Nout, 5)
    net = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=1.0)
    net.set_io(w_in, np.zeros(N1, np.float32), 1.0, w_out, np.zeros(Nout, np.float32))
    cfg = sa.default_cfg("SNICIT", threshold=2, cluster=sa.CLUSTER_LOSSY)
    srv = sa.Server(ctx, [net], [cfg], slot_samples=2048, slot_segments=16, max_slots_per_net=2,
                    group_window_ms=1.0, hbm_budget_bytes=8 << 30, pinned_bytes=pinned_mb << 20)
    r = np.random.default_rng(pinned_mb + 1)
    reqs = []
    for i in range(40):
        B = int(np.exp(r.uniform(0, np.log(1500))))  # >= 335 samples: at least 1 MiB
        x = _clustered(Nin, B, 100 + i)
        reqs.append((srv.submit(0, x), B, x))
    got = {rid: srv.wait(rid, 60000.0)[0] for rid, _, _ in reqs}
    assert srv.stats()["requests_done"] == 40
    srv.close()
    for rid, B, x in reqs:
        print(f"request {rid}: {B} samples", flush=True)
        np.testing.assert_array_equal(got[rid], _alone(net, cfg, lambda p: p.set_input(x), B))
