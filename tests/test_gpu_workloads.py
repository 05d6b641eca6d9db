"""GPU parity at the benchmark workloads themselves (the seeds and shapes bench.py runs):
  C1  N=1024 L=120, 6000 samples: BF and SNICIT bit-exact against the oracle in full;
  C2  N=16384 L=120, 60000 samples, t=30: SNICIT categories equal BF's; SNICIT on the first
      2000 samples bit-exact against the oracle (activations, clustering traces);
  C3  N=65536 L=1920, 60000 samples, t in {30, 60, 120}: SNICIT categories equal BF's; SNICIT
      on the first 48 samples bit-exact against the oracle;
  C4  64 uneven batches over net A (4096x480, exact) and net B (784->1024x24->10, lossy) in
      one captured graph: every segment equals its run alone, small segments equal the oracle;
  C5  the server over the 16384x120 net and the 3072->2048x16->10 net: every request equals
      its run alone.
Each test prints its steps, so a long one is never silent."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu

NET_SEED, IMG_SEED = 2023, 60000


def say(*a):
    print(*a, flush=True)


def _sdgc_host(N, L):
    cols = sa.gen_sdgc_cols(N, L, NET_SEED)
    return cols, np.full(cols.shape, 0.0625, np.float32)


def _plan(net, cfg, B, first=0):
    plan = sa.Plan(net, cfg, B)
    plan.set_input_images(IMG_SEED, first, B)
    st = plan.run()
    return plan, st


def _check_snicit_against_oracle(plan, st, r, t):
    assert st.live_at_threshold == r["live_at_t"] and st.centroids == r["k"]
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_CENTROIDS), r["centroids"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_ASSIGN), r["assign"])
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_NEROW)[t:], r["nerow"])
    np.testing.assert_array_equal(plan.activations(), r["yL"])
    np.testing.assert_array_equal(plan.categories(), r["cat"])


def test_c1_full_batch(ctx):
    N, L, B, bias, t = 1024, 120, 6000, -0.3, 30
    net = sa.Net.generate_sdgc(ctx, N, L, NET_SEED, bias=bias)
    cols, vals = _sdgc_host(N, L)
    y0 = sa.gen_images(N, IMG_SEED, 0, B)
    say("C1 oracle plain")
    yL, cat, live = orc.plain_f32(cols, vals, bias, 32.0, y0)
    assert 0 < cat.sum() < B
    plan, _ = _plan(net, sa.default_cfg("BF"), B)
    np.testing.assert_array_equal(plan.activations(), yL)
    np.testing.assert_array_equal(plan.categories(), cat)
    np.testing.assert_array_equal(plan.trace(sa.Plan.TRACE_LIVE)[:L], live)
    say("C1 oracle SNICIT")
    r = orc.snicit_exact(cols, vals, bias, 32.0, y0, t)
    plan, st = _plan(net, sa.default_cfg("SNICIT", threshold=t), B)
    _check_snicit_against_oracle(plan, st, r, t)
    np.testing.assert_array_equal(r["cat"], cat)


def test_c2_full_batch_and_prefix(ctx):
    N, L, B, bias, t, pre = 16384, 120, 60000, -0.4, 30, 2000
    net = sa.Net.generate_sdgc(ctx, N, L, NET_SEED, bias=bias)
    say("C2 full batch: SNICIT and BF")
    snic, st = _plan(net, sa.default_cfg("SNICIT", threshold=t), B)
    cat_s = snic.categories()
    del snic
    bf, _ = _plan(net, sa.default_cfg("BF"), B)
    cat_b = bf.categories()
    del bf
    np.testing.assert_array_equal(cat_s, cat_b)
    assert st.nerow_first > 0 and 0 < cat_s.sum() < B
    say(f"C2 oracle SNICIT on the first {pre} samples")
    cols, vals = _sdgc_host(N, L)
    r = orc.snicit_exact(cols, vals, bias, 32.0, sa.gen_images(N, IMG_SEED, 0, pre), t)
    np.testing.assert_array_equal(r["cat"], cat_b[:pre])
    plan, st = _plan(net, sa.default_cfg("SNICIT", threshold=t), pre)
    _check_snicit_against_oracle(plan, st, r, t)


def test_c3_full_batch_thresholds_and_prefix(ctx):
    N, L, B, bias, pre = 65536, 1920, 60000, -0.45, 48
    net = sa.Net.generate_sdgc(ctx, N, L, NET_SEED, bias=bias)
    say("C3 full batch: BF")
    bf, _ = _plan(net, sa.default_cfg("BF"), B)
    cat_b = bf.categories()
    del bf
    assert 0 < cat_b.sum() < B
    for t in (30, 60, 120):
        say(f"C3 full batch: SNICIT t={t}")
        snic, st = _plan(net, sa.default_cfg("SNICIT", threshold=t), B)
        np.testing.assert_array_equal(snic.categories(), cat_b)
        assert st.centroids >= 1
        del snic
    say(f"C3 oracle SNICIT t=30 on the first {pre} samples (net in host memory)")
    cols, vals = _sdgc_host(N, L)
    r = orc.snicit_exact(cols, vals, bias, 32.0, sa.gen_images(N, IMG_SEED, 0, pre), 30)
    del cols, vals
    np.testing.assert_array_equal(r["cat"], cat_b[:pre])
    plan, st = _plan(net, sa.default_cfg("SNICIT", threshold=30), pre)
    _check_snicit_against_oracle(plan, st, r, 30)


def test_c4_batches_equal_alone(ctx):
    A = dict(N=4096, L=480, bias=-0.35, t=30)
    Bn = dict(Nin=784, N=1024, L=24, K=103, Nout=10, clip=1.0, t=8)
    rng = np.random.default_rng(4)
    sizes = rng.integers(100, 8001, 64)
    owner = rng.integers(0, 2, 64)
    segments = [(int(o), int(b)) for o, b in zip(owner, sizes)]
    net_a = sa.Net.generate_sdgc(ctx, A["N"], A["L"], NET_SEED, bias=A["bias"])
    cols, vals, w_in, w_out = sa.gen_unstructured(Bn["Nin"], Bn["N"], Bn["L"], Bn["K"],
                                                  Bn["Nout"], NET_SEED)
    zin, zout = np.zeros(Bn["N"], np.float32), np.zeros(Bn["Nout"], np.float32)
    net_b = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=Bn["clip"])
    net_b.set_io(w_in, zin, 1.0, w_out, zout)
    cfgs = [sa.default_cfg("SNICIT", threshold=A["t"]),
            sa.default_cfg("SNICIT", threshold=Bn["t"], cluster=sa.CLUSTER_LOSSY)]
    feats = {i: nets.prototype_features(Bn["Nin"], b, 1000 + i)
             for i, (o, b) in enumerate(segments) if o == 1}
    say(f"C4 batch of {len(segments)} segments, {sum(sizes)} samples")
    batch = sa.Batch(ctx, [net_a, net_b], cfgs, segments)
    for i, (o, b) in enumerate(segments):
        if o == 0:
            batch.set_input_images(i, IMG_SEED, 100000 * i)
        else:
            batch.set_input(i, feats[i])
    batch.run()
    batch.run()  # the captured graph replayed
    cats = [batch.categories(i) for i in range(len(segments))]
    for i, (o, b) in enumerate(segments):
        plan = sa.Plan(net_a if o == 0 else net_b, cfgs[o], b)
        if o == 0:
            plan.set_input_images(IMG_SEED, 100000 * i, b)
        else:
            plan.set_input(feats[i])
        plan.run()
        np.testing.assert_array_equal(cats[i], plan.categories(), err_msg=f"segment {i}")
    say("C4 oracle on the smallest segment of each net")
    ia = min((i for i, (o, _) in enumerate(segments) if o == 0), key=lambda i: segments[i][1])
    ib = min((i for i, (o, _) in enumerate(segments) if o == 1), key=lambda i: segments[i][1])
    ca, va = _sdgc_host(A["N"], A["L"])
    ra = orc.snicit_exact(ca, va, A["bias"], 32.0,
                          sa.gen_images(A["N"], IMG_SEED, 100000 * ia, segments[ia][1]), A["t"],
                          want_y=False)
    np.testing.assert_array_equal(cats[ia], ra["cat"])
    np.testing.assert_array_equal(batch.activations(ib),
                                  orc.model_snicit_lossy(cols, vals, 0.0, Bn["clip"], feats[ib],
                                                         Bn["t"], w_in=w_in, b_in=zin,
                                                         w_out=w_out, b_out=zout)["yL"])


def test_c5_server_requests_equal_alone(ctx):
    A = dict(N=16384, L=120, bias=-0.4, t=30)
    Bn = dict(Nin=3072, N=2048, L=16, K=307, Nout=10, clip=1.0, t=8)
    net_a = sa.Net.generate_sdgc(ctx, A["N"], A["L"], NET_SEED, bias=A["bias"])
    cols, vals, w_in, w_out = sa.gen_unstructured(Bn["Nin"], Bn["N"], Bn["L"], Bn["K"],
                                                  Bn["Nout"], NET_SEED)
    net_b = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=Bn["clip"])
    net_b.set_io(w_in, np.zeros(Bn["N"], np.float32), 1.0, w_out,
                 np.zeros(Bn["Nout"], np.float32))
    del cols, vals
    cfgs = [sa.default_cfg("SNICIT", threshold=A["t"]),
            sa.default_cfg("SNICIT", threshold=Bn["t"], cluster=sa.CLUSTER_LOSSY)]
    srv = sa.Server(ctx, [net_a, net_b], cfgs, slot_samples=8192, slot_segments=64,
                    max_slots_per_net=2, group_window_ms=2.0, hbm_budget_bytes=64 << 30,
                    pinned_bytes=256 << 20)
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(40):
        n = int(rng.integers(0, 2))
        B = int(np.exp(rng.uniform(0, np.log(4097))))
        if n == 0:
            reqs.append((srv.submit_images(0, IMG_SEED, 10000 * i, B), n, B, 10000 * i))
        else:
            x = nets.prototype_features(Bn["Nin"], B, 500 + i)
            reqs.append((srv.submit(1, x), n, B, x))
    say(f"C5 {len(reqs)} requests submitted")
    got = {rid: srv.wait(rid, 120000.0) for rid, _, _, _ in reqs}
    assert srv.stats()["requests_done"] == len(reqs)
    assert max(info["group_requests"] for _, info in got.values()) > 1
    srv.close()
    for rid, n, B, inp in reqs:
        plan = sa.Plan(net_a if n == 0 else net_b, cfgs[n], B)
        if n == 0:
            plan.set_input_images(IMG_SEED, inp, B)
        else:
            plan.set_input(inp)
        plan.run()
        np.testing.assert_array_equal(got[rid][0], plan.categories(), err_msg=f"request {rid}")
