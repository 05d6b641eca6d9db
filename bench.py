"""Benchmark: SNICIT vs plain SpMM layer loop on one MI355X, edges/s and ms per batch.

A step is one full forward (all L layers, clustering, recovery) over one batch of
synthetic inputs already resident in HBM.  Default workload: the largest one-GPU shape,
the SDGC-like net N=65536, L=1920, 32 nnz/row of 0.0625, bias -0.45, clip 32, batch
60000 binary blob images at ~15% density, threshold layer 30 (--config c2 for N=16384,
L=120, bias -0.4).

Prints diagnostics on stderr and ONE JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "c1": dict(N=1024, L=120, B=6000, bias=-0.3, t=30),
    "c2": dict(N=16384, L=120, B=60000, bias=-0.4, t=30),
    "c3": dict(N=65536, L=1920, B=60000, bias=-0.45, t=30),
    # 64 uneven batches over two resident nets in one captured graph
    "c4": dict(batches=64, smin=100, smax=8000,
               A=dict(N=4096, L=480, bias=-0.35, t=30),
               B=dict(Nin=784, N=1024, L=24, K=103, Nout=10, clip=1.0, t=8)),
    # Poisson request stream against two resident nets
    "c5": dict(smin=1, smax=4096, loads=(0.3, 0.7, 1.0), burst=160, requests=2000,
               slots=2, window_ms=2.0, group_samples=0, pinned_mb=1024, round_ms=20.0,
               A=dict(N=16384, L=120, bias=-0.4, t=30),
               B=dict(Nin=3072, N=2048, L=16, K=307, Nout=10, clip=1.0, t=8)),
}
NET_SEED = 2023
IMG_SEED = 60000
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmm_layer_bytes(N, P, rows):
    """Algorithmic HBM bytes of one plain panel launch on an SDGC-like net: read and write
    every live activation once (rows = live columns entering the layer; dead columns a list
    carries between compactions are not counted), plus the layer in panel form (the 32 input
    ids and the one distinct weight row of each panel, panel offsets, output ids)."""
    return 8.0 * N * rows + P * 32 * 4 * 2 + (P + 1) * 4 + N * 4


def post_layer_bytes(N, P, rows):
    """Algorithmic HBM bytes of one post-threshold delta launch (panel_post_wave_kernel):
    read and write every listed delta row once and read its centroid index, plus each
    panel's 32 input ids, distinct weight row and 64-value chain record, panel offsets and
    output ids.  The next layer's centroid rows and chain records, which the same launch
    computes in a fifth wavefront (8*N*k bytes and 256 bytes per panel), are not counted."""
    return 8.0 * N * rows + 4.0 * rows + P * (32 * 4 * 2 + 64 * 4) + (P + 1) * 4 + N * 4


def measured_traffic(wl, kernel):
    """HBM bytes per launch of `kernel` on this workload from a committed counter summary
    (profiles/*_traffic.json: FETCH_SIZE and WRITE_SIZE runs, each scaled by what the
    counter reads for a copy of known size), and the kernel that summary measured: the same
    kernel, else one of its family (same name before the template arguments, e.g. another
    tile count), else (None, None)."""
    import glob
    family = kernel.split("<")[0]
    found = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
        try:
            rec = json.load(open(path))
        except (OSError, ValueError):
            continue
        w = rec.get("workload", {})
        if (w.get("N"), w.get("L"), w.get("batch"), w.get("threshold")) != (
                wl["N"], wl["L"], wl["B"], wl["t"]):
            continue
        k = rec.get("kernel", "")
        if k == kernel:
            return rec.get("traffic_bytes_per_launch"), k
        if k.split("<")[0] == family and found is None:
            found = (rec.get("traffic_bytes_per_launch"), k)
    return found or (None, None)


def live_rows(plan, B, n):
    """Live columns entering each of the first n plain layers (the lists may also carry dead
    columns between compactions; the algorithmic bytes count only the live ones)."""
    after = plan.trace(0)
    return [int(B)] + [int(after[l]) for l in range(n - 1)]


def apply_shapes(plan, path):
    """Launch shapes and post form of the single entry of an autotuner cache file, set on the
    plan directly (no measurement: a profiled process runs only the forwards it traces)."""
    entries = json.load(open(path))
    if len(entries) != 1:
        raise SystemExit(f"{path}: expected one tuned entry, found {len(entries)}")
    e = next(iter(entries.values()))
    apply_shape_dict(plan, e)
    log(f"shapes from {path}: {json.dumps(e)}")
    return e


def apply_shape_dict(plan, e):
    """Launch shapes and post form of a tuner result (cache entry or TuneInfo dict)."""
    plan.set_shape(e["grid_target"], e["narrow"], e["ell_lds_kb"], e["ell_chunked"])
    plan.set_post_shape(e["post_grid_target"], e["post_narrow"])
    plan.set_post_form(e["post_form"])
    plan.set_compact_slack(e["compact_slack"])
    plan.set_plain_tiles(e["plain_tiles"])


def time_steps(plan, steps, warmup, stages=False):
    for _ in range(warmup):
        plan.run()
    t0 = time.perf_counter()
    sts = [plan.run() for _ in range(steps)]
    wall = (time.perf_counter() - t0) * 1000.0 / steps
    ev = float(np.mean([s.ms_total for s in sts]))
    if stages:  # mean HIP-event time of the pre-threshold and post stages over the steps
        return wall, ev, sts[-1], (float(np.mean([s.ms_pre for s in sts])),
                                   float(np.mean([s.ms_post for s in sts])))
    return wall, ev, sts[-1]


def cpu_baseline(wl, gpu_cat, target_s):
    import oracle as orc
    import snicit_amd as sa

    N, L = wl["N"], wl["L"]
    cores = os.cpu_count() or 1
    cols = sa.gen_sdgc_cols(N, L, NET_SEED)
    vals = np.full(cols.shape, 0.0625, np.float32)
    # grow the sample (first n samples of the batch) until the timed loop takes about
    # target_s seconds; per-sample cost is not linear (dead rows are skipped)
    n = 4 * cores
    for _ in range(6):
        y0 = sa.gen_images(N, IMG_SEED, 0, n)
        _, cat, _, dt = orc.plain_f32(cols, vals, wl["bias"], 32.0, y0, nthreads=cores,
                                      want_y=False, timing=True)
        if dt >= 0.5 * target_s or n >= wl["B"]:
            break
        n = min(wl["B"], max(n + cores, int(n * 0.9 * target_s / max(dt, 1e-3))))
    edges = float(n) * N * 32 * L
    match = bool(np.array_equal(cat, gpu_cat[:n]))
    return {"value": edges / dt, "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of the {wl['B']} samples x all {L} layers (plain FP32 "
                      f"forward with dead-row skipping, OpenMP, {cores} threads) in {dt:.2f} s; "
                      f"edges/s extrapolated to the full batch at that rate",
            "seconds": dt, "categories_match_gpu": match, "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count()}


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def prototype_features(Nin, B, seed):
    """Samples drawn from 10 seeded prototypes of U[0,1) features; half of them get N(0, 0.1)
    noise on ~5% of the features (MNIST-like class structure for the unstructured net)."""
    import snicit_amd as sa
    r = np.random.default_rng(seed)
    proto = sa.gen_features(Nin, 77, 0, 10)
    x = proto[r.integers(0, 10, B)]
    m = (r.random((B, Nin)) < 0.05) & (r.random(B) < 0.5)[:, None]
    x[m] += r.standard_normal(int(m.sum())).astype(np.float32) * 0.1
    return np.clip(x, 0, 1).astype(np.float32)


def bench_batch(args, wl):
    """Form (a): 64 batches of seeded uneven sizes over net A (SDGC-like, exact SNICIT) and
    net B (unstructured MLP with dense ends, lossy SNICIT), one graph launch per step,
    launch shapes autotuned per net on whole-batch launches (the SNICIT batch is also timed
    with the default shapes first).  The plain (BF) batch of the same segments is timed
    beside it."""
    import snicit_amd as sa

    A, Bn = wl["A"], wl["B"]
    rng = np.random.default_rng(4)
    sizes = rng.integers(wl["smin"], wl["smax"] + 1, wl["batches"])
    owner = rng.integers(0, 2, wl["batches"])
    segments = [(int(o), int(b)) for o, b in zip(owner, sizes)]
    ctx = sa.Context(0)
    net_a = sa.Net.generate_sdgc(ctx, A["N"], A["L"], NET_SEED, bias=A["bias"])
    cols, vals, w_in, w_out = sa.gen_unstructured(Bn["Nin"], Bn["N"], Bn["L"], Bn["K"],
                                                  Bn["Nout"], NET_SEED)
    net_b = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=Bn["clip"])
    net_b.set_io(w_in, np.zeros(Bn["N"], np.float32), 1.0, w_out,
                 np.zeros(Bn["Nout"], np.float32))
    cfgs = {"SNICIT": [sa.default_cfg("SNICIT", threshold=A["t"]),
                       sa.default_cfg("SNICIT", threshold=Bn["t"], cluster=sa.CLUSTER_LOSSY)],
            "BF": [sa.default_cfg("BF"), sa.default_cfg("BF")]}
    cache = os.path.join(ROOT, "gpurun_out", "tune_cache.json")
    os.makedirs(os.path.dirname(cache), exist_ok=True)
    feats = {i: prototype_features(Bn["Nin"], b, 1000 + i) for i, (o, b) in enumerate(segments)
             if o == 1}
    res = {}
    for mode in ("SNICIT", "BF"):
        batch = sa.Batch(ctx, [net_a, net_b], cfgs[mode], segments)
        for i, (o, b) in enumerate(segments):
            if o == 0:
                batch.set_input_images(i, IMG_SEED, 100000 * i)
            else:
                batch.set_input(i, feats[i])
        untuned, tune = None, None
        if args.tune != "none":
            if mode == "SNICIT":  # the default launch shapes, for the record
                for _ in range(args.warmup):
                    batch.run()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    batch.run()
                untuned = (time.perf_counter() - t0) * 1000.0 / args.steps
            tune = batch.autotune(cache)
        for _ in range(args.warmup):
            batch.run()
        t0 = time.perf_counter()
        ev = []
        for _ in range(args.steps):
            ms, stats = batch.run()
            ev.append(ms)
        wall = (time.perf_counter() - t0) * 1000.0 / args.steps
        cats = [batch.categories(i) for i in range(len(segments))]
        res[mode] = dict(ms=wall, ev=float(np.mean(ev)), stats=stats, cats=cats, tune=tune,
                         untuned=untuned)
        log(f"{mode} batch: {wall:.3f} ms/step (graph events {np.mean(ev):.3f}); tuned "
            f"{[(t['key'], t['grid_target'], t['narrow']) for t in tune or []]}; net A k="
            f"{stats[0].centroids} nerow {stats[0].nerow_first}->{stats[0].nerow_last}; net B "
            f"k={stats[1].centroids} nerow {stats[1].nerow_first}->{stats[1].nerow_last}")
        del batch
    samples_a = sum(b for o, b in segments if o == 0)
    samples_b = sum(b for o, b in segments if o == 1)
    edges = (float(samples_a) * A["N"] * 32 * A["L"] +
             float(samples_b) * (Bn["N"] * Bn["K"] * Bn["L"] + Bn["N"] * Bn["Nin"] +
                                 Bn["Nout"] * Bn["N"]))
    same_a = all(np.array_equal(res["SNICIT"]["cats"][i], res["BF"]["cats"][i])
                 for i, (o, _) in enumerate(segments) if o == 0)
    agree_b = float(np.mean(np.concatenate([res["SNICIT"]["cats"][i] == res["BF"]["cats"][i]
                                            for i, (o, _) in enumerate(segments) if o == 1])))
    ms_s, ms_b = res["SNICIT"]["ms"], res["BF"]["ms"]
    out = {
        "metric": "batched SNICIT edges/s (sum over segments of samples x nnz(W) x layers / "
                  "time), one graph launch of all segments per step",
        "value": edges / (ms_s / 1000.0),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SDGC-like net A with blob images; seeded unstructured net B "
                "with prototype features)",
        "config": {"workload": f"{wl['batches']} batches of {wl['smin']}-{wl['smax']} samples "
                               f"over net A (N={A['N']} L={A['L']} bias={A['bias']} t={A['t']}, "
                               f"exact) and net B ({Bn['Nin']}->{Bn['N']}x{Bn['L']}->"
                               f"{Bn['Nout']}, {Bn['K']} nnz/row, clip {Bn['clip']}, t={Bn['t']}, "
                               "lossy)",
                   "segments_net_a": int((owner == 0).sum()), "samples_net_a": samples_a,
                   "segments_net_b": int((owner == 1).sum()), "samples_net_b": samples_b},
        "plain_bf_ms_per_step": ms_b,
        "plain_bf_edges_per_s": edges / (ms_b / 1000.0),
        "speedup_vs_plain_bf": ms_b / ms_s,
        "categories_identical_net_a": same_a,
        "categories_agreement_net_b_lossy": agree_b,
        "autotuned_shapes": res["SNICIT"]["tune"],
        "untuned_ms_per_step": res["SNICIT"]["untuned"],
        "roofline": None,
        "cpu_baseline": None,
    }
    return out


def short_rows(K):
    """Panel grouping on nets whose rows have K < 32 inputs (N=16384, 60 layers, 60000
    samples at density 0.3): plain BF and SNICIT (t=20) ms/step under each grouping (none,
    consecutive, search), on random rows and on banded rows listed in a shuffled order (row o
    reads o .. o+K-1); categories must agree across groupings."""
    import snicit_amd as sa

    N, L, B = 16384, 60, 60000
    rng = np.random.default_rng(5)
    random_cols = rng.integers(0, N, (L, N, K)).astype(np.int32)
    band = (np.arange(N)[:, None] + np.arange(K)[None, :]) % N
    banded_cols = np.stack([band[rng.permutation(N)] for _ in range(L)]).astype(np.int32)
    vals = (rng.standard_normal((L, N, K)) * np.sqrt(2.0 / K)).astype(np.float32)
    y0 = (rng.random((B, N)) < 0.3).astype(np.float32)
    ctx = sa.Context(0)
    names = {ctx.GROUP_NONE: "none", ctx.GROUP_CONSECUTIVE: "consecutive",
             ctx.GROUP_SEARCH: "search"}
    same = True
    for kind, cols in (("random", random_cols), ("banded, shuffled", banded_cols)):
        cats = {}
        for grouping in (ctx.GROUP_NONE, ctx.GROUP_CONSECUTIVE, ctx.GROUP_SEARCH):
            ctx.set_grouping(grouping)
            t0 = time.perf_counter()
            net = sa.Net.from_ell(ctx, cols, vals, -0.01)
            pack_s = time.perf_counter() - t0
            for mode in ("BF", "SNICIT"):
                plan = sa.Plan(net, sa.default_cfg(mode, threshold=20), B)
                plan.set_input(y0)
                ms, _, _ = time_steps(plan, 3, 1)
                cats[(grouping, mode)] = plan.categories().copy()
                log(f"K={K} {kind} rows, grouping {names[grouping]}: panels={net.panels} "
                    f"(net built in {pack_s:.1f} s) {mode}: {ms:.2f} ms/step")
                plan.close()
            net.close()
        same = same and all(np.array_equal(cats[(g, m)], cats[(ctx.GROUP_NONE, m)])
                            for g in names for m in ("BF", "SNICIT"))
    ctx.set_grouping(ctx.GROUP_SEARCH)
    log(f"categories identical under every grouping: {same}")


def stream_ceiling():
    """Device-memory rates of plain streams (best of 5): a sum reads 2 GiB, a fill writes
    2 GiB, a copy reads and writes 2 GiB each.  The ceiling the layer kernels' HBM rates are
    compared with, beside the 8 TB/s peak."""
    import torch

    n = 1 << 29
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    out = {}
    for name, fn, moved in [("read", lambda: a.sum(), 4 * n), ("write", lambda: b.fill_(2.0), 4 * n),
                            ("copy", lambda: b.copy_(a), 8 * n)]:
        best = float("inf")
        for _ in range(6):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[name] = moved / best / 1e6
        log(f"{name}: {best:.3f} ms, {out[name]:.0f} GB/s")
    print(json.dumps({"stream_GBps": out}), flush=True)


def pass_serve_net_b(args, wl):
    """One eager forward (no graph) of the serving workload's unstructured net B in the given
    mode for 1024, 4096 and 8192 samples (plan capacity 8192), for kernel traces and counters:
    the dense 3072 -> 2048 input layer, the 16 wide-row layers and the output layer."""
    import snicit_amd as sa

    Bn = wl["B"]
    ctx = sa.Context(0)
    cols, vals, w_in, w_out = sa.gen_unstructured(Bn["Nin"], Bn["N"], Bn["L"], Bn["K"],
                                                  Bn["Nout"], NET_SEED)
    net = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=Bn["clip"])
    net.set_io(w_in, np.zeros(Bn["N"], np.float32), 1.0, w_out,
               np.zeros(Bn["Nout"], np.float32))
    lossy = dict(cluster=sa.CLUSTER_LOSSY) if args.pmc_pass == "SNICIT" else {}
    cfg = sa.default_cfg(args.pmc_pass, threshold=Bn["t"], use_graph=0, **lossy)
    plan = sa.Plan(net, cfg, 8192)
    for B in (1024, 4096, 8192):
        plan.set_input(prototype_features(Bn["Nin"], B, 7))
        st = plan.run()
        log(f"net B {args.pmc_pass} pass, {B} samples: {st.ms_total:.3f} ms")


def bench_serve(args, wl):
    """Form (b): a request loop over a resident SDGC-like net (exact SNICIT) and an
    unstructured net (lossy SNICIT).  Saturated throughput comes from a burst of requests;
    then Poisson arrivals at 30/70/100% of it give p50/p99 latency (submit to results on the
    host) and the achieved edges/s."""
    # One hardware queue per slot stream (and one for the context): streams that share a
    # queue run one after another, so a small group would wait behind another slot's large
    # one.  Read when HIP initialises, i.e. before the library's first call.
    os.environ["GPU_MAX_HW_QUEUES"] = str(2 * wl["slots"] + 2)
    import snicit_amd as sa

    A, Bn = wl["A"], wl["B"]
    ctx = sa.Context(0)
    net_a = sa.Net.generate_sdgc(ctx, A["N"], A["L"], NET_SEED, bias=A["bias"])
    cols, vals, w_in, w_out = sa.gen_unstructured(Bn["Nin"], Bn["N"], Bn["L"], Bn["K"],
                                                  Bn["Nout"], NET_SEED)
    net_b = sa.Net.from_ell(ctx, cols, vals, 0.0, clip=Bn["clip"])
    net_b.set_io(w_in, np.zeros(Bn["N"], np.float32), 1.0, w_out,
                 np.zeros(Bn["Nout"], np.float32))
    del cols, vals
    cfgs = [sa.default_cfg("SNICIT", threshold=A["t"]),
            sa.default_cfg("SNICIT", threshold=Bn["t"], cluster=sa.CLUSTER_LOSSY)]
    def make_server(round_ms):
        t0 = time.perf_counter()
        s = sa.Server(ctx, [net_a, net_b], cfgs, slot_samples=8192, slot_segments=64,
                      max_slots_per_net=wl["slots"], group_window_ms=wl["window_ms"],
                      group_samples=wl["group_samples"], round_ms=round_ms,
                      hbm_budget_bytes=64 << 30, pinned_bytes=wl["pinned_mb"] << 20)
        log(f"server (round_ms {round_ms}): {s.stats()} ({time.perf_counter() - t0:.1f} s to "
            "allocate and capture)")
        return s

    # saturation is measured with the greedy launch rule (a group launches as soon as a slot
    # of its net is free), which the fixed-period rule cannot exceed
    srv = make_server(0.0)
    pool = prototype_features(Bn["Nin"], 8192, 99)
    edges_per = {0: A["N"] * 32 * A["L"],
                 1: Bn["N"] * Bn["K"] * Bn["L"] + Bn["N"] * Bn["Nin"] + Bn["Nout"] * Bn["N"]}
    rng = np.random.default_rng(5)

    def draw(n):
        sizes = np.exp(rng.uniform(np.log(wl["smin"]), np.log(wl["smax"] + 1), n)).astype(int)
        return [(int(rng.integers(0, 2)), int(min(max(s, wl["smin"]), wl["smax"])))
                for s in sizes]

    def submit(net, B, i):
        if net == 0:
            return srv.submit_images(0, IMG_SEED, 10000 * i, B)
        off = int(rng.integers(0, 8192 - B + 1))
        return srv.submit(1, pool[off:off + B])

    def run(reqs, rate):
        """Submits reqs (Poisson at `rate` requests/s, or all at once when None); returns
        latencies (ms), edges and the wall time from the first submit to the last result."""
        ids = []
        start = time.perf_counter()
        t_next = start
        for i, (net, B) in enumerate(reqs):
            if rate:
                t_next += rng.exponential(1.0 / rate)
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
            ids.append((submit(net, B, i), net, B))
        lat, edges = [], 0.0
        for rid, net, B in ids:
            _, info = srv.wait(rid, 600000.0)
            lat.append(info["ms_total"])
            edges += float(B) * edges_per[net]
        lat = np.array(lat)
        nets_of = np.array([n for _, n, _ in ids])
        sizes_of = np.array([b for _, _, b in ids])
        for n in (0, 1):
            sel = lat[nets_of == n]
            if sel.size:
                big = lat[(nets_of == n) & (sizes_of >= 1024)]
                log(f"  net {n}: {sel.size} requests, p50 {np.percentile(sel, 50):.1f} ms, "
                    f"p99 {np.percentile(sel, 99):.1f} ms; >=1024 samples: {big.size} requests"
                    + (f", p50 {np.percentile(big, 50):.1f} ms" if big.size else ""))
        return lat, edges, time.perf_counter() - start

    run(draw(16), None)  # warm-up
    lat, edges, wall = run(draw(wl["burst"]), None)
    sat_rps = wl["burst"] / wall
    if args.sat_rps:
        log(f"measured burst rate {sat_rps:.1f} req/s; loads taken from --sat-rps {args.sat_rps}")
        sat_rps = args.sat_rps
    log(f"saturated: {wl['burst']} requests in {wall:.2f} s = {sat_rps:.1f} req/s, "
        f"{edges / wall:.3e} edges/s; burst latency p50 {np.percentile(lat, 50):.1f} ms")
    table = {"saturated_burst": {"requests_per_s": sat_rps, "edges_per_s": edges / wall,
                                 "p50_ms": float(np.percentile(lat, 50)),
                                 "p99_ms": float(np.percentile(lat, 99))}}
    if wl["round_ms"] > 0:
        srv.close()
        srv = make_server(wl["round_ms"])
        run(draw(16), None)  # warm-up
    for load in wl["loads"]:
        lat, edges, wall = run(draw(wl["requests"]), load * sat_rps)
        row = {"offered_requests_per_s": load * sat_rps, "edges_per_s": edges / wall,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50))}
        table[f"load_{int(load * 100)}"] = row
        log(f"load {load:.0%}: p50 {row['p50_ms']:.1f} ms p99 {row['p99_ms']:.1f} ms "
            f"(p99/p50 {row['p99_over_p50']:.2f}), {row['edges_per_s']:.3e} edges/s")
    st = srv.stats()
    srv.close()
    out = {
        "metric": "serving edges/s at 70% of saturated load, with p50/p99 latency per load",
        "value": table["load_70"]["edges_per_s"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": wl["requests"],
        "warmup": 16,
        "ms_per_step": table["load_70"]["p50_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SDGC-like net A with device-rendered blob images; seeded "
                "unstructured net B with prototype features)",
        "config": {"workload": f"Poisson requests of {wl['smin']}-{wl['smax']} samples "
                               f"(log-uniform) over net A (N={A['N']} L={A['L']} "
                               f"bias={A['bias']} t={A['t']}, exact) and net B ({Bn['Nin']}->"
                               f"{Bn['N']}x{Bn['L']}->{Bn['Nout']}, {Bn['K']} nnz/row, t={Bn['t']}, "
                               f"lossy); slots of 8192 samples / 64 requests, {wl['slots']} "
                               f"per net, "
                               + (f"fixed grouping periods of {wl['round_ms']} ms (saturation "
                                  "measured with the greedy launch rule)"
                                  if wl["round_ms"] > 0 else
                                  f"grouping window {wl['window_ms']} ms, groups close at "
                                  f"{wl['group_samples'] or 8192} samples")
                               + f", request inputs in a {wl['pinned_mb']} MiB pinned arena",
                   "server": st},
        "serving": table,
        "roofline": None,
        "cpu_baseline": None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--bf-steps", type=int, default=5,
                    help="timed steps of the plain-BF comparison (at most --steps)")
    ap.add_argument("--threshold", type=int, default=None)
    ap.add_argument("--dense", action="store_true", help="also time the plain-dense form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--pmc-pass", default=None, choices=["BF", "SNICIT"],
                    help="run exactly one forward of this mode (for counter collection) and "
                         "write per-layer rows to gpurun_out/pmc_rows_<mode>.json")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="three 1 GiB device-to-device copies (torch), for scaling counters")
    ap.add_argument("--stream-weights", type=int, default=0, metavar="SLOTS",
                    help="keep the hidden layers' weights in pinned host memory and stream "
                         "them through SLOTS device buffers (c1-c3)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="read, write and copy rates of 2 GiB device buffers (torch)")
    ap.add_argument("--slots", type=int, default=None, help="c5: slots per net")
    ap.add_argument("--window-ms", type=float, default=None, help="c5: grouping window")
    ap.add_argument("--group-samples", type=int, default=None, help="c5: group cap")
    ap.add_argument("--round-ms", type=float, default=None,
                    help="c5: fixed grouping period (0: launch rule with window and fill)")
    ap.add_argument("--sat-rps", type=float, default=None,
                    help="c5: saturated request rate to take the loads from (default: measured "
                         "by the burst of this run)")
    ap.add_argument("--trace-steps", type=int, default=0, metavar="K",
                    help="run one untimed and K SNICIT forwards (captured graphs) and exit: the "
                         "command kernel traces are taken from")
    ap.add_argument("--tune", default="placement", choices=["none", "shape", "placement"],
                    help="autotune the SNICIT plan before timing: launch shapes and post form "
                         "(shape), plus fresh placements of the activation buffers (placement)")
    ap.add_argument("--tune-cache", default=None, metavar="PATH",
                    help="the autotuner's JSON cache (read and written)")
    ap.add_argument("--shape-from", default=None, metavar="PATH",
                    help="set the SNICIT plan's launch shapes and post form from the single "
                         "entry of an autotuner cache instead of tuning (profiled runs)")
    ap.add_argument("--no-forms", action="store_true",
                    help="c3: skip the batched (c4) and serving (c5) summaries in the line")
    ap.add_argument("--short-rows", type=int, default=0, metavar="K",
                    help="panel grouping with and without, on a random net of K-input rows")
    args = ap.parse_args()
    if args.stream_ceiling:
        stream_ceiling()
        return
    if args.short_rows:
        short_rows(args.short_rows)
        return
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 3 x 1 GiB")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); the layer loop "
              "shards by samples with no exchange inside the loop, so N GPUs would run N "
              "independent copies of this measurement.", file=sys.stderr)
        sys.exit(2)

    import snicit_amd as sa

    wl = dict(WORKLOADS[args.config])
    if args.config == "c4":
        print(json.dumps(bench_batch(args, wl)), flush=True)
        return
    for key, val in (("slots", args.slots), ("window_ms", args.window_ms),
                     ("group_samples", args.group_samples), ("round_ms", args.round_ms)):
        if val is not None and key in wl:
            wl[key] = val
    if args.config == "c5":
        if args.pmc_pass:
            pass_serve_net_b(args, wl)
        else:
            print(json.dumps(bench_serve(args, wl)), flush=True)
        return
    if args.threshold is not None:
        wl["t"] = args.threshold
    N, L, B, t = wl["N"], wl["L"], wl["B"], wl["t"]
    ctx = sa.Context(0)
    t0 = time.perf_counter()
    net = sa.Net.generate_sdgc(ctx, N, L, NET_SEED, bias=wl["bias"])
    log(f"net N={N} L={L}: {net.panels} panels, {net.device_bytes / 1e9:.2f} GB resident, "
        f"{time.perf_counter() - t0:.1f} s")
    P = net.panels // L
    weights = "resident in device memory"
    if args.stream_weights:
        net.stream_weights(args.stream_weights)
        weights = (f"streamed: pinned host blocks copied into {args.stream_weights} device "
                   f"slots per layer")
        log(f"weights {weights}; {net.device_bytes / 1e9:.2f} GB left on the device")

    if args.pmc_pass:
        # exactly one forward (the eager profiling run), with the rows each layer launch
        # processes: live columns before t, delta-list positions after it
        plan = sa.Plan(net, sa.default_cfg(args.pmc_pass, threshold=t), B)
        plan.set_input_images(IMG_SEED, 0, B)
        if args.shape_from:
            apply_shapes(plan, args.shape_from)
        ms_l, rows_l, _ = plan.profile()
        tt = t if args.pmc_pass == "SNICIT" else L
        rec = {"N": N, "P": P, "rows": [int(r) for r in rows_l[:tt]],
               "live_rows": live_rows(plan, B, tt), "ms_total": float(sum(ms_l))}
        if args.pmc_pass == "SNICIT":
            rec["post_rows"] = [int(r) for r in rows_l[tt:L]]
        os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
        with open(os.path.join(ROOT, "gpurun_out", f"pmc_rows_{args.pmc_pass}.json"), "w") as f:
            json.dump(rec, f)
        log(f"pmc pass {args.pmc_pass}: {len(rec['rows'])} plain launches")
        return

    snic = sa.Plan(net, sa.default_cfg("SNICIT", threshold=t), B)
    snic.set_input_images(IMG_SEED, 0, B)
    tune = None
    plain_tiles = 2  # the library's default
    shapes = None    # the shapes the SNICIT plan runs with, when not the defaults
    if args.shape_from:
        shapes = apply_shapes(snic, args.shape_from)
        plain_tiles = shapes["plain_tiles"]
    elif args.tune != "none":
        # untuned steps for the record (not in a traced run: the plan captures its stage
        # graphs once, after tuning), then the autotuner, outside the timed region
        ms_u, st_u = None, None
        if not args.trace_steps:
            ms_u, _, st_u = time_steps(snic, max(2, args.steps // 2), 1)
        t0 = time.perf_counter()
        info = snic.autotune(args.tune_cache, placement=args.tune == "placement")
        plain_tiles = info["plain_tiles"]
        shapes = info
        tune = {"mode": args.tune, "untuned_ms_per_step": ms_u,
                "untuned_stages_ms": st_u and {"pre": st_u.ms_pre, "post": st_u.ms_post},
                "tuning_s": time.perf_counter() - t0,
                **{k: info[k] for k in ("grid_target", "narrow", "post_grid_target",
                                        "post_narrow", "post_form", "ms_run_default",
                                        "ms_run_tuned", "ms_pre_default", "ms_pre_tuned",
                                        "ms_post_default", "ms_post_tuned", "placements",
                                        "placement", "compact_slack", "plain_tiles")}}
        log(f"autotune ({args.tune}, {tune['tuning_s']:.1f} s): untuned {ms_u} ms/step; "
            f"{json.dumps(info)}")
    if args.trace_steps:
        # the trace's marker: every dispatch after this images_kernel belongs to the one
        # untimed and K timed forwards (the tuner's runs come before it)
        snic.set_input_images(IMG_SEED, 0, B)
        ms_s, ev_s, st_s = time_steps(snic, args.trace_steps, 1)
        log(f"trace steps: SNICIT {ms_s:.3f} ms/step; pre {st_s.ms_pre:.3f} cluster "
            f"{st_s.ms_cluster:.3f} post {st_s.ms_post:.3f} recover {st_s.ms_recover:.3f}")
        return
    ms_s, ev_s, st_s, (pre_ms, post_ms) = time_steps(snic, args.steps, args.warmup, stages=True)
    cat_s = snic.categories()
    log(f"SNICIT: {ms_s:.3f} ms/step (events {ev_s:.3f}); pre {st_s.ms_pre:.3f} cluster "
        f"{st_s.ms_cluster:.3f} post {st_s.ms_post:.3f} recover {st_s.ms_recover:.3f}; "
        f"B'={st_s.live_at_threshold} k={st_s.centroids} nerow {st_s.nerow_first}->"
        f"{st_s.nerow_last} active={st_s.active}")

    bf = sa.Plan(net, sa.default_cfg("BF"), B)
    bf.set_input_images(IMG_SEED, 0, B)
    ms_b, ev_b, st_b = time_steps(bf, max(1, min(args.steps, args.bf_steps)),
                                  min(args.warmup, 1))
    cat_b = bf.categories()
    same = bool(np.array_equal(cat_s, cat_b))
    log(f"plain BF: {ms_b:.3f} ms/step (events {ev_b:.3f}); categories identical: {same}")
    del bf
    ms_d = None
    if args.dense:
        dn = sa.Plan(net, sa.default_cfg("PLAIN_DENSE"), B)
        dn.set_input_images(IMG_SEED, 0, B)
        ms_d, _, _ = time_steps(dn, max(1, args.steps // 2), 1)
        same = same and bool(np.array_equal(dn.categories(), cat_s))
        log(f"plain dense: {ms_d:.3f} ms/step")
        del dn
    sweep = None
    if args.config == "c3" and not args.no_forms:
        # the workload's threshold sweep: t = 60 and 120 with the shapes the t = 30 plan runs
        # with (no placement search), against the same plain run
        sweep = {str(t): {"ms_per_step": ms_s, "speedup_vs_plain_bf": ms_b / ms_s,
                          "categories_identical_snicit_vs_plain": same,
                          "stages_ms": {"pre": st_s.ms_pre, "cluster": st_s.ms_cluster,
                                        "post": st_s.ms_post, "recover": st_s.ms_recover},
                          "centroids": st_s.centroids, "nerow_first": st_s.nerow_first,
                          "nerow_last": st_s.nerow_last}}
        for t2 in (60, 120):
            p2 = sa.Plan(net, sa.default_cfg("SNICIT", threshold=t2), B)
            p2.set_input_images(IMG_SEED, 0, B)
            if shapes:
                apply_shape_dict(p2, shapes)
            ms2, _, st2 = time_steps(p2, max(2, args.steps // 2), 1)
            same2 = bool(np.array_equal(p2.categories(), cat_b))
            sweep[str(t2)] = {"ms_per_step": ms2, "speedup_vs_plain_bf": ms_b / ms2,
                              "categories_identical_snicit_vs_plain": same2,
                              "stages_ms": {"pre": st2.ms_pre, "cluster": st2.ms_cluster,
                                            "post": st2.ms_post, "recover": st2.ms_recover},
                              "centroids": st2.centroids, "nerow_first": st2.nerow_first,
                              "nerow_last": st2.nerow_last}
            log(f"t={t2}: SNICIT {ms2:.3f} ms/step ({ms_b / ms2:.2f}x plain); categories "
                f"identical: {same2}; k={st2.centroids} nerow {st2.nerow_first}->"
                f"{st2.nerow_last}")
            p2.close()

    # Roofline of the two layer kernels: the plain panel layer before t (stage "pre") and the
    # delta layer after it (stage "post").  Rows per launch come from one eager profiling run
    # (per-launch HIP events on the plan's stream, kept as eager_avg_launch_ms); the time is
    # each stage's HIP-event time over the timed steps divided by its layer launches, so it
    # is the captured-graph rate the kernel trace sees (the stage's few list compactions
    # are counted in it).  The one with the larger total is the roofline kernel.
    ms_l, rows_l, cms_l = snic.profile()
    live_l = live_rows(snic, B, t)
    plain_name = ((f"panel_spmm_tiles_kernel<{plain_tiles}>" if plain_tiles > 1 else
                   "panel_spmm_kernel<0>") +
                  " (plain epilogue, one chain per distinct panel row; layers < t)")
    post_name = ("panel_post_wave_kernel (delta rows, layers >= t; the next layer's centroid "
                 "rows in a fifth wavefront)")
    pre, post = list(range(t)), list(range(t, L))
    kern = {
        plain_name: {"bytes": sum(spmm_layer_bytes(N, P, live_l[l]) for l in pre),
                     "s": pre_ms / 1000.0, "eager_s": sum(ms_l[l] for l in pre) / 1000.0,
                     "launches": len(pre),
                     "per_unit": "8*N*rows (live columns) + 8*P*32 (input ids, one distinct weight row) + "
                                 "4*(P+1) + 4*N"},
        post_name: {"bytes": sum(post_layer_bytes(N, P, rows_l[l]) for l in post),
                    "s": post_ms / 1000.0, "eager_s": sum(ms_l[l] for l in post) / 1000.0,
                    "launches": len(post),
                    "per_unit": "8*N*rows + 4*rows (centroid index) + P*(8*32 + 4*64) (input "
                                "ids, distinct weight row, chain record) + 4*(P+1) + 4*N"},
    }
    dom = max(kern, key=lambda n: kern[n]["s"])

    def roof(name):
        d = kern[name]
        gbs = d["bytes"] / d["s"] / 1e9 if d["s"] > 0 else None
        traffic, traffic_kernel = measured_traffic(wl, name.split(" ")[0])
        return {"kernel": name, "bound": "hbm",
                "achieved": round(gbs, 1) if gbs else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None,
                "traffic": traffic, "traffic_profiled_kernel": traffic_kernel,
                "bytes_per_launch": d["bytes"] / max(1, d["launches"]),
                "avg_launch_ms": 1000.0 * d["s"] / max(1, d["launches"]),
                "eager_avg_launch_ms": 1000.0 * d["eager_s"] / max(1, d["launches"]),
                "launches": d["launches"], "algorithmic_bytes": d["per_unit"]}

    roofs = {n: roof(n) for n in kern}
    for n, r in roofs.items():
        log(f"roofline {n.split(' ')[0]}: {r['achieved']} GB/s ({r['frac']}), avg "
            f"{r['avg_launch_ms']:.4f} ms over {r['launches']} launches, "
            f"{kern[n]['s'] * 1000:.1f} ms in total")

    edges = float(B) * N * 32 * L
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, cat_s, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.3e} edges/s on {cpu['cores']} cores; {cpu['sample']}; "
            f"categories match GPU: {cpu['categories_match_gpu']}")
    forms = {}
    if args.config == "c3" and not args.no_forms:
        # the batched and serving forms on their own workloads, after this workload's plans
        # and net are freed (summaries in this line; full lines with --config c4 / c5)
        snic.close()
        net.close()
        a = bench_batch(args, dict(WORKLOADS["c4"]))
        forms["form_a_batches"] = {k: a[k] for k in (
            "metric", "value", "unit", "ms_per_step", "untuned_ms_per_step",
            "plain_bf_ms_per_step", "speedup_vs_plain_bf", "categories_identical_net_a",
            "categories_agreement_net_b_lossy", "config")}
        b = bench_serve(args, dict(WORKLOADS["c5"]))
        forms["form_b_serving"] = {k: b[k] for k in ("metric", "value", "unit", "config",
                                                     "serving")}
    out = {
        "metric": "SNICIT edges/s (samples x nnz(W) x layers / time), one forward per step",
        "value": edges / (ms_s / 1000.0),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SDGC-like banded+permuted net, binary blob images)",
        "config": {"workload": f"SDGC-like SNICIT vs plain, N={N} L={L} batch={B} "
                               f"bias={wl['bias']} t={t}", "N": N, "L": L, "batch": B,
                   "threshold": t, "bias": wl["bias"], "clip": 32.0, "nnz_per_row": 32,
                   "weights": weights},
        "plain_bf_ms_per_step": ms_b,
        "plain_bf_edges_per_s": edges / (ms_b / 1000.0),
        "plain_dense_ms_per_step": ms_d,
        "speedup_vs_plain_bf": ms_b / ms_s,
        "categories_identical_snicit_vs_plain": same,
        "stages_ms": {"pre": st_s.ms_pre, "cluster": st_s.ms_cluster, "post": st_s.ms_post,
                      "recover": st_s.ms_recover, "events_total": ev_s},
        "snicit": {"live_at_threshold": st_s.live_at_threshold, "centroids": st_s.centroids,
                   "nerow_first": st_s.nerow_first, "nerow_last": st_s.nerow_last,
                   "active": st_s.active},
        "roofline": roofs[dom],
        "other_kernel_roofline": roofs[plain_name if dom == post_name else post_name],
        "cpu_baseline": cpu,
        "autotune": tune,
        "threshold_sweep": sweep,
        **forms,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
