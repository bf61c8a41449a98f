"""HyperBall benchmark on one MI355X.

A step is one complete HyperBall run (init + iterations to stabilisation) over
the web-scale synthetic graph: n = 10^7 nodes, out-degree Zipf(2.1) on
[3, 10000], successors 70% uniform in a +-1000 window and 30% by
preferential attachment, sorted CSR, with its transpose (systolic frontier);
HyperLogLog log2m = 10 with 6-bit registers, Jenkins hash seed 42; sum of
distances and harmonic centrality requested.  Graph and transpose are built
and uploaded before the timed region.

metric: arc-registers per second = sum over iterations of
(visited arcs x registers per counter) / run time.

Beside the headline line the same JSON object carries:
  cpu_baseline  the oracle's OpenMP restatement of the reference engine on the
                headline graph itself (init + the first iterations), with the
                GPU timed on the same iterations;
  batched       64 independent instances in one batch against a sequential
                loop of single-instance runs;
  precision     log2m 12 on the headline graph with FP32 Kahan and plain FP32
                per-node accumulators against f64 (max / p99 relative error);
  large         n = 10^8, about 2.1e9 arcs, log2m 6, 5-bit registers, FP32
                Kahan accumulators, with the union step's roofline;
  uniform_1e5   the uniform n = 10^5 graph, GPU against the oracle engine;
  host_counters counter arrays in HBM, in host memory read in place, and in
                host memory streamed through device memory in tiles.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "webgraph-algo-rs_amd")
PEAK_HBM_GBS = 8000.0


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_bytes(s, n, counter_bytes, slab_bytes, outputs, delta_aware=True):
    """Per-iteration bytes of the no-reuse model: successor counters read,
    own counter read and written, stale copies, CSR and transpose ids,
    offsets, accumulators read-modify-write and the four frontier bitmaps.
    delta_aware: successor reads are the slabs actually gathered (delta
    gathering reads only the slabs that changed), plus the 4-byte change
    mask read per visited arc in such iterations and the 4-byte mask written
    per modified node; otherwise C per visited arc."""
    C = counter_bytes
    if delta_aware:
        succ = s["gathered_slabs"] * slab_bytes + 4 * s["modified"]
        if s["delta_gathering"]:
            succ += 4 * s["visited_arcs"]
    else:
        succ = C * s["visited_arcs"]
    return (succ + C * (2 * s["checked"] + 2 * s["copied"])
            + 4 * (s["scanned_arcs"] + s["transpose_arcs"]) + 8 * s["checked"]
            + 16 * s["modified"] * outputs + 4 * ((n + 7) // 8))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(orc, hb, g, gt, args, log2m, reg_bits):
    """The oracle's OpenMP restatement of the reference engine
    (hyperball_impl.rs:775-1213) on the headline graph itself: same graph
    bytes, hash, seed, log2m, register width and outputs.  It runs init and
    then iterations until args.cpu_budget seconds have passed (or the run
    ends); the GPU instance then runs exactly that prefix, run(k), so both
    rates are over identical work (the visited arcs are compared)."""
    threads = host_threads()
    ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=log2m,
                              reg_bits=reg_bits, hash=orc.HASH_JENKINS, seed=42,
                              sum_of_distances=True, sum_of_inverse_distances=True,
                              threads=threads)
    m = 1 << log2m
    t0 = time.perf_counter()
    ref.init()
    visited = []
    while True:
        ref.iterate()
        visited.append(ref.visited_arcs)
        if ref.modified == 0 or time.perf_counter() - t0 > args.cpu_budget:
            break
    cpu_s = time.perf_counter() - t0
    k = len(visited)
    work = sum(visited) * m
    del ref
    gpu = []
    for _ in range(3):
        t1 = time.perf_counter()
        hb.run(k)
        gpu.append(time.perf_counter() - t1)
    gpu_s = sorted(gpu)[1]
    same = [s["visited_arcs"] for s in hb.iteration_stats()] == visited
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    print(f"# cpu baseline: {k} iterations in {cpu_s:.1f}s on {threads} threads; "
          f"GPU {gpu_s * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return {"value": work / cpu_s, "unit": "arc-registers/s", "cores": threads, "kind": "port",
            "sample": f"the headline graph itself (n={g.num_nodes}, arcs={g.num_arcs}, log2m "
                      f"{log2m}, r {reg_bits}, jenkins(42), sum of distances + harmonic): init + "
                      f"the first {k} iterations in {cpu_s:.1f}s on {threads} threads",
            "cpu_model": cpu_model(), "host_cores_visible": os.cpu_count(),
            "affinity_cores": affinity,
            "threads_note": "16 threads: the host CPU share of one GPU on the measurement box",
            "gpu_same_work": {"value": work / gpu_s, "ms": gpu_s * 1e3, "iterations": k,
                              "same_visited_arcs": same},
            "gpu_over_cpu": cpu_s / gpu_s}


TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "union_step_web1e7_log2m10_r6_pmc.md")
SMALL_TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "union_step_web1e7_log2m6_r5.md")
KERNEL_SOURCE = os.path.join(PKG, "csrc", "hb_kernels.hip")


def kernel_source_sha():
    import hashlib
    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def profiled_traffic(args, profile=TRAFFIC_PROFILE, config=(10_000_000, 10, 6)):
    """HBM bytes per union step from the committed FETCH_SIZE / WRITE_SIZE
    passes (copy-calibrated), for the configuration they were taken on, and
    only when they were taken on the kernel source that is built now."""
    if (args.nodes, args.log2m, args.reg_bits) != config:
        return None, "no profile for this configuration"
    value, sha = None, None
    try:
        with open(profile) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_union_step:"):
                    value = float(line.split(":")[1])
                if line.startswith("kernel_source_sha256:"):
                    sha = line.split(":")[1].strip()
    except OSError:
        return None, "profile missing"
    if sha != kernel_source_sha():
        return None, "the profile was taken on other kernel sources"
    return value, os.path.relpath(profile, REPO) + f" (hb_kernels.hip sha256 {sha[:16]})"


def batch_workload(hbm, args):
    """64 independent instances: sizes log-uniform on [10^4, 2*10^6], each
    graph from the same generator with its own seed (1000 + i), log2m drawn
    uniformly from {6, 8, 10} (seeded), 6-bit registers, Jenkins seed per
    instance."""
    rng = np.random.default_rng(3)
    ns = np.exp(rng.uniform(np.log(1e4), np.log(2e6), args.batch)).astype(np.int64)
    log2ms = rng.choice([6, 8, 10], args.batch)
    builders = []
    for i, n in enumerate(ns):
        g = hbm.Graph.web(int(n), 2.1, 3, 10000, 1000, 0.7, 1000 + i)
        gt = g.transpose()
        builders.append(hbm.HyperBallBuilder.with_hyper_log_log(g, gt, int(log2ms[i]))
                        .register_bits(args.reg_bits).jenkins(42 + i)
                        .sum_of_distances(True).sum_of_inverse_distances(True))
    return builders


def bench_batch(hbm, args):
    """Batched run (one fused launch sequence per batched iteration, replayed
    from a HIP graph) against a sequential loop of single-instance runs of
    the same instances; both produce identical results."""
    t = time.perf_counter()
    builders = batch_workload(hbm, args)
    nodes = sum(b._graph.num_nodes for b in builders)
    arcs = sum(b._graph.num_arcs for b in builders)
    print(f"# batch of {len(builders)} graphs, {nodes} nodes, {arcs} arcs, built in "
          f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)

    def work_of(instances):
        return sum(s["visited_arcs"] * h.params()["num_registers"]
                   for h in instances for s in h.iteration_stats())

    def bytes_of(instances):
        # the headline's no-reuse model per instance: C = 4 r m / 32 bytes,
        # slabs of 4 r bytes
        total = 0.0
        for h, b in zip(instances, builders):
            C = 4 * args.reg_bits * h.params()["num_registers"] // 32
            total += sum(algorithmic_bytes(s, b._graph.num_nodes, C, 4 * args.reg_bits, 2)
                         for s in h.iteration_stats())
        return total

    batch = hbm.HyperBallBatch(builders)
    if args.heavy_threshold is None:
        chosen, table = batch.autotune()
        print("# batch autotune: " + ", ".join(f"{'none' if t == 2**32 - 1 else t}: {ms:.2f}"
                                               for t, ms in table) + f" -> {chosen}",
              file=sys.stderr, flush=True)
    else:
        chosen = args.heavy_threshold
        for h in batch.instances:
            h.set_tuning(chosen)
    for _ in range(args.warmup):
        batch.run_until_done()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run_until_done()
    batched_s = (time.perf_counter() - t0) / args.steps
    work = work_of(batch.instances)
    nbytes = bytes_of(batch.instances)
    steps_ms = batch.step_times()
    nf_batch = [h.neighborhood_function() for h in batch.instances]
    batch.close()
    print(f"# batched: {batched_s * 1e3:.1f} ms per run, {len(steps_ms)} batched iterations",
          file=sys.stderr, flush=True)

    singles = [b.build() for b in builders]
    for h in singles:
        if args.heavy_threshold is None:
            h.autotune(reps=1)
        else:
            h.set_tuning(args.heavy_threshold)
    for _ in range(args.warmup):
        for h in singles:
            h.run_until_done()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for h in singles:
            h.run_until_done()
    seq_s = (time.perf_counter() - t0) / args.steps
    same = all(np.array_equal(a, h.neighborhood_function()) for a, h in zip(nf_batch, singles))
    assert work == work_of(singles)
    for h in singles:
        h.close()
    print(f"# sequential: {seq_s * 1e3:.1f} ms per run", file=sys.stderr, flush=True)
    return {"instances": len(builders), "nodes": nodes, "arcs": arcs,
            "log2m": [6, 8, 10], "register_bits": args.reg_bits,
            "batched_value": work / batched_s, "batched_ms": batched_s * 1e3,
            "batched_iterations": len(steps_ms),
            "sequential_value": work / seq_s, "sequential_ms": seq_s * 1e3,
            "speedup": seq_s / batched_s, "identical_results": bool(same),
            "batch_heavy_row_threshold": chosen,
            "unit": "arc-registers/s",
            "roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                         "algorithmic_bytes_per_run": nbytes,
                         "batched_achieved": nbytes / batched_s / 1e9,
                         "batched_frac": nbytes / batched_s / 1e9 / PEAK_HBM_GBS,
                         "sequential_achieved": nbytes / seq_s / 1e9,
                         "sequential_frac": nbytes / seq_s / 1e9 / PEAK_HBM_GBS,
                         "note": "every instance's iterations under the headline's model, "
                                 "over the wall time of a whole run (host work included)"}}


def host_threads():
    """Host threads for the CPU legs: the cores this process may run on, at
    most 16 (the host CPU share of one GPU on the measurement box)."""
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return min(threads, 16)


def small_uniform(hbm, orc, args):
    """The uniform random graph: n = 10^5 nodes, exactly 8 distinct uniform
    successors each (self-loops excluded), log2m 8, register width from n,
    the default SipHash-1-3 hasher, NF + sum of distances + harmonic
    centrality in f64.  The GPU run (inputs resident) beside the oracle's
    OpenMP restatement of the reference engine on the same, whole workload."""
    g = hbm.Graph.uniform(100_000, 8, seed=1)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).sum_of_distances(True)
          .sum_of_inverse_distances(True).build())
    m = hb.params()["num_registers"]
    hb.run_until_done()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        hb.run_until_done()
    gpu_s = (time.perf_counter() - t0) / reps
    st = hb.iteration_stats()
    work = sum(s["visited_arcs"] for s in st) * m
    r = hb.params()["register_bits"]
    nbytes = sum(algorithmic_bytes(s, g.num_nodes, 4 * r * m // 32, 4 * r, 2) for s in st)
    nf, sd = hb.neighborhood_function(), hb.sum_of_distances()
    hb.close()
    out = {"workload": "uniform graph n=100000, 8 successors each, log2m 8, r from n, SipHash-1-3",
           "iterations": len(st), "gpu": {"value": work / gpu_s, "ms_per_run": gpu_s * 1e3,
                                          "ms_per_iteration": gpu_s * 1e3 / max(1, len(st))},
           "roofline": {"bound": "latency", "achieved": nbytes / gpu_s / 1e9, "peak": PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": nbytes / gpu_s / 1e9 / PEAK_HBM_GBS,
                        "algorithmic_bytes_per_run": nbytes,
                        "note": "the headline's byte model over the whole run's wall time; each "
                                "iteration is a few launches and one host synchronisation"}}
    if orc is not None:
        threads = host_threads()
        ref = orc.OracleHyperBall(g.offsets, g.successors, gt.offsets, gt.successors, log2m=8,
                                  sum_of_distances=True, sum_of_inverse_distances=True,
                                  threads=threads)
        t0 = time.perf_counter()
        ref.run()
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": work / cpu_s, "ms_per_run": cpu_s * 1e3, "cores": threads,
                               "kind": "port"}
        out["gpu_over_cpu"] = cpu_s / gpu_s
        out["same_sum_of_distances"] = bool(np.array_equal(sd, ref.sum_of_distances()))
        out["nf_max_rel_diff"] = float(np.max(np.abs(nf - ref.neighborhood_function()) / nf))
    out["unit"] = "arc-registers/s"
    return out


def rel_err(got, ref):
    """Max and median relative error of got against ref (absolute where ref is 0)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    err = np.abs(got - ref)
    nz = ref != 0
    err[nz] /= np.abs(ref[nz])
    return {"max": float(err.max()), "p99": float(np.quantile(err, 0.99)),
            "median": float(np.median(err)), "argmax": int(err.argmax())}


def bench_precision(hbm, args):
    """Accumulator precision on the web graph (n = args.nodes) with log2m 12:
    the same run with f64, f32 Kahan-compensated and plain f32 per-node
    accumulators.  Registers are integers in every mode; reachable counts
    come from the integer register sums, so they are compared too."""
    t = time.perf_counter()
    g = hbm.Graph.web(args.nodes, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    print(f"# graph n={g.num_nodes} arcs={g.num_arcs} built in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    out = {}
    ref = None
    for name, mode in (("f64", hbm.ACC_F64), ("f32_kahan", hbm.ACC_F32_KAHAN), ("f32", hbm.ACC_F32)):
        hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 12).register_bits(5).jenkins(42)
              .sum_of_distances(True).sum_of_inverse_distances(True).accumulator_precision(mode)
              .build())
        t0 = time.perf_counter()
        hb.run_until_done()
        dt = time.perf_counter() - t0
        res = {"sum_of_distances": hb.sum_of_distances(),
               "harmonic_centrality": hb.harmonic_centralities(),
               "reachable_nodes": hb.reachable_nodes(),
               "neighborhood_function": hb.neighborhood_function()}
        iters = len(hb.iteration_stats())
        hb.close()
        print(f"# {name}: {iters} iterations in {dt * 1e3:.0f} ms", file=sys.stderr, flush=True)
        if ref is None:
            ref = res
            out["f64"] = {"run_ms": dt * 1e3, "iterations": iters}
            continue
        out[name] = {"run_ms": dt * 1e3, "iterations": iters,
                     **{k: rel_err(res[k], ref[k]) for k in res}}
    return {"workload": f"web graph n={g.num_nodes} arcs={g.num_arcs}, log2m 12, r 5, "
                        "Jenkins 42, run to stabilisation",
            "reference": "f64 accumulators on the device (bit-exact with the oracle's "
                         "f64 engine in the parity tests)", "modes": out}


def bench_small_counters(hbm, args):
    """The headline graph (n = 10^7, the same generator and seeds) at log2m 6
    with 5-bit registers: 40-byte counters, whose iterations that check every
    node run as the window kernel (counters around each block of consecutive
    nodes staged in LDS).  One warmup run, then args.steps timed runs to
    stabilisation; the union step's roofline uses the headline's model with
    C = 40 bytes, and the traffic comes from the committed profile of this
    configuration when it matches the kernel source."""
    g = hbm.Graph.web(args.nodes, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 6).register_bits(5).jenkins(42)
          .sum_of_distances(True).sum_of_inverse_distances(True).build())
    hb.set_tuning(512)
    hb.set_timing(True)
    hb.run_until_done()
    work, union_ms, nbytes, launches, windows, iters = 0, 0.0, 0.0, 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hb.run_until_done()
        st = hb.iteration_stats()
        iters = len(st)
        for s_ in st:
            work += s_["visited_arcs"] * 64
            union_ms += s_["union_ms"]
            nbytes += algorithmic_bytes(s_, g.num_nodes, 40, 4 * 5, 2)
            launches += 1
            windows += s_["window"]
    elapsed = time.perf_counter() - t0
    hb.close()
    achieved = nbytes / (union_ms * 1e-3) / 1e9 if union_ms > 0 else 0.0
    small = argparse.Namespace(nodes=args.nodes, log2m=6, reg_bits=5)
    traffic, src = profiled_traffic(small, SMALL_TRAFFIC_PROFILE, (10_000_000, 6, 5))
    per_launch = nbytes / max(1, launches)
    print(f"# small counters (log2m 6): {elapsed * 1e3 / args.steps:.1f} ms per run", file=sys.stderr,
          flush=True)
    return {"workload": f"web graph n={g.num_nodes} arcs={g.num_arcs}, log2m 6, r 5, jenkins(42), "
                        "systolic frontier, run to stabilisation",
            "value": work / elapsed, "unit": "arc-registers/s",
            "ms_per_step": elapsed * 1e3 / args.steps, "iterations": iters,
            "window_iterations": windows // max(1, args.steps),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM_GBS, "avg_launch_ms": union_ms / max(1, launches),
                         "algorithmic_bytes_per_launch": per_launch, "traffic": traffic,
                         "traffic_over_algorithmic": traffic / per_launch if traffic else None,
                         "traffic_source": src, "kernel": "union step (window_kernel in the "
                         "iterations that check every node, union_kernel in the others)"}}


def bench_host_counters(hbm, args):
    """Counter arrays in HBM against pinned host memory, read and written by
    the kernels over the host link (COUNTERS_HOST), and against host memory
    streamed through device memory in tiles (COUNTERS_STREAM: a tile of n/2,
    then of n/8 nodes, so that 2 or 8 tiles' counters never share the device):
    the web graph with n = args.host_nodes, log2m 8, 6-bit registers,
    systolic frontier, sum of distances and harmonic; one run to
    stabilisation each, after a warmup run, results compared."""
    g = hbm.Graph.web(args.host_nodes, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    out = {"workload": f"web graph n={g.num_nodes} arcs={g.num_arcs}, log2m 8, r 6, jenkins(42)",
           "unit": "arc-registers/s"}
    results = []
    half = (g.num_nodes + 1) // 2
    eighth = (g.num_nodes + 7) // 8
    for name, where, tile in (("device", hbm.COUNTERS_DEVICE, 0), ("host", hbm.COUNTERS_HOST, 0),
                              ("stream_2_tiles", hbm.COUNTERS_STREAM, half),
                              ("stream_8_tiles", hbm.COUNTERS_STREAM, eighth)):
        hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, 8).register_bits(6).jenkins(42)
              .sum_of_distances(True).sum_of_inverse_distances(True).counter_memory(where, tile)
              .build())
        hb.run_until_done()
        t0 = time.perf_counter()
        hb.run_until_done()
        secs = time.perf_counter() - t0
        st = hb.iteration_stats()
        work = sum(s["visited_arcs"] for s in st) * 256
        out[name] = {"ms_per_run": secs * 1e3, "value": work / secs, "iterations": len(st)}
        if tile:
            out[name]["tile_nodes"] = tile
        results.append((hb.counters(0), hb.sum_of_distances(), hb.neighborhood_function()))
        hb.close()
        print(f"# {name} counters: {secs * 1e3:.1f} ms per run", file=sys.stderr, flush=True)
    out["identical_results"] = all(bool(np.array_equal(a, b)) for r in results[1:]
                                   for a, b in zip(results[0], r))
    out["host_over_device"] = out["host"]["ms_per_run"] / out["device"]["ms_per_run"]
    for k in ("stream_2_tiles", "stream_8_tiles"):
        out[k]["speedup_over_host"] = out["host"]["ms_per_run"] / out[k]["ms_per_run"]
    return out


def bench_large(hbm, args, compare_transposes=True):
    """One large instance: n = 10^8 nodes, the same generator with out-degree
    Zipf(2.1) on [4, 10000] and a +-10,000 window (about 2e9 arcs), log2m 6,
    5-bit registers, systolic frontier, f32 Kahan-compensated accumulators.
    The transpose is built on the device inside build(); with
    compare_transposes the host counting sort and hb_graph_transpose_device
    are timed beside it and compared.  The union step's roofline uses the
    same algorithmic-bytes model as the headline (40-byte counters)."""
    t = time.perf_counter()
    g = hbm.Graph.web(args.large_nodes, 2.1, 4, 10000, 10000, 0.7, 5)
    print(f"# graph n={g.num_nodes} arcs={g.num_arcs} built in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    setup = {}
    if compare_transposes:
        t = time.perf_counter()
        gt = g.transpose()
        host_t = time.perf_counter() - t
        t = time.perf_counter()
        gtd = g.transpose(device=0)
        dev_t = time.perf_counter() - t
        same = bool(np.array_equal(gt.offsets, gtd.offsets) and
                    np.array_equal(gt.successors, gtd.successors))
        print(f"# transpose: host {host_t:.1f}s, device (with both copies) {dev_t:.1f}s, "
              f"identical {same}", file=sys.stderr, flush=True)
        del gt, gtd
        setup = {"host_transpose_s": host_t, "device_transpose_with_copies_s": dev_t,
                 "transpose_identical": same}
    t = time.perf_counter()
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, None, 6).transpose_on_device()
          .register_bits(5).jenkins(42).sum_of_distances(True).sum_of_inverse_distances(True)
          .accumulator_precision(hbm.ACC_F32_KAHAN).build())
    hb.set_timing(True)
    setup["build_with_device_transpose_s"] = time.perf_counter() - t
    print(f"# instance built (upload + device transpose) in {setup['build_with_device_transpose_s']:.1f}s",
          file=sys.stderr, flush=True)
    nodes, arcs = g.num_nodes, g.num_arcs
    del g
    p = hb.params()
    C = p["words_per_counter"] * 8
    for i in range(args.warmup):
        hb.run_until_done()
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    work, union_ms, iters, nbytes, dense_bytes, dense_ms = 0, 0.0, 0, 0.0, 0.0, 0.0
    steps = max(1, min(args.steps, args.large_steps))
    t0 = time.perf_counter()
    for i in range(steps):
        hb.run_until_done()
        st = hb.iteration_stats()
        iters = len(st)
        for s_ in st:
            work += s_["visited_arcs"] * 64
            union_ms += s_["union_ms"]
            b = algorithmic_bytes(s_, nodes, C, 4 * 5, 2)
            nbytes += b
            if not s_["systolic"]:
                dense_bytes += b
                dense_ms += s_["union_ms"]
        print(f"# step {i}: {iters} iterations", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    hb.close()
    achieved = nbytes / (union_ms * 1e-3) / 1e9 if union_ms > 0 else 0.0
    return {"workload": f"web graph n={nodes} arcs={arcs} (+-10000 window, out-degree Zipf(2.1) "
                        "on [4, 10000]), log2m 6, r 5, f32 Kahan accumulators",
            "value": work / elapsed, "unit": "arc-registers/s", "ms_per_step": elapsed * 1e3 / steps,
            "steps": steps, "iterations": iters, "union_ms_per_step": union_ms / steps,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM_GBS,
                         "dense_iterations_frac": (dense_bytes / (dense_ms * 1e-3) / 1e9 / PEAK_HBM_GBS
                                                   if dense_ms > 0 else None),
                         "kernel": "union step", "bytes_model": "as the headline's, C = 40 bytes"},
            "setup": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=10_000_000)
    ap.add_argument("--log2m", type=int, default=10)
    ap.add_argument("--reg-bits", type=int, default=6)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU iterations on the headline graph (whole iterations)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="instances in the batched leg (0 = skip)")
    ap.add_argument("--only-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the uniform n=1e5 leg")
    ap.add_argument("--only-small", action="store_true", help="only the uniform n=1e5 leg")
    ap.add_argument("--heavy-threshold", type=int, default=None,
                    help="fixed heavy-row threshold instead of the autotuners (for profiling)")
    ap.add_argument("--precision", action="store_true",
                    help="only the accumulator precision study (log2m 12 on the web graph)")
    ap.add_argument("--large", action="store_true",
                    help="only the large instance (n = --large-nodes, log2m 6, r 5)")
    ap.add_argument("--large-nodes", type=int, default=100_000_000)
    ap.add_argument("--large-steps", type=int, default=3, help="timed runs of the large instance")
    ap.add_argument("--no-large", action="store_true", help="skip the large instance leg")
    ap.add_argument("--no-small-counters", action="store_true",
                    help="skip the log2m-6 leg on the headline graph")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision leg")
    ap.add_argument("--host-counters", action="store_true",
                    help="only the counters-in-host-memory comparison (n = --host-nodes)")
    ap.add_argument("--host-nodes", type=int, default=1_000_000)
    ap.add_argument("--no-host-counters", action="store_true",
                    help="skip the host / streamed counter-memory leg")
    ap.add_argument("--stats", action="store_true", help="print per-iteration statistics to stderr")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (a known-size "
                         "kernel that scales profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on exactly one GPU; multi-GPU sharding is "
                                   "designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy_mib:
        import torch
        nbytes = args.calibrate_copy_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(f"# calibration copy of {args.calibrate_copy_mib} MiB done", file=sys.stderr, flush=True)

    hbm = _load("hyperball", os.path.join(PKG, "hyperball.py"))
    if not os.path.exists(hbm.LIB_PATH):
        _load("graft_entry", os.path.join(REPO, "__graft_entry__.py")).build()

    if args.only_batch:
        print(json.dumps({"batched": bench_batch(hbm, args)}), flush=True)
        return
    if args.only_small:
        orc = None if args.no_cpu_baseline else _load("hb_oracle_py", os.path.join(REPO, "oracle", "oracle.py"))
        print(json.dumps({"uniform_1e5": small_uniform(hbm, orc, args)}), flush=True)
        return
    if args.precision:
        print(json.dumps({"precision": bench_precision(hbm, args)}), flush=True)
        return
    if args.large:
        print(json.dumps({"large": bench_large(hbm, args)}), flush=True)
        return
    if args.host_counters:
        print(json.dumps({"host_counters": bench_host_counters(hbm, args)}), flush=True)
        return

    t = time.perf_counter()
    g = hbm.Graph.web(args.nodes, 2.1, 3, 10000, 1000, 0.7, 2)
    gt = g.transpose()
    print(f"# graph n={g.num_nodes} arcs={g.num_arcs} built in {time.perf_counter() - t:.1f}s",
          file=sys.stderr, flush=True)
    hb = (hbm.HyperBallBuilder.with_hyper_log_log(g, gt, args.log2m).register_bits(args.reg_bits)
          .jenkins(42).sum_of_distances(True).sum_of_inverse_distances(True).build())
    if args.heavy_threshold is None:
        chosen, table = hb.autotune(reps=2)
        print("# autotune (heavy-row threshold, dense-iteration ms): " +
              ", ".join(f"{'none' if t == 2**32 - 1 else t}: {ms:.2f}" for t, ms in table) +
              f" -> {chosen}", file=sys.stderr, flush=True)
    else:
        chosen = args.heavy_threshold
        hb.set_tuning(chosen)
    hb.set_timing(True)
    p = hb.params()
    m, counter_bytes = p["num_registers"], p["words_per_counter"] * 8
    slab_bytes = 4 * p["register_bits"]  # 32 registers

    for i in range(args.warmup):
        hb.run_until_done()
        print(f"# warmup {i} done", file=sys.stderr, flush=True)

    work = 0
    bytes_total = 0.0
    bytes_no_delta = 0.0
    union_ms = 0.0
    dense_bytes, dense_ms = 0.0, 0.0
    launches = 0
    iters = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        hb.run_until_done()
        st = hb.iteration_stats()
        iters = len(st)
        for s in st:
            work += s["visited_arcs"] * m
            b = algorithmic_bytes(s, g.num_nodes, counter_bytes, slab_bytes, 2)
            bytes_total += b
            bytes_no_delta += algorithmic_bytes(s, g.num_nodes, counter_bytes, slab_bytes, 2,
                                                delta_aware=False)
            union_ms += s["union_ms"]
            if not s["systolic"]:
                dense_bytes += b
                dense_ms += s["union_ms"]
            launches += 1
        print(f"# step {i}: {iters} iterations", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    if args.stats:
        for s in hb.iteration_stats():
            print("# " + json.dumps(s), file=sys.stderr)

    achieved = bytes_total / (union_ms * 1e-3) / 1e9 if union_ms > 0 else 0.0
    traffic, traffic_src = profiled_traffic(args)
    result = {
        "metric": "HyperBall arc-registers/s (visited arcs x registers per counter, run to stabilisation)",
        "value": work / elapsed,
        "unit": "arc-registers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "web-scale HyperBall (n=1e7, Zipf(2.1) out-degree on [3,10000], "
                               "+-1000 window 70% / preferential attachment 30%, systolic frontier)",
                   "nodes": g.num_nodes, "arcs": g.num_arcs, "log2m": args.log2m,
                   "register_bits": args.reg_bits, "hash": "jenkins(42)", "iterations": iters,
                   "heavy_row_threshold": chosen,
                   "outputs": ["neighborhood_function", "sum_of_distances", "harmonic_centrality"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "union step (union_kernel: heavy-row phase, then chunks; one "
                               "launch per iteration)", "launches": launches,
                     "avg_launch_ms": union_ms / max(1, launches),
                     "algorithmic_bytes_per_launch": bytes_total / max(1, launches),
                     "dense_iterations_frac": (dense_bytes / (dense_ms * 1e-3) / 1e9 / PEAK_HBM_GBS
                                               if dense_ms > 0 else None),
                     "bytes_model": "slabs actually gathered (delta gathering); the model with C "
                                    "bytes per visited arc gives no_delta_model_frac",
                     "no_delta_model_frac": (bytes_no_delta / (union_ms * 1e-3) / 1e9 / PEAK_HBM_GBS
                                             if union_ms > 0 else None)},
    }
    orc = None
    if not args.no_cpu_baseline:
        orc = _load("hb_oracle_py", os.path.join(REPO, "oracle", "oracle.py"))
        result["cpu_baseline"] = cpu_baseline(orc, hb, g, gt, args, args.log2m, args.reg_bits)
    hb.close()
    del g, gt
    if not args.no_small_counters:
        result["small_counters"] = bench_small_counters(hbm, args)
    if args.batch:
        result["batched"] = bench_batch(hbm, args)
    if not args.no_precision:
        result["precision"] = bench_precision(hbm, args)
    if not args.no_large:
        result["large"] = bench_large(hbm, args, compare_transposes=False)
    if not args.no_small:
        result["uniform_1e5"] = small_uniform(hbm, orc, args)
    if not args.no_host_counters:
        result["host_counters"] = bench_host_counters(hbm, args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
