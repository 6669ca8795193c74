#!/usr/bin/env python3
"""HyperBall throughput on one MI355X.

Workload (default): R-MAT scale 24, edge factor 16 (a, b, c, d = .57, .19, .19, .05),
seeded, symmetrized, deduplicated, self-loops removed; HLL++ precision p = 7
(128 u8 registers per node); run to the reference's convergence rule.  One step
= one complete HyperBall run (register init + every iteration) over the graph,
which is staged in HBM before the timed region.  `value` is iterations per
second over the K timed steps; GTEPS and the roofline of the dominant kernel
are reported beside it, and the CPU baseline is the oracle's faithful
restatement of the reference loop (packed 6-bit registers, every node every
iteration, static thread blocks) timed on this host's cores over a bounded
sample of the same workload.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

CONFIGS = {
    # name: (generator kwargs, p, max_iters)
    "rmat24": (dict(kind="rmat", scale=24, edge_factor=16, seed=1), 7, 0),
    "rmat20": (dict(kind="rmat", scale=20, edge_factor=16, seed=1), 7, 0),
    "er": (dict(kind="er", n=65536, samples=1 << 20, seed=1), 6, 0),
    "kmer": (dict(kind="kmer", genome_len=4_000_000, copies=8, sub_rate=0.005, k=21, seed=1), 6, 64),
    "batch64": (dict(kind="batch", count=64, seed=1), 0, 0),
    "surrogate": (dict(kind="surrogate", train_graphs=32, train_scale=20, seed=1), 7, 0),
    "exact": (dict(kind="exact", scale=20, edge_factor=16, seed=1, sources=2048), 0, 0),
    "ingest": (dict(kind="ingest", scale=22, edge_factor=16, seed=1), 0, 0),
    "many64": (dict(kind="many", count=64, seed=1, threads=8), 0, 0),
    "walks": (dict(kind="walks", scale=24, edge_factor=16, seed=1), 7, 0),
    "copy": (dict(kind="copy", gib=1), 0, 0),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_graph(spec, cache=None):
    import numpy as np

    from sparkx_amd import Graph
    if cache and os.path.exists(cache + ".off.npy"):
        return Graph(np.load(cache + ".off.npy"), np.load(cache + ".nbr.npy", mmap_mode="r"), "cached")
    g = _generate(spec)
    if cache:
        np.save(cache + ".off.npy", g.offsets)
        np.save(cache + ".nbr.npy", g.neighbours)
    return g


def _generate(spec):
    from sparkx_amd import Graph
    kind = spec["kind"]
    if kind == "rmat":
        return Graph.rmat(spec["scale"], spec["edge_factor"], 0.57, 0.19, 0.19, seed=spec["seed"])
    if kind == "er":
        return Graph.erdos_renyi(spec["n"], spec["samples"], seed=spec["seed"])
    return Graph.kmer_dbg(spec["genome_len"], spec["copies"], spec["sub_rate"], spec["k"], seed=spec["seed"])


def algorithmic_bytes(slot, arcs, nodes, items, p, gathered, first_nodes=0):
    """Bytes a kernel slot must move: per recomputed node its offsets (8), own
    registers read + new registers written (2R), estimate read + write (8), dist/inv
    read + write (32); per neighbour id scanned 4; per neighbour register block read R
    (`gathered`: none in an iteration 1 that hashes ids, and only the neighbours that
    changed outside dense iterations).  Hub chunks add one partial sketch write per
    task.  Where iteration 1 hashes (p <= 10, every bin but one that k_union_rows runs)
    the node's own block is not read either (first_nodes)."""
    R = 1 << p
    b = nodes * (8 + 2 * R + 8 + 32) + arcs * 4 + gathered * R - first_nodes * R
    if slot == 3:
        b += items * R
    return b


def kernel_name(slot, p, team, layout):
    """The kernel that does a slot's work in the dense iterations.  In the chain layout at
    p = 5 and 6 the single bin of one-copy teams (an id range from node 0) is run by
    k_union_rows after the first iteration, which k_union still runs."""
    if slot == 0 and layout == "chain" and p in (5, 6) and team == (1 << p) // 16:
        return f"hb::k_union_rows<{p}>"
    if slot < 3:
        return f"hb::k_union<{p}, {team}>"
    return f"hb::k_chunk<{p}>" if slot == 3 else f"hb::k_merge<{p}>"


def measured_traffic(kernel, workload, p, layout):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/*_traffic.txt, copy-calibrated) of the same
    workload, precision and node layout, or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.txt"))):
        kv = {}
        with open(path) as fh:
            for line in fh:
                if line.startswith("#") or ":" not in line:
                    continue
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        if (kv.get("kernel") == kernel and kv.get("workload") == workload and kv.get("p") == str(p)
                and kv.get("layout") == layout):
            return int(kv["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def host_cpu():
    """CPU model and the threads this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(g, p, threads, iters=3):
    """The oracle (C restatement of hyper_ball.rs:754-874 with the reference's packed
    6-bit registers and static contiguous node blocks) on `threads` host threads for
    `iters` iterations, each timed (value: 1 / median); and one thread on a
    pseudo-random sample of 1/64 of the nodes, scaled by the arcs it covers (three
    timings, median).  The reference recomputes every node in every iteration
    (hyper_ball.rs:803-831), so every iteration does the same union work and one
    iteration's time stands for each of them."""
    import numpy as np

    from oracle import hb_oracle as O
    off = g.offsets.astype(np.uint64)
    r = O.hyperball(off, g.neighbours, p, threads=threads, packed=True, max_iters=iters)
    per = list(r.iter_seconds)
    med = float(np.median(per))
    secs, nodes, arcs = O.node_sample(off, g.neighbours, p, 64, 3)
    t1 = float(np.median(secs)) * g.width() / max(arcs, 1)
    return {"value": round(1.0 / med, 5), "unit": "iter/s", "cores": threads, "kind": "port",
            "sample": f"iterations 1..{len(per)} of the same workload on {threads} threads, each timed "
                      f"({', '.join(f'{x:.2f}' for x in per)} s; value = 1 / median; register init "
                      f"{r.init_seconds:.1f}s excluded); every reference iteration recomputes every node, so "
                      f"each stands for any of them",
            "per_iteration_s": [round(x, 3) for x in per],
            "threads_1": {"value": round(1.0 / t1, 5), "unit": "iter/s",
                          "sample": f"one thread, first-iteration work of {nodes} pseudo-random nodes "
                                    f"({arcs} arcs, 1/64 of the nodes), median of 3 timings "
                                    f"({', '.join(f'{x:.3f}' for x in secs)} s) scaled by arcs"},
            "host": host_cpu()}


def bench_surrogate(args, spec, eval_graphs=None, graphs=None):
    """Learned surrogate: MLP trained on GPU (Adam, lr 1e-3, MSE) on seeded R-MAT
    scale-20 graphs labelled by full p=7 HyperBall (inputs and target log(1+H)
    standardised per graph over linked nodes);
    evaluated on R-MAT-24 and the k-mer graph against full HyperBall: Spearman rho
    (reference tie semantics), top-1% overlap, GPU time of register init + 4
    iterations (p=4) + features + MLP against register init + the full run."""
    import numpy as np

    from sparkx_amd import Graph, HyperBall, metrics
    from sparkx_amd.surrogate import Surrogate, labelled_graph, standardize, walk_features
    t0 = time.perf_counter()
    xs, ys = [], []
    xws = {2: [], 3: []}
    for i in range(spec["train_graphs"]):
        g = Graph.rmat(spec["train_scale"], 16, seed=spec["seed"] * 1000 + i)
        x, y, _, _, _ = labelled_graph(g)
        linked = g.degrees() > 0
        with HyperBall(g, 4, _run=False) as h:
            for passes in xws:
                xws[passes].append(standardize(walk_features(h, passes)[0], linked)[linked])
        xs.append(x[linked])
        ys.append(y[linked])
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    log(f"surrogate: {len(xs)} training graphs, {x.shape[0]} samples, labelled in {time.perf_counter() - t0:.1f}s")
    s = Surrogate(seed=spec["seed"])
    t0 = time.perf_counter()
    losses = s.train(x, y, epochs=4, batch=8192, lr=1e-3, seed=spec["seed"])
    train_s = time.perf_counter() - t0
    log(f"trained (snapshot inputs) in {train_s:.1f}s, epoch losses {[round(float(v), 5) for v in losses]}")
    walk_nets, losses_w = {}, {}
    for passes, parts in xws.items():
        walk_nets[passes] = Surrogate(seed=spec["seed"])
        losses_w[passes] = walk_nets[passes].train(np.concatenate(parts), y, epochs=4, batch=8192, lr=1e-3,
                                                   seed=spec["seed"])
        log(f"trained (walk inputs, {passes} passes), epoch losses {[round(float(v), 5) for v in losses_w[passes]]}")
    del xs, xws, ys, x, y
    results = {}
    evals = eval_graphs or [("rmat24", CONFIGS["rmat24"][0], 7, 0), ("kmer", CONFIGS["kmer"][0], 6, 64)]
    for name, gspec, p_full, mi in evals:
        g = (graphs or {}).get(name) or make_graph(gspec, args.graph_cache if name == "rmat24" else None)
        with HyperBall(g, p_full, max_iters=mi) as full:
            full.run(max_iters=mi)  # warm
            st = full.run(max_iters=mi)
            full_ms = st.ms_total + st.ms_init
            truth = full.harmonic_geodesic_distances()
        # the same run with a finer sketch (relative error 1.04 / sqrt(2^p) falls 4x): how
        # much of the target ranking is the configured sketch's own estimation noise
        p_fine = min(p_full + 4, 10)
        with HyperBall(g, p_fine, max_iters=mi) as fine:
            truth_fine = fine.harmonic_geodesic_distances()
        with HyperBall(g, 4, _run=False) as h:
            s.rank(h)  # warm
            pred, _, sur_ms = s.rank(h)
            trunc = h.harmonic_geodesic_distances()
            walk_out = {}
            for passes, net in walk_nets.items():
                net.rank_walks(h, passes)  # warm
                walk_out[passes] = net.rank_walks(h, passes)
        linked = g.degrees() > 0
        arcs = g.width()
        t0 = time.perf_counter()
        m_all = metrics.evaluate(pred, truth, 0.01)  # on the GPU, reference tie semantics
        m_lnk = metrics.evaluate(pred[linked], truth[linked], 0.01)
        b_lnk = metrics.evaluate(trunc[linked], truth[linked], 0.01)
        b_all = metrics.evaluate(trunc, truth, 0.01)
        eval_s = time.perf_counter() - t0
        fine_lnk = metrics.evaluate(truth[linked], truth_fine[linked], 0.01)
        f_lnk = metrics.evaluate(pred[linked], truth_fine[linked], 0.01)
        walks = {}
        for passes, (pred_w, _, walk_ms) in walk_out.items():
            w_all = metrics.evaluate(pred_w, truth, 0.01)
            w_lnk = metrics.evaluate(pred_w[linked], truth[linked], 0.01)
            wf_lnk = metrics.evaluate(pred_w[linked], truth_fine[linked], 0.01)
            walks[f"{passes}_passes"] = {
                "surrogate_ms": round(walk_ms, 3), "time_ratio": round(walk_ms / full_ms, 4),
                "spearman_all_nodes": round(w_all["spearman"], 5),
                "spearman_linked_nodes": round(w_lnk["spearman"], 5),
                "spearman_linked_vs_finer_sketch": round(wf_lnk["spearman"], 5),
                "top1pct_overlap": round(w_all["top_overlap"], 5)}
        del g
        results[name] = {
            "nodes": int(linked.size), "arcs": int(arcs), "full_p": p_full, "full_iterations": int(st.iterations),
            "full_ms": round(full_ms, 3), "surrogate_ms": round(sur_ms, 3),
            "time_ratio": round(sur_ms / full_ms, 4),
            "spearman_all_nodes": round(m_all["spearman"], 5),
            "spearman_linked_nodes": round(m_lnk["spearman"], 5),
            "top1pct_overlap": round(m_all["top_overlap"], 5),
            "baseline_4iter_spearman_linked": round(b_lnk["spearman"], 5),
            "baseline_4iter_top1pct_overlap": round(b_all["top_overlap"], 5),
            "evaluation_seconds_4_metric_sets": round(eval_s, 3),
            "walks": walks,
            "finer_sketch": {"p": p_fine,
                             "spearman_linked_target_vs_finer": round(fine_lnk["spearman"], 5),
                             "spearman_linked_snapshot_net_vs_finer": round(f_lnk["spearman"], 5)},
        }
        log(f"surrogate on {name}: {results[name]}")
    s.close()
    for net in walk_nets.values():
        net.close()
    return {"train_graphs": spec["train_graphs"], "train_scale": spec["train_scale"],
            "train_seconds": round(train_s, 2), "epoch_losses": [round(float(v), 6) for v in losses],
            "walks_epoch_losses": {f"{k}_passes": [round(float(v), 6) for v in w] for k, w in losses_w.items()},
            "eval": results}


def bench_batch(args, spec):
    """64 independent graphs (32 ER + 32 R-MAT, n log-uniform in [2^12, 2^18], average
    degree U[4, 32], p uniform in {4..8}) advanced together by the batched engine: per
    iteration one persistent work-queue launch per (precision, team width) and hub
    chunks, hipGraph-replayed, schedule autotuned.  value = graph-iterations per second
    (wall time of each complete run); the same graphs run as 64 engine contexts
    (hb_run_many) are reported beside it."""
    from sparkx_amd.batch import HyperBallBatch, batched_workload
    t0 = time.perf_counter()
    items = batched_workload(spec["count"], seed=spec["seed"])
    nodes = sum(g.size() for g, _ in items)
    arcs = sum(g.width() for g, _ in items)
    log(f"batch: {len(items)} graphs, {nodes} nodes, {arcs} arcs, built in {time.perf_counter() - t0:.1f}s")
    bt = HyperBallBatch(items)
    knobs, tms = bt.autotune()
    log(f"autotuned schedule (small_max, mid_max, tile_arcs) = {knobs}, trial {tms:.2f} ms")
    for _ in range(args.warmup):
        bt.run()
    walls, its, arcs_g, alg, dev_ms = [], 0, 0, 0, 0.0
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = bt.run()
        walls.append(time.perf_counter() - t0)
        its += st.iterations_total
        arcs_g += st.arcs_gathered
        alg += st.alg_bytes
        dev_ms += st.ms_total
        log(f"step {i}: {st.iterations_total} graph-iterations (longest {st.iterations_max}) in "
            f"{walls[-1] * 1e3:.2f} ms wall, {st.ms_total:.2f} ms device, {st.launches} launches")
    total = sum(walls)
    achieved = alg / (dev_ms * 1e-3) / 1e9 if dev_ms else 0.0
    # the same graphs as one engine context and stream per graph (hb_run_many, 8 host threads)
    from sparkx_amd.batch import HyperBallMany
    many = HyperBallMany(items, threads=8)
    many.run()
    m_walls, m_its = [], 0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        mst = many.run()
        m_walls.append(time.perf_counter() - t0)
        m_its += sum(x.iterations for x in mst)
    many.close()
    log(f"same graphs as 64 contexts (hb_run_many, 8 threads): {sum(m_walls) / args.steps * 1e3:.2f} ms per step")
    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np

        from oracle import hb_oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        threads = max(1, min(threads, os.cpu_count() or 1))
        loop, n_it = 0.0, 0
        for g, p in items[::8]:  # every 8th graph, run to convergence
            r = O.hyperball(g.offsets.astype(np.uint64), g.neighbours, p, threads=threads, packed=True)
            loop += r.loop_seconds
            n_it += r.iterations
        cpu = {"value": round(n_it / loop, 3), "unit": "graph-iter/s", "cores": threads, "kind": "port",
               "sample": f"8 of the 64 graphs (every 8th) run to convergence, {n_it} iterations in "
                         f"{loop:.1f}s, packed 6-bit registers, register init excluded"}
    out = {
        "metric": "Batched HyperBall graph-iterations/s (64 graphs, mixed p=4..8)",
        "value": round(its / total, 2), "unit": "graph-iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "batch64", "graphs": len(items), "nodes": nodes, "arcs": arcs,
                   "schedule": list(knobs), "tiles": int(st.tiles)},
        "gteps": round(arcs_g / total / 1e9, 3),
        "ms_device_per_step": round(dev_ms / args.steps, 3),
        "many_contexts": {"graph_iter_per_s": round(m_its / sum(m_walls), 2),
                          "ms_per_step": round(sum(m_walls) / args.steps * 1e3, 3),
                          "how": "one engine context and stream per graph, hb_run_many from 8 host threads"},
        "roofline": {"bound": "hbm", "kernel": "hb::k_batch_nodes / k_batch_chunk / k_batch_merge (whole iteration)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
    }
    bt.close()
    print(json.dumps(out), flush=True)


def bench_walks(args, spec):
    """Device time of the walk-count surrogate (hb_surrogate_walks: the gather passes,
    the inputs, standardisation, inference) on R-MAT-24 for 2 and 3 passes, against
    the full p=7 run (register init + iterations).  The network's weights do not change
    the work, so it is left at its seeded initialisation; rankings are measured by
    --config surrogate."""
    from sparkx_amd import HyperBall
    from sparkx_amd.surrogate import Surrogate
    g = make_graph(dict(kind="rmat", scale=spec["scale"], edge_factor=spec["edge_factor"], seed=spec["seed"]),
                   args.graph_cache)
    log(f"walks: R-MAT {spec['scale']} n={g.size()} arcs={g.width()}")
    with HyperBall(g, 7) as full:
        st = full.run()
        full_ms = st.ms_total + st.ms_init
    net = Surrogate(seed=spec["seed"])
    out = {"metric": "walk-count surrogate device ms (R-MAT-24)", "unit": "ms", "higher_is_better": False,
           "full_run_ms": round(full_ms, 3), "config": {"workload": "walks", **spec, "nodes": g.size(),
                                                       "arcs": g.width()}}
    with HyperBall(g, 4, _run=False) as h:
        for passes in (2, 3):
            for _ in range(args.warmup):
                net.rank_walks(h, passes)
            ms = [net.rank_walks(h, passes)[2] for _ in range(args.steps)]
            avg = sum(ms) / len(ms)
            log(f"{passes} passes: {[round(v, 3) for v in ms]} ms")
            out[f"{passes}_passes"] = {"ms": round(avg, 3), "time_ratio": round(avg / full_ms, 4)}
    net.close()
    print(json.dumps(out), flush=True)


def bench_copy(args, spec):
    """Plain device-to-device copies of 1 GiB (the HIP runtime's copy kernel): the
    reference point that rocprofv3's FETCH_SIZE / WRITE_SIZE readings of the HyperBall
    kernels are scaled by (each copy reads and writes exactly `gib` GiB)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = spec["gib"] << 30
    src, dst = C.c_void_p(), C.c_void_p()
    ev = [C.c_void_p(), C.c_void_p()]
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    try:
        assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
        for e in ev:
            assert hip.hipEventCreate(C.byref(e)) == 0
        for _ in range(args.warmup):
            assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # device to device
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipEventRecord(ev[0], None) == 0
        for _ in range(args.steps):
            assert hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None) == 0
        assert hip.hipEventRecord(ev[1], None) == 0
        assert hip.hipEventSynchronize(ev[1]) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1]) == 0
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    per = ms.value / max(args.steps, 1)
    out = {"metric": "HBM copy (read + write) GB/s", "value": round(2 * nbytes / per / 1e6, 1), "unit": "GB/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
           "data": "synthetic", "config": {"workload": "copy", "bytes_read_per_copy": nbytes,
                                           "bytes_written_per_copy": nbytes},
           "roofline": {"bound": "hbm", "achieved": round(2 * nbytes / per / 1e6, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(2 * nbytes / per / 1e6 / HBM_PEAK_GBS, 4), "traffic": None},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def bench_many(args, spec):
    """The same 64 graphs (32 ER + 32 R-MAT, n log-uniform in [2^12, 2^18], average degree
    U[4, 32], p uniform in {4..8}) run to convergence together: one engine context and
    stream per graph, driven by `threads` host threads (hb_run_many).  value =
    graph-iterations per second of the whole batch; the same contexts run one after the
    other (one thread) are reported beside it."""
    from sparkx_amd.batch import HyperBallMany, batched_workload
    t0 = time.perf_counter()
    items = batched_workload(spec["count"], seed=spec["seed"])
    nodes = sum(g.size() for g, _ in items)
    arcs = sum(g.width() for g, _ in items)
    log(f"many: {len(items)} graphs, {nodes} nodes, {arcs} arcs, built in {time.perf_counter() - t0:.1f}s")
    many = HyperBallMany(items, threads=spec["threads"])
    for _ in range(args.warmup):
        many.run()
    walls, its, arcs_g = [], 0, 0
    for i in range(args.steps):
        t0 = time.perf_counter()
        stats = many.run()
        walls.append(time.perf_counter() - t0)
        its += sum(st.iterations for st in stats)
        arcs_g += sum(st.arcs_gathered for st in stats)
        log(f"step {i}: {sum(st.iterations for st in stats)} graph-iterations (longest "
            f"{max(st.iterations for st in stats)}) in {walls[-1] * 1e3:.2f} ms wall")
    total = sum(walls)
    many.threads = 1
    t0 = time.perf_counter()
    many.run()
    seq = time.perf_counter() - t0
    many.threads = spec["threads"]
    log(f"one after the other: {seq * 1e3:.2f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np

        from oracle import hb_oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        threads = max(1, min(threads, os.cpu_count() or 1))
        loop, n_it = 0.0, 0
        for g, p in items[::8]:  # every 8th graph, run to convergence
            r = O.hyperball(g.offsets.astype(np.uint64), g.neighbours, p, threads=threads, packed=True)
            loop += r.loop_seconds
            n_it += r.iterations
        cpu = {"value": round(n_it / loop, 3), "unit": "graph-iter/s", "cores": threads, "kind": "port",
               "sample": f"8 of the 64 graphs (every 8th) run to convergence, {n_it} iterations in "
                         f"{loop:.1f}s, packed 6-bit registers, register init excluded"}
    out = {
        "metric": "Batched HyperBall graph-iterations/s (64 graphs, mixed p=4..8)",
        "value": round(its / total, 2), "unit": "graph-iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "many64", "graphs": len(items), "nodes": nodes, "arcs": arcs,
                   "host_threads": spec["threads"]},
        "gteps": round(arcs_g / total / 1e9, 3),
        "ms_one_after_the_other": round(seq * 1e3, 3),
        "roofline": None, "cpu_baseline": cpu,
    }
    many.close()
    print(json.dumps(out), flush=True)


def bench_exact(args, spec):
    """Exact BFS closeness (exact.rs semantics, bfs.rs source quirk) of the first
    `sources` nodes of R-MAT scale 20: 64 bit-lane searches per pass on the GPU.
    value = sources per second; TEPS counts every arc once per source."""
    import numpy as np

    from sparkx_amd.exact import exact_bfs
    g = make_graph(dict(kind="rmat", scale=spec["scale"], edge_factor=spec["edge_factor"], seed=spec["seed"]),
                   args.graph_cache)
    k = spec["sources"]
    log(f"exact: R-MAT {spec['scale']} n={g.size()} arcs={g.width()}, {k} sources")
    for _ in range(args.warmup):
        exact_bfs(g, (0, 64), inv_sum=True)
    walls = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        reach, total, inv = exact_bfs(g, (0, k), inv_sum=True)
        walls.append(time.perf_counter() - t0)
        log(f"step {i}: {k} sources in {walls[-1]:.2f}s (host-timed, graph staging included)")
    tot = sum(walls)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import hb_oracle as O
        ns = 4
        t0 = time.perf_counter()
        r, t, v = O.exact_bfs(g.offsets.astype(np.uint64), g.neighbours.astype(np.uint32), 0, ns)
        dt = time.perf_counter() - t0
        assert np.array_equal(r, reach[:ns]) and np.array_equal(v, inv[:ns])
        cpu = {"value": round(ns / dt, 3), "unit": "sources/s", "cores": 1, "kind": "port",
               "sample": f"first {ns} sources, one queue BFS each (bfs.rs:199-234) plus the node-order "
                         f"harmonic sum, {dt:.1f}s"}
    out = {"metric": f"Exact BFS centrality sources/s (R-MAT-{spec['scale']})", "value": round(args.steps * k / tot, 2),
           "unit": "sources/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(tot / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u64", "data": "synthetic",
           "config": {"workload": "exact", **spec, "nodes": g.size(), "arcs": g.width()},
           "gteps": round(args.steps * k * g.width() / tot / 1e9, 3),
           "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_ingest(args, spec):
    """MatrixMarket ingestion on the GPU (GraphCache::from_mtx_file semantics): the
    symmetric pattern text of R-MAT scale `scale` (one line per undirected edge) parsed,
    emitted, sorted and turned into the staged CSR.  value = text GB/s over the device
    time of that pipeline (text already in HBM); the host-timed call, which adds the
    PCIe copy of the text, the symmetry check and the degree layout, is reported
    beside it."""
    import ctypes as C

    import numpy as np

    from sparkx_amd import _lib
    from sparkx_amd.ingest import write_mtx
    g = make_graph(dict(kind="rmat", scale=spec["scale"], edge_factor=spec["edge_factor"], seed=spec["seed"]),
                   args.graph_cache)
    t0 = time.perf_counter()
    text = write_mtx(g, symmetric=True)
    log(f"ingest: R-MAT {spec['scale']} n={g.size()} arcs={g.width()}, text {len(text) / 1e9:.3f} GB "
        f"written in {time.perf_counter() - t0:.1f}s")
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), 0))
    info = _lib.HbMtxInfo()
    try:
        for _ in range(args.warmup):
            _lib.check(L.hb_stage_mtx(ctx, text, len(text), C.byref(info)))
        dev_ms, walls = [], []
        for i in range(args.steps):
            t0 = time.perf_counter()
            _lib.check(L.hb_stage_mtx(ctx, text, len(text), C.byref(info)))
            walls.append(time.perf_counter() - t0)
            dev_ms.append(info.ms_device)
            log(f"step {i}: {info.entries} entries -> {info.arcs} arcs, device {info.ms_device:.2f} ms, "
                f"call {walls[-1] * 1e3:.1f} ms")
        assert info.arcs == g.width()
        off = np.empty(g.size() + 1, np.uint64)
        _lib.check(L.hb_read_csr(ctx, _lib.ptr(off, C.c_uint64), None))
        assert np.array_equal(off, g.offsets.astype(np.uint64))
    finally:
        L.hb_destroy(ctx)
    dev_s = sum(dev_ms) / 1e3
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import hb_oracle as O
        cut = text.rfind(b"\n", 0, 64 << 20) + 1  # a prefix of whole lines
        t0 = time.perf_counter()
        _, _, ci = O.parse_mtx(text[:cut])
        dt = time.perf_counter() - t0
        cpu = {"value": round(cut / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"first {cut / 1e6:.0f} MB ({ci.entries} entries) of the same text, C restatement of "
                         f"parser_mtx (header, entries, emission, degree count, scatter, per-list sort), {dt:.2f}s"}
        log(f"cpu baseline: {cpu['value']} GB/s")
    out = {"metric": f"MatrixMarket ingest GB/s (R-MAT-{spec['scale']} symmetric text)",
           "value": round(args.steps * len(text) / dev_s / 1e9, 3), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dev_s / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "ingest", **spec, "nodes": g.size(), "arcs": g.width(), "text_bytes": len(text),
                      "entries": int(info.entries)},
           "entries_per_s": round(args.steps * info.entries / dev_s, 1),
           "call_ms_with_pcie_and_staging": round(sum(walls) / args.steps * 1e3, 2),
           "roofline": None, "cpu_baseline": cpu,
           "speedup_vs_cpu": round(args.steps * len(text) / dev_s / 1e9 / cpu["value"], 1) if cpu else None}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="rmat24", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-iters", type=int, default=3, help="iterations timed in the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="leave out the surrogate field of the R-MAT-24 line (trained after the timed steps)")
    ap.add_argument("--graph-cache", default=None, help="path prefix to save/load the generated CSR (.npy)")
    ap.add_argument("--layout", default="auto", choices=["auto", "degree", "chain", "input"],
                    help="device node numbering (results are identical)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of HyperBall is a design, not implemented here", file=sys.stderr)
        sys.exit(2)

    from sparkx_amd import HyperBall, _lib

    spec, p, max_iters = CONFIGS[args.config]
    if spec["kind"] == "batch":
        return bench_batch(args, spec)
    if spec["kind"] == "exact":
        return bench_exact(args, spec)
    if spec["kind"] == "ingest":
        return bench_ingest(args, spec)
    if spec["kind"] == "many":
        return bench_many(args, spec)
    if spec["kind"] == "copy":
        return bench_copy(args, spec)
    if spec["kind"] == "walks":
        return bench_walks(args, spec)
    if spec["kind"] == "surrogate":
        print(json.dumps(bench_surrogate(args, spec)), flush=True)
        return None
    t0 = time.perf_counter()
    g = make_graph(spec, args.graph_cache)
    log(f"graph {args.config}: n={g.size()} arcs={g.width()} max_deg={int(g.degrees().max())} "
        f"built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    h = HyperBall(g, p, _run=False, layout=args.layout)
    log(f"staged CSR into HBM in {time.perf_counter() - t0:.2f}s (pinned staging, validation, symmetry "
        f"check and the {args.layout} node layout)")

    for i in range(args.warmup):
        st = h.run(max_iters=max_iters)
        log(f"warmup {i}: {st.iterations} iterations, {st.ms_total:.1f} ms device")

    walls, iters, arcs_g = [], 0, 0
    slot_ms = [0.0] * _lib.HB_NKERNELS
    slot_launch = [0] * _lib.HB_NKERNELS
    slot_bytes = [0] * _lib.HB_NKERNELS
    per_iter_ms = None
    st = None
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = h.run(max_iters=max_iters, ktiming=True)
        walls.append(time.perf_counter() - t0)
        iters += st.iterations
        arcs_g += st.arcs_gathered
        for s in range(_lib.HB_NKERNELS):
            slot_ms[s] += st.kernel_ms[s]
            slot_launch[s] += st.kernel_launches[s]
            slot_bytes[s] += algorithmic_bytes(s, st.kernel_arcs[s], st.kernel_nodes[s],
                                               st.kernel_items[s] * st.kernel_launches[s], p,
                                               st.kernel_gathered[s], st.kernel_first_nodes[s])
        per_iter_ms = [round(float(x), 3) for x in st.iter_ms[:st.iterations]]
        log(f"step {i}: {st.iterations} iterations in {walls[-1] * 1e3:.1f} ms wall "
            f"({st.ms_total:.1f} ms device), per-iteration ms {per_iter_ms}")
    total = sum(walls)
    value = iters / total
    dom = max(range(_lib.HB_NKERNELS), key=lambda s: slot_ms[s])
    dom_avg_ms = slot_ms[dom] / max(slot_launch[dom], 1)
    dom_bytes_per_launch = slot_bytes[dom] / max(slot_launch[dom], 1)
    achieved = dom_bytes_per_launch / (dom_avg_ms * 1e-3) / 1e9 if dom_avg_ms > 0 else 0.0
    kernels = {}
    for s in range(_lib.HB_NKERNELS):
        if slot_launch[s]:
            kernels[_lib.KERNEL_SLOTS[s]] = {
                "launches": slot_launch[s], "avg_ms": round(slot_ms[s] / slot_launch[s], 4),
                "team": int(st.kernel_team[s]), "items": int(st.kernel_items[s]),
                "alg_bytes_per_launch": int(slot_bytes[s] / slot_launch[s]),
                # algorithmic bytes per second; it can exceed the HBM peak where gathers hit
                # in L2 / the memory-side cache (the hub chunks do)
                "GBps_algorithmic": round(slot_bytes[s] / (slot_ms[s] * 1e-3) / 1e9, 1) if slot_ms[s] else None}
    layout_used = {v: k for k, v in HyperBall.LAYOUTS.items()}[int(st.layout)]
    dom_kernel = kernel_name(dom, p, int(st.kernel_team[dom]), layout_used)
    traffic, traffic_src = measured_traffic(dom_kernel, args.config, p, layout_used)
    all_ms = sum(slot_ms)
    iter_gbs = sum(slot_bytes) / (all_ms * 1e-3) / 1e9 if all_ms else 0.0

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host_cpu()["usable"]
        threads = max(1, min(threads, host_cpu()["usable"]))
        log(f"cpu baseline: oracle, packed 6-bit registers, {threads} threads, {args.cpu_iters} iterations, "
            "then one thread on a 1/64 node sample")
        cpu = cpu_baseline(g, p, threads, args.cpu_iters)
        log(f"cpu baseline: {cpu['value']} iter/s on {threads} threads, {cpu['threads_1']['value']} iter/s on one "
            f"({cpu['host']['cpu']})")

    out = {
        "metric": "HyperBall iter/s at R-MAT-24 p=7 (to convergence)" if args.config == "rmat24"
                  else f"HyperBall iter/s ({args.config}, p={p})",
        "value": round(value, 4),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.config, **{k: v for k, v in spec.items()}, "p": p,
                   "nodes": g.size(), "arcs": g.width(), "iterations_per_step": st.iterations,
                   "max_iters": max_iters},
        "layout": args.layout,
        "layout_used": layout_used,
        "layout_build_ms": round(float(st.ms_layout), 2),
        "gteps": round(arcs_g / total / 1e9, 3),
        "gteps_dense_equivalent": round(iters * g.width() / total / 1e9, 3),
        "iteration_GBps_alg": round(iter_gbs, 1),
        "per_iteration_ms": per_iter_ms,
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "alg_bytes_per_launch": int(dom_bytes_per_launch),
                     "avg_launch_ms": round(dom_avg_ms, 4)},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    h.close()
    if args.config == "rmat24" and not args.no_surrogate:
        out["surrogate"] = surrogate_on_run(args, g)
    print(json.dumps(out), flush=True)


def surrogate_on_run(args, g):
    """The second half of the headline metric, after the timed steps: the learned
    surrogate as configured -- an MLP trained on 32 seeded R-MAT-20 graphs labelled by
    full p=7 runs -- with each input set (HyperBall snapshots N_1..N_4 with degree; walk
    counts from 2 or 3 gather passes), ranking the benchmark graph and the k-mer graph
    (t <= 64): Spearman rho (reference tie semantics, on the GPU), top-1% overlap and
    device time against the full run's.  `best` names the input set with the highest
    rho on the benchmark graph among those within 1/5 of the full run's time."""
    spec = CONFIGS["surrogate"][0]
    t0 = time.perf_counter()
    study = bench_surrogate(args, spec, graphs={"rmat24": g})
    rows = {}
    for name, r in study["eval"].items():
        sets = {"snapshots_N1_N4": {k: r[k] for k in ("surrogate_ms", "time_ratio", "spearman_linked_nodes",
                                                       "spearman_all_nodes", "top1pct_overlap")}}
        sets["snapshots_N1_N4"]["spearman_linked_vs_finer_sketch"] = \
            r["finer_sketch"]["spearman_linked_snapshot_net_vs_finer"]
        for k, w in r["walks"].items():
            sets[f"walks_{k}"] = dict(w)
        for v in sets.values():
            v["meets_target"] = bool(v["spearman_linked_nodes"] >= 0.95 and v["time_ratio"] <= 0.2)
        # the target's own rank agreement with the same run at a finer sketch: a ceiling on
        # what any ranking can reach against the configured-precision target
        rows[name] = {"full_ms": r["full_ms"], "full_iterations": r["full_iterations"], "inputs": sets,
                      "target_vs_finer_sketch": {"p": r["finer_sketch"]["p"], "spearman_linked_nodes":
                                                 r["finer_sketch"]["spearman_linked_target_vs_finer"]},
                      "no_network_4_iterations": {"spearman_linked_nodes": r["baseline_4iter_spearman_linked"],
                                                  "top1pct_overlap": r["baseline_4iter_top1pct_overlap"]}}
    fast = {k: v for k, v in rows["rmat24"]["inputs"].items() if v["time_ratio"] <= 0.2}
    best = max(fast, key=lambda k: fast[k]["spearman_linked_nodes"]) if fast else None
    log(f"surrogate study: best within 1/5 of the run on rmat24: {best} "
        f"({time.perf_counter() - t0:.1f}s with training)")
    out = {"train": f"{spec['train_graphs']} R-MAT-{spec['train_scale']} graphs labelled by full p=7 runs, "
                    "Adam lr 1e-3, 4 epochs, MSE on log(1 + H) standardised per graph",
           "target": "Spearman >= 0.95 in <= 1/5 of the full run's GPU time",
           "best": best, "eval": rows}
    if best:
        b = rows["rmat24"]["inputs"][best]
        out.update({"inputs": best, "ms": b["surrogate_ms"], "full_run_ms": rows["rmat24"]["full_ms"],
                    "time_ratio": b["time_ratio"], "spearman_linked_nodes": b["spearman_linked_nodes"],
                    "spearman_all_nodes": b["spearman_all_nodes"], "top1pct_overlap": b["top1pct_overlap"]})
    return out


if __name__ == "__main__":
    main()
