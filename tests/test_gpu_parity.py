"""GPU parity tests: the HIP engine (through the C ABI) against the CPU oracle
on the same seeded inputs.  Registers, f32 estimates, f64 accumulators,
iteration counts and per-iteration increase counts must be bit-identical."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
THREADS = min(8, os.cpu_count() or 1)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    from sparkx_amd import _lib
    _lib.load()
    import sparkx_amd
    return sparkx_amd


def test_tested_library_was_built_from_these_sources(hb):
    """Run with the GPU tests, on the machine that runs them: the library every test in
    this run calls was compiled from the source files in this tree."""
    from sparkx_amd import _lib, build
    digest = _lib.load().hb_build_digest().decode()
    say("library", _lib.LIB_PATH, "digest", digest)
    assert digest == build.source_digest()


def _oracle_run(oracle, g, p, **kw):
    return oracle.hyperball(g.offsets.astype(np.uint64), g.neighbours.astype(np.uint32), p,
                            threads=THREADS, **kw)


def _assert_same(hbobj, ref, oracle, p, ctx=""):
    from sparkx_amd import _lib
    st = hbobj.stats
    assert st.iterations == ref.iterations, ctx
    assert list(st.increases[:st.iterations]) == list(ref.increases), ctx
    assert bool(st.final_parity_latest) == ref.final_parity_latest, ctx
    n = hbobj.g.size()
    assert np.array_equal(hbobj.registers(_lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref), ctx
    assert np.array_equal(hbobj.registers(_lib.HB_REGS_LATEST).ravel(), ref.regs_last), ctx
    dist, inv = hbobj._accumulators()
    assert np.array_equal(dist, ref.dist), ctx
    assert np.array_equal(inv, ref.inv), ctx
    est = hbobj.estimates(_lib.HB_REGS_REFERENCE)
    assert np.array_equal(est, oracle.estimates(p, ref.regs_ref)), ctx
    assert est.size == n


@pytest.mark.parametrize("p", list(range(4, 17)))
def test_er_small_all_precisions(hb, oracle, p):
    n = 3000 if p <= 12 else 600
    g = hb.Graph.erdos_renyi(n, 8 * n, seed=10 + p)
    say("er_small", p, g.size(), g.width())
    ref = _oracle_run(oracle, g, p)
    with hb.HyperBall(g, p) as h:
        _assert_same(h, ref, oracle, p, f"p={p}")


def test_er_config_every_iteration(hb, oracle):
    """ER n=65,536, m=2^20 samples, p=6: registers and accumulators after every
    iteration t equal the oracle's (GPU truncated at t vs oracle trace)."""
    g = hb.Graph.erdos_renyi(65536, 1 << 20, seed=1)
    p = 6
    say("er_config", g.size(), g.width())
    trace = {}
    ref = _oracle_run(oracle, g, p, on_iter=lambda t, r, d, i: trace.__setitem__(t, (r, d, i)))
    from sparkx_amd import _lib
    with hb.HyperBall(g, p) as h:
        _assert_same(h, ref, oracle, p, "full run")
        for t in range(1, ref.iterations + 1):
            say("er_config iteration", t)
            st = h.run(max_iters=t)
            assert st.iterations == t
            regs, dist, inv = trace[t]
            assert np.array_equal(h.registers(_lib.HB_REGS_LATEST).ravel(), regs), t
            d, i = h._accumulators()
            assert np.array_equal(d, dist) and np.array_equal(i, inv), t


@pytest.mark.parametrize("sigma,k,fixture", [(20, 2, "exact_closeness_dbg_sigma20_k2.f64"),
                                             (4, 7, "exact_closeness_dbg_sigma4_k7.f64")])
@pytest.mark.parametrize("p", [8, 9, 10, 11, 12, 13])
def test_fixture_graph_closeness_and_metrics(hb, oracle, sigma, k, fixture, p):
    """The reference's hyperball_profile (main.rs:855-944) on the de Bruijn
    fixture graphs, whose CSR keeps MTX duplicates and self-loops."""
    off, nb = oracle.dbg_mtx_csr(sigma, k)
    g = hb.Graph(off, nb.astype(np.uint64), f"dbg{sigma}_{k}")
    say("dbg", sigma, k, p)
    ref = oracle.hyperball(off, nb, p, threads=THREADS)
    ex = np.fromfile(os.path.join(GOLDEN, fixture), dtype="<f8")
    with hb.HyperBall(g, p) as h:
        _assert_same(h, ref, oracle, p)
        c = h.compute_closeness_centrality(True)
        c_ref = oracle.centrality(oracle.CLOSENESS, True, oracle.estimates(p, ref.regs_ref), ref.dist, ref.inv)
        assert np.array_equal(c, c_ref)
        assert h.get_closeness_centrality(True) is c
        metrics = (oracle.mae(c, ex), oracle.mape(c, ex), oracle.spearman(c, ex))
        if sigma == 20 and p >= 10:
            # linear-counting regime: the reference's published numbers, to the last bit
            from tests.golden.pin_hll_variant import PUBLISHED
            assert metrics == (PUBLISHED[p]["mae"][0], PUBLISHED[p]["mape"][0], PUBLISHED[p]["rho"][0])
        # the same metrics on the GPU: rho bit-exact (ssd < 2^53 here, so the reference's
        # sequential sum is exact too), MAE / MAPE within 1e-12 relative
        from sparkx_amd import metrics as M
        m = M.evaluate(c, ex)
        assert m["spearman"] == metrics[2]
        assert abs(m["mae"] - metrics[0]) <= 1e-12 * metrics[0]
        assert abs(m["mape"] - metrics[1]) <= 1e-12 * metrics[1]


def test_schedules_frontier_dense_and_graph_capture_agree(hb, oracle):
    """Scheduling never changes results: dense vs frontier, hipGraph replay, and
    degree-bin / hub-chunk settings that route nodes through every kernel."""
    g = hb.Graph.rmat(13, 16, seed=9)
    p = 7
    say("rmat13", g.size(), g.width(), int(g.degrees().max()))
    ref = _oracle_run(oracle, g, p)
    variants = [dict(), dict(dense=True), dict(graph_capture=True), dict(ktiming=True),
                dict(schedule=(2, 16, 64)), dict(schedule=(1, 4, 16), dense=True),
                dict(schedule=(64, 64, 4096)), dict(schedule=(4, 100, 7), graph_capture=True)]
    for kw in variants:
        say("variant", kw)
        with hb.HyperBall(g, p, **kw) as h:
            _assert_same(h, ref, oracle, p, str(kw))
            if "schedule" in kw and kw["schedule"][1] < g.degrees().max():
                assert h.stats.kernel_items[3] > 0  # hub chunks were used
            if kw.get("ktiming"):
                # every executed iteration was timed once, each slot once per iteration it
                # launched, from events read at the host's usual checks
                st = h.stats
                it = st.iterations
                assert all(st.iter_ms[t] > 0 for t in range(it)) and st.iter_ms[it] == 0
                assert sum(st.iter_ms[:it]) <= st.ms_total * 1.001
                for s in range(5):
                    assert st.kernel_launches[s] == (it if st.kernel_items[s] else 0), s
                    assert (st.kernel_ms[s] > 0) == (st.kernel_items[s] > 0), s


def test_push_iterations_on_symmetric_graphs(hb, oracle):
    """Late iterations on a symmetric graph list the nodes to recompute from the
    changed nodes' neighbour lists; results equal the oracle and the dense run.
    A directed graph never takes that path."""
    g = hb.Graph.rmat(14, 16, seed=5)
    p = 7
    ref = _oracle_run(oracle, g, p)
    for kw in (dict(), dict(graph_capture=True), dict(schedule=(2, 16, 64))):
        say("push", kw)
        with hb.HyperBall(g, p, **kw) as h:
            _assert_same(h, ref, oracle, p, str(kw))
            assert h.stats.symmetric == 1
            assert h.stats.push_iterations >= 1, h.stats.push_iterations
            # iteration 1 hashes ids and later sparse iterations read only changed rows
            gathered = sum(h.stats.kernel_gathered[s] for s in range(5))
            scanned = sum(h.stats.kernel_arcs[s] for s in range(5))
            assert 0 < gathered < scanned - g.width() // 2, (gathered, scanned)
    rng = np.random.default_rng(3)
    n = 4000
    src = rng.integers(0, n, 12000)
    dst = rng.integers(0, n, 12000)
    order = np.lexsort((dst, src))
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(src, minlength=n), out=off[1:])
    gd = hb.Graph(off, dst[order].astype(np.uint64), "directed")
    ref = oracle.hyperball(off, dst[order].astype(np.uint32), 6)
    with hb.HyperBall(gd, 6) as h:
        _assert_same(h, ref, oracle, 6, "directed")
        assert h.stats.symmetric == 0 and h.stats.push_iterations == 0


def test_kmer_graph_low_degree_bins(hb, oracle):
    g = hb.Graph.kmer_dbg(60_000, 8, 0.005, 21, seed=4)
    p = 6
    say("kmer", g.size(), g.width())
    ref = _oracle_run(oracle, g, p, max_iters=40)
    with hb.HyperBall(g, p, max_iters=40) as h:
        _assert_same(h, ref, oracle, p)


def _csr(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in adj])
    nb = np.array([v for x in adj for v in x], np.uint64)
    return off, nb


@pytest.mark.parametrize("name", ["single_node", "isolated", "one_edge", "star", "path", "directed_cycle",
                                  "self_loops"])
def test_edge_case_graphs(hb, oracle, name):
    if name == "single_node":
        off, nb = _csr(1, [])
    elif name == "isolated":
        off, nb = _csr(50, [(3, 4), (4, 3)])
    elif name == "one_edge":
        off, nb = _csr(2, [(0, 1), (1, 0)])
    elif name == "star":
        n = 5000
        off, nb = _csr(n, [(0, i) for i in range(1, n)] + [(i, 0) for i in range(1, n)])
    elif name == "path":
        n = 300
        off, nb = _csr(n, [(i, i + 1) for i in range(n - 1)] + [(i + 1, i) for i in range(n - 1)])
    elif name == "directed_cycle":
        n = 64
        off, nb = _csr(n, [(i, (i + 1) % n) for i in range(n)])
    else:
        off, nb = _csr(4, [(0, 0), (0, 1), (1, 0), (1, 1), (2, 2), (2, 3), (3, 2), (3, 3), (3, 3)])
    g = hb.Graph(off, nb, name)
    for p in (4, 7):
        say("edge case", name, p)
        ref = oracle.hyperball(off, nb.astype(np.uint32), p)
        with hb.HyperBall(g, p, schedule=(2, 8, 16)) as h:
            _assert_same(h, ref, oracle, p, name)


def _ball_sketches(oracle, n, p, t):
    """u8 registers of the HLL sketch of every node's ball [u - t, u + t] on the path
    0 - 1 - ... - n-1: each node inserts its own id; a sketch is the register-wise max
    of its members' insertions (sliding-window maxima from a sparse table)."""
    m = 1 << p
    one = np.zeros((n, m), np.uint8)
    for u in range(n):
        idx, rank = oracle.index_rank(p, u)
        one[u, idx] = rank
    table = [one]
    while (1 << len(table)) <= 2 * t + 1:
        prev, h = table[-1], 1 << (len(table) - 1)
        table.append(np.maximum(prev[:-h], prev[h:]) if prev.shape[0] > h else prev[:0])
    lo = np.maximum(np.arange(n) - t, 0)
    hi = np.minimum(np.arange(n) + t, n - 1)
    k = np.floor(np.log2(hi - lo + 1)).astype(int)
    out = np.empty((n, m), np.uint8)
    for lev in np.unique(k):
        rows = np.flatnonzero(k == lev)
        tab = table[lev]
        out[rows] = np.maximum(tab[lo[rows]], tab[hi[rows] - (1 << lev) + 1])
    return out


def test_run_longer_than_65534_iterations(hb, oracle):
    """No iteration cap (the reference has none, hyper_ball.rs:710-712): a path of
    70,000 nodes runs until the reference's stopping rule, past 65,534 iterations, with
    the per-iteration records drained from the device ring along the way.  After t
    iterations node u's registers are the sketch of its exact ball [u - t, u + t]:
    the last iteration's output buffer holds that for t = iterations, the reference's
    counters buffer for the iteration count rounded down to even."""
    from sparkx_amd import _lib
    n, p = 70_000, 4
    off, nb = _csr(n, [(i, i + 1) for i in range(n - 1)] + [(i + 1, i) for i in range(n - 1)])
    g = hb.Graph(off, nb, "path70k")
    say("long path", n, p)
    with hb.HyperBall(g, p) as h:
        it = h.stats.iterations
        say("long path iterations", it)
        assert it > 65_534
        last = h.registers(_lib.HB_REGS_LATEST)
        refbuf = h.registers(_lib.HB_REGS_REFERENCE)
        assert np.array_equal(last, _ball_sketches(oracle, n, p, it))
        assert np.array_equal(refbuf, last if it % 2 == 0 else _ball_sketches(oracle, n, p, it - 1))
        assert np.array_equal(h.estimates(), oracle.estimates(p, refbuf.ravel()))


def test_empty_graph(hb):
    g = hb.Graph(np.zeros(1, np.uint64), np.zeros(0, np.uint64))
    say("empty graph")
    with hb.HyperBall(g, 6) as h:
        assert h.get_iters() == 1
        assert h.registers().size == 0
        assert h.compute_closeness_centrality(True).size == 0


def test_truncation_and_reference_parity(hb, oracle):
    g = hb.Graph.erdos_renyi(5000, 12000, seed=3)
    p = 6
    with hb.HyperBall(g, p, _run=False) as h:
        for it in (1, 2, 3, 4, 5):
            say("truncate", it)
            ref = _oracle_run(oracle, g, p, max_iters=it)
            h.run(max_iters=it)
            _assert_same(h, ref, oracle, p, f"max_iters={it}")
            assert h.stats.converged == (ref.increases[-1] == 0)


def test_centralities_match_oracle(hb, oracle):
    g = hb.Graph.erdos_renyi(4000, 10000, seed=21)  # sparse: isolated and small components
    p = 8
    say("centralities")
    ref = _oracle_run(oracle, g, p)
    est = oracle.estimates(p, ref.regs_ref)
    with hb.HyperBall.new(g) as h:  # HyperBall8 default (GraphMemoryMap::hyperball)
        assert h.p == 8
        for kind, norm, fn in ((oracle.CLOSENESS, False, h.compute_closeness_centrality),
                               (oracle.HARMONIC, True, h.compute_harmonic_centrality),
                               (oracle.HARMONIC, False, h.compute_harmonic_centrality)):
            assert np.array_equal(fn(norm), oracle.centrality(kind, norm, est, ref.dist, ref.inv))
        assert np.array_equal(h.compute_lins_centrality(),
                              oracle.centrality(oracle.LIN, False, est, ref.dist, ref.inv))
        assert h.get_or_compute_lins_centrality() is h.get_lins_centrality()
        with pytest.raises(LookupError):
            h.get_closeness_centrality(True)
        with pytest.raises(AssertionError):
            h.get_geodesic_distance(g.size())
        assert h.get_geodesic_distance(5) == ref.dist[5]

    # get_or_compute_closeness tests the harmonic slot (hyper_ball.rs:332), reproduced
    with hb.HyperBall(g, p) as h:
        c1 = h.get_or_compute_closeness_centrality(True)     # harmonic slot empty: computes
        c2 = h.get_or_compute_closeness_centrality(True)     # still empty: computes again
        assert c1 is not c2 and np.array_equal(c1, c2)
        h.compute_harmonic_centrality(False)
        with pytest.raises(LookupError):
            h.get_or_compute_closeness_centrality(False)     # harmonic set, closeness not
        h.compute_harmonic_centrality(True)
        assert h.get_or_compute_closeness_centrality(True) is h.get_closeness_centrality(True)


def test_neighbourhood_function_exact(hb, oracle):
    g = hb.Graph.erdos_renyi(20000, 50000, seed=8)
    p = 7
    say("nf")
    trace = {}
    _oracle_run(oracle, g, p, on_iter=lambda t, r, d, i: trace.__setitem__(t, r))
    with hb.HyperBall(g, p, nf=True) as h:
        nf = h.neighbourhood_function()
        assert nf.size == len(trace)
        for t in range(1, nf.size + 1):
            est = oracle.estimates(p, trace[t]).astype(np.float64)
            exact = sum(int(x * 2**40) for x in est)  # every f32 estimate >= 1 is exact in 2^-40
            assert nf[t - 1] == float(exact) / 2**40, t


def test_abi_error_paths(hb):
    from sparkx_amd import _lib
    L = _lib.load()
    ctx = C.c_void_p()
    assert L.hb_create(C.byref(ctx), 0) == 0
    try:
        assert L.hb_run(ctx, 6, 0, 0, None) == _lib.HB_ESTATE
        assert L.hb_read_registers(ctx, 0, (C.c_uint8 * 4)()) == _lib.HB_ESTATE
        off = np.array([0, 1, 2], np.uint64)
        bad = np.array([1, 2], np.uint64)  # 2 >= n
        assert L.hb_stage_csr(ctx, off.ctypes.data, 8, 2, bad.ctypes.data, 8, 2) == _lib.HB_EINVAL
        assert b"not < n" in L.hb_last_error()
        dec = np.array([0, 2, 1], np.uint64)
        assert L.hb_stage_csr(ctx, dec.ctypes.data, 8, 2, bad.ctypes.data, 8, 1) == _lib.HB_EINVAL
        good = np.array([1, 0], np.uint32)
        assert L.hb_stage_csr(ctx, off.ctypes.data, 8, 2, good.ctypes.data, 4, 2) == 0
        assert L.hb_run(ctx, 3, 0, 0, None) == _lib.HB_EINVAL
        assert L.hb_run(ctx, 6, -1, 0, None) == _lib.HB_EINVAL
        assert L.hb_run(ctx, 6, 0, _lib.HB_FLAG_KTIMING | _lib.HB_FLAG_GRAPH, None) == _lib.HB_EINVAL
        assert L.hb_run(ctx, 6, 0, 0, None) == 0
        assert L.hb_centrality(ctx, 7, 0, (C.c_double * 2)()) == _lib.HB_EINVAL
        assert L.hb_read_registers(ctx, 5, (C.c_uint8 * 128)()) == _lib.HB_EINVAL
    finally:
        L.hb_destroy(ctx)


def test_cpp_mirror_matches_python_mirror(hb, oracle, tmp_path):
    """The C++ HyperBallInner mirror (include/sparkx_hb.hpp) gives the oracle's results."""
    import subprocess
    from tests.test_cpu_host import _build_cpp_mirror
    exe = _build_cpp_mirror(tmp_path)
    say("cpp mirror")
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "sparkx_amd"))
    res = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.split("\n")
    off = np.array([0, 1, 3, 5, 7, 8, 10, 12, 14], np.uint64)
    nb = np.array([1, 0, 2, 1, 3, 2, 4, 3, 6, 7, 5, 7, 5, 6], np.uint32)
    ref = oracle.hyperball(off, nb, 6)
    est = oracle.estimates(6, ref.regs_ref)
    c = oracle.centrality(oracle.CLOSENESS, True, est, ref.dist, ref.inv)
    lin = oracle.centrality(oracle.LIN, False, est, ref.dist, ref.inv)
    assert lines[0] == f"iters {ref.iterations}"
    for u in range(8):
        _, uu, d, cc, ll = lines[1 + u].split()
        assert (float(d), float(cc), float(ll)) == (ref.dist[u], c[u], lin[u])


@pytest.fixture
def batch_trace(monkeypatch):
    """Progress lines from the batched engine's host loop on stderr (SPARKX_HB_TRACE)."""
    monkeypatch.setenv("SPARKX_HB_TRACE", "1")


def _compare_batch_graph(bt, i, h, ctx):
    from sparkx_amd import _lib
    it, conv, inc = bt.graph_stats(i)
    assert it == h.stats.iterations, ctx
    assert list(inc) == list(h.stats.increases[:it]), ctx
    for which in (_lib.HB_REGS_REFERENCE, _lib.HB_REGS_LATEST):
        assert np.array_equal(bt.registers(i, which), h.registers(which)), ctx
    d, v = bt.accumulators(i)
    hd, hv = h._accumulators()
    assert np.array_equal(d, hd) and np.array_equal(v, hv), ctx
    assert np.array_equal(bt.estimates(i), h.estimates()), ctx


def _hub_graph(hb, n, hub_deg, extra, seed):
    """Symmetric CSR: node 0 joined to nodes 1..hub_deg, plus `extra` random edges
    among the other nodes (duplicates and loops dropped): a hub whose list is cut
    into chunks."""
    rng = np.random.default_rng(seed)
    src = [np.zeros(hub_deg, np.int64), rng.integers(1, n, extra)]
    dst = [np.arange(1, hub_deg + 1), rng.integers(1, n, extra)]
    a, b = np.concatenate(src), np.concatenate(dst)
    keep = a != b
    a, b = a[keep], b[keep]
    pairs = np.unique(np.stack([np.concatenate([a, b]), np.concatenate([b, a])], 1), axis=0)
    off = np.zeros(n + 1, np.uint32)
    np.add.at(off, pairs[:, 0] + 1, 1)
    return hb.Graph(np.cumsum(off).astype(np.uint32), pairs[:, 1].astype(np.uint32), f"hub{n}_{hub_deg}")


def test_batch_small_graphs_against_oracle_and_schedules(hb, oracle, batch_trace):
    from sparkx_amd.batch import HyperBallBatch
    items = []
    for i, p in enumerate((4, 5, 6, 7, 8, 6)):
        g = hb.Graph.erdos_renyi(700 + 100 * i, 2000 + 500 * i, seed=50 + i) if i % 2 else hb.Graph.rmat(10, 8, seed=60 + i)
        items.append((g, p))
    # hubs above the wavefront bound (2048 arcs at p >= 6, 16384 at p <= 5): chunked lists
    items.append((_hub_graph(hb, 9000, 6000, 20000, 71), 6))
    items.append((_hub_graph(hb, 9000, 5000, 30000, 72), 8))
    items.append((_hub_graph(hb, 20000, 17000, 10000, 73), 4))
    refs = [_oracle_run(oracle, g, p) for g, p in items]
    # the second run replays the iterations the first captured (kept per set of running
    # precisions until the schedule changes)
    variants = [dict(graph_capture=True), dict(graph_capture=True), dict(graph_capture=False), dict(dense=True),
                dict(schedule=(2, 16, 512)), dict(schedule=(16, 256, 65536))]
    from sparkx_amd import _lib
    with HyperBallBatch(items) as bt:
        for kw in variants:
            say("batch variant", kw)
            sched = kw.pop("schedule", None)
            if sched:
                bt.set_schedule(*sched)
            bt.run(**kw)
            for i, ((g, p), ref) in enumerate(zip(items, refs)):
                it, conv, inc = bt.graph_stats(i)
                assert it == ref.iterations and list(inc) == list(ref.increases), (kw, i)
                assert np.array_equal(bt.registers(i, _lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref), (kw, i)
                assert np.array_equal(bt.registers(i, _lib.HB_REGS_LATEST).ravel(), ref.regs_last), (kw, i)
                d, v = bt.accumulators(i)
                assert np.array_equal(d, ref.dist) and np.array_equal(v, ref.inv), (kw, i)
        knobs, ms = bt.autotune()
        say("batch autotune", knobs, ms)
        bt.run()
        for i, ((g, p), ref) in enumerate(zip(items, refs)):
            assert np.array_equal(bt.registers(i, _lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref), i


def test_batch_truncation_and_errors(hb, oracle, batch_trace):
    from sparkx_amd import HyperBallError, _lib
    from sparkx_amd.batch import HyperBallBatch
    g = hb.Graph.erdos_renyi(2000, 6000, seed=77)
    ref = _oracle_run(oracle, g, 6, max_iters=3)
    with HyperBallBatch([(g, 6), (g, 7)], max_iters=3) as bt:
        bt.run()
        it, conv, inc = bt.graph_stats(0)
        assert it == 3 and list(inc) == list(ref.increases)
        assert np.array_equal(bt.registers(0, _lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref)
    with pytest.raises(HyperBallError) as ei:
        HyperBallBatch([(g, 9)])
    assert ei.value.code == _lib.HB_ERANGE


def test_batch_64_graphs_match_single_graph_runs(hb, batch_trace):
    """Batched configuration: 64 seeded graphs (32 ER + 32 R-MAT, n in [2^12, 2^18],
    mixed p = 4..8) in one work-queue launch per iteration, hipGraph-replayed; every
    graph bit-identical to its own single-graph run."""
    from sparkx_amd.batch import HyperBallBatch, batched_workload
    items = batched_workload(64, seed=1)
    say("batch64", sum(g.size() for g, _ in items), sum(g.width() for g, _ in items))
    with HyperBallBatch(items) as bt:
        st = bt.run(graph_capture=True)
        assert st.tiles > 0 and st.iterations_max > 0
        for i, (g, p) in enumerate(items):
            with hb.HyperBall(g, p) as h:
                _compare_batch_graph(bt, i, h, f"graph {i} p={p}")


def test_autotune_keeps_results(hb, oracle):
    g = hb.Graph.rmat(13, 16, seed=12)
    p = 7
    ref = _oracle_run(oracle, g, p)
    with hb.HyperBall(g, p, _run=False) as h:
        knobs, ms = h.autotune(trial_iters=2)
        say("autotune", knobs, ms)
        assert ms > 0
        h.run()
        _assert_same(h, ref, oracle, p, f"autotuned {knobs}")


# ---------------------------------------------------------------------------
# learned surrogate: BF16 MFMA kernels against the CPU oracle
# ---------------------------------------------------------------------------
def _mlp_case(seed=3, n=1000):
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.25, M.N_PARAMS).astype(np.float32)
    x = np.column_stack([rng.integers(0, 50, n), rng.normal(1, 1, n), rng.normal(2, 1, (n, 4))]).astype(np.float32)
    t = rng.normal(1.0, 0.5, n).astype(np.float32)
    norm = np.concatenate([x.mean(0), x.std(0)]).astype(np.float32)
    return params, norm, x, t


def test_mlp_gradient_and_forward_match_oracle(hb):
    from oracle import mlp_oracle as M
    from sparkx_amd.surrogate import Surrogate
    params, norm, x, t = _mlp_case()
    s = Surrogate()
    s.set_params(params, norm)
    say("mlp gradient")
    g, loss = s.gradient(x, t)
    ref_loss, ref_g = M.loss_and_grad(params.astype(np.float64), x, t, norm[:6], norm[6:])
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    # f32 MFMA accumulation vs f64: tolerance relative to the gradient's scale
    scale = np.abs(ref_g).max()
    assert np.max(np.abs(g - ref_g)) <= 2e-4 * scale, np.max(np.abs(g - ref_g)) / scale
    y = s.predict(x)
    ref_y, _ = M.forward(params.astype(np.float64), x, norm[:6], norm[6:])
    err = np.abs(y - ref_y)
    scale = max(1.0, np.abs(ref_y).max())
    # f32 accumulation can move a hidden pre-activation across a bf16 rounding
    # boundary of the oracle's f64 value; that sample's output then moves by about
    # one bf16 step of one activation times its weight.  All others agree to f32.
    assert np.mean(err <= 1e-5 * scale) >= 0.98, np.sort(err)[-20:]
    assert err.max() <= 2e-3 * scale, err.max()
    s.close()


def test_mlp_adam_step_matches_oracle(hb):
    from oracle import mlp_oracle as M
    from sparkx_amd.surrogate import Surrogate
    params, norm, x, t = _mlp_case(seed=4, n=512)
    s = Surrogate()
    s.set_params(params, norm)
    g, _ = s.gradient(x, t)
    # one epoch with one minibatch = one Adam step; training recomputes the normalisation
    s.train(x, t, epochs=1, batch=x.shape[0], lr=1e-3, seed=1)
    p1, n1 = s.params()
    mean = x.astype(np.float64).mean(0)
    assert np.allclose(n1[:6], mean, rtol=1e-5, atol=1e-6)
    s.set_params(params, n1)
    g1, _ = s.gradient(x, t)
    ref, _, _ = M.adam(params.astype(np.float64), g1.astype(np.float64), np.zeros(M.N_PARAMS),
                       np.zeros(M.N_PARAMS), 1, lr=1e-3)
    assert np.max(np.abs(p1 - ref)) <= 1e-6 + 1e-5 * np.abs(ref).max()
    s.close()


def test_snapshots_match_oracle_estimates(hb, oracle):
    import ctypes as C
    from sparkx_amd import _lib
    g = hb.Graph.rmat(12, 8, seed=31)
    trace = {}
    _oracle_run(oracle, g, 4, on_iter=lambda t, r, d, i: trace.__setitem__(t, r))
    with hb.HyperBall(g, 4, _run=False) as h:
        _lib.check(h._L.hb_set_snapshots(h._ctx, 4))
        h.run(max_iters=4)
        snaps = np.empty((4, g.size()), np.float32)
        _lib.check(h._L.hb_read_snapshots(h._ctx, _lib.ptr(snaps, C.c_float)))
    for t in range(1, 5):
        assert np.array_equal(snaps[t - 1], oracle.estimates(4, trace[t])), t


def test_surrogate_beats_truncated_hyperball(hb, oracle):
    """Trained on small R-MAT graphs, the MLP ranks a held-out graph's harmonic
    centrality better than the 4-iteration HyperBall accumulator it starts from."""
    from sparkx_amd.surrogate import Surrogate, labelled_graph
    xs, ys = [], []
    for seed in range(8):
        say("surrogate train graph", seed)
        g = hb.Graph.rmat(14, 16, seed=100 + seed)
        x, y, _, _, _ = labelled_graph(g)
        linked = g.degrees() > 0
        xs.append(x[linked])
        ys.append(y[linked])
    s = Surrogate(seed=7)
    losses = s.train(np.concatenate(xs), np.concatenate(ys), epochs=20, batch=4096, lr=1e-3)
    say("surrogate losses", losses)
    assert losses[-1] < losses[0]
    g = hb.Graph.rmat(15, 16, seed=999)
    x, y, h4, _, truth = labelled_graph(g)
    with hb.HyperBall(g, 4, _run=False) as h:
        pred, feat, ms = s.rank(h, want_features=True)
    assert ms > 0
    linked = g.degrees() > 0
    assert np.allclose(feat, x, rtol=1e-5, atol=1e-5)
    assert np.array_equal(pred[linked], s.predict(feat[linked]))
    assert np.all(pred[~linked] == np.float32(-3e38))
    rho_mlp = oracle.spearman(pred[linked].astype(np.float64), truth[linked])
    rho_base = oracle.spearman(h4[linked], truth[linked])
    say("surrogate rho", rho_mlp, "truncated", rho_base)
    assert rho_mlp > rho_base
    s.close()


def test_results_written_under_reference_cache_names(hb, oracle, tmp_path):
    from sparkx_amd import cache
    g = hb.Graph.erdos_renyi(3000, 9000, seed=5)
    cache.save_graph_cache(g, str(tmp_path), "g1")
    g2 = cache.load_cached_graph(str(tmp_path), "g1")  # staged from the u64 memory maps
    ref = _oracle_run(oracle, g, 8)
    with hb.HyperBall(g2, 8) as h:
        c = h.compute_closeness_centrality(True)
        hh = h.compute_harmonic_centrality(False)
        paths = cache.write_results(h, str(tmp_path), "g1")
    assert len(paths) == 5
    regs = np.fromfile(cache.cache_file(str(tmp_path), "registers", "g1", 8), np.uint8)
    assert np.array_equal(regs, ref.regs_ref)
    assert np.array_equal(cache.read_result(str(tmp_path), "distances", "g1"), ref.dist)
    assert np.array_equal(cache.read_result(str(tmp_path), "inv_distances", "g1"), ref.inv)
    assert np.array_equal(cache.read_result(str(tmp_path), "closeness", "g1", True), c)
    assert np.array_equal(cache.read_result(str(tmp_path), "harmonic", "g1", False), hh)


def test_restore_errors_and_repeats(hb, oracle):
    """hb_restore rejects a precision outside 4..16, missing buffers and a context with
    no graph, leaving the context usable; restoring twice on one context gives the
    saved run both times."""
    import ctypes as C
    from sparkx_amd import _lib
    L = _lib.load()
    g = hb.Graph.rmat(11, 8, seed=17)
    p = 6
    ref = _oracle_run(oracle, g, p)
    regs = ref.regs_ref.copy()
    dist, inv = ref.dist.copy(), ref.inv.copy()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), 0))
    try:
        assert L.hb_restore(ctx, p, _lib.ptr(regs, C.c_uint8), _lib.ptr(dist, C.c_double),
                            _lib.ptr(inv, C.c_double)) == _lib.HB_ESTATE  # nothing staged
    finally:
        L.hb_destroy(ctx)
    with hb.HyperBall(g, p, _run=False) as h:
        assert L.hb_restore(h._ctx, 3, _lib.ptr(regs, C.c_uint8), _lib.ptr(dist, C.c_double),
                            _lib.ptr(inv, C.c_double)) == _lib.HB_EINVAL
        assert L.hb_restore(h._ctx, 17, _lib.ptr(regs, C.c_uint8), _lib.ptr(dist, C.c_double),
                            _lib.ptr(inv, C.c_double)) == _lib.HB_EINVAL
        assert L.hb_restore(h._ctx, p, None, _lib.ptr(dist, C.c_double), _lib.ptr(inv, C.c_double)) == _lib.HB_EINVAL
        assert L.hb_restore(h._ctx, p, _lib.ptr(regs, C.c_uint8), None, _lib.ptr(inv, C.c_double)) == _lib.HB_EINVAL
        for _ in range(2):
            _lib.check(L.hb_restore(h._ctx, p, _lib.ptr(regs, C.c_uint8), _lib.ptr(dist, C.c_double),
                                    _lib.ptr(inv, C.c_double)))
            h._ran(None)
            assert np.array_equal(h.registers(_lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref)
            d, i = h._accumulators()
            assert np.array_equal(d, ref.dist) and np.array_equal(i, ref.inv)
            assert np.array_equal(h.estimates(), oracle.estimates(p, ref.regs_ref))


@pytest.mark.parametrize("layout", ["auto", "input"])
def test_get_or_compute_reopens_cached_results(hb, oracle, tmp_path, layout):
    """get_or_compute (hyper_ball.rs:165-201): the first call runs and saves registers
    and accumulators; computed centralities go to their files; the second call
    rebuilds the run from the files (0 iterations, as the reference's restored struct)
    and finds the cached centralities; every reader and epilogue of the reopened run
    gives the oracle's values."""
    g = hb.Graph.erdos_renyi(3000, 9000, seed=15)
    p = 8
    ref = _oracle_run(oracle, g, p)
    est = oracle.estimates(p, ref.regs_ref)
    d = str(tmp_path)
    with hb.HyperBall.get_or_compute(g, d, "g3", p=p, layout=layout) as h:
        assert h.get_iters() == ref.iterations
        c1 = h.compute_closeness_centrality(True)
        l1 = h.compute_lins_centrality()
    with hb.HyperBall.get_or_compute(g, d, "g3", p=p, layout=layout) as h:
        assert h.get_iters() == 0
        assert np.array_equal(h.get_closeness_centrality(True), c1)
        assert np.array_equal(h.get_lins_centrality(), l1)
        with pytest.raises(LookupError):
            h.get_hamonic_centrality(False)
        assert np.array_equal(h.registers().ravel(), ref.regs_ref)
        assert np.array_equal(h.estimates(), est)
        dist, inv = h._accumulators()
        assert np.array_equal(dist, ref.dist) and np.array_equal(inv, ref.inv)
        for kind, norm, fn in ((oracle.HARMONIC, True, h.compute_harmonic_centrality),
                               (oracle.CLOSENESS, False, h.compute_closeness_centrality)):
            assert np.array_equal(fn(norm), oracle.centrality(kind, norm, est, ref.dist, ref.inv))
    with hb.HyperBall.get_or_compute(g, d, "g3", p=p, layout=layout) as h:
        assert np.array_equal(h.get_hamonic_centrality(True),  # written by the second object
                              oracle.centrality(oracle.HARMONIC, True, est, ref.dist, ref.inv))
    # the same files for another precision are not there: that one runs
    with hb.HyperBall.get_or_compute(g, d, "g3", p=6, layout=layout) as h:
        assert h.get_iters() == _oracle_run(oracle, g, 6).iterations


def test_cpp_mirror_get_or_compute(hb, oracle, tmp_path):
    """The C++ mirror's get_or_compute against a cache directory: run and save, then
    reopen; the reopened run's accumulators, cached closeness and a new Lin epilogue
    are the oracle's."""
    import subprocess
    from tests.test_cpu_host import _build_cpp_mirror
    exe = _build_cpp_mirror(tmp_path, "hyperball_cpp_cache")
    cache_dir = tmp_path / "cache"
    cache_dir.mkdir()
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "sparkx_amd"))
    res = subprocess.run([exe, str(cache_dir)], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.split("\n")
    off = np.array([0, 1, 3, 5, 7, 8, 10, 12, 14], np.uint64)
    nb = np.array([1, 0, 2, 1, 3, 2, 4, 3, 6, 7, 5, 7, 5, 6], np.uint32)
    ref = oracle.hyperball(off, nb, 6)
    est = oracle.estimates(6, ref.regs_ref)
    c = oracle.centrality(oracle.CLOSENESS, True, est, ref.dist, ref.inv)
    lin = oracle.centrality(oracle.LIN, False, est, ref.dist, ref.inv)
    assert lines[0] == f"first iters {ref.iterations}" and lines[1] == "second iters 0"
    for u in range(8):
        _, uu, d, c1, c2, ll = lines[2 + u].split()
        assert (float(d), float(c1), float(c2), float(ll)) == (ref.dist[u], c[u], c[u], lin[u])


def _layout_graph(hb, name):
    rng = np.random.default_rng(21)
    if name == "rmat13":
        return hb.Graph.rmat(13, 16, seed=5)
    if name == "kmer":
        return hb.Graph.kmer_dbg(20000, 4, 0.01, 11, seed=3)
    n = 4000 if name == "directed" else 1500
    if name == "directed":  # no reverse arcs: every run takes the scanning path
        src = rng.integers(0, n, 12000)
        dst = rng.integers(0, n, 12000)
    else:  # "multi": symmetric with duplicates, self-loops, isolated nodes and a hub
        a = rng.integers(0, n // 2, 5000)
        b = rng.integers(0, n // 2, 5000)
        hub = np.full(900, 7)
        leaves = rng.integers(0, n // 2, 900)
        loops = np.arange(0, n // 2, 5)
        src = np.concatenate([a, b, hub, leaves, loops, a[:300]])
        dst = np.concatenate([b, a, leaves, hub, loops, b[:300]])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    return hb.Graph(np.cumsum(off).astype(np.uint64), dst.astype(np.uint32), name)


@pytest.mark.parametrize("name,p", [("rmat13", 4), ("rmat13", 7), ("rmat13", 10), ("rmat13", 12),
                                    ("directed", 7), ("multi", 6), ("multi", 11), ("kmer", 6)])
def test_device_layouts_match_input_layout_and_oracle(hb, oracle, name, p):
    """The device numberings (degree order, chain order) change nothing a caller can
    read: every layout matches the oracle bit for bit, with the same neighbourhood
    function and centralities, and the layout is built once per staged graph; "auto"
    takes the chain order for low-degree sparse graphs (the k-mer and the random
    directed graph) and the degree order for the others."""
    g = _layout_graph(hb, name)
    say("layout", name, p, g.size(), g.width())
    ref = _oracle_run(oracle, g, p)
    out = {}
    # the "multi" hub (degree 918) goes through the chunk / merge kernels
    sched = (4, 256, 128) if name == "multi" else None
    chain_auto = int(g.degrees().max()) <= 64 and g.width() <= 4 * g.size()  # the auto rule
    codes = {"input": 0, "degree": 1, "chain": 2, "auto": 2 if chain_auto else 1}
    for layout in ("input", "degree", "chain", "auto"):
        with hb.HyperBall(g, p, layout=layout, nf=True, schedule=sched) as h:
            st = h.stats
            assert st.layout == codes[layout], layout
            assert (st.ms_layout > 0) == (layout != "input")
            _assert_same(h, ref, oracle, p, f"{name} p={p} {layout}")
            nf = list(st.nf[:st.iterations])
            cent = [h.compute_closeness_centrality(True), h.compute_harmonic_centrality(False),
                    h.compute_lins_centrality()]
            st2 = h.run(nf=True)  # same staged graph: the layout is reused, not rebuilt
            assert st2.ms_layout == st.ms_layout and st2.iterations == ref.iterations
            out[layout] = (nf, cent)
    for layout in ("degree", "chain", "auto"):
        assert out["input"][0] == out[layout][0]
        for a, b in zip(out["input"][1], out[layout][1]):
            assert np.array_equal(a, b)


def _chain_graph(hb, kind):
    """Graphs whose chain numbering keeps every linked node in one id-range bin, so the
    dense iterations read in-window neighbours from LDS: a de Bruijn graph with
    duplicated arcs, self-loops and isolated nodes appended, and bare directed cycles
    whose lengths do not divide a wavefront's teams."""
    if kind == "kmer":
        g = hb.Graph.kmer_dbg(12000, 4, 0.02, 9, seed=8)
        off, nb = g.offsets.astype(np.int64), g.neighbours.astype(np.int64)
        n = g.size()
        src = np.repeat(np.arange(n), np.diff(off))
        extra = np.arange(0, n, 7)  # self-loops and a repeated arc per 7th node
        src = np.concatenate([src, extra, src[: n // 3]])
        dst = np.concatenate([nb, extra, nb[: n // 3]])
        n_all = n + 37  # isolated nodes after the linked ones
    else:
        lens = [5, 63, 1000, 17, 2]
        src, dst, at = [], [], 0
        for m in lens:
            ids = np.arange(at, at + m)
            src.append(ids)
            dst.append(np.roll(ids, -1))
            at += m
        src, dst = np.concatenate(src), np.concatenate(dst)
        n_all = at + 3
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n_all + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    return hb.Graph(np.cumsum(off).astype(np.uint64), dst.astype(np.uint32), kind)


@pytest.mark.parametrize("kind", ["kmer", "cycles"])
@pytest.mark.parametrize("p", [4, 5, 6, 7, 8, 10])
def test_chain_layout_window_matches_oracle(hb, oracle, kind, p):
    """Chain numbering, dense iterations: the union reads neighbours inside each
    wavefront's window of consecutive nodes from LDS (team widths 1..64; at p = 5 and 6
    one node per lane, k_union_rows, with a 64-node window).  Default and dense
    schedules, hipGraph replay included, equal the oracle bit for bit."""
    g = _chain_graph(hb, kind)
    say("chain window", kind, p, g.size(), g.width())
    ref = _oracle_run(oracle, g, p, max_iters=300)
    linked = int((g.degrees() > 0).sum())
    for kw in (dict(), dict(dense=True), dict(dense=True, graph_capture=True)):
        with hb.HyperBall(g, p, layout="chain", max_iters=300, **kw) as h:
            assert h.stats.layout == 2
            _assert_same(h, ref, oracle, p, f"{kind} p={p} {kw}")
            if kw.get("dense"):
                # every iteration recomputes every linked node over every arc, and the
                # kernels' node and arc counts say so
                it = min(h.stats.iterations, len(h.stats.iter_nodes))
                assert list(h.stats.iter_nodes[:it]) == [linked] * it, kw
                assert list(h.stats.iter_arcs[:it]) == [g.width()] * it, kw


def test_layout_switch_and_snapshots(hb, oracle):
    """Switching layouts on one context discards the run; snapshots come back in the
    caller's order under both."""
    from sparkx_amd import _lib
    g = _layout_graph(hb, "rmat13")

    def snapshots(h):
        out = np.empty((3, g.size()), np.float32)
        _lib.check(h._L.hb_read_snapshots(h._ctx, _lib.ptr(out, C.c_float)))
        return out

    with hb.HyperBall(g, 6, _run=False) as h:
        _lib.check(h._L.hb_set_snapshots(h._ctx, 3))
        h.set_layout("input")
        h.run()
        snaps_in = snapshots(h)
        h.set_layout("degree")
        with pytest.raises(hb.HyperBallError):
            h.registers(0)
        h.run()
        assert np.array_equal(snapshots(h), snaps_in)
        h.set_layout("chain")
        h.run()
        assert np.array_equal(snapshots(h), snaps_in)
        with pytest.raises(ValueError):
            h.set_layout("random")


def test_snapshot_rows_follow_the_completed_run(hb, oracle):
    """hb_set_snapshots after a run changes the next run only: reading returns the rows
    the completed run recorded, and hb_result_info reports their count."""
    from sparkx_amd import _lib
    g = hb.Graph.erdos_renyi(3000, 9000, seed=21)
    with hb.HyperBall(g, 6, _run=False) as h:
        _lib.check(h._L.hb_set_snapshots(h._ctx, 2))
        h.run()
        _lib.check(h._L.hb_set_snapshots(h._ctx, 16))
        p, it, rows = h.result_info()
        assert (p, rows) == (6, 2) and it == h.stats.iterations
        out = np.full((16, g.size()), -1.0, np.float32)
        _lib.check(h._L.hb_read_snapshots(h._ctx, _lib.ptr(out, C.c_float)))
        assert (out[2:] == -1.0).all()
        ref = oracle.hyperball(g.offsets.astype(np.uint64), g.neighbours, 6, max_iters=1)
        assert np.array_equal(out[0], oracle.estimates(6, ref.regs_last))


def test_surrogate_replaces_the_mirrors_run(hb):
    """After an anytime ranking the Python mirror describes the surrogate's run: its
    precision sizes the register reads."""
    from sparkx_amd.surrogate import Surrogate
    g = hb.Graph.rmat(12, 8, seed=5)
    with hb.HyperBall(g, 7) as h:
        s = Surrogate()
        try:
            s.rank(h, p_feat=5)
        finally:
            s.close()
        assert h.p == 5 and h.stats is None
        p, it, rows = h.result_info()
        assert (p, rows) == (5, 4) and it <= 4
        assert h.registers().shape == (g.size(), 32)


def test_run_many_matches_oracle(hb, oracle):
    """hb_run_many: graphs of different sizes, densities and p (4..12) on their own
    streams from 4 host threads; each one bit-identical to the oracle."""
    from sparkx_amd.batch import HyperBallMany
    items = []
    for i, p in enumerate((4, 5, 6, 7, 8, 6, 7, 12, 9, 5)):
        if i % 3 == 0:
            g = hb.Graph.erdos_renyi(600 + 150 * i, 2500 + 700 * i, seed=70 + i)
        elif i % 3 == 1:
            g = hb.Graph.rmat(9 + i % 4, 8, seed=80 + i)
        else:
            g = _layout_graph(hb, "directed")
        items.append((g, p))
    refs = [_oracle_run(oracle, g, p) for g, p in items]
    with HyperBallMany(items, threads=4) as many:
        for rep in range(2):
            say("run_many", rep)
            stats = many.run()
            for i, (h, ref, (g, p)) in enumerate(zip(many.hbs, refs, items)):
                assert stats[i].iterations == ref.iterations
                _assert_same(h, ref, oracle, p, f"graph {i} p={p} rep {rep}")


def test_run_many_reports_the_failing_context(hb):
    from sparkx_amd import _lib
    L = _lib.load()
    g = hb.Graph.erdos_renyi(500, 2000, seed=3)
    ok = hb.HyperBall(g, 6, _run=False)
    empty = C.c_void_p()
    _lib.check(L.hb_create(C.byref(empty), 0))  # nothing staged
    try:
        ctxs = (C.c_void_p * 2)(ok._ctx.value, empty.value)
        ps = (C.c_int * 2)(6, 6)
        rc = L.hb_run_many(ctxs, ps, 2, 0, 0, None, 2)
        assert rc == _lib.HB_ESTATE
        assert b"context 1" in L.hb_last_error()
        assert ok.registers(0).shape == (500, 64)  # the other context completed
    finally:
        L.hb_destroy(empty)
        ok.close()
