"""CPU tests: pin the oracle (what the HIP path is checked against) against the
reference's fixtures, published numbers and the mathematics, and check the
host-side logic.  No GPU calls."""
import functools
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


# ---------------------------------------------------------------------------
# hashing: SipHash-1-3 known answers from OpenSSL; SipHash-2-4 paper vector
# ---------------------------------------------------------------------------
def test_siphash13_matches_openssl(oracle):
    with open(os.path.join(GOLDEN, "siphash13_kat.txt")) as fh:
        rows = [line.split() for line in fh if line.strip()]
    assert len(rows) >= 10
    for v, mac in rows:
        expected = int.from_bytes(bytes.fromhex(mac), "little")
        assert oracle.siphash(int(v).to_bytes(8, "little")) == expected, v
        assert oracle.lib().hbo_hash_usize(oracle.pinned(), int(v)) == expected


def test_siphash24_paper_vector(oracle):
    # Aumasson & Bernstein, SipHash paper appendix A: key 00..0f, message 00..0e
    k0 = int.from_bytes(bytes(range(8)), "little")
    k1 = int.from_bytes(bytes(range(8, 16)), "little")
    assert oracle.siphash(bytes(range(15)), 2, 4, k0, k1) == 0xA129CA6149BE45E5


def test_index_rank_definition(oracle):
    # index = top p bits, rank = 1 + clz((h << p) | 1 << (p-1)), checked on the KAT hashes
    for x in (0, 1, 400, 12345678):
        h = oracle.siphash(x.to_bytes(8, "little"))
        for p in (4, 6, 7, 8, 13, 16):
            idx, rank = oracle.index_rank(p, x)
            assert idx == h >> (64 - p)
            w = ((h << p) & (2**64 - 1)) | (1 << (p - 1))
            assert rank == 64 - w.bit_length() + 1
            assert 1 <= rank <= 65 - p


# ---------------------------------------------------------------------------
# fixture graphs and exact BFS closeness (bit-exact against the reference's files)
# ---------------------------------------------------------------------------
def test_dbg_fixture_graph_sizes(oracle):
    # datasets_size.csv: synth_gen_dbg_protein20_k2_edges16k 400/15980,
    # synth_gen_dbg_dna_k7_edges130k 16384/131068
    off, nb = oracle.dbg_mtx_csr(20, 2)
    assert (off.size - 1, nb.size) == (400, 15980)
    off, nb = oracle.dbg_mtx_csr(4, 7)
    assert (off.size - 1, nb.size) == (16384, 131068)
    assert np.all(np.diff(nb[off[100]:off[101]].astype(np.int64)) >= 0)


def test_exact_closeness_reproduces_fixtures(oracle):
    off, nb = oracle.dbg_mtx_csr(20, 2)
    fx = np.fromfile(os.path.join(GOLDEN, "exact_closeness_dbg_sigma20_k2.f64"), dtype="<f8")
    assert np.array_equal(oracle.exact_closeness(off, nb), fx)
    off, nb = oracle.dbg_mtx_csr(4, 7)
    fx = np.fromfile(os.path.join(GOLDEN, "exact_closeness_dbg_sigma4_k7.f64"), dtype="<f8")
    for lo in (0, 5000, 16384 - 300):
        assert np.array_equal(oracle.exact_closeness(off, nb, lo, lo + 300), fx[lo:lo + 300])


def _bfs_brute(n, arcs):
    """Independent distances for the BFS sums: Floyd-Warshall over path lengths >= 1.
    D[s][v] = shortest walk of length >= 1 from s to v, so D[s][s] is the shortest
    cycle through s -- the distance at which the reference's BFS rediscovers its
    source (it is queued but not marked visited)."""
    INF = 1 << 40
    A = np.full((n, n), INF, np.int64)
    for a, b in arcs:
        A[a, b] = 1
    D = A.copy()
    for k in range(n):
        D = np.minimum(D, D[:, k:k + 1] + A[k:k + 1, :])          # walks ending with an arc from k
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    D[D >= INF] = 0
    return D


def _csr_of(n, arcs):
    adj = [[] for _ in range(n)]
    for a, b in arcs:
        adj[a].append(b)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in adj])
    return off, np.array([v for x in adj for v in x], np.uint32)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_exact_bfs_against_brute_force_shortest_walks(oracle, seed):
    """reachable, total and the node-order sum of 1/d (exact.rs:386-396) from an
    independent all-pairs computation, on random directed graphs with self-loops,
    duplicate arcs, sinks and unreachable parts."""
    rng = np.random.default_rng(seed)
    n = 40
    arcs = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(90, 2))]
    arcs += [(5, 5), (7, 8), (7, 8)]
    off, nb = _csr_of(n, arcs)
    D = _bfs_brute(n, arcs)
    reach, total, inv = oracle.exact_bfs(off, nb)
    for s in range(n):
        d = D[s]
        assert reach[s] == np.count_nonzero(d), s
        assert total[s] == float(d.sum()), s
        acc = 0.0
        for v in range(n):
            acc += 1.0 / float(d[v]) if d[v] else 0.0
        assert inv[s] == acc, s
    # the self-loop makes node 5 its own neighbour at distance 1
    assert D[5][5] == 1 and reach[5] >= 1


def test_exact_bfs_closed_forms(oracle):
    # undirected path 0-1-...-(n-1): from 0, d(v) = v and the source returns at 2
    n = 50
    arcs = [(i, i + 1) for i in range(n - 1)] + [(i + 1, i) for i in range(n - 1)]
    off, nb = _csr_of(n, arcs)
    reach, total, inv = oracle.exact_bfs(off, nb, 0, 1)
    assert reach[0] == n and total[0] == n * (n - 1) / 2 + 2
    assert abs(inv[0] - (0.5 + math.fsum(1.0 / v for v in range(1, n)))) < 1e-12
    # star with centre 0: centre sees leaves at 1 and itself at 2; a leaf sees the
    # centre at 1, the other leaves and itself at 2
    n = 30
    arcs = [(0, i) for i in range(1, n)] + [(i, 0) for i in range(1, n)]
    off, nb = _csr_of(n, arcs)
    reach, total, inv = oracle.exact_bfs(off, nb, 0, 2)
    assert list(reach) == [n, n] and list(total) == [n + 1, 2 * n - 1]
    # isolated node and a node with only a self-loop
    off, nb = _csr_of(3, [(1, 1)])
    reach, total, inv = oracle.exact_bfs(off, nb)
    assert list(reach) == [0, 1, 0] and list(total) == [0, 1, 0] and list(inv) == [0.0, 1.0, 0.0]


def test_exact_centrality_formulas(oracle):
    """exact.rs: closeness None/false/true, harmonic None/false/true, Lin; zero rules."""
    reach = np.array([0, 1, 3, 5], np.uint64)
    total = np.array([0.0, 1.0, 5.0, 9.0])
    inv = np.array([0.0, 1.0, 2.0, 2.75])
    n = 10
    c = [oracle.exact_centrality(0, nz, reach, total, inv, n) for nz in (None, False, True)]
    assert list(c[0]) == [0.0, 0.0, 5.0, 9.0]
    assert list(c[1]) == [0.0, 0.0, 9.0 / 5.0, 1.0]
    assert list(c[2]) == [0.0, 0.0, 3.0 / 5.0, 5.0 / 9.0]
    h = [oracle.exact_centrality(1, nz, reach, total, inv, n) for nz in (None, False, True)]
    assert list(h[0]) == [0.0, 0.0, 2.0, 2.75]
    assert list(h[1]) == [0.0, 0.0, 4.5, 9.0 / 2.75]
    assert list(h[2]) == [0.0, 0.0, 1.5, 5.0 / 2.75]
    lin = oracle.exact_centrality(2, None, reach, total, inv, n)
    assert list(lin) == [0.0, 1.0, 9.0 / 5.0, 25.0 / 9.0]


# ---------------------------------------------------------------------------
# the reference's published accuracy numbers (plot_hyperball_results.py:38-57)
# ---------------------------------------------------------------------------
def _profile(oracle, sigma, k, p, fixture):
    off, nb = oracle.dbg_mtx_csr(sigma, k)
    ex = np.fromfile(os.path.join(GOLDEN, fixture), dtype="<f8")
    r = oracle.hyperball(off, nb, p, threads=4)
    c = oracle.centrality(oracle.CLOSENESS, True, oracle.estimates(p, r.regs_ref), r.dist, r.inv)
    return oracle.mae(c, ex), oracle.mape(c, ex), oracle.spearman(c, ex)


@pytest.mark.parametrize("p", [10, 11, 12, 13])
def test_published_numbers_exact_in_linear_counting_regime(oracle, p):
    """On the 400-node graph every estimate stays in the linear-counting regime for
    p >= 10: the published MAE, MAPE and rho are reproduced to the last bit."""
    from tests.golden.pin_hll_variant import PUBLISHED
    mae, mape, rho = _profile(oracle, 20, 2, p, "exact_closeness_dbg_sigma20_k2.f64")
    assert mae == PUBLISHED[p]["mae"][0]
    assert mape == PUBLISHED[p]["mape"][0]
    assert rho == PUBLISHED[p]["rho"][0]


def _recorded_rows():
    with open(os.path.join(GOLDEN, "published_accuracy.json")) as fh:
        return json.load(fh)["rows"]


_ROWS = _recorded_rows()
_FIXTURE = {"dbg_sigma20_k2": (20, 2, "exact_closeness_dbg_sigma20_k2.f64"),
            "dbg_sigma4_k7": (4, 7, "exact_closeness_dbg_sigma4_k7.f64")}
_METRICS = ("mae", "mape", "rho")


@functools.lru_cache(maxsize=None)
def _profile_row(p, graph):
    from oracle import hb_oracle as O
    sigma, k, fixture = _FIXTURE[graph]
    return dict(zip(_METRICS, _profile(O, sigma, k, p, fixture)))


def _row_cases(published: bool):
    cases = []
    for row in _ROWS:
        for key in _METRICS:
            marks = ()
            if published and row[key]["rel_diff"] != 0.0:
                # the estimate passes through the HLL++ bias-corrected regime, whose
                # empirical table lives only in the absent crate (parity unpinned)
                marks = (pytest.mark.xfail(strict=True, reason=(
                    f"bias-regime miss: ours {row[key]['ours']!r} vs published {row[key]['published']!r} "
                    f"(rel {row[key]['rel_diff']:.2e})")),)
            cases.append(pytest.param(row["p"], row["graph"], key, id=f"p{row['p']}-{row['graph']}-{key}",
                                      marks=marks))
    return cases


@pytest.mark.parametrize("p,graph,key", _row_cases(published=False))
def test_published_rows_reproduce_recorded_values(oracle, p, graph, key):
    """Regression guard on the estimator: every one of the 36 numbers equals, to the
    last bit, the value recorded in published_accuracy.json, so any change of the
    bias-corrected branch shows up here."""
    row = next(r for r in _ROWS if r["p"] == p and r["graph"] == graph)
    assert _profile_row(p, graph)[key] == row[key]["ours"]


@pytest.mark.parametrize("p,graph,key", _row_cases(published=True))
def test_published_rows_against_the_reference(oracle, p, graph, key):
    """Each published number (plot_hyperball_results.py:38-57) against the oracle.
    Numbers whose estimates stay in the linear-counting regime match exactly; the
    known misses are listed one by one as strict xfails with their values, so a
    change that makes one match (or breaks a matching one) is reported."""
    from tests.golden.pin_hll_variant import PUBLISHED
    gi = 0 if graph == "dbg_sigma20_k2" else 1
    assert _profile_row(p, graph)[key] == PUBLISHED[p][key][gi]


def test_published_accuracy_table_is_complete():
    assert len(_ROWS) == 12 and {(r["p"], r["graph"]) for r in _ROWS} == {
        (p, gname) for p in range(8, 14) for gname in _FIXTURE}
    matched = sum(r[k]["rel_diff"] == 0.0 for r in _ROWS for k in _METRICS)
    assert matched == 14  # 12 linear-counting numbers of the 400-node graph + rho at p = 8, 9


# ---------------------------------------------------------------------------
# the iteration: registers after t iterations are the sketch of the exact ball
# ---------------------------------------------------------------------------
def _ball_sketch(oracle, off, nb, u, t, p):
    m = 1 << p
    seen = {u}
    frontier = [u]
    for _ in range(t):
        nxt = []
        for x in frontier:
            for y in nb[off[x]:off[x + 1]]:
                y = int(y)
                if y not in seen:
                    seen.add(y)
                    nxt.append(y)
        frontier = nxt
    regs = np.zeros(m, np.uint8)
    for x in seen:
        i, r = oracle.index_rank(p, x)
        regs[i] = max(regs[i], r)
    return regs


@pytest.mark.parametrize("p", [4, 6, 7])
def test_registers_equal_sketch_of_ball(oracle, p):
    rng = np.random.default_rng(p)
    n = 40
    edges = rng.integers(0, n, size=(60, 2))
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(int(b))
        adj[b].add(int(a))
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in adj])
    nb = np.array([v for s in adj for v in sorted(s)], np.uint32)
    hist = {}
    oracle.hyperball(off, nb, p, on_iter=lambda t, regs, d, i: hist.__setitem__(t, regs.reshape(n, -1)))
    for t, regs in hist.items():
        for u in range(n):
            assert np.array_equal(regs[u], _ball_sketch(oracle, off, nb, u, t, p)), (t, u)


def test_duplicates_and_self_loops_do_not_change_registers(oracle):
    off = np.array([0, 2, 4, 5], np.uint64)
    nb = np.array([1, 2, 0, 2, 0], np.uint32)
    off2 = np.array([0, 4, 7, 9], np.uint64)
    nb2 = np.array([0, 1, 1, 2, 0, 1, 2, 0, 0], np.uint32)
    a = oracle.hyperball(off, nb, 6)
    b = oracle.hyperball(off2, nb2, 6)
    assert np.array_equal(a.regs_last, b.regs_last)


def test_threads_and_packed_layout_give_identical_results(oracle):
    off, nb = oracle.dbg_mtx_csr(4, 6)
    ref = oracle.hyperball(off, nb, 7, threads=1)
    for threads, packed in ((3, False), (1, True), (5, True)):
        r = oracle.hyperball(off, nb, 7, threads=threads, packed=packed)
        assert r.iterations == ref.iterations
        assert np.array_equal(r.regs_ref, ref.regs_ref)
        assert np.array_equal(r.dist, ref.dist) and np.array_equal(r.inv, ref.inv)


def test_reference_buffer_parity(oracle):
    off, nb = oracle.dbg_mtx_csr(4, 5)
    for it in (1, 2, 3, 4):
        r = oracle.hyperball(off, nb, 6, max_iters=it)
        assert r.final_parity_latest == (it % 2 == 0)
        if it % 2 == 0:
            assert np.array_equal(r.regs_ref, r.regs_last)
        else:
            prev = oracle.hyperball(off, nb, 6, max_iters=it - 1) if it > 1 else None
            if prev is not None:
                assert np.array_equal(r.regs_ref, prev.regs_last)


# ---------------------------------------------------------------------------
# estimator: closed forms and the exact-sum property the GPU relies on
# ---------------------------------------------------------------------------
def test_estimator_linear_counting_closed_form(oracle):
    for p in (4, 6, 7, 8, 12):
        m = 1 << p
        regs = np.zeros(m, np.uint8)
        regs[3] = 5
        assert oracle.estimate(p, regs) == np.float32(m * math.log(m / (m - 1)))


def test_estimator_raw_closed_form_above_5m(oracle):
    # all registers = r: S = m 2^-r exactly, raw = alpha m 2^r (> 5m for r >= 3)
    for p, alpha in ((4, np.float32(0.673)), (6, np.float32(0.709))):
        m = 1 << p
        for r in (4, 9, 20):
            regs = np.full(m, r, np.uint8)
            am = np.float32(alpha * np.float32(m * m))
            assert oracle.estimate(p, regs) == np.float32(am / np.float32(m * 2.0 ** -r))


def test_estimator_order_invariant_when_sum_is_exact(oracle):
    """k = sum 2^(max-M) < 2^24 makes the sequential f32 sum exact and order-free."""
    rng = np.random.default_rng(7)
    p = 7
    for _ in range(200):
        regs = rng.integers(10, 25, size=1 << p).astype(np.uint8)
        perm = rng.permutation(regs)
        assert oracle.estimate(p, regs) == oracle.estimate(p, perm)


def test_estimator_order_matters_when_sum_is_inexact(oracle):
    """Spread beyond 24 bits: the sequential order changes the f32 sum; the
    estimator must follow register order 0..m-1."""
    p = 6
    regs = np.full(64, 24, np.uint8)  # S ~ 8: raw ~ 363 > 5m, LC = 133 > 40
    regs[:8] = 0
    diffs = 0
    for seed in range(50):
        perm = np.random.default_rng(seed).permutation(regs)
        seq = np.float32(0)
        for r in perm:
            seq = np.float32(seq + np.float32(2.0 ** -int(r)))
        est = oracle.estimate(p, perm)
        alpha_mm = np.float32(np.float32(0.709) * np.float32(64 * 64))
        assert est == np.float32(alpha_mm / seq)
        diffs += est != oracle.estimate(p, regs)
    assert diffs > 0


def test_packed_and_u8_estimates_agree(oracle):
    rng = np.random.default_rng(3)
    for p in (4, 5, 6, 7, 8, 10):
        for _ in range(50):
            regs = rng.integers(0, 30, size=1 << p).astype(np.uint8)
            regs[rng.random(1 << p) < 0.3] = 0
            assert oracle.estimate(p, regs) == oracle.estimate_packed(p, oracle.pack6(p, regs))


def _bias_table():
    path = os.path.join(os.path.dirname(HERE), "sparkx_amd", "csrc", "hll_bias_table.inc")
    tables = {}
    with open(path) as fh:
        name = None
        for line in fh:
            if line.startswith("HB_BIAS_TABLE_QUAL"):
                name = line.split()[2].split("[")[0]
                tables[name] = []
            elif name and line.strip().startswith("{"):
                tables[name].append([float.fromhex(x) for x in line.strip().strip("{},").split(", ")])
    return np.array(tables["HB_BIAS_RAW"]), np.array(tables["HB_BIAS_VAL"])


def test_bias_table_against_independent_checks():
    """HLL++ bias points (p = 4..16, n_i = i*6m/199): at n = 0 the raw estimate is
    alpha*m exactly (every register empty); a mid point agrees with a Monte-Carlo
    simulation of the Poisson register model; bias = raw - n."""
    raw, val = _bias_table()
    assert raw.shape == (13, 200)
    for p in range(4, 17):
        m = 1 << p
        alpha = {16: 0.673, 32: 0.697, 64: 0.709}.get(m, 0.7213 / (1 + 1.079 / m))
        assert abs(raw[p - 4][0] - alpha * m) <= 1e-9 * alpha * m
        ns = np.arange(200) * 6.0 * m / 199
        assert np.allclose(val[p - 4], raw[p - 4] - ns, rtol=0, atol=1e-9 * m)
        assert np.all(np.diff(raw[p - 4]) > 0)
    rng = np.random.default_rng(11)
    p, m, i = 6, 64, 66
    lam = i * 6.0 / 199
    u = rng.random((200000, m))
    k = np.where(u < math.exp(-lam), 0,
                 np.ceil(-np.log2(np.maximum(np.log(u) / -lam, 1e-300))).clip(1, 60))
    mc = (0.709 * m * m / np.sum(2.0 ** -k, axis=1)).mean()
    assert abs(mc - raw[p - 4][i]) / raw[p - 4][i] < 3e-3


# ---------------------------------------------------------------------------
# epilogues and metrics (closed forms, reference semantics)
# ---------------------------------------------------------------------------
def test_centrality_formulas(oracle):
    est = np.array([5.0, 3.0, 9.0, 2.0], np.float32)
    dist = np.array([8.0, 0.0, 5e-320, 4.0])
    inv = np.array([2.0, 1.0, 1.0, 1.5])
    n = 4
    c1 = oracle.centrality(oracle.CLOSENESS, True, est, dist, inv)
    assert list(c1) == [(5.0 - 1) / 8.0, 0.0, 0.0, (2.0 - 1) / 4.0]
    c0 = oracle.centrality(oracle.CLOSENESS, False, est, dist, inv)
    assert list(c0) == [(n - 1) / 8.0, 0.0, 0.0, (n - 1) / 4.0]
    h1 = oracle.centrality(oracle.HARMONIC, True, est, dist, inv)
    assert list(h1) == [2.0 / 4.0, 0.0, 0.0, 1.5 / 1.0]  # guard on dist, not inv
    h0 = oracle.centrality(oracle.HARMONIC, False, est, dist, inv)
    assert list(h0) == [2.0 / 3.0, 0.0, 0.0, 1.5 / 3.0]
    lin = oracle.centrality(oracle.LIN, False, est, dist, inv)
    assert list(lin) == [16.0 / 8.0, 1.0, 1.0, 1.0 / 4.0]


def test_metrics_reference_semantics(oracle):
    a = np.array([1.0, 1.0, 1.0])
    b = np.array([1.0, 2.0, 3.0])
    assert list(oracle.ranks(a)) == [0, 1, 2]  # stable: ties keep index order
    assert list(oracle.ranks(b)) == [2, 1, 0]  # descending
    assert oracle.spearman(a, b) == 1.0 - 6.0 * 8.0 / (3 * 8.0)
    assert oracle.spearman(np.array([1.0]), np.array([2.0])) == 1.0
    assert oracle.mae(np.array([1.0, 3.0]), np.array([2.0, 1.0])) == 1.5
    assert oracle.mape(np.array([1.0, 3.0]), np.array([2.0, 0.0])) == 100 * 0.5 / 2


# ---------------------------------------------------------------------------
# surrogate MLP oracle and evaluation metrics
# ---------------------------------------------------------------------------
def test_mlp_oracle_backward_matches_finite_differences():
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(5)
    params = rng.normal(0, 0.3, M.N_PARAMS)
    x = rng.normal(0, 1, (40, 6)).astype(np.float32)
    t = rng.normal(0, 1, 40)
    mean, std = np.zeros(6, np.float32), np.ones(6, np.float32)
    _, g = M.loss_and_grad(params, x, t, mean, std, emulate_bf16=False)
    idx = rng.choice(M.N_PARAMS, 60, replace=False).tolist() + [M.B3, M.W1, M.B1 + 3, M.W3 + 7]
    fd = M.finite_difference(params, x, t, mean, std, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-7)


def test_mlp_oracle_bf16_rounding():
    from oracle import mlp_oracle as M
    v = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.14159, 0.0], np.float32)
    # bf16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie -> even (1.0); 1 + 2^-9 rounds down;
    # 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert list(M.bf16(v)[:4]) == [1.0, 1.0, 1.0, 1.0078125]
    assert M.bf16(v)[4] == np.float32(-3.140625)


def test_oracle_ranks_match_stable_sort(oracle):
    """The oracle's merge sort against numpy's stable sort (a library routine) on
    inputs full of ties, signed zeros and infinities."""
    rng = np.random.default_rng(9)
    for n in (1, 2, 17, 1000, 4097):
        a = rng.integers(-3, 4, n).astype(np.float64)  # many ties
        a[rng.random(n) < 0.1] = -0.0
        a[rng.random(n) < 0.02] = np.inf
        a[rng.random(n) < 0.02] = -np.inf
        order = np.argsort(-(a + 0.0), kind="stable")  # a + 0.0 maps -0.0 to +0.0
        want = np.empty(n, np.int64)
        want[order] = np.arange(n)
        assert np.array_equal(oracle.ranks(a).astype(np.int64), want)
        b = rng.normal(size=n)
        rb = oracle.ranks(b).astype(np.int64)
        d = oracle.ranks(a).astype(np.int64) - rb
        ssd = float(sum(int(x) * int(x) for x in d))  # exact below 2^53
        assert oracle.spearman(a, b) == 1.0 - 6.0 * ssd / (n * max(n * n - 1.0, 1.0))


def test_oracle_top_overlap(oracle):
    assert oracle.top_overlap([3, 2, 1, 0], [3, 2, 0, 1], 0.5) == 1.0
    assert oracle.top_overlap([0, 1, 2, 3], [3, 2, 1, 0], 0.25) == 0.0
    assert oracle.top_overlap([1, 1, 1, 1], [0, 0, 0, 1], 0.25) == 0.0  # ties: lowest index first
    assert oracle.top_overlap([1, 1, 1, 1], [1, 0, 0, 0], 0.25) == 1.0
    assert oracle.top_overlap([5.0], [7.0], 0.01) == 1.0  # k = max(1, ceil(0.01))
    v = np.random.default_rng(2).normal(size=1000)
    assert oracle.top_overlap(v, v, 0.013) == 1.0
    assert oracle.top_overlap(v, -v, 0.5) == 0.0


# ---------------------------------------------------------------------------
# walk counts (surrogate inputs): brute force, exact sparse powers, saturation
# ---------------------------------------------------------------------------
def _random_multigraph(rng, n, m):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)  # self-loops and repeated arcs included
    order = np.argsort(src, kind="stable")
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(src, minlength=n))
    return off, dst[order].astype(np.uint32)


def test_walk_counts_brute_force(oracle):
    """S_k(u) counts the walks u -> ... of k + 1 arcs (each arc of a multigraph a
    separate walk step): enumerate them on small multigraphs with self-loops."""
    rng = np.random.default_rng(3)
    for n, m in ((1, 0), (5, 9), (12, 30), (9, 40)):
        off, nb = _random_multigraph(rng, n, m)
        adj = [[int(v) for v in nb[int(off[u]):int(off[u + 1])]] for u in range(n)]

        def walks(u, length):
            return 1 if length == 0 else sum(walks(v, length - 1) for v in adj[u])

        deg, s1, s2, s3, m1 = oracle.walk_counts(off, nb)
        for u in range(n):
            assert int(deg[u]) == walks(u, 1)
            assert int(s1[u]) == walks(u, 2)
            assert int(s2[u]) == walks(u, 3)
            assert int(s3[u]) == walks(u, 4)
            assert int(m1[u]) == max((walks(v, 1) for v in adj[u]), default=0)


def test_walk_counts_match_sparse_powers(oracle):
    """S_k = A^k deg on an R-MAT graph (integer sparse products, far below 2^63);
    M_1 the row maxima of A diag(deg)."""
    from scipy.sparse import csr_matrix

    from sparkx_amd import Graph
    g = Graph.rmat(12, 8, seed=4)
    n = g.size()
    off = g.offsets.astype(np.int64)
    A = csr_matrix((np.ones(g.width(), np.int64), g.neighbours.astype(np.int64), off), shape=(n, n))
    deg = np.diff(off)
    got = oracle.walk_counts(g.offsets, g.neighbours)
    v = deg
    for k in range(1, 4):
        v = A @ v
        assert v.max() < 2 ** 62
        assert np.array_equal(got[k].astype(np.int64), v), k
    m = A.multiply(deg[None, :]).max(axis=1).toarray()[:, 0]
    assert np.array_equal(got[4].astype(np.int64), m)
    assert np.array_equal(got[0].astype(np.int64), deg)
    for passes in (2, 3):  # the inputs are f64 log1p values of these counts
        x = oracle.walk_inputs(g.offsets, g.neighbours, passes)
        ld, l1, l2 = (np.log1p(got[k].astype(np.float64)) for k in range(3))
        assert np.array_equal(x[:, 0], ld.astype(np.float32))
        assert np.array_equal(x[:, 2], l2.astype(np.float32))
        last = (l1 - ld) if passes == 2 else (np.log1p(got[3].astype(np.float64)) - l2)
        assert np.array_equal(x[:, 5], last.astype(np.float32))


def test_walk_counts_saturate():
    """Two nodes joined by d parallel arcs each way: S_k = d^(k+1), clipped at 2^64 - 1
    (d = 2.7e6: S_2 = 1.97e19 and S_3 saturate, S_1 = 7.29e12 does not); M_1 = d."""
    from oracle import hb_oracle as O
    d = 2_700_000
    off = np.array([0, d, 2 * d], np.uint64)
    nb = np.concatenate([np.ones(d, np.uint32), np.zeros(d, np.uint32)])
    deg, s1, s2, s3, m1 = O.walk_counts(off, nb)
    assert [int(x) for x in (deg[0], s1[0], s2[0], s3[0], m1[0])] == [d, d ** 2, O.U64_MAX, O.U64_MAX, d]
    x = O.walk_inputs(off, nb, 3)
    assert x[0, 3] == np.float32(np.log1p(float(O.U64_MAX))) and x[0, 5] == 0.0
    assert x[1, 1] == np.float32(math.log1p(d ** 2))
