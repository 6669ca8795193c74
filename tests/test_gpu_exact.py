"""GPU parity of the exact BFS centralities (hb_exact_bfs / hb_exact_centrality)
against the oracle's BFS (trails/bfs.rs:199-234 with its source quirk) and the
reference's committed exact-closeness fixtures.  Integer sums and every f64 value
must be bit-identical."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    from sparkx_amd import _lib
    _lib.load()
    import sparkx_amd
    return sparkx_amd


def _check_sources(hb, oracle, g, lo, hi, ctx):
    from sparkx_amd.exact import exact_bfs
    off = g.offsets.astype(np.uint64)
    nb = g.neighbours.astype(np.uint32)
    r_ref, t_ref, i_ref = oracle.exact_bfs(off, nb, lo, hi)
    r, t, i = exact_bfs(g, (lo, hi))
    assert np.array_equal(r, r_ref), ctx
    assert np.array_equal(t.astype(np.float64), t_ref), ctx
    assert np.array_equal(i, i_ref), ctx
    return r_ref, t_ref, i_ref


@pytest.mark.parametrize("sigma,k,fixture,ranges", [
    (20, 2, "exact_closeness_dbg_sigma20_k2.f64", [(0, 400)]),
    (4, 7, "exact_closeness_dbg_sigma4_k7.f64", [(0, 130), (5001, 5300), (16384 - 200, 16384)])])
def test_fixture_exact_closeness_bit_exact(hb, oracle, sigma, k, fixture, ranges):
    """The reference's own exact-closeness files (test_common.rs:195-249), reproduced
    by the GPU BFS to the last bit, plus harmonic and Lin against the oracle."""
    from sparkx_amd.exact import ExactClosenessCentrality, ExactHarmonicCentrality, ExactLinCentrality
    off, nb = oracle.dbg_mtx_csr(sigma, k)
    g = hb.Graph(off, nb.astype(np.uint64), f"dbg{sigma}_{k}")
    fx = np.fromfile(os.path.join(GOLDEN, fixture), dtype="<f8")
    n = g.size()
    for lo, hi in ranges:
        say("fixture", sigma, k, lo, hi)
        c = ExactClosenessCentrality(g, True, sources=(lo, hi)).centralities()
        assert np.array_equal(c, fx[lo:hi])
        r, t, i = _check_sources(hb, oracle, g, lo, hi, (sigma, k, lo))
        for nz in (None, False, True):
            assert np.array_equal(ExactClosenessCentrality(g, nz, sources=(lo, hi)).centralities(),
                                  oracle.exact_centrality(0, nz, r, t, i, n)), nz
            assert np.array_equal(ExactHarmonicCentrality(g, nz, sources=(lo, hi)).centralities(),
                                  oracle.exact_centrality(1, nz, r, t, i, n)), nz
        assert np.array_equal(ExactLinCentrality(g, sources=(lo, hi)).centralities(),
                              oracle.exact_centrality(2, None, r, t, i, n))


def _csr(n, arcs):
    adj = [[] for _ in range(n)]
    for a, b in arcs:
        adj[a].append(b)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in adj])
    return off, np.array([v for x in adj for v in x], np.uint64)


@pytest.mark.parametrize("name", ["directed_random", "directed_cycle", "self_loops", "isolated", "path", "star",
                                  "long_path", "directed_hub"])
def test_exact_bfs_edge_cases(hb, oracle, name):
    rng = np.random.default_rng(7)
    if name == "directed_random":
        n = 3000
        arcs = [(int(a), int(b)) for a, b in rng.integers(0, n, size=(7000, 2))]
    elif name == "directed_cycle":
        n = 200
        arcs = [(i, (i + 1) % n) for i in range(n)]
    elif name == "self_loops":
        n = 130
        arcs = [(i, i) for i in range(0, n, 3)] + [(i, i + 1) for i in range(n - 1)]
    elif name == "isolated":
        n = 100
        arcs = [(3, 4), (4, 3)]
    elif name in ("path", "long_path"):
        # 4500 levels: the 1/d table of the harmonic sums no longer fits in LDS
        n = 700 if name == "path" else 4500
        arcs = [(i, i + 1) for i in range(n - 1)] + [(i + 1, i) for i in range(n - 1)]
    elif name == "directed_hub":
        # out-degree 9000 (> 4096: its list is pushed in chunks) and a random sparse rest
        n = 9001
        arcs = [(0, i) for i in range(1, n)] + [(int(a), int(b)) for a, b in rng.integers(1, n, size=(6000, 2))]
    else:
        n = 5000
        arcs = [(0, i) for i in range(1, n)] + [(i, 0) for i in range(1, n)]
    off, nb = _csr(n, arcs)
    g = hb.Graph(off, nb, name)
    say("exact edge case", name, n)
    _check_sources(hb, oracle, g, 0, min(n, 200), name)
    if n > 200:
        _check_sources(hb, oracle, g, n - 70, n, name)


def test_exact_bfs_rmat_ranges(hb, oracle):
    """R-MAT 2^14 (hubs, many isolated nodes): source ranges that start and end off
    the 64-source boundaries."""
    g = hb.Graph.rmat(14, 16, seed=21)
    say("exact rmat14", g.size(), g.width())
    for lo, hi in ((0, 64), (5, 200), (16000, 16384)):
        _check_sources(hb, oracle, g, lo, hi, (lo, hi))


def test_exact_errors(hb):
    from sparkx_amd import HyperBallError, _lib
    from sparkx_amd.exact import ExactClosenessCentrality, exact_bfs
    g = hb.Graph.erdos_renyi(500, 1500, seed=5)
    with pytest.raises(ValueError):
        ExactClosenessCentrality(g, True, sources=(0, 501))
    with pytest.raises(HyperBallError) as ei:
        exact_bfs(g, (10, 600))
    assert ei.value.code == _lib.HB_EINVAL
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), 0))
    try:
        out = np.zeros(4)
        assert L.hb_exact_centrality(ctx, 0, 1, 0, 4, _lib.ptr(out, C.c_double)) == _lib.HB_ESTATE
        assert L.hb_exact_centrality(ctx, 7, 1, 0, 4, _lib.ptr(out, C.c_double)) == _lib.HB_EINVAL
        assert L.hb_exact_centrality(ctx, 0, 2, 0, 4, _lib.ptr(out, C.c_double)) == _lib.HB_EINVAL
    finally:
        L.hb_destroy(ctx)
    # empty range
    r, t, i = exact_bfs(g, (3, 3))
    assert r.size == 0
