"""GPU parity of the walk-count surrogate inputs (hb_walk_features /
hb_surrogate_walks) against the oracle's walk_counts / walk_inputs.

The counts are saturating u64 sums, exact and independent of how a list is split
over lanes, wavefronts and hub chunks, so every integer must match; each input is
an f64 log1p of an integer (or a difference of two) rounded to f32, where the
device's and the host's log1p may differ in the last f64 bit: logs within 1 f32
ulp, differences within 1e-6 relative + 1e-9.  Cases cover the lane
(degree <= 64), wavefront (<= 4096) and hub-chunk (> 4096) paths, both node
numberings, self-loops, repeated arcs, isolated nodes and the saturating case."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    from sparkx_amd import _lib
    _lib.load()
    import sparkx_amd
    return sparkx_amd


def _csr(n, src, dst):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    order = np.argsort(src, kind="stable")
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(src, minlength=n))
    return off, dst[order].astype(np.uint32)


def _graphs(hb):
    rng = np.random.default_rng(11)
    yield "rmat14", hb.Graph.rmat(14, 16, seed=5)
    n = 3000
    off, nb = _csr(n, rng.integers(0, n, 20000), rng.integers(0, n, 20000))
    yield "directed_multigraph", hb.Graph(off, nb, "dm")
    d = 2_700_000  # two nodes, d parallel arcs each way: S_2 = d^3 saturates
    yield "saturating", hb.Graph(np.array([0, d, 2 * d], np.uint64),
                                 np.concatenate([np.ones(d, np.uint32), np.zeros(d, np.uint32)]), "sat")
    n = 10001  # a star: the centre's list is three hub chunks
    src = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, n)])
    dst = np.concatenate([np.arange(1, n), np.zeros(n - 1, np.int64)])
    off, nb = _csr(n, src, dst)
    yield "star", hb.Graph(off, nb, "star")
    off, nb = _csr(50, [3, 3, 7], [7, 3, 3])  # isolated nodes and a self-loop
    yield "isolated", hb.Graph(off, nb, "iso")


@pytest.mark.parametrize("layout", ["degree", "chain", "input"])
@pytest.mark.parametrize("passes", [2, 3])
def test_walk_features_match_oracle(hb, oracle, layout, passes):
    from sparkx_amd.surrogate import walk_features
    diff_cols = [4, 5] if passes == 2 else [5]
    for name, g in _graphs(hb):
        say("walk features", name, layout, passes, g.size(), g.width())
        want = oracle.walk_inputs(g.offsets, g.neighbours, passes)
        logs = [k for k in range(6) if k not in diff_cols]
        with hb.HyperBall(g, 4, _run=False, layout=layout) as h:
            got, ms = walk_features(h, passes)
            assert ms >= 0
            np.testing.assert_array_max_ulp(got[:, logs], want[:, logs], maxulp=1)
            np.testing.assert_allclose(got[:, diff_cols], want[:, diff_cols], rtol=1e-6, atol=1e-9)
            # again on the same context (its hub table kept), then in the other numbering
            assert np.array_equal(walk_features(h, passes)[0], got)
            h.set_layout("input" if layout == "degree" else "degree")
            other, _ = walk_features(h, passes)
            np.testing.assert_array_max_ulp(other[:, logs], want[:, logs], maxulp=1)
            np.testing.assert_allclose(other[:, diff_cols], want[:, diff_cols], rtol=1e-6, atol=1e-9)


def test_walk_features_empty_graph(hb):
    from sparkx_amd.surrogate import walk_features
    g = hb.Graph(np.zeros(1, np.uint64), np.zeros(0, np.uint32), "empty")
    with hb.HyperBall(g, 4, _run=False) as h:
        x, ms = walk_features(h)
    assert x.shape == (0, 6) and ms == 0.0


def test_surrogate_walks_matches_host_pipeline(hb, oracle):
    """Device standardisation + inference equal the host-standardised inputs through
    hb_mlp_predict; the context's HyperBall run is left as it was."""
    from sparkx_amd.surrogate import Surrogate, standardize
    g = hb.Graph.rmat(13, 16, seed=8)
    s = Surrogate(seed=3)
    linked = g.degrees() > 0
    with hb.HyperBall(g, 6) as h:
        est0 = h.estimates().copy()
        pred, feat, ms = s.rank_walks(h, 2, want_features=True)
        assert np.array_equal(h.estimates(), est0)
    assert ms > 0
    host = standardize(oracle.walk_inputs(g.offsets, g.neighbours, 2), linked)
    assert np.allclose(feat, host, rtol=1e-5, atol=1e-5)
    assert np.array_equal(pred[linked], s.predict(feat[linked]))
    assert np.all(pred[~linked] == np.float32(-3e38))
    s.close()


@pytest.mark.parametrize("passes", [2, 3])
def test_surrogate_walks_ranks_held_out_graph(hb, oracle, passes):
    """Trained on small R-MAT graphs, the walk-count network ranks a held-out
    graph's harmonic centrality with Spearman rho >= 0.95, and better than ranking by
    any one of its own inputs alone (degree, the walk counts S_1..S_passes, the
    largest neighbour degree, the differences): the network adds to what it is fed."""
    from sparkx_amd.surrogate import Surrogate, labelled_walks
    xs, ys = [], []
    for seed in range(8):
        say("walks train graph", seed, passes)
        g = hb.Graph.rmat(14, 16, seed=100 + seed)
        x, y, _, _ = labelled_walks(g, passes)
        linked = g.degrees() > 0
        xs.append(x[linked])
        ys.append(y[linked])
    s = Surrogate(seed=7)
    losses = s.train(np.concatenate(xs), np.concatenate(ys), epochs=20, batch=4096, lr=1e-3)
    say("walks losses", losses)
    assert losses[-1] < losses[0]
    g = hb.Graph.rmat(15, 16, seed=999)
    _, _, _, truth = labelled_walks(g, passes)
    with hb.HyperBall(g, 4, _run=False) as h:
        pred, _, _ = s.rank_walks(h, passes)
    linked = g.degrees() > 0
    rho = oracle.spearman(pred[linked].astype(np.float64), truth[linked])
    raw = oracle.walk_inputs(g.offsets, g.neighbours, passes)[linked].astype(np.float64)
    rho_inputs = [oracle.spearman(raw[:, k], truth[linked]) for k in range(raw.shape[1])]
    say("walks rho", rho, "single inputs", rho_inputs)
    assert rho >= 0.95 and rho > max(rho_inputs)
    s.close()


def test_walk_errors(hb):
    """Bad arguments come back as status codes with a message, nothing runs."""
    import ctypes as C
    from sparkx_amd import _lib
    from sparkx_amd.surrogate import Surrogate
    L = _lib.load()
    g = hb.Graph.erdos_renyi(500, 1500, seed=3)
    x = np.empty((g.size(), 6), np.float32)
    ms = C.c_double()
    s = Surrogate(seed=1)
    with hb.HyperBall(g, 4, _run=False) as h:
        for passes in (0, 1, 4):
            assert L.hb_walk_features(h._ctx, passes, _lib.ptr(x, C.c_float), C.byref(ms)) == _lib.HB_EINVAL
            assert b"passes" in L.hb_last_error()
            assert L.hb_surrogate_walks(h._ctx, s._m, passes, None, None, C.byref(ms)) == _lib.HB_EINVAL
        assert L.hb_surrogate_walks(h._ctx, None, 2, None, None, C.byref(ms)) == _lib.HB_EINVAL
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), 0))
    try:  # nothing staged
        assert L.hb_walk_features(ctx, 2, None, C.byref(ms)) == _lib.HB_ESTATE
        assert L.hb_surrogate_walks(ctx, s._m, 2, None, None, C.byref(ms)) == _lib.HB_ESTATE
    finally:
        L.hb_destroy(ctx)
        s.close()
