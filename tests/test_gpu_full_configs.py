"""Parity at the full sizes of the measured configurations: the R-MAT scale-24 run
the headline figure is quoted on (to convergence, p = 7, 16.8M nodes, 520.8M arcs,
hubs of degree 405,863 cut into hundreds of chunks) and the k-mer de Bruijn run
(4 Mbp genome, 8 copies with 0.5% substitutions, k = 21, p = 6, truncated at 64
iterations).  The GPU run (default schedule: degree layout, frontier, selective
gathers, push iterations, hub chunks) is compared with the oracle's restatement of
hyper_ball.rs:754-874 run on the host's threads: iteration count, per-iteration
increase counts, both register buffers, estimates and both f64 accumulators, all
bit for bit.  The oracle also reports how many of the (node, iteration) estimates
fell in each regime of the estimator (linear counting, raw with the HLL++ bias
subtracted, raw as is)."""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _full_parity(oracle, g, p, max_iters, name):
    from sparkx_amd import HyperBall, _lib
    say(name, "nodes", g.size(), "arcs", g.width(), "p", p, "max_iters", max_iters)
    t0 = time.time()
    with HyperBall(g, p, max_iters=max_iters) as h:
        st = h.stats
        say(name, "gpu", st.iterations, "iterations")
        t1 = time.time()
        ref = oracle.hyperball(g.offsets.astype(np.uint64), g.neighbours, p, threads=THREADS,
                               max_iters=max_iters)
        lc, bias, raw = ref.regime
        total = max(1, lc + bias + raw)
        say(name, f"oracle {ref.iterations} iterations in {time.time() - t1:.1f}s on {THREADS} threads; "
            f"estimates after the union: linear counting {lc / total:.4f}, raw - bias {bias / total:.4f}, "
            f"raw {raw / total:.4f} of {total}")
        assert st.iterations == ref.iterations
        assert list(st.increases[:st.iterations]) == list(ref.increases)
        assert bool(st.final_parity_latest) == ref.final_parity_latest
        assert np.array_equal(h.registers(_lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref)
        assert np.array_equal(h.registers(_lib.HB_REGS_LATEST).ravel(), ref.regs_last)
        dist, inv = h._accumulators()
        assert np.array_equal(dist, ref.dist)
        assert np.array_equal(inv, ref.inv)
        est = h.estimates(_lib.HB_REGS_REFERENCE)
        assert np.array_equal(est, oracle.estimates(p, ref.regs_ref))
    say(name, f"bit-identical ({time.time() - t0:.1f}s)")
    return ref


def test_rmat24_p7_config(oracle):
    """R-MAT scale 24, edge factor 16, (a, b, c) = (.57, .19, .19), seed 1, symmetrised,
    deduplicated, loop-free; p = 7 to convergence."""
    from sparkx_amd import Graph
    g = Graph.rmat(24, 16, 0.57, 0.19, 0.19, seed=1)
    assert (g.size(), g.width()) == (16_777_216, 520_768_654)
    deg = np.diff(g.offsets)
    assert int(deg.max()) == 405_863
    ref = _full_parity(oracle, g, 7, 0, "rmat24_p7")
    assert ref.iterations == 8


def test_kmer_config_t64(oracle):
    """k-mer de Bruijn graph: 4,000,000 bp genome, 8 copies with 0.5% substitutions,
    k = 21, seed 1; p = 6 truncated at t = 64."""
    from sparkx_amd import Graph
    g = Graph.kmer_dbg(4_000_000, 8, 0.005, 21, seed=1)
    assert (g.size(), g.width()) == (7_183_108, 14_653_044)
    ref = _full_parity(oracle, g, 6, 64, "kmer_p6_t64")
    assert ref.iterations == 64
