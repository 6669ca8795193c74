"""HyperBall on MI355X behind the reference's HyperBallInner interface.

Mirrors `HyperBallInner<'a, N, E, Ix, P, B = 6>` (src/centralities/hyper_ball.rs:81-108)
and its aliases HyperBall4..HyperBall16 (:46-58): same constructor names,
argument meaning, accessors, epilogues and error behaviour, over the C ABI of
libsparkx_hb.so.  All numerics run in the HIP engine; this module only moves
host buffers and keeps the reference's caching rules for centralities.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _lib
from .graph import Graph


class HyperBall:
    """Register-exact HyperBall (HLL++ counters with 2^p six-bit registers).

    new(g) / new_with_conf(g, max_depth, threads) run the algorithm at once,
    like the reference (hyper_ball.rs:124-157).  `max_depth` and `threads` are
    accepted for signature parity; the reference ignores max_depth (:710-712 is
    dead code) and threads only partition its CPU loop, so neither changes
    results.  `max_iters` (an extension) truncates the run.
    """

    PRECISION = 8  # GraphMemoryMap::hyperball() uses HyperBall8 (impl_algos.rs:45-48)

    def __init__(self, g: Graph, p: int, *, max_iters: int = 0, dense: bool = False,
                 graph_capture: bool = False, timing: bool = False, nf: bool = False,
                 ktiming: bool = False, layout: str = "auto",
                 device: int = 0, schedule: Optional[tuple] = None, _run: bool = True, _stage_fn=None):
        if not 4 <= p <= 16:
            raise ValueError(f"precision p={p} outside 4..16")
        self.g = g
        self.p = p
        self._L = _lib.load()
        self._ctx = C.c_void_p()
        _lib.check(self._L.hb_create(C.byref(self._ctx), device))
        self._closeness = [None, None]
        self._harmonic = [None, None]
        self._lin = None
        self.stats = None
        self._cache = None  # (cache_dir, graph id) when opened through get_or_compute
        try:
            self.set_layout(layout)
            if schedule is not None:
                _lib.check(self._L.hb_set_schedule(self._ctx, *[int(x) for x in schedule]))
            if _stage_fn is None:
                self._stage()
            else:  # staged on the device (MatrixMarket text, edge list); keep the host CSR
                self.g, self.ingest_info = _stage_fn(self._L, self._ctx)
            if _run:
                self.run(max_iters=max_iters, dense=dense, graph_capture=graph_capture, timing=timing,
                         nf=nf, ktiming=ktiming)
        except Exception:
            self.close()
            raise

    # ---- constructors (hyper_ball.rs:124-201) ---------------------------------
    @classmethod
    def new(cls, g: Graph, p: Optional[int] = None, **kw) -> "HyperBall":
        return cls(g, cls.PRECISION if p is None else p, **kw)

    @classmethod
    def new_with_conf(cls, g: Graph, max_depth: Optional[int], threads: int,
                      p: Optional[int] = None, **kw) -> "HyperBall":
        del max_depth, threads  # no effect on results in the reference either
        return cls(g, cls.PRECISION if p is None else p, **kw)

    @classmethod
    def get_or_compute(cls, g: Graph, cache_dir: str, gid: str, p: Optional[int] = None, **kw) -> "HyperBall":
        """HyperBallInner::get_or_compute (hyper_ball.rs:165-201).  When cache_dir holds
        the registers and both accumulators of graph `gid` (hyperballregisters_<p>_<id>,
        this library's own file -- the reference's counters file holds hyperloglog-rs
        structs whose layout is not available, so it is neither written nor read --
        and hyperballdistances_<id> / hyperballinvdistances_<id>, the reference's
        names), the run is rebuilt from them on the device and every cached
        centrality is searched for, as the reference does; the rebuilt run reports 0
        iterations, as the reference's does.  Otherwise HyperBall runs and its
        registers and accumulators are saved there (the reference's run creates its
        files).  Either way the object keeps the cache: compute_* write their files."""
        from . import cache
        p = cls.PRECISION if p is None else p
        n = g.size()
        files = [cache.cache_file(cache_dir, "registers", gid, p), cache.cache_file(cache_dir, "distances", gid),
                 cache.cache_file(cache_dir, "inv_distances", gid)]
        if all(os.path.exists(f) for f in files):
            regs = np.fromfile(files[0], np.uint8)
            dist = np.fromfile(files[1], "<f8")
            inv = np.fromfile(files[2], "<f8")
            if regs.size == n << p and dist.size == n and inv.size == n:
                h = cls(g, p, _run=False, **kw)
                try:
                    _lib.check(h._L.hb_restore(h._ctx, p, _lib.ptr(regs, C.c_uint8), _lib.ptr(dist, C.c_double),
                                               _lib.ptr(inv, C.c_double)))
                    h._ran(None)
                    h._cache = (cache_dir, gid)
                    for norm in (False, True):
                        for search in (h.search_cache_closeness_centrality, h.search_cache_harmonic_centrality):
                            try:
                                search(norm)
                            except LookupError:
                                pass
                    try:
                        h.search_cache_lins_centrality()
                    except LookupError:
                        pass
                except Exception:
                    h.close()
                    raise
                return h
        h = cls(g, p, **kw)
        h._cache = (cache_dir, gid)
        cache.write_results(h, cache_dir, gid)
        return h

    @classmethod
    def from_mtx(cls, text, p: Optional[int] = None, **kw) -> "HyperBall":
        """Parses a MatrixMarket text (bytes, or a path) on the GPU, stages the CSR it
        describes and runs, as GraphCache::from_mtx_file followed by new() would
        (parser_mtx/mod.rs:34-162).  `self.ingest_info` holds the parse summary."""
        from .ingest import stage_mtx_fn
        return cls(None, cls.PRECISION if p is None else p, _stage_fn=stage_mtx_fn(text), **kw)

    # ---- engine ---------------------------------------------------------------
    LAYOUTS = {"input": _lib.HB_LAYOUT_INPUT, "degree": _lib.HB_LAYOUT_DEGREE, "chain": _lib.HB_LAYOUT_CHAIN,
               "auto": _lib.HB_LAYOUT_AUTO}

    def set_layout(self, layout: str) -> None:
        """Node numbering on the device, built on the GPU when the graph is staged:
        "degree" (nodes by descending degree), "chain" (nodes along the paths of a
        multi-source breadth-first search, for long low-degree paths such as de Bruijn
        graphs), "auto" (default: "chain" when every degree is <= 64 and the graph has at
        most 4 arcs per node, "degree" otherwise) or "input".  Results are the same
        either way and always come back in the graph's own node order."""
        if layout not in self.LAYOUTS:
            raise ValueError(f"layout {layout!r} not in {sorted(self.LAYOUTS)}")
        _lib.check(self._L.hb_set_layout(self._ctx, self.LAYOUTS[layout]))

    def _stage(self):
        g = self.g
        off, nb = g.offsets, g.neighbours
        nbp = nb.ctypes.data if nb.size else None
        _lib.check(self._L.hb_stage_csr(self._ctx, off.ctypes.data, off.itemsize, g.size(),
                                        nbp, nb.itemsize, g.width()))

    def run(self, *, max_iters: int = 0, dense: bool = False, graph_capture: bool = False,
            timing: bool = False, nf: bool = False, ktiming: bool = False) -> "_lib.HbStats":
        flags = ((_lib.HB_FLAG_DENSE if dense else 0) | (_lib.HB_FLAG_GRAPH if graph_capture else 0)
                 | (_lib.HB_FLAG_TIMING if timing else 0) | (_lib.HB_FLAG_NF if nf else 0)
                 | (_lib.HB_FLAG_KTIMING if ktiming else 0))
        st = _lib.HbStats()
        _lib.check(self._L.hb_run(self._ctx, self.p, max_iters, flags, C.byref(st)))
        return self._ran(st)

    def _ran(self, st: "_lib.HbStats") -> "_lib.HbStats":
        """A run of this context completed (here or through hb_run_many): new results."""
        self.stats = st
        self._closeness = [None, None]
        self._harmonic = [None, None]
        self._lin = None
        return st

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.hb_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- results --------------------------------------------------------------
    def get_iters(self) -> int:
        """Iterations executed, including the final one with no increase (hyper_ball.rs:557-560)."""
        return int(self.result_info()[1])

    def _replaced(self) -> None:
        """The engine's run on this graph may have been replaced by one whose statistics
        are not returned (the surrogate's 4 iterations): drop the cached results and
        take the precision of whatever run the engine now holds."""
        self._ran(None)
        try:
            self.p = self.result_info()[0]
        except _lib.HyperBallError:
            pass  # no completed run: the accessors raise HB_ESTATE

    def result_info(self):
        """(p, iterations, snapshot rows) of the results the engine holds for this graph."""
        p, it, snaps = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self._L.hb_result_info(self._ctx, C.byref(p), C.byref(it), C.byref(snaps)))
        return p.value, it.value, snaps.value

    def registers(self, which: int = _lib.HB_REGS_REFERENCE) -> np.ndarray:
        # sized from the engine's completed run, whose p is the one the registers have
        out = np.empty((self.g.size(), 1 << self.result_info()[0]), np.uint8)
        _lib.check(self._L.hb_read_registers(self._ctx, which, _lib.ptr(out, C.c_uint8)))
        return out

    def estimates(self, which: int = _lib.HB_REGS_REFERENCE) -> np.ndarray:
        out = np.empty(self.g.size(), np.float32)
        _lib.check(self._L.hb_read_estimates(self._ctx, which, _lib.ptr(out, C.c_float)))
        return out

    def _accumulators(self):
        n = self.g.size()
        dist = np.empty(n, np.float64)
        inv = np.empty(n, np.float64)
        _lib.check(self._L.hb_read_accumulators(self._ctx, _lib.ptr(dist, C.c_double), _lib.ptr(inv, C.c_double)))
        return dist, inv

    # hyper_ball.rs:272-280
    def geodesic_distances(self) -> np.ndarray:
        return self._accumulators()[0]

    def harmonic_geodesic_distances(self) -> np.ndarray:
        return self._accumulators()[1]

    # hyper_ball.rs:260-270: assert!(node_id < self.g.size())
    def get_geodesic_distance(self, node_id: int) -> float:
        if not 0 <= node_id < self.g.size():
            raise AssertionError(f"node_id {node_id} >= |V| = {self.g.size()}")
        return float(self.geodesic_distances()[node_id])

    def get_harmonic_geodesic_distance(self, node_id: int) -> float:
        if not 0 <= node_id < self.g.size():
            raise AssertionError(f"node_id {node_id} >= |V| = {self.g.size()}")
        return float(self.harmonic_geodesic_distances()[node_id])

    def _centrality(self, kind: int, normalize: bool) -> np.ndarray:
        out = np.empty(self.g.size(), np.float64)
        _lib.check(self._L.hb_centrality(self._ctx, kind, 1 if normalize else 0, _lib.ptr(out, C.c_double)))
        return out

    def _saved(self, kind: str, normalize: bool, values: np.ndarray) -> np.ndarray:
        """A computed centrality also goes to its cache file when the object has a cache
        (the reference's compute_* write into their file, hyper_ball.rs:441-445)."""
        if self._cache is not None:
            from . import cache
            cache.write_centrality(self._cache[0], kind, self._cache[1], normalize, values)
        return values

    # hyper_ball.rs:434-477
    def compute_closeness_centrality(self, normalize: bool) -> np.ndarray:
        self._closeness[int(normalize)] = self._saved("closeness", normalize,
                                                      self._centrality(_lib.HB_CLOSENESS, normalize))
        return self._closeness[int(normalize)]

    # hyper_ball.rs:487-529
    def compute_harmonic_centrality(self, normalize: bool) -> np.ndarray:
        self._harmonic[int(normalize)] = self._saved("harmonic", normalize,
                                                     self._centrality(_lib.HB_HARMONIC, normalize))
        return self._harmonic[int(normalize)]

    # hyper_ball.rs:532-555
    def compute_lins_centrality(self) -> np.ndarray:
        self._lin = self._saved("lin", False, self._centrality(_lib.HB_LIN, False))
        return self._lin

    # hyper_ball.rs:302-317, :359-374, :405-415: the cached file into its slot, or Err
    def _search(self, kind: str, normalize: bool) -> Optional[np.ndarray]:
        if self._cache is None:
            return None
        from . import cache
        path = cache.cache_file(self._cache[0], kind, self._cache[1], cache.centrality_seq(kind, normalize))
        if not os.path.exists(path):
            return None
        v = np.fromfile(path, "<f8")
        return v if v.size == self.g.size() else None

    def search_cache_closeness_centrality(self, normalization: bool) -> np.ndarray:
        v = self._search("closeness", normalization)
        if v is None:
            raise LookupError("error no such cached centrality")
        self._closeness[int(normalization)] = v
        return v

    def search_cache_harmonic_centrality(self, normalization: bool) -> np.ndarray:
        v = self._search("harmonic", normalization)
        if v is None:
            raise LookupError("error no such cached centrality")
        self._harmonic[int(normalization)] = v
        return v

    def search_cache_lins_centrality(self) -> np.ndarray:
        v = self._search("lin", False)
        if v is None:
            raise LookupError("error no such cached centrality")
        self._lin = v
        return v

    # hyper_ball.rs:290-299, :347-356, :397-402: Err when not computed
    def get_closeness_centrality(self, normalization: bool) -> np.ndarray:
        v = self._closeness[int(normalization)]
        if v is None:
            raise LookupError("error no such cached centrality")
        return v

    def get_hamonic_centrality(self, normalization: bool) -> np.ndarray:
        v = self._harmonic[int(normalization)]
        if v is None:
            raise LookupError("no such cached centrality")
        return v

    get_harmonic_centrality = get_hamonic_centrality

    def get_lins_centrality(self) -> np.ndarray:
        if self._lin is None:
            raise LookupError("no such cached centrality")
        return self._lin

    # hyper_ball.rs:327-337, :384-394, :418-424.  The closeness variant tests the
    # *harmonic* slot (:332), reproduced: with that slot empty it recomputes, otherwise
    # it returns the cached closeness and fails like get_closeness_centrality if there
    # is none.
    def get_or_compute_closeness_centrality(self, normalization: bool) -> np.ndarray:
        if self._harmonic[int(normalization)] is None:
            return self.compute_closeness_centrality(normalization)
        return self.get_closeness_centrality(normalization)

    def get_or_compute_harmonic_centrality(self, normalization: bool) -> np.ndarray:
        if self._harmonic[int(normalization)] is None:
            return self.compute_harmonic_centrality(normalization)
        return self._harmonic[int(normalization)]

    def get_or_compute_lins_centrality(self) -> np.ndarray:
        if self._lin is None:
            return self.compute_lins_centrality()
        return self._lin

    def autotune(self, trial_iters: int = 3):
        """Picks the fastest degree-bin / hub-chunk schedule for this graph (results are
        unchanged by any choice); call before run()."""
        s, w, k, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
        _lib.check(self._L.hb_autotune(self._ctx, self.p, trial_iters, C.byref(s), C.byref(w), C.byref(k),
                                       C.byref(ms)))
        return (s.value, w.value, k.value), ms.value

    def neighbourhood_function(self) -> np.ndarray:
        """N(t) = sum_u |B(u, t)| estimates, t = 1..iterations (needs nf=True)."""
        it = self.get_iters()
        out = np.empty(max(it, 1), np.float64)
        k = _lib.check(self._L.hb_neighbourhood_function(self._ctx, _lib.ptr(out, C.c_double), it))
        return out[:k]


def _alias(p: int):
    class _HB(HyperBall):
        PRECISION = p

    _HB.__name__ = _HB.__qualname__ = f"HyperBall{p}"
    return _HB


# hyper_ball.rs:46-58
HyperBall4, HyperBall5, HyperBall6, HyperBall7, HyperBall8, HyperBall9, HyperBall10 = (
    _alias(p) for p in range(4, 11))
HyperBall11, HyperBall12, HyperBall13, HyperBall14, HyperBall15, HyperBall16 = (
    _alias(p) for p in range(11, 17))
