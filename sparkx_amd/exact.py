"""Exact BFS centralities on MI355X behind the reference's exact-centrality interface.

Mirrors ExactClosenessCentrality, ExactHarmonicCentrality and ExactLinCentrality
(src/centralities/exact.rs:10-226, :228-466, :468-644): `new(g, normalized)` runs
at once, `centralities()` and `get_node_centrality(idx)` read the result.  The
searches run in the HIP engine (hb_exact_bfs / hb_exact_centrality of
libsparkx_hb.so), 64 sources per pass; `sources` (an extension) restricts the run
to a contiguous range of source nodes, e.g. a sample for rank correlation on
graphs where one BFS per node is too long.  Values follow BFSDists
(src/trails/bfs.rs:199-234) including its quirk: the source is rediscovered and
counted.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .graph import Graph

_NORM = {None: -1, False: 0, True: 1}


class _Exact:
    KIND = -1

    def __init__(self, g: Graph, normalized: Optional[bool] = None, *,
                 sources: Optional[Tuple[int, int]] = None, device: int = 0):
        if normalized not in _NORM:
            raise TypeError("normalized must be None, False or True")
        self.g = g
        self.normalized = normalized
        n = g.size()
        self.sources = (0, n) if sources is None else (int(sources[0]), int(sources[1]))
        if not 0 <= self.sources[0] <= self.sources[1] <= n:
            raise ValueError(f"sources {self.sources} outside [0, {n}]")
        self._L = _lib.load()
        self._ctx = C.c_void_p()
        _lib.check(self._L.hb_create(C.byref(self._ctx), device))
        try:
            # the searches walk the caller's CSR: no degree-ordered copy is built
            _lib.check(self._L.hb_set_layout(self._ctx, _lib.HB_LAYOUT_INPUT))
            off, nb = g.offsets, g.neighbours
            _lib.check(self._L.hb_stage_csr(self._ctx, off.ctypes.data, off.itemsize, n,
                                            nb.ctypes.data if nb.size else None, nb.itemsize, g.width()))
            lo, hi = self.sources
            self._c = np.zeros(hi - lo, np.float64)
            _lib.check(self._L.hb_exact_centrality(self._ctx, self.KIND, _NORM[normalized], lo, hi,
                                                   _lib.ptr(self._c, C.c_double)))
        finally:
            self._L.hb_destroy(self._ctx)
            self._ctx = C.c_void_p()

    @classmethod
    def new(cls, g: Graph, normalized: Optional[bool] = None, **kw):
        return cls(g, normalized, **kw)

    def centralities(self) -> np.ndarray:
        return self._c

    def get_node_centrality(self, idx: int) -> float:
        lo, hi = self.sources
        if not lo <= idx < hi:
            raise IndexError(f"node {idx} outside the computed sources [{lo}, {hi})")
        return float(self._c[idx - lo])


class ExactClosenessCentrality(_Exact):
    """None: total distance; Some(false): (n-1)/total; Some(true): reachable/total
    (exact.rs:160-201); 0 when reachable <= 1 or total == 0."""
    KIND = _lib.HB_CLOSENESS


class ExactHarmonicCentrality(_Exact):
    """None: sum of 1/d in node order; Some(false): (n-1)/sum; Some(true):
    reachable/sum (exact.rs:379-441); 0 when reachable <= 1."""
    KIND = _lib.HB_HARMONIC


class ExactLinCentrality(_Exact):
    """reachable^2 / total, 0 when total == 0 (exact.rs:610-619); takes no
    normalisation flag in the reference."""
    KIND = _lib.HB_LIN

    def __init__(self, g: Graph, *, sources=None, device: int = 0):
        super().__init__(g, None, sources=sources, device=device)

    @classmethod
    def new(cls, g: Graph, **kw):
        return cls(g, **kw)


def exact_bfs(g: Graph, sources: Optional[Tuple[int, int]] = None, inv_sum: bool = True, device: int = 0):
    """Per-source (reachable u64, total distance u64, node-order sum of 1/d f64 or None)."""
    L = _lib.load()
    n = g.size()
    lo, hi = (0, n) if sources is None else sources
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), device))
    try:
        _lib.check(L.hb_set_layout(ctx, _lib.HB_LAYOUT_INPUT))
        off, nb = g.offsets, g.neighbours
        _lib.check(L.hb_stage_csr(ctx, off.ctypes.data, off.itemsize, n, nb.ctypes.data if nb.size else None,
                                  nb.itemsize, g.width()))
        reach = np.zeros(hi - lo, np.uint64)
        total = np.zeros(hi - lo, np.uint64)
        inv = np.zeros(hi - lo, np.float64) if inv_sum else None
        _lib.check(L.hb_exact_bfs(ctx, lo, hi, _lib.ptr(reach, C.c_uint64), _lib.ptr(total, C.c_uint64),
                                  _lib.ptr(inv, C.c_double) if inv_sum else None))
        return reach, total, inv
    finally:
        L.hb_destroy(ctx)
