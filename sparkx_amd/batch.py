"""Batched HyperBall: many independent graphs run together.

- HyperBallMany: one engine context (own stream) per (graph, p), all advanced at once
  by hb_run_many from a pool of host threads; every graph's results are exactly its
  own single-graph run's.
- HyperBallBatch: the batched engine (hb_batch_*): every graph's CSR and registers in one
  set of device arrays, advanced by persistent kernels per (precision, degree bin) over
  the running graphs' tiles, one hipGraph replay per iteration; every graph's results are
  exactly its own single-graph run's.
- batched_workload: the seeded 64-graph workload (half Erdos-Renyi, half R-MAT, n
  log-uniform in [2^12, 2^18], average degree uniform in [4, 32], p uniform in {4..8}).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .graph import Graph


def batched_workload(count: int = 64, seed: int = 1, n_range=(12, 18), deg_range=(4.0, 32.0),
                     p_choices=(4, 5, 6, 7, 8), threads: int = 0):
    """[(Graph, p)] for the batched configuration; deterministic in `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        logn = rng.uniform(*n_range)
        deg = rng.uniform(*deg_range)
        p = int(rng.choice(p_choices))
        gseed = int(rng.integers(1, 2**62))
        if i < count // 2:
            n = int(round(2.0 ** logn))
            g = Graph.erdos_renyi(n, int(n * deg / 2), seed=gseed, threads=threads)
        else:
            scale = int(round(logn))
            g = _rmat(scale, deg, gseed, threads)
        out.append((g, p))
    return out


def _rmat(scale, deg, seed, threads):
    h = C.c_void_p()
    n = 1 << scale
    _lib.check(_lib.load().hb_graph_rmat(C.byref(h), scale, int(n * deg / 2), 0.57, 0.19, 0.19, seed, threads))
    return Graph._take(h, f"rmat{scale}_d{deg:.1f}_s{seed}")


class HyperBallMany:
    """Runs HyperBall on every (graph, p) of `items` concurrently; `hbs[i]` is the
    HyperBall of item i (its accessors, epilogues and caches as usual)."""

    def __init__(self, items, device: int = 0, threads: int = 0, layout: str = "auto"):
        from .hyper_ball import HyperBall
        self.items = list(items)
        self.threads = threads
        self.hbs = []
        try:
            for g, p in self.items:
                self.hbs.append(HyperBall(g, p, _run=False, device=device, layout=layout))
        except Exception:
            self.close()
            raise

    def run(self, max_iters: int = 0, dense: bool = False, timing: bool = False):
        n = len(self.hbs)
        ctxs = (C.c_void_p * n)(*[h._ctx.value for h in self.hbs])
        ps = (C.c_int * n)(*[h.p for h in self.hbs])
        stats = (_lib.HbStats * n)()
        flags = (_lib.HB_FLAG_DENSE if dense else 0) | (_lib.HB_FLAG_TIMING if timing else 0)
        L = _lib.load()
        _lib.check(L.hb_run_many(ctxs, ps, n, max_iters, flags, stats, self.threads))
        out = []
        for h, st in zip(self.hbs, stats):
            own = _lib.HbStats()
            C.memmove(C.byref(own), C.byref(st), C.sizeof(own))
            out.append(h._ran(own))
        return out

    def close(self):
        for h in self.hbs:
            h.close()
        self.hbs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class HyperBallBatch:
    """Runs HyperBall on every (graph, p) of `items`; per-graph results as HyperBall."""

    def __init__(self, items, max_iters: int = 0, device: int = 0):
        self.items = list(items)
        self._L = _lib.load()
        descs = (_lib.HbGraphDesc * len(self.items))()
        self._keep = []
        for d, (g, p) in zip(descs, self.items):
            off, nb = g.offsets, g.neighbours
            self._keep += [off, nb]
            d.offsets = off.ctypes.data
            d.off_bytes = off.itemsize
            d.n = g.size()
            d.nbrs = nb.ctypes.data if nb.size else None
            d.nbr_bytes = nb.itemsize
            d.arcs = g.width()
            d.p = p
            d.max_iters = max_iters
        self._b = C.c_void_p()
        _lib.check(self._L.hb_batch_create(C.byref(self._b), device, descs, len(self.items)))
        self.stats = None

    def set_schedule(self, small_max_deg=0, mid_max_deg=0, tile_arcs=0):
        _lib.check(self._L.hb_batch_set_schedule(self._b, small_max_deg, mid_max_deg, tile_arcs))

    def autotune(self):
        s, m, t, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
        _lib.check(self._L.hb_batch_autotune(self._b, C.byref(s), C.byref(m), C.byref(t), C.byref(ms)))
        return (s.value, m.value, t.value), ms.value

    def run(self, dense: bool = False, graph_capture: bool = True):
        flags = (_lib.HB_FLAG_DENSE if dense else 0) | (_lib.HB_FLAG_GRAPH if graph_capture else 0)
        st = _lib.HbBatchStats()
        _lib.check(self._L.hb_batch_run(self._b, flags, C.byref(st)))
        self.stats = st
        return st

    def graph_stats(self, i: int):
        it, conv = C.c_int32(), C.c_int32()
        inc = np.zeros(4096, np.uint64)
        k = _lib.check(self._L.hb_batch_graph_stats(self._b, i, C.byref(it), C.byref(conv),
                                                    _lib.ptr(inc, C.c_uint64), inc.size))
        return it.value, bool(conv.value), inc[:k]

    def registers(self, i: int, which: int = _lib.HB_REGS_REFERENCE) -> np.ndarray:
        g, p = self.items[i]
        out = np.empty((g.size(), 1 << p), np.uint8)
        _lib.check(self._L.hb_batch_read_registers(self._b, i, which, _lib.ptr(out, C.c_uint8)))
        return out

    def estimates(self, i: int, which: int = _lib.HB_REGS_REFERENCE) -> np.ndarray:
        g, _ = self.items[i]
        out = np.empty(g.size(), np.float32)
        _lib.check(self._L.hb_batch_read_estimates(self._b, i, which, _lib.ptr(out, C.c_float)))
        return out

    def accumulators(self, i: int):
        g, _ = self.items[i]
        d = np.empty(g.size(), np.float64)
        v = np.empty(g.size(), np.float64)
        _lib.check(self._L.hb_batch_read_accumulators(self._b, i, _lib.ptr(d, C.c_double), _lib.ptr(v, C.c_double)))
        return d, v

    def close(self):
        if getattr(self, "_b", None):
            self._L.hb_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
