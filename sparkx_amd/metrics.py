"""Evaluation metrics with the reference's semantics (src/utils.rs:57-100), computed
on the GPU through libsparkx_hb (hb_eval_ranks / hb_eval_metrics):

- mae, mape (terms whose truth is exactly 0 count 0), in the argument order of
  utils::mae / utils::mape;
- ranks by a stable descending sort (ties keep index order, no averaging; -0.0
  equals +0.0), bit-identical to utils::ranks;
- Spearman rho = 1 - 6 ssd / (n max(n^2 - 1, 1)) over those ranks, with ssd summed
  exactly and rounded once;
- the top-k overlap |top_k(pred) & top_k(truth)| / k, k = max(1, ceil(frac n)), used to
  score anytime rankings (no reference counterpart).

A NaN in any input raises HyperBallError (the reference's order for NaN depends on
its sort's internals).  There is no host implementation: without the library these
raise.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, np.float64).ravel())


def evaluate(pred, truth, top_frac: float = 0.01, device: int = 0) -> dict:
    """All four metrics of pred against truth in one pass on the device."""
    a, b = _f64(pred), _f64(truth)
    if a.size != b.size:
        raise ValueError(f"length mismatch: {a.size} != {b.size}")
    out = np.zeros(4, np.float64)
    L = _lib.load()
    _lib.check(L.hb_eval_metrics(device, _lib.ptr(a, C.c_double), _lib.ptr(b, C.c_double), a.size,
                                 float(top_frac), _lib.ptr(out, C.c_double)))
    return {"mae": float(out[0]), "mape": float(out[1]), "spearman": float(out[2]),
            "top_overlap": float(out[3])}


def ranks(v, device: int = 0) -> np.ndarray:
    """utils.rs:76-85: rank of each element in a stable descending order."""
    x = _f64(v)
    r = np.zeros(x.size, np.uint32)
    L = _lib.load()
    _lib.check(L.hb_eval_ranks(device, _lib.ptr(x, C.c_double), x.size, _lib.ptr(r, C.c_uint32)))
    return r


def mae(a, b) -> float:
    """utils.rs:57-59."""
    return evaluate(a, b)["mae"]


def mape(a, b) -> float:
    """utils.rs:60-74 (percent; terms with b == 0 count 0)."""
    return evaluate(a, b)["mape"]


def spearman_rho(a, b) -> float:
    """utils.rs:87-100."""
    return evaluate(a, b)["spearman"]


def top_overlap(pred, truth, frac: float = 0.01) -> float:
    """|top-k(pred) & top-k(truth)| / k with k = max(1, ceil(frac * n)), ties broken by index."""
    return evaluate(pred, truth, frac)["top_overlap"]
