"""Evaluation metrics with the reference's semantics (src/utils.rs:57-100):
MAE, MAPE (skipping exact zeros), ranks by a stable descending sort (ties keep
index order, no averaging) and Spearman rho = 1 - 6 sum d^2 / (n (n^2 - 1));
plus the top-k overlap used to score anytime rankings."""
from __future__ import annotations

import numpy as np


def mae(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sum(np.abs(a - b)) / a.size)


def mape(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nz = b != 0
    return float(100.0 * np.sum(np.abs(a[nz] - b[nz]) / np.abs(b[nz])) / a.size)


def ranks(v) -> np.ndarray:
    """utils.rs:76-85: rank of each element in a stable descending order."""
    v = np.asarray(v, np.float64)
    order = np.argsort(-v, kind="stable")
    r = np.empty(v.size, np.int64)
    r[order] = np.arange(v.size)
    return r


def spearman_rho(a, b) -> float:
    """utils.rs:87-100."""
    ra, rb = ranks(a), ranks(b)
    n = float(ra.size)
    d = (ra - rb).astype(np.float64)
    return 1.0 - (6.0 * float(np.dot(d, d))) / (n * max(n * n - 1.0, 1.0))


def top_overlap(pred, truth, frac: float = 0.01) -> float:
    """|top-k(pred) & top-k(truth)| / k with k = ceil(frac * n), ties broken by index."""
    n = len(truth)
    k = max(1, int(np.ceil(frac * n)))
    tp = np.argsort(-np.asarray(pred, np.float64), kind="stable")[:k]
    tt = np.argsort(-np.asarray(truth, np.float64), kind="stable")[:k]
    return float(np.intersect1d(tp, tt).size) / k
