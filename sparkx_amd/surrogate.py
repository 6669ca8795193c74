"""Learned surrogate for anytime centrality rankings.

An MLP 6 -> 64 -> 64 -> 1 (GELU; BF16 MFMA matmuls with f32 accumulation; f32
master weights and Adam) maps each node's degree, log(1 + degree) and log N_1..N_4
-- N_t its HyperBall estimate after t iterations -- to its log harmonic
centrality log(1 + H), H the full run's inverse-distance accumulator
(hyper_ball.rs:824-825).  Inputs and target are standardised per graph (mean and
standard deviation over the nodes that have a neighbour), so that a network trained
on small graphs ranks the nodes of a larger one: the prediction is the node's
standardised log harmonic centrality, which orders the graph's nodes.  Training,
inference and the inputs run on the GPU (include/sparkx_hb.h, hb_mlp_*,
hb_surrogate)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .graph import Graph
from .hyper_ball import HyperBall

N_PARAMS = 4673


def raw_features(g: Graph, snaps: np.ndarray) -> np.ndarray:
    """degree, log(1 + degree), log N_1..N_4 (f32, as the device computes them)."""
    deg = g.degrees().astype(np.float32)
    x = np.empty((g.size(), 6), np.float32)
    x[:, 0] = deg
    x[:, 1] = np.log1p(deg)
    x[:, 2:] = np.log(np.asarray(snaps[:4], np.float32).T)
    return x


def standardize(v: np.ndarray, linked: np.ndarray) -> np.ndarray:
    """(v - mean) / std over the linked rows (f64 statistics, population std;
    a constant column maps to 0), returned as f32."""
    v = np.asarray(v, np.float64)
    mean = v[linked].mean(0)
    sd = v[linked].std(0)
    safe = np.where(sd > 0, sd, 1.0)
    return np.where(sd > 0, (v - mean) / safe, 0.0).astype(np.float32)


def features(g: Graph, snaps: np.ndarray) -> np.ndarray:
    """Host form of the surrogate inputs: raw_features standardised per graph."""
    return standardize(raw_features(g, snaps), g.degrees() > 0)


class Surrogate:
    def __init__(self, device: int = 0, seed: int = 1):
        self._L = _lib.load()
        self._m = C.c_void_p()
        _lib.check(self._L.hb_mlp_create(C.byref(self._m), device, seed))
        self.device = device

    def params(self):
        p = np.empty(N_PARAMS, np.float32)
        norm = np.empty(12, np.float32)
        _lib.check(self._L.hb_mlp_get_params(self._m, _lib.ptr(p, C.c_float), _lib.ptr(norm, C.c_float)))
        return p, norm

    def set_params(self, params=None, norm=None):
        pp = None if params is None else _lib.ptr(np.ascontiguousarray(params, np.float32), C.c_float)
        nn = None if norm is None else _lib.ptr(np.ascontiguousarray(norm, np.float32), C.c_float)
        keep = (params, norm)  # noqa: F841 (buffers alive across the call)
        _lib.check(self._L.hb_mlp_set_params(self._m, pp, nn))

    def gradient(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_float()
        _lib.check(self._L.hb_mlp_gradient(self._m, _lib.ptr(x, C.c_float), _lib.ptr(y, C.c_float), y.size,
                                           _lib.ptr(g, C.c_float), C.byref(loss)))
        return g, loss.value

    def train(self, x, y, epochs=4, batch=8192, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.empty(epochs, np.float32)
        _lib.check(self._L.hb_mlp_train(self._m, _lib.ptr(x, C.c_float), _lib.ptr(y, C.c_float), y.size,
                                        epochs, batch, lr, seed, _lib.ptr(losses, C.c_float)))
        return losses

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(x.shape[0], np.float32)
        _lib.check(self._L.hb_mlp_predict(self._m, _lib.ptr(x, C.c_float), out.size, _lib.ptr(out, C.c_float)))
        return out

    def rank(self, h: HyperBall, p_feat: int = 4, want_features: bool = False):
        """Anytime ranking of the graph staged in `h` (its run is replaced): returns
        (predicted standardised log harmonic centrality, nodes without neighbours at
        -3e38; standardised inputs or None; device ms)."""
        n = h.g.size()
        pred = np.empty(n, np.float32)
        feat = np.empty((n, 6), np.float32) if want_features else None
        ms = C.c_double()
        _lib.check(self._L.hb_surrogate(h._ctx, self._m, p_feat, _lib.ptr(pred, C.c_float),
                                        _lib.ptr(feat, C.c_float) if feat is not None else None, C.byref(ms)))
        return pred, feat, ms.value

    def close(self):
        if getattr(self, "_m", None):
            self._L.hb_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def labelled_graph(g: Graph, p_label: int = 7, p_feat: int = 4, device: int = 0):
    """(inputs, target, H_4, full-run ms, H) of one graph: H from the full HyperBall
    at p_label, the target its standardised log(1 + H); inputs from the first 4
    iterations at p_feat, whose harmonic accumulator H_4 is the truncated baseline.
    Rows of nodes without neighbours are included; training uses the linked ones."""
    with HyperBall(g, p_label, device=device) as full:
        inv = full.harmonic_geodesic_distances()
        ms_full = full.stats.ms_total + full.stats.ms_init
    with HyperBall(g, p_feat, device=device, _run=False) as h:
        _lib.check(h._L.hb_set_snapshots(h._ctx, 4))
        h.run(max_iters=4)
        snaps = np.empty((4, g.size()), np.float32)
        _lib.check(h._L.hb_read_snapshots(h._ctx, _lib.ptr(snaps, C.c_float)))
        h4 = h.harmonic_geodesic_distances()
    linked = g.degrees() > 0
    y = standardize(np.log1p(np.asarray(inv, np.float64))[:, None], linked)[:, 0]
    return features(g, snaps), y, h4, ms_full, inv
