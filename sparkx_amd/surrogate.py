"""Learned surrogate for anytime centrality rankings.

An MLP 6 -> 64 -> 64 -> 1 (GELU; BF16 MFMA matmuls with f32 accumulation; f32
master weights and Adam) maps six inputs per node to its log harmonic centrality
log(1 + H), H the full run's inverse-distance accumulator (hyper_ball.rs:824-825).
Two input sets:

* snapshots: degree, log(1 + degree) and log N_1..N_4, N_t the node's HyperBall
  estimate after t iterations (hb_surrogate);
* walks: logs of S_1..S_passes (S_k the number of walks of k + 1 arcs leaving the
  node), of its degree and of its largest neighbour degree, with differences of
  them -- two or three gather passes, no HyperBall iteration (hb_surrogate_walks;
  the sets are listed in include/sparkx_hb.h).

Inputs and target are standardised per graph (mean and standard deviation over the
nodes that have a neighbour), so that a network trained on small graphs ranks the
nodes of a larger one: the prediction is the node's standardised log harmonic
centrality, which orders the graph's nodes.  Training, inference and the inputs run
on the GPU (include/sparkx_hb.h, hb_mlp_*, hb_surrogate*, hb_walk_features)."""
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
        """Anytime ranking of the graph staged in `h` (its run is replaced by the 4
        iterations at p_feat, and `h` then describes that run): returns (predicted
        standardised log harmonic centrality, nodes without neighbours at -3e38;
        standardised inputs or None; device ms)."""
        n = h.g.size()
        pred = np.empty(n, np.float32)
        feat = np.empty((n, 6), np.float32) if want_features else None
        ms = C.c_double()
        try:
            _lib.check(self._L.hb_surrogate(h._ctx, self._m, p_feat, _lib.ptr(pred, C.c_float),
                                            _lib.ptr(feat, C.c_float) if feat is not None else None,
                                            C.byref(ms)))
        finally:
            h._replaced()  # the context's previous run is gone (also when the call failed midway)
        return pred, feat, ms.value

    def rank_walks(self, h: HyperBall, passes: int = 3, want_features: bool = False):
        """Anytime ranking of the graph staged in `h` from walk-count inputs of
        `passes` gather passes (the network must be trained on the same set; its
        HyperBall run, if any, is kept): (prediction as in rank(); standardised inputs
        or None; device ms)."""
        n = h.g.size()
        pred = np.empty(n, np.float32)
        feat = np.empty((n, 6), np.float32) if want_features else None
        ms = C.c_double()
        _lib.check(self._L.hb_surrogate_walks(h._ctx, self._m, passes, _lib.ptr(pred, C.c_float),
                                              _lib.ptr(feat, C.c_float) if feat is not None else None,
                                              C.byref(ms)))
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


def walk_features(h: HyperBall, passes: int = 3):
    """Unstandardised walk-count inputs of the graph staged in `h` (f32 [n, 6], the
    caller's node order) and the device ms of the passes."""
    x = np.empty((h.g.size(), 6), np.float32)
    ms = C.c_double()
    _lib.check(h._L.hb_walk_features(h._ctx, passes, _lib.ptr(x, C.c_float), C.byref(ms)))
    return x, ms.value


def labelled_walks(g: Graph, passes: int = 3, p_label: int = 7, device: int = 0):
    """(standardised walk-count inputs, target, full-run ms, H) of one graph, the
    target as in labelled_graph."""
    with HyperBall(g, p_label, device=device) as full:
        inv = full.harmonic_geodesic_distances()
        ms_full = full.stats.ms_total + full.stats.ms_init
        x, _ = walk_features(full, passes)
    linked = g.degrees() > 0
    y = standardize(np.log1p(np.asarray(inv, np.float64))[:, None], linked)[:, 0]
    return standardize(x, linked), y, ms_full, inv


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
