"""Learned surrogate for anytime centrality rankings.

After 4 HyperBall iterations each node has its estimates N_1..N_4 and its
truncated harmonic accumulator H_4 = sum_{t<=4} (1/t)(N_t - N_{t-1}) (the
reference's inverse-distance accumulator, hyper_ball.rs:824-825, stopped early).
An MLP 6 -> 64 -> 64 -> 1 (GELU; BF16 MFMA matmuls with f32 accumulation; f32
master weights and Adam) predicts the remaining log-gap to the full run,
log(1 + H) - log(1 + H_4), from features that do not grow with the graph:
log(1 + degree), log N_2/N_1, log N_3/N_2, log N_4/N_3, log n/N_4 and
log((1 + H_4)/n).  The ranking score is log(1 + H_4) plus that prediction.
Training, inference and features run on the GPU (include/sparkx_hb.h, hb_mlp_*,
hb_surrogate)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .graph import Graph
from .hyper_ball import HyperBall

N_PARAMS = 4673


def features(g: Graph, snaps: np.ndarray, h4: np.ndarray) -> np.ndarray:
    """Host form of the surrogate inputs (the device computes the same in f32)."""
    deg = g.degrees().astype(np.float32)
    lg = np.log(np.asarray(snaps[:4], np.float32))
    ln_n = np.log(np.float32(g.size()))
    x = np.empty((g.size(), 6), np.float32)
    x[:, 0] = np.log1p(deg)
    x[:, 1] = lg[1] - lg[0]
    x[:, 2] = lg[2] - lg[1]
    x[:, 3] = lg[3] - lg[2]
    x[:, 4] = ln_n - lg[3]
    x[:, 5] = np.log1p(np.asarray(h4, np.float64)).astype(np.float32) - ln_n
    return x


def base_score(h4: np.ndarray) -> np.ndarray:
    """log(1 + H_4) in f32, the part of the score the network does not predict."""
    return np.log1p(np.asarray(h4, np.float64)).astype(np.float32)


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
        (score = log(1 + H_4) + predicted residual, features or None, device ms)."""
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
    """(features, target residual log(1+H) - log(1+H_4), H_4, full-run ms, H) of one
    graph: H from the full HyperBall at p_label, features and H_4 from 4
    iterations at p_feat."""
    with HyperBall(g, p_label, device=device) as full:
        inv = full.harmonic_geodesic_distances()
        ms_full = full.stats.ms_total + full.stats.ms_init
    with HyperBall(g, p_feat, device=device, _run=False) as h:
        _lib.check(h._L.hb_set_snapshots(h._ctx, 4))
        h.run(max_iters=4)
        snaps = np.empty((4, g.size()), np.float32)
        _lib.check(h._L.hb_read_snapshots(h._ctx, _lib.ptr(snaps, C.c_float)))
        h4 = h.harmonic_geodesic_distances()
    y = (np.log1p(np.asarray(inv, np.float64)) - np.log1p(np.asarray(h4, np.float64))).astype(np.float32)
    return features(g, snaps, h4), y, h4, ms_full, inv
