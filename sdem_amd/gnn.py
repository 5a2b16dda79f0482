"""ctypes mirror of include/sdem_gnn.h: the learned surrogate."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from .native import SDEM_OK, SdemError, load_hip_library

_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)
_LP = C.POINTER(C.c_int64)
_UP = C.POINTER(C.c_uint16)


class Config(C.Structure):
    _fields_ = [("vel_scale", C.c_double), ("acc_scale", C.c_double), ("len_scale", C.c_double),
                ("radius", C.c_double), ("cutoff_factor", C.c_double), ("floor_z", C.c_double),
                ("dx_scale", C.c_double),
                ("dv_scale", C.c_double), ("hidden", C.c_int32), ("layers", C.c_int32),
                ("node_in", C.c_int32), ("edge_in", C.c_int32), ("out", C.c_int32), ("nsub", C.c_int32)]


def _declare(lib) -> None:
    if getattr(lib, "_gnn_declared", False):
        return
    V = C.c_void_p
    sig = {
        "config_default": ([C.POINTER(Config)], C.c_int),
        "create": ([C.POINTER(Config), C.c_uint64, C.POINTER(V)], C.c_int),
        "get_config": ([V, C.POINTER(Config)], C.c_int),
        "set_scales": ([V, C.POINTER(Config)], C.c_int),
        "num_params": ([V, _LP], C.c_int),
        "get_params": ([V, _FP], C.c_int),
        "set_params": ([V, _FP], C.c_int),
        "get_grads": ([V, _FP], C.c_int),
        "get_layout": ([V, C.c_int32, _IP, _LP, _IP], C.c_int),
        "set_graph": ([V, C.c_int32, _FP, C.c_int64, _IP, _IP, _FP], C.c_int),
        "graph_from_dem": ([V, V, C.c_int32, _LP], C.c_int),
        "get_node_ids": ([V, C.c_int32, _IP], C.c_int),
        "get_num_nodes": ([V, _IP, _LP], C.c_int),
        "get_graph": ([V, C.c_int64, _IP, _IP, _FP, _FP], C.c_int),
        "forward": ([V, _FP], C.c_int),
        "backward": ([V, _FP], C.c_int),
        "train_step": ([V, _FP, C.c_float, C.c_float, _FP], C.c_int),
        "advance": ([V, V, C.c_int32], C.c_int),
        "mfma_probe": ([_UP, _UP, _FP], C.c_int),
        "gemm": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _FP, _FP, _FP, C.c_int32, C.c_int32,
                  _FP, _FP], C.c_int),
        "last_error": ([V], C.c_char_p),
        "destroy": ([V], None),
    }
    for k, (a, r) in sig.items():
        fn = getattr(lib, "sdem_gnn_" + k)
        fn.argtypes, fn.restype = a, r
    lib._gnn_declared = True


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return None if a is None else a.ctypes.data_as(_FP)


def default_config(lib=None, **kw) -> Config:
    lib = lib or load_hip_library()
    _declare(lib)
    c = Config()
    lib.sdem_gnn_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Model:
    """One surrogate network (sdem_gnn)."""

    def __init__(self, cfg: Optional[Config] = None, seed: int = 1, lib=None):
        self.lib = lib or load_hip_library()
        _declare(self.lib)
        self.cfg = cfg or default_config(self.lib)
        h = C.c_void_p()
        rc = self.lib.sdem_gnn_create(C.byref(self.cfg), C.c_uint64(seed), C.byref(h))
        if rc != SDEM_OK:
            raise SdemError(rc, "sdem_gnn_create", "model creation failed")
        self.h = h

    def _check(self, rc, where):
        if rc != SDEM_OK:
            msg = self.lib.sdem_gnn_last_error(self.h)
            raise SdemError(rc, "sdem_gnn_" + where, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdem_gnn_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def num_params(self) -> int:
        n = C.c_int64(0)
        self._check(self.lib.sdem_gnn_num_params(self.h, C.byref(n)), "num_params")
        return n.value

    def params(self) -> np.ndarray:
        w = np.zeros(self.num_params(), np.float32)
        self._check(self.lib.sdem_gnn_get_params(self.h, _fp(w)), "get_params")
        return w

    def set_params(self, w) -> None:
        w = _f32(w)
        assert w.size == self.num_params()
        self._check(self.lib.sdem_gnn_set_params(self.h, _fp(w)), "set_params")

    def grads(self) -> np.ndarray:
        g = np.zeros(self.num_params(), np.float32)
        self._check(self.lib.sdem_gnn_get_grads(self.h, _fp(g)), "get_grads")
        return g

    def layout(self):
        n = C.c_int32(0)
        offs = np.zeros(6 * 64, np.int64)
        dims = np.zeros(3 * 64, np.int32)
        self._check(self.lib.sdem_gnn_get_layout(self.h, 64, C.byref(n), offs.ctypes.data_as(_LP),
                                                 dims.ctypes.data_as(_IP)), "get_layout")
        return offs[:6 * n.value].reshape(-1, 6), dims[:3 * n.value].reshape(-1, 3)

    def set_scales(self, **kw) -> None:
        for k, v in kw.items():
            setattr(self.cfg, k, v)
        self._check(self.lib.sdem_gnn_set_scales(self.h, C.byref(self.cfg)), "set_scales")

    def set_graph(self, node_feat, src, dst, edge_feat) -> None:
        fn, fe = _f32(node_feat), _f32(edge_feat)
        src, dst = np.ascontiguousarray(src, np.int32), np.ascontiguousarray(dst, np.int32)
        self._check(self.lib.sdem_gnn_set_graph(self.h, fn.shape[0], _fp(fn), len(src), src.ctypes.data_as(_IP),
                                                dst.ctypes.data_as(_IP), _fp(fe)), "set_graph")
        self.n = fn.shape[0]

    def graph_from_dem(self, eng, for_training: bool = True) -> int:
        """Neighbour graph and features of the DEM context's state; nodes in
        storage order (node_ids() gives their particle ids)."""
        ne = C.c_int64(0)
        self._check(self.lib.sdem_gnn_graph_from_dem(self.h, eng.h, int(for_training), C.byref(ne)),
                    "graph_from_dem")
        n = C.c_int32(0)
        self._check(self.lib.sdem_gnn_get_num_nodes(self.h, C.byref(n), None), "get_num_nodes")
        self.n = n.value
        self.ne = ne.value
        return ne.value

    def node_ids(self) -> np.ndarray:
        ids = np.zeros(self.n, np.int32)
        self._check(self.lib.sdem_gnn_get_node_ids(self.h, self.n, ids.ctypes.data_as(_IP)), "get_node_ids")
        return ids

    def get_graph(self):
        ne = self.ne
        src, dst = np.zeros(ne, np.int32), np.zeros(ne, np.int32)
        fn = np.zeros((self.n, self.cfg.node_in), np.float32)
        fe = np.zeros((ne, self.cfg.edge_in), np.float32)
        self._check(self.lib.sdem_gnn_get_graph(self.h, ne, src.ctypes.data_as(_IP), dst.ctypes.data_as(_IP),
                                                _fp(fn), _fp(fe)), "get_graph")
        return fn, src, dst, fe

    def forward(self) -> np.ndarray:
        out = np.zeros((self.n, self.cfg.out), np.float32)
        self._check(self.lib.sdem_gnn_forward(self.h, _fp(out)), "forward")
        return out

    def backward(self, dout) -> None:
        d = _f32(dout)
        self._check(self.lib.sdem_gnn_backward(self.h, _fp(d)), "backward")

    def train_step(self, target, lr=1e-3, clip=1.0) -> float:
        t = _f32(target)
        loss = C.c_float(0)
        self._check(self.lib.sdem_gnn_train_step(self.h, _fp(t), C.c_float(lr), C.c_float(clip), C.byref(loss)),
                    "train_step")
        return loss.value

    def advance(self, eng, calls: int = 1) -> None:
        """`calls` surrogate calls on a DEM context (each nsub steps)."""
        self._check(self.lib.sdem_gnn_advance(self.h, eng.h, int(calls)), "advance")
        n, ne = C.c_int32(0), C.c_int64(0)
        self._check(self.lib.sdem_gnn_get_num_nodes(self.h, C.byref(n), C.byref(ne)), "get_num_nodes")
        self.n, self.ne = n.value, ne.value


def mfma_probe(a_bits: np.ndarray, b_bits: np.ndarray, lib=None) -> np.ndarray:
    lib = lib or load_hip_library()
    _declare(lib)
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    d = np.zeros((64, 16), np.float32)
    rc = lib.sdem_gnn_mfma_probe(a.ctypes.data_as(_UP), b.ctypes.data_as(_UP), _fp(d))
    if rc != SDEM_OK:
        raise SdemError(rc, "sdem_gnn_mfma_probe", "")
    return d


def gemm(a, b, ta=False, tb=False, bias=None, flags=0, splits=1, lib=None):
    """C = op(a) op(b) on the matrix cores (inputs rounded to bf16); a, b as
    stored (a^T when ta, b^T when tb).  Returns (C, pre-activation)."""
    lib = lib or load_hip_library()
    _declare(lib)
    a, b = _f32(a), _f32(b)
    M = a.shape[1] if ta else a.shape[0]
    K = a.shape[0] if ta else a.shape[1]
    N = b.shape[0] if tb else b.shape[1]
    c = np.zeros((M, N), np.float32)
    pre = np.zeros((M, N), np.float32)
    bb = None if bias is None else _f32(bias)
    rc = lib.sdem_gnn_gemm(int(ta), int(tb), M, N, K, _fp(a), _fp(b), _fp(bb), int(flags), int(splits), _fp(c),
                           _fp(pre))
    if rc != SDEM_OK:
        raise SdemError(rc, "sdem_gnn_gemm", "")
    return c, pre
