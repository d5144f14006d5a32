"""ctypes front end of the CPU oracle (oracle2d.c) and of the reference build
(_ref/libref2dfmm.so).  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
product in 2dfmm_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle2d.so")
REF_PATH = os.path.join(HERE, "_ref", "libref2dfmm.so")
REF_SRC = "/root/reference/ver-c/src"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build(ref: bool = True) -> None:
    """Compile liboracle2d.so, and the reference into _ref/ when its sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle2d.so"], check=True)
    if ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_build.restype = C.c_void_p
        L.orc_build.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_num_nodes.argtypes = [C.c_void_p]
        L.orc_depth.argtypes = [C.c_void_p]
        L.orc_nodes.argtypes = [C.c_void_p, _up, _ip, _dp, _dp, _dp, _ip, _ip]
        L.orc_node_members.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_coeffs.argtypes = [C.c_void_p, _dp]
        L.orc_eval.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.orc_list.argtypes = [C.c_void_p, C.c_double, C.c_double, _up, _ip, C.c_int]
        L.orc_direct.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int]
        L.orc_read_params.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_global2local.restype = C.c_int64
        L.orc_global2local.argtypes = [C.c_int64, C.c_int]
        L.orc_local2global.restype = C.c_int64
        L.orc_local2global.argtypes = [C.c_int64, C.c_int]
        L.orc_get_level.argtypes = [C.c_int64]
        L.orc_get_parent.restype = C.c_int64
        L.orc_get_parent.argtypes = [C.c_int64]
        L.orc_get_child.argtypes = [C.c_int64, C.POINTER(C.c_int64)]
        L.orc_quadrant_pos.argtypes = [_dp, _dp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def reflib():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_make_tree.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double]
        R.ref_nodes.argtypes = [_up, _ip, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp]
        R.ref_eval.argtypes = [C.c_int, _dp, _dp]
        R.ref_list.argtypes = [C.c_double, C.c_double, _up, _ip, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_theta.restype = C.c_double
        R.ref_getChild.argtypes = [C.c_int, C.POINTER(C.c_int)]
        R.ref_quadrant_pos.argtypes = [C.c_double] * 4
        _ref = R
    return _ref


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Tree:
    """Oracle tree: build, introspect, evaluate (see oracle2d.h)."""

    def __init__(self, xy, cir, cor, nterms, num_p, theta, max_depth=31):
        L = lib()
        self.xy = _f64(xy).reshape(-1, 2)
        self.n = self.xy.shape[0]
        self.nterms, self.num_p, self.theta = int(nterms), int(num_p), float(theta)
        h = L.orc_build(self.n, self.xy.ravel(), _f64(cir), _f64(cor), self.nterms,
                        self.num_p, self.theta, int(max_depth))
        if not h:
            raise RuntimeError("oracle build failed: " + L.orc_last_error().decode())
        self._h = C.c_void_p(h)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.orc_free(self._h)
            self._h = None

    @property
    def num_nodes(self) -> int:
        return lib().orc_num_nodes(self._h)

    @property
    def depth(self) -> int:
        return lib().orc_depth(self._h)

    def nodes(self) -> dict:
        n = self.num_nodes
        d = dict(gid=np.zeros(n, np.uint64), level=np.zeros(n, np.int32), cx=np.zeros(n),
                 cy=np.zeros(n), region=np.zeros(n), count=np.zeros(n, np.int32),
                 is_leaf=np.zeros(n, np.int32))
        lib().orc_nodes(self._h, d["gid"], d["level"], d["cx"], d["cy"], d["region"],
                        d["count"], d["is_leaf"])
        lo = np.zeros(n + 1, np.int32)
        ids = np.zeros(self.n, np.int32)
        lib().orc_node_members(self._h, lo, ids)
        d["lo"], d["members"] = lo, ids
        return d

    def coeffs(self) -> np.ndarray:
        out = np.zeros(self.num_nodes * self.nterms * 2)
        lib().orc_coeffs(self._h, out)
        return out.view(np.complex128).reshape(self.num_nodes, self.nterms)

    def eval(self, txy, nthreads=None) -> np.ndarray:
        txy = _f64(txy).reshape(-1, 2)
        uv = np.zeros_like(txy)
        lib().orc_eval(self._h, txy.shape[0], txy.ravel(), uv.ravel(), int(nthreads or os.cpu_count()))
        return uv

    def list(self, x, y, cap=1 << 16):
        g = np.zeros(cap, np.uint64)
        k = np.zeros(cap, np.int32)
        n = lib().orc_list(self._h, float(x), float(y), g, k, cap)
        if n > cap:
            return self.list(x, y, cap=n)
        return g[:n], k[:n]


def direct(xy, cir, cor, txy, style=1, nthreads=None) -> np.ndarray:
    """FP64 direct sum; style 0 = biot_b (cut-off 1e-6), 1 = direct_evaluate rule (1e-10)."""
    xy = _f64(xy).reshape(-1, 2)
    txy = _f64(txy).reshape(-1, 2)
    uv = np.zeros_like(txy)
    lib().orc_direct(xy.shape[0], xy.ravel(), _f64(cir), _f64(cor), txy.shape[0], txy.ravel(),
                     uv.ravel(), int(style), int(nthreads or os.cpu_count()))
    return uv


def read_params(path):
    a, b = C.c_int(0), C.c_int(0)
    c, d = C.c_double(0), C.c_double(0)
    got = lib().orc_read_params(path.encode(), C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return got, a.value, b.value, c.value, d.value


class RefRun:
    """One run of the reference make_tree (oracle/_ref); the reference keeps a
    single global tree, so only the most recent RefRun is valid."""

    def __init__(self, xy, cir, cor, nterms, num_p, theta):
        R = reflib()
        self.xy = _f64(xy).reshape(-1, 2)
        self.n = self.xy.shape[0]
        self.nterms = int(nterms)
        rc = R.ref_make_tree(self.n, self.xy.ravel(), _f64(cir), _f64(cor), int(nterms), int(num_p),
                             float(theta))
        if rc != 0:
            raise RuntimeError(f"reference make_tree harness failed ({rc})")

    def nodes(self) -> dict:
        R = reflib()
        n = R.ref_num_nodes()
        d = dict(gid=np.zeros(n, np.uint64), level=np.zeros(n, np.int32), cx=np.zeros(n),
                 cy=np.zeros(n), region=np.zeros(n), count=np.zeros(n, np.int32),
                 is_leaf=np.zeros(n, np.int32))
        lo = np.zeros(n + 1, np.int32)
        mem = np.zeros(self.n, np.int32)
        co = np.zeros(n * self.nterms * 2)
        R.ref_nodes(d["gid"], d["level"], d["cx"], d["cy"], d["region"], d["count"], d["is_leaf"],
                    lo, mem, co)
        d["leaf_lo"], d["leaf_members"] = lo, mem
        d["coeffs"] = co.view(np.complex128).reshape(n, self.nterms)
        return d

    def eval(self, txy) -> np.ndarray:
        txy = _f64(txy).reshape(-1, 2)
        uv = np.zeros_like(txy)
        reflib().ref_eval(txy.shape[0], txy.ravel(), uv.ravel())
        return uv

    def list(self, x, y, cap=1 << 16):
        g = np.zeros(cap, np.uint64)
        k = np.zeros(cap, np.int32)
        u, v = C.c_double(0), C.c_double(0)
        n = reflib().ref_list(float(x), float(y), g, k, cap, C.byref(u), C.byref(v))
        return g[:n], k[:n], (u.value, v.value)
