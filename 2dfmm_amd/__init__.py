"""MI355X (gfx950) 2-D vortex-blob treecode: Python front end of libfmm2d.so.

The library is the product; this module only marshals numpy arrays (or device
pointers) into its C ABI (include/fmm2d.h).  The two reference entry points
keep their names and meaning:

    make_tree(vor_b, cir_b, cor_b)     ver-c/src/quadtree.c:102-171
    biot_tree(x, y) -> (u, v)          ver-c/src/fmm.c:69-118

Where the reference prints and carries on (it has no error path), this module
raises FmmError with the library's message.  There is no CPU fallback: if the
library is not built, importing this module fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmm2d.so")

MODE_TREECODE = 0
MODE_FMM64 = 1
MODE_MIXED = 2


class FmmError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                      f"python 2dfmm_amd/_build.py")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

lib = C.CDLL(LIB_PATH)
lib.make_tree.restype = None
lib.make_tree.argtypes = [C.c_int, _dp, _dp, _dp]
lib.biot_tree.restype = None
lib.biot_tree.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
lib.fmm_set_params.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
lib.fmm_use_param_file.argtypes = [C.c_char_p]
lib.fmm_get_params.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]
lib.fmm_set_mode.argtypes = [C.c_int]
lib.fmm_set_fmm_params.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
lib.fmm_selftest_mfma.argtypes = [C.c_int, C.POINTER(C.c_double)]
lib.fmm_set_prefetch.argtypes = [C.c_int]
lib.fmm_last_error.restype = C.c_char_p
lib.fmm_make_tree_device.argtypes = [C.c_int, _vp, _vp, _vp]
lib.fmm_eval_sources.argtypes = [_dp]
lib.fmm_eval_sources_device.argtypes = [_vp]
lib.fmm_eval_batch.argtypes = [C.c_int, _dp, _dp]
lib.fmm_eval_batch_device.argtypes = [C.c_int, _vp, _vp]
lib.fmm_tree_size.argtypes = [C.POINTER(C.c_int)] * 3
lib.fmm_dump_tree.argtypes = [_up, _ip, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _dp]
lib.fmm_dump_lists.argtypes = [C.c_int, _dp, C.c_int, _up, _ip, _ip]
lib.fmm_direct.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp, C.c_int]
lib.fmm_direct_device.argtypes = [C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int]
lib.fmm_direct_device_noalloc.argtypes = [C.c_int, _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int, _vp]
lib.fmm_make_forest.argtypes = [C.c_int, _ip, _dp, _dp, _dp]
lib.fmm_make_forest_device.argtypes = [C.c_int, _ip, _vp, _vp, _vp]
lib.fmm_set_persistent.argtypes = [C.c_int, C.c_int]
lib.fmm_set_split.argtypes = [C.c_int]
lib.fmm_set_far_burst.argtypes = [C.c_int]
lib.fmm_set_treecode_walk.argtypes = [C.c_int]
lib.fmm_set_l2l.argtypes = [C.c_int]
lib.fmm_set_start_lists.argtypes = [C.c_int]
lib.fmm_get_split.argtypes = []
lib.fmm_set_smooth.argtypes = [C.c_double, C.c_int]
lib.fmm_get_smooth.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int)]
lib.fmm_count_work.argtypes = [C.POINTER(C.c_longlong)] * 3
lib.fmm_count_work_detail.argtypes = [np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                      C.c_int]
lib.fmm_count_work_leaves.argtypes = [np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                      C.c_int]
lib.fmm_device_cus.argtypes = []
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
lib.fmm_set_learned.argtypes = [C.c_void_p, C.c_float, C.c_float]
lib.fmm_set_learned_floor.argtypes = [C.c_float]
lib.fmm_num_target_leaves.argtypes = []
lib.fmm_leaf_features.argtypes = [_fp]
lib.fmm_blob_leaf.argtypes = [_ip]
lib.fmm_leaf_params.argtypes = [_fp]
lib.fmm_set_leaf_params.argtypes = [C.c_void_p, C.c_int]
lib.fmm_mlp_forward.argtypes = [C.c_int, _fp, _fp, _fp]
lib.fmm_mlp_grad.argtypes = [C.c_int, _fp, _fp, _fp, _fp, C.POINTER(C.c_float)]
lib.fmm_mlp_train.argtypes = [C.c_int, _fp, _fp, _fp, C.c_int, C.c_int, C.c_float, C.c_uint, _fp]
MLP_WEIGHTS = 4866
lib.fmm_make_tree_mirror_device.argtypes = [C.c_int, _vp, _vp, _vp, C.c_double]
lib.fmm_drift_ab3_device.argtypes = [C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_double]
lib.fmm_history_shift_device.argtypes = [C.c_int, _vp, C.c_int]
lib.fmm_store_velocity_device.argtypes = [C.c_int, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double]
lib.fmm_core_spread_device.argtypes = [C.c_int, _vp, C.c_double, C.c_double]
lib.fmm_set_timing.argtypes = [C.c_int]
lib.fmm_get_timing.argtypes = [_dp, C.c_int]
lib.fmm_count_interactions.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
lib.fmm_synchronize.argtypes = []
lib.fmm_release.argtypes = []
lib.fmm_release.restype = None


_state = {"n": -1}


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise FmmError(f"{what}: {lib.fmm_last_error().decode(errors='replace')}")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def last_error() -> str:
    return lib.fmm_last_error().decode(errors="replace")


def set_params(nterms: int, num_p: int, theta: float, r_min: float = 0.01) -> None:
    _check(lib.fmm_set_params(int(nterms), int(num_p), float(theta), float(r_min)), "set_params")


def use_param_file(path: str | None = None) -> None:
    _check(lib.fmm_use_param_file(path.encode() if path else None), "use_param_file")


def get_params():
    a, b = C.c_int(), C.c_int()
    c, d = C.c_double(), C.c_double()
    _check(lib.fmm_get_params(C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "get_params")
    return a.value, b.value, c.value, d.value


def set_mode(mode: int) -> None:
    _check(lib.fmm_set_mode(int(mode)), "set_mode")


def set_fmm_params(p: int = 16, leaf: int = 64, theta: float = 0.5, cut_sigmas: float = 4.0) -> None:
    _check(lib.fmm_set_fmm_params(int(p), int(leaf), float(theta), float(cut_sigmas)),
           "set_fmm_params")


def selftest_mfma(fp32: bool) -> float:
    e = C.c_double()
    _check(lib.fmm_selftest_mfma(1 if fp32 else 0, C.byref(e)), "selftest_mfma")
    return e.value


def set_prefetch(on: bool) -> None:
    _check(lib.fmm_set_prefetch(1 if on else 0), "set_prefetch")


def make_tree(vor_b, cir_b, cor_b) -> None:
    """vor_b: (N, 2) positions (memory x0,y0,x1,y1,...); cir_b, cor_b: (N,)."""
    xy = _f64(vor_b).reshape(-1, 2)
    n = xy.shape[0]
    cir, cor = _f64(cir_b).ravel(), _f64(cor_b).ravel()
    if cir.shape[0] != n or cor.shape[0] != n:
        raise ValueError("vor_b, cir_b, cor_b disagree on N")
    if n == 0:  # the C ABI takes no array for N = 0; pass placeholders
        xy, cir, cor = np.zeros((1, 2)), np.zeros(1), np.zeros(1)
    lib.make_tree(n, xy.ravel(), cir, cor)
    _state["n"] = -1
    _check(lib.fmm_tree_size(None, None, None), "make_tree")
    _state["n"] = n


def biot_tree(x: float, y: float):
    u, v = C.c_double(), C.c_double()
    lib.biot_tree(float(x), float(y), C.byref(u), C.byref(v))
    return u.value, v.value


def make_tree_device(n: int, xy_ptr: int, cir_ptr: int, cor_ptr: int) -> None:
    """Device-resident inputs (e.g. torch tensors' data_ptr())."""
    _state["n"] = -1
    _check(lib.fmm_make_tree_device(int(n), xy_ptr, cir_ptr, cor_ptr), "make_tree_device")
    _state["n"] = int(n)


def concat_systems(systems):
    """[(xy, cir, cor), ...] -> (offsets int32 (S+1,), xy, cir, cor) concatenated."""
    sizes = [np.asarray(s[1]).shape[0] for s in systems]
    off = np.zeros(len(systems) + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    xy = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.float64).reshape(-1, 2)
                                              for s in systems]))
    cir = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.float64) for s in systems]))
    cor = np.ascontiguousarray(np.concatenate([np.asarray(s[2], np.float64) for s in systems]))
    return off, xy, cir, cor


def make_forest(systems) -> np.ndarray:
    """One tree per system, built in one pass; returns the offsets used."""
    off, xy, cir, cor = concat_systems(systems)
    n = int(off[-1])
    _state["n"] = -1
    _check(lib.fmm_make_forest(len(systems), off, xy.ravel() if n else np.zeros(2),
                               cir if n else np.zeros(1), cor if n else np.zeros(1)), "make_forest")
    _state["n"] = n
    return off


def make_forest_device(offsets, xy_ptr: int, cir_ptr: int, cor_ptr: int) -> None:
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    _state["n"] = -1
    _check(lib.fmm_make_forest_device(len(off) - 1, off, xy_ptr, cir_ptr, cor_ptr),
           "make_forest_device")
    _state["n"] = int(off[-1])


def make_tree_mirror_device(n: int, xy_ptr: int, cir_ptr: int, cor_ptr: int, d_mirror: float):
    """Tree over n blobs plus their mirror images about y = d_mirror."""
    _state["n"] = -1
    _check(lib.fmm_make_tree_mirror_device(int(n), xy_ptr, cir_ptr, cor_ptr, float(d_mirror)),
           "make_tree_mirror_device")
    _state["n"] = int(n)


def drift_ab3_device(n, xy_ptr, hist_ptr, dt, mirror=False, d_mirror=0.0) -> None:
    _check(lib.fmm_drift_ab3_device(int(n), xy_ptr, hist_ptr, float(dt), 1 if mirror else 0,
                                    float(d_mirror)), "drift_ab3_device")


def history_shift_device(n, hist_ptr, first_step=False) -> None:
    _check(lib.fmm_history_shift_device(int(n), hist_ptr, 1 if first_step else 0),
           "history_shift_device")


def store_velocity_device(n, xy_ptr, uv_ptr, hist_ptr, uinf=0.0, vinf=0.0, omz=0.0) -> None:
    _check(lib.fmm_store_velocity_device(int(n), xy_ptr, uv_ptr, hist_ptr, float(uinf), float(vinf),
                                         float(omz)), "store_velocity_device")


def core_spread_device(n, cor_ptr, dt, re) -> None:
    _check(lib.fmm_core_spread_device(int(n), cor_ptr, float(dt), float(re)), "core_spread_device")


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def set_learned_floor(cut_min: float) -> None:
    """Floor of the learned near-field radius in core radii (default 2.5)."""
    _check(lib.fmm_set_learned_floor(float(cut_min)), "set_learned_floor")


def set_learned(weights, p_margin: float = 0.0, cut_margin: float = 0.0) -> None:
    """Per-leaf p and near radius from the MLP (weights: 4866 floats), or None to stop."""
    if weights is None:
        _check(lib.fmm_set_learned(None, 0.0, 0.0), "set_learned")
        return
    w = _f32(weights).ravel()
    if w.shape[0] != MLP_WEIGHTS:
        raise ValueError(f"need {MLP_WEIGHTS} weights")
    _check(lib.fmm_set_learned(w.ctypes.data, float(p_margin), float(cut_margin)), "set_learned")


def num_target_leaves() -> int:
    k = lib.fmm_num_target_leaves()
    if k < 0:
        raise FmmError("num_target_leaves: " + last_error())
    return k


def leaf_features() -> np.ndarray:
    out = np.zeros((max(num_target_leaves(), 1), 8), np.float32)
    _check(lib.fmm_leaf_features(out.ravel()), "leaf_features")
    return out[:num_target_leaves()]


def blob_leaf() -> np.ndarray:
    out = np.zeros(max(_state["n"], 1), np.int32)
    _check(lib.fmm_blob_leaf(out), "blob_leaf")
    return out[:_state["n"]]


def leaf_params() -> np.ndarray:
    out = np.zeros((max(num_target_leaves(), 1), 2), np.float32)
    _check(lib.fmm_leaf_params(out.ravel()), "leaf_params")
    return out[:num_target_leaves()]


def set_leaf_params(par) -> None:
    if par is None:
        _check(lib.fmm_set_leaf_params(None, 0), "set_leaf_params")
        return
    p = _f32(par).reshape(-1, 2)
    _check(lib.fmm_set_leaf_params(p.ctypes.data, p.shape[0]), "set_leaf_params")


def mlp_forward(x, weights) -> np.ndarray:
    x = _f32(x).reshape(-1, 8)
    out = np.zeros((max(x.shape[0], 1), 2), np.float32)
    if x.shape[0]:
        _check(lib.fmm_mlp_forward(x.shape[0], x.ravel(), _f32(weights).ravel(), out.ravel()),
               "mlp_forward")
    return out[:x.shape[0]]


def mlp_grad(x, y, weights):
    x = _f32(x).reshape(-1, 8)
    y = _f32(y).reshape(-1, 2)
    g = np.zeros(MLP_WEIGHTS, np.float32)
    loss = C.c_float()
    _check(lib.fmm_mlp_grad(x.shape[0], x.ravel(), y.ravel(), _f32(weights).ravel(), g,
                            C.byref(loss)), "mlp_grad")
    return g, loss.value


def mlp_train(x, y, weights, epochs=100, batch=1024, lr=1e-3, seed=1):
    x = _f32(x).reshape(-1, 8)
    y = _f32(y).reshape(-1, 2)
    w = _f32(weights).ravel().copy()
    hist = np.zeros(int(epochs), np.float32)
    _check(lib.fmm_mlp_train(x.shape[0], x.ravel(), y.ravel(), w, int(epochs), int(batch),
                             float(lr), int(seed), hist), "mlp_train")
    return w, hist


def set_persistent(blocks: int, graph: bool = False) -> None:
    _check(lib.fmm_set_persistent(int(blocks), 1 if graph else 0), "set_persistent")


def set_l2l(min_children) -> None:
    """Local expansions inherited from the parent (accurate modes): parents
    with at least min_children target-leaf children (True = 2, False = off)."""
    if min_children is True:
        min_children = 2
    _check(lib.fmm_set_l2l(int(min_children)), "set_l2l")


def set_start_lists(cap=True) -> None:
    """Children's walks from the parent's start lists of at most `cap` entries
    (True = 128, the default and most) or from the root (False / 0, and
    parents whose list would be longer); the same interactions either way."""
    if cap is True:
        cap = 128
    _check(lib.fmm_set_start_lists(int(cap)), "set_start_lists")


def set_treecode_walk(per_leaf: bool) -> None:
    """Treecode at the blobs: a wave per 64 consecutive blobs (True) or a lane per target."""
    _check(lib.fmm_set_treecode_walk(1 if per_leaf else 0), "set_treecode_walk")


def set_far_burst(cells: int) -> None:
    _check(lib.fmm_set_far_burst(int(cells)), "set_far_burst")


def set_split(budget: int) -> None:
    """Near work units per chunk of a split target leaf (0: no splitting)."""
    _check(lib.fmm_set_split(int(budget)), "set_split")


def get_split() -> int:
    return int(lib.fmm_get_split())


def set_smooth(eta: float, min_blobs: int = 48) -> None:
    """Smooth cells of the mixed mode (one core radius): near nodes of half side
    <= eta * sigma with >= min_blobs blobs are summed through a Taylor expansion
    of the regularised kernel (eta 0: off).  Takes effect at the next build."""
    _check(lib.fmm_set_smooth(float(eta), int(min_blobs)), "set_smooth")


def get_smooth():
    e, m = C.c_double(), C.c_int()
    _check(lib.fmm_get_smooth(C.byref(e), C.byref(m)), "get_smooth")
    return e.value, m.value


def device_cus() -> int:
    return int(lib.fmm_device_cus())


def eval_sources() -> np.ndarray:
    """Velocities at all blobs of the last tree, caller's order, shape (N, 2)."""
    n = _state["n"]
    if n < 0:
        raise FmmError("no valid tree: call make_tree first")
    uv = np.zeros((max(n, 1), 2))
    _check(lib.fmm_eval_sources(uv.ravel()), "eval_sources")
    return uv[:n]


def eval_sources_device(uv_ptr: int) -> None:
    _check(lib.fmm_eval_sources_device(uv_ptr), "eval_sources_device")


def eval_batch(txy) -> np.ndarray:
    txy = _f64(txy).reshape(-1, 2)
    nt = txy.shape[0]
    uv = np.zeros((max(nt, 1), 2))
    if nt:
        _check(lib.fmm_eval_batch(nt, txy.ravel(), uv.ravel()), "eval_batch")
    return uv[:nt]


def eval_batch_device(nt: int, txy_ptr: int, uv_ptr: int) -> None:
    _check(lib.fmm_eval_batch_device(int(nt), txy_ptr, uv_ptr), "eval_batch_device")


def tree_size() -> dict:
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    _check(lib.fmm_tree_size(C.byref(a), C.byref(b), C.byref(c)), "tree_size")
    return {"nnodes": a.value, "depth": b.value, "nleaves": c.value}


def dump_tree(n_blobs: int, nterms: int) -> dict:
    """Nodes in ascending reference global id, leaf members, coefficients."""
    m = tree_size()["nnodes"]
    d = dict(gid=np.zeros(m, np.uint64), level=np.zeros(m, np.int32), cx=np.zeros(m), cy=np.zeros(m),
             region=np.zeros(m), count=np.zeros(m, np.int32), is_leaf=np.zeros(m, np.int32))
    lo = np.zeros(m + 1, np.int32)
    mem = np.zeros(max(n_blobs, 1), np.int32)
    co = np.zeros(max(m * nterms * 2, 1))
    if m:
        _check(lib.fmm_dump_tree(d["gid"], d["level"], d["cx"], d["cy"], d["region"], d["count"],
                                 d["is_leaf"], lo, mem, co), "dump_tree")
    d["leaf_lo"], d["leaf_members"] = lo, mem[:n_blobs]
    d["coeffs"] = co[:m * nterms * 2].view(np.complex128).reshape(m, nterms)
    return d


def dump_lists(txy, cap: int = 4096):
    """Per target: (gids, kinds) in visit order (0 = P2P leaf, 1 = M2S cell)."""
    txy = _f64(txy).reshape(-1, 2)
    nt = txy.shape[0]
    gid = np.zeros(nt * cap, np.uint64)
    kind = np.zeros(nt * cap, np.int32)
    ln = np.zeros(max(nt, 1), np.int32)
    _check(lib.fmm_dump_lists(nt, txy.ravel(), cap, gid, kind, ln), "dump_lists")
    if nt and ln[:nt].max(initial=0) > cap:
        return dump_lists(txy, cap=int(ln.max()))
    out = []
    for t in range(nt):
        L = ln[t]
        out.append((gid[t * cap:t * cap + L].copy(), kind[t * cap:t * cap + L].copy()))
    return out


def direct(xy, cir, cor, txy, style: int = 1) -> np.ndarray:
    """FP64 direct sum on the GPU (style 0: biot_b rule; 1: direct_evaluate rule)."""
    xy = _f64(xy).reshape(-1, 2)
    txy = _f64(txy).reshape(-1, 2)
    ns, nt = xy.shape[0], txy.shape[0]
    uv = np.zeros((max(nt, 1), 2))
    if nt:
        _check(lib.fmm_direct(ns, xy.ravel() if ns else np.zeros(2), _f64(cir).ravel() if ns else np.zeros(1),
                              _f64(cor).ravel() if ns else np.zeros(1), nt, txy.ravel(), uv.ravel(),
                              int(style)), "direct")
    return uv[:nt]


def direct_device(ns, xy_ptr, cir_ptr, cor_ptr, nt, txy_ptr, uv_ptr, style: int = 1) -> None:
    _check(lib.fmm_direct_device(int(ns), xy_ptr, cir_ptr, cor_ptr, int(nt), txy_ptr, uv_ptr,
                                 int(style)), "direct_device")


NOALLOC_MAX_TARGETS = 1 << 20


def direct_device_noalloc(ns, xy_ptr, cir_ptr, cor_ptr, nt, txy_ptr, uv_ptr, style: int = 1,
                          stream: int = 0) -> None:
    """direct_device with its partial sums in the library's fixed device array
    (no allocation; nt <= 2^20); runs on `stream` (0: the null stream) and
    returns without waiting."""
    _check(lib.fmm_direct_device_noalloc(int(ns), xy_ptr, cir_ptr, cor_ptr, int(nt), txy_ptr,
                                         uv_ptr, int(style), stream or None),
           "direct_device_noalloc")


PHASES = ("bbox_keys", "radix_sort", "tree_levels", "upward", "evaluate",
          "levels", "leaf_order_gather", "s2m_m2m", "fused_pass", "split_passes")


def set_timing(on: bool) -> None:
    _check(lib.fmm_set_timing(1 if on else 0), "set_timing")


def get_timing() -> dict:
    """Milliseconds of the phases of the last build and evaluation."""
    ms = np.zeros(len(PHASES))
    _check(lib.fmm_get_timing(ms, len(PHASES)), "get_timing")
    return dict(zip(PHASES, ms.tolist()))


def count_interactions():
    """(far cells evaluated by M2S, near pairs) over all blobs as targets."""
    a, b = C.c_longlong(), C.c_longlong()
    _check(lib.fmm_count_interactions(C.byref(a), C.byref(b)), "count_interactions")
    return a.value, b.value


WORK_KEYS = ("far", "pairs", "smooth", "far_terms2", "l2p_terms", "target_leaves",
             "parent_far", "parent_far_terms2", "parents", "l2l_shifts")


def count_work_detail() -> dict:
    """The evaluation's work in detail (see fmm_count_work_detail)."""
    out = np.zeros(len(WORK_KEYS), np.int64)
    _check(lib.fmm_count_work_detail(out, len(WORK_KEYS)), "count_work_detail")
    return dict(zip(WORK_KEYS, (int(v) for v in out)))


def count_work_leaves() -> np.ndarray:
    """Per target leaf: (far cells, near pairs, smooth evaluations, terms, targets)."""
    nl = num_target_leaves()
    out = np.zeros((max(nl, 1), 5), np.int64)
    _check(lib.fmm_count_work_leaves(out.ravel(), nl), "count_work_leaves")
    return out[:nl]


def count_work():
    """(far cells, pairs summed directly, (target, smooth cell) evaluations)."""
    a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
    _check(lib.fmm_count_work(C.byref(a), C.byref(b), C.byref(c)), "count_work")
    return a.value, b.value, c.value


def synchronize() -> None:
    _check(lib.fmm_synchronize(), "synchronize")


def release() -> None:
    lib.fmm_release()
