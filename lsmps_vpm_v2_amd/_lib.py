"""ctypes binding of the C ABI in include/vpmfmm.h (libvpmfmm.so, in-tree).

There is no fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpmfmm.so")

# Symbols declared in include/vpmfmm.h (checked by the CPU tests).
EXPORTS = (
    "vpmfmm_default_opts", "vpmfmm_create", "vpmfmm_destroy", "vpmfmm_set_opts",
    "vpmfmm_get_opts", "vpmfmm_set_stream", "vpmfmm_field", "vpmfmm_field_device",
    "vpmfmm_velocity", "vpmfmm_velocity_device", "vpmfmm_velocity_device_noalloc",
    "vpmfmm_advect_device",
    "vpmfmm_step_device", "vpmfmm_direct_device", "vpmfmm_potential",
    "vpmfmm_potential_device", "vpmfmm_type1_default_opts", "vpmfmm_velocity_type1",
    "vpmfmm_velocity_type1_device", "vpmfmm_fmm3d_default_opts", "vpmfmm_velocity3d",
    "vpmfmm_velocity3d_device", "vpmfmm_velocity3d_alpha", "vpmfmm_fmm3d_last_stats",
    "vpmfmm_fmm3d_cells", "vpmfmm_fmm3d_perm",
    "vpmfmm_last_stats", "vpmfmm_last_error", "vpmfmm_build_tree", "vpmfmm_tree_counts",
    "vpmfmm_tree_geometry", "vpmfmm_tree_cells", "vpmfmm_tree_leaves", "vpmfmm_tree_lists",
    "vpmfmm_tree_expansions", "vpmfmm_set_level_orders", "vpmfmm_set_selector",
    "vpmfmm_last_selection", "vpmfmm_level_features", "vpmfmm_mlp_forward",
    "vpmfmm_mlp_loss_grad", "vpmfmm_mlp_train",
)

STATUS = {
    0: "OK", 1: "BAD_ARG", 2: "NONFINITE_INPUT", 3: "DEGENERATE_DOMAIN",
    4: "TREE_DEPTH_EXCEEDED", 5: "OUT_OF_MEMORY", 6: "INTERNAL", 7: "DUPLICATE_POINTS",
    8: "HIP_ERROR", 9: "OUTSIDE_BOXES", 10: "LEVEL1_LEAF",
}

FP64, MIXED = 0, 1
ORDER_FIXED, ORDER_LEARNED = 0, 1
SEL_FEATURES, SEL_HIDDEN, SEL_PARAMS = 9, 64, 4930
# limits of vpmfmm_velocity_device_noalloc (include/vpmfmm.h)
NOALLOC_SCRATCH_BYTES, NOALLOC_MAX_N = 256 << 20, 1 << 20


class Opts(C.Structure):
    _fields_ = [
        ("p", C.c_int32), ("max_par", C.c_int32), ("max_src", C.c_int32),
        ("exp_tree", C.c_double), ("min_level", C.c_int32), ("max_level", C.c_int32),
        ("precision", C.c_int32), ("binom_mode", C.c_int32), ("list4_mode", C.c_int32),
        ("u_inf", C.c_double), ("v_inf", C.c_double), ("device", C.c_int32),
        ("timing", C.c_int32), ("stage_mask", C.c_int32), ("order_select", C.c_int32),
        ("reselect_every", C.c_int32), ("judge_every", C.c_int32),
        ("icutoff", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class Type1Opts(C.Structure):
    """vpmfmm_type1_opts: the reference's type-1 FMM knobs (global.cpp:316-328)."""
    _fields_ = [
        ("n_s", C.c_int32), ("lmax", C.c_int32), ("ndp", C.c_int32), ("icutoff", C.c_int32),
        ("l2l_mode", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class Fmm3dOpts(C.Structure):
    """vpmfmm_fmm3d_opts: the 3-D treecodes' knobs (fast3DFMM.cpp:4, fmm3D.cpp:3,
    velocity_calc.cpp:31, global.cpp:609-611, 633-635, 769, 778-779)."""
    _fields_ = [
        ("crit_num", C.c_int32), ("max_depth", C.c_int32), ("exp_tree", C.c_double),
        ("center", C.c_double * 3), ("support", C.c_double), ("u_inf", C.c_double),
        ("v_inf", C.c_double), ("w_inf", C.c_double), ("type", C.c_int32),
        ("max_par", C.c_int32), ("max_src", C.c_int32), ("reserved", C.c_int32),
    ]


class Fmm3dStats(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("n_active", C.c_int32), ("n_cells", C.c_int32),
        ("n_leaves", C.c_int32), ("depth", C.c_int32), ("sorted_levels", C.c_int32),
        ("root_centre", C.c_double * 3), ("root_side", C.c_double), ("pairs", C.c_uint64),
        ("m2p", C.c_uint64), ("wave_pairs", C.c_uint64), ("wave_m2p", C.c_uint64),
        ("ms_tree", C.c_double), ("ms_moments", C.c_double), ("ms_eval", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["root_centre"] = tuple(self.root_centre)
        return d


class Selector(C.Structure):
    _fields_ = [
        ("n_features", C.c_int32), ("hidden", C.c_int32), ("n_params", C.c_int32),
        ("p_min", C.c_int32), ("p_max", C.c_int32), ("n_thresholds", C.c_int32),
        ("thresholds", C.c_int32 * 8), ("target_err", C.c_float), ("margin", C.c_float),
        ("mu", C.c_float * 9), ("sd", C.c_float * 9), ("lo", C.c_float * 9),
        ("hi", C.c_float * 9),
    ]


class Selection(C.Structure):
    _fields_ = [
        ("source", C.c_int32), ("threshold", C.c_int32), ("p", C.c_int32 * 17),
        ("reselected", C.c_int32), ("pred_err", C.c_float), ("pred_ms", C.c_float),
        ("ms_reselect", C.c_float), ("judge_err", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {"source": self.source, "threshold": self.threshold, "p": list(self.p),
                "reselected": self.reselected, "pred_err": self.pred_err,
                "pred_ms": self.pred_ms, "ms_reselect": self.ms_reselect,
                "judge_err": self.judge_err}


class Stats(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("n_active", C.c_int32), ("n_cells", C.c_int32),
        ("n_leaves", C.c_int32), ("min_leaf_level", C.c_int32), ("max_leaf_level", C.c_int32),
        ("n_list1", C.c_int64), ("n_list2", C.c_int64), ("n_list3", C.c_int64),
        ("n_list4", C.c_int64), ("p2p_pairs", C.c_int64), ("m2p_evals", C.c_int64),
        ("ms_total", C.c_double),
        ("ms_tree", C.c_double), ("ms_lists", C.c_double), ("ms_upward", C.c_double),
        ("ms_m2l", C.c_double), ("ms_downward", C.c_double), ("ms_eval", C.c_double),
        ("m2p_terms", C.c_int64), ("l2p_terms", C.c_int64), ("scratch_bytes", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class VpmFmmError(RuntimeError):
    """A non-OK status from the C ABI (``code``, ``status`` name, message)."""

    def __init__(self, code: int, where: str, msg: str):
        self.code = code
        self.status = STATUS.get(code, str(code))
        super().__init__(f"{where}: {self.status}: {msg}")


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint8)
_lib = None


def load() -> C.CDLL:
    """Load libvpmfmm.so (raises FileNotFoundError if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m lsmps_vpm_v2_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.vpmfmm_default_opts.argtypes = [C.POINTER(Opts)]
    L.vpmfmm_default_opts.restype = None
    L.vpmfmm_create.argtypes = [C.POINTER(Opts), C.POINTER(_vp)]
    L.vpmfmm_destroy.argtypes = [_vp]
    L.vpmfmm_set_opts.argtypes = [_vp, C.POINTER(Opts)]
    L.vpmfmm_get_opts.argtypes = [_vp, C.POINTER(Opts)]
    L.vpmfmm_set_stream.argtypes = [_vp, _vp]
    for name in ("vpmfmm_field", "vpmfmm_field_device"):
        getattr(L, name).argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]
    for name in ("vpmfmm_velocity", "vpmfmm_velocity_device", "vpmfmm_velocity_device_noalloc"):
        getattr(L, name).argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    for name in ("vpmfmm_potential", "vpmfmm_potential_device"):
        getattr(L, name).argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp]
    L.vpmfmm_type1_default_opts.argtypes = [C.POINTER(Type1Opts)]
    L.vpmfmm_type1_default_opts.restype = None
    for name in ("vpmfmm_velocity_type1", "vpmfmm_velocity_type1_device"):
        getattr(L, name).argtypes = [_vp, C.POINTER(Type1Opts), C.c_int32, _vp, _vp, _vp, _vp,
                                     _vp, _vp]
    L.vpmfmm_fmm3d_default_opts.argtypes = [C.POINTER(Fmm3dOpts)]
    L.vpmfmm_fmm3d_default_opts.restype = None
    for name in ("vpmfmm_velocity3d", "vpmfmm_velocity3d_device"):
        getattr(L, name).argtypes = [_vp, C.POINTER(Fmm3dOpts), C.c_int32] + [_vp] * 11
    L.vpmfmm_velocity3d_alpha.argtypes = [_vp, C.POINTER(Fmm3dOpts), C.c_int32] + [_vp] * 10
    L.vpmfmm_fmm3d_last_stats.argtypes = [_vp, C.POINTER(Fmm3dStats)]
    L.vpmfmm_fmm3d_cells.argtypes = [_vp] * 8
    L.vpmfmm_fmm3d_perm.argtypes = [_vp, _vp]
    L.vpmfmm_advect_device.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_double,
                                       C.c_int32]
    L.vpmfmm_step_device.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     C.c_double, C.c_int32]
    L.vpmfmm_direct_device.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32, _vp,
                                       _vp, _vp]
    L.vpmfmm_last_stats.argtypes = [_vp, C.POINTER(Stats)]
    L.vpmfmm_last_error.argtypes = [_vp]
    L.vpmfmm_last_error.restype = C.c_char_p
    L.vpmfmm_build_tree.argtypes = [_vp, C.c_int32, _vp, _vp, _vp]
    L.vpmfmm_tree_counts.argtypes = [_vp, _ip, _ip, _ip, _ip]
    L.vpmfmm_tree_geometry.argtypes = [_vp, _dp, _dp]
    L.vpmfmm_tree_cells.argtypes = [_vp, _lp, _ip, _ip, _ip, _ip, _ip, _up, _up]
    L.vpmfmm_tree_leaves.argtypes = [_vp, _lp, _ip, _ip]
    L.vpmfmm_tree_lists.argtypes = [_vp, C.c_int64, _lp, _ip, _lp, _ip, _lp, _ip, _lp, _ip,
                                    C.c_int32]
    L.vpmfmm_tree_expansions.argtypes = [_vp, _lp, _dp, _dp]
    L.vpmfmm_set_level_orders.argtypes = [_vp, _ip, C.c_int32]
    L.vpmfmm_set_selector.argtypes = [_vp, C.POINTER(Selector), _vp]
    L.vpmfmm_last_selection.argtypes = [_vp, C.POINTER(Selection)]
    L.vpmfmm_level_features.argtypes = [_vp, _vp, _vp, _ip]
    L.vpmfmm_mlp_forward.argtypes = [_vp, _vp, _vp, C.c_int32, _vp]
    L.vpmfmm_mlp_loss_grad.argtypes = [_vp, _vp, _vp, C.c_int32, _vp, C.c_int32, _vp, C.c_float,
                                       _vp, _vp]
    L.vpmfmm_mlp_train.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int32, _vp, C.c_int32, _vp,
                                   C.c_float, C.c_int32, C.c_int32, C.c_float, C.c_float, _vp]
    _lib = L
    return L


def default_opts() -> Opts:
    o = Opts()
    load().vpmfmm_default_opts(C.byref(o))
    return o
