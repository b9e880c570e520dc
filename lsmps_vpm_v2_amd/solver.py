"""Host-side interface of the MI355X FMM velocity solver.

``FmmSolver`` wraps one C-ABI handle (include/vpmfmm.h).  The classes at the
bottom mirror the reference's interface for this path so code written against
it reads the same:

* ``Particle``      -- the fields of ``struct Particle`` the path touches
                       (reference Utils.hpp:77-159);
* ``VelocityCalc``  -- ``get_velocity(particle, base_grid, step)``
                       (velocity_calc.hpp:42, velocity_calc.cpp:149-251), its
                       FMM (type 2) and type-1 FMM branches;
* ``velocity_biot_savart`` -- ``biotsavart_fmm_2d`` (biotsavart_fmm_2d.cpp:4);
* ``TreeCell``      -- ``initializeTree`` / ``updateTree`` / ``intList`` /
                       ``extList`` and the tree summary (treeCell.hpp:83-195);
* ``fmm2D``         -- ``calcField`` and ``get_Field_x/y`` (fmm2D.hpp:15-68).

Host arrays are numpy; device entry points take torch tensors on the GPU
(torch only provides device memory and streams here).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FP64, MIXED, ORDER_FIXED, ORDER_LEARNED, Opts, Stats, VpmFmmError

_PRECISION = {"fp64": FP64, "mixed": MIXED, FP64: FP64, MIXED: MIXED}
_ORDER = {"fixed": ORDER_FIXED, "learned": ORDER_LEARNED, ORDER_FIXED: ORDER_FIXED,
          ORDER_LEARNED: ORDER_LEARNED}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u8(a):
    return np.ascontiguousarray(np.asarray(a) != 0, dtype=np.uint8)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _dptr(t):
    return C.c_void_p(int(t.data_ptr()))


def _torch_first():
    """torch carries its own HIP runtime, which finds no GPU once the
    library's runtime has initialised the device in the same process; when
    torch is installed and sees a GPU, it initialises first."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.cuda.init()


class FmmSolver:
    """One solver handle: persistent device buffers, one HIP stream."""

    def __init__(self, p=10, max_par=75, max_src=50, exp_tree=1.0, min_level=1, max_level=0,
                 precision="fp64", binom_mode=1, list4_mode=1, u_inf=1.0, v_inf=0.0,
                 device=0, timing=False, stage_mask=0, order="fixed", reselect_every=0,
                 selector=None, icutoff=0):
        """order = "learned" (mixed path only) installs ``selector`` (a
        ``selector.Model`` or a model file path; None: the packaged model) and
        lets it pick p per level and the leaf threshold, with ``p`` as the
        fixed-order fallback.  icutoff 1 / 2 / 3 regularizes the near field
        as the reference's Pars::icutoff (regul_func_2d; velocity calls only)."""
        _torch_first()
        self._L = _lib.load()
        o = _lib.default_opts()
        o.p, o.max_par, o.max_src, o.exp_tree = p, max_par, max_src, exp_tree
        o.min_level, o.max_level = min_level, max_level
        o.precision = _PRECISION[precision]
        o.binom_mode, o.list4_mode = binom_mode, list4_mode
        o.u_inf, o.v_inf, o.device, o.timing = u_inf, v_inf, device, 1 if timing else 0
        o.stage_mask = stage_mask
        o.order_select = _ORDER[order]
        o.reselect_every = reselect_every
        o.icutoff = icutoff
        self._o = o
        h = C.c_void_p()
        st = self._L.vpmfmm_create(C.byref(o), C.byref(h))
        if st != 0:
            raise VpmFmmError(st, "vpmfmm_create", "invalid options or device")
        self._h = h
        self._stream = None  # torch stream the handle runs on (set at the first device call)
        self._side = None    # stream used while torch runs on its legacy default stream
        if o.order_select == ORDER_LEARNED:
            from . import selector as _sel
            self.set_selector(_sel.load_model() if selector is None else selector)

    # ------------------------------------------------------------------ utils
    def _check(self, st, where):
        if st != 0:
            msg = self._L.vpmfmm_last_error(self._h).decode(errors="replace")
            raise VpmFmmError(st, where, msg)

    def close(self):
        if getattr(self, "_h", None):
            self._L.vpmfmm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def opts(self) -> Opts:
        return self._o

    def set_opts(self, **kw):
        for k, v in kw.items():
            if k == "precision":
                v = _PRECISION[v]
            if k == "timing":
                v = 1 if v else 0
            setattr(self._o, k, v)
        self._check(self._L.vpmfmm_set_opts(self._h, C.byref(self._o)), "vpmfmm_set_opts")

    def set_level_orders(self, orders):
        """Per-level expansion orders of the mixed path (None: opts.p everywhere)."""
        if orders is None:
            st = self._L.vpmfmm_set_level_orders(self._h, None, 0)
        else:
            a = np.ascontiguousarray(orders, dtype=np.int32)
            st = self._L.vpmfmm_set_level_orders(self._h, a.ctypes.data_as(_lib._ip), len(a))
        self._check(st, "vpmfmm_set_level_orders")

    def set_selector(self, model):
        """Install a learned selection model (``selector.Model`` or file path;
        None removes it)."""
        if model is None:
            self._check(self._L.vpmfmm_set_selector(self._h, None, None), "vpmfmm_set_selector")
            return
        from . import selector as _sel
        if isinstance(model, str):
            model = _sel.Model.load(model)
        sel = model.to_struct()
        params = np.ascontiguousarray(model.params, dtype=np.float32)
        self._check(self._L.vpmfmm_set_selector(self._h, C.byref(sel), _ptr(params)),
                    "vpmfmm_set_selector")

    def last_selection(self) -> dict:
        """Orders, leaf threshold and predictions of the last mixed-path call."""
        s = _lib.Selection()
        self._check(self._L.vpmfmm_last_selection(self._h, C.byref(s)), "vpmfmm_last_selection")
        return s.as_dict()

    def level_features(self):
        """(feats[n_levels, 9] float32 with order column 0, raw[n_levels, 5])
        of the last tree."""
        f = np.zeros((17, _lib.SEL_FEATURES), np.float32)
        r = np.zeros((17, 5), np.float64)
        n = C.c_int32()
        self._check(self._L.vpmfmm_level_features(self._h, _ptr(f), _ptr(r), C.byref(n)),
                    "vpmfmm_level_features")
        return f[: n.value].copy(), r[: n.value].copy()

    # MLP kernels (device tensors, float32 / int32, contiguous)
    def mlp_forward(self, params, x):
        import torch
        out = torch.empty((x.shape[0], 2), dtype=torch.float32, device=x.device)
        self._enter(x)
        self._check(self._L.vpmfmm_mlp_forward(self._h, _dptr(params), _dptr(x), x.shape[0],
                                               _dptr(out)), "vpmfmm_mlp_forward")
        return out

    def mlp_loss_grad(self, params, x, seg, y, w_time):
        import torch
        grad = torch.empty_like(params)
        loss = torch.empty(1, dtype=torch.float32, device=params.device)
        self._enter(params)
        self._check(self._L.vpmfmm_mlp_loss_grad(self._h, _dptr(params), _dptr(x), x.shape[0],
                                                 _dptr(seg), seg.shape[0] - 1, _dptr(y),
                                                 float(w_time), _dptr(grad), _dptr(loss)),
                    "vpmfmm_mlp_loss_grad")
        return float(loss.item()), grad

    def mlp_train(self, params, m, v, x, seg, y, w_time, steps, step0=0, lr=3e-3, lr_min=1e-4):
        """Adam steps on the device, in place; returns the loss of every step."""
        import torch
        hist = torch.empty(steps, dtype=torch.float32, device=params.device)
        self._enter(params)
        self._check(self._L.vpmfmm_mlp_train(self._h, _dptr(params), _dptr(m), _dptr(v),
                                             _dptr(x), x.shape[0], _dptr(seg), seg.shape[0] - 1,
                                             _dptr(y), float(w_time), int(steps), int(step0),
                                             float(lr), float(lr_min), _dptr(hist)),
                    "vpmfmm_mlp_train")
        return hist.cpu().numpy()

    def set_stream(self, stream):
        """Run on a torch.cuda.Stream (None: the handle's own stream)."""
        ptr = None if stream is None else C.c_void_p(int(stream.cuda_stream))
        self._check(self._L.vpmfmm_set_stream(self._h, ptr), "vpmfmm_set_stream")
        self._stream = stream

    def _enter(self, t):
        """Order the call after the work torch has queued for the device
        tensors (the calls themselves return with their results complete).
        On a stream the caller created, the handle runs on that stream
        itself; on torch's legacy default stream it runs on a side stream
        that first waits for it."""
        import torch
        cur = torch.cuda.current_stream(t.device)
        if cur.cuda_stream:
            if self._stream is None or self._stream.cuda_stream != cur.cuda_stream:
                self.set_stream(cur)
            return
        if self._side is None:
            self._side = torch.cuda.Stream(device=t.device)
        if self._stream is None or self._stream.cuda_stream != self._side.cuda_stream:
            self.set_stream(self._side)
        self._side.wait_stream(cur)

    def stats(self) -> dict:
        s = Stats()
        self._check(self._L.vpmfmm_last_stats(self._h, C.byref(s)), "vpmfmm_last_stats")
        return s.as_dict()

    # ------------------------------------------------------------- host paths
    def field(self, x, y, q, active):
        """(ex, ey) of fmm2D::calcField for host arrays."""
        x, y, q, a = _f64(x), _f64(y), _f64(q), _u8(active)
        n = len(x)
        ex, ey = np.empty(n), np.empty(n)
        self._check(self._L.vpmfmm_field(self._h, n, _ptr(x), _ptr(y), _ptr(q), _ptr(a),
                                         _ptr(ex), _ptr(ey)), "vpmfmm_field")
        return ex, ey

    def potential(self, x, y, q, active):
        """phi of fmm2D::calcPotential for host arrays (FP64, singular kernel)."""
        x, y, q, a = _f64(x), _f64(y), _f64(q), _u8(active)
        n = len(x)
        phi = np.empty(n)
        self._check(self._L.vpmfmm_potential(self._h, n, _ptr(x), _ptr(y), _ptr(q), _ptr(a),
                                             _ptr(phi)), "vpmfmm_potential")
        return phi

    def potential_device(self, x, y, q, active, phi=None):
        """phi = sum_{j != i} q_j log r_ij on device tensors (float64 / uint8)."""
        n = x.numel()
        if phi is None:
            phi = x.new_empty(n)
        self._enter(x)
        self._check(self._L.vpmfmm_potential_device(self._h, n, _dptr(x), _dptr(y), _dptr(q),
                                                    _dptr(active), _dptr(phi)),
                    "vpmfmm_potential_device")
        return phi

    def velocity(self, x, y, s, omega, active):
        """(u, v) of VelocityCalc::get_velocity (FMM branch) for host arrays."""
        x, y, s, w, a = _f64(x), _f64(y), _f64(s), _f64(omega), _u8(active)
        n = len(x)
        u, v = np.empty(n), np.empty(n)
        self._check(self._L.vpmfmm_velocity(self._h, n, _ptr(x), _ptr(y), _ptr(s), _ptr(w),
                                            _ptr(a), _ptr(u), _ptr(v)), "vpmfmm_velocity")
        return u, v

    @staticmethod
    def type1_opts(n_s=50, lmax=7, ndp=10, icutoff=2, l2l_mode=1) -> "_lib.Type1Opts":
        """vpmfmm_type1_opts; defaults are the reference's global.cpp:316-328
        (l2l_mode 1: the L2L shift as the reference runs it)."""
        o = _lib.Type1Opts()
        o.n_s, o.lmax, o.ndp, o.icutoff, o.l2l_mode = n_s, lmax, ndp, icutoff, l2l_mode
        return o

    def velocity_type1(self, x, y, s, gz, **type1):
        """(u, v) of the type-1 FMM (VelocityCalc::velocity_old ->
        biotsavart_fmm_2d, plus the handle's free stream) for host arrays;
        keyword arguments are the fields of type1_opts."""
        o = self.type1_opts(**type1)
        x, y, g = _f64(x), _f64(y), _f64(gz)
        s = _f64(s) if s is not None else None
        n = len(x)
        u, v = np.empty(n), np.empty(n)
        self._check(self._L.vpmfmm_velocity_type1(self._h, C.byref(o), n, _ptr(x), _ptr(y),
                                                  _ptr(s) if s is not None else None, _ptr(g),
                                                  _ptr(u), _ptr(v)), "vpmfmm_velocity_type1")
        return u, v

    def velocity_type1_device(self, x, y, s, gz, u=None, v=None, **type1):
        """velocity_type1 on device tensors (s may be None for icutoff 0)."""
        o = self.type1_opts(**type1)
        n = x.numel()
        if u is None:
            u = x.new_empty(n)
        if v is None:
            v = x.new_empty(n)
        self._enter(x)
        self._check(self._L.vpmfmm_velocity_type1_device(
            self._h, C.byref(o), n, _dptr(x), _dptr(y), _dptr(s) if s is not None else None,
            _dptr(gz), _dptr(u), _dptr(v)), "vpmfmm_velocity_type1_device")
        return u, v

    # ------------------------------------------------------------- 3-D path
    @staticmethod
    def fmm3d_opts(crit_num=30, max_depth=20, exp_tree=1.0, center=(0.0, 0.0, 0.0), support=2.0,
                   u_inf=1.0, v_inf=0.0, w_inf=0.0, type=1, max_par=50,
                   max_src=30) -> "_lib.Fmm3dOpts":
        """vpmfmm_fmm3d_opts; defaults are the reference's DIM == 3 parameters
        (critNum = src_count_max 30, expTree 1.0, centre 0, free stream
        (1, 0, 0), par_count_max 50: global.cpp:609-611, 633-635, 769,
        778-779), SUPPORT_RADIUS_FACTOR 2.0 (fast3DFMM.cpp:4) and
        VELOCITY_3D_TYPE 1 (velocity_calc.cpp:31; type 2 is fmm3D::calcVelocity
        over the TreeCell, with the max_par / max_src leaf rule)."""
        o = _lib.Fmm3dOpts()
        o.crit_num, o.max_depth, o.exp_tree, o.support = crit_num, max_depth, exp_tree, support
        o.type, o.max_par, o.max_src = type, max_par, max_src
        for d in range(3):
            o.center[d] = center[d]
        o.u_inf, o.v_inf, o.w_inf = u_inf, v_inf, w_inf
        return o

    def velocity3d(self, x, y, z, s, wx, wy, wz, active, **fmm3d):
        """(u, v, w) of the 3-D treecode (velocity_fmm_3d_fast plus the free
        stream of the keyword options) for host arrays."""
        o = self.fmm3d_opts(**fmm3d)
        arrs = [_f64(t) for t in (x, y, z, s, wx, wy, wz)]
        a = _u8(active)
        n = len(arrs[0])
        u, v, w = np.empty(n), np.empty(n), np.empty(n)
        self._check(self._L.vpmfmm_velocity3d(self._h, C.byref(o), n, *(_ptr(t) for t in arrs),
                                              _ptr(a), _ptr(u), _ptr(v), _ptr(w)),
                    "vpmfmm_velocity3d")
        return u, v, w

    def velocity3d_alpha(self, x, y, z, ax, ay, az, active, **fmm3d):
        """Rotational (u, v, w) from the vortex strengths themselves
        (fmm3D::calcVelocityFast) for host arrays."""
        o = self.fmm3d_opts(**fmm3d)
        arrs = [_f64(t) for t in (x, y, z, ax, ay, az)]
        a = _u8(active)
        n = len(arrs[0])
        u, v, w = np.empty(n), np.empty(n), np.empty(n)
        self._check(self._L.vpmfmm_velocity3d_alpha(self._h, C.byref(o), n,
                                                    *(_ptr(t) for t in arrs), _ptr(a), _ptr(u),
                                                    _ptr(v), _ptr(w)), "vpmfmm_velocity3d_alpha")
        return u, v, w

    def velocity3d_device(self, x, y, z, s, wx, wy, wz, active, u=None, v=None, w=None,
                          **fmm3d):
        """velocity3d on device tensors (float64 / uint8, contiguous)."""
        o = self.fmm3d_opts(**fmm3d)
        n = x.numel()
        u = x.new_empty(n) if u is None else u
        v = x.new_empty(n) if v is None else v
        w = x.new_empty(n) if w is None else w
        self._enter(x)
        self._check(self._L.vpmfmm_velocity3d_device(
            self._h, C.byref(o), n, *(_dptr(t) for t in (x, y, z, s, wx, wy, wz, active)),
            _dptr(u), _dptr(v), _dptr(w)), "vpmfmm_velocity3d_device")
        return u, v, w

    def fmm3d_stats(self) -> dict:
        st = _lib.Fmm3dStats()
        self._check(self._L.vpmfmm_fmm3d_last_stats(self._h, C.byref(st)),
                    "vpmfmm_fmm3d_last_stats")
        return st.as_dict()

    def fmm3d_cells(self) -> dict:
        """Cells of the last 3-D solve in depth-first order."""
        nc = self.fmm3d_stats()["n_cells"]
        out = {"level": np.zeros(nc, np.int32), "key": np.zeros(nc, np.int64),
               "count": np.zeros(nc, np.int32), "leaf": np.zeros(nc, np.uint8),
               "active": np.zeros(nc, np.uint8), "centre": np.zeros((nc, 3)),
               "size": np.zeros(nc)}
        self._check(self._L.vpmfmm_fmm3d_cells(
            self._h, *(_ptr(out[k]) for k in ("level", "key", "count", "leaf", "active",
                                              "centre", "size"))), "vpmfmm_fmm3d_cells")
        return out

    def fmm3d_perm(self) -> np.ndarray:
        n = self.fmm3d_stats()["n"]
        perm = np.zeros(n, np.int32)
        self._check(self._L.vpmfmm_fmm3d_perm(self._h, _ptr(perm)), "vpmfmm_fmm3d_perm")
        return perm

    # ----------------------------------------------------------- device paths
    def field_device(self, x, y, q, active, ex=None, ey=None):
        """Device tensors in, device tensors out (float64 / uint8, contiguous)."""
        n = x.numel()
        if ex is None:
            ex = x.new_empty(n)
        if ey is None:
            ey = x.new_empty(n)
        self._enter(x)
        self._check(self._L.vpmfmm_field_device(self._h, n, _dptr(x), _dptr(y), _dptr(q),
                                                _dptr(active), _dptr(ex), _dptr(ey)),
                    "vpmfmm_field_device")
        return ex, ey

    def velocity_device(self, x, y, s, omega, active, u=None, v=None):
        n = x.numel()
        if u is None:
            u = x.new_empty(n)
        if v is None:
            v = x.new_empty(n)
        self._enter(x)
        self._check(self._L.vpmfmm_velocity_device(self._h, n, _dptr(x), _dptr(y), _dptr(s),
                                                   _dptr(omega), _dptr(active), _dptr(u),
                                                   _dptr(v)), "vpmfmm_velocity_device")
        return u, v

    def velocity_device_noalloc(self, x, y, s, omega, active, u, v):
        """velocity_device with every scratch buffer in the library's static
        device array (vpmfmm_velocity_device_noalloc): no device allocation,
        n <= NOALLOC_MAX_N; the caller passes the outputs."""
        n = x.numel()
        self._enter(x)
        self._check(self._L.vpmfmm_velocity_device_noalloc(
            self._h, n, _dptr(x), _dptr(y), _dptr(s), _dptr(omega), _dptr(active), _dptr(u),
            _dptr(v)), "vpmfmm_velocity_device_noalloc")
        return u, v

    _SCHEMES = {"euler": 0, "rk2": 1}

    def advect_device(self, x, y, u, v, dt, scheme="euler"):
        """Advection sub-step in place on device tensors x, y
        (main_advection.cpp:6-83; "euler" is the scheme the reference compiles)."""
        n = x.numel()
        self._enter(x)
        self._check(self._L.vpmfmm_advect_device(self._h, n, _dptr(x), _dptr(y), _dptr(u),
                                                 _dptr(v), float(dt), self._SCHEMES[scheme]),
                    "vpmfmm_advect_device")

    def step_device(self, x, y, s, omega, active, u, v, dt, scheme="euler"):
        """One loop step on device tensors: advect x, y with (u, v), then the
        velocity at the moved particles into (u, v) (main.cpp:373, 410)."""
        n = x.numel()
        self._enter(x)
        self._check(self._L.vpmfmm_step_device(self._h, n, _dptr(x), _dptr(y), _dptr(s),
                                               _dptr(omega), _dptr(active), _dptr(u), _dptr(v),
                                               float(dt), self._SCHEMES[scheme]),
                    "vpmfmm_step_device")

    def direct_device(self, x, y, q, src_mask=None, targets=None):
        """FP64 direct sum on the GPU at `targets` (int32 device tensor) or all."""
        n = x.numel()
        nt = n if targets is None else targets.numel()
        ex, ey = x.new_empty(nt), x.new_empty(nt)
        self._enter(x)
        self._check(self._L.vpmfmm_direct_device(
            self._h, n, _dptr(x), _dptr(y), _dptr(q),
            None if src_mask is None else _dptr(src_mask), nt,
            None if targets is None else _dptr(targets), _dptr(ex), _dptr(ey)),
            "vpmfmm_direct_device")
        return ex, ey

    # --------------------------------------------------------- tree inspection
    def build_tree(self, x, y, active):
        x, y, a = _f64(x), _f64(y), _u8(active)
        self._check(self._L.vpmfmm_build_tree(self._h, len(x), _ptr(x), _ptr(y), _ptr(a)),
                    "vpmfmm_build_tree")

    def tree_counts(self):
        v = [C.c_int32() for _ in range(4)]
        self._check(self._L.vpmfmm_tree_counts(self._h, *[C.byref(t) for t in v]),
                    "vpmfmm_tree_counts")
        return tuple(t.value for t in v)

    def tree_geometry(self):
        r = C.c_double()
        piv = np.zeros(2)
        self._check(self._L.vpmfmm_tree_geometry(self._h, C.byref(r),
                                                 piv.ctypes.data_as(C.POINTER(C.c_double))),
                    "vpmfmm_tree_geometry")
        return r.value, piv

    def tree_cells(self) -> dict:
        nc = self.tree_counts()[0]
        out = dict(id=np.zeros(nc, np.int64), level=np.zeros(nc, np.int32),
                   ix=np.zeros(nc, np.int32), iy=np.zeros(nc, np.int32),
                   par_num=np.zeros(nc, np.int32), src_num=np.zeros(nc, np.int32),
                   is_leaf=np.zeros(nc, np.uint8), is_active=np.zeros(nc, np.uint8))
        self._check(self._L.vpmfmm_tree_cells(
            self._h, *[out[k].ctypes.data_as(t) for k, t in (
                ("id", _lib._lp), ("level", _lib._ip), ("ix", _lib._ip), ("iy", _lib._ip),
                ("par_num", _lib._ip), ("src_num", _lib._ip), ("is_leaf", _lib._up),
                ("is_active", _lib._up))]), "vpmfmm_tree_cells")
        return out

    def tree_leaves(self, n):
        nl = self.tree_counts()[1]
        lid = np.zeros(nl, np.int64)
        poff = np.zeros(nl + 1, np.int32)
        pidx = np.zeros(n, np.int32)
        self._check(self._L.vpmfmm_tree_leaves(self._h, lid.ctypes.data_as(_lib._lp),
                                               poff.ctypes.data_as(_lib._ip),
                                               pidx.ctypes.data_as(_lib._ip)),
                    "vpmfmm_tree_leaves")
        return lid, poff, pidx

    def tree_expansions(self):
        """(ids, ak, bk) of the last FP64 call; ak/bk (n_cells, p+1) complex."""
        nc = self.tree_counts()[0]
        P1 = self._o.p + 1
        ids = np.zeros(nc, np.int64)
        ak = np.zeros((nc, P1), np.complex128)
        bk = np.zeros((nc, P1), np.complex128)
        self._check(self._L.vpmfmm_tree_expansions(
            self._h, ids.ctypes.data_as(_lib._lp), ak.ctypes.data_as(_lib._dp),
            bk.ctypes.data_as(_lib._dp)), "vpmfmm_tree_expansions")
        return ids, ak, bk

    def tree_lists(self, cell_id, cap=4096):
        bufs = [np.zeros(cap, np.int64) for _ in range(4)]
        cnt = [C.c_int32() for _ in range(4)]
        self._check(self._L.vpmfmm_tree_lists(
            self._h, int(cell_id),
            bufs[0].ctypes.data_as(_lib._lp), C.byref(cnt[0]),
            bufs[1].ctypes.data_as(_lib._lp), C.byref(cnt[1]),
            bufs[2].ctypes.data_as(_lib._lp), C.byref(cnt[2]),
            bufs[3].ctypes.data_as(_lib._lp), C.byref(cnt[3]), cap), "vpmfmm_tree_lists")
        if max(c.value for c in cnt) > cap:
            return self.tree_lists(cell_id, cap=max(c.value for c in cnt))
        return tuple(b[: c.value].copy() for b, c in zip(bufs, cnt))


# ---------------------------------------------------------------------------
# mirrors of the reference interface
# ---------------------------------------------------------------------------
@dataclass
class Particle:
    """Fields of the reference ``Particle`` used by the velocity path
    (``gz``, the vortex strength vorticity s^2, feeds the type-1 branch)."""
    x: np.ndarray
    y: np.ndarray
    s: np.ndarray
    vorticity: np.ndarray
    isActive: np.ndarray
    u: np.ndarray = field(default_factory=lambda: np.zeros(0))
    v: np.ndarray = field(default_factory=lambda: np.zeros(0))
    gz: np.ndarray | None = None
    # 3-D fields (DIM 3): z, the vorticity vector, the z velocity
    z: np.ndarray | None = None
    vortx: np.ndarray | None = None
    vorty: np.ndarray | None = None
    vortz: np.ndarray | None = None
    w: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def num(self) -> int:
        return len(self.x)


class VelocityCalc:
    """``VelocityCalc`` with the FMM branches on the GPU: VELOCITY_2D_TYPE 2
    (``velocity_fmm_2d``, the default) and 1 (``velocity_old``, the type-1
    FMM, with the type-1 knobs n_s, lmax, ndp, icutoff, l2l_mode); with
    ``dim=3``, VELOCITY_3D_TYPE 1 (``velocity_fmm_3d_fast``, the knobs of
    FmmSolver.fmm3d_opts; u_inf and v_inf default to the solver's)."""

    def __init__(self, velocity_type: int | None = None, type1: dict | None = None,
                 dim: int = 2, fmm3d: dict | None = None, **solver_opts):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if velocity_type is None:  # VELOCITY_2D_TYPE 2, VELOCITY_3D_TYPE 1
            velocity_type = 2 if dim == 2 else 1
        if dim == 2 and velocity_type not in (1, 2):
            raise ValueError("velocity_type must be 1 (type-1 FMM) or 2 (FMM)")
        if dim == 3 and velocity_type not in (1, 2):
            raise ValueError("the 3-D velocity type must be 1 (velocity_fmm_3d_fast) or 2 "
                             "(velocity_fmm_3d)")
        self.dim = dim
        self.velocity_type = velocity_type
        self.type1 = dict(type1 or {})
        self._solver = FmmSolver(**solver_opts)
        o = self._solver.opts
        self.fmm3d = {"u_inf": o.u_inf, "v_inf": o.v_inf, **(fmm3d or {})}
        if dim == 3:
            self.fmm3d.setdefault("type", velocity_type)

    @property
    def solver(self) -> FmmSolver:
        return self._solver

    def get_velocity(self, particle: Particle, base_grid=None, step: int = 0) -> None:
        """velocity_calc.cpp:149-251: p.u, p.v = rotational part + free stream.

        ``base_grid`` is accepted and unused, as on the reference's FMM
        branches; ``step`` is accepted for signature parity (the tree is
        rebuilt every call, as TreeCell::updateTree does,
        treeCell.cpp:1245-1249).  Type 1 reads ``gz`` (vorticity s^2 when it
        is None, redistribute_particles.cpp:186) and ignores ``isActive``
        (velocity_calc.cpp:404-424: every particle is a source)."""
        if self.dim == 3:
            particle.u, particle.v, particle.w = self._solver.velocity3d(
                particle.x, particle.y, particle.z, particle.s, particle.vortx, particle.vorty,
                particle.vortz, particle.isActive, **self.fmm3d)
            return
        if self.velocity_type == 1:
            gz = particle.gz if particle.gz is not None else \
                np.asarray(particle.vorticity) * np.asarray(particle.s) ** 2
            particle.u, particle.v = self._solver.velocity_type1(
                particle.x, particle.y, particle.s, gz, **self.type1)
            return
        particle.u, particle.v = self._solver.velocity(
            particle.x, particle.y, particle.s, particle.vorticity, particle.isActive)


class velocity_biot_savart:  # noqa: N801 -- reference class name
    """``velocity_biot_savart`` (velocity_biot_savart.hpp:12-63): the type-1
    FMM ``biotsavart_fmm_2d(pi, pj, icutoff, n_s, n_inter, ndp)``
    (biotsavart_fmm_2d.cpp:4).  Like the reference, targets and sources are
    one particle set (its box lists index pj with pi's particles; velocity_old
    passes two copies of the same particles), n_inter 1 is the only mode, and
    the result is the rotational part alone, written to pi.u, pi.v.  The
    near-field kernel is ``icutoff`` (the reference reads Pars::icutoff, the
    same value in velocity_old); lmax and the L2L mode are constructor knobs
    (Pars::lmax = 7; l2l_mode 1 is the reference's shift)."""

    def __init__(self, lmax: int = 7, l2l_mode: int = 1, solver: FmmSolver | None = None):
        self.lmax, self.l2l_mode = lmax, l2l_mode
        self._solver = solver or FmmSolver(u_inf=0.0, v_inf=0.0)

    def biotsavart_fmm_2d(self, pi: Particle, pj: Particle, icutoff: int = 2, n_s: int = 50,
                          n_inter: int = 1, ndp: int = 10) -> None:
        if n_inter != 1:
            raise ValueError("n_inter must be 1 (targets and sources are one particle set)")
        if pj is not pi and not (np.array_equal(pi.x, pj.x) and np.array_equal(pi.y, pj.y)):
            raise ValueError("targets and sources must be the same particles")
        gz = pj.gz if pj.gz is not None else np.asarray(pj.vorticity) * np.asarray(pj.s) ** 2
        o = self._solver.opts
        free = (o.u_inf, o.v_inf)
        if free != (0.0, 0.0):
            self._solver.set_opts(u_inf=0.0, v_inf=0.0)
        try:
            pi.u, pi.v = self._solver.velocity_type1(pj.x, pj.y, pj.s, gz, n_s=n_s,
                                                     lmax=self.lmax, ndp=ndp, icutoff=icutoff,
                                                     l2l_mode=self.l2l_mode)
        finally:
            if free != (0.0, 0.0):
                self._solver.set_opts(u_inf=free[0], v_inf=free[1])


class TreeCell:
    """Tree summary and lists of the last build (treeCell.hpp:83-195)."""

    def __init__(self, solver: FmmSolver | None = None, **solver_opts):
        self._solver = solver or FmmSolver(**solver_opts)
        self.min_level = 2
        self.max_level = 2
        self.cellCount = 0
        self.leafList = np.zeros(0, np.int64)
        self.rootLength = 0.0
        self.pivotCoord = np.zeros(2)

    # the tree's own knobs (TreeCell reads Pars::par_count_max, src_count_max
    # and expTree; treeCell.hpp:174-185), applied by fmm2D on this tree
    TREE_KNOBS = ("max_par", "max_src", "exp_tree", "min_level", "max_level")

    def initializeTree(self, parPos, activeFlag):
        pos = np.asarray(parPos, dtype=np.float64)
        self._n = len(pos)
        self._fingerprint = _fingerprint(pos[:, 0], pos[:, 1], activeFlag)
        self._solver.build_tree(pos[:, 0], pos[:, 1], activeFlag)
        nc, nl, lo, hi = self._solver.tree_counts()
        self.cellCount, self.min_level, self.max_level = nc, lo, hi
        self.rootLength, self.pivotCoord = self._solver.tree_geometry()
        self.leafList, self._poff, self._pidx = self._solver.tree_leaves(self._n)

    def updateTree(self, parPos, activeFlag):
        self.initializeTree(parPos, activeFlag)

    def cells(self) -> dict:
        return self._solver.tree_cells()

    def parIDList(self, leaf_pos: int) -> np.ndarray:
        return self._pidx[self._poff[leaf_pos]:self._poff[leaf_pos + 1]]

    def intList(self, cell_id):
        l1, _, l3, _ = self._solver.tree_lists(cell_id)
        return l1, l3

    def extList(self, cell_id):
        _, l2, _, l4 = self._solver.tree_lists(cell_id)
        return l2, l4


def _fingerprint(x, y, active):
    h = hashlib.blake2b(digest_size=16)
    for a in (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64),
              np.ascontiguousarray(active, np.uint8)):
        h.update(a.tobytes())
    return h.hexdigest()


class fmm2D:  # noqa: N801 -- reference class name
    """``fmm2D::calcField`` / ``calcPotential`` and their getters
    (fmm2D.hpp:59-67).  The expansion order and binomials are this object's
    (Pars::P_max); the tree is the given ``cellTree``: the library builds the
    tree inside each call, deterministically, so the call runs with the
    tree's leaf rule on the positions the tree was built from (others raise
    ValueError, where the reference would mix the two)."""

    def __init__(self, solver: FmmSolver | None = None, **solver_opts):
        self._solver = solver or FmmSolver(**solver_opts)
        self._ex = np.zeros(0)
        self._ey = np.zeros(0)
        self._phi = np.zeros(0)

    @contextlib.contextmanager
    def _on_tree(self, cellTree, x, y, activeFlag):
        if cellTree is None:
            yield
            return
        if getattr(cellTree, "_fingerprint", None) != _fingerprint(x, y, activeFlag):
            raise ValueError("fmm2D: the tree was built from other positions or active flags "
                             "(call cellTree.initializeTree on these particles first)")
        keep = {k: getattr(self._solver.opts, k) for k in TreeCell.TREE_KNOBS}
        self._solver.set_opts(**{k: getattr(cellTree._solver.opts, k)
                                 for k in TreeCell.TREE_KNOBS})
        try:
            yield
        finally:
            self._solver.set_opts(**keep)

    def calcField(self, cellTree: TreeCell | None, parPos, activeFlag, srcVal):
        pos = np.asarray(parPos, dtype=np.float64)
        with self._on_tree(cellTree, pos[:, 0], pos[:, 1], activeFlag):
            self._ex, self._ey = self._solver.field(pos[:, 0], pos[:, 1], srcVal, activeFlag)

    def calcPotential(self, cellTree: TreeCell | None, parPos, activeFlag, srcVal):
        pos = np.asarray(parPos, dtype=np.float64)
        with self._on_tree(cellTree, pos[:, 0], pos[:, 1], activeFlag):
            self._phi = self._solver.potential(pos[:, 0], pos[:, 1], srcVal, activeFlag)

    def get_Potential(self):
        return self._phi.copy()

    def get_Field_x(self):
        return self._ex.copy()

    def get_Field_y(self):
        return self._ey.copy()


class fmm3D:  # noqa: N801 -- reference class name
    """``fmm3D``'s fast entry (fast3DFMM.hpp, fast3DFMM.cpp:32-92):
    ``calcVelocityFast(parPos, activeFlag, targetFlag, alphaX, alphaY,
    alphaZ)`` then ``get_Field_x/y/z()``.  As in the reference, every particle
    is a target (``targetFlag`` is accepted and unused, fast3DFMM.cpp:341),
    inactive particles are not sources, the result is the rotational part
    alone, and the octree and opening knobs are the reference's
    (``critNum`` = src_count_max, expTree, the domain centre,
    SUPPORT_RADIUS_FACTOR), overridable as keyword arguments."""

    def __init__(self, solver: FmmSolver | None = None, **fmm3d):
        self._solver = solver or FmmSolver()
        self.fmm3d = dict(fmm3d)
        self.parField_x = self.parField_y = self.parField_z = np.zeros(0)

    def calcVelocityFast(self, parPos, activeFlag, targetFlag, alphaX, alphaY, alphaZ):  # noqa: N802
        P = np.asarray(parPos, dtype=np.float64)
        if P.ndim != 2 or P.shape[1] != 3:
            raise ValueError("parPos must be (n, 3)")
        self.parField_x, self.parField_y, self.parField_z = self._solver.velocity3d_alpha(
            P[:, 0], P[:, 1], P[:, 2], alphaX, alphaY, alphaZ, activeFlag, **self.fmm3d)

    def get_Field_x(self):  # noqa: N802
        return self.parField_x

    def get_Field_y(self):  # noqa: N802
        return self.parField_y

    def get_Field_z(self):  # noqa: N802
        return self.parField_z
