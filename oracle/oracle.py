"""ctypes access to the CPU checker (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.  The product package never does.

Two libraries are wrapped:

* ``liboracle_fmm2d.so`` -- our C restatement of the reference FMM path
  (fmm2d_oracle.c), always built by ``make -C oracle``;
* ``_ref/libref_fmm2d.so`` -- the reference's own ``TreeCell``/``fmm2D``
  sources compiled with ``ref_harness.cpp`` (``make -C oracle ref``, only where
  the reference tree exists; the built file travels with the repository).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_fmm2d.so")
REF_SO = os.path.join(HERE, "_ref", "libref_fmm2d.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_int64)
_up = C.POINTER(C.c_uint8)


class OrcOpts(C.Structure):
    _fields_ = [
        ("p", C.c_int),
        ("max_par", C.c_int),
        ("max_src", C.c_int),
        ("exp_tree", C.c_double),
        ("min_level", C.c_int),
        ("max_level", C.c_int),
        ("binom_mode", C.c_int),
        ("list4_mode", C.c_int),
        ("stage_mask", C.c_int),
        ("icutoff", C.c_int),
    ]


class O3Opts(C.Structure):
    """Options of the 3-D oracle (fmm3d_oracle.h)."""
    _fields_ = [
        ("crit_num", C.c_int),
        ("exp_tree", C.c_double),
        ("center", C.c_double * 3),
        ("support", C.c_double),
        ("max_depth", C.c_int),
        ("type", C.c_int),
        ("max_par", C.c_int),
        ("max_src", C.c_int),
    ]


class T1Opts(C.Structure):
    """Options of the type-1 oracle (fmm2d_type1_oracle.h)."""
    _fields_ = [
        ("n_s", C.c_int),
        ("lmax", C.c_int),
        ("ndp", C.c_int),
        ("icutoff", C.c_int),
        ("l2l_mode", C.c_int),
    ]


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref and os.path.isdir("/root/reference/src/FMM"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        L.orc_default_opts.argtypes = [C.POINTER(OrcOpts)]
        L.orc_tree_build.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _up,
                                     C.POINTER(OrcOpts)]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_counts.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip]
        L.orc_tree_geometry.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tree_cells.argtypes = [C.c_void_p, _lp, _ip, _ip, _ip, _ip, _ip, _up, _up]
        L.orc_tree_leaves.argtypes = [C.c_void_p, _lp, _ip, _ip]
        L.orc_int_list.argtypes = [C.c_void_p, C.c_int64, _lp, _ip, _lp, _ip, C.c_int]
        L.orc_ext_list.argtypes = [C.c_void_p, C.c_int64, _lp, _ip, _lp, _ip, C.c_int]
        L.orc_fmm_field.argtypes = [C.c_int, _dp, _dp, _dp, _up, C.POINTER(OrcOpts), _dp, _dp]
        L.orc_velocity.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _up, C.POINTER(OrcOpts),
                                   C.c_double, C.c_double, _dp, _dp]
        L.orc_fmm_expansions.argtypes = [C.c_int, _dp, _dp, _dp, _up, C.POINTER(OrcOpts),
                                         _lp, _dp, _dp]
        L.orc_direct_field.argtypes = [C.c_int, _dp, _dp, _dp, _up, C.c_int, _ip, _dp, _dp]
        L.orc_direct_field.restype = None
        L.orc_fmm_potential.argtypes = [C.c_int, _dp, _dp, _dp, _up, C.POINTER(OrcOpts), _dp]
        L.orc_direct_potential.argtypes = [C.c_int, _dp, _dp, _dp, _up, C.c_int, _ip, _dp]
        L.orc_direct_potential.restype = None
        L.orc_m2l0.argtypes = [_dp, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_m2l0.restype = None
        L.orc_binom_ref.argtypes = [C.c_int, C.c_int]
        L.orc_binom_exact.argtypes = [C.c_int, C.c_int]
        L.orc_binom_exact.restype = C.c_double
        for name in ("orc_p2m",):
            getattr(L, name).argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double,
                                         C.c_int, _dp]
        L.orc_m2m.argtypes = [_dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_m2l.argtypes = [_dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_l2l.argtypes = [_dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_m2p.argtypes = [_dp, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_l2p.argtypes = [_dp, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.orc_t1_default_opts.argtypes = [C.POINTER(T1Opts)]
        L.orc_t1_bico.argtypes = [C.c_int, _ip]
        L.orc_t1_bico.restype = C.c_double
        L.orc_t1_l2l_table.argtypes = [C.c_int, C.c_int, _dp, _ip]
        L.orc_t1_build.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, C.POINTER(T1Opts)]
        L.orc_t1_free.argtypes = [C.c_void_p]
        for name in ("orc_t1_nboxes", "orc_t1_depth", "orc_t1_nlost"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_t1_root.argtypes = [C.c_void_p, _dp]
        L.orc_t1_boxes.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip, _dp]
        L.orc_t1_box_particles.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        L.orc_t1_list.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, C.c_int]
        L.orc_t1_velocity.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.POINTER(T1Opts), _dp, _dp]
        L.orc_t1_direct.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, _ip, _dp, _dp]
        L.orc3_default_opts.argtypes = [C.POINTER(O3Opts)]
        L.orc3_build.argtypes = [C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _dp, _up,
                                 C.POINTER(O3Opts)]
        L.orc3_free.argtypes = [C.c_void_p]
        L.orc3_ncells.argtypes = [C.c_void_p]
        L.orc3_depth.argtypes = [C.c_void_p]
        L.orc3_cells.argtypes = [C.c_void_p, _ip, _lp, _ip, _up, _up, _dp, _dp]
        L.orc3_leaf_particles.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        L.orc3_velocity_alpha.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _up,
                                          C.POINTER(O3Opts), _dp, _dp, _dp]
        L.orc3_velocity.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _up,
                                    C.POINTER(O3Opts), C.c_double, C.c_double, C.c_double,
                                    _dp, _dp, _dp]
        L.orc3_m2p_mul.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc3_p2m.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc3_m2m.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc3_direct.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _up, C.c_int, _ip,
                                  _dp, _dp, _dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        for name in ("orc3_default_opts", "orc3_free", "orc3_cells", "orc3_m2p_mul",
                     "orc3_p2m", "orc3_m2m", "orc3_direct"):
            getattr(L, name).restype = None
        for name in ("orc_t1_default_opts", "orc_t1_l2l_table", "orc_t1_free", "orc_t1_root",
                     "orc_t1_boxes", "orc_t1_direct"):
            getattr(L, name).restype = None
        for name in ("orc_p2m", "orc_m2m", "orc_m2l", "orc_l2l", "orc_m2p", "orc_l2p",
                     "orc_tree_free", "orc_tree_geometry", "orc_tree_cells",
                     "orc_tree_leaves", "orc_default_opts"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference build (raises if it was never compiled)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_SO)
        L = C.CDLL(REF_SO)
        L.ref_set_params.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_set_threads.restype = None
        L.ref_binom.argtypes = [C.c_int, C.c_int]
        L.ref_build_tree.argtypes = [C.c_int, _dp, _dp, _up]
        L.ref_tree_counts.argtypes = [_ip, _ip, _ip, _ip]
        L.ref_tree_geometry.argtypes = [_dp, _dp]
        L.ref_tree_cells.argtypes = [_lp, _ip, _ip, _ip, _ip, _ip, _up, _up]
        L.ref_tree_leaves.argtypes = [_lp, _ip, _ip]
        L.ref_int_list.argtypes = [C.c_int64, _lp, _ip, _lp, _ip, C.c_int]
        L.ref_ext_list.argtypes = [C.c_int64, _lp, _ip, _lp, _ip, C.c_int]
        L.ref_fmm_field.argtypes = [C.c_int, _dp, _dp, _dp, _up, _dp, _dp, _dp]
        L.ref_fmm_potential.argtypes = [C.c_int, _dp, _dp, _dp, _up, _dp]
        L.ref_type1_velocity.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                         C.c_int, C.c_int, _dp, _dp]
        _ref = L
    return _ref


# --------------------------------------------------------------------------
# options
# --------------------------------------------------------------------------
def opts(p=10, max_par=75, max_src=50, exp_tree=1.0, min_level=1, max_level=0,
         binom_mode=1, list4_mode=1, stage_mask=0, icutoff=0) -> OrcOpts:
    return OrcOpts(p, max_par, max_src, exp_tree, min_level, max_level, binom_mode,
                   list4_mode, stage_mask, icutoff)


def regul_factor(icutoff, rho2):
    """2 pi q of regul_func_2d (regularization_function_2d.cpp:11-37) as a
    function of rho^2 = r^2 / sij^2: 0 singular (1), 1 high-order algebraic,
    2 Gaussian, 3 super-Gaussian."""
    rho2 = np.asarray(rho2, dtype=np.float64)
    if icutoff == 0:
        return np.ones_like(rho2)
    if icutoff == 1:
        return rho2 * (rho2 + 2.0) / (rho2 + 1.0) ** 2
    qt = -rho2 / 2.0
    if icutoff == 2:
        return 1.0 - np.exp(qt)
    return 1.0 - (1.0 - qt) * np.exp(qt)


def regularized_direct_velocity(x, y, s, omega, icutoff, targets=None, r_min=0.0):
    """Regularized Biot-Savart direct sum (biotsavart_direct_2d.cpp:5-41):
    u_i -= dy gz_j q / r^2, v_i += dx gz_j q / r^2 over all sources j with
    r > r_min, gz = omega s^2, q = regul_factor / (2 pi) at
    sij = sqrt(si^2 + sj^2) / 2.  Rotational part only (no free stream)."""
    x, y, s, w = _d(x), _d(y), _d(s), _d(omega)
    gz = w * s ** 2
    idx = np.arange(len(x)) if targets is None else np.asarray(targets)
    u, v = np.zeros(len(idx)), np.zeros(len(idx))
    for c0 in range(0, len(idx), 512):
        t = idx[c0:c0 + 512]
        dx = x[t, None] - x[None, :]
        dy = y[t, None] - y[None, :]
        r2 = dx * dx + dy * dy
        sij2 = 0.25 * (s[t, None] ** 2 + s[None, :] ** 2)
        keep = r2 > r_min * r_min
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(keep, regul_factor(icutoff, r2 / sij2) / (2 * np.pi) / r2, 0.0)
        u[c0:c0 + 512] = -np.sum(dy * gz[None, :] * f, axis=1)
        v[c0:c0 + 512] = np.sum(dx * gz[None, :] * f, axis=1)
    return u, v


class OracleError(RuntimeError):
    pass


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what} failed with status {st}")


# --------------------------------------------------------------------------
# oracle: tree, lists, field, direct sum
# --------------------------------------------------------------------------
class Tree:
    """A tree built by the oracle (reference semantics)."""

    def __init__(self, x, y, active, o: OrcOpts | None = None):
        L = lib()
        self._x, self._y, self._a = _d(x), _d(y), _u8(active)
        self.o = o or opts()
        h = C.c_void_p()
        _check(L.orc_tree_build(C.byref(h), len(self._x), _ptr(self._x, _dp),
                                _ptr(self._y, _dp), _ptr(self._a, _up), C.byref(self.o)),
               "orc_tree_build")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def counts(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().orc_tree_counts(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def geometry(self):
        r = C.c_double()
        piv = np.zeros(2)
        lib().orc_tree_geometry(self.h, C.byref(r), _ptr(piv, _dp))
        return r.value, piv

    def cells(self):
        return _cells(lib().orc_tree_cells, self.counts()[0], self.h)

    def leaves(self):
        nc, nl, _, _ = self.counts()
        lid = np.zeros(nl, np.int64)
        poff = np.zeros(nl + 1, np.int32)
        pidx = np.zeros(len(self._x), np.int32)
        lib().orc_tree_leaves(self.h, _ptr(lid, _lp), _ptr(poff, _ip), _ptr(pidx, _ip))
        return lid, poff, pidx

    def int_list(self, cid):
        return _lists(lib().orc_int_list, self.h, cid)

    def ext_list(self, cid):
        return _lists(lib().orc_ext_list, self.h, cid)


def _cells(fn, n, *pre):
    out = dict(id=np.zeros(n, np.int64), level=np.zeros(n, np.int32),
               ix=np.zeros(n, np.int32), iy=np.zeros(n, np.int32),
               par_num=np.zeros(n, np.int32), src_num=np.zeros(n, np.int32),
               is_leaf=np.zeros(n, np.uint8), is_active=np.zeros(n, np.uint8))
    fn(*pre, _ptr(out["id"], _lp), _ptr(out["level"], _ip), _ptr(out["ix"], _ip),
       _ptr(out["iy"], _ip), _ptr(out["par_num"], _ip), _ptr(out["src_num"], _ip),
       _ptr(out["is_leaf"], _up), _ptr(out["is_active"], _up))
    return out


def _lists(fn, *pre_and_id, cap=4096):
    a = np.zeros(cap, np.int64)
    b = np.zeros(cap, np.int64)
    na, nb = C.c_int(), C.c_int()
    st = fn(*pre_and_id, _ptr(a, _lp), C.byref(na), _ptr(b, _lp), C.byref(nb), cap)
    _check(st, fn.__name__)
    if na.value > cap or nb.value > cap:
        return _lists(fn, *pre_and_id, cap=max(na.value, nb.value))
    return a[: na.value].copy(), b[: nb.value].copy()


def fmm_field(x, y, q, active, o: OrcOpts | None = None):
    L = lib()
    x, y, q, a = _d(x), _d(y), _d(q), _u8(active)
    n = len(x)
    ex, ey = np.zeros(n), np.zeros(n)
    o = o or opts()
    _check(L.orc_fmm_field(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), _ptr(a, _up),
                           C.byref(o), _ptr(ex, _dp), _ptr(ey, _dp)), "orc_fmm_field")
    return ex, ey


def fmm_expansions(x, y, q, active, o: OrcOpts | None = None):
    """(ids sorted, ak, bk) with ak/bk of shape (n_cells, p+1) complex."""
    L = lib()
    o = o or opts()
    x, y, q, a = _d(x), _d(y), _d(q), _u8(active)
    nc = Tree(x, y, a, o).counts()[0]
    ids = np.zeros(nc, np.int64)
    ak = np.zeros((nc, o.p + 1), np.complex128)
    bk = np.zeros((nc, o.p + 1), np.complex128)
    _check(L.orc_fmm_expansions(len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp),
                                _ptr(a, _up), C.byref(o), _ptr(ids, _lp),
                                ak.ctypes.data_as(_dp), bk.ctypes.data_as(_dp)),
           "orc_fmm_expansions")
    return ids, ak, bk


def velocity(x, y, s, omega, active, o: OrcOpts | None = None, u_inf=1.0, v_inf=0.0):
    L = lib()
    x, y, s, w, a = _d(x), _d(y), _d(s), _d(omega), _u8(active)
    n = len(x)
    u, v = np.zeros(n), np.zeros(n)
    o = o or opts()
    _check(L.orc_velocity(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(w, _dp),
                          _ptr(a, _up), C.byref(o), u_inf, v_inf, _ptr(u, _dp),
                          _ptr(v, _dp)), "orc_velocity")
    return u, v


def advect(x, y, u, v, dt, scheme="euler"):
    """Advection sub-step, returning new (x, y) (main_advection.cpp:44-83).

    "euler": x + u dt (:44-61, the scheme the reference compiles).  "rk2":
    x + (u + 2 u1 + 2 u2 + u3) dt / 6 with u1 = u/2, u2 = u1/2, u3 = u2
    (:64-83), evaluated left to right as the reference's loop does; numpy
    rounds every operation once (no FMA), like the reference's host build."""
    x, y, u, v = _d(x), _d(y), _d(u), _d(v)
    if scheme == "rk2":
        u1 = u / 2
        u2 = u1 / 2
        u3 = u2
        v1 = v / 2
        v2 = v1 / 2
        v3 = v2
        return x + (u + 2 * u1 + 2 * u2 + u3) * dt / 6, y + (v + 2 * v1 + 2 * v2 + v3) * dt / 6
    if scheme != "euler":
        raise ValueError(scheme)
    return x + u * dt, y + v * dt


def inviscid_loop(x, y, s, omega, active, dt, steps, o: OrcOpts | None = None, u_inf=1.0,
                  v_inf=0.0, scheme="euler"):
    """The reference's loop restricted to the velocity path (main.cpp:296-429:
    advection at :373, the velocity solve at :410), from the velocity at the
    initial positions.  Returns (x, y, u, v) after `steps` steps."""
    u, v = velocity(x, y, s, omega, active, o, u_inf, v_inf)
    x, y = _d(x), _d(y)
    for _ in range(steps):
        x, y = advect(x, y, u, v, dt, scheme)
        u, v = velocity(x, y, s, omega, active, o, u_inf, v_inf)
    return x, y, u, v


def fmm_potential(x, y, q, active, o: OrcOpts | None = None):
    """fmm2D::calcPotential + get_Potential (fmm2D.cpp:681-896): phi per particle."""
    L = lib()
    o = o or opts()
    x, y, q, a = _d(x), _d(y), _d(q), _u8(active)
    phi = np.zeros(len(x))
    _check(L.orc_fmm_potential(len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), _ptr(a, _up),
                               C.byref(o), _ptr(phi, _dp)), "orc_fmm_potential")
    return phi


def direct_potential(x, y, q, targets=None, src_mask=None):
    """phi_i = sum_{j != i} q_j log(r_ij^2) / 2 at `targets` (indices) or all."""
    L = lib()
    x, y, q = _d(x), _d(y), _d(q)
    n = len(x)
    if targets is None:
        nt, tp = n, None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = len(t), _ptr(t, _ip)
    phi = np.zeros(nt)
    m = None if src_mask is None else _ptr(_u8(src_mask), _up)
    L.orc_direct_potential(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), m, nt, tp,
                           _ptr(phi, _dp))
    return phi


def direct_field(x, y, q, targets=None, src_mask=None):
    """Singular FP64 direct sum (j == i skipped) at `targets` (indices) or all."""
    L = lib()
    x, y, q = _d(x), _d(y), _d(q)
    n = len(x)
    if targets is None:
        nt, tp = n, None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int32)
        nt, tp = len(t), _ptr(t, _ip)
    ex, ey = np.zeros(nt), np.zeros(nt)
    m = None if src_mask is None else _ptr(_u8(src_mask), _up)
    L.orc_direct_field(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), m, nt, tp,
                       _ptr(ex, _dp), _ptr(ey, _dp))
    return ex, ey


# --------------------------------------------------------------------------
# the reference build
# --------------------------------------------------------------------------
class RefTree:
    """A tree built by the reference's own TreeCell (oracle/_ref)."""

    def __init__(self, x, y, active, p=10, max_par=75, max_src=50, exp_tree=1.0):
        R = ref()
        R.ref_set_params(p, max_par, max_src, exp_tree)
        self._x, self._y, self._a = _d(x), _d(y), _u8(active)
        _check(R.ref_build_tree(len(self._x), _ptr(self._x, _dp), _ptr(self._y, _dp),
                                _ptr(self._a, _up)), "ref_build_tree")

    def counts(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ref().ref_tree_counts(C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def geometry(self):
        r = C.c_double()
        piv = np.zeros(2)
        ref().ref_tree_geometry(C.byref(r), _ptr(piv, _dp))
        return r.value, piv

    def cells(self):
        return _cells(ref().ref_tree_cells, self.counts()[0])

    def leaves(self):
        nc, nl, _, _ = self.counts()
        lid = np.zeros(nl, np.int64)
        poff = np.zeros(nl + 1, np.int32)
        pidx = np.zeros(len(self._x), np.int32)
        ref().ref_tree_leaves(_ptr(lid, _lp), _ptr(poff, _ip), _ptr(pidx, _ip))
        return lid, poff, pidx

    def int_list(self, cid):
        return _lists(ref().ref_int_list, cid)

    def ext_list(self, cid):
        return _lists(ref().ref_ext_list, cid)


def ref_fmm_field(x, y, q, active, p=10, max_par=75, max_src=50, exp_tree=1.0,
                  threads=None):
    """Field from the reference's own FMM; also returns (tree_s, fmm_s)."""
    R = ref()
    R.ref_set_params(p, max_par, max_src, exp_tree)
    if threads:
        R.ref_set_threads(int(threads))
    x, y, q, a = _d(x), _d(y), _d(q), _u8(active)
    n = len(x)
    ex, ey, sec = np.zeros(n), np.zeros(n), np.zeros(2)
    _check(R.ref_fmm_field(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), _ptr(a, _up),
                           _ptr(ex, _dp), _ptr(ey, _dp), _ptr(sec, _dp)), "ref_fmm_field")
    return ex, ey, (float(sec[0]), float(sec[1]))


def ref_fmm_potential(x, y, q, active, p=10, max_par=75, max_src=50, exp_tree=1.0):
    """Potential from the reference's own fmm2D::calcPotential."""
    R = ref()
    R.ref_set_params(p, max_par, max_src, exp_tree)
    x, y, q, a = _d(x), _d(y), _d(q), _u8(active)
    phi = np.zeros(len(x))
    _check(R.ref_fmm_potential(len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(q, _dp), _ptr(a, _up),
                               _ptr(phi, _dp)), "ref_fmm_potential")
    return phi


def ref_binom(n, k) -> int:
    return int(ref().ref_binom(n, k))


# --------------------------------------------------------------------------
# type-1 ("old") FMM: biotsavart_fmm_2d and its subroutines
# --------------------------------------------------------------------------
T1_STATUS = {1: "bad argument", 2: "non-finite input", 5: "out of memory",
             7: "domain too small for the absolute box tolerance"}


def t1_opts(n_s=50, lmax=7, ndp=10, icutoff=2, l2l_mode=1) -> T1Opts:
    """Defaults: global.cpp:316-328; l2l_mode 1 is the shift as the reference
    runs it."""
    return T1Opts(n_s, lmax, ndp, icutoff, l2l_mode)


def t1_bico(n, k):
    """(value, k after the call) of the reference's bico."""
    kk = C.c_int(k)
    c = lib().orc_t1_bico(n, C.byref(kk))
    return c, kk.value


def t1_l2l_table(ndp, l2l_mode=1):
    m = ndp + 1
    coef = np.zeros(m * m)
    expo = np.zeros(m * m, np.int32)
    lib().orc_t1_l2l_table(ndp, l2l_mode, _ptr(coef, _dp), _ptr(expo, _ip))
    return coef.reshape(m, m), expo.reshape(m, m)


class T1Tree:
    """The type-1 box hierarchy (hierarchy_mesh) and its lists."""

    def __init__(self, x, y, o: T1Opts | None = None):
        L = lib()
        self._x, self._y = _d(x), _d(y)
        self.o = o or t1_opts()
        h = C.c_void_p()
        st = L.orc_t1_build(C.byref(h), len(self._x), _ptr(self._x, _dp), _ptr(self._y, _dp),
                            C.byref(self.o))
        if st:
            raise OracleError(f"orc_t1_build: {T1_STATUS.get(st, st)}")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_t1_free(self.h)
            self.h = None

    @property
    def nboxes(self):
        return lib().orc_t1_nboxes(self.h)

    @property
    def depth(self):
        return lib().orc_t1_depth(self.h)

    @property
    def nlost(self):
        return lib().orc_t1_nlost(self.h)

    def root(self):
        g = np.zeros(5)
        lib().orc_t1_root(self.h, _ptr(g, _dp))
        return g

    def boxes(self):
        nb = self.nboxes
        out = {k: np.zeros(nb, np.int32) for k in ("level", "ix", "iy", "count", "nchild",
                                                   "parent")}
        bnd = np.zeros((nb, 4))
        lib().orc_t1_boxes(self.h, *(_ptr(out[k], _ip) for k in ("level", "ix", "iy", "count",
                                                                "nchild", "parent")),
                           _ptr(bnd, _dp))
        out["bounds"] = bnd
        return out

    def particles(self, b):
        cap = len(self._x)
        buf = np.zeros(max(cap, 1), np.int32)
        m = lib().orc_t1_box_particles(self.h, int(b), _ptr(buf, _ip), cap)
        return buf[:m].copy()

    def list(self, b, which):
        cap = 1024
        while True:
            buf = np.zeros(cap, np.int32)
            m = lib().orc_t1_list(self.h, int(b), int(which), _ptr(buf, _ip), cap)
            if m < 0:
                raise OracleError("orc_t1_list failed")
            if m <= cap:
                return buf[:m].copy()
            cap = m


def t1_velocity(x, y, s, g, o: T1Opts | None = None):
    """Rotational velocity of the type-1 FMM (no free stream)."""
    x, y, s, g = _d(x), _d(y), _d(s), _d(g)
    n = len(x)
    u, v = np.zeros(n), np.zeros(n)
    o = o or t1_opts()
    st = lib().orc_t1_velocity(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(g, _dp),
                               C.byref(o), _ptr(u, _dp), _ptr(v, _dp))
    if st:
        raise OracleError(f"orc_t1_velocity: {T1_STATUS.get(st, st)}")
    return u, v


def t1_direct(x, y, s, g, icutoff=2, targets=None):
    """Direct sum of the type-1 pair kernel over every pair at r > 0."""
    x, y, s, g = _d(x), _d(y), _d(s), _d(g)
    n = len(x)
    if targets is None:
        m, tp = n, None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int32)
        m, tp = len(t), _ptr(t, _ip)
    u, v = np.zeros(m), np.zeros(m)
    lib().orc_t1_direct(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(g, _dp), int(icutoff),
                        m, tp, _ptr(u, _dp), _ptr(v, _dp))
    return u, v


def ref_type1_velocity(x, y, s, g, n_s=50, lmax=7, ndp=10, icutoff=2):
    """The reference's own biotsavart_fmm_2d (rotational part)."""
    R = ref()
    x, y, s, g = _d(x), _d(y), _d(s), _d(g)
    n = len(x)
    u, v = np.zeros(n), np.zeros(n)
    _check(R.ref_type1_velocity(n, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(g, _dp),
                                int(n_s), int(lmax), int(ndp), int(icutoff), _ptr(u, _dp),
                                _ptr(v, _dp)), "ref_type1_velocity")
    return u, v


# --------------------------------------------------------------------------
# the 3-D vector FMM (fmm3d_oracle.c; reference fast3DFMM.cpp)
# --------------------------------------------------------------------------
def set_threads(n: int) -> None:
    """OpenMP threads of this thread's later oracle calls (and of the
    reference build's, which shares the OpenMP runtime)."""
    lib().orc_set_threads(int(n))


def o3_opts(crit_num=30, exp_tree=1.0, center=(0.0, 0.0, 0.0), support=2.0,
            max_depth=21, type=1, max_par=50, max_src=30) -> O3Opts:
    o = O3Opts()
    lib().orc3_default_opts(C.byref(o))
    o.crit_num, o.exp_tree, o.support, o.max_depth = crit_num, exp_tree, support, max_depth
    o.type, o.max_par, o.max_src = type, max_par, max_src
    for d in range(3):
        o.center[d] = center[d]
    return o


class O3Tree:
    """The octree of FMMCell::generateTree (fast3DFMM.cpp:485-745): cells in
    creation order with level, path key, count, leaf, active, centre, size."""

    def __init__(self, x, y, z, active, o: O3Opts | None = None):
        L = lib()
        x, y, z, a = _d(x), _d(y), _d(z), _u8(active)
        o = o or o3_opts()
        self._h = C.c_void_p()
        _check(L.orc3_build(C.byref(self._h), len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                            _ptr(a, _up), C.byref(o)), "orc3_build")
        nc = L.orc3_ncells(self._h)
        self.depth = L.orc3_depth(self._h)
        self.level = np.zeros(nc, np.int32)
        self.key = np.zeros(nc, np.int64)
        self.count = np.zeros(nc, np.int32)
        self.leaf = np.zeros(nc, np.uint8)
        self.active = np.zeros(nc, np.uint8)
        self.centre = np.zeros((nc, 3))
        self.size = np.zeros(nc)
        L.orc3_cells(self._h, _ptr(self.level, _ip), _ptr(self.key, _lp), _ptr(self.count, _ip),
                     _ptr(self.leaf, _up), _ptr(self.active, _up), _ptr(self.centre, _dp),
                     _ptr(self.size, _dp))

    def particles(self, c):
        out = np.zeros(max(int(self.count[c]), 1), np.int32)
        k = lib().orc3_leaf_particles(self._h, int(c), _ptr(out, _ip), len(out))
        return out[:max(k, 0)].copy()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc3_free(self._h)
            self._h = None


def o3_velocity_alpha(x, y, z, ax, ay, az, active, o: O3Opts | None = None):
    """Rotational velocity from vortex strengths (fmm3D::calcVelocityFast)."""
    L = lib()
    x, y, z, ax, ay, az, a = _d(x), _d(y), _d(z), _d(ax), _d(ay), _d(az), _u8(active)
    n = len(x)
    u, v, w = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(L.orc3_velocity_alpha(n, *(_ptr(t, _dp) for t in (x, y, z, ax, ay, az)), _ptr(a, _up),
                                 C.byref(o or o3_opts()), _ptr(u, _dp), _ptr(v, _dp),
                                 _ptr(w, _dp)), "orc3_velocity_alpha")
    return u, v, w


def o3_velocity(x, y, z, s, wx, wy, wz, active, o: O3Opts | None = None, u_inf=0.0,
                v_inf=0.0, w_inf=0.0):
    """Velocity of every particle as VelocityCalc::velocity_fmm_3d_fast plus the
    free stream (velocity_calc.cpp:2384-2525, 221-236)."""
    L = lib()
    arrs = [_d(t) for t in (x, y, z, s, wx, wy, wz)]
    a = _u8(active)
    n = len(arrs[0])
    u, v, w = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(L.orc3_velocity(n, *(_ptr(t, _dp) for t in arrs), _ptr(a, _up),
                           C.byref(o or o3_opts()), u_inf, v_inf, w_inf, _ptr(u, _dp),
                           _ptr(v, _dp), _ptr(w, _dp)), "orc3_velocity")
    return u, v, w


def o3_direct(x, y, z, ax, ay, az, active=None, targets=None):
    """Direct 3-D Biot-Savart sum over active sources (the target skipped)."""
    L = lib()
    x, y, z, ax, ay, az = _d(x), _d(y), _d(z), _d(ax), _d(ay), _d(az)
    a = _u8(active) if active is not None else None
    t = np.ascontiguousarray(targets, np.int32) if targets is not None else None
    m = len(t) if t is not None else len(x)
    u, v, w = np.zeros(m), np.zeros(m), np.zeros(m)
    L.orc3_direct(len(x), *(_ptr(q, _dp) for q in (x, y, z, ax, ay, az)),
                  _ptr(a, _up) if a is not None else None, m,
                  _ptr(t, _ip) if t is not None else None, _ptr(u, _dp), _ptr(v, _dp),
                  _ptr(w, _dp))
    return u, v, w


def o3_m2p_mul(dx, dy, dz):
    mx, my, mz = np.zeros(10), np.zeros(10), np.zeros(10)
    lib().orc3_m2p_mul(dx, dy, dz, _ptr(mx, _dp), _ptr(my, _dp), _ptr(mz, _dp))
    return mx, my, mz


def o3_p2m(mp, src, dx, dy, dz):
    mp = _d(mp).copy()
    lib().orc3_p2m(_ptr(mp, _dp), src, dx, dy, dz)
    return mp


def o3_m2m(mp, mpc, dx, dy, dz):
    mp, mpc = _d(mp).copy(), _d(mpc)
    lib().orc3_m2m(_ptr(mp, _dp), _ptr(mpc, _dp), dx, dy, dz)
    return mp
