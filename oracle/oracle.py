"""ctypes front-end of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  It loads

  oracle/libsister_oracle.so   the scalar C restatement (oracle/sister_oracle.c)
  oracle/_ref/libsister_ref.so the reference's own census.cpp / sgm.cpp /
                               postprocess.cpp behind the same orchestration
                               (built by `make -C oracle ref` where the
                               reference checkout exists)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsister_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsister_ref.so")

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)


class CpuDebug(C.Structure):
    _fields_ = [
        ("census", _u64p * 8),
        ("wta_l", _f32p * 4),
        ("wta_r", _f32p * 4),
        ("med_l", _f32p * 4),
        ("med_r", _f32p * 4),
        ("mask", _u8p * 4),
        ("msum", _u16p * 3),
        ("sum", _u16p * 3),
        ("disp_pad", _f32p * 3),
    ]


def build(ref=True):
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference/cpp"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_libs = {}


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle` / `make -C oracle ref`)")
        lib = C.CDLL(path)
        _libs[path] = lib
    return _libs[path]


def lib():
    l = _load(LIB_PATH)
    l.oracle_census.argtypes = [_u8p, C.c_int, C.c_int, _u64p]
    l.oracle_cost.argtypes = [_u64p, _u64p, C.c_int, C.c_int, C.c_int, _u16p]
    l.oracle_wta_left.argtypes = [_u16p, C.c_int, C.c_int, C.c_int, _f32p]
    l.oracle_wta_right.argtypes = [_u16p, C.c_int, C.c_int, C.c_int, _f32p]
    l.oracle_median_inplace.argtypes = [_f32p, C.c_int, C.c_int]
    l.oracle_lrcheck.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int]
    l.oracle_sgm.argtypes = [_u16p, C.c_int, C.c_int, C.c_int, _u16p]
    l.oracle_compute.argtypes = [_u8p * 5, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _u16p * 3, _f32p, _u16p, C.POINTER(CpuDebug)]
    l.oracle_compute.restype = C.c_int
    return l


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    l = _load(REF_PATH)
    l.ref_compute.argtypes = [_u8p * 5, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              _u16p * 3, _f32p, _u16p, C.POINTER(CpuDebug)]
    l.ref_compute.restype = C.c_int
    l.ref_census.argtypes = [_u8p, C.c_int, C.c_int, _u64p]
    return l


def _ptr(a, t):
    return a.ctypes.data_as(t)


def census(img, use_ref=False):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    out = np.empty(img.shape, np.uint64)
    if use_ref:
        ref_lib().ref_census(_ptr(img, _u8p), img.shape[0], img.shape[1], _ptr(out, _u64p))
    else:
        lib().oracle_census(_ptr(img, _u8p), img.shape[0], img.shape[1], _ptr(out, _u64p))
    return out


def cost(cc, cn, D):
    R, Cc = cc.shape
    out = np.empty((R, Cc, D), np.uint16)
    lib().oracle_cost(_ptr(np.ascontiguousarray(cc), _u64p), _ptr(np.ascontiguousarray(cn), _u64p),
                      R, Cc, D, _ptr(out, _u16p))
    return out


def wta_left(dsi):
    R, Cc, D = dsi.shape
    out = np.empty((R, Cc), np.float32)
    lib().oracle_wta_left(_ptr(np.ascontiguousarray(dsi), _u16p), R, Cc, D, _ptr(out, _f32p))
    return out


def wta_right(dsi):
    R, Cc, D = dsi.shape
    out = np.empty((R, Cc), np.float32)
    lib().oracle_wta_right(_ptr(np.ascontiguousarray(dsi), _u16p), R, Cc, D, _ptr(out, _f32p))
    return out


def median_inplace(a):
    a = np.array(a, dtype=np.float32, copy=True)
    lib().oracle_median_inplace(_ptr(a, _f32p), a.shape[0], a.shape[1])
    return a


def sgm(msum):
    H, W, D = msum.shape
    out = np.zeros((H, W, D), np.uint16)
    lib().oracle_sgm(_ptr(np.ascontiguousarray(msum), _u16p), H, W, D, _ptr(out, _u16p))
    return out


def lrcheck(L, R, thr=5):
    """doLRCheck (postprocess.cpp:318-341) on copies; returns the checked L."""
    L = np.array(L, dtype=np.float32, copy=True)
    R = np.array(R, dtype=np.float32, copy=True)
    lib().oracle_lrcheck(_ptr(L, _f32p), _ptr(R, _f32p), L.shape[0], L.shape[1], thr)
    return L


class _Kernels:
    """The stage kernels of doMultiStereo one by one, on numpy arrays, from
    either the oracle's restatement or the reference's compiled kernels
    (oracle/_ref).  Glue written outside orchestrate.c drives these."""

    def __init__(self, use_ref):
        self.use_ref = use_ref
        if use_ref:
            l = ref_lib()
            l.ref_cost.argtypes = [_u8p, _u8p, C.c_int, C.c_int, C.c_int, _u16p]
            l.ref_wta_left.argtypes = [_u16p, C.c_int, C.c_int, C.c_int, _f32p]
            l.ref_wta_right.argtypes = [_u16p, C.c_int, C.c_int, C.c_int, _f32p]
            l.ref_median.argtypes = [_f32p, C.c_int, C.c_int]
            l.ref_lrcheck.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int]
            l.ref_sgm.argtypes = [_u16p, _u8p, C.c_int, C.c_int, C.c_int, _u16p]
            self.l = l

    def census(self, img):
        return census(img, use_ref=self.use_ref)

    def cost(self, imc, imn, D):
        """ad_census (census.cpp:149-158) of two frame images; rows R-2, R-1
        of the volume are never written and stay 0 (zeroed allocation)."""
        imc = np.ascontiguousarray(imc, dtype=np.uint8)
        imn = np.ascontiguousarray(imn, dtype=np.uint8)
        if not self.use_ref:
            return cost(census(imc), census(imn), D)
        R, Cc = imc.shape
        out = np.zeros((R, Cc, D), np.uint16)
        self.l.ref_cost(_ptr(imc, _u8p), _ptr(imn, _u8p), R, Cc, D, _ptr(out, _u16p))
        return out

    def _wta(self, dsi, right):
        if not self.use_ref:
            return wta_right(dsi) if right else wta_left(dsi)
        dsi = np.array(dsi, dtype=np.uint16, copy=True)
        R, Cc, D = dsi.shape
        out = np.empty((R, Cc), np.float32)
        (self.l.ref_wta_right if right else self.l.ref_wta_left)(_ptr(dsi, _u16p), R, Cc, D, _ptr(out, _f32p))
        return out

    def wta_left(self, dsi):
        return self._wta(dsi, False)

    def wta_right(self, dsi):
        return self._wta(dsi, True)

    def median_inplace(self, a):
        if not self.use_ref:
            return median_inplace(a)
        a = np.array(a, dtype=np.float32, copy=True)
        self.l.ref_median(_ptr(a, _f32p), a.shape[0], a.shape[1])
        return a

    def lrcheck(self, L, R, thr=5):
        if not self.use_ref:
            return lrcheck(L, R, thr)
        L = np.array(L, dtype=np.float32, copy=True)
        R = np.array(R, dtype=np.float32, copy=True)
        self.l.ref_lrcheck(_ptr(L, _f32p), _ptr(R, _f32p), L.shape[0], L.shape[1], thr)
        return L

    def sgm(self, msum, img):
        if not self.use_ref:
            return sgm(msum)
        msum = np.array(msum, dtype=np.uint16, copy=True)
        img = np.ascontiguousarray(img, dtype=np.uint8)
        H, W, D = msum.shape
        out = np.zeros((H, W, D), np.uint16)
        self.l.ref_sgm(_ptr(msum, _u16p), _ptr(img, _u8p), H, W, D, _ptr(out, _u16p))
        return out


def kernels(use_ref=False):
    return _Kernels(use_ref)


def compute(views, D, modes=7, debug=False, use_ref=False, subpix=True):
    """Run compute_disparities on the CPU.

    views: 5 uint8 HxW arrays (center, right, top, left, bottom).
    Returns dict with 'multiview', 'horizontal', 'vertical' (uint16, x255),
    'disp' (uint16 integer multiview disparity), 'subpix' (float32) and, with
    debug=True, the intermediates named in orchestrate.h.
    """
    views = [np.ascontiguousarray(v, dtype=np.uint8) for v in views]
    H, W = views[0].shape
    Wp, Hp = W + 2 * D, H + 2 * D
    res = {}
    outs = []
    names = ("multiview", "horizontal", "vertical")
    for m in range(3):
        if modes & (1 << m):
            a = np.zeros((H, W), np.uint16)
            res[names[m]] = a
            outs.append(_ptr(a, _u16p))
        else:
            outs.append(_u16p())
    vp = (_u8p * 5)(*[_ptr(v, _u8p) for v in views])
    op = (_u16p * 3)(*outs)
    sp = None
    dp = np.zeros((H, W), np.uint16)
    res["disp"] = dp
    if subpix and (modes & 1):
        sp = np.zeros((H, W), np.float32)
        res["subpix"] = sp
    dbg = None
    keep = []
    if debug:
        dbg = CpuDebug()
        npx = Wp * Hp
        shapes = [(Hp, Wp), (Wp, Hp), (Hp, Wp), (Wp, Hp)]
        cen = []
        for F in range(4):
            for j in range(2):
                a = np.zeros(shapes[F], np.uint64)
                cen.append(a)
                dbg.census[2 * F + j] = _ptr(a, _u64p)
        res["census"] = cen
        for name in ("wta_l", "wta_r", "med_l", "med_r"):
            arrs = []
            for F in range(4):
                a = np.zeros(shapes[F], np.float32)
                arrs.append(a)
                getattr(dbg, name)[F] = _ptr(a, _f32p)
            res[name] = arrs
        masks = []
        for F in range(4):
            a = np.zeros((Hp, Wp), np.uint8)
            masks.append(a)
            dbg.mask[F] = _ptr(a, _u8p)
        res["mask"] = masks
        for name, dt, shp in (("msum", np.uint16, (Hp, Wp, D)), ("sum", np.uint16, (Hp, Wp, D)),
                              ("disp_pad", np.float32, (Hp, Wp))):
            arrs = []
            for m in range(3):
                if modes & (1 << m):
                    a = np.zeros(shp, dt)
                    arrs.append(a)
                    getattr(dbg, name)[m] = _ptr(a, _u16p if dt == np.uint16 else _f32p)
                else:
                    arrs.append(None)
            res[name] = arrs
        keep.append(npx)
    fn = ref_lib().ref_compute if use_ref else lib().oracle_compute
    rc = fn(vp, W, H, W, D, modes, op, _ptr(sp, _f32p) if sp is not None else _f32p(),
            _ptr(dp, _u16p), C.byref(dbg) if dbg is not None else C.POINTER(CpuDebug)())
    if rc != 0:
        raise ValueError(f"compute returned {rc} (W={W} H={H} D={D})")
    return res
