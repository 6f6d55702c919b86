"""sister_amd -- MI355X-native SiSter 5-view disparity path.

Python host mirror of the reference's helper class
(cpp/include/sister/SisterMultiviewDisparities.hpp:18-26) over the C ABI of
include/sister.h, implemented by sister_amd/libsister_hip.so (hand-written
gfx950 HIP kernels).  There is no CPU fallback: importing the compute API
without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsister_hip.so")

SISTER_OK, SISTER_EINVAL, SISTER_ENOMEM, SISTER_EDEVICE, SISTER_EBUSY, SISTER_ENOTFOUND = range(6)
MODE_MULTIVIEW, MODE_HORIZONTAL, MODE_VERTICAL, MODE_ALL = 1, 2, 4, 7
NOALLOC_MAX_BYTES = 256 << 20   # SISTER_NOALLOC_MAX_BYTES

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_f32p = C.POINTER(C.c_float)


class SisterError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sister error {code}: {msg}")
        self.code = code


class RigStruct(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("stride", C.c_int), ("disp_count", C.c_int),
        ("view", C.c_void_p * 5),
        ("out_multiview", C.c_void_p), ("out_horizontal", C.c_void_p), ("out_vertical", C.c_void_p),
        ("out_disp", C.c_void_p), ("out_subpix", C.c_void_p),
    ]


class ConfigStruct(C.Structure):
    _fields_ = [("device", C.c_int), ("flags", C.c_int), ("hbm_budget", C.c_size_t)]


class StatsStruct(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
        ("rigs_per_s", C.c_double), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t),
    ]


_lib = None

# every entry point of include/sister.h with its ctypes signature
_SIGS = {
    "sister_create": (C.c_int, [C.POINTER(ConfigStruct), C.POINTER(C.c_void_p)]),
    "sister_destroy": (None, [C.c_void_p]),
    "sister_last_error": (C.c_char_p, [C.c_void_p]),
    "sister_compute_rig": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int]),
    "sister_run_device": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.c_int]),
    "sister_run_device_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.c_int]),
    "sister_sync": (C.c_int, [C.c_void_p]),
    "sister_stream": (C.c_void_p, [C.c_void_p]),
    "sister_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "sister_outside_reference_domain": (C.c_int, [C.c_int]),
    "sister_graph_capture": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.c_int]),
    "sister_graph_launch": (C.c_int, [C.c_void_p]),
    "sister_autotune": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.c_int, C.c_char_p]),
    "sister_tune_describe": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "sister_serve_start": (C.c_int, [C.c_void_p, C.c_int]),
    "sister_submit": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.POINTER(C.c_uint64)]),
    "sister_submit_device": (C.c_int, [C.c_void_p, C.POINTER(RigStruct), C.c_int, C.POINTER(C.c_uint64)]),
    "sister_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]),
    "sister_wait": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sister_serve_stop": (C.c_int, [C.c_void_p]),
    "sister_stats_get": (C.c_int, [C.c_void_p, C.POINTER(StatsStruct)]),
    "sister_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "sister_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.c_int]),
    "sister_debug_fetch": (C.c_int, [C.c_void_p, C.c_int, C.c_char_p, C.c_int, C.c_void_p, C.c_size_t]),
    "sister_debug_keep_sum": (C.c_int, [C.c_void_p, C.c_int]),
    "sister_debug_vsweep": (C.c_int, [C.c_void_p, C.c_int]),
    "sister_debug_sweep_in_place": (C.c_int, [C.c_void_p, C.c_int]),
}


def lib():
    """The loaded HIP library (raises if it is not built: no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m sister_amd.build`")
        # PyTorch-ROCm ships its own libamdhip64 with the same soname as the
        # system one.  Loading torch first makes this library bind to the
        # runtime torch uses, so device tensors and this library share one
        # HIP runtime in the process (two copies cannot both own the GPU).
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def workspace_bytes(W, H, D):
    return int(lib().sister_workspace_bytes(W, H, D))


def outside_reference_domain(D):
    return bool(lib().sister_outside_reference_domain(D))


def bgr_to_gray(img):
    """cv::cvtColor(COLOR_BGR2GRAY) for 8-bit input (hpp:29-33): OpenCV's
    fixed-point weights 1868 (B), 9617 (G), 4899 (R) with 14-bit rounding.
    Gray-replicated inputs are unchanged (the weights sum to 2^14)."""
    a = np.asarray(img)
    if a.ndim == 2:
        return np.ascontiguousarray(a, dtype=np.uint8)
    a = a.astype(np.uint32)
    y = (a[..., 0] * 1868 + a[..., 1] * 9617 + a[..., 2] * 4899 + (1 << 13)) >> 14
    return y.astype(np.uint8)


class Context:
    """One device context (stream, HBM pool).  Not thread safe."""

    def __init__(self, device=0, hbm_budget=0):
        cfg = ConfigStruct(device, 0, hbm_budget)
        h = C.c_void_p()
        rc = lib().sister_create(C.byref(cfg), C.byref(h))
        if rc != SISTER_OK:
            raise SisterError(rc, "sister_create failed (no HIP device?)")
        self._h = h
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            lib().sister_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != SISTER_OK:
            raise SisterError(rc, lib().sister_last_error(self._h).decode())

    @property
    def stream(self):
        return lib().sister_stream(self._h)

    # -- host API ----------------------------------------------------------
    def compute_rig(self, views, D, modes=MODE_ALL, disp=True, subpix=True):
        """Host-buffer form: 5 uint8 HxW views -> dict of numpy outputs."""
        views = [np.ascontiguousarray(v, dtype=np.uint8) for v in views]
        H, W = views[0].shape
        r = RigStruct()
        r.width, r.height, r.stride, r.disp_count = W, H, W, D
        for i, v in enumerate(views):
            if v.shape != (H, W):
                raise ValueError("all five views must have the same shape")
            r.view[i] = v.ctypes.data
        out = {}
        for bit, name, field in ((1, "multiview", "out_multiview"), (2, "horizontal", "out_horizontal"),
                                 (4, "vertical", "out_vertical")):
            if modes & bit:
                a = np.zeros((H, W), np.uint16)
                out[name] = a
                setattr(r, field, a.ctypes.data)
        if disp and (modes & 1):
            out["disp"] = np.zeros((H, W), np.uint16)
            r.out_disp = out["disp"].ctypes.data
        if subpix and (modes & 1):
            out["subpix"] = np.zeros((H, W), np.float32)
            r.out_subpix = out["subpix"].ctypes.data
        self._check(lib().sister_compute_rig(self._h, C.byref(r), modes))
        return out

    # -- device API (torch tensors in HBM) -----------------------------------
    @staticmethod
    def rig_struct(views, D, outs):
        """views: 5 uint8 (H, W) device tensors; outs: dict of device tensors."""
        H, W = views[0].shape
        r = RigStruct()
        r.width, r.height, r.disp_count = W, H, D
        r.stride = views[0].stride(0)
        for i, v in enumerate(views):
            r.view[i] = v.data_ptr()
        for name, field in (("multiview", "out_multiview"), ("horizontal", "out_horizontal"),
                            ("vertical", "out_vertical"), ("disp", "out_disp"), ("subpix", "out_subpix")):
            t = outs.get(name)
            setattr(r, field, t.data_ptr() if t is not None else None)
        return r

    def run_device(self, rigs, modes=MODE_MULTIVIEW):
        arr = (RigStruct * len(rigs))(*rigs)
        self._check(lib().sister_run_device(self._h, arr, len(rigs), modes))

    def run_device_noalloc(self, rigs, modes=MODE_MULTIVIEW):
        """run_device with no device allocation: workspaces in the library's
        fixed scratch array (NOALLOC_MAX_BYTES per GPU)."""
        arr = (RigStruct * len(rigs))(*rigs)
        self._check(lib().sister_run_device_noalloc(self._h, arr, len(rigs), modes))

    def sync(self):
        self._check(lib().sister_sync(self._h))

    def graph_capture(self, rigs, modes=MODE_MULTIVIEW):
        arr = (RigStruct * len(rigs))(*rigs)
        self._check(lib().sister_graph_capture(self._h, arr, len(rigs), modes))

    def graph_launch(self):
        self._check(lib().sister_graph_launch(self._h))

    def autotune(self, rigs, modes=MODE_MULTIVIEW, path=None):
        """Time the launch forms of this batch (graph parts, then each D
        group's SGM and cost-sum forms) and keep the fastest; the batch is
        left captured in its tuned form.  path: forms file read and rewritten."""
        arr = (RigStruct * len(rigs))(*rigs)
        self._check(lib().sister_autotune(self._h, arr, len(rigs), modes,
                                          path.encode() if path else None))

    def tune_table(self):
        buf = C.create_string_buffer(1 << 16)
        self._check(lib().sister_tune_describe(self._h, buf, len(buf)))
        return buf.value.decode()

    # -- serving -------------------------------------------------------------
    def serve_start(self, max_inflight=4):
        self._check(lib().sister_serve_start(self._h, max_inflight))

    def submit(self, rig, modes=MODE_MULTIVIEW, device=False):
        """Queue a rig; device=True when every buffer of the rig is device
        memory (computed in place, no staging copies)."""
        t = C.c_uint64()
        f = lib().sister_submit_device if device else lib().sister_submit
        self._check(f(self._h, C.byref(rig), modes, C.byref(t)))
        return t.value

    def poll(self, ticket):
        d = C.c_int()
        self._check(lib().sister_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        self._check(lib().sister_wait(self._h, ticket))

    def serve_stop(self):
        self._check(lib().sister_serve_stop(self._h))

    def stats(self):
        s = StatsStruct()
        self._check(lib().sister_stats_get(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in StatsStruct._fields_}

    # -- instrumentation -------------------------------------------------------
    def profile(self, on=True):
        self._check(lib().sister_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        n = 16
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int * n)()
        k = lib().sister_profile_read(self._h, names, ms, cnt, n)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k)}

    def keep_sum(self, on=True):
        self._check(lib().sister_debug_keep_sum(self._h, 1 if on else 0))

    def vsweep(self, mode):
        """Vertical SGM families: 0 separate sweeps, 1 fused sweep, -1 automatic."""
        self._check(lib().sister_debug_vsweep(self._h, int(mode)))

    def sweep_in_place(self, on=True):
        """Plans keep the fused sweep's partial sums over the rows family's
        bytes (the layout of plans short of memory)."""
        self._check(lib().sister_debug_sweep_in_place(self._h, 1 if on else 0))

    def fetch(self, index, what, which, shape, dtype):
        a = np.empty(shape, dtype)
        self._check(lib().sister_debug_fetch(self._h, index, what.encode(), which, a.ctypes.data, a.nbytes))
        return a


class SisterMultiviewDisparities:
    """Mirror of the reference helper class (hpp:18-26).

    SisterMultiviewDisparities(center, right, top, left, bottom)
    compute_disparities(dispCount) -> (disp_multiview, disp_horizontal, disp_vertical)

    Inputs are 8-bit gray or 3-channel BGR arrays (converted like
    cvtColor(BGR2GRAY), hpp:29-33).  Outputs are uint16 W x H maps holding
    min(255*d, 65535) (hpp:111-118).  The reference returns through Mat&
    out-parameters; Python returns the three maps.
    """

    def __init__(self, center, right, top, left, bottom, context=None):
        self.center, self.right, self.top, self.left, self.bottom = center, right, top, left, bottom
        self._ctx = context

    def compute_disparities(self, dispCount):
        ctx = self._ctx or _default_context()
        views = [bgr_to_gray(v) for v in (self.center, self.right, self.top, self.left, self.bottom)]
        r = ctx.compute_rig(views, int(dispCount), MODE_ALL, disp=False, subpix=False)
        return r["multiview"], r["horizontal"], r["vertical"]


_ctx = None


def _default_context():
    global _ctx
    if _ctx is None:
        _ctx = Context()
    return _ctx
