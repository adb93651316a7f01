"""ctypes binding of libgstools_core_amd.so (declarations in include/gstools_core_amd.h).

The library is loaded lazily on first use.  There is no CPU fallback: if the
library is missing or no gfx950 device is visible, the call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgstools_core_amd.so")

GSC_OK = 0
GSC_ERR_INVALID = 1
GSC_ERR_DIM = 2
GSC_ERR_NODEVICE = 3
GSC_ERR_HIP = 4
GSC_ERR_ALLOC = 5
GSC_ERR_WORKSPACE = 6

GSC_F64 = 0
GSC_F32 = 1

GSC_FIELD_SCALAR = 0
GSC_FIELD_INCOMPR = 1
GSC_FIELD_FOURIER = 2

GSC_PATH_AUTO = 0
GSC_PATH_DIRECT = 1
GSC_PATH_FACTORED = 2

PRECISIONS = {
    "auto": 0, "fp16x3": 1, "bf16x3": 2, "bf16x6": 3, "fp16x1": 4, "bf16x1": 5,
}


class GscOpts(ctypes.Structure):
    _fields_ = [
        ("out_dtype", ctypes.c_int32),
        ("path", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("num_threads", ctypes.c_int32),
        ("accum", ctypes.c_int32),
        ("direct_variant", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("scale", ctypes.c_double),
        ("tile", ctypes.c_int32),
        ("table_bytes", ctypes.c_int64),
    ]


_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_optp = ctypes.POINTER(GscOpts)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vpp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every function declared in the public header.
SIGNATURES = {
    "gsc_opts_init": (None, [_optp]),
    "gsc_version": (ctypes.c_char_p, []),
    "gsc_last_error": (ctypes.c_char_p, []),
    "gsc_device_count": (_int, []),
    "gsc_workspace_bytes": (_sz, [_int, _i64, _i64, _optp]),
    "gsc_summate": (_int, [_int, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                           _i64, _vp, _i64, _optp]),
    "gsc_summate_incompr": (_int, [_int, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp,
                                   _i64, _i64, _vp, _i64, _i64, _optp]),
    "gsc_summate_fourier": (_int, [_int, _i64, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp,
                                   _i64, _vp, _i64, _i64, _vp, _i64, _optp]),
    "gsc_summate_dev": (_int, [_int, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz, _optp,
                               _vp]),
    "gsc_summate_dev_noalloc": (_int, [_int, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _optp,
                                       _vp]),
    "gsc_summate_incompr_dev": (_int, [_int, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _vp,
                                       _sz, _optp, _vp]),
    "gsc_summate_fourier_dev": (_int, [_int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                       _sz, _optp, _vp]),
    "gsc_grid_path": (_int, [_int, _i64, _i64p, _optp]),
    "gsc_grid_workspace_bytes": (_sz, [_int, _i64, _i64p, _optp]),
    "gsc_summate_grid": (_int, [_int, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vpp, _i64p,
                                _i64p, _vp, _optp]),
    "gsc_summate_grid_dev": (_int, [_int, _i64, _vp, _vp, _vp, _vpp, _i64p, _vp, _vp, _sz, _optp,
                                    _vp]),
    "gsc_grid_field_workspace_bytes": (_sz, [_int, _int, _i64, _i64p, _optp]),
    "gsc_summate_grid_field": (_int, [_int, _int, _i64, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _vp,
                                      _i64, _vpp, _i64p, _i64p, _vp, _optp]),
    "gsc_summate_grid_field_dev": (_int, [_int, _int, _i64, _vp, _vp, _vp, _vp, _vpp, _i64p, _vp,
                                          _vp, _sz, _optp, _vp]),
    "gsc_tune_grid_time": (_int, [_int, _int, _i64, _i64p, _optp, _int,
                                  ctypes.POINTER(ctypes.c_double)]),
    "gsc_ensemble_create_tuned": (_int, [ctypes.POINTER(_vp), _int, _int, _i64p, _i64p,
                                         ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_double), _vp, _vp, _vp, _int,
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_uint64), _vp, _vp, _optp]),
    "gsc_ensemble_create": (_int, [ctypes.POINTER(_vp), _int, _int, _i64p, _i64p,
                                   ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                   _vp, _vp, _vp, _vp, _vp, _optp]),
    "gsc_ensemble_run": (_int, [_vp, _vp]),
    "gsc_ensemble_time": (_int, [_vp, _int, ctypes.POINTER(ctypes.c_double)]),
    "gsc_ensemble_info": (_int, [_vp, _i64p, _i64p, ctypes.POINTER(_sz),
                                 ctypes.POINTER(ctypes.c_double)]),
    "gsc_ensemble_destroy": (None, [_vp]),
    "gsc_sample_modes": (_int, [_int, _int, _i64, ctypes.c_double, ctypes.c_uint64,
                                ctypes.c_uint64, _vp, _vp, _vp, _optp]),
    "gsc_sample_modes_dev": (_int, [_int, _int, _i64, ctypes.c_double, ctypes.c_uint64,
                                    ctypes.c_uint64, _vp, _vp, _vp, _optp, _vp]),
    "gsc_ensemble_create_seeded": (_int, [ctypes.POINTER(_vp), _int, _int, _i64p, _i64p,
                                          ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_double), _int,
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_uint64), _vp, _vp, _optp]),
    "gsc_ensemble_reseed": (_int, [_vp, ctypes.POINTER(ctypes.c_uint64), _vp]),
    "gsc_krige": (_int, [_i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64,
                         _vp, _i64, _optp]),
    "gsc_krige_workspace_bytes": (_sz, [_i64, _i64, _int]),
    "gsc_krige_dev": (_int, [_i64, _i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _sz, _optp, _vp]),
    "gsc_direct_phase_times": (_int, [_int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp,
                                      _i64, _vp, _sz, _optp, _vp, _int,
                                      ctypes.POINTER(ctypes.c_double)]),
    "gsc_calibrate_sine": (_int, [_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]),
    "gsc_calibrate_local_mix": (_int, [_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    "gsc_device_alloc": (_vp, [_sz, _int]),
    "gsc_device_free": (None, [_vp]),
    "gsc_memcpy_to_device": (_int, [_vp, _vp, _sz]),
    "gsc_memcpy_to_host": (_int, [_vp, _vp, _sz]),
    "gsc_synchronize": (_int, []),
}

_lib = None
_lock = threading.Lock()


def _preload_hip_runtime() -> None:
    """Chooses the HIP runtime the library binds to.

    PyTorch-ROCm ships its own libamdhip64.so with the same SONAME
    (libamdhip64.so.7) as /opt/rocm's; the first one loaded serves every later
    request for that SONAME in the process.  GSC_HIP_RUNTIME:
      auto (default): when torch is already imported its runtime is loaded,
              and this library binds to it; otherwise the system runtime
              (/opt/rocm) is used.  torch does not initialise on a runtime
              other than its own, so import torch first to use both;
      torch:  preload torch's copy even when torch is not imported yet;
      system: preload nothing.
    """
    mode = os.environ.get("GSC_HIP_RUNTIME", "auto")
    if mode != "torch":
        return
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    cand = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(cand):
        ctypes.CDLL(cand, mode=ctypes.RTLD_GLOBAL)


def hip_runtime_path() -> str:
    """File of the HIP runtime mapped into this process ("" if none)."""
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    return line.split()[-1]
    except OSError:
        pass
    return ""


def lib():
    """Returns the loaded library, loading it on first use (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"native library {LIB_PATH} is not built; run "
                    "`python gstools-core_amd/build_native.py` (hipcc, gfx950)")
            _preload_hip_runtime()
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


ACCUM = {"f64": 0, "f32x2": 1, "f32": 2}
TILES = {"auto": 0, "256": 1, "128": 2}
DIRECT_VARIANTS = {"mfma": 0, "valu": 1, "local": 2, "auto": 3}


def make_opts(out_dtype=GSC_F64, scale=1.0, device=-1, num_threads=0, path=GSC_PATH_AUTO,
              precision=0, accum=1, direct_variant=3, tile=0) -> GscOpts:
    o = GscOpts()
    lib().gsc_opts_init(ctypes.byref(o))
    o.out_dtype = out_dtype
    o.scale = float(scale)
    o.device = int(device)
    o.num_threads = int(num_threads)
    o.path = int(path)
    o.precision = int(precision)
    o.accum = int(accum)
    o.direct_variant = int(direct_variant)
    o.tile = int(tile)
    return o


class DeviceBuffer:
    """A device allocation owned by Python (gsc_device_alloc / gsc_device_free)."""

    def __init__(self, nbytes: int, device: int = -1):
        self.nbytes = int(nbytes)
        self.ptr = lib().gsc_device_alloc(self.nbytes, int(device))
        if not self.ptr:
            raise MemoryError(last_error() or "device allocation failed")

    @classmethod
    def from_array(cls, a, device: int = -1) -> "DeviceBuffer":
        import numpy as np
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes, device)
        check(lib().gsc_memcpy_to_device(buf.ptr, a.ctypes.data, a.nbytes))
        return buf

    def to_array(self, dtype, count: int, offset_bytes: int = 0):
        import numpy as np
        out = np.empty(count, dtype=dtype)
        check(lib().gsc_memcpy_to_host(out.ctypes.data, self.ptr + offset_bytes, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().gsc_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    check(lib().gsc_synchronize())


def last_error() -> str:
    msg = lib().gsc_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Maps a gsc_status to the Python exception the drop-in raises."""
    if rc == GSC_OK:
        return
    msg = last_error()
    if rc in (GSC_ERR_INVALID, GSC_ERR_DIM, GSC_ERR_WORKSPACE):
        raise ValueError(msg)
    if rc == GSC_ERR_ALLOC:
        raise MemoryError(msg)
    raise RuntimeError(msg or f"gstools_core native call failed with status {rc}")
