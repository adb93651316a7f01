"""Device-resident entry points (the *_dev functions of the C-ABI).

Arguments are raw device addresses (ints), e.g. ``tensor.data_ptr()`` of a
torch tensor, and a HIP stream handle (``torch.cuda.current_stream().cuda_stream``
or 0 for the null stream).  Arrays must be dense row-major: cov (d, M),
z1/z2/sf (M), pos (d, N) with row pitch pos_ld, out (N) or (d, N) with row
pitch out_ld.  Work is enqueued and the call returns without synchronising.
"""
from __future__ import annotations

import ctypes

from . import _native


def workspace_bytes(dim: int, m: int, n: int, variant: str = "auto") -> int:
    """Device scratch (bytes) the *_dev summators need for `m` modes and `n`
    points in `dim` dimensions with direct kernel `variant`."""
    o = _opts(False, 1.0, variant)
    return int(_native.lib().gsc_workspace_bytes(int(dim), int(m), int(n), ctypes.byref(o)))


def _opts(out_f32, scale, variant="auto"):
    return _native.make_opts(out_dtype=_native.GSC_F32 if out_f32 else _native.GSC_F64,
                             scale=scale, direct_variant=_native.DIRECT_VARIANTS[variant])


def summate_dev(dim, m, n, cov, z1, z2, pos, pos_ld, out, work, work_bytes, *, out_f32=False,
                scale=1.0, stream=0, variant="auto"):
    o = _opts(out_f32, scale, variant)
    rc = _native.lib().gsc_summate_dev(dim, m, n, cov, z1, z2, pos, pos_ld, out, work, work_bytes,
                                       ctypes.byref(o), ctypes.c_void_p(stream or None))
    _native.check(rc)


def summate_incompr_dev(dim, m, n, cov, z1, z2, pos, pos_ld, out, out_ld, work, work_bytes, *,
                        out_f32=False, scale=1.0, stream=0, variant="auto"):
    o = _opts(out_f32, scale, variant)
    rc = _native.lib().gsc_summate_incompr_dev(dim, m, n, cov, z1, z2, pos, pos_ld, out, out_ld,
                                               work, work_bytes, ctypes.byref(o),
                                               ctypes.c_void_p(stream or None))
    _native.check(rc)


def summate_fourier_dev(dim, m, n, sf, modes, z1, z2, pos, pos_ld, out, work, work_bytes, *,
                        out_f32=False, scale=1.0, stream=0, variant="auto"):
    o = _opts(out_f32, scale, variant)
    rc = _native.lib().gsc_summate_fourier_dev(dim, m, n, sf, modes, z1, z2, pos, pos_ld, out,
                                               work, work_bytes, ctypes.byref(o),
                                               ctypes.c_void_p(stream or None))
    _native.check(rc)
