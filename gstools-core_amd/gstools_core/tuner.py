"""Autotuner: direct kernel vs factored MFMA contraction per grid shape class.

For a structured grid the two paths compute the same field (to within the
factored format's error); which is faster depends on the grid shape (padding
of the 128 x 128 GEMM tiles, table size relative to the output) and on M.
The tuner measures both on synthetic, seeded modes for a shape class (every
axis length rounded up to a power of two, M rounded up to a multiple of 256),
keeps the faster one, and caches the decision as JSON keyed by device, library
version, class, precision and accumulation.

Cache file: $GSC_TUNE_CACHE, else ~/.cache/gstools_core_amd/tuning.json.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Sequence

import numpy as np

from . import _native
from .grid import summate_grid_dev, grid_workspace_bytes

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "gstools_core_amd",
                              "tuning.json")


def _pow2(n: int) -> int:
    n = max(int(n), 16)
    return 1 << (n - 1).bit_length()


def shape_class(shape: Sequence[int], m: int):
    return tuple(_pow2(s) for s in shape), max(256, (int(m) + 255) // 256 * 256)


class Tuner:
    def __init__(self, path: str = None, repeats: int = 3):
        self.path = path or os.environ.get("GSC_TUNE_CACHE", _DEFAULT_CACHE)
        self.repeats = repeats
        self.table: Dict[str, dict] = {}
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}
        self.device_key = f"gfx950|{_native.lib().gsc_version().decode()}"

    def _key(self, dim, shape, m, precision, accum) -> str:
        cls, mb = shape_class(shape, m)
        return f"{self.device_key}|{dim}|{'x'.join(map(str, cls))}|{mb}|{precision}|{accum}"

    def measure(self, dim: int, shape: Sequence[int], m: int, precision: str = "auto",
                accum: str = "f32x2") -> dict:
        """Times both paths on the shape class of (shape, m); returns ms per path."""
        cls, mb = shape_class(shape, m)
        rng = np.random.default_rng(12345)
        cov = rng.normal(0, 0.05, size=(dim, mb))
        z1, z2 = rng.normal(size=mb), rng.normal(size=mb)
        axes = [np.arange(n, dtype=np.float64) for n in cls]
        bufs = [_native.DeviceBuffer.from_array(a) for a in (cov, z1, z2, *axes)]
        n = int(np.prod(cls))
        out = _native.DeviceBuffer(n * 4)
        res = {}
        try:
            for path in ("direct", "factored"):
                nb = grid_workspace_bytes(dim, mb, cls, path=path, precision=precision)
                work = _native.DeviceBuffer(nb)

                def step():
                    summate_grid_dev(dim, mb, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                     [b.ptr for b in bufs[3:]], cls, out.ptr, work.ptr, nb,
                                     path=path, precision=precision, out_f32=True, accum=accum)

                step()
                _native.synchronize()
                t0 = time.perf_counter()
                for _ in range(self.repeats):
                    step()
                _native.synchronize()
                res[path] = (time.perf_counter() - t0) * 1e3 / self.repeats
                work.free()
        finally:
            out.free()
            for b in bufs:
                b.free()
        return res

    def choose(self, dim: int, shape: Sequence[int], m: int, precision: str = "auto",
               accum: str = "f32x2") -> str:
        """'direct' or 'factored' for this grid, measuring its class on first use."""
        if dim < 2 or min(shape) == 0:
            return "direct"
        key = self._key(dim, shape, m, precision, accum)
        hit = self.table.get(key)
        if hit is None:
            ms = self.measure(dim, shape, m, precision, accum)
            hit = {"path": min(ms, key=ms.get), "ms": ms}
            self.table[key] = hit
            self.save()
        return hit["path"]

    def save(self) -> None:
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home only loses the cache, not the decision
