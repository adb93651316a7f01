"""Autotuner for structured grids: evaluation path, split format, accumulation
and GEMM tile shape, per shape class.

For a structured grid the direct kernel and the factored MFMA contraction
compute the same field (within the factored format's error); which is faster,
and which split format, accumulation and block tile of the contraction, depends
on the grid shape (padding of the GEMM tiles, table size relative to the
output) and on M.  The candidates are only those that meet 1e-5 sigma:

  direct                                   (the per-point kernel on the grid)
  factored  x  {fp16x3, bf16x6}            (bf16x3 and the single products miss the bar)
            x  {f32x2, f64}                (two-level accumulation)
               + f32 up to M = 1024        (one FP32 sum: 6.1e-6 sigma at M = 1000 on
                                            a 256^3 grid, 1.04e-5 at M = 4000)
            x  {256, 128}                  (GEMM block rows; 128 runs two blocks per CU)

A shape class is the grid with every axis rounded up to a power of two (at
least 16) and M rounded up to a multiple of 256.  Each candidate is timed on
the class by the library itself (gsc_tune_grid_time: HIP events around
`repeats` calls after a warm-up, synthetic seeded modes), and the table of
times is cached as JSON keyed by device, library version, field and class.
A grid inside a class is estimated from the class time scaled by its padded
work (the GEMM's padded rows x columns x modes, or points x modes for the
direct kernel), so the choice also follows each grid's own tile padding.

Ensembles are timed in their own setting: a single call of a small grid is
dominated by launch latency and by how many CUs its few tiles occupy, while
in a persistent plan only the throughput per tile counts.  For a plan the
tuner therefore times every candidate on a plan of k copies of the class
grid (k chosen for about 2e10 point-modes per replay, at most 64; gsc_ensemble_time),
per split format and accumulation, and picks the pair with the smallest
summed estimate, then per realisation its path and tile shape.  A tuned
``Ensemble`` then times that plan as a whole against two simpler ones (the
same paths with one tile shape for all, and the library rule) and keeps the
fastest: per-class times do not see that a second tile shape adds a second
persistent launch, with its own tail.

Cache file: $GSC_TUNE_CACHE, else ~/.cache/gstools_core_amd/tuning.json.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, List, Optional, Sequence, Tuple

from . import _native
from .grid import FIELDS, PATHS

_DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "gstools_core_amd",
                              "tuning.json")

FORMATS = ("fp16x3", "bf16x6")  # the split formats that meet 1e-5 sigma
ACCUMS = ("f32x2", "f64")
F32_MAX_MODES = 1024  # plain FP32 accumulation is a candidate up to this M
TILES = ("256", "128")

_TILE_ROWS = {"256": 256, "128": 128}
_TILE_COLS = {("fp16x3", "f32x2"): 128, ("bf16x6", "f32x2"): 64, ("fp16x3", "f64"): 32,
              ("bf16x6", "f64"): 32, ("fp16x3", "f32"): 128, ("bf16x6", "f32"): 64}


def _pow2(n: int) -> int:
    n = max(int(n), 16)
    return 1 << (n - 1).bit_length()


def shape_class(shape: Sequence[int], m: int):
    """(axes rounded up to powers of two (>= 16), M rounded up to 256)."""
    return tuple(_pow2(s) for s in shape), max(256, (int(m) + 255) // 256 * 256)


def config_key(path: str, precision: str = "", accum: str = "", tile: str = "") -> str:
    return "direct" if path == "direct" else f"factored/{precision}/{accum}/{tile}"


def parse_config(key: str) -> dict:
    if key == "direct":
        return {"path": "direct"}
    _, precision, accum, tile = key.split("/")
    return {"path": "factored", "precision": precision, "accum": accum, "tile": tile}


def padded_work(shape: Sequence[int], m: int, key: str) -> float:
    """Point-mode work of one evaluation as the kernels pad it: the direct
    kernel N M; the factored GEMM (rows padded to the block rows, columns to
    the block tile, modes to 32), in the orientation of a 2-D plan that pads
    less."""
    shape = [int(s) for s in shape]
    if key == "direct":
        n = 1
        for s in shape:
            n *= s
        return float(n) * max(int(m), 1)
    c = parse_config(key)
    br = _TILE_ROWS[c["tile"]]
    tc = _TILE_COLS[(c["precision"], c["accum"])]
    mp = (max(int(m), 1) + 31) // 32 * 32

    def pad(n, k):
        return (n + k - 1) // k * k

    if len(shape) == 2:
        a = pad(shape[0], br) * ((shape[1] + 127) // 128 * 128)
        b = pad(shape[1], br) * ((shape[0] + 127) // 128 * 128)
        rows, cols = (shape[1], shape[0]) if b < a else (shape[0], shape[1])
        return float(pad(rows, br)) * pad(cols, tc) * mp
    return float(shape[0]) * pad(shape[1], br) * pad(shape[2], tc) * mp


class Tuner:
    def __init__(self, path: Optional[str] = None, repeats: int = 10,
                 formats: Sequence[str] = FORMATS, accums: Sequence[str] = ACCUMS,
                 tiles: Sequence[str] = TILES):
        self.path = path or os.environ.get("GSC_TUNE_CACHE", _DEFAULT_CACHE)
        self.repeats = int(repeats)
        self.formats, self.accums, self.tiles = tuple(formats), tuple(accums), tuple(tiles)
        self.table: Dict[str, dict] = {}
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}
        self.device_key = f"gfx950|{_native.lib().gsc_version().decode()}"
        self.measured = 0  # classes measured (not found in the cache) by this tuner

    def _key(self, dim, shape, m, field) -> str:
        cls, mb = shape_class(shape, m)
        return f"{self.device_key}|{field}|{dim}|{'x'.join(map(str, cls))}|{mb}"

    def accums_for(self, m: int) -> Tuple[str, ...]:
        """Accumulations that meet 1e-5 sigma at M modes."""
        extra = ("f32",) if "f32" not in self.accums and int(m) <= F32_MAX_MODES else ()
        return self.accums + extra

    def candidates(self, dim: int, m: int) -> List[str]:
        """Configurations of a class with M (rounded) modes."""
        keys = ["direct"]
        if dim >= 2:
            keys += [config_key("factored", p, a, t) for p in self.formats
                     for a in self.accums_for(m) for t in self.tiles]
        return keys

    def time_config(self, dim: int, shape: Sequence[int], m: int, key: str,
                    field: str = "scalar") -> float:
        """ms of one call (library timing, HIP events) for this exact shape."""
        c = parse_config(key)
        opts = _native.make_opts(
            out_dtype=_native.GSC_F32, path=PATHS[c["path"]],
            precision=_native.PRECISIONS[c.get("precision", "auto")],
            accum=_native.ACCUM[c.get("accum", "f32x2")],
            direct_variant=_native.DIRECT_VARIANTS["valu"],
            tile=_native.TILES[c.get("tile", "auto")])
        c_shape = (ctypes.c_int64 * dim)(*[int(s) for s in shape])
        ms = ctypes.c_double()
        _native.check(_native.lib().gsc_tune_grid_time(FIELDS[field], dim, int(m), c_shape,
                                                       ctypes.byref(opts), self.repeats,
                                                       ctypes.byref(ms)))
        return ms.value

    def measure(self, dim: int, shape: Sequence[int], m: int, field: str = "scalar") -> dict:
        """Times every candidate on the class of (shape, m): config -> ms."""
        cls, mb = shape_class(shape, m)
        return {k: self.time_config(dim, cls, mb, k, field) for k in self.candidates(dim, mb)}

    def class_times(self, dim: int, shape: Sequence[int], m: int,
                    field: str = "scalar") -> dict:
        """The class's table of times, measured on first use and cached."""
        key = self._key(dim, shape, m, field)
        hit = self.table.get(key)
        _, mb = shape_class(shape, m)
        missing = hit is None or any(k not in hit["ms"] for k in self.candidates(dim, mb))
        if missing:
            ms = self.measure(dim, shape, m, field)
            cls, mb = shape_class(shape, m)
            hit = {"class": [list(cls), mb], "ms": ms, "best": min(ms, key=ms.get)}
            self.table[key] = hit
            self.measured += 1
            self.save()
        return hit

    def estimate(self, dim: int, shape: Sequence[int], m: int, key: str,
                 field: str = "scalar") -> float:
        """Estimated ms of `key` on this grid: the class time scaled by padded work."""
        hit = self.class_times(dim, shape, m, field)
        cls, mb = hit["class"]
        return hit["ms"][key] * padded_work(shape, m, key) / padded_work(cls, mb, key)

    def choose(self, dim: int, shape: Sequence[int], m: int, field: str = "scalar",
               precision: Optional[str] = None, accum: Optional[str] = None) -> dict:
        """Fastest configuration for this grid: {"path", "precision", "accum",
        "tile", "ms"} (precision / accum fixed when given)."""
        if dim < 2 or min(shape) == 0 or m == 0:
            return {"path": "direct", "ms": 0.0}
        best, best_ms = None, None
        for k in self.candidates(dim, shape_class(shape, m)[1]):
            c = parse_config(k)
            if c["path"] == "factored" and ((precision not in (None, "auto") and
                                             c["precision"] != precision) or
                                            (accum is not None and c["accum"] != accum)):
                continue
            t = self.estimate(dim, shape, m, k, field)
            if best_ms is None or t < best_ms:
                best, best_ms = k, t
        out = parse_config(best)
        out["ms"] = best_ms
        return out

    def ensemble_class_times(self, dim: int, shape: Sequence[int], m: int, precision: str,
                             accum: str) -> dict:
        """ms per realisation of the class grid inside a persistent plan, for
        the direct kernel and the factored path with each tile shape (one
        format and accumulation), measured on first use and cached."""
        import numpy as np

        from .ensemble import Ensemble

        key = f"ens|{self._key(dim, shape, m, 'scalar')}|{precision}|{accum}"
        hit = self.table.get(key)
        if hit is None:
            cls, mb = shape_class(shape, m)
            npm = float(np.prod(cls)) * mb
            rng = np.random.default_rng(12345)
            cov = rng.normal(0.0, 8.0 / max(cls), size=(dim, mb))
            z1, z2 = rng.normal(size=mb), rng.normal(size=mb)
            axes = [np.arange(n, dtype=np.float64) for n in cls]
            times = {}
            for cfg in ["direct"] + [config_key("factored", precision, accum, t)
                                     for t in self.tiles]:
                c = parse_config(cfg)
                # the direct kernel is ~10x slower: fewer copies keep it short
                target = 2e9 if cfg == "direct" else 2e10
                k = int(min(64, max(1, round(target / npm))))
                ens = Ensemble([(cov, z1, z2, axes)] * k, paths=[c["path"]] * k,
                               tiles=[c.get("tile", "256")] * k, precision=precision,
                               accum=accum, out_dtype=np.float32)
                try:
                    times[cfg] = ens.time(self.repeats) / k
                finally:
                    ens.close()
            hit = {"class": [list(cls), mb], "ms": times}
            self.table[key] = hit
            self.measured += 1
            self.save()
        return hit

    def plan(self, dim: int, shapes: Sequence[Sequence[int]], ms: Sequence[int],
             precision: Optional[str] = None, accum: Optional[str] = None,
             field: str = "scalar") -> Tuple[str, str, List[str], List[str]]:
        """Ensemble plan: the (format, accumulation) pair that minimises the
        summed estimated time over the realisations (one pair per plan), and
        for every realisation its path and tile shape under that pair, from
        per-class times measured inside persistent plans."""
        m_top = max([int(m) for m in ms] + [0])
        pairs = [(p, a) for p in self.formats
                 for a in self.accums_for(shape_class((16,), m_top)[1])
                 if precision in (None, "auto", p) and accum in (None, a)]
        if not pairs:
            pairs = [(precision, accum or "f32x2")]
        best = None
        for p, a in pairs:
            total, paths, tiles = 0.0, [], []
            for sh, m in zip(shapes, ms):
                if dim < 2 or min(sh) == 0 or m == 0:
                    paths.append("direct")
                    tiles.append("256")
                    continue
                hit = self.ensemble_class_times(dim, sh, m, p, a)
                cls, mb = hit["class"]
                est = {k: t * padded_work(sh, m, k) / padded_work(cls, mb, k)
                       for k, t in hit["ms"].items()}
                k = min(est, key=est.get)
                total += est[k]
                c = parse_config(k)
                paths.append(c["path"])
                tiles.append(c.get("tile", "256"))
            if best is None or total < best[0]:
                best = (total, p, a, paths, tiles)
        return best[1], best[2], best[3], best[4]

    def save(self) -> None:
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # a read-only home only loses the cache, not the decision
