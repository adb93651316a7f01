"""Ensembles of realisations on structured grids in one graph-captured launch set.

GSTools generates an ensemble by calling the summator once per realisation
(one seed, one set of modes each).  ``Ensemble`` evaluates many realisations,
each with its own modes and its own grid, through one plan of the C-ABI
(gsc_ensemble_create): the inputs are uploaded once, the library builds flat
work lists over all realisations (GEMM tiles for factored ones, point blocks
for direct ones) and captures the launches in a HIP graph; ``run`` replays it.

The per-realisation choice between the direct kernel and the factored MFMA
contraction, and the GEMM tile shape of each factored realisation, come from
the library rule, explicit ``paths`` / ``tiles`` lists, or the autotuner
(``tune=True``, see gstools_core.tuner), which also picks the plan's split
format and accumulation when they are left to it.

``table_bytes`` bounds the factored tables the plan holds at once: above it
the realisations are split into batches that build and use their tables in
turn in one region (0 = the library's 4 GiB), trading a little time per
batch for workspace.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .grid import PATHS


class Ensemble:
    """Device-resident ensemble plan.

    realisations: sequence of (cov_samples (d, M_r), z1 (M_r), z2 (M_r), axes)
    with axes a list of d 1-D coordinate arrays; all realisations share d (2 or 3).
    ``Ensemble.seeded`` builds a plan whose modes are drawn on the device from
    per-realisation seeds (gstools_core.sampling) at the start of every run.
    """

    def __init__(self, realisations: Sequence, *, precision: str = "auto",
                 out_dtype=np.float32, accum: Optional[str] = None,
                 paths: Optional[Sequence[str]] = None,
                 scales: Optional[Sequence[float]] = None, tune: bool = False, tuner=None,
                 device: int = -1, blocks_per_cu: int = 0, tile: str = "auto",
                 tiles: Optional[Sequence[str]] = None, table_bytes: int = 0):
        if len(realisations) == 0:
            raise ValueError("an ensemble needs at least one realisation")
        covs, z1s, z2s, grids, ms = [], [], [], [], []
        for cov, z1, z2, axes in realisations:
            cov = np.ascontiguousarray(cov, dtype=np.float64)
            if cov.ndim != 2 or cov.shape[1] != len(z1) or cov.shape[1] != len(z2):
                raise ValueError("cov_samples must be (d, M) with len(z1) == len(z2) == M")
            if cov.shape[0] != len(axes):
                raise ValueError("cov_samples must have one row per axis")
            covs.append(cov.ravel())
            z1s.append(np.asarray(z1, dtype=np.float64))
            z2s.append(np.asarray(z2, dtype=np.float64))
            grids.append(axes)
            ms.append(cov.shape[1])
        args = self._setup(grids, ms, precision, out_dtype, accum, paths, scales, tune, tuner,
                           device, blocks_per_cu, tile, tiles, table_bytes)
        cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0)  # noqa: E731
        self._cov = _native.DeviceBuffer.from_array(cat(covs), device)
        self._z1 = _native.DeviceBuffer.from_array(cat(z1s), device)
        self._z2 = _native.DeviceBuffer.from_array(cat(z2s), device)
        c_shapes, c_ms, path_codes, tile_codes, scale_arr, opts = args

        def create(pc, tc):
            plan = ctypes.c_void_p()
            rc = _native.lib().gsc_ensemble_create_tuned(
                ctypes.byref(plan), self.count, self.dim, c_shapes, c_ms, pc, tc, scale_arr,
                self._cov.ptr, self._z1.ptr, self._z2.ptr, 0, None, None, self._ax.ptr,
                self._out.ptr, ctypes.byref(opts))
            _native.check(rc)
            return plan

        self._plan = self._build(create, path_codes, tile_codes)

    @classmethod
    def seeded(cls, model: str, grids: Sequence, m_counts: Sequence[int],
               len_scales: Sequence[float], seeds: Sequence[int], **kw) -> "Ensemble":
        """Plan whose realisation r draws m_counts[r] modes of `model`
        ("gaussian" or "exponential") with len_scales[r] and seeds[r] on the
        device (gsc_ensemble_create_seeded); grids[r] is its list of axes.
        Keyword arguments as for the constructor."""
        from .sampling import MODELS
        if model not in MODELS:
            raise ValueError(f"model must be one of {sorted(MODELS)}")
        n = len(grids)
        if n == 0:
            raise ValueError("an ensemble needs at least one realisation")
        if not (len(m_counts) == len(len_scales) == len(seeds) == n):
            raise ValueError("one mode count, length scale and seed per realisation")
        self = cls.__new__(cls)
        self._cov = self._z1 = self._z2 = None
        c_shapes, c_ms, path_codes, tile_codes, scale_arr, opts = self._setup(
            grids, [int(m) for m in m_counts], kw.pop("precision", "auto"),
            kw.pop("out_dtype", np.float32), kw.pop("accum", None), kw.pop("paths", None),
            kw.pop("scales", None), kw.pop("tune", False), kw.pop("tuner", None),
            kw.pop("device", -1), kw.pop("blocks_per_cu", 0), kw.pop("tile", "auto"),
            kw.pop("tiles", None), kw.pop("table_bytes", 0))
        if kw:
            raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
        c_len = (ctypes.c_double * n)(*[float(v) for v in len_scales])
        c_seeds = (ctypes.c_uint64 * n)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in seeds])

        def create(pc, tc):
            plan = ctypes.c_void_p()
            rc = _native.lib().gsc_ensemble_create_tuned(
                ctypes.byref(plan), self.count, self.dim, c_shapes, c_ms, pc, tc, scale_arr,
                None, None, None, MODELS[model], c_len, c_seeds, self._ax.ptr, self._out.ptr,
                ctypes.byref(opts))
            _native.check(rc)
            return plan

        self._plan = self._build(create, path_codes, tile_codes)
        return self

    def _build(self, create, path_codes, tile_codes):
        """The plan.  A tuned plan is checked as a whole: the per-class
        choices can lose to simpler plans once the realisations share one
        launch (a second GEMM launch for the other tile shape has its own
        tail), so the tuned plan, the same paths with one tile shape for all
        (256 or 128 rows) and the library rule's paths and tiles ("library"),
        all in the tuned split format and accumulation, are each built and
        timed (gsc_ensemble_time) and the fastest is kept; ``tuned`` records
        the times and the choice."""
        if self.tuned is None:
            return create(path_codes, tile_codes)
        tall = (ctypes.c_int32 * self.count)(*[_native.TILES["256"]] * self.count)
        tshort = (ctypes.c_int32 * self.count)(*[_native.TILES["128"]] * self.count)
        cands = (("tuned", path_codes, tile_codes), ("tuned_256", path_codes, tall),
                 ("tuned_128", path_codes, tshort), ("library", None, None))
        best, times = None, {}
        for name, pc, tc in cands:
            plan = create(pc, tc)
            ms = ctypes.c_double()
            _native.check(_native.lib().gsc_ensemble_time(plan, 5, ctypes.byref(ms)))
            times[name] = ms.value
            if best is None or ms.value < best[1]:
                if best is not None:
                    _native.lib().gsc_ensemble_destroy(best[2])
                best = (name, ms.value, plan)
            else:
                _native.lib().gsc_ensemble_destroy(plan)
        self.tuned["plan_ms"] = times
        self.tuned["kept"] = best[0]
        if best[0] != "tuned":
            self.tuned["tiles"] = ["128" if best[0] == "tuned_128" else "256"] * self.count
            if best[0] == "library":
                self.tuned["paths"] = None
        return best[2]

    def reseed(self, seeds: Sequence[int], stream: int = 0) -> None:
        """New per-realisation seeds for the following runs of a seeded plan."""
        if len(seeds) != self.count:
            raise ValueError("one seed per realisation")
        c_seeds = (ctypes.c_uint64 * self.count)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in seeds])
        _native.check(_native.lib().gsc_ensemble_reseed(self._plan, c_seeds,
                                                        ctypes.c_void_p(stream or None)))

    def _setup(self, grids, ms, precision, out_dtype, accum, paths, scales, tune, tuner, device,
               blocks_per_cu, tile="auto", tiles=None, table_bytes=0):
        """Geometry, output buffer and options shared by both constructors."""
        dims = {len(g) for g in grids}
        if len(dims) != 1:
            raise ValueError("all realisations must have the same number of axes")
        self.dim = dims.pop()
        if self.dim not in (2, 3):
            raise ValueError("ensembles support 2-D and 3-D grids")
        self.count = len(grids)
        self.out_dtype = np.dtype(out_dtype)
        if self.out_dtype not in (np.float32, np.float64):
            raise ValueError("out_dtype must be float32 or float64")
        axs = [np.asarray(a, dtype=np.float64) for g in grids for a in g]
        shapes = [[len(a) for a in g] for g in grids]
        self.shapes = [tuple(s) for s in shapes]
        self.m_counts = list(ms)
        self.tuned = None
        if tune:
            from .tuner import Tuner
            tuner = tuner or Tuner()
            precision, accum, paths, tiles = tuner.plan(self.dim, shapes, ms, precision, accum)
            self.tuned = {"precision": precision, "accum": accum, "paths": list(paths),
                          "tiles": list(tiles)}
        accum = accum or "f32x2"
        path_codes = None
        if paths is not None:
            if len(paths) != self.count:
                raise ValueError("one path per realisation")
            path_codes = (ctypes.c_int32 * self.count)(*[PATHS[p] for p in paths])
        tile_codes = None
        if tiles is not None:
            if len(tiles) != self.count:
                raise ValueError("one tile shape per realisation")
            tile_codes = (ctypes.c_int32 * self.count)(*[_native.TILES[str(t)] for t in tiles])
        scale_arr = None
        if scales is not None:
            if len(scales) != self.count:
                raise ValueError("one scale per realisation")
            scale_arr = (ctypes.c_double * self.count)(*[float(s) for s in scales])
        self._ax = _native.DeviceBuffer.from_array(
            np.concatenate(axs) if axs else np.zeros(0), device)
        self.n_points = [int(np.prod(s)) for s in shapes]
        self._out = _native.DeviceBuffer(max(1, sum(self.n_points)) * self.out_dtype.itemsize,
                                         device)
        opts = _native.make_opts(
            out_dtype=_native.GSC_F32 if self.out_dtype == np.float32 else _native.GSC_F64,
            precision=_native.PRECISIONS[precision], accum=_native.ACCUM[accum], device=device,
            tile=_native.TILES[str(tile)])
        opts.blocks_per_cu = int(blocks_per_cu)
        if int(table_bytes) < 0:
            raise ValueError("table_bytes must be >= 0")
        opts.table_bytes = int(table_bytes)
        c_shapes = (ctypes.c_int64 * (self.count * self.dim))(*[v for s in shapes for v in s])
        c_ms = (ctypes.c_int64 * self.count)(*ms)
        return c_shapes, c_ms, path_codes, tile_codes, scale_arr, opts

    @property
    def info(self) -> dict:
        nf, nd = ctypes.c_int64(), ctypes.c_int64()
        wb, pm = ctypes.c_size_t(), ctypes.c_double()
        _native.check(_native.lib().gsc_ensemble_info(self._plan, ctypes.byref(nf),
                                                      ctypes.byref(nd), ctypes.byref(wb),
                                                      ctypes.byref(pm)))
        return {"factored": nf.value, "direct": nd.value, "workspace_bytes": wb.value,
                "point_modes": pm.value}

    def run(self, stream: int = 0) -> None:
        """Replays the captured graph on `stream` (0: null stream); does not synchronise."""
        _native.check(_native.lib().gsc_ensemble_run(self._plan, ctypes.c_void_p(stream or None)))

    def time(self, repeats: int = 10) -> float:
        """Mean ms of one replay (HIP events, private stream, after a warm-up)."""
        ms = ctypes.c_double()
        _native.check(_native.lib().gsc_ensemble_time(self._plan, int(repeats),
                                                      ctypes.byref(ms)))
        return ms.value

    @property
    def out_ptr(self) -> int:
        return self._out.ptr

    def results(self) -> List[np.ndarray]:
        """Synchronises and copies every realisation's grid to the host."""
        _native.synchronize()
        flat = self._out.to_array(self.out_dtype, max(1, sum(self.n_points)))
        res, off = [], 0
        for s, n in zip(self.shapes, self.n_points):
            res.append(flat[off:off + n].reshape(s))
            off += n
        return res

    def close(self) -> None:
        if getattr(self, "_plan", None):
            _native.lib().gsc_ensemble_destroy(self._plan)
            self._plan = None
        for b in ("_cov", "_z1", "_z2", "_ax", "_out"):
            buf = getattr(self, b, None)
            if buf is not None:
                buf.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
