#!/usr/bin/env python3
"""Throughput benchmark of the direct randomization summator on one MI355X.

Workload: 3-D unstructured field, N = 10,000,000 points uniform in [0,500]^3,
M = 1000 Gaussian-model modes (l = 20), seeded (gstools_core.workloads.c2),
FP64 inputs resident in HBM, FP64 output.  One step = one full summation
through the device entry point gsc_summate_dev on torch's current HIP stream:
mode-table preparation, the spatial counting sort of the points and the
cell-local direct kernel (the default variant).

Prints ONE JSON line: metric, value (point*modes/s), roofline of the
summation (transcendental-issue bound: one v_sin_f32 per point-mode), the
CPU FP64 oracle timed on a bounded sample of the same workload on this host,
the other direct variants, the structured-grid leg (256^3, factored MFMA vs
direct), the ensemble leg (512 realisations), the other configurations' legs
(2-D C1, incompressible, Fourier method, heavy-tailed 512^3 grid, kriging)
and, per configuration, the CPU oracle on a bounded sample beside the GPU
rate of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gstools-core_amd"))
sys.path.insert(0, ROOT)

METRIC = "point·modes/s: direct vs factored-MFMA summator at ≤1e-5·σ err, 1 GPU"
# max|err|/sigma of every timed grid format against the FP64 oracle on the
# seeded 10^6-point subsets, as printed by tests/test_gpu_grid.py (the bench
# itself does not evaluate the oracle); the bar is 1e-5 sigma
GRID_ERRORS = {
    "structured_grid": {"factored_fp16x3_f32x2": 2.201e-06, "factored_fp16x3_f64": 2.201e-06,
                        "factored_fp16x3_f32": 6.077e-06, "factored_bf16x6_f32x2": 3.036e-06,
                        "factored_fp16x1_f32": 1.253e-03, "direct": 7.951e-07},
    "heavy_tail_grid": {"factored_fp16x3": 1.735e-06, "factored_bf16x3": 1.872e-05,
                        "factored_bf16x6": 2.486e-06, "direct": 7.464e-07},
}


def label_errors(res: dict, leg: str) -> None:
    """Attaches the measured error and whether it meets 1e-5 sigma to each timed format."""
    errs = GRID_ERRORS[leg]
    for key, r in res.items():
        if not isinstance(r, dict):
            continue
        base = key.replace("_tile128", "")
        if base in errs and errs[base] is not None:
            r["max_err_sigma"] = errs[base]
            r["meets_1e-5_sigma"] = errs[base] <= 1e-5
# gfx950 FP32 transcendental issue peak: 256 CUs x 32 lanes/clk (half rate) x 2.4 GHz
TRANS_PEAK = 256 * 32 * 2.4e9


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, n_sample: int) -> dict:
    import numpy as np

    import oracle

    md = w.modes
    pos = np.ascontiguousarray(w.pos[:, :n_sample])
    threads = oracle.max_threads()
    t0 = time.perf_counter()
    oracle.summate(md.cov_samples, md.z1, md.z2, pos, num_threads=threads)
    dt = time.perf_counter() - t0
    # one thread on a sixteenth of the sample: the per-core rate beside it
    n1 = max(1, n_sample // 16)
    t0 = time.perf_counter()
    oracle.summate(md.cov_samples, md.z1, md.z2, np.ascontiguousarray(pos[:, :n1]), num_threads=1)
    dt1 = time.perf_counter() - t0
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": n_sample * md.m / dt,
        "unit": "point·modes/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n_sample} points of the workload x {md.m} modes, FP64 oracle "
                  f"(C restatement of field.rs:53-62, OpenMP), {dt:.2f} s wall",
        "single_core": {"value": n1 * md.m / dt1,
                        "sample": f"first {n1} points x {md.m} modes, 1 thread, {dt1:.2f} s wall"},
        "host": {"cpu_model": cpu_model, "cpu_count": os.cpu_count()},
    }


def sine_leg(repeats: int = 10) -> dict:
    """Calibrated v_sin_f32 issue rate of this device (gsc_calibrate_sine):
    the measured peak of the direct kernels' roofline, beside the rated one
    (256 CUs x 4 SIMDs x 64 lanes / 8 cycles per wave64 sine x 2.4 GHz)."""
    import ctypes

    from gstools_core import _native

    a, b = ctypes.c_double(), ctypes.c_double()
    _native.check(_native.lib().gsc_calibrate_sine(int(repeats), ctypes.byref(a),
                                                   ctypes.byref(b)))
    ms, mv = ctypes.c_double(), ctypes.c_double()
    _native.check(_native.lib().gsc_calibrate_local_mix(max(1, int(repeats) // 2),
                                                        ctypes.byref(ms), ctypes.byref(mv)))
    return {"sin_per_s": a.value, "sin_fma_pairs_per_s": b.value,
            "rated_sin_per_s": TRANS_PEAK, "frac_of_rated": a.value / TRANS_PEAK,
            "kernel": "k_sine_rate: 2048 blocks x 256 threads, 8 independent chains x 4096 "
                      "iterations per lane",
            "local_mix": {"scalar_pm_per_s": ms.value, "vector_pm_per_s": mv.value,
                          "kernel": "k_local_mix: per 1024 point-modes of a wave 1 x "
                                    "32x32x16 f16 MFMA + 16 sines + 16 FMAs (scalar) or 4 x "
                                    "16x16x32 bf16 MFMA + 16 sines + 48 FMAs (vector), 4 "
                                    "waves per SIMD, no memory traffic"}}


def time_steps(step, steps: int, warmup: int, stream, settle_ms: float = 30.0):
    """Per-step times (ms) of `step` on `stream`, HIP events around each step,
    after `warmup` untimed steps and further untimed steps until about
    settle_ms of GPU time have passed (the clock rises under sustained load)."""
    import torch

    for _ in range(warmup):
        step()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    done = 0.0
    for _ in range(1000):
        step()
        t1.record(stream)
        t1.synchronize()
        done = t0.elapsed_time(t1)
        if done >= settle_ms:
            break
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        step()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    return [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)], ev[0].elapsed_time(ev[-1])


def grid_leg(steps: int, warmup: int) -> dict:
    """Structured 256^3 grid, M = 1000 (workload c3): factored MFMA vs direct kernel."""
    import numpy as np
    import torch

    from gstools_core import grid, workloads

    w = workloads.c3()
    md = w.modes
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    axes = [torch.from_numpy(a).to(dev) for a in w.axes]
    shape = [len(a) for a in w.axes]
    n = w.n_points
    out = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {"workload": f"3-D grid {shape[0]}x{shape[1]}x{shape[2]}, M={md.m} Gaussian (l=16), "
                       "FP32 output", "n_points": n, "n_modes": md.m}
    for path, prec, acc, tile in (("factored", "fp16x3", "f32x2", "256"),
                                  ("factored", "fp16x3", "f32x2", "128"),
                                  ("factored", "fp16x3", "f32", "256"),
                                  ("factored", "fp16x3", "f64", "256"),
                                  ("factored", "bf16x6", "f32x2", "256"),
                                  ("factored", "bf16x6", "f32x2", "128"),
                                  ("factored", "fp16x1", "f32", "256"),
                                  ("direct", "auto", "f32x2", "256")):
        nb = grid.grid_workspace_bytes(3, md.m, shape, path=path, precision=prec, tile=tile)
        work = torch.empty(nb, dtype=torch.uint8, device=dev)

        def step():
            grid.summate_grid_dev(3, md.m, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                                  [a.data_ptr() for a in axes], shape, out.data_ptr(),
                                  work.data_ptr(), nb, path=path, precision=prec, out_f32=True,
                                  stream=stream.cuda_stream, accum=acc, tile=tile)

        st = steps if path == "factored" else max(3, steps // 4)
        per, total = time_steps(step, st, warmup, stream)
        ms = total / st
        key = (f"{path}_{prec}_{acc}" + ("_tile128" if tile == "128" else "")
               if path == "factored" else "direct")
        res[key] = {"ms_per_step": ms, "point_modes_per_s": n * md.m / (ms * 1e-3)}
        if path == "factored":
            nprod = {"fp16x3": 3, "bf16x6": 6, "fp16x1": 1}[prec]
            # 2 flops per real MAC, K = 2M -> 4 M flops per point per product
            tflops = 4.0 * n * md.m * nprod / (ms * 1e-3) / 1e12
            res[key]["mfma_tflops"] = tflops
            res[key]["mfma_frac_of_2500"] = tflops / 2500.0
        del work
    res["factored_fp16x3_over_direct"] = (res["direct"]["ms_per_step"] /
                                          res["factored_fp16x3_f32x2"]["ms_per_step"])
    label_errors(res, "structured_grid")
    return res


def c5_leg(steps: int, warmup: int) -> dict:
    """Heavy-tailed 512^3 grid, M = 4000 exponential-model modes (workload c5),
    FP32 output: the factored path per operand format vs the direct kernel on
    the grid.  Accuracy against the FP64 oracle is checked by the GPU tests
    (tests/test_gpu_grid.py), not here; each format carries the error they
    measure (bf16x3 misses the 1e-5 sigma bar and is labelled so)."""
    import numpy as np
    import torch

    from gstools_core import grid, workloads

    w = workloads.c5()
    md = w.modes
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    axes = [torch.from_numpy(a).to(dev) for a in w.axes]
    shape = [len(a) for a in w.axes]
    n = w.n_points
    out = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    res = {"workload": f"3-D grid {shape[0]}x{shape[1]}x{shape[2]}, M={md.m} exponential-model "
                       "modes (multivariate t, 1 dof, l=8), FP32 output",
           "n_points": n, "n_modes": md.m}
    for path, prec, tile in (("factored", "fp16x3", "256"), ("factored", "fp16x3", "128"),
                             ("factored", "bf16x3", "256"), ("factored", "bf16x6", "256"),
                             ("factored", "bf16x6", "128"), ("direct", "auto", "256")):
        nb = grid.grid_workspace_bytes(3, md.m, shape, path=path, precision=prec, tile=tile)
        work = torch.empty(nb, dtype=torch.uint8, device=dev)

        def step():
            grid.summate_grid_dev(3, md.m, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                                  [a.data_ptr() for a in axes], shape, out.data_ptr(),
                                  work.data_ptr(), nb, path=path, precision=prec, out_f32=True,
                                  stream=stream.cuda_stream, accum="f32x2", tile=tile)

        st = steps if path == "factored" else 2
        _, total = time_steps(step, st, warmup, stream)
        ms = total / st
        key = (f"factored_{prec}" + ("_tile128" if tile == "128" else "")
               if path == "factored" else "direct")
        res[key] = {"ms_per_step": ms, "point_modes_per_s": n * md.m / (ms * 1e-3)}
        if path == "factored":
            nprod = {"fp16x3": 3, "bf16x3": 3, "bf16x6": 6}[prec]
            tflops = 4.0 * n * md.m * nprod / (ms * 1e-3) / 1e12
            res[key]["mfma_tflops"] = tflops
            res[key]["mfma_frac_of_2500"] = tflops / 2500.0
        del work
    res["factored_fp16x3_over_direct"] = (res["direct"]["ms_per_step"] /
                                          res["factored_fp16x3"]["ms_per_step"])
    label_errors(res, "heavy_tail_grid")
    return res


def ensemble_leg(steps: int, warmup: int, plan_only: bool = False) -> dict:
    """512 seeded 2-D realisations (workload c4) as one graph-captured plan;
    plan_only (leg c4plan) stops after the library-rule plan (for counter runs)."""
    import numpy as np
    import torch

    from gstools_core import grid, workloads
    from gstools_core.ensemble import Ensemble
    from gstools_core.tuner import Tuner

    ens_w = workloads.c4()
    reals = [(w.modes.cov_samples, w.modes.z1, w.modes.z2, w.axes) for w in ens_w]
    scales = [w.modes.scale for w in ens_w]
    stream = torch.cuda.current_stream()
    # the library rule (every realisation factored, fp16x3, 256-row tiles)
    t0 = time.perf_counter()
    ens = Ensemble(reals, scales=scales, out_dtype=np.float32, precision="fp16x3")
    t_plan = time.perf_counter() - t0
    per, total = time_steps(lambda: ens.run(stream.cuda_stream), steps, warmup, stream)
    ms = total / steps
    info = ens.info
    ens.close()
    del ens
    if plan_only:
        return {"point_modes": info["point_modes"], "workspace_gb": info["workspace_bytes"] / 1e9,
                "ms_per_step": ms, "point_modes_per_s": info["point_modes"] / (ms * 1e-3)}
    # the autotuner's plan: path and tile shape per realisation, split format
    # and accumulation per plan, from per-class timings (cached)
    t0 = time.perf_counter()
    tuner = Tuner(path=os.path.join("/tmp", "gsc_tuning.json"))
    tuned = Ensemble(reals, scales=scales, out_dtype=np.float32, tune=True, tuner=tuner)
    t_tune = time.perf_counter() - t0
    _, total_t = time_steps(lambda: tuned.run(stream.cuda_stream), steps, warmup, stream)
    ms_t = total_t / steps
    tinfo = tuned.info
    choice = tuned.tuned
    tuned.close()
    del tuned
    from collections import Counter
    tuned_res = {"ms_per_step": ms_t, "point_modes_per_s": tinfo["point_modes"] / (ms_t * 1e-3),
                 "workspace_gb": tinfo["workspace_bytes"] / 1e9, "tune_and_plan_s": t_tune,
                 "classes_measured": tuner.measured, "precision": choice["precision"],
                 "accum": choice["accum"], "plan_ms": choice["plan_ms"], "kept": choice["kept"],
                 "choices": (dict(Counter(f"{p}/{t}" if p == "factored" else p
                                          for p, t in zip(choice["paths"], choice["tiles"])))
                             if choice["paths"] is not None
                             else "library rule's paths and tiles, tuned format")}
    res = {"workload": "512 2-D grids, sides uniform in {64..1024}, M=1000 Gaussian modes "
                       "(l uniform in [5,50]) per realisation, FP32 output, fp16x3",
           "point_modes": info["point_modes"], "factored": info["factored"],
           "direct": info["direct"], "workspace_gb": info["workspace_bytes"] / 1e9,
           "plan_s": t_plan, "ms_per_step": ms,
           "point_modes_per_s": info["point_modes"] / (ms * 1e-3), "tuned": tuned_res}
    # the library rule with at most 1 GiB of factored tables held at once
    # (table batches that reuse one region)
    bounded = Ensemble(reals, scales=scales, out_dtype=np.float32, precision="fp16x3",
                       table_bytes=1 << 30)
    _, total_b = time_steps(lambda: bounded.run(stream.cuda_stream), steps, warmup, stream)
    ms_b = total_b / steps
    res["bounded_tables"] = {"table_bytes": 1 << 30, "ms_per_step": ms_b,
                             "point_modes_per_s": info["point_modes"] / (ms_b * 1e-3),
                             "workspace_gb": bounded.info["workspace_bytes"] / 1e9}
    bounded.close()
    del bounded
    # the same grids with their modes drawn on the device from per-realisation
    # seeds inside the graph (sampler + tables + GEMM per replay)
    seeded = Ensemble.seeded("gaussian", [w.axes for w in ens_w], [w.modes.m for w in ens_w],
                             [w.modes.len_scale for w in ens_w],
                             [19031977 + r for r in range(len(ens_w))], scales=scales,
                             out_dtype=np.float32, precision=choice["precision"],
                             accum=choice["accum"], paths=choice["paths"],
                             tiles=choice["tiles"])
    _, total_s = time_steps(lambda: seeded.run(stream.cuda_stream), steps, warmup, stream)
    ms_s = total_s / steps
    res["seeded_on_device"] = {"ms_per_step": ms_s,
                               "point_modes_per_s": seeded.info["point_modes"] / (ms_s * 1e-3)}
    seeded.close()
    del seeded
    # the same realisations one grid call at a time (the caller-side loop)
    dev = torch.device("cuda:0")
    sub = ens_w[:64]
    bufs = []
    for w in sub:
        md = w.modes
        shape = [len(a) for a in w.axes]
        nb = grid.grid_workspace_bytes(2, md.m, shape, path="factored", precision="fp16x3")
        bufs.append((md.m, shape,
                     torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev),
                     torch.from_numpy(md.z1).to(dev), torch.from_numpy(md.z2).to(dev),
                     [torch.from_numpy(a).to(dev) for a in w.axes],
                     torch.empty(w.n_points, dtype=torch.float32, device=dev),
                     torch.empty(nb, dtype=torch.uint8, device=dev), nb, md.scale))

    def loop():
        for m, shape, cov, z1, z2, axes, out, work, nb, sc in bufs:
            grid.summate_grid_dev(2, m, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                                  [a.data_ptr() for a in axes], shape, out.data_ptr(),
                                  work.data_ptr(), nb, path="factored", precision="fp16x3",
                                  out_f32=True, scale=sc, stream=stream.cuda_stream)

    _, tl = time_steps(loop, max(2, steps // 4), 1, stream)
    loop_ms = tl / max(2, steps // 4)
    pm_sub = float(sum(w.n_points * w.modes.m for w in sub))
    res["per_realisation_loop_64"] = {"ms_per_step": loop_ms,
                                      "point_modes_per_s": pm_sub / (loop_ms * 1e-3)}
    return res


def incompr_leg(steps: int, warmup: int, n: int = 4_000_000, m: int = 1000,
                sine: dict | None = None) -> dict:
    """summate_incompr on the 3-D unstructured workload at N = 4e6, M = 1000:
    the cell-local kernel (one sine per point-mode shared by the three
    components) and the FP64-phase VALU kernel, device-resident, FP64 out.
    The local kernel's launches are also timed alone (HIP events around the
    summation kernel, gsc_direct_phase_times) against the sine peak: one
    v_sin_f32 and three FP32 FMAs per point-mode."""
    import ctypes
    import numpy as np
    import torch

    from gstools_core import device, workloads

    w = workloads.c2(n=n, m=m)
    md = w.modes
    d = md.dim
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    out = torch.empty((d, n), dtype=torch.float64, device=dev)
    nb = max(device.workspace_bytes(d, m, n, v) for v in ("local", "valu"))
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    res = {"workload": f"3-D unstructured summate_incompr, N={n} uniform in [0,500]^3, "
                       f"M={m} Gaussian modes (l=20)"}
    for vname in ("local", "valu"):
        def step(vname=vname):
            device.summate_incompr_dev(d, m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                                       pos.data_ptr(), n, out.data_ptr(), n, work.data_ptr(), nb,
                                       stream=stream.cuda_stream, variant=vname)

        _, tot = time_steps(step, steps, warmup, stream)
        ms = tot / steps
        res[f"{vname}_phase"] = {"ms_per_step": ms, "point_modes_per_s": n * m / (ms * 1e-3)}
    from gstools_core import _native

    ph = (ctypes.c_double * 3)()
    _native.check(_native.lib().gsc_direct_phase_times(
        _native.GSC_FIELD_INCOMPR, d, m, n, None, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
        pos.data_ptr(), n, out.data_ptr(), n, work.data_ptr(), nb,
        ctypes.byref(_native.make_opts()), ctypes.c_void_p(stream.cuda_stream), steps, ph))
    tsin = n * m / (ph[1] * 1e-3) / 1e12
    r = {"kernel": "k_direct_local_vec<3,double>", "sort_ms": ph[0], "kernel_ms": ph[1],
         "call_ms": ph[2], "achieved_tsin": tsin, "frac_of_rated": tsin / (TRANS_PEAK / 1e12)}
    if sine is not None:
        r["frac_of_calibrated"] = tsin / (sine["sin_per_s"] / 1e12)
        r["frac_of_mix_ceiling"] = tsin / (sine["local_mix"]["vector_pm_per_s"] / 1e12)
    res["local_phase"]["kernel"] = r
    return res


def fourier_modes(side: int = 10):
    """Modes of the Fourier-method workload: M = side^3 wave vectors on a
    regular k-grid (spacing 2 pi / 200, centred on 0), spectrum factors of a
    Gaussian spectrum (l = 20, sum of sf^2 = 1), seeded z1, z2."""
    import numpy as np

    m = side ** 3
    dk = 2.0 * np.pi / 200.0
    ax = (np.arange(side) - (side - 1) / 2.0) * dk
    kg = np.ascontiguousarray(np.stack(np.meshgrid(ax, ax, ax, indexing="ij")).reshape(3, m))
    sf = np.exp(-((kg ** 2).sum(axis=0)) * 20.0 ** 2 / 8.0)
    sf = sf / np.sqrt((sf ** 2).sum())
    rng = np.random.default_rng(20240607)
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    return sf, kg, z1, z2


def fourier_leg(steps: int, warmup: int, n: int = 10_000_000, side: int = 10) -> dict:
    """summate_fourier (field.rs:219-249) on the 3-D unstructured positions of
    the direct workload (N = 1e7 uniform in [0,500]^3) with M = side^3 modes
    on a regular k-grid (spacing 2 pi / 200, centred on 0) weighted by a
    Gaussian spectrum (l = 20, sum of sf^2 = 1): the cell-local kernel and the
    FP64-phase VALU kernel, device-resident, FP64 out."""
    import numpy as np
    import torch

    from gstools_core import device, workloads

    w = workloads.c2(n=n)
    d, m = 3, side ** 3
    sf, kg, z1, z2 = fourier_modes(side)
    dev = torch.device("cuda:0")
    t_sf = torch.from_numpy(sf).to(dev)
    t_k = torch.from_numpy(np.ascontiguousarray(kg)).to(dev)
    t_z1, t_z2 = torch.from_numpy(z1).to(dev), torch.from_numpy(z2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nb = max(device.workspace_bytes(d, m, n, v) for v in ("local", "valu"))
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    res = {"workload": f"3-D unstructured summate_fourier, N={n} uniform in [0,500]^3, "
                       f"M={m} modes on a {side}^3 k-grid (spacing 2 pi/200), Gaussian spectrum "
                       "(l=20)"}
    for vname in ("local", "valu"):
        def step(vname=vname):
            device.summate_fourier_dev(d, m, n, t_sf.data_ptr(), t_k.data_ptr(), t_z1.data_ptr(),
                                       t_z2.data_ptr(), pos.data_ptr(), n, out.data_ptr(),
                                       work.data_ptr(), nb, stream=stream.cuda_stream,
                                       variant=vname)

        _, tot = time_steps(step, steps, warmup, stream)
        ms = tot / steps
        res[f"{vname}_phase"] = {"ms_per_step": ms, "point_modes_per_s": n * m / (ms * 1e-3)}
    return res


def c1_leg(steps: int, warmup: int) -> dict:
    """The 2-D configuration the reference runs on a CPU (N = 1e5 points
    uniform in [0,100]^2, M = 1000 Gaussian modes, l = 10): one device call
    (default variant), device-resident, FP64 out."""
    import numpy as np
    import torch

    from gstools_core import device, workloads

    w = workloads.c1()
    md = w.modes
    d, m, n = md.dim, md.m, w.n_points
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nb = device.workspace_bytes(d, m, n)
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        device.summate_dev(d, m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(), pos.data_ptr(),
                           n, out.data_ptr(), work.data_ptr(), nb, stream=stream.cuda_stream)

    _, tot = time_steps(step, steps, warmup, stream)
    ms = tot / steps
    return {"workload": f"2-D unstructured summate, N={n} uniform in [0,100]^2, M={m} Gaussian "
                        "modes (l=10)", "ms_per_step": ms, "point_modes_per_s": n * m / (ms * 1e-3)}


def config_cpu_baselines(gpu: dict, threads: int) -> dict:
    """The FP64 oracle (the C restatement of the reference, OpenMP over this
    host's threads) on bounded samples of every configuration, each beside
    the GPU rate of the same workload measured in this run (gpu: config ->
    point-modes/s): C1 in full; C2i in the reference's mode-outer chunked
    order (field.rs:130-163) on its first 2.5e5 points; C3 and C5 on seeded
    random subsets of their grids; C4 on its first four realisations in full;
    the Fourier-method workload on its first 2.5e5 points.  threads: the
    OpenMP threads of the main baseline (set explicitly on every call: the
    oracle's thread count persists between calls)."""
    import numpy as np

    import oracle
    from gstools_core import workloads

    res = {}

    def timed(name, fn, pm, sample):
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        r = {"value": pm / dt, "unit": "point·modes/s", "cores": threads, "kind": "port",
             "sample": f"{sample}, {dt:.2f} s wall"}
        g = gpu.get(name)
        if g:
            r["gpu_point_modes_per_s"] = g
            r["gpu_over_cpu"] = g / r["value"]
        res[name] = r

    w = workloads.c1()
    md = w.modes
    timed("c1", lambda: oracle.summate(md.cov_samples, md.z1, md.z2, w.pos, num_threads=threads),
          w.point_modes, f"all {w.n_points} points x {md.m} modes (2-D)")
    ni = 250_000
    w = workloads.c2(n=4_000_000)
    md = w.modes
    pos = np.ascontiguousarray(w.pos[:, :ni])
    timed("c2i", lambda: oracle.summate_incompr_modechunk(md.cov_samples, md.z1, md.z2, pos,
                                                           num_threads=threads),
          float(ni) * md.m, f"summate_incompr, first {ni} points x {md.m} modes, mode-outer order")
    for name, wl, k in (("c3", workloads.c3(), 250_000), ("c5", workloads.c5(), 100_000)):
        md = wl.modes
        sub = workloads.grid_points(wl.axes, workloads.subset_indices(wl.n_points, k))
        timed(name, lambda md=md, sub=sub: oracle.summate(md.cov_samples, md.z1, md.z2, sub,
                                                          num_threads=threads),
              float(k) * md.m, f"seeded {k}-point subset of the grid x {md.m} modes")
    ens = workloads.c4(count=4)
    pts = [oracle.grid_points(r.axes) for r in ens]

    def run_c4():
        for r, p in zip(ens, pts):
            oracle.summate(r.modes.cov_samples, r.modes.z1, r.modes.z2, p, num_threads=threads)

    timed("c4", run_c4, sum(r.point_modes for r in ens),
          f"first 4 realisations in full ({sum(r.n_points for r in ens)} points x 1000 modes)")
    sf, kg, z1, z2 = fourier_modes()
    pos = np.ascontiguousarray(workloads.c2(n=ni).pos)
    timed("fourier", lambda: oracle.summate_fourier(sf, kg, z1, z2, pos, num_threads=threads),
          float(ni) * sf.size, f"summate_fourier, first {ni} points x {sf.size} modes")
    return res


def krige_leg(steps: int, warmup: int, n: int = 1000, N: int = 1_000_000) -> dict:
    """Kriging calculators (krige.rs:24-118) on device-resident inputs: n
    conditions, N targets, FP64.  field+error: 2 n^2 N flops on the FP64
    matrix cores; field alone: one read of krig_vecs (8 n N bytes)."""
    import ctypes

    import torch

    from gstools_core import _native

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(19031977)
    a = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g) / n ** 0.5
    mat = (a @ a.T).contiguous()
    vecs = torch.randn(n, N, dtype=torch.float64, device=dev, generator=g)
    cond = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    out = torch.empty(2 * N, dtype=torch.float64, device=dev)
    lib = _native.lib()
    nb = max(lib.gsc_krige_workspace_bytes(n, N, 1), lib.gsc_krige_workspace_bytes(n, N, 0))
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    opts = _native.make_opts()
    stream = torch.cuda.current_stream()
    res = {"workload": f"n={n} conditions, N={N} targets, FP64, synthetic SPD krige_mat"}
    for name, err in (("field_and_variance", True), ("field", False)):
        def step(err=err):
            _native.check(lib.gsc_krige_dev(n, N, mat.data_ptr(), vecs.data_ptr(), N,
                                            cond.data_ptr(), out.data_ptr(),
                                            out.data_ptr() + 8 * N if err else None,
                                            work.data_ptr(), nb, ctypes.byref(opts),
                                            ctypes.c_void_p(stream.cuda_stream)))

        _, tot = time_steps(step, steps, warmup, stream)
        ms = tot / steps
        r = {"ms_per_step": ms, "targets_per_s": N / (ms * 1e-3)}
        if err:
            tf = 2.0 * n * n * N / (ms * 1e-3) / 1e12
            r.update({"fp64_mfma_tflops": tf, "frac_of_78_6": tf / 78.6})
        else:
            gbs = 8.0 * n * N / (ms * 1e-3) / 1e9
            r.update({"vecs_read_gb_s": gbs, "frac_of_8000": gbs / 8000.0})
        res[name] = r
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000,
                    help="points of the workload timed on the CPU oracle (0 = skip)")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--m", type=int, default=1000)
    ap.add_argument("--no-grid", action="store_true", help="skip the structured-grid leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble leg")
    ap.add_argument("--legs", default="sine,direct,c1,incompr,fourier,grid,c5,ensemble,krige",
                    help="comma list of legs: sine (v_sin_f32 issue-rate calibration), direct, "
                         "c1 (2-D, N = 1e5), incompr, fourier, grid, c5, ensemble, c4plan (the "
                         "ensemble's plan "
                         "alone), krige, copy "
                         "(a 1 GiB device copy, the reference for HBM byte counters)")
    args = ap.parse_args()
    legs = set(args.legs.split(","))
    if args.no_grid:
        legs.discard("grid")
    if args.no_ensemble:
        legs.discard("ensemble")
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the summation would shard by points, with no exchange between GPUs)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from gstools_core import device, workloads

    if "copy" in legs:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0")  # 1 GiB
        a.fill_(1.0)
        b = torch.empty_like(a)
        per, tot = time_steps(lambda: b.copy_(a), args.steps, args.warmup,
                              torch.cuda.current_stream())
        print(json.dumps({"copy_bytes": a.numel() * 8, "ms_per_copy": tot / args.steps}),
              flush=True)
        if legs == {"copy"}:
            return 0
    sine = None
    if "sine" in legs:
        sine = sine_leg()
        log(f"sine calibration: {sine['sin_per_s'] / 1e12:.2f} T/s "
            f"({sine['frac_of_rated']:.3f} of rated)")
    if "direct" not in legs:
        res = {"sine_calibration": sine} if sine else {}
        if "c1" in legs:
            res["c1"] = c1_leg(args.steps, args.warmup)
        if "incompr" in legs:
            res["incompr"] = incompr_leg(max(3, args.steps // 2), 1, sine=sine)
        if "fourier" in legs:
            res["fourier"] = fourier_leg(max(3, args.steps // 2), 1)
        if "grid" in legs:
            res["structured_grid"] = grid_leg(args.steps, args.warmup)
        if "c5" in legs:
            res["heavy_tail_grid"] = c5_leg(max(3, args.steps // 4), 1)
        if "krige" in legs:
            res["krige"] = krige_leg(max(3, args.steps // 4), 1)
        if "ensemble" in legs:
            res["ensemble"] = ensemble_leg(max(3, args.steps // 2), 1)
        if "c4plan" in legs:
            res["ensemble_plan"] = ensemble_leg(max(3, args.steps // 2), 1, plan_only=True)
        print(json.dumps(res), flush=True)
        return 0

    log(f"generating workload c2: n={args.n} m={args.m}")
    w = workloads.c2(n=args.n, m=args.m)
    md = w.modes
    n, m, d = w.n_points, md.m, md.dim
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nb = max(device.workspace_bytes(d, m, n, v) for v in ("local", "valu", "mfma"))
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def step():
        device.summate_dev(d, m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(), pos.data_ptr(),
                           n, out.data_ptr(), work.data_ptr(), nb, stream=sh)

    import ctypes

    from gstools_core import _native

    def phase_times(repeats):
        """(sort, kernel, call) mean ms of `repeats` calls: HIP events before the
        call, between the cell sort and the summation kernel, and after it."""
        ph = (ctypes.c_double * 3)()
        _native.check(_native.lib().gsc_direct_phase_times(
            _native.GSC_FIELD_SCALAR, d, m, n, None, cov.data_ptr(), z1.data_ptr(),
            z2.data_ptr(), pos.data_ptr(), n, out.data_ptr(), n, work.data_ptr(), nb,
            ctypes.byref(_native.make_opts()), ctypes.c_void_p(sh), int(repeats), ph))
        return ph[0], ph[1], ph[2]

    # the GPU raises its clock over ~30 ms of sustained load (kernel times fall
    # by ~20 % over the first ~20 back-to-back calls): an untimed pass of
    # phase-timed calls brings it to its working clock before the warmup
    settle = phase_times(30)
    log(f"clock-settling pass: sort {settle[0]:.3f} ms, kernel {settle[1]:.3f} ms")
    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    log(f"timing {args.steps} steps")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        step()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    ms_per_step = total_ms / args.steps
    value = n * m / (ms_per_step * 1e-3)
    log(f"ms/step {ms_per_step:.3f} (min {min(per):.3f}, wall {wall * 1e3 / args.steps:.3f})")

    # roofline: one hardware sine per point-mode (amplitude-phase form), i.e.
    # N*M transcendentals per launch of the summation kernel
    # (k_direct_local32), whose launches are timed here with HIP events on
    # `stream` (gsc_direct_phase_times: an event between the cell sort and the
    # kernel); the whole device call gives the step's fraction beside it
    sort_ms, kernel_ms, call_ms = phase_times(args.steps)
    log(f"phases: sort {sort_ms:.3f} ms, kernel {kernel_ms:.3f} ms, call {call_ms:.3f} ms")
    avg_launch_s = kernel_ms * 1e-3
    achieved = n * m / avg_launch_s / 1e12
    traffic = None  # HBM bytes per step, measured by rocprofv3 --pmc (profiles/)
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            t = json.load(f)["direct_local_step_c2"]
        if n == 10_000_000 and m == 1000 and d == 3:
            traffic = t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        pass
    step_tsin = n * m / (ms_per_step * 1e-3) / 1e12
    roofline = {
        "bound": "valu-transcendental",
        "achieved": round(achieved, 4),
        "peak": round(TRANS_PEAK / 1e12, 4),
        "unit": "Tsin/s",
        "frac": round(achieved / (TRANS_PEAK / 1e12), 4),
        "traffic": traffic,
        "kernel": "k_direct_local32<3,double> (the summation kernel of gsc_summate_dev)",
        "kernel_ms": round(kernel_ms, 4),
        "work_per_launch": f"{n}*{m} point-modes, 1 v_sin_f32 + 1 v_fma_f32 each on the VALU; "
                           "phases on v_mfma_f32_32x32x16_f16 (1024 point-modes per instruction)",
        "step": {"ms": round(ms_per_step, 4), "sort_ms": round(sort_ms, 4),
                 "achieved": round(step_tsin, 4),
                 "frac": round(step_tsin / (TRANS_PEAK / 1e12), 4)},
        "traffic_note": "bytes per step (whole device call) from FETCH_SIZE x2 + WRITE_SIZE x1 "
                        "(copy-calibrated); algorithmic 320 MB",
    }
    if sine is not None:
        cal = sine["sin_per_s"] / 1e12
        roofline["calibrated_peak"] = round(cal, 4)
        roofline["frac_of_calibrated"] = round(achieved / cal, 4)
        mix = sine["local_mix"]["scalar_pm_per_s"] / 1e12
        roofline["mix_ceiling"] = round(mix, 4)
        roofline["frac_of_mix_ceiling"] = round(achieved / mix, 4)
        roofline["step"]["frac_of_calibrated"] = round(step_tsin / cal, 4)

    inc_res = None
    if "incompr" in legs:
        log("incompressible leg (N = 4e6)")
        inc_res = incompr_leg(max(3, args.steps // 2), 1, sine=sine)

    fourier_res = None
    if "fourier" in legs:
        log("fourier leg (N = 1e7, 10^3 k-grid)")
        fourier_res = fourier_leg(max(3, args.steps // 2), 1)

    grid_res = None
    if "grid" in legs:
        log("grid leg (256^3, factored vs direct)")
        grid_res = grid_leg(args.steps, args.warmup)

    c5_res = None
    if "c5" in legs:
        log("heavy-tailed grid leg (512^3, M = 4000)")
        c5_res = c5_leg(max(3, args.steps // 4), 1)

    krige_res = None
    if "krige" in legs:
        log("kriging leg (n = 1000, N = 1e6)")
        krige_res = krige_leg(max(3, args.steps // 4), 1)

    ens_res = None
    if "ensemble" in legs:
        log("ensemble leg (512 realisations)")
        ens_res = ensemble_leg(max(3, args.steps // 2), 1)

    # the FP64-phase variants of the direct kernel, for comparison
    direct_variants = {"local_phase": {"ms_per_step": ms_per_step, "point_modes_per_s": value}}
    for vname, key in (("valu", "valu_phase"), ("mfma", "mfma_phase")):
        def step_v(vname=vname):
            device.summate_dev(d, m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                               pos.data_ptr(), n, out.data_ptr(), work.data_ptr(), nb, stream=sh,
                               variant=vname)

        _, tot_v = time_steps(step_v, max(3, args.steps // 2), 1, stream)
        v_ms = tot_v / max(3, args.steps // 2)
        direct_variants[key] = {"ms_per_step": v_ms, "point_modes_per_s": n * m / (v_ms * 1e-3)}

    # the drop-in on host arrays (numpy in, numpy out): the same sum plus the
    # PCIe copies of 24 N bytes in and 8 N bytes out; wall clock per call
    import gstools_core

    host_pos = np.ascontiguousarray(w.pos)
    gstools_core.summate(md.cov_samples, md.z1, md.z2, host_pos)
    calls = 5
    t_h = time.perf_counter()
    for _ in range(calls):
        gstools_core.summate(md.cov_samples, md.z1, md.z2, host_pos)
    h_ms = (time.perf_counter() - t_h) * 1e3 / calls
    host_arrays = {"ms_per_call": h_ms, "point_modes_per_s": n * m / (h_ms * 1e-3),
                   "note": "gstools_core.summate on numpy arrays, wall clock, copies included"}
    ni = 4_000_000
    host_pos_i = np.ascontiguousarray(w.pos[:, :ni])
    gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, host_pos_i)
    t_h = time.perf_counter()
    for _ in range(calls):
        gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, host_pos_i)
    hi_ms = (time.perf_counter() - t_h) * 1e3 / calls
    host_arrays["incompr_4e6"] = {"ms_per_call": hi_ms,
                                  "point_modes_per_s": ni * m / (hi_ms * 1e-3),
                                  "note": "summate_incompr, (3, N) Fortran-order result"}

    c1_res = None
    if "c1" in legs:
        log("c1 leg (2-D, N = 1e5)")
        c1_res = c1_leg(args.steps, args.warmup)

    cpu = None
    cpu_cfg = None
    if args.cpu_sample > 0:
        log(f"cpu baseline on {args.cpu_sample} points")
        cpu = cpu_baseline(w, min(args.cpu_sample, n))
        log("cpu baselines of the other configurations")
        gpu_rates = {}
        if c1_res:
            gpu_rates["c1"] = c1_res["point_modes_per_s"]
        if inc_res:
            gpu_rates["c2i"] = inc_res["local_phase"]["point_modes_per_s"]
        if grid_res:
            gpu_rates["c3"] = grid_res["factored_fp16x3_f32x2"]["point_modes_per_s"]
        if c5_res:
            gpu_rates["c5"] = c5_res["factored_fp16x3"]["point_modes_per_s"]
        if ens_res:
            gpu_rates["c4"] = ens_res["point_modes_per_s"]
        if fourier_res:
            gpu_rates["fourier"] = fourier_res["local_phase"]["point_modes_per_s"]
        cpu_cfg = config_cpu_baselines(gpu_rates, cpu["cores"])

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "point·modes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64-cell-phase/f32-local-phase/f32-sin",
        "data": "synthetic",
        "config": {"workload": f"3-D unstructured summate, N={n} uniform in [0,500]^3, "
                               f"M={m} Gaussian modes (l=20), seed 19031977",
                   "path": "direct", "n_points": n, "n_modes": m, "dim": d},
        "roofline": roofline,
        "sine_calibration": sine,
        "cpu_baseline": cpu,
        "cpu_baselines_by_config": cpu_cfg,
        "c1": c1_res,
        "direct_variants": direct_variants,
        "host_arrays": host_arrays,
        "incompr": inc_res,
        "fourier": fourier_res,
        "structured_grid": grid_res,
        "heavy_tail_grid": c5_res,
        "krige": krige_res,
        "ensemble": ens_res,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
