"""Seeded synthetic workloads for the summators (the caller side of the path).

GSTools' RandMeth (the reference's caller, benches/gen_benchmark_inputs.py:20-29)
draws spectral samples k (d x M) from the covariance model's spectral density,
z1, z2 ~ N(0, 1), calls summate and scales by sqrt(var / M).  GSTools is not
part of this project, so the workloads below draw the distributions stated
for each benchmark configuration directly, with numpy's PCG64:

  * root seed 19031977 (benches/gen_benchmark_inputs.py:70), spawned into one
    SeedSequence child per configuration;
  * draw order per child: k (d x M), z1 (M), z2 (M), then positions;
  * Gaussian model with length scale l: k ~ N(0, l^-2 I);
  * exponential model (heavy-tailed): k = g / (l |w|), g ~ N(0, I_d),
    w ~ N(0, 1), a multivariate Student-t with one degree of freedom;
  * var = 1, so the scaled field u * sqrt(1 / M) has unit variance (sigma = 1).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

ROOT_SEED = 19031977

# name -> index of the SeedSequence child
CONFIG_INDEX = {"c1": 0, "c2": 1, "c3": 2, "c4": 3, "c5": 4}


@dataclass
class Modes:
    cov_samples: np.ndarray  # (d, M)
    z1: np.ndarray           # (M,)
    z2: np.ndarray           # (M,)
    len_scale: float

    @property
    def dim(self) -> int:
        return self.cov_samples.shape[0]

    @property
    def m(self) -> int:
        return self.cov_samples.shape[1]

    @property
    def scale(self) -> float:
        """sqrt(var / M) with var = 1."""
        return float(np.sqrt(1.0 / self.m))


@dataclass
class Workload:
    name: str
    modes: Modes
    pos: Optional[np.ndarray] = None                     # (d, N) unstructured points
    axes: Optional[List[np.ndarray]] = field(default=None)  # structured grid axes

    @property
    def n_points(self) -> int:
        if self.pos is not None:
            return self.pos.shape[1]
        return int(np.prod([len(a) for a in self.axes]))

    @property
    def point_modes(self) -> float:
        return float(self.n_points) * self.modes.m


def config_rng(name: str) -> np.random.Generator:
    children = np.random.SeedSequence(ROOT_SEED).spawn(len(CONFIG_INDEX))
    return np.random.default_rng(children[CONFIG_INDEX[name]])


def gaussian_modes(rng: np.random.Generator, dim: int, m: int, len_scale: float) -> Modes:
    k = rng.normal(0.0, 1.0 / len_scale, size=(dim, m))
    z1 = rng.normal(size=m)
    z2 = rng.normal(size=m)
    return Modes(k, z1, z2, len_scale)


def exponential_modes(rng: np.random.Generator, dim: int, m: int, len_scale: float) -> Modes:
    g = rng.normal(size=(dim, m))
    w = rng.normal(size=m)
    k = g / (len_scale * np.abs(w))
    z1 = rng.normal(size=m)
    z2 = rng.normal(size=m)
    return Modes(k, z1, z2, len_scale)


def c1(n: int = 100_000, m: int = 1000) -> Workload:
    """2-D unstructured: N points uniform in [0,100]^2, Gaussian l = 10."""
    rng = config_rng("c1")
    modes = gaussian_modes(rng, 2, m, 10.0)
    pos = rng.uniform(0.0, 100.0, size=(2, n))
    return Workload("c1_2d_unstructured", modes, pos=pos)


def c2(n: int = 10_000_000, m: int = 1000) -> Workload:
    """3-D unstructured: N points uniform in [0,500]^3, Gaussian l = 20.

    The incompressible variant of this configuration uses n = 4,000,000.
    """
    rng = config_rng("c2")
    modes = gaussian_modes(rng, 3, m, 20.0)
    pos = rng.uniform(0.0, 500.0, size=(3, n))
    return Workload("c2_3d_unstructured", modes, pos=pos)


def c3(side: int = 256, m: int = 1000) -> Workload:
    """3-D structured grid side^3, spacing 1, Gaussian l = 16."""
    rng = config_rng("c3")
    modes = gaussian_modes(rng, 3, m, 16.0)
    axes = [np.arange(side, dtype=np.float64) for _ in range(3)]
    return Workload(f"c3_grid{side}^3", modes, axes=axes)


def c4(count: int = 512, m: int = 1000) -> List[Workload]:
    """Ensemble of 2-D grids, sides uniform in {64..1024} per axis, l ~ U[5, 50].

    Per realisation child: (nx, ny, l) are drawn first, then k, z1, z2.
    """
    parent = np.random.SeedSequence(ROOT_SEED).spawn(len(CONFIG_INDEX))[CONFIG_INDEX["c4"]]
    out = []
    for i, child in enumerate(parent.spawn(count)):
        rng = np.random.default_rng(child)
        nx, ny = (int(v) for v in rng.integers(64, 1025, size=2))
        ls = float(rng.uniform(5.0, 50.0))
        modes = gaussian_modes(rng, 2, m, ls)
        axes = [np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64)]
        out.append(Workload(f"c4_real{i}", modes, axes=axes))
    return out


def c5(side: int = 512, m: int = 4000) -> Workload:
    """3-D grid side^3, exponential-model (multivariate t, 1 dof) modes, l = 8."""
    rng = config_rng("c5")
    modes = exponential_modes(rng, 3, m, 8.0)
    axes = [np.arange(side, dtype=np.float64) for _ in range(3)]
    return Workload(f"c5_grid{side}^3_heavy", modes, axes=axes)


def subset_indices(n_total: int, k: int, seed: int = 1) -> np.ndarray:
    """Seeded sorted sample of k distinct point indices out of n_total."""
    rng = np.random.default_rng(np.random.SeedSequence([ROOT_SEED, seed]))
    k = min(k, n_total)
    return np.sort(rng.choice(n_total, size=k, replace=False))


def grid_points(axes, flat_index: np.ndarray) -> np.ndarray:
    """Coordinates (d, len(flat_index)) of C-order flat grid indices."""
    shape = tuple(len(a) for a in axes)
    idx = np.unravel_index(flat_index, shape)
    return np.stack([np.asarray(a, dtype=np.float64)[i] for a, i in zip(axes, idx)])
