"""Randomization-method field generators: the caller side of the summators.

GSTools' ``RandMeth`` / ``IncomprRandMeth`` (the callers of ``summate`` and
``summate_incompr``, referenced at gen_benchmark_inputs.py:7, 17-29 and
compare_cython_rust.py:9-18; GSTools itself is not part of the reference)
draw the spectral modes, call the summator and finish the field on the host:

  scalar:         u(x) = sqrt(var / M) * summate(k, z1, z2, x) + nugget noise
  incompressible: u(x) = U e_1 + U sqrt(var / M) * summate_incompr(k, z1, z2, x)
                         + nugget noise          (U = mean velocity, e_1 = (1, 0, ..))

Here the modes come from the device sampler (gstools_core.sampling; Gaussian
or exponential covariance model), the scale sqrt(var / M) is applied inside
the summation kernels, and structured meshes go through the grid entry point
(the factored MFMA path) instead of being expanded to point lists.  The nugget
noise is drawn on the host from numpy's Philox generator keyed (seed, stream +
1), so a generator with a fixed seed is reproducible.

Covariance models, as GSTools parametrises them (len_scale l, var s^2):
  "gaussian"     C(r) = s^2 exp(-pi r^2 / (4 l^2))  (GSTools' Gaussian with its
                 default rescale sqrt(pi) / 2): k ~ N(0, (pi / 2) l^-2 I), i.e.
                 the sampler's Gaussian law with length l sqrt(2 / pi);
  "exponential"  C(r) = s^2 exp(-r / l): k = g / (l |w|), the sampler's
                 multivariate t law with one degree of freedom.
rescale=False takes the sampler's laws as they are (Gaussian: C(r) =
s^2 exp(-r^2 / (2 l^2)), the law the benchmark workloads state).

What is not GSTools': the random streams (GSTools seeds numpy RandomState its
own way), and the covariance models (only these two).  The distributions, the
scaling and the e_1 term are the same.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import grid as _grid
from .sampling import MODELS, sample_modes


class RandMeth:
    """Scalar random field generator.

    model: "gaussian" or "exponential"; dim: 1..3; var, len_scale, nugget: the
    covariance model's variance, length scale and nugget; mode_no: number of
    modes M; seed: Philox key (None: drawn from numpy's entropy); stream:
    second key word (independent streams for one seed); rescale: GSTools'
    length-scale convention of the model (see the module docstring).
    """

    def __init__(self, model: str = "gaussian", dim: int = 2, *, var: float = 1.0,
                 len_scale: float = 1.0, nugget: float = 0.0, mode_no: int = 1000,
                 seed: Optional[int] = None, stream: int = 0, device: int = -1,
                 rescale: bool = True):
        if model not in MODELS:
            raise ValueError(f"model must be one of {sorted(MODELS)}")
        if dim < 1 or dim > 3:
            raise ValueError("dim must be 1, 2 or 3")
        if var < 0 or nugget < 0:
            raise ValueError("var and nugget must be non-negative")
        if not len_scale > 0:
            raise ValueError("len_scale must be positive")
        if mode_no < 1:
            raise ValueError("mode_no must be positive")
        self.model, self.dim = model, int(dim)
        self.var, self.len_scale, self.nugget = float(var), float(len_scale), float(nugget)
        self.mode_no = int(mode_no)
        self.stream = int(stream)
        self.device = device
        self.rescale = bool(rescale)
        self.reset_seed(seed)

    @property
    def sampler_len_scale(self) -> float:
        """Length scale handed to the sampler: l sqrt(2 / pi) for GSTools'
        Gaussian model (k ~ N(0, (pi / 2) l^-2 I)), else l."""
        if self.rescale and self.model == "gaussian":
            return self.len_scale * float(np.sqrt(2.0 / np.pi))
        return self.len_scale

    def reset_seed(self, seed: Optional[int] = None) -> None:
        """Draws new modes for `seed` (None: a fresh random seed)."""
        if seed is None:
            seed = int(np.random.SeedSequence().generate_state(1, np.uint64)[0])
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.cov_samples, self.z_1, self.z_2 = sample_modes(
            self.model, self.dim, self.mode_no, self.sampler_len_scale, self.seed, self.stream,
            self.device)

    @property
    def scale(self) -> float:
        """sqrt(var / M), the factor between the summator's sum and the field."""
        return float(np.sqrt(self.var / self.mode_no))

    def get_nugget(self, shape) -> np.ndarray | float:
        """sqrt(nugget) * N(0, 1) noise of `shape` (0.0 without a nugget)."""
        if self.nugget == 0.0:
            return 0.0
        key = np.array([self.seed, (self.stream + 1) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
        rng = np.random.Generator(np.random.Philox(key=key))
        return np.sqrt(self.nugget) * rng.standard_normal(size=shape)

    def _axes(self, pos) -> Sequence[np.ndarray]:
        axes = [np.asarray(a, dtype=np.float64) for a in pos]
        if len(axes) != self.dim or any(a.ndim != 1 for a in axes):
            raise ValueError(f"a structured mesh needs {self.dim} 1-D axes")
        return axes

    def _points(self, pos) -> np.ndarray:
        pos = np.asarray(pos, dtype=np.float64)
        if pos.ndim == 1 and self.dim == 1:
            pos = pos[None, :]
        if pos.ndim != 2 or pos.shape[0] != self.dim:
            raise ValueError(f"pos must be ({self.dim}, N)")
        return pos

    def __call__(self, pos, mesh_type: str = "unstructured", add_nugget: bool = True,
                 **summate_kw) -> np.ndarray:
        """The field at the points pos ((dim, N), mesh_type "unstructured") or on
        the mesh axes[0] x .. x axes[dim-1] (a sequence of 1-D axes,
        mesh_type "structured", result shaped like the mesh)."""
        from . import summate
        if mesh_type == "structured":
            field = _grid.summate_grid(self.cov_samples, self.z_1, self.z_2, self._axes(pos),
                                       scale=self.scale, device=self.device, **summate_kw)
        elif mesh_type == "unstructured":
            field = summate(self.cov_samples, self.z_1, self.z_2, self._points(pos),
                            scale=self.scale, device=self.device, **summate_kw)
        else:
            raise ValueError("mesh_type must be 'structured' or 'unstructured'")
        if add_nugget:
            field = field + self.get_nugget(field.shape)
        return field


class IncomprRandMeth(RandMeth):
    """Incompressible (divergence-free) vector field generator, dim 2 or 3,
    with mean velocity U along the first axis."""

    def __init__(self, model: str = "gaussian", dim: int = 2, *, mean_velocity: float = 1.0,
                 **kw):
        if dim not in (2, 3):
            raise ValueError("Only 2D and 3D incompressible fields can be generated.")
        super().__init__(model, dim, **kw)
        self.mean_u = float(mean_velocity)

    def __call__(self, pos, mesh_type: str = "unstructured", add_nugget: bool = True,
                 **summate_kw) -> np.ndarray:
        """(dim, N) at the points pos, or (dim, n_0, .., n_{dim-1}) on a
        structured mesh (the grid entry point: the factored path folds the
        projector into the last axis' table)."""
        from . import summate_incompr
        if mesh_type == "structured":
            field = _grid.summate_incompr_grid(self.cov_samples, self.z_1, self.z_2,
                                               self._axes(pos), scale=self.mean_u * self.scale,
                                               device=self.device, **summate_kw)
        elif mesh_type == "unstructured":
            field = summate_incompr(self.cov_samples, self.z_1, self.z_2, self._points(pos),
                                    scale=self.mean_u * self.scale, device=self.device,
                                    **summate_kw)
        else:
            raise ValueError("mesh_type must be 'structured' or 'unstructured'")
        field[0] += self.mean_u
        if add_nugget:
            field = field + self.get_nugget(field.shape)
        return field
