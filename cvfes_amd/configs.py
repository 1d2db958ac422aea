"""Named ensemble workloads: tube meshes, GMRF modulus samples, material, load and time step.

Each :class:`EnsembleCase` fixes everything a run needs so that the GPU path, the CPU oracle,
the benchmark and the tests all see the same inputs.  Sizes follow the five workloads the solid
step is measured on:

* ``tube-64x32-s16``      : 2,048-node tube (R 0.3 cm, L 3 cm), 16 samples, 1,000 steps (CPU-runnable).
* ``tube-256x400-s128``   : 102,400-node tube (R 0.5 cm, L 6 cm), 128 samples, Rayleigh beta 1e-4,
                            2,000 steps, stored block-CSR - the throughput workload.
* ``tube-192x256-s32``    : 49,152-node tube (R 0.5 cm, L 6 cm), 32 samples, 2,000 steps (adjoint).
* ``stenosed-512x1000-s256``: 512,000-node tapered tube 1.2 -> 0.8 cm over 12 cm with a 50 %
                            Gaussian stenosis at mid-length, 256 samples, beta 1e-4.
* ``batch-64``            : 64 uneven tubes (see :func:`batch_cases`).

Common material: thickness 0.07 cm, density 1.0 g/cm^3, Poisson ratio 0.5, E0 = 7.0e6 dyn/cm^2 with
per-element log-normal GMRF samples at 10 % coefficient of variation.  Load: uniform internal
pressure p(t) = 1.33e4 (1 + 0.1 sin(2 pi t / 0.8)) dyn/cm^2 applied as p(t) * f1 on the reference
geometry; both end rings fixed.

Time steps: explicit central differences are stable for ``dt^2 lam + 2 dt beta lam < 4`` with
``lam = lambda_max(M^-1 K)``.  With the CPU path's lumped mass the 64x32 tube has
``lam = 1.62e11 s^-2`` at its stiffest of 16 samples (dt_crit = 4.98e-6 s; 1.50e11-1.54e11 for
the first four), so its step is 4e-6 s instead of 1e-5 s.  The
102,400-node tube has ``lam = 1.36e12 s^-2``; with beta = 1e-4 its bound is 1.47e-8 s and it runs at
1.2e-8 s.  The 49,152-node tube (no damping) has ``lam = 6.80e11 s^-2``, bound 2.43e-6 s, step
2e-6 s.  (Largest eigenvalues from ARPACK on the assembled pencil, first samples.)  Cases with
``dt=None`` take 0.8 x the bound from a power iteration on the assembled system.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .gmrf import sample_element_E
from .mesh import stenosed_radius, tube_geometry

__all__ = ["EnsembleCase", "CASES", "get_case", "sine_pressure", "ramp_pressure", "batch_cases"]


def sine_pressure(t, mean=1.33e4, rel_amp=0.1, period=0.8):
    """``p(t) = mean (1 + rel_amp sin(2 pi t / period))``."""
    t = np.asarray(t, np.float64)
    return mean * (1.0 + rel_amp * np.sin(2.0 * math.pi * t / period))


def ramp_pressure(t, p0, ramp_T):
    """Cosine ramp to ``p0`` over ``ramp_T`` then constant (``physicsSolverSDampGPUs.py:316-322``)."""
    t = np.asarray(t, np.float64)
    a = b = p0 / 2.0
    n = math.pi / ramp_T
    return np.where(t > ramp_T, p0, a - b * np.cos(n * t))


@dataclass
class EnsembleCase:
    name: str
    ntheta: int
    nz: int
    n_smp: int
    radius: float = 0.5
    length: float = 6.0
    radius_out: Optional[float] = None      # tapered when set
    stenosis: float = 0.0                   # Gaussian narrowing at mid-length (fraction)
    seed: int = 2024
    thickness: float = 0.07
    density: float = 1.0
    nu: float = 0.5
    E0: float = 7.0e6
    cov: float = 0.10
    beta: float = 0.0
    damp: float = 0.0
    dt: Optional[float] = None
    nsteps: int = 1000
    p_mean: float = 1.33e4
    p_rel_amp: float = 0.1
    p_period: float = 0.8
    notes: str = field(default="", repr=False)

    # -- geometry and fields -------------------------------------------------------------------
    def geometry(self):
        profile = None
        if self.radius_out is not None or self.stenosis:
            r_out = self.radius if self.radius_out is None else self.radius_out
            profile = stenosed_radius(self.radius, r_out, self.length, severity=self.stenosis)
        return tube_geometry(self.ntheta, self.nz, self.radius, self.length, profile)

    def sample_E(self, elements=None):
        """Per-element modulus samples ``[E, S]`` (GMRF, seeded by the case)."""
        el = self.geometry()[1] if elements is None else elements
        return sample_element_E(el, self.n_smp, seed=self.seed, E0=self.E0, cov=self.cov)

    def pressures(self, n0, n, dt=None):
        """Pressure at ``t_k = (n0 + k) dt`` for ``k = 0..n-1``."""
        dt = self.dt if dt is None else dt
        t = (n0 + np.arange(n, dtype=np.float64)) * dt
        return sine_pressure(t, self.p_mean, self.p_rel_amp, self.p_period)

    @property
    def n_nodes(self):
        return self.ntheta * self.nz

    @property
    def n_elms(self):
        return 2 * self.ntheta * (self.nz - 1)

    @property
    def nnzb(self):
        return self.n_nodes + 2 * self.ntheta * (3 * self.nz - 2)


CASES = {
    c.name: c for c in [
        EnsembleCase("tube-64x32-s16", 64, 32, 16, radius=0.3, length=3.0, dt=4.0e-6, nsteps=1000,
                     seed=1, notes="specified dt 1e-5 exceeds dt_crit 4.98e-6"),
        EnsembleCase("tube-256x400-s128", 256, 400, 128, radius=0.5, length=6.0, beta=1.0e-4,
                     dt=1.2e-8, nsteps=2000, seed=2,
                     notes="lam 1.358e12; dt_crit(beta 1e-4) = 1.47e-8"),
        EnsembleCase("tube-192x256-s32", 192, 256, 32, radius=0.5, length=6.0, dt=2.0e-6,
                     nsteps=2000, seed=4, notes="lam 6.80e11; dt_crit = 2.43e-6"),
        EnsembleCase("stenosed-512x1000-s256", 512, 1000, 256, radius=1.2, radius_out=0.8,
                     length=12.0, stenosis=0.5, beta=1.0e-4, nsteps=2000, seed=5),
    ]
}


def get_case(name: str) -> EnsembleCase:
    return CASES[name]


def batch_cases(n_mesh=64, seed=3, byte_budget=200e9, beta=0.0):
    """Uneven ensemble of tubes: ``ntheta ~ U[32, 512]``, ``nz ~ U[64, 600]``, radius ``U[0.2, 1.2]``,
    length 10 x radius, samples ``U{8..256}``; sample counts are scaled down uniformly until the
    stored block values of all meshes fit ``byte_budget``."""
    rng = np.random.default_rng(seed)
    nth = rng.integers(32, 513, n_mesh)
    nzs = rng.integers(64, 601, n_mesh)
    rad = rng.uniform(0.2, 1.2, n_mesh)
    smp = rng.integers(8, 257, n_mesh)
    nnzb = nth * nzs + 2 * nth * (3 * nzs - 2)
    total = float(np.sum(nnzb * 9 * 8 * smp))
    if total > byte_budget:
        smp = np.maximum(2, np.floor(smp * byte_budget / total).astype(int) // 2 * 2)
    return [EnsembleCase(f"batch-{i:02d}", int(nth[i]), int(nzs[i]), int(smp[i]), radius=float(rad[i]),
                         length=10.0 * float(rad[i]), seed=1000 + i, beta=beta)
            for i in range(n_mesh)]
