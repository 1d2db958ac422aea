"""Time-dependent loads of the solid step: pressure waveforms and fluid wall-traction series.

The native step takes the nodal load of step ``n`` as ``f_n = p[n] f1 + qa[n] ga + qb[n] gb``
(``cvfes_load_t``).  This module turns CVFES's load definitions into those tables:

* pressure: the cosine ramp of ``GPUSolidSolver.RefreshContext``
  (``physicsSolverSDampGPUs.py:316-322``) or a sine waveform ``mean (1 + amp sin(2 pi t / period))``;
* wall traction from fluid stress snapshots ``<name><k>.npy`` (``[nNodes, 3]``):
  ``constant`` - one field with a cosine ramp (``physicsSolverSDampGPUs.py:288-294``);
  ``series`` - linear interpolation between consecutive snapshots, in the damped solver's form
  (ramp on the first snapshot, then snapshots indexed from ``ramp_T``, ``:296-309``) or the legacy
  form (interpolation from ``t = 0`` starting at a zero field, ``physicsSolverGPUs.py:275-283``).
The traction is zeroed on fixed nodes (``physicsSolverSDampGPUs.py:313-314``; the legacy solver
overwrites those dofs with the boundary values afterwards, which has the same effect).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["PressureWaveform", "TractionSchedule"]


class PressureWaveform:
    """``ramp``: ``p0/2 (1 - cos(pi t / T))`` for ``t <= T`` else ``p0``; ``sine``: see module doc."""

    def __init__(self, kind="ramp", p0=0.0, ramp_T=0.0, mean=None, amplitude=0.0, period=1.0):
        if kind not in ("ramp", "sine", "constant"):
            raise ValueError(f"unknown pressure waveform {kind!r}")
        self.kind, self.p0, self.ramp_T = kind, float(p0), float(ramp_T)
        self.mean = self.p0 if mean is None else float(mean)
        self.amplitude, self.period = float(amplitude), float(period)

    @classmethod
    def from_config(cls, config):
        kind = getattr(config, "pressure_waveform", "ramp") or "ramp"
        p0 = getattr(config, "constant_pressure", 0.0) or 0.0
        return cls(kind, p0=p0, ramp_T=getattr(config, "constant_T", 0.0) or 0.0,
                   mean=getattr(config, "pressure_mean", p0),
                   amplitude=getattr(config, "pressure_amplitude", 0.0) or 0.0,
                   period=getattr(config, "pressure_period", 1.0) or 1.0)

    def __call__(self, t):
        t = np.asarray(t, np.float64)
        if self.kind == "sine":
            return self.mean * (1.0 + self.amplitude * np.sin(2.0 * math.pi * t / self.period))
        if self.kind == "constant":
            return np.full_like(t, self.p0)
        if self.ramp_T <= 0.0:
            return np.full_like(t, self.p0)
        a = b = self.p0 / 2.0
        n = math.pi / self.ramp_T
        return np.where(t > self.ramp_T, self.p0, a - b * np.cos(n * t))


class TractionSchedule:
    """Fluid wall-traction snapshots turned into ``(ga, qa, gb, qb)`` per time.

    ``fields(t)`` returns ``(key, ga, qa, gb, qb)``; ``key`` changes whenever a new pair of
    snapshot fields becomes active, so callers can keep the fields resident and only re-upload on
    a change.
    """

    def __init__(self, filename, constant, ramp_T, dt_f, n_nodes, boundary, legacy=False,
                 loader=np.load):
        self.filename, self.constant = filename, bool(constant)
        self.ramp_T, self.dt_f = float(ramp_T), dt_f
        self.legacy = legacy
        self.loader = loader
        self.n_nodes = n_nodes
        self.boundary = np.asarray(boundary, np.int64)
        self._cache = {}

    def _snapshot(self, k):
        if k not in self._cache:
            name = f"{self.filename}.npy" if k is None else f"{self.filename}{k}.npy"
            a = np.array(self.loader(name), dtype=np.float64).reshape(self.n_nodes, 3)
            a[self.boundary, :] = 0.0
            self._cache = {kk: v for kk, v in self._cache.items() if kk is None or k is None or abs(kk - k) <= 2}
            self._cache[k] = a
        return self._cache[k]

    def fields(self, t):
        t = float(t)
        T = self.ramp_T
        ramp = (lambda: 0.5 - 0.5 * math.cos(math.pi / T * t)) if T > 0 else (lambda: 1.0)
        if self.constant:
            q = 1.0 if t > T else ramp()
            return ("c",), self._snapshot(None), q, None, 0.0
        if self.legacy:
            nt = int(t / self.dt_f)
            w = (t - nt * self.dt_f) / self.dt_f
            start = np.zeros((self.n_nodes, 3)) if nt == 0 else self._snapshot(nt - 1)
            return ("s", nt), start, 1.0 - w, self._snapshot(nt), w
        if t <= T:
            return ("r",), self._snapshot(0), ramp(), None, 0.0
        nt = int((t - T) / self.dt_f)
        w = (t - T - nt * self.dt_f) / self.dt_f
        return ("s", nt), self._snapshot(nt), 1.0 - w, self._snapshot(nt + 1), w
