"""``GPUSolidSolver`` for MI355X: drop-in for CVFES's legacy ``physicsSolverGPUs`` module.

The legacy PyOpenCL solver (``physicsSolverGPUs.py:45-496``) has the same contract as the damped
one but applies no damping (``physicsSolverGPUs.cl:419-432``) and interpolates the fluid wall
traction from ``t = 0`` starting at a zero field (``physicsSolverGPUs.py:275-283``).  This module
keeps that behaviour on the same native step (``damp`` is ignored; Rayleigh ``beta`` from the
configuration still applies).
"""
from .physicsSolverSDampGPUs import GPUSolidSolver as _DampedGPUSolidSolver

__all__ = ["GPUSolidSolver"]


class GPUSolidSolver(_DampedGPUSolidSolver):
    damped = False
    legacy_traction = True
