"""Outer time loop of the solid-only run (``solver.py:41-158``).

``TransientSolverGPU`` reproduces CVFES's segregated loop with the fluid side reduced to the time
keeper of ``physicsSolver.FluidSolver``: per step, the fluid stand-in stores ``round(t, 4)``
(``solver.py:118``), the solid solver refreshes its load from it, then advances one step; results
are saved at ``save_result_num`` evenly spaced steps (``solver.py:108,133-136``).

``Solve(fast=True)`` instead advances the whole interval between two save points in one native call
with the load evaluated at the exact step times ``t_n = t_0 + n dt`` (no 4-decimal rounding).
"""
from __future__ import annotations

from timeit import default_timer as timer

import numpy as np

from .physicsSolver import FluidSolver

__all__ = ["Solver", "TransientSolver", "TransientSolverGPU"]


class Solver:
    def __init__(self, comm, meshes, config):
        self.comm = comm
        self.rank = comm.Get_rank()
        self.meshes = meshes
        self.saveResNum = config.saveResNum or 0
        self.saveStressFilename = config.saveStressFilename

    def Solve(self):
        pass


class TransientSolver(Solver):
    """Time-looping solver (``solver.py:58-144``) with a pluggable solid solver class."""

    solid_class = None

    def __init__(self, comm, meshes, config, verbose=False):
        Solver.__init__(self, comm, meshes, config)
        self.time = config.time
        self.dt = config.dt
        self.endtime = config.endtime
        self.restart = getattr(config, "restart", False)
        self.restartFilename = getattr(config, "restartFilename", None)
        self.restartTimestep = getattr(config, "restartTimestep", 0) or 0
        self.verbose = verbose
        self.t = np.append(np.arange(self.endtime, step=self.dt), self.endtime)
        self.nTimeSteps = len(self.t) - 1
        self.__initPhysicSolver__(comm, meshes, config)

    def __initPhysicSolver__(self, comm, meshes, config):
        self.fluidSolver = FluidSolver(comm, meshes.get("lumen"), config)
        self.solidSolver = self.solid_class(comm, meshes["wall"], config)

    def _save_steps(self):
        return set(np.linspace(0, self.nTimeSteps, self.saveResNum + 1, dtype=int).tolist()) \
            if self.saveResNum else set()

    def Solve(self, fast=False):
        if self.rank == 0 and self.verbose:
            print("Time\tResidual norm\tTime used")
        solverStart = timer()
        saveSteps = self._save_steps()
        if fast:
            self._solve_fast(saveSteps)
        else:
            for timeStep in range(self.restartTimestep, self.nTimeSteps):
                start = timer()
                dt = self.t[timeStep + 1] - self.t[timeStep]
                self.fluidSolver.Solve(round(self.t[timeStep], 4), dt)
                self.solidSolver.RefreshContext(self.fluidSolver)
                self.solidSolver.Solve(self.t[timeStep], dt)
                self.fluidSolver.RefreshContext(self.solidSolver)
                end = timer()
                if self.rank == 0 and self.verbose:
                    print("{:3.5e}\t{:10.1f} ms".format(self.t[timeStep], (end - start) * 1000.0))
                if timeStep + 1 in saveSteps and self.saveStressFilename:
                    self.solidSolver.Save(self.saveStressFilename, timeStep + 1)
                if self.fluidSolver.Steady():
                    break
        self.elapsed = timer() - solverStart
        if self.verbose:
            print("Rank: {}\t{:10.1f} ms".format(self.rank, self.elapsed * 1000.0))

    def _solve_fast(self, saveSteps):
        marks = sorted(s for s in saveSteps if s > self.restartTimestep) + [self.nTimeSteps]
        step = self.restartTimestep
        for mark in marks:
            if mark > step:
                self.solidSolver.Run(self.t[step], mark - step, self.dt)
                step = mark
            if mark in saveSteps and self.saveStressFilename:
                self.solidSolver.Save(self.saveStressFilename, mark)


class TransientSolverGPU(TransientSolver):
    """Structure solver on the GPU (``solver.py:148-158``): solid class ``GPUSolidSolver``."""

    from .physicsSolverSDampGPUs import GPUSolidSolver as solid_class
