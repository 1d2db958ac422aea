"""Solver-class contract of CVFES (``physicsSolver.py:53-99``) and a single-process communicator.

``PhysicsSolver`` is the base class every time-step plugin derives from; ``FluidSolver`` is the
minimal fluid stand-in of the solid-only loop (it only stores the time, ``physicsSolver.py:97-99``).
``SerialComm`` provides the ``Get_size``/``Get_rank``/``Barrier`` subset of an MPI communicator the
solid solvers call when mpi4py is not used (one process per GPU, one GPU here).
"""
from __future__ import annotations

__all__ = ["PhysicsSolver", "FluidSolver", "SerialComm"]


class SerialComm:
    """One-rank communicator (``mpi4py.MPI.COMM_WORLD`` stand-in)."""

    def Get_size(self):
        return 1

    def Get_rank(self):
        return 0

    def Barrier(self):
        return None

    def bcast(self, obj, root=0):
        return obj


class PhysicsSolver:
    """One time step solver inside the time loop (``physicsSolver.py:53-84``)."""

    Dof = 3

    def __init__(self, comm, mesh, config):
        self.comm = comm
        self.size = comm.Get_size()
        self.rank = comm.Get_rank()
        self.mesh = mesh
        self.Dof = mesh.dof
        self.dt = config.dt

    def RefreshContext(self, physicSolver):
        pass

    def Solve(self, t, dt):
        pass

    def Save(self, filename, counter):
        pass

    def Steady(self):
        return False


class FluidSolver(PhysicsSolver):
    """Stores the current time for the solid solver's ``RefreshContext`` (``physicsSolver.py:87-99``)."""

    def __init__(self, comm, mesh, config):
        self.comm = comm
        self.size = comm.Get_size()
        self.rank = comm.Get_rank()
        self.mesh = mesh
        self.dt = config.dt
        self.t = getattr(config, "time", 0.0)

    def Solve(self, t, dt):
        self.t = t
