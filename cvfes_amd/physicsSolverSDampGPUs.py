"""``GPUSolidSolver`` for MI355X: drop-in for CVFES's ``physicsSolverSDampGPUs`` module.

Same class name, constructor ``(comm, mesh, config)`` and methods as
``physicsSolverSDampGPUs.py:45-496`` (``RefreshContext``, ``Solve``, ``Save``, ``Steady``,
``ApplyBoundaryCondition``, ``SyncCommNodes``, ``UnionDisplacement``), so ``solver.py``'s
``TransientSolverGPU`` can instantiate it unchanged.  The step itself differs from the PyOpenCL
kernels it replaces:

* the stiffness is the linear membrane stiffness on the reference geometry, stored as a 3x3
  block-CSR with sample-fastest values (``stiffness_method = stored``) or rebuilt per step from the
  per-element coefficients (``matfree``), instead of the matrix-free re-formation on the deformed
  geometry of ``assemble_K_P`` (``physicsSolverSDampGPUs.cl:232-391``);
* the pressure acts as ``p(t) f1`` on the reference geometry (``f1`` = unit-pressure nodal load);
* product, central-difference update, damping (``damp`` as ``physicsSolverSDampGPUs.cl:426``, plus
  optional Rayleigh ``beta``), traction and Dirichlet values are one kernel per step on the GPU,
  with no host round trip; ``Run`` advances many steps in one call;
* the lumped mass follows the CPU path (``mass_lumping = cpu``, oracle parity) unless
  ``mass_lumping = gpu`` selects the ``/3`` form of the OpenCL kernels;
* with ``stiffness_update_interval = k`` in the configuration (``config.update_interval``), the
  stiffness and the pressure load are those of the moved geometry ``X + u``, re-formed every ``k``
  steps as the reference CPU path does (``physicsSolver.py:294-324``); ``k = 1`` is the
  every-step ``K(X+u)`` of ``assemble_K_P``.  This mode is matrix-free;
* the start ``u_{-1} = u_0 - dt du_0 + dt^2/2 ddu_0`` takes ``ddu_0 = (f_0 - K u_0)/LHS`` without
  the damping force, as ``calc_ddu`` does (``physicsSolverSDampGPUs.py:232-277``);
* ``CalculateStress`` / ``SaveStress`` give the per-element stresses of
  ``SolidSolver.Save`` (``physicsSolver.py:384-424``) from a GPU kernel.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .loads import PressureWaveform, TractionSchedule
from .physicsSolver import PhysicsSolver
from .system import LoadTables, MembraneSystem
from . import vtkio

__all__ = ["GPUSolidSolver"]


def update_interval(config):
    """Re-formation interval of the updated-geometry mode, or None for the linear mode.

    The reference re-forms K(X + u) when its step counter reaches ``stiffness_update_interval``;
    an interval no smaller than the run's step count (``len(np.arange(endtime, step=dt))``,
    solver.py:48) never does, which is the linear mode (the shipped cfgs say 1000000000)."""
    interval = getattr(config, "update_interval", None)
    if interval is None:
        return None
    endtime = getattr(config, "endtime", None)
    n_run = len(np.arange(float(endtime), step=float(config.dt))) if endtime is not None else 1 << 40
    return int(interval) if int(interval) < n_run else None


class GPUSolidSolver(PhysicsSolver):
    """Ensemble membrane solid step on one MI355X (one process per GPU)."""

    damped = True          # physicsSolverSDampGPUs: mass-independent damping `damp` is applied
    legacy_traction = False

    def __init__(self, comm, mesh, config):
        PhysicsSolver.__init__(self, comm, mesh, config)
        self.InitializeSync()
        if self.InitializeGPU() < 0:
            raise SystemExit(-1)

        self.dt_f = getattr(config, "dt_f", None)
        self.stressFilename = getattr(config, "exportBdyStressFilename", None)
        self.useConstantStress = bool(getattr(config, "useConstantStress", False))
        self.constant_T = getattr(config, "constant_T", 0.0) or 0.0
        self.constantPressure = getattr(config, "constant_pressure", 0.0) or 0.0
        self.waveform = PressureWaveform.from_config(config)
        self.pressure = 0.0

        self.nSmp = int(config.nSmp)
        self.ndof = mesh.ndof
        self.nNodes = mesh.nNodes
        self.nElms = mesh.nElements
        self.lclNNodes = mesh.lclNNodes
        self.lclNodeIds = mesh.lclNodeIds
        self.lclElmNodeIds = mesh.lclElmNodeIds
        self.lclNDof = self.lclNNodes * mesh.dof
        self.lclNCommNodes = mesh.lclNCommNodes
        self.lclNCommDof = self.lclNCommNodes * mesh.dof
        self.lclNSpecialHeadDof = mesh.lclNSpecialHead * mesh.dof
        self.lclBoundary = mesh.lclBoundary

        self.damp = float(mesh.damp) if self.damped else 0.0
        self.beta = float(getattr(config, "beta", 0.0) or 0.0)
        self.method = getattr(config, "stiffness_method", "stored") or "stored"
        self.updateInterval = update_interval(config)
        self.massLumping = getattr(config, "mass_lumping", "cpu") or "cpu"
        self.kernelAutotune = bool(getattr(config, "kernel_autotune", True))
        # Solve() checks the displacement for non-finite values every this many steps (0 = never)
        self.finiteCheckEvery = int(getattr(config, "finite_check_every", 100) or 0)
        self._nSolve = 0
        self.tuned = None
        self.time0 = float(getattr(config, "time", 0.0) or 0.0)

        self.dofs = (3 * np.asarray(self.lclNodeIds)[:, None] + np.arange(3)).ravel()
        self.du = np.asarray(mesh.iniDu)[self.dofs]
        self.u = np.asarray(mesh.iniU)[self.dofs]

        self.traction = None
        if self.stressFilename is not None:
            self.traction = TractionSchedule(self.stressFilename, self.useConstantStress,
                                             self.constant_T, self.dt_f, self.lclNNodes,
                                             self.lclBoundary, legacy=self.legacy_traction)
        self.InitializeSolver()

    # ------------------------------------------------------------------------------------------
    def InitializeGPU(self):
        """Select ``cuda:<rank>`` and open the native library (``physicsSolverSDampGPUs.py:107-131``)."""
        if not torch.cuda.is_available():
            print("GPUs is not enough! Actural size: 0, need: {}".format(self.size))
            return -1
        ndevices = torch.cuda.device_count()
        if ndevices < self.size:
            print("GPUs is not enough! Actural size: {}, need: {}".format(ndevices, self.size))
            return -1
        self.device = torch.device("cuda", self.rank)
        torch.cuda.set_device(self.device)
        self.lib = _lib.lib()
        return 0

    def InitializeSync(self):
        self.bdyDofs = (3 * np.asarray(self.mesh.lclBoundary, np.int64)[:, None] + np.arange(3)).ravel()
        if self.size > 1:
            self.totalCommDofs = (3 * np.asarray(self.mesh.totalCommNodeIds)[:, None] + np.arange(3)).ravel()
            self.commDofs = (3 * np.asarray(self.mesh.commNodeIds)[:, None] + np.arange(3)).ravel()

    def InitializeSolver(self):
        """Assemble on the GPU and set ``u_0``, ``u_{-1}`` (``physicsSolverSDampGPUs.py:143-277``)."""
        mesh = self.mesh
        elmE = getattr(mesh, "elmE", None)
        self.system = MembraneSystem(
            np.asarray(mesh.nodes)[self.lclNodeIds], mesh.lclElmNodeIds, self.lclBoundary, self.nSmp,
            vthickness=mesh.vthickness[self.lclNodeIds], vE=None if elmE is not None else mesh.vE[self.lclNodeIds],
            elmE=elmE, density=mesh.density, nu=mesh.v, mass_form=self.massLumping,
            bdy_u=float(mesh.bdyU), store_values=(self.method == "stored" and self.updateInterval is None),
            device=self.device)
        self.t = self.time0
        if self.updateInterval is None and self.kernelAutotune:
            # occupancy cap of the step kernel, timed on a constant load (the displacements do not
            # depend on it); the initial state below replaces the tuning steps' state
            tune = self.system.loads(np.full(40, 1.0))
            self.tuned = self.system.autotune_occupancy(tune, self.dt, beta=self.beta, damp=self.damp,
                                                        method=self.method, steps=6)
            if self.method == "stored":
                # row visiting order: strips bring the mirrored reads of the symmetric storage
                # within L2 reach on tubular walls (the result does not depend on the order)
                self.tuned["row_order"] = self.system.autotune_row_order(
                    tune, self.dt, beta=self.beta, damp=self.damp, method=self.method, steps=6)
        load0 = self._load_tables(np.array([self.time0]))
        # calc_ddu forms ddu_0 = (P_0 - K u_0) / LHS with no damping force (physicsSolverSDampGPUs.py
        # :232-246, .cl:398-409), so damp enters only the steps; beta (an extension, 0 unless set)
        # adds beta K du_0 to the stiffness force of the start
        self.system.initial_state(self.u, self.du, load0[0][1], self.dt, beta=self.beta,
                                  damp=0.0, method=self.method)
        self.pressure = float(self.waveform(self.time0))
        self._step_load = None
        self._solve_tables = None

    # ------------------------------------------------------------------------------------------
    def _load_tables(self, times):
        """Split ``times`` into runs with one traction-field pair: ``[(n_steps, LoadTables), ...]``."""
        times = np.asarray(times, np.float64)
        p = self.waveform(times)
        if self.traction is None:
            return [(times.size, self.system.loads(p))]
        runs, start = [], 0
        key, ga, qa, gb, qb = self._resident(*self.traction.fields(times[0]))
        qas, qbs = [qa], [qb]
        for k in range(1, times.size):
            nk, nga, nqa, ngb, nqb = self._resident(*self.traction.fields(times[k]))
            if nk != key:
                runs.append((k - start, self._tables(p[start:k], ga, qas, gb, qbs)))
                start, key, ga, gb, qas, qbs = k, nk, nga, ngb, [], []
            qas.append(nqa)
            qbs.append(nqb)
        runs.append((times.size - start, self._tables(p[start:], ga, qas, gb, qbs)))
        return runs

    def _resident(self, key, ga, qa, gb, qb):
        """Keep the active snapshot pair on the device; upload only when the pair changes."""
        cache = getattr(self, "_fields", None)
        if cache is None or cache[0] != key:
            to_dev = (lambda g: None if g is None else
                      torch.as_tensor(np.ascontiguousarray(g, np.float64)).to(self.device))
            cache = (key, to_dev(ga), to_dev(gb))
            self._fields = cache
        return key, cache[1], qa, cache[2], qb

    def _tables(self, p, ga, qa, gb, qb):
        return self.system.loads(p, ga=ga, qa=np.asarray(qa), gb=gb,
                                 qb=None if gb is None else np.asarray(qb))

    def RefreshContext(self, physicSolver):
        """Pressure (and wall traction) at the fluid solver's time (``physicsSolverSDampGPUs.py:280-322``)."""
        self.t = float(physicSolver.t)
        self.pressure = float(self.waveform(self.t))

    def Solve(self, t, dt):
        """One step of all samples with the load set by the last ``RefreshContext``."""
        if self.traction is None:
            # one resident single-step table; the pressure is set by a fill on the launch stream,
            # so there is no allocation and no host-to-device copy per step
            if self._solve_tables is None:
                self._solve_tables = self.system.loads(np.zeros(1))
            tables = self._solve_tables
        else:
            tables = self._load_tables(np.array([self.t]))[0][1]
        tables.p.fill_(self.pressure)
        self._advance(tables, dt)
        self._step_load = tables
        self._nSolve += 1
        if self.finiteCheckEvery > 0 and self._nSolve % self.finiteCheckEvery == 0:
            self.CheckFinite(dt)

    def _advance(self, tables, dt):
        if self.updateInterval is None:
            self.system.step(tables, dt, beta=self.beta, damp=self.damp, method=self.method)
        else:
            self.system.step_updated(tables, dt, self.updateInterval, beta=self.beta, damp=self.damp)

    def Run(self, t0, nsteps, dt):
        """``nsteps`` steps from ``t0`` in as few native calls as the traction series allows;
        step ``n`` uses the load at ``t0 + n dt`` (no host round trip between steps)."""
        times = t0 + np.arange(int(nsteps), dtype=np.float64) * dt
        for n, tables in self._load_tables(times):
            self._advance(tables, dt)
        self.t = float(t0 + nsteps * dt)
        self.CheckFinite(dt)
        return self.system.u

    def CheckFinite(self, dt=None):
        """Raise ``FloatingPointError`` if the displacement holds a NaN or an infinity (one
        reduction and one synchronisation).  An explicit step past its stability bound grows
        without limit, so this is where a too-large ``dt`` shows."""
        if not bool(torch.isfinite(self.system.u).all()):
            step = "" if dt is None else f" with dt = {dt:g}"
            raise FloatingPointError(
                f"the displacement is not finite at t = {self.t:g}{step}: the explicit step is "
                "unstable (dt above the stability bound of the mesh, modulus and damping?)")

    # ------------------------------------------------------------------------------------------
    def GetDisplacement(self):
        """Current displacement as the reference's ``[ndof, nSmp]`` host array."""
        return self.system.u.cpu().numpy().reshape(self.lclNDof, self.nSmp)

    def CalculateStress(self):
        """Per-element stresses ``[nElms, nSmp, 5]`` (xx, yy, xy, xz, yz) of the current
        displacement on the current geometry (``physicsSolver.py:416-424``)."""
        return self.system.element_stress().permute(0, 2, 1).cpu().numpy()

    def SaveStress(self, filename, counter):
        """Write displacements and per-element stresses, as ``SolidSolver.Save``
        (``physicsSolver.py:384-398``; arrays ``displacement_###`` and ``xx_###`` ...)."""
        resU = self.UnionDisplacement(self.GetDisplacement())
        stress = self.CalculateStress()
        if self.rank == 0:
            u = resU.transpose().reshape(self.nSmp, self.mesh.nNodes, self.Dof)
            vtkio.save_solution(filename, counter, self.mesh.nodes, self.mesh.elementNodeIds, u, stress)
        self.comm.Barrier()

    def ApplyBoundaryCondition(self, quant):
        quant[self.bdyDofs, :] = self.mesh.bdyU

    def SyncCommNodes(self, quant):
        if self.size == 1:
            return
        raise NotImplementedError(
            "multi-rank domain decomposition is not part of the single-GPU solid step; shard the "
            "ensemble by samples instead (one GPUSolidSolver per GPU, no per-step exchange)")

    def Save(self, filename, counter):
        """Write ``u_000 ...`` per sample to ``<filename><counter>.vtp`` (``physicsSolverSDampGPUs.py:452-464``)."""
        resU = self.UnionDisplacement(self.GetDisplacement())
        if self.rank == 0:
            u = resU.transpose().reshape(self.nSmp, self.mesh.nNodes, self.Dof)
            vtkio.save_displacement(filename, counter, self.mesh.nodes, self.mesh.elementNodeIds, u)
        self.comm.Barrier()

    def UnionDisplacement(self, quant):
        if self.size == 1:
            resU = np.empty((self.ndof, self.nSmp))
            resU[self.dofs, :] = quant
            return resU
        raise NotImplementedError("multi-rank union is not part of the single-GPU solid step")
