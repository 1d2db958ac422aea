"""Batched ensemble of uneven meshes advanced by one persistent launch per step.

``BatchedEnsemble`` takes a list of assembled :class:`MembraneSystem` (each with its own sample
count, operator, time step, damping and load tables), numbers their wave tasks once
(``cvfes_batch_plan``) and advances all of them together:

* ``run(n, graph=True)`` replays HIP graphs that hold ``block`` batched steps (a multiple of 3, so
  the rotating state roles come back to the start) and a device step counter increment; the
  remainder runs as eager launches;
* ``autotune()`` times candidate lane widths (two samples per lane where the count is even; one
  everywhere; two only where the rows are wave-uniform) and grid sizes (persistent grids of a few
  wavefronts per SIMD up to one wavefront per task) and keeps the fastest.  The launch uses a
  kernel instance holding only the paths the plan needs (operators, damping, lane widths), so
  the registers, and with them the waves in flight, follow the ensemble;
* ``run_per_mesh(n, method)`` is the comparison path: each mesh's own single-mesh step kernel
  launched per mesh per step (with its own operator, or the one given).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import CvfesBatchMesh, check, ptr
from .system import LoadTables, MembraneSystem

__all__ = ["BatchedEnsemble"]


class BatchedEnsemble:
    def __init__(self, systems, loads, dts, methods="matfree", betas=0.0, damps=0.0, device=None):
        self.systems = list(systems)
        n = len(self.systems)
        if n < 1:
            raise ValueError("empty ensemble")
        as_list = lambda v: list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v] * n
        self.loads = list(loads)
        self.dts = [float(x) for x in as_list(dts)]
        self.methods = as_list(methods)
        self.betas = [float(x) for x in as_list(betas)]
        self.damps = [float(x) for x in as_list(damps)]
        self.device = torch.device(device) if device is not None else self.systems[0].device
        self.lib = _lib.lib()
        self.n_waves = 0
        # lanes: 0 two samples per lane for even counts, 1 one everywhere, 2 two only where the
        # rows are wave-uniform with two (packed rows take one)
        self.vw_hint = 0
        self.cap = False          # cap the kernel at 64 registers (8 waves per SIMD, possible spills)
        self.waves_per_simd = 0   # cap on the waves per SIMD (2..7, unused LDS; 0 = none)
        # rows wider than a wavefront: visit them sample group first (all rows for the first 64
        # lanes' samples, then the next group) instead of row by row
        self.groups_first = False
        self.kinds = 0
        self.phase = 0
        self.step_index = 0
        self._graphs = {}
        self.counter = torch.zeros(1, dtype=torch.int64, device=self.device)

    @property
    def stream(self):
        return _lib.stream_handle(self.device)

    def initial_state(self, u0s=None, du0s=None):
        """Each mesh's ``u_0``, ``u_{-1}`` (zero start by default), then plan the batch."""
        for i, s in enumerate(self.systems):
            u0 = np.zeros(3 * s.n_nodes) if u0s is None else u0s[i]
            du0 = np.zeros(3 * s.n_nodes) if du0s is None else du0s[i]
            s.initial_state(u0, du0, self.loads[i], self.dts[i], beta=self.betas[i],
                            damp=self.damps[i], method=self.methods[i])
        self._plan()
        self.phase = 0
        self.step_index = 0

    def _plan(self):
        n = len(self.systems)
        arr = (CvfesBatchMesh * n)()
        for i, s in enumerate(self.systems):
            m = arr[i]
            m.sys = s._make_struct()
            m.method = MembraneSystem.method_id(self.methods[i])
            u, up, work = s.state
            m.state[0], m.state[1], m.state[2] = u.data_ptr(), work.data_ptr(), up.data_ptr()
            m.load = self.loads[i].struct(0)
            m.dt, m.beta, m.damp = self.dts[i], self.betas[i], self.damps[i]
            m.vw = self.vw_hint
        self._bufs = [(s.state[0], s.state[2], s.state[1]) for s in self.systems]   # state[0..2]
        nbytes = int(self.lib.cvfes_batch_plan_bytes(n))
        self.dev_meshes = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        n_tasks, kinds = ctypes.c_int64(0), ctypes.c_int64(0)
        check(self.lib.cvfes_batch_plan(self.stream, arr, n, self.dev_meshes.data_ptr(),
                                        ctypes.byref(n_tasks), ctypes.byref(kinds)))
        self.host_meshes = arr
        # the kernels read load-table entry (absolute step) with no bound of their own
        self.max_steps = min(int(ld.n) for ld in self.loads)
        self.n_tasks = int(n_tasks.value)
        self.kinds = int(kinds.value)
        self._drop_graphs()

    def _kinds(self):
        return (self.kinds | (32 if self.cap else 0) | (self.waves_per_simd << 6)
                | (512 if self.groups_first else 0))

    def _drop_graphs(self):
        for g in self._graphs.values():
            self.lib.cvfes_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self._drop_graphs()
        except Exception:
            pass

    # ------------------------------------------------------------------------------------------
    def _check_steps(self, n):
        """Refuse ``n`` more steps when they would run past any mesh's load tables."""
        if n < 0 or self.step_index + n > self.max_steps:
            raise ValueError(f"steps {self.step_index}..{self.step_index + n - 1} exceed the load "
                             f"tables ({self.max_steps} entries in the shortest)")

    def _eager(self, n):
        self._check_steps(n)
        for _ in range(n):
            check(self.lib.cvfes_batch_step(self.stream, self.dev_meshes.data_ptr(),
                                            len(self.systems), self.n_tasks, self.phase, None,
                                            self.step_index, self.n_waves, self._kinds()))
            self.phase = (self.phase + 1) % 3
            self.step_index += 1

    def _graph(self, block):
        key = (block, self.phase, self.n_waves, self._kinds())
        if key not in self._graphs:
            g = ctypes.c_void_p()
            check(self.lib.cvfes_batch_graph_create(self.dev_meshes.data_ptr(), len(self.systems),
                                                    self.n_tasks, self.counter.data_ptr(), block,
                                                    self.phase, self.n_waves, self._kinds(),
                                                    ctypes.byref(g)))
            self._graphs[key] = g.value
        return self._graphs[key]

    def run(self, n, graph=True, block=99):
        """Advance every mesh by ``n`` steps."""
        self._check_steps(n)
        if graph and n >= block:
            if block % 3:
                raise ValueError("graph block must be a multiple of 3")
            self.counter.fill_(self.step_index)
            g = self._graph(block)
            for _ in range(n // block):
                check(self.lib.cvfes_graph_launch(g, self.stream))
                self.step_index += block
            n -= (n // block) * block
        self._eager(n)
        self._sync_roles()

    def _sync_roles(self):
        p = self.phase
        for s, bufs in zip(self.systems, self._bufs):
            s.state = [bufs[p], bufs[(p + 2) % 3], bufs[(p + 1) % 3]]

    def run_per_mesh(self, n, method=None):
        """Comparison path: every mesh's own single-mesh kernel, one launch per mesh per step
        (``method`` overrides the meshes' operators, e.g. ``"stored"`` for the plain step)."""
        for k in range(n):
            for i, s in enumerate(self.systems):
                s.step(self.loads[i], self.dts[i], beta=self.betas[i], damp=self.damps[i],
                       method=method or self.methods[i], offset=self.step_index + k, nsteps=1)
        self.step_index += n
        self._plan()          # descriptors follow each mesh's current buffer roles
        self.phase = 0

    # ------------------------------------------------------------------------------------------
    def autotune(self, candidates=None, steps=6, lane_widths=(0, 1, 2), caps=(False, True),
                 occupancies=(0, 6, 5, 4)):
        """Pick the lane width, register cap and grid size (wavefronts) with the shortest step,
        then the cap on the waves per SIMD for that choice, then the wave order of rows wider
        than a wavefront (row by row or sample group first)."""
        n_grid = len(candidates) if candidates is not None else 5
        n_occ = sum(1 for o in occupancies if o != 0)
        # every timed candidate runs 1 + steps eager steps of the ensemble's load tables
        self._check_steps((len(lane_widths) * len(caps) * n_grid + n_occ + 1) * (steps + 1))
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        best = None
        saved = (self.phase, self.step_index)
        for vw in lane_widths:
            self.vw_hint = vw
            self._sync_roles()
            self._replan()
            # persistent grids of 16 and 64 wavefronts per CU, a wavefront per 16 or 4 tasks (the
            # descriptor lookup amortised over consecutive rows of one mesh), one per task
            grid = candidates if candidates is not None else [
                cus * 16, cus * 64, max(cus * 64, self.n_tasks // 16), max(cus * 64, self.n_tasks // 4),
                self.n_tasks]
            for cap in caps:
                self.cap = cap
                for w in grid:
                    self.n_waves = w
                    dt = self._time(steps)
                    if best is None or dt < best[3]:
                        best = (vw, cap, w, dt, 0)
        self.vw_hint, self.cap, self.n_waves = best[0], best[1], best[2]
        self._sync_roles()
        self._replan()
        for occ in occupancies:
            if occ == 0:
                continue
            self.waves_per_simd = occ
            dt = self._time(steps)
            if dt < best[3]:
                best = best[:3] + (dt, occ)
        self.waves_per_simd = best[4]
        if any(int(m.lanes) > 64 and not int(m.packed) for m in self.host_meshes):
            self.groups_first = True
            dt = self._time(steps)
            if dt < best[3]:
                best = best[:3] + (dt, best[4])
            else:
                self.groups_first = False
        # the tuning steps advanced the state: restart the bookkeeping from the current roles
        self._sync_roles()
        self._replan()
        names = {0: "two where even", 1: "one everywhere", 2: "two where rows are uniform"}
        self.tuned = {"lane_width": names[best[0]], "register_cap": best[1], "n_waves": best[2],
                      "waves_per_simd": best[4], "sample_groups_first": self.groups_first,
                      "ms_per_step": best[3] * 1e3, "from_step": saved[1]}
        return self.tuned

    def _time(self, steps):
        self._eager(1)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        self._eager(steps)
        torch.cuda.synchronize(self.device)
        return (time.perf_counter() - t0) / steps

    def _replan(self):
        """Re-plan from the meshes' current buffer roles (keeps the step index)."""
        step = self.step_index
        self._plan()
        self.phase = 0
        self.step_index = step
