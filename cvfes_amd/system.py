"""Device-resident membrane system: structure, element blocks, mass, load and stiffness values.

``MembraneSystem`` owns every device buffer of one mesh and its sample ensemble as torch tensors
and drives the native kernels through the C ABI.  The host side does integer preprocessing only
(block structure, element/block/node index maps); all floating-point work - element matrices,
stiffness coefficients, lumped mass, unit load, block assembly, products and time steps - runs in
the HIP kernels of ``libcvfes_mi355x.so``.
"""
from __future__ import annotations

import copy
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import CvfesGeometry, CvfesLoad, CvfesSystem, CvfesUpdate, check, ptr
from .sparse import BlockMaps, bisection_order, block_structure, strip_order

__all__ = ["MembraneSystem", "LoadTables", "MASS_SCALE"]

# Lumped-mass normalisation: the CPU path's double quadrature weight gives rho*A/9*(...)
# (optimizedSolidAssemble.pyx:313-323 with physicsSolver.py:197); the GPU kernels use rho*A/3*(...)
# (physicsSolverSDampGPUs.cl:172-221).
MASS_SCALE = {"cpu": 1.0 / 9.0, "gpu": 1.0 / 3.0}


def _dev(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


class LoadTables:
    """Device load tables ``f_n = p[n] f1 + qa[n] ga + qb[n] gb`` for a run of steps."""

    def __init__(self, p, device, ga=None, qa=None, gb=None, qb=None, n_nodes=None):
        f64 = torch.float64
        self.p = p.to(device, f64).contiguous() if isinstance(p, torch.Tensor) else _dev(np.atleast_1d(np.asarray(p, np.float64)), f64, device)
        self.n = int(self.p.numel())
        self.ga = self._field(ga, device)
        self.gb = self._field(gb, device)
        self.qa = None if qa is None else _dev(np.asarray(qa, np.float64).reshape(-1), f64, device)
        self.qb = None if qb is None else _dev(np.asarray(qb, np.float64).reshape(-1), f64, device)
        for q in (self.qa, self.qb):
            if q is not None and q.numel() != self.n:
                raise ValueError("load tables must all have one entry per step")
        if (self.ga is None) != (self.qa is None) or (self.gb is None) != (self.qb is None):
            raise ValueError("a traction field needs its table and vice versa")
        # every step kernel reads g[3*row + i] for all rows: a field of another mesh would be read
        # out of bounds on the device
        for g in (self.ga, self.gb):
            if g is not None and n_nodes is not None and tuple(g.shape) != (int(n_nodes), 3):
                raise ValueError(f"traction field of shape {tuple(g.shape)}, expected ({n_nodes}, 3)")

    @staticmethod
    def _field(g, device):
        if g is None:
            return None
        if isinstance(g, torch.Tensor):
            return g.to(device, torch.float64).reshape(-1, 3).contiguous()
        return _dev(np.asarray(g, np.float64).reshape(-1, 3), torch.float64, device)

    def struct(self, offset=0):
        """``cvfes_load_t`` whose step 0 is entry ``offset`` of the tables."""
        if not 0 <= offset <= self.n:
            raise ValueError("load table offset out of range")
        o = 8 * offset
        return CvfesLoad(p=self.p.data_ptr() + o,
                         ga=ptr(self.ga), qa=None if self.qa is None else self.qa.data_ptr() + o,
                         gb=ptr(self.gb), qb=None if self.qb is None else self.qb.data_ptr() + o)


class MembraneSystem:
    """Assembled ensemble membrane system on one GPU.

    Parameters
    ----------
    nodes, elements, boundary : reference coordinates ``[N,3]``, triangles ``[E,3]``, fixed nodes.
    n_smp : number of samples ``S``.
    vthickness : ``[N,S]`` or ``[N]`` node thickness.
    vE / elmE : node modulus ``[N,S]`` or element modulus ``[E,S]`` (one of them).
    coef : optional precomputed stiffness coefficient ``[E,S]`` (overrides vE/elmE).
    density, nu, kshear : material constants (``kshear`` = 5/6 as ``physicsSolver.py:177``).
    mass_form : ``"cpu"`` (oracle parity, default) or ``"gpu"`` lumping.
    store_values : form the stored block-CSR values.
    storage : how the stored values are kept: ``"sym"`` (default) the upper blocks only,
        ``uvalues [nnzu][9][S]`` (K is symmetric: a lower block is its mirror read transposed;
        57 % of the full bytes), or ``"full"`` every block, ``values [nnzb][9][S]``.
    """

    def __init__(self, nodes, elements, boundary, n_smp, *, vthickness, vE=None, elmE=None,
                 coef=None, density=1.0, nu=0.5, kshear=5.0 / 6.0, mass_form="cpu", bdy_u=0.0,
                 store_values=True, storage="sym", device=None):
        if storage not in ("sym", "full"):
            raise ValueError("storage must be 'sym' or 'full'")
        self.storage = storage
        if not torch.cuda.is_available():
            raise RuntimeError("MembraneSystem needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.lib()
        self.device = torch.device(device if device is not None else "cuda")
        dev = self.device
        nodes = np.asarray(nodes, np.float64)
        elements = np.asarray(elements, np.int64)
        self.n_nodes = int(nodes.shape[0])
        self.n_elms = int(elements.shape[0])
        self.n_smp = int(n_smp)
        self.nu, self.kshear, self.density = float(nu), float(kshear), float(density)
        # --- integer preprocessing (host) --------------------------------------------------------
        self.indptr_h, self.indices_h = block_structure(elements, self.n_nodes)
        self.nnzb = int(self.indices_h.size)
        maps = BlockMaps(elements, self.n_nodes, self.indptr_h, self.indices_h)
        self.maps = maps
        # exact structure maxima for the row-load kernels
        row_of = np.repeat(np.arange(self.n_nodes), np.diff(self.indptr_h))
        nc = np.diff(maps.cptr.astype(np.int64))
        diag = self.indices_h == row_of
        self.max_row_blocks = int(np.diff(self.indptr_h).max())
        self.max_offdiag_contrib = int(nc[~diag].max()) if np.any(~diag) else 1
        self.max_diag_contrib = int(nc[diag].max()) if np.any(diag) else 1
        self.variant = _lib.VARIANT_AUTO
        i32 = torch.int32
        self.nodes = _dev(nodes, torch.float64, dev)
        self.elems = _dev(elements, i32, dev)
        self.indptr = _dev(self.indptr_h, i32, dev)
        self.indices = _dev(self.indices_h, i32, dev)
        self.cptr = _dev(maps.cptr, i32, dev)
        self.celem = _dev(maps.celem, i32, dev)
        self.cpos = _dev(maps.cpos, i32, dev)
        self.nptr = _dev(maps.nptr, i32, dev)
        self.nent = _dev(maps.nent, i32, dev)
        self.cslot = _dev(maps.cslot, i32, dev)
        self.max_node_elems = maps.max_node_elems
        self.ublk = _dev(maps.ublk, i32, dev)
        self.vblk = _dev(maps.vblk, i32, dev)
        self.nnzu = maps.nnzu
        self.uvalues = None
        self._nodes_h = nodes
        self._elems_h = np.asarray(elements, np.int64)
        self.row_order = None
        self._order_h = None                          # host copy of the visiting order (None: natural)
        self.row_records = None                       # symmetric kernels' row records (on demand)
        self.record_width = 0
        self.color_ptr = self.color_elems = None      # element colouring (EBE variant), on demand
        self.n_colors = self.max_color_elems = 0
        self.waves_per_simd = 0                       # occupancy cap of the step kernels (0 = none)
        fixed = np.zeros(self.n_nodes, np.uint8)
        fixed[np.asarray(boundary, np.int64)] = 1
        self.fixed = _dev(fixed, torch.uint8, dev)
        self.boundary = np.asarray(boundary, np.int64)
        self.bdy_u = float(bdy_u)
        f64 = torch.float64
        S = self.n_smp
        with torch.cuda.device(dev):
            st = _lib.stream_handle(dev)
            # --- element blocks ------------------------------------------------------------------
            self.khat = torch.empty((9 * self.n_elms, 9), dtype=f64, device=dev)
            self.area = torch.empty(self.n_elms, dtype=f64, device=dev)
            self.normal = torch.empty((self.n_elms, 3), dtype=f64, device=dev)
            check(self.lib.cvfes_element_matrices(st, self.n_elms, ptr(self.nodes), ptr(self.elems),
                                                  self.nu, self.kshear, ptr(self.cpos),
                                                  ptr(self.khat), ptr(self.area), ptr(self.normal)))
            # --- thickness and stiffness coefficient ---------------------------------------------
            vthick = np.asarray(vthickness, np.float64)
            if vthick.ndim == 2 and vthick.shape[1] == S and np.all(vthick == vthick[:, :1]):
                vthick = vthick[:, 0]
            self.thick_sstride = 1 if vthick.ndim == 2 else 0
            self.vthick = _dev(vthick, f64, dev)
            # element-average modulus for the stress recovery (mean over the Gauss points,
            # physicsSolver.py:416-424)
            self.ave_E = None
            if coef is None and elmE is not None:
                self.ave_E = _dev(np.asarray(elmE, np.float64).reshape(self.n_elms, S), f64, dev)
            elif coef is None and vE is not None:
                vEa = np.asarray(vE, np.float64).reshape(self.n_nodes, S)
                gE = (0.5 * (vEa[elements[:, 0]] + vEa[elements[:, 1]]) + 0.5 * (vEa[elements[:, 0]] + vEa[elements[:, 2]])
                      + 0.5 * (vEa[elements[:, 1]] + vEa[elements[:, 2]])) / 3.0
                self.ave_E = _dev(gE, f64, dev)
            if coef is not None:
                self.coef = _dev(np.asarray(coef, np.float64).reshape(self.n_elms, S), f64, dev)
            else:
                vE_d = None if vE is None else _dev(np.asarray(vE, np.float64).reshape(self.n_nodes, S), f64, dev)
                elmE_d = None if elmE is None else _dev(np.asarray(elmE, np.float64).reshape(self.n_elms, S), f64, dev)
                if vE_d is None and elmE_d is None:
                    raise ValueError("one of coef, vE, elmE is required")
                self.coef = torch.empty((self.n_elms, S), dtype=f64, device=dev)
                check(self.lib.cvfes_element_coef(st, self.n_elms, S, ptr(self.elems), ptr(self.vthick),
                                                  self.thick_sstride, ptr(vE_d), ptr(elmE_d), 0,
                                                  ptr(self.coef)))
                torch.cuda.current_stream(dev).synchronize()
            # --- lumped mass and unit-pressure load -----------------------------------------------
            self.mass_form = mass_form
            lm_shape = (self.n_nodes, S) if self.thick_sstride else (self.n_nodes,)
            self.lm = torch.empty(lm_shape, dtype=f64, device=dev)
            self.f1 = torch.empty((self.n_nodes, 3), dtype=f64, device=dev)
            check(self.lib.cvfes_lumped_mass_load(st, self.n_nodes, S, ptr(self.nptr), ptr(self.nent),
                                                  ptr(self.elems), ptr(self.area), ptr(self.normal),
                                                  ptr(self.vthick), self.thick_sstride, self.density,
                                                  MASS_SCALE[mass_form], ptr(self.lm), ptr(self.f1)))
            # inverse masses of the adjoint and gradient kernels
            self.minv = torch.empty_like(self.lm)
            check(self.lib.cvfes_inverse_mass(st, self.lm.numel(), ptr(self.lm), ptr(self.minv)))
            # --- stored block-CSR values ---------------------------------------------------------
            self.values = None
            if store_values:
                self.assemble_values()
        self._sys = self._make_struct()
        self.state = None
        self.update = None
        self.ug = None

    # ------------------------------------------------------------------------------------------
    def set_log_modulus(self, theta):
        """Re-form ``coef`` from per-element ``theta = log E`` (device tensor ``[E,S]``) and, for the
        stored operator, the block values."""
        theta = theta.detach().to(self.device, torch.float64).contiguous()
        if theta.shape != (self.n_elms, self.n_smp):
            raise ValueError("theta must be [n_elms, n_smp]")
        check(self.lib.cvfes_element_coef(self.stream, self.n_elms, self.n_smp, ptr(self.elems),
                                          ptr(self.vthick), self.thick_sstride, None, ptr(theta), 1,
                                          ptr(self.coef)))
        self.ave_E = torch.exp(theta)
        if self.values is not None or self.uvalues is not None:
            self.assemble_values()

    def assemble_values(self):
        """Form (or re-form) the stored block values from ``coef``: the upper blocks ``uvalues``
        with symmetric storage, the full ``values [nnzb][9][S]`` with full storage (and the upper
        blocks as well when they exist)."""
        S = self.n_smp
        if self.storage == "full":
            if self.values is None:
                self.values = torch.empty((self.nnzb, 9, S), dtype=torch.float64, device=self.device)
            check(self.lib.cvfes_assemble_bcsr(self.stream, self.nnzb, S, ptr(self.cptr),
                                               ptr(self.celem), ptr(self.khat), ptr(self.coef),
                                               ptr(self.values)))
        if self.storage == "sym" or self.uvalues is not None:
            self.assemble_upper()
        if hasattr(self, "_sys"):
            self._sys = self._make_struct()

    def set_storage(self, storage):
        """Switch the stored values to ``"sym"`` (upper blocks only; the full values are dropped)
        or ``"full"`` (every block; existing upper blocks are kept for the symmetric variants)."""
        if storage not in ("sym", "full"):
            raise ValueError("storage must be 'sym' or 'full'")
        self.storage = storage
        if storage == "sym":
            self.values = None
        self.assemble_values()

    def assemble_upper(self):
        """Form (or re-form) the symmetric storage ``uvalues [nnzu][9][S]`` (upper blocks only) used
        by the ``sym`` step variants."""
        S = self.n_smp
        if self.uvalues is None:
            self.uvalues = torch.empty((self.nnzu, 9, S), dtype=torch.float64, device=self.device)
        check(self.lib.cvfes_assemble_upper(self.stream, self.nnzu, S, ptr(self.ublk), ptr(self.cptr),
                                            ptr(self.celem), ptr(self.khat), ptr(self.coef),
                                            ptr(self.uvalues)))
        if self.row_records is None:
            self._build_records()
        if hasattr(self, "_sys"):
            self._sys = self._make_struct()

    def _record_slots(self):
        """Block slots of the symmetric kernels for this structure (as the C side picks them), or
        0 when they do not apply."""
        if self.max_offdiag_contrib > 2:
            return 0
        for maxb in (7, 10):
            if self.max_row_blocks <= maxb and self.max_diag_contrib <= maxb - 1:
                return maxb
        return 0

    def _build_records(self):
        """Row records of the symmetric kernels in the current visiting order: row id, block count,
        the blocks' columns and upper-storage entries, padded with the row's last block."""
        maxb = self._record_slots()
        if maxb == 0:
            self.row_records, self.record_width = None, 0
            return
        ip = self.indptr_h.astype(np.int64)
        rows = np.arange(self.n_nodes) if self._order_h is None else self._order_h.astype(np.int64)
        p0 = ip[rows]
        nb = ip[rows + 1] - p0
        slot = p0[:, None] + np.minimum(np.arange(maxb)[None, :], np.maximum(nb - 1, 0)[:, None])
        slot = np.minimum(slot, max(self.nnzb - 1, 0))
        rec = np.empty((self.n_nodes, 2 + 2 * maxb), np.int32)
        rec[:, 0] = rows
        rec[:, 1] = nb
        rec[:, 2:2 + maxb] = self.indices_h[slot]
        rec[:, 2 + maxb:] = self.maps.vblk[slot]
        self.row_records = _dev(rec, torch.int32, self.device)
        self.record_width = 2 + 2 * maxb

    def sample_block(self, s0, s1):
        """The system of samples ``s0 .. s1-1`` alone, for the matrix-free operator: it shares the
        structure, element blocks, loads and per-node masses with this one and holds its own
        copies of the block's coefficients (and per-sample thickness and masses), no stored
        values and no state.  Samples do not interact, so a step of the block system computes
        exactly (bit for bit) the block's samples of a step of this one.  :meth:`refresh_block`
        copies the coefficients again after this system's change."""
        s0, s1 = int(s0), int(s1)
        if not 0 <= s0 < s1 <= self.n_smp:
            raise ValueError("sample block must satisfy 0 <= s0 < s1 <= n_smp")
        b = copy.copy(self)
        b.n_smp = s1 - s0
        b.block_of = (s0, s1)
        b.coef = self.coef[:, s0:s1].contiguous()
        b.ave_E = None if self.ave_E is None else self.ave_E[:, s0:s1].contiguous()
        if self.thick_sstride:
            b.vthick = self.vthick[:, s0:s1].contiguous()
            b.lm = self.lm[:, s0:s1].contiguous()
            b.minv = None if self.minv is None else self.minv[:, s0:s1].contiguous()
        b.values = b.uvalues = None
        b.row_records, b.record_width = None, 0
        b.state, b.update, b.ug = None, None, None
        b._sys = b._make_struct()
        return b

    def refresh_block(self, parent):
        """Copy this sample block's coefficients (and per-sample thickness and masses) from the
        system it was taken from (:meth:`sample_block`)."""
        s0, s1 = self.block_of
        self.coef.copy_(parent.coef[:, s0:s1])
        if parent.ave_E is not None and self.ave_E is not None:
            self.ave_E.copy_(parent.ave_E[:, s0:s1])
        if parent.thick_sstride:
            self.vthick.copy_(parent.vthick[:, s0:s1])
            self.lm.copy_(parent.lm[:, s0:s1])
            if self.minv is not None:
                self.minv.copy_(parent.minv[:, s0:s1])

    def release_values(self):
        """Drop the stored values (matrix-free runs only need coef and khat)."""
        self.values = None
        self.uvalues = None
        self._sys = self._make_struct()

    def _make_struct(self):
        return CvfesSystem(
            n_nodes=self.n_nodes, n_smp=self.n_smp, nnzb=self.nnzb, n_elms=self.n_elms,
            indptr=ptr(self.indptr), indices=ptr(self.indices), values=ptr(self.values),
            cptr=ptr(self.cptr), celem=ptr(self.celem), khat=ptr(self.khat), coef=ptr(self.coef),
            lm=ptr(self.lm), lm_sstride=self.thick_sstride, f1=ptr(self.f1), fixed=ptr(self.fixed),
            bdy_u=self.bdy_u, max_row_blocks=self.max_row_blocks,
            max_offdiag_contrib=self.max_offdiag_contrib, max_diag_contrib=self.max_diag_contrib,
            variant=self.variant, nptr=ptr(self.nptr), nent=ptr(self.nent), cslot=ptr(self.cslot),
            max_node_elems=self.max_node_elems, uvalues=ptr(self.uvalues), vblk=ptr(self.vblk),
            nnzu=self.nnzu, row_order=ptr(self.row_order), color_ptr=ptr(self.color_ptr),
            color_elems=ptr(self.color_elems), elems=ptr(self.elems), cpos=ptr(self.cpos),
            n_colors=self.n_colors, max_color_elems=self.max_color_elems,
            waves_per_simd=self.waves_per_simd, row_records=ptr(self.row_records),
            record_width=self.record_width, minv=ptr(self.minv))

    def set_variant(self, variant):
        """Select the step kernel variant (names in ``_lib.VARIANTS``); the symmetric ones form the
        upper-block storage on first use, the element-by-element one the element colouring."""
        self.variant = _lib.VARIANTS[variant] if isinstance(variant, str) else int(variant)
        if self.variant in (_lib.VARIANT_SYM, _lib.VARIANT_SYM_NT, _lib.VARIANT_SYM_NTL) and self.uvalues is None \
                and self.values is not None:
            self.assemble_upper()
        if self.variant == _lib.VARIANT_EBE and self.color_ptr is None:
            self.set_coloring()
        self._sys = self._make_struct()

    def set_coloring(self, groups=None):
        """Element colouring for the element-by-element variant: ``groups`` (element-id arrays, no
        two elements of a group sharing a node; default: the greedy ``mesh_coloring``)."""
        if groups is None:
            from .coloring import mesh_coloring
            groups = mesh_coloring(self._elems_h, self.n_nodes, n_colors=64)[1]
        groups = [np.asarray(g, np.int64) for g in groups if len(g)]
        flat = np.concatenate(groups) if groups else np.zeros(0, np.int64)
        if not np.array_equal(np.sort(flat), np.arange(self.n_elms)):
            raise ValueError("the colour groups must partition the elements")
        for g in groups:
            nodes = self._elems_h[g].ravel()
            if np.unique(nodes).size != nodes.size:
                raise ValueError("two elements of one colour share a node")
        ptrs = np.zeros(len(groups) + 1, np.int64)
        np.cumsum([len(g) for g in groups], out=ptrs[1:])
        self.color_ptr = _dev(ptrs.astype(np.int32), torch.int32, self.device)
        self.color_elems = _dev(flat.astype(np.int32), torch.int32, self.device)
        self.n_colors = len(groups)
        self.max_color_elems = max((len(g) for g in groups), default=0)
        self._sys = self._make_struct()

    def set_row_order(self, order="bisection", leaf=64, width=64):
        """Visiting order of the block rows for the wave-uniform kernels: ``"natural"``,
        ``"bisection"`` (recursive coordinate bisection, parts of ``leaf`` nodes), ``"strips"``
        (angular strips of about ``width`` ring nodes swept along the tube axis; the fastest order
        for the symmetric-storage kernels) or an explicit permutation.  The result does not depend
        on it."""
        if order is None or (isinstance(order, str) and order == "natural"):
            self.row_order = None
        else:
            if isinstance(order, str) and order == "strips":
                perm = strip_order(self._nodes_h, self._elems_h, width)
            elif isinstance(order, str) and order == "bisection":
                perm = bisection_order(self._nodes_h, leaf)
            elif isinstance(order, str):
                raise ValueError(f"unknown row order {order!r}")
            else:
                perm = np.asarray(order)
            if perm.shape != (self.n_nodes,) or not np.array_equal(np.sort(perm), np.arange(self.n_nodes)):
                raise ValueError("row order must be a permutation of the nodes")
            self.row_order = _dev(perm.astype(np.int32), torch.int32, self.device)
        self._order_h = None if self.row_order is None else perm.astype(np.int32)
        if self.uvalues is not None:
            self._build_records()
        self._sys = self._make_struct()

    def set_occupancy(self, waves_per_simd=0):
        """Cap the step kernels that stage nothing in LDS at ``waves_per_simd`` (2..7; 0 = no cap)
        wavefronts per SIMD.  The results do not depend on it."""
        w = int(waves_per_simd)
        if w != 0 and not 2 <= w <= 7:
            raise ValueError("waves_per_simd must be 0 or 2..7")
        self.waves_per_simd = w
        self._sys = self._make_struct()

    def _time_steps(self, load, dt, beta, damp, method, steps):
        z = np.zeros(3 * self.n_nodes)
        self.initial_state(z, z, load, dt, beta=beta, damp=damp, method=method)
        self.step(load, dt, beta=beta, damp=damp, method=method, offset=0, nsteps=2)
        st = torch.cuda.current_stream(self.device)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        self.step(load, dt, beta=beta, damp=damp, method=method, offset=2, nsteps=steps)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / steps

    def autotune_row_order(self, load, dt, beta=0.0, damp=0.0, method="stored", steps=20,
                           candidates=(("natural", 0), ("strips", 32), ("strips", 48), ("strips", 64))):
        """Time ``steps`` steps under each visiting order in ``candidates`` ((order, strip width))
        and keep the fastest; returns ``{"order", "width", "ms"}``.  The displacements do not depend
        on the order (every order gives the same result bit for bit), only the time does.  Leaves
        a fresh zero initial state behind."""
        best = None
        for order, width in candidates:
            self.set_row_order(order, width=width)
            ms = self._time_steps(load, dt, beta, damp, method, steps)
            if best is None or ms < best["ms"]:
                best = {"order": order, "width": width, "ms": ms}
        self.set_row_order(best["order"], width=best["width"])
        z = np.zeros(3 * self.n_nodes)
        self.initial_state(z, z, load, dt, beta=beta, damp=damp, method=method)
        return best

    def autotune_occupancy(self, load, dt, beta=0.0, damp=0.0, method="stored", steps=10,
                           candidates=(0, 6, 5, 4, 3, 2)):
        """Time ``steps`` steps under each occupancy cap in ``candidates`` (waves per SIMD, 0 = no
        cap) and keep the fastest; returns ``{"waves_per_simd", "ms"}``.  Large rows (many
        samples, long rings) run faster with fewer rows in flight per XCD; the displacements do not
        depend on it.  Leaves a fresh zero initial state behind."""
        best = None
        for w in candidates:
            self.set_occupancy(w)
            ms = self._time_steps(load, dt, beta, damp, method, steps)
            if best is None or ms < best["ms"]:
                best = {"waves_per_simd": w, "ms": ms}
        self.set_occupancy(best["waves_per_simd"])
        z = np.zeros(3 * self.n_nodes)
        self.initial_state(z, z, load, dt, beta=beta, damp=damp, method=method)
        return best

    @property
    def stream(self):
        return _lib.stream_handle(self.device)

    @staticmethod
    def method_id(method):
        if isinstance(method, int):
            return method
        return _lib.METHODS[method]

    def state_shape(self):
        return (self.n_nodes, 3, self.n_smp)

    # ------------------------------------------------------------------------------------------
    def geometry(self):
        return CvfesGeometry(nodes=ptr(self.nodes), elems=ptr(self.elems), nu=self.nu, kshear=self.kshear)

    def element_stress(self, u=None, ug=None, ave_E=None):
        """Per-element stress ``[E, 5, S]`` (xx, yy, xy, xz, yz) of ``u`` (default: the current
        displacement) on the geometry ``X + ug`` (default: the reference geometry, or in the
        updated-geometry mode the geometry of the last re-formation, as ``mesh.updateNodes``)."""
        u = self.u if u is None else u.contiguous()
        if ug is None and self.update is not None and self.update.has_geometry:
            # geometry of the last re-formation: with interval 1 the displacement the last step
            # started from (u_{n-1}), otherwise the saved copy
            ug = self.state[1] if self.update.interval == 1 else self.ug
        ave_E = self.ave_E if ave_E is None else ave_E.contiguous()
        if ave_E is None:
            raise ValueError("the system was built from coef: pass ave_E")
        out = torch.empty((self.n_elms, 5, self.n_smp), dtype=torch.float64, device=self.device)
        geo = self.geometry()
        check(self.lib.cvfes_element_stress(self.stream, self.n_elms, self.n_smp, ctypes.byref(geo),
                                            ptr(u), ptr(ug), ptr(ave_E), ptr(out)))
        return out

    def step_updated(self, load, dt, interval, beta=0.0, damp=0.0, offset=0, nsteps=None):
        """Advance ``nsteps`` steps of the updated-geometry mode (stiffness and unit-traction load
        re-formed on ``X + u`` every ``interval`` steps, matrix-free); the schedule continues
        across calls until the next :meth:`initial_state`."""
        if self.state is None:
            raise RuntimeError("call initial_state() first")
        if not isinstance(load, LoadTables):
            load = self.loads(load)
        nsteps = load.n - offset if nsteps is None else int(nsteps)
        if offset < 0 or nsteps < 0 or offset + nsteps > load.n:
            raise ValueError("steps exceed the load tables")
        interval = int(interval)
        if interval < 1:
            raise ValueError("interval must be >= 1")
        if self.update is None:
            self.update = CvfesUpdate(interval=interval, counter=0, has_geometry=0, ug=None)
        elif self.update.interval != interval:
            raise ValueError("the re-formation interval is fixed until the next initial_state()")
        if interval > 1 and self.ug is None:
            self.ug = torch.empty(self.state_shape(), dtype=torch.float64, device=self.device)
        self.update.ug = ptr(self.ug)
        arr = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in self.state])
        ls = load.struct(offset)
        geo = self.geometry()
        check(self.lib.cvfes_step_updated(self.stream, ctypes.byref(self._sys), ctypes.byref(geo),
                                          ctypes.byref(self.update), arr, ctypes.byref(ls),
                                          int(nsteps), float(dt), float(beta), float(damp)))
        by_ptr = {b.data_ptr(): b for b in self.state}
        self.state = [by_ptr[arr[i]] for i in range(3)]
        self._load_keepalive = load
        return self.state[0]

    def spmv(self, x, method="stored", out=None):
        """``y = K x`` for ``x [N,3,S]`` (device tensor)."""
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        check(self.lib.cvfes_bcsr_spmv(self.stream, ctypes.byref(self._sys), self.method_id(method),
                                       ptr(x), ptr(y)))
        return y

    def loads(self, p, ga=None, qa=None, gb=None, qb=None):
        """Upload load tables (see :class:`LoadTables`)."""
        return LoadTables(p, self.device, ga=ga, qa=qa, gb=gb, qb=qb, n_nodes=self.n_nodes)

    def initial_state(self, u0, du0, load, dt, beta=0.0, damp=0.0, method="stored", offset=0):
        """Allocate the three state buffers and set ``u_0``, ``u_{-1}`` (``u0``/``du0`` are ``[3N]``).

        ``load`` is a :class:`LoadTables` (or a pressure value/array); its entry ``offset`` is the
        load at the start time.
        """
        f64 = torch.float64
        if not isinstance(load, LoadTables):
            load = self.loads(load)
        shp = self.state_shape()
        bufs = [torch.empty(shp, dtype=f64, device=self.device) for _ in range(3)]
        u0_d = _dev(np.asarray(u0, np.float64).reshape(-1), f64, self.device)
        du0_d = _dev(np.asarray(du0, np.float64).reshape(-1), f64, self.device)
        ls = load.struct(offset)
        check(self.lib.cvfes_initial_state(self.stream, ctypes.byref(self._sys), self.method_id(method),
                                           ptr(u0_d), ptr(du0_d), ctypes.byref(ls), float(dt),
                                           float(beta), float(damp), ptr(bufs[0]), ptr(bufs[1]),
                                           ptr(bufs[2])))
        torch.cuda.current_stream(self.device).synchronize()
        self.state = bufs
        self.update = None
        return bufs

    def step(self, load, dt, beta=0.0, damp=0.0, method="stored", offset=0, nsteps=None):
        """Advance ``nsteps`` steps using load-table entries ``offset ... offset+nsteps-1``.

        ``load`` is a :class:`LoadTables` or a host/device array of pressures.
        """
        if self.state is None:
            raise RuntimeError("call initial_state() first")
        if not isinstance(load, LoadTables):
            load = self.loads(load)
        n = load.n - offset if nsteps is None else int(nsteps)
        if offset < 0 or n < 0 or offset + n > load.n:
            raise ValueError("steps exceed the load tables")
        ls = load.struct(offset)
        arr = (ctypes.c_void_p * 3)(*[b.data_ptr() for b in self.state])
        check(self.lib.cvfes_step(self.stream, ctypes.byref(self._sys), self.method_id(method), arr,
                                  ctypes.byref(ls), n, float(dt), float(beta), float(damp)))
        addr = {b.data_ptr(): b for b in self.state}
        self.state = [addr[arr[k]] for k in range(3)]
        self._load_keepalive = load
        return self.state[0]

    @property
    def u(self):
        return self.state[0]

    @property
    def up(self):
        return self.state[1]

    # ------------------------------------------------------------------------------------------
    def values_reference_layout(self):
        """Stored values as the reference's ``[nnzb, S, 3, 3]`` (host numpy); with symmetric
        storage every lower block is its mirror transposed."""
        if self.values is not None:
            v = self.values.cpu().numpy()
            return np.ascontiguousarray(v.reshape(self.nnzb, 3, 3, self.n_smp).transpose(0, 3, 1, 2))
        u = self.uvalues.cpu().numpy().reshape(self.nnzu, 3, 3, self.n_smp)
        blk = u[self.maps.vblk]
        row = np.repeat(np.arange(self.n_nodes), np.diff(self.indptr_h))
        low = self.indices_h < row
        blk[low] = blk[low].transpose(0, 2, 1, 3)
        return np.ascontiguousarray(blk.transpose(0, 3, 1, 2))

    def lumped_mass_dofs(self):
        """Lumped mass as the reference's ``[ndof, S]`` (host numpy)."""
        lm = self.lm.cpu().numpy()
        if lm.ndim == 1:
            lm = np.repeat(lm[:, None], self.n_smp, axis=1)
        return np.repeat(lm, 3, axis=0)

    def critical_dt(self, beta=0.0, iters=60, seed=0, method="stored", damp=0.0):
        """Largest stable step from a power iteration on ``M^-1 K`` (max over samples).

        A mode of ``M^-1 K`` with eigenvalue ``lam`` and mass-independent damping rate
        ``g = damp / m`` obeys ``z^2 - (2 - dt^2 lam - c) z + (1 - c) = 0`` with
        ``c = dt (beta lam + g)`` (backward-difference damping forces), whose roots stay in the
        unit circle for ``dt^2 lam + 2 dt (beta lam + g) < 4``.  The bound is taken with ``lam``
        and ``g`` both at their maxima (``g`` at the smallest non-zero mass), which for
        ``damp = 0`` is ``dt < -beta + sqrt(beta^2 + 4/lam)``.  The full operator (fixed rows
        included) bounds the constrained one from above, but the Rayleigh quotient approaches
        ``lam_max`` from below, so the returned step can exceed the true limit by the iteration's
        error (a few tenths of a percent after 60 iterations on the tubes): run at a fraction of
        it.  Returns ``(dt_max, lam_max)``.
        """
        g = torch.Generator(device="cpu").manual_seed(seed)
        x = torch.rand(self.state_shape(), generator=g, dtype=torch.float64).to(self.device)
        lm = self.lm
        lm_b = lm[:, None, :] if lm.ndim == 2 else lm[:, None, None]
        lam = None
        for _ in range(iters):
            kx = self.spmv(x, method)
            # Rayleigh quotient of the pencil (K, M): never above lam_max, converges from below.
            lam = (x * kx).sum(dim=(0, 1)) / (lm_b * x * x).sum(dim=(0, 1))
            y = kx / lm_b
            x = y / torch.linalg.vector_norm(y, dim=(0, 1))
        lam_max = float(lam.max())
        g = 0.0
        if damp != 0.0:
            m = lm[lm > 0]
            g = float(damp) / float(m.min())
        b = beta * lam_max + g
        return float((-b + np.sqrt(b * b + 4.0 * lam_max)) / lam_max), lam_max
