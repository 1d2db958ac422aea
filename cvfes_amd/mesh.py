"""Synthetic membrane-tube meshes carrying the attribute set of CVFES's ``SolidMesh``.

The solid solvers of CVFES read a wall mesh from VTK files (``mesh.py:130-173``) and per-node
Young's modulus / thickness fields from ``.npy`` files (``mesh.py:620-655``).  This module builds
the same attribute set for structured tubes (straight, tapered, stenosed), so the GPU solver can
be driven without VTK:

* ``nodes [N,3]``, ``elementNodeIds [E,3]`` (int64, as ``mesh.py:157``),
* ``vE``/``vthickness [N,S]`` (``mesh.py:605-614``), plus the additive per-element field
  ``elmE [E,S]`` used when Young's modulus is sampled per element,
* ``boundary`` = both end rings (``mesh.py:658-665``), ``iniU``/``iniDu`` (``mesh.py:667-671``),
  ``bdyU`` (``mesh.py:673-676``),
* the single-partition ``lcl*`` attributes that ``meshPartition.py:240-283`` would produce on one
  rank (identity numbering), and ``colorGroups`` (``meshColoring.py:7-63``) on demand.

Node numbering is ring-major: node ``iz*ntheta + it`` sits at angle ``2*pi*it/ntheta`` on ring
``iz``.  Each quad is split into two triangles wound counter-clockwise seen from outside, so the
element frame's third axis ``e3 = e1 x e2`` (``optimizedSolidAssemble.pyx:111-114``) points
outward and a positive pressure inflates the tube.
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np

__all__ = ["tube_geometry", "stenosed_radius", "SolidMesh", "tube_solid_mesh", "Element"]


def tube_geometry(ntheta: int, nz: int, radius: float, length: float,
                  radius_profile: Optional[Callable[[np.ndarray], np.ndarray]] = None):
    """Structured tube: returns ``nodes [N,3]`` f64, ``elements [E,3]`` int64, ``boundary`` int64.

    ``E = 2*ntheta*(nz-1)``; ``boundary`` lists the ``iz = 0`` ring then the ``iz = nz-1`` ring.
    ``radius_profile(z)`` overrides the constant radius (tapered / stenosed tubes).
    """
    if ntheta < 3 or nz < 2:
        raise ValueError("tube needs ntheta >= 3 and nz >= 2")
    it = np.arange(ntheta)
    iz = np.arange(nz)
    theta = 2.0 * math.pi * it / ntheta
    z = length * iz / (nz - 1)
    r = np.full(nz, float(radius)) if radius_profile is None else np.asarray(radius_profile(z), float)
    nodes = np.empty((nz, ntheta, 3))
    nodes[:, :, 0] = r[:, None] * np.cos(theta)[None, :]
    nodes[:, :, 1] = r[:, None] * np.sin(theta)[None, :]
    nodes[:, :, 2] = z[:, None]
    nodes = nodes.reshape(-1, 3)

    qz, qt = np.meshgrid(np.arange(nz - 1), it, indexing="ij")
    n00 = qz * ntheta + qt
    n10 = qz * ntheta + (qt + 1) % ntheta
    n01 = (qz + 1) * ntheta + qt
    n11 = (qz + 1) * ntheta + (qt + 1) % ntheta
    tri_a = np.stack([n00, n10, n11], axis=-1)
    tri_b = np.stack([n00, n11, n01], axis=-1)
    elements = np.stack([tri_a, tri_b], axis=2).reshape(-1, 3).astype(np.int64)
    boundary = np.concatenate([it, (nz - 1) * ntheta + it]).astype(np.int64)
    return nodes, elements, boundary


def stenosed_radius(r_in: float, r_out: float, length: float, severity: float = 0.5,
                    width: Optional[float] = None):
    """Linear taper ``r_in -> r_out`` times a Gaussian narrowing of ``severity`` at mid-length.

    ``width`` is the Gaussian standard deviation (default ``length/20``).
    """
    sigma = length / 20.0 if width is None else width

    def profile(z):
        z = np.asarray(z, float)
        base = r_in + (r_out - r_in) * z / length
        return base * (1.0 - severity * np.exp(-0.5 * ((z - 0.5 * length) / sigma) ** 2))

    return profile


class Element:
    """Minimal stand-in for the VTK-cell wrapper of ``mesh.py`` (only ``.nodes`` is used)."""

    __slots__ = ("nodes",)

    def __init__(self, nodes):
        self.nodes = nodes


class SolidMesh:
    """Wall mesh with the attribute contract of CVFES's ``SolidMesh`` (``mesh.py:591-680``).

    Parameters mirror the solid ``[[equation]]`` section of the ``.cfg`` (``cvconfig.py:110-133``):
    ``thickness``, ``density``, ``E`` (Young's modulus), ``v`` (Poisson ratio), ``damp``.  The
    per-node fields default to uniform values as ``mesh.py:628-629`` does when no property file is
    given.  ``elmE`` (``[E,S]``) is an additive attribute: when present the stiffness coefficient of
    an element is ``elmE * (Gauss-weighted thickness)`` instead of the node-interpolated product.
    """

    dof = 3

    def __init__(self, nodes, elements, boundary, nSmp, *, thickness=0.07, density=1.0, E=7.0e6,
                 v=0.5, damp=0.0, vE=None, vthickness=None, elmE=None, iniU=None, iniDu=None,
                 bdyU=0.0, name="wall"):
        self.name = name
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elementNodeIds = np.ascontiguousarray(elements, dtype=np.int64)
        self.nNodes = int(self.nodes.shape[0])
        self.nElements = int(self.elementNodeIds.shape[0])
        self.gnElements = self.nElements
        self.ndof = self.dof * self.nNodes
        self.nSmp = int(nSmp)
        self.thickness = float(thickness)
        self.density = float(density)
        self.E = float(E)
        self.v = float(v)
        self.damp = float(damp)
        self.vthickness = (np.full((self.nNodes, self.nSmp), self.thickness) if vthickness is None
                           else np.ascontiguousarray(vthickness, dtype=np.float64))
        self.vE = (np.full((self.nNodes, self.nSmp), self.E) if vE is None
                   else np.ascontiguousarray(vE, dtype=np.float64))
        self.elmE = None if elmE is None else np.ascontiguousarray(elmE, dtype=np.float64)
        if self.vthickness.shape != (self.nNodes, self.nSmp):
            raise ValueError("vthickness must be [nNodes, nSmp]")
        if self.vE.shape != (self.nNodes, self.nSmp):
            raise ValueError("vE must be [nNodes, nSmp]")
        if self.elmE is not None and self.elmE.shape != (self.nElements, self.nSmp):
            raise ValueError("elmE must be [nElements, nSmp]")
        self.boundary = np.ascontiguousarray(boundary, dtype=np.int64)
        self.iniU = np.zeros(self.ndof) if iniU is None else np.asarray(iniU, float).ravel().copy()
        self.iniDu = np.zeros(self.ndof) if iniDu is None else np.asarray(iniDu, float).ravel().copy()
        self.bdyU = float(bdyU)
        # Single-partition view: what meshPartition.py:240-283 yields on one rank.
        self.lclNNodes = self.nNodes
        self.lclNodeIds = np.arange(self.nNodes, dtype=np.int64)
        self.lclElmNodeIds = self.elementNodeIds
        self.lclBoundary = self.boundary
        self.lclNCommNodes = 0
        self.lclNSpecialHead = int(self.boundary.size)
        self.commNodeIds = None
        self.totalCommNodeIds = None
        self.partition = None
        self._colors = None
        self.updateNodes = None

    # --- lazily built attributes -------------------------------------------------------------
    @property
    def elements(self):
        """List of ``Element`` objects (``mesh.py:155``); built on first use."""
        return [Element(row) for row in self.elementNodeIds]

    @property
    def colorGroups(self):
        if self._colors is None:
            from .coloring import mesh_coloring
            self._colors = mesh_coloring(self.elementNodeIds, self.nNodes)
        return self._colors[1]

    @property
    def elmColors(self):
        if self._colors is None:
            self.colorGroups
        return self._colors[0]

    def UpdateCoordinates(self, u):
        """``updateNodes[s] = nodes + u[:, s]`` (``mesh.py:678-680``); ``u`` is ``[ndof, nSmp]``."""
        u = np.asarray(u).reshape(self.nNodes, self.dof, self.nSmp)
        self.updateNodes = self.nodes[None, :, :] + np.transpose(u, (2, 0, 1))
        return self.updateNodes


def tube_solid_mesh(ntheta: int, nz: int, nSmp: int, *, radius=0.3, length=3.0,
                    radius_profile=None, **props) -> SolidMesh:
    """Tube mesh wrapped as a :class:`SolidMesh` (both end rings fixed)."""
    nodes, elements, boundary = tube_geometry(ntheta, nz, radius, length, radius_profile)
    mesh = SolidMesh(nodes, elements, boundary, nSmp, **props)
    mesh.ntheta, mesh.nz = ntheta, nz
    return mesh
