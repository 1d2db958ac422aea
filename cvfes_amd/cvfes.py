"""Solid-only run driver with the ``CVFES`` call sequence (``cvfes.py:26-119``, ``main.py:21-49``).

``CVFES(comm).ReadInputFile(cfg, nSmp)`` builds the wall mesh either from a VTK ``.vtp`` or ``.vtu``
file with ``faces`` (boundary = union of the face nodes matched by ``GlobalNodeID``,
``mesh.py:50-67,175-184,658-665``) or, with
``generator = tube`` in the wall mesh section, from a synthetic tube.  Young's modulus and thickness
come from the solid equation section: uniform values, ``.npy`` property files (``mesh.py:620-655``)
or, with ``Youngs Modulus model = gmrf_element``, per-element GMRF samples.  ``Solve()`` runs
``TransientSolverGPU``.  Partitioning and colouring are single-rank no-ops here: the ensemble is
sharded by samples across GPUs, not by domain.
"""
from __future__ import annotations

import os

import numpy as np

from .cvconfig import CVConfig, read_cfg
from .gmrf import sample_element_E
from .mesh import SolidMesh, stenosed_radius, tube_geometry
from .physicsSolver import SerialComm
from .solver import TransientSolverGPU
from . import vtkio

__all__ = ["CVFES", "build_wall_mesh", "main"]


def _prop_file(name, rho):
    base, ext = os.path.splitext(name)
    path = f"{base}{rho}{ext}"
    if not os.path.exists(path):
        raise FileNotFoundError(f"Solid mesh, load random fields failed: File '{path}' does not exist!")
    return np.load(path)


def build_wall_mesh(mc, eq, nSmp):
    """``SolidMesh`` from the wall mesh config ``mc`` and solid equation config ``eq``."""
    if getattr(mc, "generator", None) == "tube":
        profile = None
        if mc.radius_out is not None or mc.stenosis:
            r_out = mc.radius if mc.radius_out is None else mc.radius_out
            profile = stenosed_radius(mc.radius, r_out, mc.length, severity=mc.stenosis)
        nodes, elements, boundary = tube_geometry(mc.ntheta, mc.nz, mc.radius, mc.length, profile)
    else:
        wall = vtkio.read_vtk(mc.file_path)
        nodes, elements = wall.points, wall.polys
        gids = wall.point_data.get("GlobalNodeID")
        boundary = []
        for face in mc.faces:
            if face.name != "boundaries":
                continue
            if gids is None:
                raise ValueError("wall mesh has no GlobalNodeID array to match its faces")
            # first wall node carrying each face GlobalNodeID; ids absent from the wall are skipped
            # (Face.adjustNodeIds, mesh.py:59-67)
            lookup = {}
            for i, g in enumerate(np.asarray(gids).ravel()):
                lookup.setdefault(int(g), i)
            paths = face.file_path if isinstance(face.file_path, list) else [face.file_path]
            for p in paths:
                fgids = np.asarray(vtkio.read_vtk(p).point_data["GlobalNodeID"]).ravel()
                boundary.extend(lookup[int(g)] for g in fgids if int(g) in lookup)
        boundary = np.array(sorted(set(boundary)), dtype=np.int64)
    n = nodes.shape[0]
    vthick = (np.full((n, nSmp), eq.thickness) if eq.thicknessFilename is None
              else _prop_file(eq.thicknessFilename, eq.rhoThickness))
    elmE = None
    vE = None
    if getattr(eq, "EModel", None) == "gmrf_element":
        elmE = sample_element_E(elements, nSmp, seed=eq.gmrfSeed, E0=eq.E, cov=eq.gmrfCov).E
    elif eq.YoungsModulusFilename is not None:
        vE = _prop_file(eq.YoungsModulusFilename, eq.rhoE)
    iniU = iniDu = None
    ic = eq.initialConditions
    if ic is not None and isinstance(ic.displacement, str):
        iniU = np.asarray(vtkio.read_vtk(ic.displacement).point_data["displacement"], np.float64).ravel()
    if ic is not None and isinstance(ic.velocity, str):
        iniDu = np.asarray(vtkio.read_vtk(ic.velocity).point_data["velocity"], np.float64).ravel()
    if ic is not None and not isinstance(ic.displacement, str) and ic.displacement:
        iniU = np.full(3 * n, float(ic.displacement))
    if ic is not None and not isinstance(ic.velocity, str) and ic.velocity:
        iniDu = np.full(3 * n, float(ic.velocity))
    bdyU = eq.boundaryConditions.bdyDisplacement if eq.boundaryConditions is not None else 0.0
    mesh = SolidMesh(nodes, elements, boundary, nSmp, thickness=eq.thickness, density=eq.density,
                     E=eq.E, v=eq.v, damp=eq.damp, vE=vE, vthickness=vthick, elmE=elmE,
                     iniU=iniU, iniDu=iniDu, bdyU=float(bdyU))
    return mesh


class CVFES:
    """Solid-only CVFES driver (``cvfes.py:26-119``)."""

    def __init__(self, comm=None):
        self.comm = comm or SerialComm()
        self.size = self.comm.Get_size()
        self.rank = self.comm.Get_rank()

    def ReadInputFile(self, filename, nSmp=None):
        self.cvConfig = CVConfig(read_cfg(filename))
        if nSmp is not None:
            self.cvConfig.nSmp = nSmp
            self.cvConfig.meshes["wall"].nSmp = nSmp
            self.cvConfig.solver.nSmp = nSmp
        cfg = self.cvConfig
        self.meshes = {"lumen": None,
                       "wall": build_wall_mesh(cfg.meshes["wall"], cfg.equations["solid"], cfg.nSmp)}
        self.name = cfg.name
        return self

    def Distribute(self):
        return 0

    def Coloring(self):
        return 0

    def Solve(self, fast=False, verbose=False):
        method = self.cvConfig.solver.method
        if method != "transient GPU":
            print("Unknown method: {}".format(method))
            return None
        self.solver = TransientSolverGPU(self.comm, self.meshes, self.cvConfig.solver, verbose=verbose)
        self.solver.Solve(fast=fast)
        return self.solver


def main(argv=None):
    import argparse
    parser = argparse.ArgumentParser()
    parser.add_argument("-s", "--samples", type=int, help="Number of samples.")
    parser.add_argument("-f", "--file", default="cfg/input.cfg", help="Configuration file.")
    parser.add_argument("--fast", action="store_true", help="advance between save points in one call")
    args = parser.parse_args(argv)
    cv = CVFES()
    cv.ReadInputFile(args.file, nSmp=args.samples)
    cv.Solve(fast=args.fast, verbose=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
