"""The drop-in ``GPUSolidSolver`` driven like CVFES drives it, against the CPU oracle (GPU).

* ``TransientSolverGPU`` per-step loop (fluid time rounded to 4 decimals, ``solver.py:118``);
* the fast ``Run`` path (exact step times);
* fluid wall-traction input (snapshot series and constant stress) and the legacy module;
* ``Save`` writes the per-sample displacement arrays;
* ``stiffness_update_interval`` selects the updated-geometry mode (oracle ``run_updated``), and
  ``SaveStress`` writes the oracle's per-element stresses.
Bar: 1e-10 relative max-norm on the displacements.
"""
import os

import numpy as np
import pytest

from oracle import solid_oracle as so

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "cvfes_amd", "cfg", "tube_64x32_s16.cfg")
REL = 1e-10
# 2^-18 s (< the 4.98e-6 s stability bound): a power of two, so `np.arange(endtime, step=dt)` in the
# time loop (solver.py:75) gives exactly endtime/dt steps of exactly dt.
DT = 2.0 ** -18


def endtime(n):
    return repr(n * DT)


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def write_cfg(tmp_path, **repl):
    text = open(CFG).read()
    for k, v in repl.items():
        lines = []
        for line in text.splitlines():
            stripped = line.strip()
            if stripped.startswith(k + " =") or stripped.startswith(k + "="):
                indent = line[: len(line) - len(line.lstrip())]
                line = f"{indent}{k} = {v}"
            lines.append(line)
        text = "\n".join(lines)
    path = tmp_path / "case.cfg"
    path.write_text(text)
    return str(path)


class OracleModel:
    """Oracle assembly of the mesh a CVFES driver built."""

    def __init__(self, mesh):
        el, n, S = mesh.elementNodeIds, mesh.nNodes, mesh.nSmp
        self.indptr, self.indices = so.sparse_info_loops(el, n)
        gE, gth = so.gauss_properties(mesh.vE, mesh.vthickness, el)
        coefK = so.coef_k(gth, gE) if mesh.elmE is None else so.coef_k_element_E(gth, mesh.elmE)
        _, self.K, self.f1, self.LM = so.assemble_reference(mesh.nodes, el, 1.0, so.d_matrix(mesh.v),
                                                            coefK, gth, mesh.density, self.indptr,
                                                            self.indices)
        self.bdy_dofs = (3 * mesh.boundary[:, None] + np.arange(3)).ravel()

    def run(self, pressures, dt, beta=0.0, damp=0.0, u0=None, du0=None, **kw):
        n = self.indptr.size - 1
        u0 = np.zeros(3 * n) if u0 is None else u0
        du0 = np.zeros(3 * n) if du0 is None else du0
        u, _, _ = so.run(self.indptr, self.indices, self.K, self.LM, self.f1, pressures, u0, du0, dt,
                         beta=beta, damp=damp, bdy_dofs=self.bdy_dofs, **kw)
        return u


@pytest.fixture()
def driver(require_gpu):
    from cvfes_amd.cvfes import CVFES
    return CVFES


def test_transient_loop_matches_oracle(tmp_path, driver):
    cfg = write_cfg(tmp_path, endtime=endtime(200), dt=repr(DT))
    cv = driver().ReadInputFile(cfg)
    solver = cv.Solve()
    mesh = cv.meshes["wall"]
    n_steps = solver.nTimeSteps
    assert n_steps == 200
    wf = solver.solidSolver.waveform
    p = np.array([float(wf(round(solver.t[k], 4))) for k in range(n_steps)])
    om = OracleModel(mesh)
    u_ref = om.run(p, DT, p_init=float(wf(0.0)))
    assert rel_err(solver.solidSolver.GetDisplacement(), u_ref) < REL


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_fast_run_matches_oracle_exact_times(tmp_path, driver, method):
    cfg = write_cfg(tmp_path, endtime=endtime(300), dt=repr(DT), stiffness_method=method,
                    beta="2.0e-7", Damp="20.0")
    cv = driver().ReadInputFile(cfg)
    solver = cv.Solve(fast=True)
    mesh = cv.meshes["wall"]
    n_steps = solver.nTimeSteps
    wf = solver.solidSolver.waveform
    assert n_steps == 300
    p = wf(np.arange(n_steps) * DT)
    om = OracleModel(mesh)
    u_ref = om.run(p, DT, beta=2e-7, damp=20.0)
    assert rel_err(solver.solidSolver.GetDisplacement(), u_ref) < REL


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_case_read_from_files_matches_oracle(tmp_path, driver, method):
    """A case whose wall, faces, property fields and initial state come from files (VTK's default
    appended/base64/zlib encoding, a .vtu face, .npy fields with the rho suffix) solved on the GPU."""
    import vtkfiles as vf
    cfg, want = vf.write_cvfes_case(str(tmp_path), nsteps=80, dt=DT, method=method)
    cv = driver().ReadInputFile(cfg)
    mesh = cv.meshes["wall"]
    assert np.array_equal(mesh.vE, want["E"]) and np.array_equal(mesh.iniU, want["u0"])
    solver = cv.Solve(fast=True)
    n_steps = solver.nTimeSteps
    assert n_steps == 80
    p = solver.solidSolver.waveform(np.arange(n_steps) * DT)
    om = OracleModel(mesh)
    assert np.array_equal(np.sort(om.bdy_dofs // 3), np.repeat(want["boundary"], 3))
    u_ref = om.run(p, DT, u0=mesh.iniU, du0=mesh.iniDu)
    assert np.max(np.abs(u_ref)) > 0
    assert rel_err(solver.solidSolver.GetDisplacement(), u_ref) < REL


@pytest.mark.parametrize("legacy", [False, True])
def test_wall_traction_series(tmp_path, driver, legacy):
    from cvfes_amd.loads import TractionSchedule
    rng = np.random.default_rng(4)
    n = 64 * 32
    base = str(tmp_path / "wallStress")
    for k in range(6):
        np.save(f"{base}{k}.npy", rng.standard_normal((n, 3)) * 50.0)
    cfg = write_cfg(tmp_path, endtime=endtime(150), dt=repr(DT), Damp="10.0", constant_T="2.0e-4",
                    dt_f="1.2e-4")
    text = open(cfg).read().replace("sample_num = 16", f"sample_num = 16\nexport_stress_filename = '{base}'")
    open(cfg, "w").write(text)
    from cvfes_amd.cvfes import CVFES
    cv = CVFES().ReadInputFile(cfg)
    if legacy:
        from cvfes_amd.physicsSolverGPUs import GPUSolidSolver as Legacy
        from cvfes_amd.solver import TransientSolver

        class LegacyLoop(TransientSolver):
            solid_class = Legacy
        solver = LegacyLoop(cv.comm, cv.meshes, cv.cvConfig.solver)
    else:
        from cvfes_amd.solver import TransientSolverGPU
        solver = TransientSolverGPU(cv.comm, cv.meshes, cv.cvConfig.solver)
    solver.Solve(fast=True)
    mesh = cv.meshes["wall"]
    steps = solver.nTimeSteps
    sched = TractionSchedule(base, False, 2.0e-4, 1.2e-4, n, mesh.boundary, legacy=legacy)
    assert steps == 150
    times = np.arange(steps) * DT
    extra = []
    for t in times:
        _, ga, qa, gb, qb = sched.fields(t)
        f = qa * ga + (0.0 if gb is None else qb * gb)
        extra.append(f.ravel())
    extra = np.array(extra)
    wf = solver.solidSolver.waveform
    om = OracleModel(mesh)
    u_ref = om.run(wf(times), DT, damp=0.0 if legacy else 10.0, extra_forces=extra)
    assert rel_err(solver.solidSolver.GetDisplacement(), u_ref) < REL


def test_save_writes_per_sample_displacements(tmp_path, driver):
    from cvfes_amd import vtkio
    out = str(tmp_path / "disp.vtp")
    cfg = write_cfg(tmp_path, endtime=endtime(50), dt=repr(DT), save_result_num="2",
                    save_stress_filename=f"'{out}'")
    cv = driver().ReadInputFile(cfg)
    solver = cv.Solve()
    path = str(tmp_path / "disp50.vtp")
    assert os.path.exists(path) and os.path.exists(str(tmp_path / "disp25.vtp"))
    back = vtkio.read_vtp(path)
    u = solver.solidSolver.GetDisplacement()
    assert len([k for k in back.point_data if k.startswith("u_")]) == 16
    assert np.array_equal(back.point_data["u_003"].ravel(), u[:, 3])


def test_updated_geometry_solver_and_stress(tmp_path, driver):
    from cvfes_amd import vtkio
    cfg = write_cfg(tmp_path, endtime=endtime(120), dt=repr(DT), stiffness_method="matfree",
                    stiffness_update_interval="2")
    cv = driver().ReadInputFile(cfg)
    solver = cv.Solve(fast=True)
    mesh = cv.meshes["wall"]
    wf = solver.solidSolver.waveform
    p = wf(np.arange(120) * DT)
    om = OracleModel(mesh)
    D = so.d_matrix(mesh.v)
    _, gth = so.gauss_properties(mesh.vE, mesh.vthickness, mesh.elementNodeIds)
    coefK = so.coef_k_element_E(gth, mesh.elmE)
    z = np.zeros(3 * mesh.nNodes)
    u, _, _, nup = so.run_updated(mesh.nodes, mesh.elementNodeIds, D, coefK, om.indptr, om.indices,
                                  om.K, om.LM, om.f1, p, z, z, DT, 2, bdy_dofs=om.bdy_dofs)
    assert nup == 59
    ug = solver.solidSolver.GetDisplacement()
    assert rel_err(ug, u) < REL
    # stresses on the geometry of the last re-formation (the displacement before step 118)
    _, _, snaps, _ = so.run_updated(mesh.nodes, mesh.elementNodeIds, D, coefK, om.indptr, om.indices,
                                    om.K, om.LM, om.f1, p[:118], z, z, DT, 2, bdy_dofs=om.bdy_dofs,
                                    record=(118,))
    S = mesh.nSmp
    moved = mesh.nodes[None] + snaps[118].T.reshape(S, mesh.nNodes, 3)
    ref = so.calculate_stress(moved, mesh.elementNodeIds, D, mesh.elmE, ug)
    st = solver.solidSolver.CalculateStress()
    assert rel_err(st, ref) < 1e-9
    out = str(tmp_path / "wall.vtp")
    solver.solidSolver.SaveStress(out, 7)
    back = vtkio.read_vtp(str(tmp_path / "wall7.vtp"))
    assert np.array_equal(back.cell_data["xz_005"].ravel(), st[:, 5, 3])
    assert np.array_equal(back.point_data["displacement_002"].reshape(-1, 3),
                          ug[:, 2].reshape(-1, 3))


def test_refuses_without_library(monkeypatch, tmp_path):
    from cvfes_amd import _lib
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(_lib.CvfesError):
        _lib.lib()
