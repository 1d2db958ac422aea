"""Stress recovery and the updated-geometry mode on the GPU against the CPU oracle.

* element stresses of a displacement field equal ``calculate_stress`` of the oracle
  (``OptimizedCalculateStress``) on the reference geometry and on a moved geometry, for element
  and node modulus fields (1e-12 relative max-norm);
* the updated-geometry step (stiffness and pressure load re-formed on ``X + u`` every
  ``interval`` steps) follows ``run_updated`` of the oracle (1e-10 relative max-norm) with and
  without Rayleigh and mass-independent damping, at intervals 1, 4 and beyond the run (the
  linear limit), and differs measurably from the linear run;
* the re-formation schedule continues across calls (split runs equal one run, bit for bit);
* the drop-in solver with ``stiffness_update_interval`` and its stress output.
"""
import numpy as np
import pytest
import torch

from cvfes_amd.mesh import tube_geometry
from oracle import solid_oracle as so

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def small_system(S=5, seed=0, node_field=False, nt=16, nz=8, R=0.3):
    from cvfes_amd.system import MembraneSystem
    nodes, el, bdy = tube_geometry(nt, nz, R, 10 * R)
    n = nodes.shape[0]
    rng = np.random.default_rng(seed)
    if node_field:
        vE = 7e6 * np.exp(0.1 * rng.standard_normal((n, S)))
        sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(n, 0.07), vE=vE, density=1.0,
                              nu=0.45, store_values=False)
        gE, _ = so.gauss_properties(vE, np.full((n, S), 0.07), el)
        aveE = gE.mean(axis=2)
    else:
        E = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
        sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(n, 0.07), elmE=E, density=1.0,
                              nu=0.45, store_values=False)
        aveE = E
    return nodes, el, bdy, sysm, aveE


@pytest.mark.parametrize("node_field", [False, True])
def test_stress_matches_oracle(require_gpu, node_field):
    nodes, el, bdy, sysm, aveE = small_system(S=6, node_field=node_field)
    n, S = nodes.shape[0], 6
    rng = np.random.default_rng(3)
    u = 1e-3 * rng.standard_normal((n, 3, S))
    ug = 2e-2 * rng.standard_normal((n, 3, S))
    ud, ugd = torch.as_tensor(u).cuda(), torch.as_tensor(ug).cuda()
    D = so.d_matrix(0.45)
    for geom in (None, ug):
        st = sysm.element_stress(u=ud, ug=None if geom is None else ugd).cpu().numpy()
        moved = nodes[None] + (0.0 if geom is None else np.moveaxis(geom, 2, 0))
        ref = so.calculate_stress(np.broadcast_to(moved, (S, n, 3)), el, D, aveE, u.reshape(-1, S))
        assert rel_err(np.moveaxis(st, 1, 2), ref) < 1e-12


def oracle_updated(nodes, el, bdy, sysm, p, dt, interval, beta, damp, record):
    S = sysm.n_smp
    coef = sysm.coef.cpu().numpy()
    ip, ix = so.sparse_info_loops(el, nodes.shape[0])
    gth = np.full((el.shape[0], S, 3), 0.07)
    D = so.d_matrix(0.45)
    _, K0, f1, LM = so.assemble_reference(nodes, el, 1.0, D, coef, gth, 1.0, ip, ix)
    bd = (3 * bdy[:, None] + np.arange(3)).ravel()
    z = np.zeros(3 * nodes.shape[0])
    return so.run_updated(nodes, el, D, coef, ip, ix, K0, LM, f1, p, z, z, dt, interval, beta=beta,
                          damp=damp, bdy_dofs=bd, record=record)


@pytest.mark.parametrize("interval,beta,damp", [(1, 0.0, 0.0), (4, 2e-7, 0.0), (1, 1e-7, 20.0),
                                                (1000, 0.0, 0.0)])
def test_updated_geometry_matches_oracle(require_gpu, interval, beta, damp):
    nodes, el, bdy, sysm, _ = small_system(S=5, seed=interval)
    nsteps, dt = 80, 4e-6
    p = 1.33e4 * (1 + 0.5 * np.sin(2 * np.pi * np.arange(nsteps) * dt / 1e-4))
    record = (1, 17, 40, 80)
    uo, _, snaps, nup = oracle_updated(nodes, el, bdy, sysm, p, dt, interval, beta, damp, record)
    assert nup == (0 if interval >= nsteps else len(range(interval, nsteps, interval)))
    z = np.zeros(3 * nodes.shape[0])
    load = sysm.loads(p)
    sysm.initial_state(z, z, load, dt, beta=beta, damp=damp, method="matfree")
    done = 0
    for k in record:
        sysm.step_updated(load, dt, interval, beta=beta, damp=damp, offset=done, nsteps=k - done)
        done = k
        ug = sysm.u.cpu().numpy().reshape(-1, 5)
        assert rel_err(ug, snaps[k]) < REL, f"step {k}"
    if interval < nsteps:
        ul, _, _ = so.run(*so.sparse_info_loops(el, nodes.shape[0]),
                          *oracle_linear(nodes, el, sysm), p, z, z, dt, beta=beta, damp=damp,
                          bdy_dofs=(3 * bdy[:, None] + np.arange(3)).ravel())
        assert rel_err(uo, ul) > 1e-6          # the moved geometry matters at this load


@pytest.mark.parametrize("form", ["generic", "rowload1", "rowload"])
@pytest.mark.parametrize("S,interval,beta,damp", [(6, 1, 1e-7, 20.0), (128, 1, 2e-7, 0.0),
                                                  (64, 4, 0.0, 10.0), (7, 1, 0.0, 0.0)])
def test_updated_kernel_forms_match_oracle(require_gpu, form, S, interval, beta, damp):
    """Every form of the updated-geometry kernel (divided frames with one sample per lane;
    reciprocal frames with one or two samples per lane; odd counts fall back to one) against the
    oracle, at packed and wave-uniform rows."""
    nodes, el, bdy, sysm, _ = small_system(S=S, seed=S + interval)
    sysm.set_variant(form)
    nsteps, dt = 60, 4e-6
    p = 1.33e4 * (1 + 0.5 * np.sin(2 * np.pi * np.arange(nsteps) * dt / 1e-4))
    record = (1, 25, 60)
    _, _, snaps, _ = oracle_updated(nodes, el, bdy, sysm, p, dt, interval, beta, damp, record)
    z = np.zeros(3 * nodes.shape[0])
    load = sysm.loads(p)
    sysm.initial_state(z, z, load, dt, beta=beta, damp=damp, method="matfree")
    done = 0
    for k in record:
        sysm.step_updated(load, dt, interval, beta=beta, damp=damp, offset=done, nsteps=k - done)
        done = k
        ug = sysm.u.cpu().numpy().reshape(-1, S)
        assert np.all(np.isfinite(ug))
        assert rel_err(ug, snaps[k]) < REL, f"step {k}"


def oracle_linear(nodes, el, sysm):
    S = sysm.n_smp
    ip, ix = so.sparse_info_loops(el, nodes.shape[0])
    gth = np.full((el.shape[0], S, 3), 0.07)
    _, K0, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(0.45), sysm.coef.cpu().numpy(),
                                          gth, 1.0, ip, ix)
    return K0, LM, f1


def test_updated_schedule_continues_across_calls(require_gpu):
    nodes, el, bdy, sysm, _ = small_system(S=4, seed=2)
    nsteps, dt = 50, 4e-6
    load = sysm.loads(np.full(nsteps, 1.33e4))
    z = np.zeros(3 * nodes.shape[0])
    outs = []
    for split in ((50,), (7, 13, 30)):
        sysm.initial_state(z, z, load, dt, method="matfree")
        done = 0
        for k in split:
            sysm.step_updated(load, dt, 3, offset=done, nsteps=k)
            done += k
        outs.append(sysm.u.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_updated_interval_one_stress_uses_last_geometry(require_gpu):
    nodes, el, bdy, sysm, aveE = small_system(S=3, seed=5)
    nsteps, dt = 30, 4e-6
    load = sysm.loads(np.full(nsteps, 1.33e4))
    z = np.zeros(3 * nodes.shape[0])
    sysm.initial_state(z, z, load, dt, method="matfree")
    sysm.step_updated(load, dt, 1)
    u = sysm.u.cpu().numpy()
    up = sysm.state[1].cpu().numpy()
    st = sysm.element_stress().cpu().numpy()
    moved = nodes[None] + np.moveaxis(up, 2, 0)
    ref = so.calculate_stress(moved, el, so.d_matrix(0.45), aveE, u.reshape(-1, 3))
    assert rel_err(np.moveaxis(st, 1, 2), ref) < 1e-12
