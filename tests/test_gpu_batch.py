"""Batched ensemble (one persistent launch per step over uneven meshes) against per-mesh runs (GPU).

* a mixed ensemble (sample counts 5, 16, 33, 100, 128, 7; both operators; per-mesh dt, beta and
  damp) stepped in one launch per step equals each mesh stepped alone with the generic
  single-mesh kernel, bit for bit, at both lane widths;
* the autotuner (lane widths x grid sizes) leaves the trajectory unchanged;
* graph replays (with eager work queued between the replays) equal eager batched steps, bit for
  bit, over several replays;
* one batched mesh matches the CPU oracle (1e-10).
"""
import numpy as np
import pytest
import torch

from cvfes_amd.mesh import tube_geometry

pytestmark = pytest.mark.gpu

SPECS = [  # ntheta, nz, R, S, method, dt, beta, damp
    (20, 8, 0.3, 5, "stored", 4e-6, 0.0, 0.0),
    (32, 12, 0.4, 16, "matfree", 5e-6, 1e-7, 0.0),
    (24, 30, 0.3, 33, "stored", 3e-6, 0.0, 12.0),
    (48, 16, 0.5, 100, "matfree", 4e-6, 2e-7, 5.0),
    (64, 20, 0.3, 128, "stored", 3e-6, 1e-7, 0.0),
    (16, 40, 0.2, 7, "matfree", 2e-6, 0.0, 0.0),
]


# rows wider than one wavefront: 256 and 200 samples at two per lane (128 lanes, uniform rows)
WIDE = [
    (24, 10, 0.3, 256, "matfree", 3e-6, 1e-7, 0.0),
    (20, 12, 0.4, 200, "stored", 4e-6, 0.0, 10.0),
    (16, 8, 0.3, 16, "matfree", 5e-6, 0.0, 0.0),
]


def make_ensemble(nsteps, seed=0, storage="full", specs=SPECS):
    from cvfes_amd.system import MembraneSystem
    rng = np.random.default_rng(seed)
    systems, loads = [], []
    for (nt, nz, R, S, method, dt, beta, damp) in specs:
        nodes, el, bdy = tube_geometry(nt, nz, R, 10 * R)
        E = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
        s = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(nodes.shape[0], 0.07), elmE=E,
                           density=1.0, nu=0.5, store_values=(method == "stored"), storage=storage)
        s.set_variant("generic")
        systems.append(s)
        t = np.arange(nsteps) * dt
        loads.append(s.loads(1.33e4 * (1 + 0.1 * np.sin(2 * np.pi * t / 0.8))))
    return systems, loads


def states(systems):
    return [s.u.cpu().numpy().copy() for s in systems]


@pytest.mark.parametrize("vw_hint,cap,waves,storage", [
    (0, False, 0, "full"), (1, False, 0, "full"), (2, False, 0, "full"), (2, True, 0, "full"),
    (0, True, 0, "full"), (2, False, 4, "full"), (0, True, 7, "full"), (0, False, 0, "sym"),
    (1, True, 0, "sym")])
def test_batched_equals_per_mesh(require_gpu, vw_hint, cap, waves, storage):
    """Every lane-width policy (two samples per lane where even / one everywhere / two only on
    wave-uniform rows), with or without the register and waves-per-SIMD caps, and the stored
    meshes in full or symmetric storage, gives the per-mesh result bit for bit."""
    from cvfes_amd.batch import BatchedEnsemble
    nsteps = 40
    systems, loads = make_ensemble(nsteps, storage=storage)
    cols = list(zip(*SPECS))
    ens = BatchedEnsemble(systems, loads, dts=list(cols[5]), methods=list(cols[4]),
                          betas=list(cols[6]), damps=list(cols[7]))
    ens.vw_hint, ens.cap, ens.waves_per_simd = vw_hint, cap, waves
    ens.initial_state()
    assert ens.kinds == {0: 31, 1: 14, 2: 15}[vw_hint]
    ens.run(nsteps, graph=False)
    batched = states(systems)
    ens.initial_state()
    ens.run_per_mesh(nsteps)
    single = states(systems)
    for a, b in zip(batched, single):
        assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
        assert np.array_equal(a, b)


@pytest.mark.parametrize("storage", ["full", "sym"])
def test_sample_group_order_equals_per_mesh(require_gpu, storage):
    """Rows two wavefronts wide (256 and 200 samples) visited sample group first (every row for
    the first 128 samples, then the rest) give the per-mesh result bit for bit, eagerly and in
    graph replays."""
    from cvfes_amd.batch import BatchedEnsemble
    nsteps = 36
    systems, loads = make_ensemble(nsteps, seed=7, storage=storage, specs=WIDE)
    cols = list(zip(*WIDE))
    ens = BatchedEnsemble(systems, loads, dts=list(cols[5]), methods=list(cols[4]),
                          betas=list(cols[6]), damps=list(cols[7]))
    ens.initial_state()
    assert [int(m.lanes) for m in ens.host_meshes] == [128, 128, 8]
    assert [int(m.packed) for m in ens.host_meshes] == [0, 0, 1]    # 16 samples: packed rows
    ens.run_per_mesh(nsteps)
    single = states(systems)
    for graph in (False, True):
        ens.groups_first = True
        ens.initial_state()
        ens.run(nsteps, graph=graph, block=12)
        for a, b in zip(states(systems), single):
            assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
            assert np.array_equal(a, b), graph


def test_graph_replays_equal_eager(require_gpu):
    from cvfes_amd.batch import BatchedEnsemble
    nsteps = 3 * 33 + 5
    systems, loads = make_ensemble(2 * nsteps, seed=1)
    cols = list(zip(*SPECS))
    ens = BatchedEnsemble(systems, loads, dts=list(cols[5]), methods=list(cols[4]),
                          betas=list(cols[6]), damps=list(cols[7]))
    ens.initial_state()
    ens.run(nsteps, graph=False)
    eager = states(systems)
    ens.initial_state()
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    done = 0
    for _ in range(3):                  # three graph replays of 33 steps with eager work between
        ens.run(33, graph=True, block=33)
        scratch.add_(1.0)
        torch.cuda.synchronize()
        done += 33
    ens.run(nsteps - done, graph=False)
    graphed = states(systems)
    for a, b in zip(graphed, eager):
        assert np.array_equal(a, b)
    # a further full replay continues the same trajectory as eager steps
    ens.run(99, graph=True, block=99)
    g2 = states(systems)
    ens.initial_state()
    ens.run(nsteps + 99, graph=False)
    for a, b in zip(g2, states(systems)):
        assert np.array_equal(a, b)


def test_batched_mesh_matches_oracle(require_gpu):
    from cvfes_amd.batch import BatchedEnsemble
    from oracle import solid_oracle as so
    nsteps = 60
    systems, loads = make_ensemble(nsteps, seed=2)
    cols = list(zip(*SPECS))
    ens = BatchedEnsemble(systems, loads, dts=list(cols[5]), methods=list(cols[4]),
                          betas=list(cols[6]), damps=list(cols[7]))
    ens.initial_state()
    ens.run(nsteps, graph=True, block=30)
    k = 3                               # 48x16 tube, 100 samples, matrix-free, beta and damp
    nt, nz, R, S, method, dt, beta, damp = SPECS[k]
    nodes, el, bdy = tube_geometry(nt, nz, R, 10 * R)
    n = nodes.shape[0]
    coef = systems[k].coef.cpu().numpy()
    ip, ix = so.sparse_info_loops(el, n)
    gth = np.full((el.shape[0], S, 3), 0.07)
    M, K, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(0.5), coef, gth, 1.0, ip, ix)
    p = 1.33e4 * (1 + 0.1 * np.sin(2 * np.pi * np.arange(nsteps) * dt / 0.8))
    bd = (3 * bdy[:, None] + np.arange(3)).ravel()
    u, _, _ = so.run(ip, ix, K, LM, f1, p, np.zeros(3 * n), np.zeros(3 * n), dt, beta=beta, damp=damp,
                     bdy_dofs=bd)
    ug = systems[k].u.cpu().numpy().reshape(-1, S)
    assert np.max(np.abs(ug - u)) / np.max(np.abs(u)) < 1e-10


def test_gmrf_device_solves_match_host(require_gpu):
    """GMRF samples with the probe and sample solves on the GPU equal the host solves (CG
    tolerance), on a mesh large enough for the stochastic mean-diagonal estimate."""
    from cvfes_amd.gmrf import sample_element_E
    nodes, el, bdy = tube_geometry(64, 40, 0.3, 3.0)
    assert el.shape[0] > 4000
    host = sample_element_E(el, 8, seed=5)
    dev = sample_element_E(el, 8, seed=5, device="cuda")
    assert abs(dev.tau - host.tau) <= 1e-10 * host.tau
    assert np.max(np.abs(dev.z - host.z)) <= 1e-9 * np.max(np.abs(host.z))


def test_autotune_keeps_the_trajectory(require_gpu):
    from cvfes_amd.batch import BatchedEnsemble
    nsteps = 3 * 11 + 40
    systems, loads = make_ensemble(nsteps, seed=3)
    cols = list(zip(*SPECS))
    ens = BatchedEnsemble(systems, loads, dts=list(cols[5]), methods=list(cols[4]),
                          betas=list(cols[6]), damps=list(cols[7]))
    ens.initial_state()
    tuned = ens.autotune(candidates=[256, 2048], steps=1)
    assert tuned["n_waves"] in (256, 2048)
    done = ens.step_index
    ens.run(nsteps - done, graph=True, block=33)
    tuned_states = states(systems)
    ens.vw_hint, ens.n_waves = 0, 0
    ens.initial_state()
    ens.run(nsteps, graph=False)
    for a, b in zip(tuned_states, states(systems)):
        assert np.array_equal(a, b)
