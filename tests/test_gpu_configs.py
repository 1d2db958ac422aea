"""The batching and combined workloads at their configured sizes (GPU).

* Batching: the 64 uneven tubes of ``batch_cases()`` (ntheta 32..512, nz 64..600, 8..256 samples
  per mesh, Sum S = 8,183) advanced by one persistent matrix-free launch per step for 99 steps from
  one captured HIP graph; six seeded meshes equal the same meshes stepped alone with the
  single-mesh kernel bit for bit, and the smallest mesh matches the C oracle on two sample columns.
* Combined: the stenosed tapered 512,000-node tube with 256 samples and beta = 1e-4; the stored
  block-CSR and matrix-free operators agree over 30 steps (1e-12), one sample column matches the C
  oracle, and the adjoint gradient with two checkpoint levels equals the one-level gradient bit for
  bit on a short run.

The moduli are seeded i.i.d. log-normal per element and sample (10 % CoV) instead of GMRF draws:
what is tested is the operators and the batching at these sizes, and the GMRF sampler has its own
tests.  Bars: bit-exact where the same kernel arithmetic runs; 1e-10 relative max-norm against the
oracle; 1e-12 between the two operators.
"""
import numpy as np
import pytest
import torch

from oracle_cases import oracle_trajectory

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def lognormal_E(n_elms, S, seed, E0=7.0e6, cov=0.10):
    var = np.log1p(cov * cov)
    z = np.random.default_rng(seed).standard_normal((n_elms, S))
    return E0 * np.exp(np.sqrt(var) * z - 0.5 * var)


def test_batch_configured_ensemble(require_gpu):
    from cvfes_amd.batch import BatchedEnsemble
    from cvfes_amd.configs import batch_cases
    from cvfes_amd.system import MembraneSystem

    nsteps, block = 99, 99
    cases = batch_cases(byte_budget=float("inf"))   # the configured sample counts
    assert len(cases) == 64 and sum(c.n_smp for c in cases) == 8183
    systems, loads, dts, Es, geos = [], [], [], [], []
    for c in cases:
        nodes, el, bdy = c.geometry()
        E = lognormal_E(el.shape[0], c.n_smp, c.seed)
        s = MembraneSystem(nodes, el, bdy, c.n_smp, vthickness=np.full(nodes.shape[0], c.thickness),
                           elmE=E, density=c.density, nu=c.nu, store_values=False)
        s.set_variant("generic")
        dt = 0.8 * s.critical_dt(iters=30, method="matfree")[0]
        systems.append(s)
        dts.append(dt)
        loads.append(s.loads(c.pressures(0, nsteps, dt)))
        geos.append((nodes, el, bdy))
        Es.append(E if c.n_nodes * c.n_smp < 2e6 else None)
    print(f"batch: {len(systems)} meshes, {sum(s.n_nodes * s.n_smp for s in systems):.3e} "
          f"node-samples", flush=True)
    ens = BatchedEnsemble(systems, loads, dts=dts, methods="matfree")
    ens.initial_state()
    ens.run(nsteps, graph=True, block=block)
    torch.cuda.synchronize()
    batched = [s.u.cpu().numpy() for s in systems]
    assert all(np.all(np.isfinite(u)) for u in batched)

    rng = np.random.default_rng(12)
    subset = sorted(rng.choice(len(systems), 6, replace=False))
    for i in subset:
        s = systems[i]
        z = np.zeros(3 * s.n_nodes)
        s.initial_state(z, z, loads[i], dts[i], method="matfree")
        s.step(loads[i], dts[i], method="matfree", offset=0, nsteps=nsteps)
        alone = s.u.cpu().numpy()
        assert np.max(np.abs(alone)) > 0
        assert np.array_equal(batched[i], alone), i

    # the smallest mesh against the C oracle, two sample columns
    k = min(range(len(systems)), key=lambda j: systems[j].n_nodes * systems[j].n_smp)
    c, s = cases[k], systems[k]
    nodes, el, bdy = geos[k]
    cols = np.array([0, c.n_smp - 1])
    snaps = oracle_trajectory(nodes, el, bdy, Es[k][:, cols], c.pressures(0, nsteps, dts[k]),
                              dts[k], 0.0, c.E0, c.thickness, c.density, c.nu, s.indptr_h,
                              s.indices_h, (nsteps,))
    ug = batched[k][:, :, cols].reshape(3 * s.n_nodes, cols.size)
    assert np.max(np.abs(snaps[0])) > 0
    assert rel_err(ug, snaps[0]) < 1e-10


@pytest.fixture(scope="module")
def combined(require_gpu):
    from cvfes_amd.configs import get_case
    from cvfes_amd.system import MembraneSystem
    case = get_case("stenosed-512x1000-s256")
    nodes, el, bdy = case.geometry()
    N, S = nodes.shape[0], case.n_smp
    E = lognormal_E(el.shape[0], S, case.seed)
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=E,
                          density=case.density, nu=case.nu, store_values=True)
    dt = 0.8 * sysm.critical_dt(beta=case.beta, iters=40, method="matfree")[0]
    print(f"combined: N={N} E={el.shape[0]} nnzb={sysm.nnzb} S={S} dt={dt:.4e}", flush=True)
    yield case, sysm, (nodes, el, bdy, E), dt
    del sysm
    torch.cuda.empty_cache()


def test_combined_operators_and_oracle(combined):
    case, sysm, (nodes, el, bdy, E), dt = combined
    n = 30
    p = case.pressures(0, n, dt)
    load = sysm.loads(p)
    z = np.zeros(3 * sysm.n_nodes)
    out = {}
    for method in ("stored", "matfree"):
        sysm.initial_state(z, z, load, dt, beta=case.beta, method=method)
        sysm.step(load, dt, beta=case.beta, method=method, offset=0, nsteps=n)
        out[method] = sysm.u.cpu().numpy()
    assert np.all(np.isfinite(out["stored"])) and np.max(np.abs(out["stored"])) > 0
    assert rel_err(out["matfree"], out["stored"]) < 1e-12
    col = np.array([101])
    snaps = oracle_trajectory(nodes, el, bdy, E[:, col], p, dt, case.beta, case.E0, case.thickness,
                              case.density, case.nu, sysm.indptr_h, sysm.indices_h, (n,))
    ug = out["stored"][:, :, col].reshape(3 * sysm.n_nodes, 1)
    assert rel_err(ug, snaps[0]) < 1e-10


def test_combined_two_level_checkpoints_bit_identical(combined):
    from cvfes_amd.adjoint import AdjointProblem
    case, sysm, (nodes, el, bdy, E), dt = combined
    sysm.release_values()
    torch.cuda.empty_cache()
    nsteps = 40
    load = sysm.loads(case.pressures(0, nsteps, dt))
    obs = [10, 20, 30, 40]
    S, N = sysm.n_smp, sysm.n_nodes
    g = torch.Generator(device="cuda").manual_seed(5)
    u_obs = 1e-5 * torch.randn((len(obs), N, 3, S), dtype=torch.float64, device="cuda", generator=g)
    theta = torch.as_tensor(np.log(E)).cuda()
    grads = []
    # one level, two levels, and the combined workload's schedule (one level by blocks of 64
    # samples with the gradient steps on their own stream)
    for C, C2, B, conc in ((8, None, None, False), (20, 4, None, False), (8, None, 64, True)):
        prob = AdjointProblem(sysm, load, nsteps, dt, beta=case.beta, method="matfree",
                              obs_steps=obs, u_obs=u_obs, checkpoint_every=C, inner_every=C2,
                              sample_block=B, concurrent=conc)
        prob.set_theta(theta)
        prob.forward()
        grads.append(prob.backward().cpu().numpy())
        del prob
        torch.cuda.empty_cache()
    assert np.all(np.isfinite(grads[0])) and np.max(np.abs(grads[0])) > 0
    for g_other in grads[1:]:
        assert np.array_equal(g_other, grads[0])
