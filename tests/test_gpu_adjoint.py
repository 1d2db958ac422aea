"""Adjoint gradient d L / d log E (GPU) against the FP64 torch CPU reference and finite differences.

* small tube: the GPU adjoint equals torch autograd through the same recurrence (1e-9 relative),
  for both operators, with Rayleigh beta, damping, a non-zero initial displacement (gradient
  through u_{-1}) and a checkpoint interval that does not divide the step count;
* the gradient does not depend on the checkpoint interval (bit-exact: recomputation replays the
  same arithmetic);
* 49,152-node tube, 32 samples, 2,000 steps, 20 observation times, checkpoints every 100 steps:
  central finite differences on 20 seeded (element, sample) pairs.  The loss difference is formed
  from the states, sum 1/2 (u+ - u-)(u+ + u- - 2 u_obs), not as L(theta+h) - L(theta-h).  In FP64
  two 2,000-step runs that differ by a step of 1e-6 in one element's log-modulus carry a rounding
  noise of about 1e-10 in the difference quotient (measured; independent of the component and
  falling as 1/step), so at step 1e-6 the check is relative to the gradient's max-norm (< 1e-5);
  the per-component 1e-6 check uses Richardson-extrapolated differences at step 1e-3.
"""
import numpy as np
import pytest
import torch

from cvfes_amd.mesh import tube_geometry

pytestmark = pytest.mark.gpu


def build(nt, nz, R, L, S, seed, store=True):
    from cvfes_amd.system import MembraneSystem
    nodes, el, bdy = tube_geometry(nt, nz, R, L)
    rng = np.random.default_rng(seed)
    theta = np.log(7e6) + 0.1 * rng.standard_normal((el.shape[0], S))
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(nodes.shape[0], 0.07),
                          elmE=np.exp(theta), density=1.0, nu=0.5, store_values=store)
    return nodes, el, bdy, theta, sysm


def sine(n, dt):
    return 1.33e4 * (1 + 0.1 * np.sin(2 * np.pi * np.arange(n) * dt / 0.8))


@pytest.mark.parametrize("method,beta,damp,C,u0,S", [
    ("stored", 0.0, 0.0, 100, False, 4),
    ("matfree", 2e-7, 30.0, 7, False, 4),
    ("stored", 2e-7, 0.0, 13, True, 4),
    ("matfree", 0.0, 25.0, 1, True, 4),
    ("stored", 1e-7, 10.0, 20, False, 100),
])
def test_adjoint_matches_torch_autograd(require_gpu, method, beta, damp, C, u0, S):
    from cvfes_amd.adjoint import AdjointProblem, EnsembleLoss
    from oracle.torch_adjoint import TorchEnsemble
    nodes, el, bdy, theta, sysm = build(16, 6, 0.3, 1.0, S, 3)
    nsteps, dt = 60, 4e-6
    p = sine(nsteps, dt)
    obs = [12, 24, 36, 48, 60]
    rng = np.random.default_rng(5)
    u0v = None
    if u0:
        u0v = 1e-5 * rng.standard_normal(3 * nodes.shape[0])
    ref = TorchEnsemble(nodes, el, bdy, 0.07, 1.0, 0.5)
    # observations: the same model under another modulus field
    theta_obs = torch.as_tensor(theta + 0.05 * rng.standard_normal(theta.shape))
    with torch.no_grad():
        u_obs = ref.states_at(theta_obs, p, dt, beta, damp, obs, u0=u0v)
    th = torch.as_tensor(theta).clone().requires_grad_(True)
    Lref = ref.loss(th, p, dt, beta, damp, obs, u_obs, u0=u0v)
    wts = torch.as_tensor(np.resize([1.0, 0.5, 2.0, -1.0], S))
    (Lref * wts).sum().backward()
    gref = th.grad.numpy()

    prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=beta, damp=damp, method=method,
                          obs_steps=obs, u_obs=u_obs.to("cuda"), checkpoint_every=C, u0=u0v)
    thd = torch.as_tensor(theta).cuda().requires_grad_(True)
    Lgpu = EnsembleLoss.apply(thd, prob)
    assert np.allclose(Lgpu.detach().cpu().numpy(), Lref.detach().numpy(), rtol=1e-10, atol=0)
    (Lgpu * wts.cuda()).sum().backward()
    g = thd.grad.cpu().numpy()
    assert np.max(np.abs(gref)) > 0
    assert np.max(np.abs(g - gref)) / np.max(np.abs(gref)) < 1e-9


def test_gradient_independent_of_checkpoint_interval(require_gpu):
    from cvfes_amd.adjoint import AdjointProblem
    nodes, el, bdy, theta, sysm = build(24, 10, 0.3, 1.5, 8, 11)
    nsteps, dt = 90, 4e-6
    p = sine(nsteps, dt)
    obs = [30, 60, 90]
    u_obs = torch.zeros((3, nodes.shape[0], 3, 8), dtype=torch.float64, device="cuda")
    grads = []
    for C in (1, 10, 90, 1000):
        prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, method="stored",
                              obs_steps=obs, u_obs=u_obs, checkpoint_every=C)
        prob.set_theta(torch.as_tensor(theta).cuda())
        prob.forward()
        grads.append(prob.backward().cpu().numpy())
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def test_differentiation_workload_finite_differences(require_gpu):
    """49,152-node tube, 32 samples, 2,000 steps, 20 observation times, checkpoints every 100."""
    import time
    from cvfes_amd.adjoint import AdjointProblem
    from cvfes_amd.configs import get_case
    from cvfes_amd.system import MembraneSystem
    case = get_case("tube-192x256-s32")
    nodes, el, bdy = case.geometry()
    N, S, E = nodes.shape[0], case.n_smp, el.shape[0]
    theta = np.log(case.sample_E(el).E)
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=np.exp(theta),
                          density=case.density, nu=case.nu)
    nsteps, dt = case.nsteps, case.dt
    load = sysm.loads(case.pressures(0, nsteps))
    obs = list(range(100, nsteps + 1, 100))
    # observations: the same model under a second GMRF modulus field
    from cvfes_amd.gmrf import sample_element_E
    theta_obs = np.log(sample_element_E(el, S, seed=case.seed + 100).E)
    sysm.set_log_modulus(torch.as_tensor(theta_obs).cuda())
    u0 = np.zeros(3 * N)
    sysm.initial_state(u0, u0, load, dt)
    states, done = [], 0
    for k in obs:
        sysm.step(load, dt, offset=done, nsteps=k - done)
        done = k
        states.append(sysm.u.clone())
    u_obs = torch.stack(states)
    prob = AdjointProblem(sysm, load, nsteps, dt, method="stored", obs_steps=obs, u_obs=u_obs,
                          checkpoint_every=100)
    th = torch.as_tensor(theta).cuda()
    prob.set_theta(th)
    prob.forward()                      # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = prob.forward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g = prob.backward().cpu().numpy()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"forward {t1 - t0:.3f} s, adjoint {t2 - t1:.3f} s, ratio {(t2 - t1) / (t1 - t0):.2f}")
    rng = np.random.default_rng(20)
    picks = [(int(rng.integers(E)), int(rng.integers(S))) for _ in range(20)]

    def observed(e, s, dtheta):
        tp = theta.copy()
        tp[e, s] += dtheta
        prob.set_theta(torch.as_tensor(tp).cuda())
        prob.forward(record=True)
        return [u[:, :, s] for u in prob.obs_states]

    def dquot(e, s, h):
        up_, um_ = observed(e, s, h), observed(e, s, -h)
        return sum(float((0.5 * (a - b) * (a + b - 2.0 * o[:, :, s])).sum())
                   for a, b, o in zip(up_, um_, u_obs)) / (2 * h)

    gmax = max(abs(g[e, s]) for e, s in picks)
    worst = max(abs(dquot(e, s, 1e-6) - g[e, s]) for e, s in picks) / gmax
    print(f"central differences, step 1e-6: max |fd - adjoint| / max|adjoint| = {worst:.2e}")
    assert worst < 1e-5
    h = 1e-3
    worst4 = 0.0
    for e, s in picks:
        fd4 = (4.0 * dquot(e, s, h) - dquot(e, s, 2 * h)) / 3.0
        rel = abs(fd4 - g[e, s]) / abs(g[e, s])
        worst4 = max(worst4, rel)
        print(f"  element {e} sample {s}: adjoint {g[e, s]:.12e} fd4 {fd4:.12e} rel {rel:.1e}")
    print(f"fourth-order differences, step 1e-3: max relative difference {worst4:.2e}")
    assert worst4 <= 1e-6
