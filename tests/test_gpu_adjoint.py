"""Adjoint gradient d L / d log E (GPU) against the FP64 torch CPU reference and finite differences.

* small tube: the GPU adjoint equals torch autograd through the same recurrence (1e-9 relative),
  for both operators, with Rayleigh beta, damping, a non-zero initial displacement (gradient
  through u_{-1}) and a checkpoint interval that does not divide the step count; sample counts
  4, 100 and 64 cover the three gradient-accumulation kernels (unit blocks read per lane, staged
  in LDS, read as scalars);
* the gradient does not depend on the checkpoint scheme (one or two levels, observations on the
  device or in host memory, recomputation fused with the adjoint steps or not; bit-exact:
  recomputation replays the same arithmetic);
* 49,152-node tube, 32 samples, 2,000 steps, 20 observation times, checkpoints every 100 steps:
  central finite differences with step 1e-6 on 20 seeded (element, sample) pairs, each within
  1e-6 relative of the adjoint.  Two separate 2,000-step FP64 runs at theta +- h would carry a
  rounding noise of about 1e-10 in their difference quotient (about 4e-6 of the gradient's
  max-norm at h = 1e-6), so the pair is advanced as its mean u_bar = (u+ + u-)/2 and difference
  d = u+ - u-, which obey (exactly, in real arithmetic) the coupled recurrences
      M(u_bar'' ) = f - C(u_bar') - K_bar x_bar - dK x_d / 4,   M(d'') = -C(d') - K_bar x_d - dK x_bar
  with K_bar = (K+ + K-)/2 (the element coefficient scaled by cosh h), dK = K+ - K- (the element
  block times 2 sinh h) and x = u + (beta/dt)(u - u_prev); the loss difference is
  L+ - L- = sum_k d_k . (u_bar_k - u_obs,k).  The difference d is then formed without
  cancellation.  All 20 pairs run together as 20 sample columns of one system.
"""
import numpy as np
import pytest
import torch

from cvfes_amd.mesh import tube_geometry

pytestmark = pytest.mark.gpu


def build(nt, nz, R, L, S, seed, store=True, storage="full"):
    from cvfes_amd.system import MembraneSystem
    nodes, el, bdy = tube_geometry(nt, nz, R, L)
    rng = np.random.default_rng(seed)
    theta = np.log(7e6) + 0.1 * rng.standard_normal((el.shape[0], S))
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(nodes.shape[0], 0.07),
                          elmE=np.exp(theta), density=1.0, nu=0.5, store_values=store,
                          storage=storage)
    return nodes, el, bdy, theta, sysm


def method_storage(method):
    """``stored-sym``: the stored operator kept as upper blocks only."""
    return ("stored", "sym") if method == "stored-sym" else (method, "full")


def sine(n, dt):
    return 1.33e4 * (1 + 0.1 * np.sin(2 * np.pi * np.arange(n) * dt / 0.8))


@pytest.mark.parametrize("method,beta,damp,C,u0,S", [
    ("stored", 0.0, 0.0, 100, False, 4),
    ("matfree", 2e-7, 30.0, 7, False, 4),
    ("stored", 2e-7, 0.0, 13, True, 4),
    ("matfree", 0.0, 25.0, 1, True, 4),
    ("stored", 1e-7, 10.0, 20, False, 100),
    ("matfree", 1e-7, 0.0, 9, True, 64),
    ("stored-sym", 1e-7, 10.0, 11, True, 4),
    ("stored-sym", 2e-7, 0.0, 20, False, 128),
])
def test_adjoint_matches_torch_autograd(require_gpu, method, beta, damp, C, u0, S):
    from cvfes_amd.adjoint import AdjointProblem, EnsembleLoss
    from oracle.torch_adjoint import TorchEnsemble
    method, storage = method_storage(method)
    nodes, el, bdy, theta, sysm = build(16, 6, 0.3, 1.0, S, 3, storage=storage)
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


@pytest.mark.parametrize("method,S", [("stored", 8), ("matfree", 8), ("stored", 5), ("matfree", 66),
                                      ("stored-sym", 8), ("stored-sym", 5), ("matfree", 128)])
def test_fused_reverse_sweep_is_bit_identical(require_gpu, method, S):
    """The pipelined sweep (recomputation of a segment fused with the adjoint steps of the next)
    gives the gradient of the separate kernels bit for bit, with beta, damping, an initial
    displacement, observations on both kinds of step and checkpoint intervals that do not
    divide the step count, with one and with two checkpoint levels, also with the kernels capped
    at 4 waves per SIMD; and it does fuse steps."""
    from cvfes_amd.adjoint import AdjointProblem
    method, storage = method_storage(method)
    nodes, el, bdy, theta, sysm = build(20, 9, 0.3, 1.2, S, 7, storage=storage)
    nsteps, dt = 75, 4e-6
    p = sine(nsteps, dt)
    obs = [8, 30, 31, 64, 75]
    rng = np.random.default_rng(2)
    u_obs = torch.as_tensor(1e-5 * rng.standard_normal((len(obs), nodes.shape[0], 3, S))).cuda()
    u0 = 1e-6 * rng.standard_normal(3 * nodes.shape[0])
    grads = []
    for C, C2, n_fused in ((11, None, 59), (30, 7, 47)):      # fused steps of 75 steps
        for fuse, waves in ((False, 0), (True, 0), (True, 4)):
            sysm.set_occupancy(waves)
            prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, damp=15.0,
                                  method=method, obs_steps=obs, u_obs=u_obs, checkpoint_every=C,
                                  inner_every=C2, u0=u0, fuse=fuse)
            prob.set_theta(torch.as_tensor(theta).cuda())
            prob.forward()
            grads.append(prob.backward().cpu().numpy())
            assert prob.fused() == fuse
            assert prob.stats["fused_steps"] == (n_fused if fuse else 0)
    sysm.set_occupancy(0)
    assert np.max(np.abs(grads[0])) > 0
    for g in grads[1:]:
        assert np.array_equal(grads[0], g)


@pytest.mark.parametrize("method,S", [("matfree", 128), ("stored-sym", 128), ("matfree", 64)])
def test_node_mass_scales_match_per_sample_masses(require_gpu, method, S):
    """The adjoint and gradient kernels scale by the inverse masses, read per node or (with the
    same masses stored per sample) per sample: loss and gradient are the same bit for bit, for
    separate and fused sweeps.  Without the inverse masses (the kernels divide by the masses)
    they agree to rounding."""
    from cvfes_amd.adjoint import AdjointProblem
    method, storage = method_storage(method)
    nodes, el, bdy, theta, sysm = build(20, 9, 0.3, 1.2, S, 7, storage=storage)
    nsteps, dt = 40, 4e-6
    p = sine(nsteps, dt)
    obs = [8, 25, 40]
    rng = np.random.default_rng(4)
    u_obs = torch.as_tensor(1e-5 * rng.standard_normal((len(obs), nodes.shape[0], 3, S))).cuda()
    u0 = 1e-6 * rng.standard_normal(3 * nodes.shape[0])
    out = []
    for per_sample in (False, True):
        if per_sample:
            sysm.lm = sysm.lm[:, None].expand(sysm.n_nodes, S).contiguous()
            sysm.minv = sysm.minv[:, None].expand(sysm.n_nodes, S).contiguous()
            sysm.vthick = sysm.vthick[:, None].expand(sysm.n_nodes, S).contiguous()
            sysm.thick_sstride = 1
            sysm._sys = sysm._make_struct()
        for fuse in (False, True):
            prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, damp=15.0,
                                  method=method, obs_steps=obs, u_obs=u_obs, checkpoint_every=9,
                                  u0=u0, fuse=fuse)
            prob.set_theta(torch.as_tensor(theta).cuda())
            loss = prob.forward().cpu().numpy()
            out.append((loss, prob.backward().cpu().numpy()))
    assert np.max(np.abs(out[0][1])) > 0
    for loss, g in out[1:]:
        assert np.array_equal(out[0][0], loss)
        assert np.array_equal(out[0][1], g)
    sysm.minv = None
    sysm._sys = sysm._make_struct()
    prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, damp=15.0, method=method,
                          obs_steps=obs, u_obs=u_obs, checkpoint_every=9, u0=u0, fuse=False)
    prob.set_theta(torch.as_tensor(theta).cuda())
    assert np.array_equal(out[0][0], prob.forward().cpu().numpy())
    g = prob.backward().cpu().numpy()
    assert np.max(np.abs(g - out[0][1])) <= 1e-13 * np.max(np.abs(out[0][1]))


@pytest.mark.parametrize("S,B", [(10, 4), (128, 64)])
def test_sample_blocks_and_host_checkpoints_are_bit_identical(require_gpu, S, B):
    """The reverse sweep by sample blocks (blocks of B samples, the last one shorter when B does
    not divide S), with the checkpoints on the device or in host memory, with host checkpoints
    for the whole ensemble, with the gradient steps on a stream beside the adjoint steps, and
    block by block with the pipelined (fused) sweep, gives the loss and gradient of the plain
    one-level sweep bit for bit (observations on steps inside segments and on the last step, an
    initial displacement, beta and damping)."""
    from cvfes_amd.adjoint import AdjointProblem
    nodes, el, bdy, theta, sysm = build(20, 9, 0.3, 1.2, S, 5, store=False)
    nsteps, dt = 50, 4e-6
    p = sine(nsteps, dt)
    obs = [7, 20, 21, 44, 50]
    rng = np.random.default_rng(6)
    u_obs = torch.as_tensor(1e-5 * rng.standard_normal((len(obs), nodes.shape[0], 3, S))).pin_memory()
    u0 = 1e-6 * rng.standard_normal(3 * nodes.shape[0])
    wts = torch.as_tensor(np.resize([1.0, 0.5, 2.0, -1.0], S)).cuda()
    out = []
    for blk, host, conc, fuse in ((None, False, False, False), (B, False, False, False),
                                  (B, True, False, False), (None, True, False, False),
                                  (None, False, True, False), (B, False, True, False),
                                  (B, False, False, True)):
        prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, damp=15.0,
                              method="matfree", obs_steps=obs, u_obs=u_obs, checkpoint_every=9,
                              u0=u0, fuse=fuse, sample_block=blk, host_checkpoints=host,
                              concurrent=conc)
        prob.set_theta(torch.as_tensor(theta).cuda())
        loss = prob.forward().cpu().numpy()
        out.append((loss, prob.backward(wts).cpu().numpy()))
        assert (prob.stats["fused_steps"] > 0) == fuse
    # observations on the device: the blocks' copies of them are made on the helper stream into
    # freshly allocated buffers, which must wait for the work queued on the main stream
    u_dev = u_obs.cuda()
    for fuse in (False, True):
        prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, damp=15.0,
                              method="matfree", obs_steps=obs, u_obs=u_dev, checkpoint_every=9,
                              u0=u0, fuse=fuse, sample_block=B)
        prob.set_theta(torch.as_tensor(theta).cuda())
        loss = prob.forward().cpu().numpy()
        out.append((loss, prob.backward(wts).cpu().numpy()))
    assert np.max(np.abs(out[0][1])) > 0
    for loss, g in out[1:]:
        assert np.array_equal(out[0][0], loss)
        assert np.array_equal(out[0][1], g)


def test_gradient_independent_of_checkpoint_interval(require_gpu):
    from cvfes_amd.adjoint import AdjointProblem
    nodes, el, bdy, theta, sysm = build(24, 10, 0.3, 1.5, 8, 11)
    nsteps, dt = 90, 4e-6
    p = sine(nsteps, dt)
    obs = [30, 60, 90]
    u_obs = torch.zeros((3, nodes.shape[0], 3, 8), dtype=torch.float64, device="cuda")
    u_obs[1] = 1e-4                       # a non-zero observation, also kept in host memory
    u_host = u_obs.cpu().pin_memory()
    grads = []
    for C, C2, host in ((1, None, False), (10, None, False), (90, None, False), (1000, None, False),
                        (30, 7, False), (45, 10, True), (90, 1, True)):
        prob = AdjointProblem(sysm, sysm.loads(p), nsteps, dt, beta=1e-7, method="stored",
                              obs_steps=obs, u_obs=u_host if host else u_obs, checkpoint_every=C,
                              inner_every=C2)
        prob.set_theta(torch.as_tensor(theta).cuda())
        prob.forward()
        grads.append(prob.backward().cpu().numpy())
    assert np.max(np.abs(grads[0])) > 0
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
    fd = paired_central_differences(case, nodes, el, bdy, theta, picks, 1e-6, load_p=case.pressures(0, nsteps),
                                    dt=dt, nsteps=nsteps, obs=obs, u_obs=u_obs)
    worst = 0.0
    for k, (e, s) in enumerate(picks):
        rel = abs(fd[k] - g[e, s]) / abs(g[e, s])
        worst = max(worst, rel)
        print(f"  element {e} sample {s}: adjoint {g[e, s]:.12e} fd {fd[k]:.12e} rel {rel:.1e}")
    print(f"central differences, step 1e-6: max relative difference {worst:.2e}")
    assert worst <= 1e-6


def paired_central_differences(case, nodes, el, bdy, theta, picks, h, load_p, dt, nsteps, obs, u_obs,
                               beta=0.0, damp=0.0):
    """(L(theta + h e_k) - L(theta - h e_k)) / 2h for every pick k = (element, sample), with the
    +-h pair advanced as mean and difference (module docstring); all picks in one system whose
    sample column k is a copy of sample s_k."""
    from cvfes_amd.system import MembraneSystem
    N, P = nodes.shape[0], len(picks)
    ek = np.array([e for e, _ in picks])
    sk = np.array([s for _, s in picks])
    th_bar = theta[:, sk].copy()
    th_bar[ek, np.arange(P)] += np.log(np.cosh(h))
    sb = MembraneSystem(nodes, el, bdy, P, vthickness=np.full(N, case.thickness), elmE=np.exp(th_bar),
                        density=case.density, nu=case.nu, store_values=False)
    dev = sb.device
    # dK restricted to each pick's element: 2 sinh(h) coef[e, s] * khat blocks (a, b) of element e
    cpos = sb.cpos.long().reshape(-1, 9)[torch.as_tensor(ek, device=dev)]        # [P, 9]
    kb = sb.khat[cpos.reshape(-1)].reshape(P, 3, 3, 3, 3)                         # [P, a, b, i, j]
    coef = sb.coef[torch.as_tensor(ek, device=dev), torch.arange(P, device=dev)] / np.cosh(h)
    kb = kb * (2.0 * np.sinh(h) * coef)[:, None, None, None, None]
    idx = torch.as_tensor(el[ek], device=dev).long()                              # [P, 3] nodes
    cols = torch.arange(P, device=dev)

    def dk(x):                                                                    # dK x, x [N, 3, P]
        xs = x.permute(0, 2, 1)[idx, cols[:, None]]                               # [P, b, j]
        out = torch.einsum("pabij,pbj->pai", kb, xs)                              # [P, a, i]
        y = torch.zeros((N, P, 3), dtype=x.dtype, device=dev)
        y[idx, cols[:, None]] = out
        return y.permute(0, 2, 1)

    lm = sb.lm
    lm_b = lm[:, None, :] if lm.ndim == 2 else lm[:, None, None]
    nz = lm_b != 0
    safe = torch.where(nz, lm_b, torch.ones_like(lm_b))

    def solve(r):                                                                 # r / M, 0 where M = 0
        return torch.where(nz, r / safe, torch.zeros_like(r))

    free = (sb.fixed == 0)[:, None, None].to(torch.float64)
    f1 = sb.f1[:, :, None]
    dt2, c, dd = dt * dt, beta / dt, damp / dt
    p = np.asarray(load_p, np.float64)
    zero = torch.zeros((N, 3, P), dtype=torch.float64, device=dev)
    # u0 = du0 = 0: u_bar_-1 = dt^2/2 f0 / M, d_-1 = 0 (unprojected, as the forward step)
    ub, ubp = zero.clone(), solve(0.5 * dt2 * p[0] * f1 + zero)
    d, dp = zero.clone(), zero.clone()
    obs_at = {k: i for i, k in enumerate(obs)}
    dL = torch.zeros(P, dtype=torch.float64, device=dev)
    uo = u_obs[:, :, :, torch.as_tensor(sk, device=u_obs.device)]
    for n in range(nsteps):
        xb = ub + c * (ub - ubp)
        xd = d + c * (d - dp)
        kxb = sb.spmv(xb, "matfree")
        kxd = sb.spmv(xd, "matfree")
        rb = dt2 * (p[n] * f1 - dd * (ub - ubp) - kxb - 0.25 * dk(xd)) + lm_b * (2.0 * ub - ubp)
        rd = dt2 * (-dd * (d - dp) - kxd - dk(xb)) + lm_b * (2.0 * d - dp)
        ubn = solve(rb) * free + sb.bdy_u * (1.0 - free)
        dn = solve(rd) * free
        ubp, ub, dp, d = ub, ubn, d, dn
        if n + 1 in obs_at:
            dL += (d * (ub - uo[obs_at[n + 1]])).sum(dim=(0, 1))
    return (dL / (2.0 * h)).cpu().numpy()
