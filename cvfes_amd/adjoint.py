"""Gradient of a displacement-misfit loss with respect to per-element log Young's modulus.

``AdjointProblem`` runs the forward ensemble step with checkpoints every ``checkpoint_every``
steps, evaluates the per-sample loss

    L_s = 1/2 sum_{k in obs} || u_s(t_k) - u_obs,s(t_k) ||^2,

and runs the reverse sweep: for each segment between checkpoints the forward states are recomputed
from the checkpoint, then ``cvfes_adjoint_step`` (the transposed step) and ``cvfes_grad_accum``
(the element-wise derivative of the stiffness) run backwards through the segment.  Memory is one
pair of states per checkpoint plus one segment of states.  With ``inner_every = C2`` the sweep is
two-level: a segment of ``C`` steps is first recomputed keeping a pair every ``C2`` steps, and
only ``C2`` states are held at a time, so memory falls from ``2 n/C + C`` to
``2 n/C + 2 C/C2 + C2`` states for one more forward pass.  The innermost segments can be
pipelined: the segment before the current one is recomputed during the current one's adjoint
steps, a recomputed step and an adjoint step per fused kernel, in the buffers the current one
releases.  Observations may live in host memory (copied to the device when used).  Recomputation
replays the same arithmetic, so the gradient does not depend on the checkpoint scheme.

Samples do not interact, so the reverse sweep can also run sample block by sample block
(``sample_block``): each block's recomputed segment then takes ``B/S`` of the memory, which lets a
one-level scheme fit where the whole ensemble would need two levels (one forward pass less).  The
checkpoints (one pair per segment and block) can be kept in host memory (``host_checkpoints``):
a helper thread copies them out during the forward pass and back one pair ahead during the
reverse sweep, on a stream of its own, while the main thread keeps the GPU busy.  ``EnsembleLoss`` wraps it as a ``torch.autograd.Function`` of
``theta = log E [E, S]``.

All device work is in the native kernels; torch provides the buffers, the per-sample loss
reduction and the autograd plumbing.
"""
from __future__ import annotations

import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from ._lib import VARIANT_GENERIC, check, ptr
from .system import LoadTables, MembraneSystem

__all__ = ["AdjointProblem", "EnsembleLoss"]


_HELPER = None


def _helper():
    """The helper thread (one for the process) that copies checkpoints and observations while
    the main thread keeps queuing kernels."""
    global _HELPER
    if _HELPER is None:
        _HELPER = ThreadPoolExecutor(max_workers=1, thread_name_prefix="cvfes-copy")
    return _HELPER


class _ObsStream:
    """The observations restricted to samples a..b-1 on the device, in the order a reverse sweep
    takes them (last step first): the problem's helper thread copies them into a few buffers in
    rotation, each one once the kernel that read the observation it held has been issued."""

    def __init__(self, prob, a, b, depth=3):
        self.p, self.a, self.b = prob, a, b
        self.order = sorted(prob.obs_steps, reverse=True)
        self.pos = {k: i for i, k in enumerate(self.order)}
        n = min(depth, len(self.order))
        shape = (prob.sys.n_nodes, 3, b - a)
        self.ring = [torch.empty(shape, dtype=torch.float64, device=prob.sys.device) for _ in range(n)]
        # fresh buffers may reuse memory that kernels queued on the main stream still read or
        # write (the allocator orders reuse on that stream only): the first copies wait for them
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(prob.sys.device))
        self.free = [ev] * n
        self.jobs = {}
        self.next = 0
        self.last = None
        for _ in range(n):
            self._issue()

    def _issue(self):
        i = self.next
        if i >= len(self.order):
            return
        self.next += 1
        k = self.order[i]
        slot = i % len(self.ring)
        src = self.p.u_obs[self.p.obs_index[k]][..., self.a:self.b]
        self.jobs[i] = self.p._worker.submit(self.p._copy_in, self.free[slot], self.ring[slot], src)

    def get(self, k):
        i = self.pos[k]
        if self.last is not None:
            if i != self.last + 1:
                raise RuntimeError("observations must be taken last step first, each once")
            # the kernel that read the previous observation has been issued: its slot is free
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(self.p.sys.device))
            self.free[self.last % len(self.ring)] = ev
            self._issue()
        elif i != 0:
            raise RuntimeError("observations must be taken last step first, each once")
        self.jobs.pop(i).result()
        self.last = i
        return self.ring[i % len(self.ring)]

    def close(self):
        for job in self.jobs.values():
            job.result()
        self.jobs = {}


class AdjointProblem:
    """Forward + adjoint of ``nsteps`` steps on a :class:`MembraneSystem`.

    Parameters
    ----------
    system : assembled system (its ``coef``/values are replaced by :meth:`set_theta`).
    load : :class:`LoadTables` with at least ``nsteps`` entries (entry ``n`` = load of step ``n``).
    nsteps, dt, beta, damp, method : as for the forward step.
    obs_steps : sorted step indices ``k`` (1..nsteps) at which ``u_k`` is observed.
    u_obs : tensor ``[len(obs_steps), N, 3, S]`` of observations (``None`` = zero), on the device or
        in (preferably pinned) host memory.
    checkpoint_every : checkpoint interval ``C`` (1 = keep every state, no recomputation).
    inner_every : second-level interval ``C2`` inside a segment (``None`` = one level).
    fuse : recompute the segment before the current one (of the innermost level) during the
        current one's reverse sweep, each recomputed step in the same kernel as an adjoint step
        (the operator is read once for both; needs the generic forward kernel and one more state
        of memory).  ``"auto"`` times three fused kernels against three forward + adjoint pairs
        on the problem's own states first and keeps the faster form (the fused kernel carries
        two columns of states and loses where they overflow L2); the gradient is the same bit
        for bit either way.
    sample_block : run the reverse sweep for blocks of this many samples at a time (matrix-free,
        one checkpoint level): with device checkpoints one block after the other, each with the
        pipelined sweep when ``fuse`` allows it; with host checkpoints segment by segment over
        the blocks, without fused kernels.  The gradient is that of the whole-ensemble sweep bit
        for bit.
    host_checkpoints : keep the checkpoints in (pageable) host memory, one pair per segment and
        sample block, copied by a helper thread while the kernels run.
    concurrent : run each gradient step on a stream of its own beside the adjoint steps (the two
        kernels are independent and neither fills the memory system alone) in the sweeps that
        recompute a segment before reversing it (not the pipelined one).  Same result bit for
        bit.
    """

    def __init__(self, system: MembraneSystem, load: LoadTables, nsteps, dt, beta=0.0, damp=0.0,
                 method="stored", obs_steps=(), u_obs=None, checkpoint_every=100, u0=None, du0=None,
                 inner_every=None, fuse="auto", sample_block=None, host_checkpoints=False,
                 concurrent=False):
        self.sys = system
        self.load = load
        self.nsteps = int(nsteps)
        self.dt, self.beta, self.damp = float(dt), float(beta), float(damp)
        self.method = method
        self.obs_steps = [int(k) for k in obs_steps]
        if any(k < 1 or k > self.nsteps for k in self.obs_steps):
            raise ValueError("observation steps must lie in 1..nsteps")
        self.obs_index = {k: i for i, k in enumerate(self.obs_steps)}
        self.u_obs = u_obs
        self.C = max(1, int(checkpoint_every))
        self.C2 = None if inner_every is None or int(inner_every) >= self.C else max(1, int(inner_every))
        self.fuse = fuse if fuse == "auto" else bool(fuse)
        self._obs_dev = None
        N = system.n_nodes
        self.u0 = np.zeros(3 * N) if u0 is None else np.asarray(u0, np.float64)
        self.du0 = np.zeros(3 * N) if du0 is None else np.asarray(du0, np.float64)
        self.checkpoints = {}
        self.stats = {}
        self.concurrent = bool(concurrent)
        self._grad_stream = torch.cuda.Stream(device=system.device) if self.concurrent else None
        S = system.n_smp
        self.sample_block = None if sample_block is None else int(sample_block)
        self.host_checkpoints = bool(host_checkpoints)
        self._blocks = None
        if self._blocked():
            if self.sample_block is not None and not 1 <= self.sample_block <= S:
                raise ValueError("sample_block must lie in 1..n_smp")
            if self.C2 is not None:
                raise ValueError("sample blocks and host checkpoints run one checkpoint level")
            if method != "matfree" and self._ranges() != [(0, S)]:
                raise ValueError("sample blocks run the matrix-free operator")
            if self.host_checkpoints:
                self.fuse = False
            dev = system.device
            self._copy_stream = torch.cuda.Stream(device=dev)
            self._stage = [None, None]          # device staging of the checkpoint copies
            self._stage_jobs = [None, None]     # helper-thread copies out of each staging slot
            self._stage_i = 0
            self._host = {}
            self._worker = _helper()
            if self.host_checkpoints:
                for mark in range(0, self.nsteps, self.C):
                    self._host[mark] = [tuple(torch.empty((N, 3, b - a), dtype=torch.float64)
                                              for _ in range(2)) for a, b in self._ranges()]

    # ------------------------------------------------------------------------------------------
    def set_theta(self, theta):
        self.sys.set_log_modulus(theta)

    def _obs(self, k):
        if self.u_obs is None:
            return None
        ref = self.u_obs[self.obs_index[k]]
        if ref.device == self.sys.device:
            return ref
        if self._obs_dev is None:
            self._obs_dev = torch.empty(ref.shape, dtype=torch.float64, device=self.sys.device)
        self._obs_dev.copy_(ref, non_blocking=True)
        return self._obs_dev

    def _loss_term(self, u, k):
        ref = self._obs(k)
        r = u if ref is None else u - ref
        return 0.5 * (r * r).sum(dim=(0, 1))

    def _start(self):
        self.sys.initial_state(self.u0, self.du0, self.load, self.dt, beta=self.beta,
                               damp=self.damp, method=self.method)

    def _advance(self, offset, n):
        if n > 0:
            self.sys.step(self.load, self.dt, beta=self.beta, damp=self.damp, method=self.method,
                          offset=offset, nsteps=n)

    def forward(self, record=False):
        """Run the forward pass with checkpoints; returns the per-sample loss ``[S]``.

        ``record=True`` also keeps the observed states in ``self.obs_states`` (in ``obs_steps``
        order).
        """
        sysm = self.sys
        self._start()
        self.checkpoints = {}
        self.obs_states = []
        loss = torch.zeros(sysm.n_smp, dtype=torch.float64, device=sysm.device)
        marks = sorted(set(range(0, self.nsteps, self.C)) | set(self.obs_steps) | {self.nsteps})
        step = 0
        for mark in marks:
            self._advance(step, mark - step)
            step = mark
            if mark % self.C == 0 and mark < self.nsteps:
                self._store_checkpoint(mark)
            if mark in self.obs_index:
                loss += self._loss_term(sysm.u, mark)
                if record:
                    self.obs_states.append(sysm.u.clone())
        self.u_final = sysm.u.clone()
        if self.host_checkpoints:
            for job in self._stage_jobs:
                if job is not None:
                    job.result()
        return loss

    def _blocked(self):
        return self.sample_block is not None or self.host_checkpoints

    def _ranges(self):
        S = self.sys.n_smp
        B = S if self.sample_block is None else self.sample_block
        return [(a, min(a + B, S)) for a in range(0, S, B)]

    def _store_checkpoint(self, mark):
        """Keep the pair (u_mark, u_mark-1): on the device, or per sample block (device copies of
        the block's samples), or in host memory: device copies into one of two staging slots,
        then the helper thread's copy to the host while the main thread goes on with the steps
        (it waits only when a slot's previous copy is still running)."""
        sysm = self.sys
        if not self._blocked():
            self.checkpoints[mark] = (sysm.u.clone(), sysm.up.clone())
            return
        ranges = self._ranges()
        if not self.host_checkpoints:
            self.checkpoints[mark] = [(sysm.u[..., a:b].contiguous(), sysm.up[..., a:b].contiguous())
                                      for a, b in ranges]
            return
        main = torch.cuda.current_stream(sysm.device)
        slot = self._stage_i
        self._stage_i ^= 1
        if self._stage_jobs[slot] is not None:
            self._stage_jobs[slot].result()           # the slot's previous copy is out
        if self._stage[slot] is None:
            self._stage[slot] = [tuple(torch.empty((sysm.n_nodes, 3, b - a), dtype=torch.float64,
                                                   device=sysm.device) for _ in range(2))
                                 for a, b in ranges]
        stage = self._stage[slot]
        for (a, b), (su, sup) in zip(ranges, stage):
            su.copy_(sysm.u[..., a:b])
            sup.copy_(sysm.up[..., a:b])
        ready = torch.cuda.Event()
        ready.record(main)
        host = self._host[mark]
        self._stage_jobs[slot] = self._worker.submit(self._copy_pairs, ready, stage, host)
        self.checkpoints[mark] = host

    def _copy_in(self, ready, dst, src):
        """Helper thread: once ``ready`` has passed on the device, copy ``src`` into the device
        buffer ``dst`` on the copy stream."""
        with torch.cuda.device(self.sys.device):
            if ready is not None:
                ready.synchronize()
            with torch.cuda.stream(self._copy_stream):
                dst.copy_(src)
            self._copy_stream.synchronize()

    def _copy_pairs(self, ready, src, dst):
        """Helper thread: once ``ready`` has passed on the device, copy the pairs ``src`` into
        ``dst`` on the copy stream (pageable host copies return when done)."""
        dev = self.sys.device
        with torch.cuda.device(dev):
            if ready is not None:
                ready.synchronize()
            with torch.cuda.stream(self._copy_stream):
                for (su, sup), (du, dup) in zip(src, dst):
                    du.copy_(su)
                    dup.copy_(sup)
            self._copy_stream.synchronize()

    # ------------------------------------------------------------------------------------------
    def _adjoint_call(self, mu1, mu2, out, obs_u=None, obs_ref=None, w=None, project=1):
        sysm = self.sys
        check(sysm.lib.cvfes_adjoint_step(sysm.stream, ctypes.byref(sysm._sys),
                                          sysm.method_id(self.method), ptr(mu1), ptr(mu2), ptr(out),
                                          ptr(obs_u), ptr(obs_ref), ptr(w), self.dt, self.beta,
                                          self.damp, project))

    def _fused_call(self, u, up, out, n, mu1, mu2, mu_out, obs_u=None, obs_ref=None, w=None):
        """Forward step n (u, up -> out) and one projected adjoint step in one kernel."""
        sysm = self.sys
        arr = (ctypes.c_void_p * 3)(u.data_ptr(), up.data_ptr(), out.data_ptr())
        ls = self.load.struct(n)
        check(sysm.lib.cvfes_adjoint_fused_step(sysm.stream, ctypes.byref(sysm._sys),
                                                sysm.method_id(self.method), arr, ctypes.byref(ls),
                                                ptr(mu1), ptr(mu2), ptr(mu_out), ptr(obs_u),
                                                ptr(obs_ref), ptr(w), self.dt, self.beta, self.damp))

    def fused(self):
        """Whether the reverse sweep can pair recomputation and adjoint steps in one kernel (with
        ``fuse = "auto"`` it does so where that is faster, ``stats["fused"]``)."""
        leaf = self.C if self.C2 is None else self.C2
        if not self.fuse or leaf < 2 or (self.C2 is None and len(self.checkpoints) < 2):
            return False
        return self._generic_forward_kernel()

    def _generic_forward_kernel(self):
        sysm = self.sys
        kern = sysm.lib.cvfes_step_kernel(ctypes.byref(sysm._sys), sysm.method_id(self.method))
        return kern == VARIANT_GENERIC

    def _grad_call(self, mu, xc, xp, nu_scale, grad):
        sysm = self.sys
        check(sysm.lib.cvfes_grad_accum(sysm.stream, ctypes.byref(sysm._sys), ptr(sysm.elems),
                                        ptr(sysm.cpos), ptr(mu), ptr(xc), ptr(xp), self.dt,
                                        self.beta, nu_scale, ptr(grad)))

    def backward(self, grad_out=None):
        """Reverse sweep; returns ``dL/dtheta [E, S]`` for the loss weighted by ``grad_out [S]``."""
        sysm = self.sys
        dev = sysm.device
        S = sysm.n_smp
        w = (torch.ones(S, dtype=torch.float64, device=dev) if grad_out is None
             else grad_out.to(dev, torch.float64).reshape(S).contiguous())
        if self._blocked():
            return self._backward_blocks(w)
        N = self.nsteps
        self._begin_reverse(w)
        shape = sysm.state_shape()
        leaf = self.C if self.C2 is None else self.C2
        fused = self.fused()
        # states of the innermost segment (the pipelined sweep also holds the next one's, in the
        # buffers the current one releases, plus one)
        self._pool = [torch.empty(shape, dtype=torch.float64, device=dev)
                      for _ in range(leaf if fused else leaf - 1)]
        self._tmp = [torch.empty(shape, dtype=torch.float64, device=dev) for _ in range(3)] \
            if self.C2 is not None else []
        self.stats["fused_steps"] = 0
        if fused and self.fuse == "auto":
            last = max(self.checkpoints)
            fused = self._fusion_pays(self.checkpoints[last], last)
        self.stats["fused"] = fused
        if fused and self.C2 is None:
            self._reverse_pipelined(self.checkpoints, self.C, N)
        else:
            for s0 in sorted(self.checkpoints, reverse=True):
                s1 = min(s0 + self.C, N)
                if self.C2 is None:
                    self._reverse_leaf(s0, s1, self.checkpoints[s0])
                    continue
                pairs = self._recompute_pairs(s0, s1, self.checkpoints[s0], self.C2)
                if fused:
                    self._reverse_pipelined(pairs, self.C2, s1)
                else:
                    for t0 in sorted(pairs, reverse=True):
                        self._reverse_leaf(t0, min(t0 + self.C2, s1), pairs[t0])
                del pairs
        return self._end_reverse()

    def _begin_reverse(self, w):
        """Adjoint state mu_{N+1} = 0, mu_N = P g_N and a zero gradient."""
        sysm = self.sys
        dev = sysm.device
        shape = sysm.state_shape()
        self._zeros = torch.zeros(shape, dtype=torch.float64, device=dev)
        mu = [torch.zeros(shape, dtype=torch.float64, device=dev) for _ in range(2)]
        self._grad = torch.zeros((sysm.n_elms, sysm.n_smp), dtype=torch.float64, device=dev)
        N = self.nsteps
        if N in self.obs_index:
            self._adjoint_call(self._zeros, self._zeros, mu[0], self.u_final, self._obs(N), w)
        self._mu = [mu[0], self._zeros.clone(), mu[1]]        # mu_{n+1}, mu_{n+2}, spare
        self._w = w

    def _end_reverse(self):
        """The initial state's term (u_{-1} depends on K through u0 + beta du0); returns the
        gradient and releases the sweep's buffers."""
        sysm = self.sys
        dev = sysm.device
        S = sysm.n_smp
        shape = sysm.state_shape()
        grad, zeros = self._grad, self._zeros
        mu_n1 = self._mu[0]
        self._pool = self._tmp = None
        self._mu = self._grad = self._zeros = None
        sysm.state = None
        x0 = self.u0 + self.beta * self.du0
        if np.any(x0 != 0.0):
            lam_m1 = torch.empty(shape, dtype=torch.float64, device=dev)
            self._adjoint_call(zeros, mu_n1, lam_m1, project=0)
            x0_d = torch.as_tensor(np.repeat(x0.reshape(-1, 3)[:, :, None], S, axis=2)).to(dev)
            self._grad_call(lam_m1, x0_d, None, 0.5 * self.dt * self.dt, grad)
        return grad

    def _block_problems(self):
        """One single-level problem per sample block, on the block's system (this system itself
        when one block holds every sample)."""
        if self._blocks is None:
            self._blocks = []
            for a, b in self._ranges():
                sb = self.sys if (a, b) == (0, self.sys.n_smp) else self.sys.sample_block(a, b)
                if sb is not self.sys and self.sys.waves_per_simd:
                    # the occupancy cap that suits the whole ensemble need not suit a block
                    sb.autotune_occupancy(self.load, self.dt, self.beta, self.damp,
                                          method=self.method, steps=5)
                p = AdjointProblem(sb, self.load, self.nsteps, self.dt, beta=self.beta,
                                   damp=self.damp, method=self.method, obs_steps=self.obs_steps,
                                   checkpoint_every=self.C, u0=self.u0, du0=self.du0, fuse=False,
                                   concurrent=self.concurrent)
                p.block_range = (a, b)
                self._blocks.append(p)
        return self._blocks

    def _block_obs(self, k, j):
        """Observation k restricted to sample block j (device; one host transfer per step k)."""
        if self.u_obs is None:
            return None
        ranges = self._ranges()
        if len(ranges) == 1:
            return self._obs(k)
        if k not in self._obs_blocks:
            full = self._obs(k)
            while len(self._obs_blocks) >= 2:                 # the steps of the current segment
                del self._obs_blocks[next(iter(self._obs_blocks))]
            self._obs_blocks[k] = [full[..., a:b].contiguous() for a, b in ranges]
        return self._obs_blocks[k][j]

    def _backward_blocks(self, w):
        """Reverse sweep one sample block at a time: for each segment (last first), each block
        recomputes its states from the block's checkpoint pair and runs its adjoint and gradient
        steps through the segment.  Host checkpoint pairs come to the device one pair ahead on
        the copy stream, in two alternating buffer slots."""
        sysm = self.sys
        dev = sysm.device
        blocks = self._block_problems()
        self._obs_blocks = {}
        self._stage = [None, None]           # the forward pass's staging (its copies were waited on)
        if blocks[0].sys is not sysm:
            sysm.state = None                # the blocks recompute into their own buffers
        for j, p in enumerate(blocks):
            a, b = p.block_range
            if p.sys is not sysm:
                p.sys.refresh_block(sysm)
            p.u_final = self.u_final if p.sys is sysm else self.u_final[..., a:b].contiguous()
        big = blocks[0].sys.state_shape()

        def fit(buf, shape):                 # the first `shape` elements of a block-0 buffer
            return buf.view(-1)[:int(np.prod(shape))].view(shape)

        if not self.host_checkpoints:
            return self._backward_block_major(w, blocks, big, fit)
        for j, p in enumerate(blocks):
            a, b = p.block_range
            p._obs = (lambda k, j=j: self._block_obs(k, j))
            p._begin_reverse(w[a:b].contiguous())
        pool = [torch.empty(big, dtype=torch.float64, device=dev) for _ in range(self.C - 1)]

        order = [(m, j) for m in sorted(self.checkpoints, reverse=True) for j in range(len(blocks))]
        main = torch.cuda.current_stream(dev)
        slots = [None, None]
        free = [None, None]                  # main-stream events: a slot's last reader has run

        def fetch(i):                        # copy pair i to a device slot on the helper thread
            m, j = order[i]
            k = i % 2
            if slots[k] is None:
                slots[k] = tuple(torch.empty(big, dtype=torch.float64, device=dev) for _ in range(2))
                # the helper writes them on its own stream: after what the main stream has queued
                free[k] = torch.cuda.Event()
                free[k].record(main)
            shape = blocks[j].sys.state_shape()
            pair = tuple(fit(t, shape) for t in slots[k])
            return self._worker.submit(self._copy_pairs, free[k], [self.checkpoints[m][j]],
                                       [pair]), pair

        pending = fetch(0) if self.host_checkpoints and order else None
        for i, (m, j) in enumerate(order):
            p = blocks[j]
            if self.host_checkpoints:
                job, pair = pending
                job.result()
                if i + 1 < len(order):
                    pending = fetch(i + 1)
            else:
                pair = self.checkpoints[m][j]
            shape = p.sys.state_shape()
            p._pool = [fit(t, shape) for t in pool]
            p._reverse_leaf(m, min(m + self.C, self.nsteps), pair)
            if self.host_checkpoints:
                ev = torch.cuda.Event()
                ev.record(main)
                free[i % 2] = ev
        self._obs_blocks = {}
        S = sysm.n_smp
        grad = torch.empty((sysm.n_elms, S), dtype=torch.float64, device=dev)
        for p in blocks:
            a, b = p.block_range
            g = p._end_reverse()
            grad[:, a:b] = g
        self.stats["fused_steps"] = 0
        self.stats["fused"] = False
        return grad

    def _backward_block_major(self, w, blocks, big, fit):
        """Device checkpoints: the whole reverse sweep of one sample block after the other, so
        that a block can run the pipelined sweep (each segment recomputed during the reverse
        steps of the one after it, one fused kernel per pair of steps) where that is faster.  The
        block's observations come to the device one after the other, a few ahead, on the helper
        thread."""
        sysm = self.sys
        dev = sysm.device
        S = sysm.n_smp
        grad = torch.empty((sysm.n_elms, S), dtype=torch.float64, device=dev)
        marks = sorted(self.checkpoints)
        fused = bool(self.fuse) and len(marks) >= 2 and blocks[0]._generic_forward_kernel()
        pool = [torch.empty(big, dtype=torch.float64, device=dev)
                for _ in range(self.C if fused else self.C - 1)]
        self.stats["fused_steps"] = 0
        for j, p in enumerate(blocks):
            a, b = p.block_range
            obs = _ObsStream(self, a, b) if self.u_obs is not None else None
            p._obs = (lambda k, obs=obs: None if obs is None else obs.get(k))
            p._begin_reverse(w[a:b].contiguous())
            shape = p.sys.state_shape()
            p._pool = [fit(t, shape) for t in pool]
            pairs = {m: self.checkpoints[m][j] for m in marks}
            if fused and self.fuse == "auto" and j == 0:
                fused = p._fusion_pays(pairs[marks[-1]], marks[-1])
                self.stats["fusion_timing_ms"] = p.stats.get("fusion_timing_ms")
            if fused:
                p.stats["fused_steps"] = 0
                p._reverse_pipelined(pairs, self.C, self.nsteps)
                self.stats["fused_steps"] += p.stats["fused_steps"]
            else:
                p._pool = p._pool[:self.C - 1]
                for m in reversed(marks):
                    p._reverse_leaf(m, min(m + self.C, self.nsteps), pairs[m])
            grad[:, a:b] = p._end_reverse()
            if obs is not None:
                obs.close()
        self.stats["fused"] = fused
        return grad

    def _fusion_pays(self, pair, n):
        """Time three fused kernels against three forward + adjoint pairs (after one of each) on
        the checkpoint ``pair`` of step ``n`` and the current adjoint state, writing only into
        scratch buffers of the pool; True if the fused kernel is faster."""
        sysm = self.sys
        u, up = pair
        a, b = self._pool[0], self._pool[1]
        mu_n1, mu_n2, _ = self._mu

        def separate():
            sysm.state = [u, up, a]
            self._advance(n, 1)
            self._adjoint_call(mu_n1, mu_n2, b)

        def fused():
            self._fused_call(u, up, a, n, mu_n1, mu_n2, b)

        st = torch.cuda.current_stream(sysm.device)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        separate()
        fused()
        ev[0].record(st)
        for _ in range(3):
            separate()
        ev[1].record(st)
        for _ in range(3):
            fused()
        ev[2].record(st)
        ev[2].synchronize()
        t_sep, t_fused = ev[0].elapsed_time(ev[1]) / 3, ev[1].elapsed_time(ev[2]) / 3
        self.stats["fusion_timing_ms"] = {"separate": t_sep, "fused": t_fused}
        return t_fused < t_sep

    def _recompute_pairs(self, s0, s1, pair, C2):
        """Forward from the pair (u_{s0}, u_{s0-1}) to step s1 - 1, keeping (u_k, u_{k-1}) for
        k = s0 + j C2 < s1; returns {k: pair} including s0."""
        sysm = self.sys
        pairs = {s0: pair}
        cur, prev = pair
        for n in range(s0, s1 - 1):
            work = next(b for b in self._tmp if b is not cur and b is not prev)
            sysm.state = [cur, prev, work]
            self._advance(n, 1)
            prev, cur = cur, sysm.state[0]
            if (n + 1 - s0) % C2 == 0:
                pairs[n + 1] = (cur.clone(), prev.clone())
        return pairs

    def _recompute(self, s0, s1, pair, free):
        """States u_{s0-1} .. u_{s1-1} from the pair (u_{s0}, u_{s0-1}) by plain steps into
        buffers taken from the list ``free``; returns {n: state}."""
        sysm = self.sys
        u_s0, u_sm1 = pair
        states = {s0 - 1: u_sm1, s0: u_s0}
        for n in range(s0, s1 - 1):
            out = free.pop()
            sysm.state = [states[n], states[n - 1], out]
            self._advance(n, 1)
            states[n + 1] = out
        return states

    def _reverse_pipelined(self, pairs, C, end):
        """Reverse sweep over the segments that start at the keys of ``pairs`` (one every ``C``
        steps, the last ending at ``end``) in which the segment before the current one is
        recomputed during the current one's adjoint steps, one recomputed step and one adjoint
        step per fused kernel; the arithmetic of both is that of the separate kernels.  The
        state u_n of the current segment is not read again once step n's gradient and adjoint
        kernels are issued (all on one stream), so its buffer takes a recomputed state next:
        ``C`` buffers hold both segments."""
        sysm = self.sys
        free = list(self._pool)
        starts = sorted(pairs, reverse=True)
        cur = self._recompute(starts[0], min(starts[0] + C, end), pairs[starts[0]], free)
        mu_n1, mu_n2, spare = self._mu
        fused = 0
        for idx, s0 in enumerate(starts):
            s1 = min(s0 + C, end)
            nxt = starts[idx + 1] if idx + 1 < len(starts) else None
            fn = fend = 0
            if nxt is not None:
                u_c, up_c = pairs[nxt]
                nstates = {nxt - 1: up_c, nxt: u_c}
                fn, fend = nxt, min(nxt + C, end) - 1            # forward steps fn .. fend - 1
            for n in range(s1 - 1, s0 - 1, -1):
                self._grad_call(mu_n1, cur[n], cur[n - 1] if self.beta != 0.0 else None,
                                self.dt * self.dt, self._grad)
                if n == 0:
                    break
                obs_u = cur[n] if n in self.obs_index else None
                ref = self._obs(n) if obs_u is not None else None
                w = self._w if obs_u is not None else None
                if fn < fend:
                    out = free.pop()
                    self._fused_call(nstates[fn], nstates[fn - 1], out, fn, mu_n1, mu_n2, spare,
                                     obs_u, ref, w)
                    nstates[fn + 1] = out
                    fn += 1
                    fused += 1
                else:
                    self._adjoint_call(mu_n1, mu_n2, spare, obs_u, ref, w)
                mu_n2, mu_n1, spare = mu_n1, spare, mu_n2
                if n > s0:
                    free.append(cur[n])                          # a recomputed state, done with
            while fn < fend:                                     # the rest of the recomputation
                out = free.pop()
                sysm.state = [nstates[fn], nstates[fn - 1], out]
                self._advance(fn, 1)
                nstates[fn + 1] = out
                fn += 1
            if nxt is not None:
                cur = nstates
        self._mu = [mu_n1, mu_n2, spare]
        self.stats["fused_steps"] += fused

    def _reverse_leaf(self, s0, s1, pair):
        """Recompute u_{s0+1} .. u_{s1-1} from (u_{s0}, u_{s0-1}), then run the adjoint and gradient
        steps n = s1-1 .. s0."""
        sysm = self.sys
        u_s0, u_sm1 = pair
        states = {s0 - 1: u_sm1, s0: u_s0}
        for i, n in enumerate(range(s0, s1 - 1)):
            sysm.state = [states[n], states[n - 1], self._pool[i]]
            self._advance(n, 1)
            states[n + 1] = self._pool[i]
        mu_n1, mu_n2, spare = self._mu
        gs = self._grad_stream
        main = torch.cuda.current_stream(sysm.device) if gs is not None else None
        done = {}
        for n in range(s1 - 1, s0 - 1, -1):
            # gradient of step n: mu_{n+1}, x_n = u_n + (beta/dt)(u_n - u_{n-1})
            xp = states[n - 1] if self.beta != 0.0 else None
            if gs is None:
                self._grad_call(mu_n1, states[n], xp, self.dt * self.dt, self._grad)
            else:
                # on the gradient stream, once mu_{n+1} is out (main-stream order); the gradient
                # steps keep their order on that stream, so the sum is the same bit for bit
                gs.wait_stream(main)
                with torch.cuda.stream(gs):
                    self._grad_call(mu_n1, states[n], xp, self.dt * self.dt, self._grad)
                    done[n] = torch.cuda.Event()
                    done[n].record(gs)
            if n == 0:
                break
            obs_u = states[n] if n in self.obs_index else None
            if n + 2 in done:              # `spare` holds mu_{n+3}, read by step n+2's gradient
                main.wait_event(done.pop(n + 2))
            self._adjoint_call(mu_n1, mu_n2, spare, obs_u,
                               self._obs(n) if obs_u is not None else None,
                               self._w if obs_u is not None else None)
            mu_n2, mu_n1, spare = mu_n1, spare, mu_n2
        if gs is not None:
            main.wait_stream(gs)           # the segment's states are overwritten next
        self._mu = [mu_n1, mu_n2, spare]


class EnsembleLoss(torch.autograd.Function):
    """``loss[S] = EnsembleLoss.apply(theta, problem)`` with ``theta = log E [E, S]``."""

    @staticmethod
    def forward(ctx, theta, problem):
        problem.set_theta(theta)
        loss = problem.forward()
        ctx.problem = problem
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        grad = ctx.problem.backward(grad_out)
        return grad.to(grad_out.device), None
