"""FP64 torch (CPU) restatement of the ensemble step and its loss - TEST INFRASTRUCTURE ONLY.

Reference gradient for the adjoint kernels: the same recurrence as ``solid_oracle.run`` (linear
mode, CPU-path lumped mass, load ``p_n f1``, Rayleigh ``beta``, ``damp``, fixed end rings), written
with torch tensors so that ``torch.autograd`` differentiates the loss
``L_s = 1/2 sum_k |u_s(t_k) - u_obs,s(t_k)|^2`` with respect to ``theta = log E`` per element and
sample.  The stiffness acts element by element, ``K x = sum_e coef_e khat_e x_e`` with ``khat`` from
``solid_oracle.element_blocks_closed_form`` (pinned in tests/test_oracle.py) and
``coef_e = E_e * h`` for uniform thickness ``h``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import solid_oracle as so


class TorchEnsemble:
    def __init__(self, nodes, elements, boundary, thickness, density, nu, mass_scale=1.0 / 9.0):
        el = np.asarray(elements, np.int64)
        kh, area, normal = so.element_blocks_closed_form(np.asarray(nodes, float), el, nu)
        n = nodes.shape[0]
        self.n = n
        self.el = torch.as_tensor(el)
        self.khat = torch.as_tensor(kh)                        # [E, 9, 9]
        lm = np.zeros(n)
        f1 = np.zeros((n, 3))
        for a in range(3):
            np.add.at(lm, el[:, a], mass_scale * density * area * thickness)
            np.add.at(f1, el[:, a], (area / 3.0)[:, None] * normal)
        self.lm = torch.as_tensor(lm)
        self.f1 = torch.as_tensor(f1)
        free = np.ones(n)
        free[np.asarray(boundary)] = 0.0
        self.free = torch.as_tensor(free)
        self.h = float(thickness)

    def kx(self, coef, x):
        """``K x`` for ``x [N, 3, S]`` with ``coef [E, S]``."""
        xe = x[self.el].reshape(self.el.shape[0], 9, -1)       # [E, 9, S]
        ye = torch.einsum("eij,ejs->eis", self.khat, xe) * coef[:, None, :]
        y = torch.zeros_like(x)
        y.index_add_(0, self.el.reshape(-1), ye.reshape(-1, 3, x.shape[2]))
        return y

    def states_at(self, theta, pressures, dt, beta=0.0, damp=0.0, steps=(), u0=None, du0=None):
        """Displacements ``u_k`` for ``k`` in ``steps`` (stacked ``[len(steps), N, 3, S]``)."""
        out = []
        self._run(theta, pressures, dt, beta, damp, steps, u0, du0, lambda k, u: out.append(u))
        return torch.stack(out)

    def loss(self, theta, pressures, dt, beta=0.0, damp=0.0, obs_steps=(), u_obs=None, u0=None,
             du0=None):
        obs_index = {k: i for i, k in enumerate(obs_steps)}
        terms = []

        def add(k, u):
            ref = 0.0 if u_obs is None else u_obs[obs_index[k]]
            r = u - ref
            terms.append(0.5 * (r * r).sum(dim=(0, 1)))

        self._run(theta, pressures, dt, beta, damp, obs_steps, u0, du0, add)
        return torch.stack(terms).sum(dim=0)

    def _run(self, theta, pressures, dt, beta, damp, steps, u0, du0, visit):
        S = theta.shape[1]
        coef = self.h * torch.exp(theta)
        m = self.lm[:, None, None]
        f1 = self.f1[:, :, None]
        z = torch.zeros((self.n, 3, S), dtype=torch.float64)
        u0 = z if u0 is None else torch.as_tensor(np.asarray(u0)).reshape(self.n, 3, 1).expand(-1, -1, S)
        du0 = z if du0 is None else torch.as_tensor(np.asarray(du0)).reshape(self.n, 3, 1).expand(-1, -1, S)
        ddu = (pressures[0] * f1 - damp * du0 - self.kx(coef, u0 + beta * du0)) / m
        u, up = u0, u0 - dt * du0 + 0.5 * dt * dt * ddu
        steps = set(steps)
        free = self.free[:, None, None]
        for n, p in enumerate(pressures):
            x = u + (beta / dt) * (u - up)
            rhs = dt * dt * (p * f1 - damp * (u - up) / dt - self.kx(coef, x)) + m * (2.0 * u - up)
            unew = free * (rhs / m)
            up, u = u, unew
            if n + 1 in steps:
                visit(n + 1, u)
