"""Per-element Gaussian Markov random field samples of Young's modulus.

CVFES samples node-based Matern-SPDE fields offline with CHOLMOD (``GMRF.py:140-354``).  The
ensemble configurations here use a per-element log-normal model instead::

    E[e,s] = E0 * exp(z[e,s] - var/2),    z[:,s] ~ N(0, Q^-1),   Q = tau * (I + L)

where ``L = B^T B`` is the graph Laplacian of the element dual graph (elements sharing an edge)
and ``B`` its signed edge-incidence matrix.  Sampling uses perturbation-optimisation, which is
exact for this Q: with ``xi1 ~ N(0, I_E)``, ``xi2 ~ N(0, I_m)``, ``z = Q^-1 sqrt(tau)(xi1 + B^T xi2)``
has covariance ``Q^-1 Q Q^-1 = Q^-1``.  The solve ``(I + L) y = xi1 + B^T xi2`` (spectrum in
``[1, 1 + 2*deg_max]``) is done by conjugate gradients on all samples at once, and
``z = y / sqrt(tau)``.

``tau`` is set so that the mean marginal variance of ``z`` equals ``var = ln(1 + cov^2)`` (10 %
coefficient of variation gives ``var = ln 1.01``): ``tau = mean(diag((I+L)^-1)) / var``, with the
mean diagonal computed exactly (dense) on small graphs and by a seeded Hutchinson estimate with
Rademacher probes on large ones.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

__all__ = ["element_dual_edges", "dual_laplacian_system", "mean_inverse_diagonal",
           "sample_element_E", "GMRFSample"]


def element_dual_edges(elements: np.ndarray) -> np.ndarray:
    """Pairs ``(e1, e2)`` (``e1 < e2``) of elements that share a mesh edge, shape ``[m, 2]``."""
    elements = np.asarray(elements, dtype=np.int64)
    n_elm = elements.shape[0]
    a = elements[:, [0, 1, 2]]
    b = elements[:, [1, 2, 0]]
    lo = np.minimum(a, b).ravel()
    hi = np.maximum(a, b).ravel()
    owner = np.repeat(np.arange(n_elm), 3)
    order = np.lexsort((owner, hi, lo))
    lo, hi, owner = lo[order], hi[order], owner[order]
    same = (lo[1:] == lo[:-1]) & (hi[1:] == hi[:-1])
    pairs = np.stack([owner[:-1][same], owner[1:][same]], axis=1)
    return np.sort(pairs, axis=1)


def dual_laplacian_system(elements: np.ndarray):
    """Return ``(A, B)`` with ``A = I + B^T B`` (CSR) and the signed incidence ``B`` (``[m, E]``)."""
    n_elm = np.asarray(elements).shape[0]
    pairs = element_dual_edges(elements)
    m = pairs.shape[0]
    rows = np.repeat(np.arange(m), 2)
    cols = pairs.ravel()
    vals = np.tile([1.0, -1.0], m)
    B = sp.csr_matrix((vals, (rows, cols)), shape=(m, n_elm))
    A = (sp.identity(n_elm, format="csr") + (B.T @ B)).tocsr()
    return A, B


def _cg(A, rhs, tol=1e-13, maxiter=500):
    """Conjugate gradients on every column of ``rhs`` at once (per-column step lengths)."""
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy()
    rr = np.einsum("ij,ij->j", r, r)
    r0 = np.sqrt(np.maximum(rr, 1e-300))
    for _ in range(maxiter):
        if np.all(np.sqrt(rr) <= tol * r0):
            break
        ap = A @ p
        pap = np.einsum("ij,ij->j", p, ap)
        alpha = np.where(pap > 0, rr / np.where(pap > 0, pap, 1.0), 0.0)
        x += alpha * p
        r -= alpha * ap
        rr_new = np.einsum("ij,ij->j", r, r)
        beta = np.where(rr > 0, rr_new / np.where(rr > 0, rr, 1.0), 0.0)
        p = r + beta * p
        rr = rr_new
    return x


def mean_inverse_diagonal(A, seed: int = 12345, probes: int = 64, exact_limit: int = 4000,
                          device=None) -> float:
    """Mean of ``diag(A^-1)``: exact for ``n <= exact_limit``, else Hutchinson with Rademacher probes
    (the probe solves run on ``device`` when one is given)."""
    n = A.shape[0]
    if n <= exact_limit:
        return float(np.mean(np.diag(np.linalg.inv(A.toarray()))))
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 2, size=(n, probes)).astype(np.float64) * 2.0 - 1.0
    y = _cg(A, v) if device is None else _cg_torch(A, v, device)
    return float(np.mean(np.einsum("ij,ij->j", v, y)) / n)


class GMRFSample:
    """Result of :func:`sample_element_E`: fields ``E [E,S]``, ``z [E,S]``, ``tau``, ``var``."""

    def __init__(self, E, z, tau, var):
        self.E, self.z, self.tau, self.var = E, z, tau, var


def _cg_torch(A, rhs, device, tol=1e-13, maxiter=500):
    """Batched conjugate gradients with the sparse operator on ``device`` (torch CSR)."""
    import torch
    A = A.tocsr()
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr, dtype=torch.int64),
                                 torch.as_tensor(A.indices, dtype=torch.int64),
                                 torch.as_tensor(A.data), size=A.shape, device=device)
    b = torch.as_tensor(rhs).to(device)
    x = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = (r * r).sum(0)
    r0 = torch.sqrt(torch.clamp(rr, min=1e-300))
    for _ in range(maxiter):
        if bool((torch.sqrt(rr) <= tol * r0).all()):
            break
        ap = At @ p
        pap = (p * ap).sum(0)
        alpha = torch.where(pap > 0, rr / torch.where(pap > 0, pap, torch.ones_like(pap)), torch.zeros_like(pap))
        x += alpha * p
        r -= alpha * ap
        rr_new = (r * r).sum(0)
        beta = torch.where(rr > 0, rr_new / torch.where(rr > 0, rr, torch.ones_like(rr)), torch.zeros_like(rr))
        p = r + beta * p
        rr = rr_new
    return x.cpu().numpy()


def sample_element_E(elements: np.ndarray, n_smp: int, seed: int, E0: float = 7.0e6,
                     cov: float = 0.10, tau: float | None = None, device=None) -> GMRFSample:
    """Draw ``n_smp`` per-element Young's modulus samples (layout ``[E, S]``, sample-fastest).

    The random numbers always come from the seeded host generator; ``device`` (e.g. ``"cuda"``)
    only moves the conjugate-gradient solve there for large ensembles.
    """
    A, B = dual_laplacian_system(elements)
    n_elm = A.shape[0]
    var = math.log1p(cov * cov)
    if tau is None:
        tau = mean_inverse_diagonal(A, device=device) / var
    rng = np.random.default_rng(seed)
    xi1 = rng.standard_normal((n_elm, n_smp))
    xi2 = rng.standard_normal((B.shape[0], n_smp))
    rhs = xi1 + B.T @ xi2
    y = _cg(A, rhs) if device is None else _cg_torch(A, rhs, device)
    z = y / math.sqrt(tau)
    E = E0 * np.exp(z - 0.5 * var)
    return GMRFSample(np.ascontiguousarray(E), np.ascontiguousarray(z), float(tau), var)
