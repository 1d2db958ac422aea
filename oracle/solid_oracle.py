"""NumPy restatement of CVFES's CPU membrane solid path (linear mode) - TEST INFRASTRUCTURE.

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` may use this module.

Each function follows the reference file:line given in its docstring.  Extensions the reference
CPU path does not have are marked EXTENSION and follow the reference GPU kernels or the
configuration text:

* nodal load ``f(t_n) = p(t_n) * f1`` with ``f1`` assembled at unit traction (the CPU path assembles
  ``f`` at ``appTraction = 0``, ``physicsSolver.py:114,234``; ``f`` is linear in the traction,
  ``optimizedSolidAssemble.pyx:263,338-343``);
* Rayleigh stiffness damping ``C = beta*K`` with the backward-difference velocity, which puts
  ``K (u + (beta/dt)(u - up))`` in place of ``K u``;
* the mass-independent damping force ``-damp*(u - up)/dt`` of ``physicsSolverSDampGPUs.cl:426``.

Pinning (what checks each function against something other than its own formula, see
``tests/test_oracle.py``):

* ``sparse_info_loops``: brute-force neighbour sets and closed-form block counts of the tube.
* ``coordinate_transformation`` / ``assemble_reference``: the closed-form block formula of the GPU
  kernels (``physicsSolverSDampGPUs.cl:340-357``), symmetry/PSD, rigid-translation null space, the
  constant-strain patch test, a hand-computed right-triangle stiffness, and the lumped mass / load
  closed forms (total mass = rho*h*area/3, facet areas of the faceted tube, sum of loads = 0).
* ``multiply_by_vector`` / ``multiply_by_1d_vector``: dense mat-vec of the assembled matrix.
* ``initial_state`` / ``step`` / ``run``: closed-form solutions of the central-difference recurrence
  (free flight under constant load, undamped and beta-damped single-degree-of-freedom systems).
* ``gauss_properties`` / ``coef_k``: uniform-field closed form (``coefK = h*E``) and linear-field
  interpolation at edge midpoints.
* ``assemble_update``: sample ``s`` equals ``assemble_reference`` (pinned above) on the mesh moved
  by ``u_s``; rigid translation leaves it unchanged, a rigid rotation R maps every block to
  R K R^T; planted mistakes (geometry of the wrong sample) are caught.
* ``run_updated``: reduces to ``run`` when the interval exceeds the step count; the update
  schedule (first re-formation at step ``interval``) is checked by counting re-formations.
* ``calculate_stress``: zero for rigid translations; a homogeneous in-plane strain on a flat patch
  gives ``E D eps`` in every element; the nodal force ``h A B^T sigma`` (rotated to global) equals
  the element stiffness product ``K_e u_e`` of the pinned assembly.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "sparse_info_loops", "gauss_points", "gauss_properties", "d_matrix", "local_mass", "coef_k",
    "coef_k_element_E", "coordinate_transformation", "assemble_reference", "element_blocks_closed_form",
    "multiply_by_vector", "multiply_by_1d_vector", "initial_state", "step", "run",
    "assemble_update", "run_updated", "calculate_stress",
]


# ----------------------------------------------------------------------------------------------
# Structure
# ----------------------------------------------------------------------------------------------
def sparse_info_loops(elements, n_nodes):
    """Per-node sorted unique neighbour lists -> ``(indptr, indices)`` int64 (``sparse.py:38-57``)."""
    info = [[] for _ in range(n_nodes)]
    for elm in np.asarray(elements).tolist():
        for node in elm:
            info[node].extend(elm)
    indptr = [0]
    indices = []
    for node in range(n_nodes):
        indices.extend(sorted(set(info[node])))
        indptr.append(len(indices))
    return np.array(indptr, dtype=np.int64), np.array(indices, dtype=np.int64)


# ----------------------------------------------------------------------------------------------
# Material / quadrature constants (physicsSolver.py:146-202)
# ----------------------------------------------------------------------------------------------
def gauss_points():
    """Edge-midpoint rule with weights 1/3 (``physicsSolver.py:149-152``)."""
    return np.array([[0.5, 0.5, 0.0, 1.0 / 3.0],
                     [0.5, 0.0, 0.5, 1.0 / 3.0],
                     [0.0, 0.5, 0.5, 1.0 / 3.0]])


def gauss_properties(vE, vthickness, elements):
    """``elmGE``, ``elmGThick`` ``[E, S, 3]``: node fields at the Gauss points (``physicsSolver.py:154-165``)."""
    gps = gauss_points()[:, :3]
    elmVerE = np.swapaxes(vE[elements, :], 1, 2)
    elmGE = np.matmul(elmVerE, gps.T)
    elmVerThick = np.swapaxes(vthickness[elements, :], 1, 2)
    elmGThick = np.ascontiguousarray(np.matmul(elmVerThick, gps.T))
    return elmGE, elmGThick


def d_matrix(v, k=5.0 / 6.0):
    """5x5 membrane + transverse-shear constitutive matrix divided by ``1 - v^2`` (``physicsSolver.py:184-188``)."""
    return np.array([[1.0, v, 0.0, 0.0, 0.0],
                     [v, 1.0, 0.0, 0.0, 0.0],
                     [0.0, 0.0, 0.5 * (1 - v), 0.0, 0.0],
                     [0.0, 0.0, 0.0, 0.5 * k * (1 - v), 0.0],
                     [0.0, 0.0, 0.0, 0.0, 0.5 * k * (1 - v)]]) / (1 - v * v)


def local_mass(density):
    """``localM[q] = N_q^T N_q * w_q * rho`` (``physicsSolver.py:191-197``)."""
    xw = gauss_points()
    lm = np.zeros((3, 9, 9))
    for i in range(3):
        N = np.array([[xw[i][0], 0, 0, xw[i][1], 0, 0, xw[i][2], 0, 0],
                      [0, xw[i][0], 0, 0, xw[i][1], 0, 0, xw[i][2], 0],
                      [0, 0, xw[i][0], 0, 0, xw[i][1], 0, 0, xw[i][2]]])
        lm[i] = np.dot(N.T, N) * xw[i][3] * density
    return lm


def coef_k(elmGThick, elmGE):
    """``coefK[e,s] = sum_q w_q h_q E_q`` (``physicsSolver.py:200-202``)."""
    xw = gauss_points()
    gTE = elmGThick * elmGE
    return np.ascontiguousarray(gTE[:, :, 0] * xw[0][3] + gTE[:, :, 1] * xw[1][3] + gTE[:, :, 2] * xw[2][3])


def coef_k_element_E(elmGThick, elmE):
    """EXTENSION: per-element modulus - ``coefK[e,s] = E[e,s] * sum_q w_q h_q`` (same weights)."""
    xw = gauss_points()
    return np.ascontiguousarray(
        (elmGThick[:, :, 0] * elmE) * xw[0][3] + (elmGThick[:, :, 1] * elmE) * xw[1][3]
        + (elmGThick[:, :, 2] * elmE) * xw[2][3])


# ----------------------------------------------------------------------------------------------
# Element geometry and assembly (optimizedSolidAssemble.pyx)
# ----------------------------------------------------------------------------------------------
def coordinate_transformation(nodes, elements):
    """Element frames ``T [E,3,3]`` (rows e1, e2, e3) (``optimizedSolidAssemble.pyx:82-114``)."""
    X = nodes[elements]                                   # [E, 3 nodes, 3]
    e0 = X[:, 2] - X[:, 1]
    e1 = X[:, 0] - X[:, 2]
    T = np.empty((elements.shape[0], 3, 3))
    nrm = np.sqrt(e0[:, 0] * e0[:, 0] + e0[:, 1] * e0[:, 1] + e0[:, 2] * e0[:, 2])
    T[:, 0] = e0 / nrm[:, None]
    proj = e1[:, 0] * T[:, 0, 0] + e1[:, 1] * T[:, 0, 1] + e1[:, 2] * T[:, 0, 2]
    T[:, 1] = e1 - proj[:, None] * T[:, 0]
    nrm = np.sqrt(T[:, 1, 0] * T[:, 1, 0] + T[:, 1, 1] * T[:, 1, 1] + T[:, 1, 2] * T[:, 1, 2])
    T[:, 1] = T[:, 1] / nrm[:, None]
    T[:, 2, 0] = T[:, 0, 1] * T[:, 1, 2] - T[:, 0, 2] * T[:, 1, 1]
    T[:, 2, 1] = T[:, 0, 2] * T[:, 1, 0] - T[:, 0, 0] * T[:, 1, 2]
    T[:, 2, 2] = T[:, 0, 0] * T[:, 1, 1] - T[:, 0, 1] * T[:, 1, 0]
    return T


def _big_T(T):
    """``TT = blockdiag(T, T, T)`` ``[E,9,9]`` (``optimizedSolidAssemble.pyx:116-127``)."""
    TT = np.zeros((T.shape[0], 9, 9))
    for a in range(3):
        TT[:, 3 * a:3 * a + 3, 3 * a:3 * a + 3] = T
    return TT


def _local_geometry(nodes, elements, T):
    """Local 2-D coordinates, area and B (5x9) (``optimizedSolidAssemble.pyx:273-305``)."""
    X = nodes[elements]
    l0 = np.einsum("eak,ek->ea", X, T[:, 0])              # x of the 3 nodes
    l1 = np.einsum("eak,ek->ea", X, T[:, 1])              # y of the 3 nodes
    area = ((l0[:, 1] * l1[:, 2] - l0[:, 2] * l1[:, 1])
            - (l0[:, 0] * l1[:, 2] - l0[:, 2] * l1[:, 0])
            + (l0[:, 0] * l1[:, 1] - l0[:, 1] * l1[:, 0])) * 0.5
    y23 = l1[:, 1] - l1[:, 2]
    y31 = l1[:, 2] - l1[:, 0]
    y12 = l1[:, 0] - l1[:, 1]
    x32 = l0[:, 2] - l0[:, 1]
    x13 = l0[:, 0] - l0[:, 2]
    x21 = l0[:, 1] - l0[:, 0]
    c1 = 2.0 * area
    B = np.zeros((elements.shape[0], 5, 9))
    for (r, c), val in {(0, 0): y23, (2, 1): y23, (3, 2): y23,
                        (0, 3): y31, (2, 4): y31, (3, 5): y31,
                        (0, 6): y12, (2, 7): y12, (3, 8): y12,
                        (1, 1): x32, (2, 0): x32, (4, 2): x32,
                        (1, 4): x13, (2, 3): x13, (4, 5): x13,
                        (1, 7): x21, (2, 6): x21, (4, 8): x21}.items():
        B[:, r, c] = val / c1
    return area, B


def assemble_reference(nodes, elements, trac, D, coefK, elmGThick, density, indptr, indices,
                       chunk=4096):
    """``OptimizedSolidAssemble`` + ``lumping`` (``optimizedSolidAssemble.pyx:194-207,213-356``).

    Returns ``M, K [nnzb, S, 3, 3]``, ``f [ndof]`` (traction ``trac``) and ``LM [ndof, S]``, in the
    reference layout.  ``tM`` is ``local_mass(density)`` and ``wt`` the Gauss weights, as the
    reference's caller passes them (``physicsSolver.py:234-239``).
    """
    nodes = np.asarray(nodes, float)
    elements = np.asarray(elements, np.int64)
    n_nodes = nodes.shape[0]
    n_smp = coefK.shape[1]
    nnzb = indices.size
    tM = local_mass(density)
    wt = gauss_points()[:, 3]
    M = np.zeros((nnzb, n_smp, 3, 3))
    K = np.zeros((nnzb, n_smp, 3, 3))
    f = np.zeros(3 * n_nodes)
    row_keys = np.repeat(np.arange(n_nodes), np.diff(indptr)) * n_nodes + indices
    for e0 in range(0, elements.shape[0], chunk):
        el = elements[e0:e0 + chunk]
        T = coordinate_transformation(nodes, el)
        TT = _big_T(T)
        area, B = _local_geometry(nodes, el, T)
        BTD = np.einsum("eki,kj->eij", B, D)               # dotT(B, D)
        BTDB = np.einsum("eik,ekj->eij", BTD, B)           # dot(BTD, B)
        cK = coefK[e0:e0 + chunk]
        lK = BTDB[:, None] * cK[:, :, None, None] * area[:, None, None, None]
        lM = np.einsum("kij,esk,k->esij", tM, elmGThick[e0:e0 + chunk], wt) * area[:, None, None, None]
        gK = np.einsum("eki,eskl,elj->esij", TT, lK, TT)   # TT^T lK TT
        gM = np.einsum("eki,eskl,elj->esij", TT, lM, TT)
        lf = np.zeros((el.shape[0], 9))
        lf[:, 2] = lf[:, 5] = lf[:, 8] = trac / 3.0 * area
        gf = np.einsum("eji,ej->ei", TT, lf)               # dotT1(TT, lf)
        # assembling(): element order, binary search of (row, col) in the block structure.
        for a in range(3):
            for i in range(3):
                np.add.at(f, 3 * el[:, a] + i, gf[:, 3 * a + i])
            for b in range(3):
                pos = np.searchsorted(row_keys, el[:, a] * n_nodes + el[:, b])
                np.add.at(K, pos, gK[:, :, 3 * a:3 * a + 3, 3 * b:3 * b + 3])
                np.add.at(M, pos, gM[:, :, 3 * a:3 * a + 3, 3 * b:3 * b + 3])
    # lumping(): LM[3i+ik, s] = sum over the row's blocks j and columns jk of M[j, s, ik, jk].
    LM = np.zeros((3 * n_nodes, n_smp))
    row_of = np.repeat(np.arange(n_nodes), np.diff(indptr))
    msum = M.sum(axis=3)                                   # [nnzb, S, 3]
    for ik in range(3):
        np.add.at(LM[ik::3], row_of, msum[:, :, ik])
    return M, K, f, LM


def element_blocks_closed_form(nodes, elements, v, k=5.0 / 6.0):
    """Unit-coefficient element matrices ``[E, 9, 9]`` from the factorised block formula of
    ``physicsSolverSDampGPUs.cl:340-357`` (``K_e = coefK * this``)."""
    T = coordinate_transformation(nodes, elements)
    X = nodes[elements]
    lx = np.einsum("eak,ek->ea", X, T[:, 0])
    ly = np.einsum("eak,ek->ea", X, T[:, 1])
    area = 0.5 * (lx[:, 0] * (ly[:, 1] - ly[:, 2]) - lx[:, 1] * (ly[:, 0] - ly[:, 2])
                  + lx[:, 2] * (ly[:, 0] - ly[:, 1]))
    bb = np.stack([ly[:, 1] - ly[:, 2], ly[:, 2] - ly[:, 0], ly[:, 0] - ly[:, 1]], axis=1)
    cc = np.stack([lx[:, 2] - lx[:, 1], lx[:, 0] - lx[:, 2], lx[:, 1] - lx[:, 0]], axis=1)
    g = 0.5 * (1 - v)
    gk = 0.5 * k * (1 - v)
    scale = 1.0 / (4.0 * area * (1 - v * v))
    out = np.zeros((elements.shape[0], 9, 9))
    for a in range(3):
        for b in range(3):
            kl = np.zeros((elements.shape[0], 3, 3))
            kl[:, 0, 0] = bb[:, a] * bb[:, b] + g * cc[:, a] * cc[:, b]
            kl[:, 0, 1] = v * bb[:, a] * cc[:, b] + g * cc[:, a] * bb[:, b]
            kl[:, 1, 0] = v * cc[:, a] * bb[:, b] + g * bb[:, a] * cc[:, b]
            kl[:, 1, 1] = cc[:, a] * cc[:, b] + g * bb[:, a] * bb[:, b]
            kl[:, 2, 2] = gk * (bb[:, a] * bb[:, b] + cc[:, a] * cc[:, b])
            out[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] = (
                np.einsum("eki,ekl,elj->eij", T, kl, T) * scale[:, None, None])
    return out, area, T[:, 2].copy()


# ----------------------------------------------------------------------------------------------
# Sparse products (optimizedSolidAssemble.pyx:1250-1299)
# ----------------------------------------------------------------------------------------------
def multiply_by_vector(indptr, indices, K, v):
    """``res[3a+i, s] = sum_p sum_j K[p,s,i,j] v[3*col(p)+j, s]`` (``optimizedSolidAssemble.pyx:1250-1274``)."""
    n_nodes = indptr.size - 1
    n_smp = K.shape[1]
    vb = v.reshape(n_nodes, 3, n_smp)[indices]            # [nnzb, 3(j), S]
    prod = np.einsum("psij,pjs->pis", K, vb)               # [nnzb, 3(i), S]
    row_of = np.repeat(np.arange(n_nodes), np.diff(indptr))
    res = np.zeros((n_nodes, 3, n_smp))
    np.add.at(res, row_of, prod)
    return res.reshape(3 * n_nodes, n_smp)


def multiply_by_1d_vector(indptr, indices, K, v):
    """Sample-independent ``v [ndof]`` (``optimizedSolidAssemble.pyx:1280-1299``)."""
    n_smp = K.shape[1]
    return multiply_by_vector(indptr, indices, K, np.repeat(np.asarray(v, float)[:, None], n_smp, axis=1))


# ----------------------------------------------------------------------------------------------
# Time integration (physicsSolver.py:206-229,263-292)
# ----------------------------------------------------------------------------------------------
def _safe_divide(a, b):
    """``np.divide(a, b, where=b != 0)`` with zeros elsewhere (``physicsSolver.py:221,281``)."""
    return np.divide(a, b, out=np.zeros_like(a), where=b != 0)


def initial_state(indptr, indices, K, LM, f, u0, du0, dt, beta=0.0, damp=0.0):
    """``u_0``, ``u_{-1}`` and ``LHS`` (``physicsSolver.py:209-229``, LC = 0).

    ``u0``/``du0`` are ``[ndof]`` (tiled over samples as ``physicsSolver.py:214-215``); ``f`` is the
    load at the start time ``[ndof]`` or ``[ndof, S]``.  EXTENSIONS: ``beta`` adds ``beta*K du0`` to
    the stiffness force and ``damp`` the force ``-damp*du0``.
    """
    n_smp = K.shape[1]
    x = np.asarray(u0, float) + beta * np.asarray(du0, float)
    Ku = multiply_by_1d_vector(indptr, indices, K, x)
    du = np.tile(np.asarray(du0, float)[:, None], (1, n_smp))
    u = np.tile(np.asarray(u0, float)[:, None], (1, n_smp))
    f = np.asarray(f, float)
    if f.ndim == 1:
        f = f[:, None]
    RHS = f - damp * du - Ku
    LHS = np.copy(LM)
    ddu = _safe_divide(RHS, LHS)
    up = u - dt * du + 0.5 * (dt ** 2) * ddu
    return u, up, LHS


def step(indptr, indices, K, LM, LHS, f, u, up, dt, beta=0.0, damp=0.0, bdy_dofs=None, bdy_u=0.0):
    """One explicit central-difference step (``physicsSolver.py:263-292``, LC = 0, linear mode).

    ``Ku = (dt^2 K) x`` with ``x = u + (beta/dt)(u - up)`` (EXTENSION, x = u when beta = 0);
    ``RHS = dt^2 (f - damp (u - up)/dt) + 2 LM u - Ku - LM up`` (damping force from
    ``physicsSolverSDampGPUs.cl:426``); ``u+ = RHS / LHS`` where ``LHS != 0``; then the Dirichlet
    values ``u+[bdy_dofs] = bdy_u`` (``physicsSolver.py:327-328``).
    """
    f = np.asarray(f, float)
    if f.ndim == 1:
        f = f[:, None]
    x = u if beta == 0.0 else u + (beta / dt) * (u - up)
    A1 = (dt ** 2) * (f - damp * (u - up) / dt) if damp != 0.0 else (dt ** 2) * f
    Ku = multiply_by_vector(indptr, indices, (dt ** 2) * K, x)
    A2 = 2.0 * LM * u - Ku
    A3 = (0.0 - LM) * up
    RHS = A1 + A2 + A3
    unew = _safe_divide(RHS, LHS)
    if bdy_dofs is not None:
        unew[bdy_dofs, :] = bdy_u
    return unew


def run(indptr, indices, K, LM, f1, pressures, u0, du0, dt, beta=0.0, damp=0.0, bdy_dofs=None,
        bdy_u=0.0, record=None, extra_forces=None, p_init=None, extra_init=None):
    """Run ``len(pressures)`` steps; step ``n`` uses the load ``pressures[n] * f1 (+ extra_forces[n])``.

    The initial acceleration uses ``p_init * f1 + extra_init`` (default: the step-0 load).
    ``extra_forces`` ``[nsteps, ndof]`` are EXTENSIONS standing for the fluid wall traction
    (``physicsSolverSDampGPUs.py:384-390``).  Returns ``(u_N, u_{N-1}, snapshots)`` where
    ``snapshots[k]`` is ``u`` after ``k`` steps for each ``k`` in ``record``.
    """
    pressures = np.asarray(pressures, float)
    p0 = pressures[0] if p_init is None else p_init
    f0 = p0 * f1
    if extra_init is not None:
        f0 = f0 + extra_init
    elif extra_forces is not None:
        f0 = f0 + extra_forces[0]
    u, up, LHS = initial_state(indptr, indices, K, LM, f0, u0, du0, dt, beta, damp)
    snaps = {}
    record = set(record or ())
    if 0 in record:
        snaps[0] = u.copy()
    for n, p in enumerate(pressures):
        fn = p * f1 if extra_forces is None else p * f1 + extra_forces[n]
        unew = step(indptr, indices, K, LM, LHS, fn, u, up, dt, beta, damp, bdy_dofs, bdy_u)
        up, u = u, unew
        if n + 1 in record:
            snaps[n + 1] = u.copy()
    return u, up, snaps


# ----------------------------------------------------------------------------------------------
# Updated geometry (optimizedSolidAssemble.pyx:606-819; physicsSolver.py:294-324)
# ----------------------------------------------------------------------------------------------


def _frames_batched(X):
    """Frames ``T [..., 3, 3]`` of triangles ``X [..., 3 nodes, 3]`` (``pyx:606-651``)."""
    e0 = X[..., 2, :] - X[..., 1, :]
    e1 = X[..., 0, :] - X[..., 2, :]
    t0 = e0 / np.sqrt((e0 * e0).sum(-1))[..., None]
    t1 = e1 - (e1 * t0).sum(-1)[..., None] * t0
    t1 = t1 / np.sqrt((t1 * t1).sum(-1))[..., None]
    t2 = np.stack([t0[..., 1] * t1[..., 2] - t0[..., 2] * t1[..., 1],
                   t0[..., 2] * t1[..., 0] - t0[..., 0] * t1[..., 2],
                   t0[..., 0] * t1[..., 1] - t0[..., 1] * t1[..., 0]], axis=-1)
    return np.stack([t0, t1, t2], axis=-2)


def _local_batched(X, T):
    """Area and B ``[..., 5, 9]`` of triangles ``X`` in frames ``T`` (``pyx:754-779``)."""
    lx = np.einsum("...ak,...k->...a", X, T[..., 0, :])
    ly = np.einsum("...ak,...k->...a", X, T[..., 1, :])
    area = ((lx[..., 1] * ly[..., 2] - lx[..., 2] * ly[..., 1])
            - (lx[..., 0] * ly[..., 2] - lx[..., 2] * ly[..., 0])
            + (lx[..., 0] * ly[..., 1] - lx[..., 1] * ly[..., 0])) * 0.5
    c1 = 2.0 * area
    y = [ly[..., 1] - ly[..., 2], ly[..., 2] - ly[..., 0], ly[..., 0] - ly[..., 1]]
    x = [lx[..., 2] - lx[..., 1], lx[..., 0] - lx[..., 2], lx[..., 1] - lx[..., 0]]
    B = np.zeros(X.shape[:-2] + (5, 9))
    for a in range(3):
        B[..., 0, 3 * a] = y[a] / c1
        B[..., 2, 3 * a + 1] = y[a] / c1
        B[..., 3, 3 * a + 2] = y[a] / c1
        B[..., 1, 3 * a + 1] = x[a] / c1
        B[..., 2, 3 * a] = x[a] / c1
        B[..., 4, 3 * a + 2] = x[a] / c1
    return area, B


def _big_T_batched(T):
    TT = np.zeros(T.shape[:-2] + (9, 9))
    for a in range(3):
        TT[..., 3 * a:3 * a + 3, 3 * a:3 * a + 3] = T
    return TT


def assemble_update(nodes_s, elements, trac, D, coefK, indptr, indices, chunk=2048):
    """``OptimizedSolidAssembleUpdate`` (``optimizedSolidAssemble.pyx:710-819``, with
    ``assemblingUpdate`` ``:657-704``): per-sample stiffness ``K [nnzb, S, 3, 3]`` and traction load
    ``f [ndof, S]`` on the per-sample geometry ``nodes_s [S, N, 3]`` (the mass is not re-formed)."""
    nodes_s = np.asarray(nodes_s, float)
    elements = np.asarray(elements, np.int64)
    n_smp, n_nodes = nodes_s.shape[0], nodes_s.shape[1]
    nnzb = indices.size
    K = np.zeros((nnzb, n_smp, 3, 3))
    f = np.zeros((3 * n_nodes, n_smp))
    row_keys = np.repeat(np.arange(n_nodes), np.diff(indptr)) * n_nodes + indices
    for e0 in range(0, elements.shape[0], chunk):
        el = elements[e0:e0 + chunk]
        X = np.moveaxis(nodes_s[:, el], 0, 1)                # [e, S, 3 nodes, 3]
        T = _frames_batched(X)                               # [e, S, 3, 3]
        TT = _big_T_batched(T)
        area, B = _local_batched(X, T)
        BTDB = np.einsum("eski,kl,eslj->esij", B, D, B)
        lK = BTDB * (coefK[e0:e0 + chunk] * area)[..., None, None]
        gK = np.einsum("eski,eskl,eslj->esij", TT, lK, TT)
        lf = np.zeros(area.shape + (9,))
        lf[..., 2] = lf[..., 5] = lf[..., 8] = trac / 3.0 * area
        gf = np.einsum("esji,esj->esi", TT, lf)
        for a in range(3):
            for i in range(3):
                np.add.at(f, 3 * el[:, a] + i, gf[:, :, 3 * a + i])
            for b in range(3):
                pos = np.searchsorted(row_keys, el[:, a] * n_nodes + el[:, b])
                np.add.at(K, pos, gK[:, :, 3 * a:3 * a + 3, 3 * b:3 * b + 3])
    return K, f


def run_updated(nodes, elements, D, coefK, indptr, indices, K0, LM, f1, pressures, u0, du0, dt,
                interval, beta=0.0, damp=0.0, bdy_dofs=None, bdy_u=0.0, record=None):
    """Explicit steps with the stiffness and the unit-traction load re-formed on the moved geometry
    ``X + u_s`` every ``interval`` steps (``physicsSolver.py:294-324``): a step counter starts at 0
    after the initial assembly; before each step, if the counter has reached ``interval`` the
    system is re-formed from the current ``u`` and the counter set to 1, otherwise it is
    incremented (so the first re-formation happens before step ``interval``, 0-based).  The load of
    step ``n`` is ``pressures[n] * f1`` with ``f1`` the unit-traction load of the current geometry
    (EXTENSION: the reference CPU path applies its fixed ``appTraction``).  Returns
    ``(u_N, u_{N-1}, snapshots, n_updates)``."""
    nodes = np.asarray(nodes, float)
    n_smp = K0.shape[1]
    u, up, LHS = initial_state(indptr, indices, K0, LM, pressures[0] * f1, u0, du0, dt, beta, damp)
    K, fs = K0, np.asarray(f1, float)[:, None]
    counter, n_up = 0, 0
    snaps = {}
    record = set(record or ())
    for n, p in enumerate(np.asarray(pressures, float)):
        if counter >= interval:
            moved = nodes[None] + u.T.reshape(n_smp, -1, 3)
            K, fs = assemble_update(moved, elements, 1.0, D, coefK, indptr, indices)
            counter, n_up = 1, n_up + 1
        else:
            counter += 1
        unew = step(indptr, indices, K, LM, LHS, p * fs, u, up, dt, beta, damp, bdy_dofs, bdy_u)
        up, u = u, unew
        if n + 1 in record:
            snaps[n + 1] = u.copy()
    return u, up, snaps, n_up


def _to_global(T, St):
    """``T^T S T``: local stress tensor to global axes (``pyx:1207-1208``)."""
    return np.einsum("eski,eskl,eslj->esij", T, St, T)


def calculate_stress(nodes_s, elements, D, ave_E, u):
    """``OptimizedCalculateStress`` (``optimizedSolidAssemble.pyx:1116-1223``): global stress
    components ``(xx, yy, xy, xz, yz)`` ``[E, S, 5]`` of the displacement ``u [ndof, S]`` on the
    per-sample geometry ``nodes_s [S, N, 3]`` (``mesh.updateNodes``) with the element-average modulus
    ``ave_E [E, S]`` (``physicsSolver.py:416-424``)."""
    nodes_s = np.asarray(nodes_s, float)
    elements = np.asarray(elements, np.int64)
    n_smp = u.shape[1]
    X = np.moveaxis(nodes_s[:, elements], 0, 1)              # [E, S, 3 nodes, 3]
    T = _frames_batched(X)
    TT = _big_T_batched(T)
    area, B = _local_batched(X, T)
    dofs = (3 * elements[:, :, None] + np.arange(3)).reshape(-1, 9)   # [E, 9]
    gU = np.moveaxis(u[dofs], 2, 1)                           # [E, S, 9]
    lU = np.einsum("esij,esj->esi", TT, gU)                   # dot1(TT, gU)
    DB = np.einsum("ik,eskj->esij", D, B)
    ls = np.einsum("esij,esj->esi", DB, lU) * ave_E[..., None]
    St = np.zeros(ls.shape[:2] + (3, 3))
    St[..., 0, 0] = ls[..., 0]
    St[..., 1, 1] = ls[..., 1]
    St[..., 0, 1] = St[..., 1, 0] = ls[..., 2]
    St[..., 0, 2] = St[..., 2, 0] = ls[..., 3]
    St[..., 1, 2] = St[..., 2, 1] = ls[..., 4]
    G = _to_global(T, St)
    return np.stack([G[..., 0, 0], G[..., 1, 1], G[..., 0, 1], G[..., 0, 2], G[..., 1, 2]], axis=-1)
