"""Oracle-side inputs shared by the GPU tests that compare with the C oracle at large sizes."""
import numpy as np


def oracle_mass_load(nodes, el, density, thickness, nu, S):
    """Lumped mass ``LM [3N, S]`` and unit load ``f1 [3N]`` from the oracle's element areas and
    normals (closed forms pinned in test_oracle: rho A h / 9 per incident element, A / 3 * normal)."""
    from oracle import solid_oracle as so
    N = nodes.shape[0]
    _, area, normal = so.element_blocks_closed_form(nodes, el, nu)
    lm_node = np.zeros(N)
    f1 = np.zeros((N, 3))
    for a in range(3):
        np.add.at(lm_node, el[:, a], density * area * thickness / 9.0)
        np.add.at(f1, el[:, a], (area / 3.0)[:, None] * normal)
    LM = np.repeat(np.repeat(lm_node, 3)[:, None], S, axis=1)
    return LM, f1.ravel()


def oracle_operator(nodes, el, E, E0, thickness, nu, indptr, indices):
    """Reference-layout K [nnzb, S, 3, 3] for the per-element moduli ``E [E, S]`` (C oracle
    assembly; uniform thickness)."""
    from oracle import cpu_oracle as co
    from oracle import solid_oracle as so
    N, S = nodes.shape[0], E.shape[1]
    _, gth = so.gauss_properties(np.full((N, S), E0), np.full((N, S), thickness), el)
    coefK = so.coef_k_element_E(gth, E)
    return co.assemble_K(nodes, el, so.d_matrix(nu), coefK, np.asarray(indptr, np.int64),
                         np.asarray(indices, np.int64))


def oracle_trajectory(nodes, el, bdy, E, p, dt, beta, E0, thickness, density, nu, indptr, indices,
                      snap_steps):
    """C oracle run from rest (u_0 = du_0 = 0) with pressures ``p`` on the sample columns of ``E``;
    returns the states after ``snap_steps`` steps, ``[len(snap_steps), 3N, S]``."""
    from oracle import cpu_oracle as co
    N, S = nodes.shape[0], E.shape[1]
    ip = np.asarray(indptr, np.int64)
    ix = np.asarray(indices, np.int64)
    K = oracle_operator(nodes, el, E, E0, thickness, nu, ip, ix)
    LM, f1 = oracle_mass_load(nodes, el, density, thickness, nu, S)
    u0 = np.zeros((3 * N, S))
    # u_{-1} = u_0 - dt du_0 + dt^2/2 (f_0 - K u_0) / LM with u_0 = du_0 = 0
    up0 = 0.5 * dt ** 2 * (p[0] * f1[:, None] / LM)
    bd = (3 * np.asarray(bdy)[:, None] + np.arange(3)).ravel()
    _, _, snaps = co.run(ip, ix, K, LM, LM, f1, p, u0, up0, dt, beta=beta, bdy_dofs=bd,
                         snap_steps=snap_steps)
    return snaps
