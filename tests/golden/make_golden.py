"""Write the golden fixture ``tube_12x6_s2.npz``: inputs and the oracle's outputs on a small tube.

The reference ships no golden vectors, known-answer tests or fixtures for this path.  This fixture
freezes the outputs of the oracle (``oracle/solid_oracle.py``, itself pinned by closed forms and
invariants in ``tests/test_oracle.py``) on one small case, so that a later change of the oracle or
of the GPU kernels is caught against committed data.

Inputs: a 12x6 tube (R 0.3 cm, L 1 cm, both end rings fixed), 2 samples of per-element E drawn
from a fixed seed (log-normal, 10 % spread around 7e6), thickness 0.07, density 1, nu 0.5,
beta 1e-6, damping 10, pressure 1.33e4 (1 + 0.1 sin n) at step n, 40 steps of dt 4e-6.
Outputs: block structure, element coefficients, assembled K [nnzb, S, 3, 3], lumped mass, unit
load, and the displacements after 1, 10 and 40 steps.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cvfes_amd.mesh import tube_geometry  # noqa: E402
from oracle import solid_oracle as so     # noqa: E402

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tube_12x6_s2.npz")


def case_inputs():
    nodes, el, bdy = tube_geometry(12, 6, 0.3, 1.0)
    S = 2
    rng = np.random.default_rng(20261202)
    E = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
    nsteps, dt = 40, 4e-6
    p = 1.33e4 * (1.0 + 0.1 * np.sin(np.arange(nsteps)))
    return dict(nodes=nodes, elements=el, boundary=bdy, E=E, thickness=0.07, density=1.0, nu=0.5,
                beta=1e-6, damp=10.0, pressures=p, dt=dt)


def oracle_outputs(c):
    nodes, el, bdy, E = c["nodes"], c["elements"], c["boundary"], c["E"]
    N, S = nodes.shape[0], E.shape[1]
    indptr, indices = so.sparse_info_loops(el, N)
    h = np.full((N, S), c["thickness"])
    _, gth = so.gauss_properties(np.full((N, S), 7e6), h, el)
    coefK = so.coef_k_element_E(gth, E)
    _, K, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(c["nu"]), coefK, gth,
                                         c["density"], indptr, indices)
    bdy_dofs = (3 * bdy[:, None] + np.arange(3)).ravel()
    z = np.zeros(3 * N)
    _, _, snaps = so.run(indptr, indices, K, LM, f1, c["pressures"], z, z, c["dt"], beta=c["beta"],
                         damp=c["damp"], bdy_dofs=bdy_dofs, bdy_u=0.0, record={1, 10, 40})
    return dict(indptr=indptr, indices=indices, coefK=coefK, K=K, f1=f1, LM=LM,
                u1=snaps[1], u10=snaps[10], u40=snaps[40])


if __name__ == "__main__":
    c = case_inputs()
    out = oracle_outputs(c)
    np.savez_compressed(PATH, **{f"in_{k}": np.asarray(v) for k, v in c.items()},
                        **{f"out_{k}": v for k, v in out.items()})
    print(f"wrote {PATH} ({os.path.getsize(PATH)} bytes)")
