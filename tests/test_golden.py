"""The committed golden fixture (``tests/golden/tube_12x6_s2.npz``, written by
``tests/golden/make_golden.py``): inputs and the oracle's outputs on a small tube.  The reference
ships no golden vectors for this path; the fixture freezes the pinned oracle's outputs so that a
change of the oracle (CPU test) or of the GPU path (GPU test) is caught against data.

Bars: structure bit-exact; K, lumped mass, load and coefficients within 1e-13 relative; the oracle's
displacements within 1e-12 relative (same algorithm, possibly another NumPy); the GPU path's
displacements within 1e-10 relative max-norm (the parity bar).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tube_12x6_s2.npz")


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def golden():
    d = np.load(FIXTURE, allow_pickle=False)
    return {k: d[k] for k in d.files}


def inputs(g):
    return {k[3:]: (g[k].item() if g[k].ndim == 0 else g[k]) for k in g if k.startswith("in_")}


def test_oracle_reproduces_the_golden_fixture(golden):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import oracle_outputs
    out = oracle_outputs(inputs(golden))
    assert np.array_equal(out["indptr"], golden["out_indptr"])
    assert np.array_equal(out["indices"], golden["out_indices"])
    for k in ("coefK", "K", "f1", "LM"):
        assert rel_err(out[k], golden[f"out_{k}"]) < 1e-13, k
    for k in ("u1", "u10", "u40"):
        assert rel_err(out[k], golden[f"out_{k}"]) < 1e-12, k
    assert np.max(np.abs(golden["out_u40"])) > 1e-9          # the tube moves


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_gpu_path_matches_the_golden_fixture(require_gpu, golden, method):
    from cvfes_amd.system import MembraneSystem
    c = inputs(golden)
    nodes, el, bdy, E = c["nodes"], c["elements"], c["boundary"], c["E"]
    N, S = nodes.shape[0], E.shape[1]
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, c["thickness"]), elmE=E,
                          density=c["density"], nu=c["nu"])
    assert np.array_equal(sysm.indptr.cpu().numpy().astype(np.int64), golden["out_indptr"])
    assert rel_err(sysm.values_reference_layout(), golden["out_K"]) < 1e-13
    assert rel_err(sysm.lumped_mass_dofs(), golden["out_LM"]) < 1e-13
    p, dt = c["pressures"], c["dt"]
    z = np.zeros(3 * N)
    load = sysm.loads(p)
    sysm.initial_state(z, z, load, dt, beta=c["beta"], damp=c["damp"], method=method)
    done = 0
    for k in (1, 10, 40):
        sysm.step(load, dt, beta=c["beta"], damp=c["damp"], method=method, offset=done, nsteps=k - done)
        done = k
        u = sysm.u.cpu().numpy().reshape(-1, S)
        assert rel_err(u, golden[f"out_u{k}"]) < 1e-10, f"step {k}"
