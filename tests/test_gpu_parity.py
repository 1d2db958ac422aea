"""GPU path vs the CPU oracle on the 64x32 tube ensemble (and variants) through the C ABI.

Bars: block structure bit-exact; assembled values, lumped mass and load within 1e-13 relative;
displacements after N steps within 1e-10 relative max-norm (max |u_gpu - u_cpu| / max |u_cpu|).
"""
import numpy as np
import pytest

from cvfes_amd.configs import get_case
from oracle import solid_oracle as so

pytestmark = pytest.mark.gpu

REL_U = 1e-10
REL_ASSEMBLY = 1e-13


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


_CASE_IDS = iter(range(1 << 30))


class Case:
    """Inputs for one ensemble shared by the GPU system and the oracle."""

    def __init__(self, name, n_smp=None, seed=None, node_E=False, storage="full"):
        import torch  # noqa: F401
        from cvfes_amd.system import MembraneSystem

        c = get_case(name)
        self.uid = next(_CASE_IDS)       # oracle cache key (id() values are reused after collection)
        self.case = c
        self.S = n_smp or c.n_smp
        self.nodes, self.el, self.bdy = c.geometry()
        self.N = self.nodes.shape[0]
        rng = np.random.default_rng(seed if seed is not None else c.seed)
        if node_E:
            self.vE = c.E0 * np.exp(0.1 * rng.standard_normal((self.N, self.S)))
            self.elmE = None
        else:
            from cvfes_amd.gmrf import sample_element_E
            self.elmE = sample_element_E(self.el, self.S, seed=c.seed).E
            self.vE = np.full((self.N, self.S), c.E0)
        self.h = np.full((self.N, self.S), c.thickness)
        self.sys = MembraneSystem(self.nodes, self.el, self.bdy, self.S, vthickness=self.h,
                                  vE=self.vE if node_E else None, elmE=self.elmE,
                                  density=c.density, nu=c.nu, storage=storage)
        # Oracle assembly in the reference layout.
        self.indptr, self.indices = so.sparse_info_loops(self.el, self.N)
        elmGE, elmGThick = so.gauss_properties(self.vE, self.h, self.el)
        coefK = so.coef_k(elmGThick, elmGE) if node_E else so.coef_k_element_E(elmGThick, self.elmE)
        self.coefK = coefK
        self.M, self.K, self.f1, self.LM = so.assemble_reference(
            self.nodes, self.el, 1.0, so.d_matrix(c.nu), coefK, elmGThick, c.density,
            self.indptr, self.indices)
        self.bdy_dofs = (3 * self.bdy[:, None] + np.arange(3)).ravel()


@pytest.fixture(scope="module")
def case1(require_gpu):
    return Case("tube-64x32-s16")


def test_structure_bit_exact(case1):
    assert np.array_equal(case1.sys.indptr.cpu().numpy().astype(np.int64), case1.indptr)
    assert np.array_equal(case1.sys.indices.cpu().numpy().astype(np.int64), case1.indices)


def test_coef_matches_oracle(case1):
    assert rel_err(case1.sys.coef.cpu().numpy(), case1.coefK) < REL_ASSEMBLY


def test_assembled_values_match_oracle(case1):
    assert rel_err(case1.sys.values_reference_layout(), case1.K) < REL_ASSEMBLY


def test_mass_and_load_match_oracle(case1):
    assert rel_err(case1.sys.lumped_mass_dofs(), case1.LM) < REL_ASSEMBLY
    assert rel_err(case1.sys.f1.cpu().numpy().ravel(), case1.f1) < REL_ASSEMBLY


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_spmv_matches_oracle(case1, method):
    import torch
    rng = np.random.default_rng(7)
    x = rng.standard_normal((case1.N, 3, case1.S))
    y = case1.sys.spmv(torch.as_tensor(x).cuda(), method).cpu().numpy()
    yref = so.multiply_by_vector(case1.indptr, case1.indices, case1.K, x.reshape(-1, case1.S))
    assert rel_err(y.reshape(-1, case1.S), yref) < 1e-13


_ORACLE_CACHE = {}


def _run_both(case, nsteps, method, beta=0.0, damp=0.0, dt=None, record=()):
    c = case.case
    dt = c.dt if dt is None else dt
    p = c.pressures(0, nsteps, dt)
    u0 = np.zeros(3 * case.N)
    sysm = case.sys
    sysm.initial_state(u0, u0, sysm.loads(p), dt, beta=beta, damp=damp, method=method)
    out = {}
    done = 0
    for k in sorted(set(record) | {nsteps}):
        sysm.step(p, dt, beta=beta, damp=damp, method=method, offset=done, nsteps=k - done)
        done = k
        out[k] = sysm.u.cpu().numpy().reshape(-1, case.S)
    key = (case.uid, nsteps, beta, damp, dt, tuple(sorted(record)))
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = so.run(case.indptr, case.indices, case.K, case.LM, case.f1, p, u0, u0,
                                    dt, beta=beta, damp=damp, bdy_dofs=case.bdy_dofs, bdy_u=0.0,
                                    record=set(record) | {nsteps})[2]
    return out, _ORACLE_CACHE[key]


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_1000_steps_match_oracle(case1, method):
    out, snaps = _run_both(case1, 1000, method, record=(1, 10, 100))
    for k, u in out.items():
        assert np.all(np.isfinite(u))
        assert rel_err(u, snaps[k]) < REL_U, f"step {k}"
    # The motion is not trivial: the tube inflates.
    assert np.max(np.abs(snaps[1000])) > 1e-6


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_damped_steps_match_oracle(case1, method):
    out, snaps = _run_both(case1, 300, method, beta=1e-6, damp=50.0, dt=3e-6, record=(1, 50))
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U, f"step {k}"


def test_node_field_modulus_and_odd_sample_count(require_gpu):
    c = Case("tube-64x32-s16", n_smp=5, node_E=True)
    assert rel_err(c.sys.coef.cpu().numpy(), c.coefK) < REL_ASSEMBLY
    out, snaps = _run_both(c, 200, "stored", record=(7,))
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U
    out, snaps = _run_both(c, 200, "matfree", record=(7,))
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U


@pytest.fixture(scope="module")
def case128(require_gpu):
    return Case("tube-64x32-s16", n_smp=128, seed=9)


@pytest.mark.parametrize("variant", ["generic", "rowload", "rowload_nt", "rowload1", "rowload1_nt",
                                     "lds", "lds2", "sym", "sym_nt", "sym_ntl",
                                     "cached", "cached1", "ebe"])
@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_kernel_variants_match_oracle(case128, method, variant):
    """Every step-kernel variant at 128 samples (wave-uniform rows) against the oracle."""
    case128.sys.set_variant(variant)
    try:
        out, snaps = _run_both(case128, 120, method, beta=1e-6, damp=30.0, dt=3e-6, record=(1, 40))
    finally:
        case128.sys.set_variant("auto")
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U, f"step {k}"


@pytest.mark.parametrize("n_smp", [128, 100, 5])
def test_symmetric_storage_matches_oracle(require_gpu, n_smp):
    """Stored values kept as upper blocks only (no full copy): the step (symmetric kernel at 128
    samples, the generic kernel reading the mirrored blocks transposed otherwise), the generic
    kernel at 128, and the product all match the oracle; the blocks rebuilt from the upper
    storage equal the oracle's assembly."""
    import torch
    c = Case("tube-64x32-s16", n_smp=n_smp, seed=5, storage="sym")
    assert c.sys.values is None and c.sys.uvalues is not None
    assert rel_err(c.sys.values_reference_layout(), c.K) < REL_ASSEMBLY
    expect = "sym" if n_smp == 128 else "generic"
    from cvfes_amd import _lib
    assert c.sys.lib.cvfes_step_kernel(c.sys._sys, _lib.METHOD_STORED) == _lib.VARIANTS[expect]
    for variant in ("auto", "generic"):
        c.sys.set_variant(variant)
        out, snaps = _run_both(c, 120, "stored", beta=1e-6, damp=30.0, dt=3e-6, record=(1, 40))
        for k, u in out.items():
            assert rel_err(u, snaps[k]) < REL_U, (variant, k)
    c.sys.set_variant("auto")
    x = np.random.default_rng(2).standard_normal((c.N, 3, c.S))
    y = c.sys.spmv(torch.as_tensor(x).cuda(), "stored").cpu().numpy()
    yref = so.multiply_by_vector(c.indptr, c.indices, c.K, x.reshape(-1, c.S))
    assert rel_err(y.reshape(-1, c.S), yref) < 1e-13


def test_structure_limits(case128):
    s = case128.sys
    assert (s.max_row_blocks, s.max_offdiag_contrib, s.max_diag_contrib) == (7, 2, 6)


@pytest.mark.parametrize("n_smp", [97, 100])
@pytest.mark.parametrize("method", ["stored", "matfree", "matfree-ebe"])
def test_padded_wavefront_rows_match_oracle(require_gpu, method, n_smp):
    """Sample counts that are not a multiple of the wavefront (rows padded to whole wavefronts)."""
    c = Case("tube-64x32-s16", n_smp=n_smp, seed=3)
    if method == "matfree-ebe":
        method = "matfree"
        c.sys.set_variant("ebe")
    out, snaps = _run_both(c, 150, method, beta=1e-6, damp=20.0, dt=3e-6, record=(1, 60))
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U, f"step {k}"
    import torch
    x = np.random.default_rng(1).standard_normal((c.N, 3, c.S))
    y = c.sys.spmv(torch.as_tensor(x).cuda(), method).cpu().numpy()
    yref = so.multiply_by_vector(c.indptr, c.indices, c.K, x.reshape(-1, c.S))
    assert rel_err(y.reshape(-1, c.S), yref) < 1e-13


@pytest.mark.parametrize("n_smp", [1, 2, 3])
@pytest.mark.parametrize("method", ["stored", "matfree", "matfree-ebe"])
def test_smallest_sample_counts_match_oracle(require_gpu, method, n_smp):
    """One, two and three samples: one-lane rows, two samples on one lane, an odd count."""
    import torch
    c = Case("tube-64x32-s16", n_smp=n_smp, seed=11)
    if method == "matfree-ebe":
        method = "matfree"
        c.sys.set_variant("ebe")
    out, snaps = _run_both(c, 120, method, beta=1e-6, damp=20.0, dt=3e-6, record=(1, 40))
    for k, u in out.items():
        assert rel_err(u, snaps[k]) < REL_U, f"step {k}"
    # zero steps leave the state untouched
    before = [b.clone() for b in c.sys.state]
    c.sys.step(c.case.pressures(0, 1, 3e-6), 3e-6, beta=1e-6, damp=20.0, method=method, nsteps=0)
    for b0, b1 in zip(before, c.sys.state):
        assert torch.equal(b0, b1)
    x = np.random.default_rng(2).standard_normal((c.N, 3, c.S))
    y = c.sys.spmv(torch.as_tensor(x).cuda(), method).cpu().numpy()
    yref = so.multiply_by_vector(c.indptr, c.indices, c.K, x.reshape(-1, c.S))
    assert rel_err(y.reshape(-1, c.S), yref) < 1e-13


@pytest.mark.parametrize("method,variant", [("stored", "rowload_nt"), ("stored", "sym"),
                                            ("stored", "sym_nt"), ("stored", "sym_ntl"),
                                            ("stored", "generic"), ("matfree", "generic")])
def test_row_order_does_not_change_result(case128, method, variant):
    """Visiting the rows in a bisection, strip or random order gives bit-identical states."""
    s = case128.sys
    outs = []
    perm = np.random.default_rng(0).permutation(s.n_nodes)
    for order in ("natural", "bisection", "strips", perm):
        s.set_variant(variant)
        s.set_row_order(order)
        try:
            out, snaps = _run_both(case128, 40, method, beta=1e-6, damp=30.0, dt=3e-6, record=(40,))
        finally:
            s.set_row_order("natural")
            s.set_variant("auto")
        outs.append(out[40])
        assert rel_err(out[40], snaps[40]) < REL_U
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


@pytest.mark.parametrize("variant", ["sym", "sym_nt", "sym_ntl"])
def test_row_records_do_not_change_result(require_gpu, variant):
    """The symmetric kernels reading each row's structure from its record give the states of the
    kernels reading indptr/indices/vblk, bit for bit, in the natural and a strip order."""
    c = Case("tube-64x32-s16", n_smp=128, seed=13, storage="sym")
    s = c.sys
    assert s.row_records is not None and s.record_width == 16
    s.set_variant(variant)
    for order in ("natural", "strips"):
        s.set_row_order(order, width=16)
        outs = []
        for records in (True, False):
            if not records:
                s.row_records, s.record_width = None, 0
                s._sys = s._make_struct()
            out, snaps = _run_both(c, 40, "stored", beta=1e-6, damp=30.0, dt=3e-6, record=(40,))
            outs.append(out[40])
            assert rel_err(out[40], snaps[40]) < REL_U
        assert np.array_equal(outs[0], outs[1])
        s._build_records()


@pytest.mark.parametrize("method,variant", [("stored", "rowload_nt"), ("stored", "sym_ntl"),
                                            ("stored", "generic"), ("matfree", "generic")])
def test_occupancy_cap_does_not_change_result(case128, method, variant):
    """Capping the waves per SIMD (unused LDS per workgroup) gives bit-identical states."""
    s = case128.sys
    outs = []
    for w in (0, 2, 4, 7):
        s.set_variant(variant)
        s.set_occupancy(w)
        try:
            out, snaps = _run_both(case128, 40, method, beta=1e-6, damp=30.0, dt=3e-6, record=(40,))
        finally:
            s.set_occupancy(0)
            s.set_variant("auto")
        outs.append(out[40])
        assert rel_err(out[40], snaps[40]) < REL_U
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    with pytest.raises(ValueError):
        s.set_occupancy(1)


@pytest.mark.parametrize("method,variant", [("stored", "generic"), ("stored", "rowload_nt"),
                                            ("stored", "sym_ntl"), ("matfree", "generic"),
                                            ("matfree", "lds"), ("matfree", "cached"),
                                            ("matfree", "cached1"), ("matfree", "ebe")])
def test_isolated_node_every_kernel(require_gpu, method, variant):
    """A node in no element (empty block row, zero mass) next to a tube: every kernel keeps it at
    zero and matches the oracle elsewhere."""
    import torch  # noqa: F401
    from cvfes_amd.mesh import tube_geometry
    from cvfes_amd.system import MembraneSystem
    nodes, el, bdy = tube_geometry(24, 8, 0.3, 2.0)
    nodes = np.vstack([nodes, [[5.0, 5.0, 5.0]]])            # isolated last node
    N, S = nodes.shape[0], 128
    rng = np.random.default_rng(4)
    E = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, 0.07), elmE=E, density=1.0, nu=0.5)
    sysm.set_variant(variant)
    nsteps, dt = 60, 3e-6
    p = 1.33e4 * (1 + 0.1 * np.sin(np.arange(nsteps)))
    z = np.zeros(3 * N)
    sysm.initial_state(z, z, sysm.loads(p), dt, beta=1e-6, damp=10.0, method=method)
    sysm.step(p, dt, beta=1e-6, damp=10.0, method=method)
    ug = sysm.u.cpu().numpy().reshape(-1, S)
    ip, ix = so.sparse_info_loops(el, N)
    gth = np.full((el.shape[0], S, 3), 0.07)
    _, K, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(0.5), so.coef_k_element_E(gth, E),
                                         gth, 1.0, ip, ix)
    bd = (3 * bdy[:, None] + np.arange(3)).ravel()
    u, _, _ = so.run(ip, ix, K, LM, f1, p, z, z, dt, beta=1e-6, damp=10.0, bdy_dofs=bd)
    assert np.all(ug[-3:] == 0.0)
    assert rel_err(ug, u) < REL_U
