"""Host-side logic on the CPU: configuration files, loads, VTP I/O, meshes, GMRF inputs, the C ABI."""
import glob
import math
import os
import re

import numpy as np
import pytest

from cvfes_amd.cvconfig import CVConfig, parse_cfg, read_cfg
from cvfes_amd.gmrf import dual_laplacian_system, element_dual_edges, sample_element_E
from cvfes_amd.loads import PressureWaveform, TractionSchedule
from cvfes_amd.mesh import stenosed_radius, tube_geometry, tube_solid_mesh
from cvfes_amd.coloring import mesh_coloring
from cvfes_amd import vtkio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG_DIR = os.path.join(ROOT, "cvfes_amd", "cfg")


# ---------------------------------------------------------------- configuration ---------------
def test_parser_sections_lists_comments_bools():
    text = """
a = 1 # comment
name = 'x # not a comment'
lst = 'p.vtp','q.vtp'
flag = yes
[top]
    k = 2.5
    [[sub 1]]
        name = s1
        [[[deep]]]
            v = off
    [[sub 2]]
        name = s2
"""
    c = parse_cfg(text)
    assert c.as_int("a") == 1 and c["name"] == "x # not a comment"
    assert c["lst"] == ["p.vtp", "q.vtp"] and c.as_bool("flag") is True
    assert c["top"].as_float("k") == 2.5 and c["top"].sections == ["sub 1", "sub 2"]
    assert c["top"]["sub 1"]["deep"].as_bool("v") is False
    assert c["top"]["sub 1"].parent is c["top"]


def test_example_cfgs_parse():
    for f in sorted(glob.glob(os.path.join(CFG_DIR, "*.cfg"))):
        cv = CVConfig(read_cfg(f))
        assert cv.solver.method == "transient GPU"
        assert cv.meshes["wall"].generator == "tube"
        assert cv.equations["solid"].EModel == "gmrf_element"
        assert cv.solver.pressure_waveform == "sine"
    cv = CVConfig(read_cfg(os.path.join(CFG_DIR, "tube_256x400_s128.cfg")))
    assert (cv.meshes["wall"].ntheta, cv.meshes["wall"].nz, cv.nSmp) == (256, 400, 128)
    assert cv.solver.beta == 1e-4 and cv.solver.dt == 1.2e-8


@pytest.mark.skipif(not os.path.isdir("/root/reference/cfg"), reason="reference cfgs not present")
def test_reference_cfgs_parse():
    files = sorted(glob.glob("/root/reference/cfg/**/*.cfg", recursive=True))
    assert files
    for f in files:
        cv = CVConfig(read_cfg(f))
        assert "wall" in cv.meshes and cv.solver.dt > 0
        solid = cv.equations["solid"]
        assert solid.E > 0 and 0 <= solid.v < 0.5 + 1e-12


# ---------------------------------------------------------------- loads --------------------------
def test_pressure_waveforms():
    ramp = PressureWaveform("ramp", p0=100.0, ramp_T=0.2)
    assert ramp(0.0) == 0.0 and math.isclose(float(ramp(0.1)), 50.0) and ramp(0.3) == 100.0
    sine = PressureWaveform("sine", mean=1.33e4, amplitude=0.1, period=0.8)
    assert math.isclose(float(sine(0.2)), 1.33e4 * 1.1) and math.isclose(float(sine(0.0)), 1.33e4)


def test_traction_schedules(tmp_path):
    n = 5
    bdy = np.array([0, 4])
    base = str(tmp_path / "trac")
    for k in range(4):
        np.save(f"{base}{k}.npy", np.full((n, 3), float(k + 1)))
    np.save(f"{base}.npy", np.full((n, 3), 7.0))
    # damped form: ramp on snapshot 0 until T, then snapshots from T
    s = TractionSchedule(base, False, 0.2, 0.01, n, bdy)
    key, ga, qa, gb, qb = s.fields(0.1)
    assert gb is None and math.isclose(qa, 0.5) and ga[2, 0] == 1.0 and ga[0, 0] == 0.0
    key, ga, qa, gb, qb = s.fields(0.2 + 0.0125)
    f = qa * ga + qb * gb
    assert math.isclose(f[1, 0], 2.0 + 0.25 * (3.0 - 2.0)) and f[4, 1] == 0.0
    # legacy form: from a zero field at t = 0
    s = TractionSchedule(base, False, 0.2, 0.01, n, bdy, legacy=True)
    key, ga, qa, gb, qb = s.fields(0.005)
    assert math.isclose((qa * ga + qb * gb)[1, 0], 0.5)
    key, ga, qa, gb, qb = s.fields(0.0125)
    assert math.isclose((qa * ga + qb * gb)[1, 0], 1.25)
    # constant form
    s = TractionSchedule(base, True, 0.2, 0.01, n, bdy)
    key, ga, qa, gb, qb = s.fields(0.5)
    assert qa == 1.0 and ga[1, 2] == 7.0


# ---------------------------------------------------------------- VTP ---------------------------
@pytest.mark.parametrize("fmt", ["binary", "ascii"])
def test_vtp_round_trip(tmp_path, fmt):
    nodes, el, _ = tube_geometry(6, 3, 0.3, 1.0)
    u = np.random.default_rng(0).standard_normal((nodes.shape[0], 3))
    pd = vtkio.PolyData(nodes, el, {"u_000": u, "GlobalNodeID": np.arange(nodes.shape[0], dtype=np.int64)},
                        {"xx_000": np.arange(len(el), dtype=float)})
    path = str(tmp_path / "m.vtp")
    vtkio.write_vtp(path, pd, fmt)
    back = vtkio.read_vtp(path)
    assert np.array_equal(back.points, nodes) and np.array_equal(back.polys, el)
    assert np.array_equal(back.point_data["u_000"], u)
    assert np.array_equal(back.point_data["GlobalNodeID"], np.arange(nodes.shape[0]))
    assert np.array_equal(back.cell_data["xx_000"], np.arange(len(el), dtype=float))


def test_save_displacement_names(tmp_path):
    nodes, el, _ = tube_geometry(5, 3, 0.3, 1.0)
    u = np.zeros((2, nodes.shape[0], 3))
    u[1, :, 2] = 1.0
    path = vtkio.save_displacement(str(tmp_path / "disp.vtp"), 7, nodes, el, u)
    assert path.endswith("disp7.vtp")
    back = vtkio.read_vtp(path)
    assert set(back.point_data) == {"u_000", "u_001"} and back.point_data["u_001"][3, 2] == 1.0


# ---------------------------------------------------------------- meshes ------------------------
def test_tube_orientation_counts_boundary():
    nt, nz = 24, 7
    nodes, el, bdy = tube_geometry(nt, nz, 0.4, 2.0)
    assert nodes.shape == (nt * nz, 3) and el.shape == (2 * nt * (nz - 1), 3)
    X = nodes[el]
    n = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
    c = X.mean(axis=1)
    assert np.all(np.einsum("ij,ij->i", n[:, :2], c[:, :2]) > 0)          # outward normals
    assert np.allclose(nodes[bdy[:nt], 2], 0.0) and np.allclose(nodes[bdy[nt:], 2], 2.0)


def test_stenosed_profile():
    prof = stenosed_radius(1.2, 0.8, 12.0, severity=0.5)
    assert math.isclose(float(prof(6.0)), 0.5 * 1.0) and math.isclose(float(prof(0.0)), 1.2, rel_tol=1e-6)
    nodes, el, _ = tube_geometry(16, 50, 1.2, 12.0, prof)
    r = np.hypot(nodes[:, 0], nodes[:, 1])
    assert r.min() < 0.51 and r.max() <= 1.2 + 1e-12


def test_solid_mesh_attributes_and_coloring():
    m = tube_solid_mesh(8, 4, 3)
    for attr in ("ndof", "nNodes", "nElements", "nodes", "elementNodeIds", "vE", "vthickness", "density",
                 "v", "damp", "iniU", "iniDu", "bdyU", "boundary", "lclNNodes", "lclNodeIds",
                 "lclElmNodeIds", "lclNCommNodes", "lclNSpecialHead", "lclBoundary"):
        assert hasattr(m, attr), attr
    groups = m.colorGroups
    colors = m.elmColors
    for g in groups:                                   # no two elements of a colour share a node
        nodes_in = m.elementNodeIds[g].ravel()
        assert nodes_in.size == np.unique(nodes_in).size
    assert sum(len(g) for g in groups) == m.nElements and colors.max() < 23
    with pytest.raises(RuntimeError):
        mesh_coloring(m.elementNodeIds, m.nNodes, n_colors=2)


# ---------------------------------------------------------------- GMRF --------------------------
def test_dual_graph_edges():
    nodes, el, _ = tube_geometry(6, 3, 0.3, 1.0)
    pairs = element_dual_edges(el)
    # each interior triangle edge is shared by 2 elements: interior edges = 3E/2 - boundary/2
    assert pairs.shape[0] == (3 * len(el) - 2 * 6) // 2
    A, B = dual_laplacian_system(el)
    assert np.allclose(A.sum(axis=1), 1.0)


def test_gmrf_covariance_matches_precision_inverse():
    nodes, el, _ = tube_geometry(4, 3, 0.3, 1.0)           # 16 elements
    A, _ = dual_laplacian_system(el)
    g = sample_element_E(el, 40000, seed=11, cov=0.1)
    cov = np.cov(g.z)
    target = np.linalg.inv(g.tau * A.toarray())
    assert np.abs(cov - target).max() < 0.05 * target.max()
    assert math.isclose(float(np.mean(np.diag(target))), g.var, rel_tol=1e-10)
    cv = g.E.std(axis=1) / g.E.mean(axis=1)
    assert abs(cv.mean() - 0.1) < 0.01 and abs(g.E.mean() / 7e6 - 1.0) < 0.01


# ---------------------------------------------------------------- C ABI -------------------------
def header_functions():
    text = open(os.path.join(ROOT, "include", "cvfes_mi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cvfes_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from cvfes_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libcvfes_mi355x.so not built")
    h = ctypes.CDLL(str(_lib.LIB_PATH))
    names = header_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(h, name), name
    assert set(names) == set(_lib.EXPORTED)
    assert h.cvfes_abi_version() == _lib.ABI_VERSION


def test_ctypes_mirrors_match_compiled_struct_sizes():
    """Every ctypes mirror has the size of the C struct it stands for (a field added on one side
    only would shift every later field of arrays such as the batch descriptors)."""
    import ctypes
    from cvfes_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libcvfes_mi355x.so not built")
    lib = _lib.lib()
    mirrors = [_lib.CvfesSystem, _lib.CvfesLoad, _lib.CvfesBatchMesh, _lib.CvfesGeometry, _lib.CvfesUpdate]
    for which, cls in enumerate(mirrors):
        assert lib.cvfes_struct_size(which) == ctypes.sizeof(cls), cls.__name__
    assert lib.cvfes_struct_size(99) == -1


def test_abi_argument_errors_without_gpu():
    """Argument validation runs on the host: null/invalid inputs return CVFES_ERR_ARG, no launch."""
    import ctypes
    from cvfes_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libcvfes_mi355x.so not built")
    lib = _lib.lib()
    assert lib.cvfes_assemble_bcsr(None, 10, 4, None, None, None, None, None) == -1
    assert b"null pointer" in lib.cvfes_last_error()
    assert lib.cvfes_bcsr_spmv(None, None, 0, None, None) == -1
    sysd = _lib.CvfesSystem(n_nodes=4, n_smp=2, nnzb=4)
    assert lib.cvfes_bcsr_spmv(None, ctypes.byref(sysd), 0, None, None) == -1
    assert b"block structure" in lib.cvfes_last_error()
    assert lib.cvfes_element_matrices(None, 3, None, None, 0.3, 5 / 6, None, None, None, None) == -1
    assert lib.cvfes_element_matrices(None, 0, None, None, 0.3, 5 / 6, None, None, None, None) == 0
    # the updated-geometry step's buffer resources hold a node's 3*S doubles in 31 bits
    dummy = ctypes.c_void_p(16)
    upd_sys = _lib.CvfesSystem(n_nodes=4, n_smp=(1 << 31) // 24, nptr=dummy, nent=dummy, coef=dummy,
                               lm=dummy)
    geo = _lib.CvfesGeometry(nodes=dummy, elems=dummy, nu=0.3, kshear=5 / 6)
    upd = _lib.CvfesUpdate(interval=1, counter=0, has_geometry=0, ug=None)
    state = (ctypes.c_void_p * 3)(16, 32, 48)
    load = _lib.CvfesLoad(p=dummy)
    assert lib.cvfes_step_updated(None, ctypes.byref(upd_sys), ctypes.byref(geo), ctypes.byref(upd),
                                  state, ctypes.byref(load), 0, 1e-6, 0.0, 0.0) == -1
    assert b"too many samples" in lib.cvfes_last_error()


def test_step_kernel_resolution_without_gpu():
    """cvfes_step_kernel (host only) names the kernel cvfes_step would launch: AUTO and the
    variants a structure, sample count or operator does not support fall back as documented."""
    import ctypes
    from cvfes_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libcvfes_mi355x.so not built")
    lib = _lib.lib()
    V = _lib.VARIANTS
    dummy = ctypes.c_void_p(16)                 # never dereferenced: resolution reads the struct only

    def kern(S, variant, method, uvalues=False, lds=False, limits=(7, 2, 6), colors=False,
             full=True):
        sysd = _lib.CvfesSystem(n_nodes=4, n_smp=S, nnzb=4, variant=V[variant],
                                max_row_blocks=limits[0], max_offdiag_contrib=limits[1],
                                max_diag_contrib=limits[2])
        if full:
            sysd.values = dummy
        if uvalues:
            sysd.uvalues = sysd.vblk = dummy
        if lds:
            sysd.nptr = sysd.nent = sysd.cslot = dummy
            sysd.max_node_elems = 6
        if colors:
            sysd.color_ptr = sysd.color_elems = sysd.elems = sysd.cpos = dummy
            sysd.n_elms, sysd.n_colors, sysd.max_color_elems = 10, 3, 4
        return lib.cvfes_step_kernel(ctypes.byref(sysd), _lib.METHODS[method])

    assert kern(128, "auto", "stored") == V["rowload_nt"]
    assert kern(128, "auto", "matfree") == V["generic"]
    assert kern(100, "auto", "stored") == V["generic"]          # row-load needs S % 128 == 0
    assert kern(128, "auto", "stored", limits=(7, 3, 6)) == V["generic"]
    assert kern(128, "rowload_nt", "matfree") == V["rowload"]   # no non-temporal matrix-free form
    assert kern(128, "rowload1_nt", "matfree") == V["rowload1"]
    assert kern(128, "sym_ntl", "stored") == V["rowload_nt"]    # no symmetric storage formed
    assert kern(128, "sym_ntl", "stored", uvalues=True) == V["sym_ntl"]
    assert kern(128, "sym", "matfree", uvalues=True) == V["generic"]
    assert kern(128, "lds2", "matfree") == V["generic"]         # no node-element lists
    assert kern(128, "lds2", "matfree", lds=True) == V["lds2"]
    assert kern(127, "lds2", "matfree", lds=True) == V["lds"]   # odd S: one sample per lane
    assert kern(128, "lds", "stored", lds=True) == V["rowload_nt"]
    assert kern(128, "cached", "matfree", lds=True) == V["cached"]
    assert kern(66, "cached", "matfree", lds=True) == V["generic"]    # 33 of 64 lanes: not uniform
    assert kern(100, "cached", "matfree", lds=True) == V["cached"]    # 50 of 64 lanes: uniform
    assert kern(128, "generic", "stored") == V["generic"]
    assert kern(128, "ebe", "matfree") == V["generic"]          # no element colouring
    assert kern(128, "ebe", "matfree", colors=True) == V["ebe"]
    # symmetric storage only (no full values): the symmetric kernel, or the generic one
    for v in ("auto", "rowload_nt", "rowload1", "sym", "lds"):
        assert kern(128, v, "stored", uvalues=True, full=False) == V["sym"], v
    assert kern(128, "sym_ntl", "stored", uvalues=True, full=False) == V["sym_ntl"]
    # a block across the samples must fit the symmetric kernel's 31-bit buffer resources
    assert kern(1 << 25, "auto", "stored", uvalues=True, full=False) == V["generic"]
    assert kern(1 << 24, "auto", "stored", uvalues=True, full=False) == V["sym"]
    assert kern(128, "generic", "stored", uvalues=True, full=False) == V["generic"]
    assert kern(100, "auto", "stored", uvalues=True, full=False) == V["generic"]
    assert kern(128, "auto", "stored", uvalues=True, full=False, limits=(7, 3, 6)) == V["generic"]
    assert kern(128, "auto", "matfree", uvalues=True, full=False) == V["generic"]
    assert kern(97, "ebe", "matfree", colors=True) == V["ebe"]
    assert kern(128, "ebe", "stored", colors=True) == V["rowload_nt"]
    assert lib.cvfes_step_kernel(None, 0) == -1
    assert kern(128, "auto", "stored") == kern(128, "auto", "bcsr")
    sysd = _lib.CvfesSystem(n_nodes=4, n_smp=8, nnzb=4, variant=99)
    assert lib.cvfes_step_kernel(ctypes.byref(sysd), 0) == -1


@pytest.mark.parametrize("tilt", [False, True])
def test_strip_order_is_a_local_permutation(tilt):
    """The strip visiting order is a permutation, and a node's ring neighbour (same angle, next
    ring) sits about one strip width back instead of a whole ring, also for a tube whose axis is
    not a coordinate axis."""
    from cvfes_amd.sparse import bisection_order, strip_order
    nth, nz = 256, 120
    nodes, el, _ = tube_geometry(nth, nz, 0.5, 6.0 * nz / 400)
    if tilt:
        a, b = 0.7, -0.4
        ra = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
        rb = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
        nodes = nodes @ (rb @ ra).T + np.array([3.0, -1.0, 2.0])
    N = nodes.shape[0]
    for perm in (strip_order(nodes, el, 64), bisection_order(nodes, 64)):
        assert perm.dtype == np.int32 and np.array_equal(np.sort(perm), np.arange(N))
    perm = strip_order(nodes, el, 64)
    pos = np.empty(N, np.int64)
    pos[perm] = np.arange(N)
    ring = np.abs(pos[nth:] - pos[:-nth])           # node (z, t) and (z + 1, t)
    assert np.median(ring) <= 2 * 64 and np.percentile(ring, 90) <= 3 * 64
    same = np.abs(pos[1:] - pos[:-1])[np.arange(N - 1) % nth != nth - 1]   # (z, t) and (z, t + 1)
    assert np.percentile(same, 90) <= 2
    assert strip_order(nodes[:0], el[:0], 64).size == 0


def test_update_interval_selects_linear_mode_from_the_run_length():
    """An interval at least the run's step count never re-forms the stiffness: linear mode."""
    from types import SimpleNamespace as NS
    from cvfes_amd.physicsSolverSDampGPUs import update_interval
    dt = 2.0 ** -18
    run = NS(dt=dt, endtime=200 * dt)                          # 200 steps
    assert update_interval(NS(**vars(run), update_interval=1)) == 1
    assert update_interval(NS(**vars(run), update_interval=199)) == 199
    assert update_interval(NS(**vars(run), update_interval=200)) is None
    assert update_interval(NS(**vars(run), update_interval=1000000000)) is None
    assert update_interval(NS(**vars(run), update_interval=None)) is None
    assert update_interval(NS(dt=dt, update_interval=1 << 62)) is None      # no end time known
    assert update_interval(NS(dt=dt, update_interval=5)) == 5
    cfg = CVConfig(read_cfg(os.path.join(os.path.dirname(__file__), "..", "cvfes_amd", "cfg",
                                         "tube_64x32_s16.cfg")))
    assert update_interval(cfg.solver) is None                 # the shipped cfg: linear mode


# ---------------------------------------------------------------- adjoint sweep schedules -------
class _ToySystem:
    """CPU stand-in for the device system: the state rotation of ``MembraneSystem.step``."""

    def __init__(self, n, s):
        import torch
        self.n_nodes, self.n_smp, self.n_elms = n, s, n
        self.device = torch.device("cpu")
        self.state = None

    def state_shape(self):
        return (self.n_nodes, 3, self.n_smp)

    @property
    def u(self):
        return self.state[0]

    @property
    def up(self):
        return self.state[1]


def _toy_problem(nsteps, C, C2, fuse, obs, u_obs, beta):
    """An ``AdjointProblem`` whose kernels are a small linear recurrence on CPU tensors: a forward
    step, its adjoint, the gradient accumulation and the fused pair (forward + adjoint), each the
    same arithmetic as the separate calls."""
    import torch
    from cvfes_amd.adjoint import AdjointProblem

    def fwd(u, up, n):
        return 1.6 * u - 0.7 * up + 0.1 * torch.roll(u, 1, 0) + 1e-2 * math.cos(0.3 * n)

    def adj(mu1, mu2, out, obs_u, ref, w):
        r = 1.6 * mu1 - 0.7 * mu2 + 0.1 * torch.roll(mu1, -1, 0)
        if obs_u is not None:
            r = r + w * (obs_u - (0.0 if ref is None else ref))
        out.copy_(r)

    class Toy(AdjointProblem):
        def _generic_forward_kernel(self):
            return True

        def _start(self):
            z = torch.linspace(-1e-3, 1e-3, 3 * self.sys.n_nodes * self.sys.n_smp,
                               dtype=torch.float64).reshape(self.sys.state_shape())
            self.sys.state = [z.clone(), 0.5 * z, torch.empty_like(z)]

        def _advance(self, offset, n):
            for k in range(n):
                st = self.sys.state
                st[2].copy_(fwd(st[0], st[1], offset + k))
                self.sys.state = [st[2], st[0], st[1]]

        def _adjoint_call(self, mu1, mu2, out, obs_u=None, obs_ref=None, w=None, project=1):
            adj(mu1, mu2, out, obs_u, obs_ref, w)

        def _fused_call(self, u, up, out, n, mu1, mu2, mu_out, obs_u=None, obs_ref=None, w=None):
            assert out is not u and out is not up and out is not obs_u
            out.copy_(fwd(u, up, n))
            adj(mu1, mu2, mu_out, obs_u, obs_ref, w)

        def _grad_call(self, mu, xc, xp, nu_scale, grad):
            g = (mu * xc).sum(dim=1)
            if xp is not None:
                g = g + 0.25 * (mu * xp).sum(dim=1)
            grad += nu_scale * g

    sysm = _ToySystem(7, 4)
    return Toy(sysm, None, nsteps, 0.5, beta=beta, obs_steps=obs, u_obs=u_obs,
               checkpoint_every=C, inner_every=C2, fuse=fuse)


@pytest.mark.parametrize("beta", [0.0, 0.3])
def test_reverse_sweep_schedules_agree_on_a_toy_recurrence(beta):
    """Every checkpoint schedule (one or two levels, the recomputation fused with the adjoint
    steps or not, intervals that do not divide the run) replays the same arithmetic on the same
    states, so the gradients agree bit for bit; a buffer handed to the next segment's
    recomputation while its state is still needed would change them."""
    import torch
    nsteps, obs = 75, [8, 30, 31, 64, 75]
    u_obs = torch.linspace(-1, 1, len(obs) * 7 * 3 * 4, dtype=torch.float64).reshape(len(obs), 7, 3, 4)
    ref = None
    # fused steps: every adjoint step of a segment that has an earlier segment to recompute,
    # up to that segment's length - 1 (e.g. C = 11: 9 in [66, 75) + 5 x 10 = 59; C = 30 with
    # C2 = 7: 7 in [60, 75) + 20 + 20 = 47)
    for C, C2, fuse, n_fused in ((1, None, False, 0), (11, None, False, 0), (11, None, True, 59),
                                 (75, None, True, 0), (1000, None, True, 0), (30, 7, False, 0),
                                 (30, 7, True, 47), (40, 9, True, 52), (30, 1, True, 0),
                                 (2, None, True, 37), (3, 2, True, 25)):
        prob = _toy_problem(nsteps, C, C2, fuse, obs, u_obs, beta)
        loss = prob.forward()
        g = prob.backward(torch.ones(4, dtype=torch.float64))
        assert prob.stats["fused_steps"] == n_fused, (C, C2, fuse, prob.stats["fused_steps"])
        if ref is None:
            ref = (loss, g)
            assert float(g.abs().max()) > 0
        else:
            assert torch.equal(loss, ref[0]) and torch.equal(g, ref[1]), (C, C2, fuse)


def test_bench_refuses_more_than_one_gpu():
    """The measurement is one GPU's share; a multi-GPU request exits with the reason (no GPU
    touched)."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_cpu_baseline_is_the_reference_step():
    """The CPU baseline leg of bench.py (C restatement in the reference layout, timed on the host)
    advances the same recurrence as the NumPy oracle, on a small damped tube."""
    import sys
    import types
    import torch
    sys.path.insert(0, ROOT)
    import bench
    from oracle import solid_oracle as so

    nodes, el, bdy = tube_geometry(16, 6, 0.5, 2.0)
    N, S, nu, dt, beta, steps = nodes.shape[0], 4, 0.45, 2e-6, 1e-6, 5
    rng = np.random.default_rng(4)
    elmE = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
    h = np.full((N, S), 0.07)
    indptr, indices = so.sparse_info_loops(el, N)
    _, gth = so.gauss_properties(np.full((N, S), 7e6), h, el)
    coefK = so.coef_k_element_E(gth, elmE)
    _, K, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(nu), coefK, gth, 1.0,
                                         indptr, indices)
    pres = 1.33e4 * (1.0 + 0.1 * np.sin(np.arange(steps + 3)))
    sysm = types.SimpleNamespace(indptr_h=indptr, indices_h=indices, n_nodes=N,
                                 f1=torch.as_tensor(f1), lumped_mass_dofs=lambda: LM)
    case = types.SimpleNamespace(n_smp=S, name="tube-16x6", nu=nu,
                                 pressures=lambda n0, n, dt_: pres[n0:n0 + n])
    res, u = bench.cpu_baseline(case, sysm, nodes, el, bdy, coefK, dt, beta, steps, 2)
    bdy_dofs = (3 * bdy[:, None] + np.arange(3)).ravel()
    z = np.zeros(3 * N)
    uo, _, _ = so.run(indptr, indices, K, LM, f1, pres[:steps], z, z, dt, beta=beta,
                      bdy_dofs=bdy_dofs)
    assert np.max(np.abs(uo)) > 0
    assert np.max(np.abs(u - uo)) / np.max(np.abs(uo)) < 1e-12
    assert res["kind"] == "port" and res["cores"] == 2 and res["value"] > 0
    assert res["one_thread"]["cores"] == 1 and res["one_thread"]["value"] > 0
    # two warm-up steps, the rest timed; three more steps of the recurrence on one thread
    assert len(res["step_s"]) == steps - 2 and "3 timed steps after 2 warm-up" in res["sample"]
    assert res["one_thread"]["steps"] == 3


def test_solver_reports_a_non_finite_displacement():
    """``Solve`` checks the displacement every ``finite_check_every`` steps and raises
    ``FloatingPointError`` once it holds a NaN (a stand-in system on the CPU, NaN from step 7)."""
    import types
    import torch
    from cvfes_amd.physicsSolverSDampGPUs import GPUSolidSolver

    pressures = []

    class Sys:
        u = torch.zeros(4, 3, 2, dtype=torch.float64)

        def step(self, tables, dt, **kw):
            pressures.append(float(tables.p[0]))
            if len(pressures) == 7:
                self.u[1, 2, 0] = float("nan")

    s = object.__new__(GPUSolidSolver)
    s.system, s.traction, s.updateInterval = Sys(), None, None
    s.beta = s.damp = 0.0
    s.method = "stored"
    s._solve_tables = types.SimpleNamespace(p=torch.zeros(1, dtype=torch.float64))
    s.t, s.pressure, s.finiteCheckEvery, s._nSolve = 0.0, 5.0, 4, 0
    for _ in range(7):
        s.Solve(0.0, 1e-6)                       # checked after step 4: finite
    with pytest.raises(FloatingPointError, match="not finite"):
        s.Solve(0.0, 1e-6)                       # checked after step 8
    assert pressures == [5.0] * 8
    s.system.u.zero_()
    s.CheckFinite(1e-6)
    cfg = CVConfig(read_cfg(os.path.join(CFG_DIR, "tube_64x32_s16.cfg")))
    assert cfg.solver.finite_check_every == 100


def test_critical_dt_power_iteration_and_bound():
    """``MembraneSystem.critical_dt`` (power iteration on M^-1 K) on a stand-in whose product is
    the oracle's: its eigenvalue approaches ARPACK's largest from below, and the returned step is
    the stability limit of the damped central-difference recurrence, whose characteristic roots
    leave the unit circle just above it."""
    import types
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    import torch
    from cvfes_amd.system import MembraneSystem
    from oracle import solid_oracle as so

    nodes, el, bdy = tube_geometry(32, 16, 0.3, 1.5)
    N, S = nodes.shape[0], 4
    rng = np.random.default_rng(7)
    E = 7e6 * np.exp(0.1 * rng.standard_normal((el.shape[0], S)))
    h = np.full((N, S), 0.07)
    indptr, indices = so.sparse_info_loops(el, N)
    _, gth = so.gauss_properties(np.full((N, S), 7e6), h, el)
    coefK = so.coef_k_element_E(gth, E)
    _, K, f1, LM = so.assemble_reference(nodes, el, 1.0, so.d_matrix(0.5), coefK, gth, 1.0,
                                         indptr, indices)
    # largest eigenvalue of the pencil (K, M) over the samples, by ARPACK on M^-1/2 K M^-1/2
    row = np.repeat(np.arange(N), indptr[1:] - indptr[:-1])
    r = np.broadcast_to(3 * row[:, None, None] + np.arange(3)[None, :, None], (row.size, 3, 3))
    c = np.broadcast_to(3 * indices[:, None, None] + np.arange(3)[None, None, :], (row.size, 3, 3))
    lam_exact = 0.0
    for s in range(S):
        A = sp.csr_matrix((K[:, s].ravel(), (r.ravel(), c.ravel())), shape=(3 * N, 3 * N))
        d = sp.diags(1.0 / np.sqrt(LM[:, s]))
        lam_exact = max(lam_exact, sla.eigsh(d @ A @ d, k=1, which="LA",
                                             return_eigenvectors=False)[0])
    stand_in = types.SimpleNamespace(
        device=torch.device("cpu"), lm=torch.as_tensor(LM[0::3].copy()),
        state_shape=lambda: (N, 3, S),
        spmv=lambda x, method: torch.as_tensor(so.multiply_by_vector(
            indptr, indices, K, x.reshape(3 * N, S).numpy()).reshape(N, 3, S)))
    m_min = float(LM[LM > 0].min())
    for beta, damp in ((0.0, 0.0), (2e-6, 0.0), (0.0, 40.0), (2e-6, 40.0)):
        dt, lam = MembraneSystem.critical_dt(stand_in, beta=beta, iters=200, damp=damp)
        assert 0.995 * lam_exact <= lam <= lam_exact * (1 + 1e-9)
        g = damp / m_min
        # u+ = 2u - up - dt^2 lam (u + (beta/dt)(u - up)) - dt g (u - up): roots of
        # z^2 - (2 - dt^2 lam - c) z + (1 - c),  c = dt (beta lam + g)
        for f, stable in ((0.99, True), (1.01, False)):
            h_ = f * dt
            c = h_ * (beta * lam + g)
            z = np.roots([1.0, -(2.0 - h_ * h_ * lam - c), 1.0 - c])
            assert (np.max(np.abs(z)) <= 1.0 + 1e-12) == stable, (beta, damp, f, np.abs(z))
        if damp:
            assert dt < MembraneSystem.critical_dt(stand_in, beta=beta, iters=200)[0]


def test_load_tables_step_offsets_and_validation():
    """``LoadTables.struct(k)`` points every per-step table at entry k (the fields stay whole),
    and inconsistent tables are refused (CPU tensors; no GPU call)."""
    import torch
    from cvfes_amd.system import LoadTables
    cpu = torch.device("cpu")
    p = np.linspace(1.0, 2.0, 6)
    g = np.arange(12.0).reshape(4, 3)
    t = LoadTables(p, cpu, ga=g, qa=np.arange(6.0), gb=2 * g, qb=np.ones(6))
    assert t.n == 6 and t.ga.shape == (4, 3) and t.qb.dtype == torch.float64
    for k in (0, 3, 6):
        s = t.struct(k)
        assert s.p == t.p.data_ptr() + 8 * k
        assert s.qa == t.qa.data_ptr() + 8 * k and s.qb == t.qb.data_ptr() + 8 * k
        assert s.ga == t.ga.data_ptr() and s.gb == t.gb.data_ptr()
    plain = LoadTables(3.5, cpu).struct(0)
    assert plain.ga is None and plain.qa is None and plain.gb is None and plain.qb is None
    with pytest.raises(ValueError):
        t.struct(7)
    with pytest.raises(ValueError):
        t.struct(-1)
    with pytest.raises(ValueError):
        LoadTables(p, cpu, ga=g, qa=np.arange(5.0))            # a table one entry short
    with pytest.raises(ValueError):
        LoadTables(p, cpu, ga=g)                                # a field without its table
    # fields are checked against the mesh's node count (MembraneSystem.loads passes it)
    assert LoadTables(p, cpu, ga=g, qa=np.arange(6.0), n_nodes=4).ga.shape == (4, 3)
    with pytest.raises(ValueError, match="expected"):
        LoadTables(p, cpu, ga=g, qa=np.arange(6.0), n_nodes=5)  # a field of another mesh
    with pytest.raises(ValueError, match="expected"):
        LoadTables(p, cpu, gb=np.arange(15.0), qb=np.ones(6), n_nodes=4)


def test_single_rank_contract_helpers():
    """The reference helpers the drop-in keeps for one rank: boundary values, the union of the
    local rows into the global ``[ndof, nSmp]`` array, no-op synchronisation (multi-rank
    decomposition is refused), ``Steady`` false, and the one-rank communicator."""
    from cvfes_amd.physicsSolver import SerialComm
    from cvfes_amd.physicsSolverSDampGPUs import GPUSolidSolver

    s = object.__new__(GPUSolidSolver)
    s.size, s.ndof, s.nSmp = 1, 12, 2
    s.dofs = np.array([3, 4, 5, 0, 1, 2, 9, 10, 11, 6, 7, 8])        # local rows -> global dofs
    s.bdyDofs = np.array([0, 1, 2])
    s.mesh = type("M", (), {"bdyU": 0.25})()
    q = np.arange(24.0).reshape(12, 2)
    s.ApplyBoundaryCondition(q)
    assert np.all(q[:3] == 0.25) and q[3, 0] == 6.0
    u = s.UnionDisplacement(q)
    assert u.shape == (12, 2) and np.array_equal(u[s.dofs], q)
    assert s.SyncCommNodes(q) is None and s.Steady() is False
    s.size = 2
    with pytest.raises(NotImplementedError):
        s.SyncCommNodes(q)
    with pytest.raises(NotImplementedError):
        s.UnionDisplacement(q)
    comm = SerialComm()
    assert (comm.Get_size(), comm.Get_rank(), comm.Barrier(), comm.bcast("x")) == (1, 0, None, "x")


def test_batch_refuses_steps_past_load_tables():
    """The batched kernel reads load-table entry (absolute step) with no bound of its own, so
    ``BatchedEnsemble`` checks every run, eager block and tuning sweep against the shortest
    mesh's tables before it launches anything (stand-in library and CPU buffers; no GPU call)."""
    import types
    import torch
    from cvfes_amd import batch as B
    from cvfes_amd._lib import CvfesSystem
    from cvfes_amd.system import LoadTables

    launched = []

    class Lib:
        def cvfes_batch_plan_bytes(self, n):
            return 64 * n

        def cvfes_batch_plan(self, stream, arr, n, dev, n_tasks, kinds):
            n_tasks._obj.value, kinds._obj.value = 8 * n, 1
            return 0

        def cvfes_batch_step(self, *args):
            launched.append(args[4])
            return 0

        def cvfes_graph_destroy(self, g):
            return 0

    class CpuEnsemble(B.BatchedEnsemble):
        stream = None

    cpu = torch.device("cpu")

    def mesh(n_nodes):
        bufs = [torch.zeros(n_nodes, 3, 2, dtype=torch.float64) for _ in range(3)]
        return types.SimpleNamespace(n_nodes=n_nodes, state=bufs, _make_struct=CvfesSystem)

    tables = [LoadTables(np.ones(9), cpu), LoadTables(np.ones(6), cpu)]
    ens = object.__new__(CpuEnsemble)
    ens.systems, ens.loads = [mesh(5), mesh(7)], tables
    ens.dts, ens.methods, ens.betas, ens.damps = [1e-6] * 2, ["matfree"] * 2, [0.0] * 2, [0.0] * 2
    ens.device, ens.lib, ens._graphs = cpu, Lib(), {}
    ens.vw_hint, ens.cap, ens.waves_per_simd, ens.n_waves, ens.groups_first = 0, False, 0, 0, False
    ens._plan()
    ens.phase, ens.step_index = 0, 0
    assert ens.max_steps == 6 and ens.n_tasks == 16
    with pytest.raises(ValueError, match="exceed the load tables"):
        ens.run(7, graph=False)
    assert launched == [] and ens.step_index == 0
    ens.run(4, graph=False)
    assert launched == [0, 1, 2, 0] and ens.step_index == 4
    with pytest.raises(ValueError, match="exceed the load tables"):
        ens._eager(3)
    with pytest.raises(ValueError, match="exceed the load tables"):
        ens.run(99, graph=True)
    with pytest.raises(ValueError, match="exceed the load tables"):
        ens.autotune(steps=2, lane_widths=(0,), caps=(False,), occupancies=(0,),
                     candidates=[64])                       # 1 candidate x (1 + 2) steps > 2 left
    assert launched == [0, 1, 2, 0] and ens.step_index == 4
    ens.run(2, graph=False)
    assert ens.step_index == 6

