"""cvfes_initial_state_static (static device scratch, no allocation) against cvfes_initial_state.

Both entry points run the same three kernels, so u_0 and u_{-1} must agree bit for bit.  Two calls in
a row on one GPU share the static scratch: the second, on a smaller system, must not see the first
call's contents.  A system larger than the scratch is refused with an argument error.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(ntheta, nz, n_smp, seed):
    from cvfes_amd.gmrf import sample_element_E
    from cvfes_amd.mesh import tube_geometry
    from cvfes_amd.system import MembraneSystem

    nodes, el, bdy = tube_geometry(ntheta, nz, 0.3, 3.0)
    N = nodes.shape[0]
    E = sample_element_E(el, n_smp, seed=seed).E
    sysm = MembraneSystem(nodes, el, bdy, n_smp, vthickness=np.full((N, n_smp), 0.07), elmE=E,
                          density=1.0, nu=0.5)
    rng = np.random.default_rng(seed)
    u0 = 1e-4 * rng.standard_normal(3 * N)
    du0 = 1e-2 * rng.standard_normal(3 * N)
    return sysm, u0, du0


def _both(sysm, u0, du0, method, beta, damp):
    import torch
    from cvfes_amd._lib import check, ptr
    from cvfes_amd.system import _dev

    f64 = torch.float64
    dt = 4e-6
    load = sysm.loads(np.array([1.33e4, 1.2e4]))
    ls = load.struct(0)
    shp = sysm.state_shape()
    u0_d = _dev(u0, f64, sysm.device)
    du0_d = _dev(du0, f64, sysm.device)
    u_a, up_a, work = (torch.empty(shp, dtype=f64, device=sysm.device) for _ in range(3))
    u_b = torch.full(shp, np.nan, dtype=f64, device=sysm.device)
    up_b = torch.full(shp, np.nan, dtype=f64, device=sysm.device)
    m = sysm.method_id(method)
    lib = sysm.lib
    check(lib.cvfes_initial_state(sysm.stream, ctypes.byref(sysm._sys), m, ptr(u0_d), ptr(du0_d),
                                  ctypes.byref(ls), dt, beta, damp, ptr(u_a), ptr(up_a), ptr(work)))
    check(lib.cvfes_initial_state_static(sysm.stream, ctypes.byref(sysm._sys), m, ptr(u0_d),
                                         ptr(du0_d), ctypes.byref(ls), dt, beta, damp, ptr(u_b),
                                         ptr(up_b)))
    torch.cuda.synchronize()
    return (u_a.cpu().numpy(), up_a.cpu().numpy()), (u_b.cpu().numpy(), up_b.cpu().numpy())


@pytest.mark.parametrize("method", ["stored", "matfree"])
def test_static_initial_state_matches_workspace_form(require_gpu, method):
    big = _system(64, 32, 16, seed=11)
    small = _system(24, 12, 8, seed=12)
    for sysm, u0, du0 in (big, small):
        (ua, upa), (ub, upb) = _both(sysm, u0, du0, method, beta=1e-4 if method == "stored" else 0.0,
                                     damp=2.0)
        assert np.all(np.isfinite(upb))
        assert np.array_equal(ua, ub)
        assert np.array_equal(upa, upb)
        # the start is not trivial: u_{-1} differs from u_0 through du0 and the acceleration
        assert np.max(np.abs(upa - ua)) > 0


def test_static_initial_state_refuses_oversized_system(require_gpu):
    import torch
    from cvfes_amd._lib import STATIC_SCRATCH_DOUBLES, CvfesError, check, ptr

    sysm, u0, du0 = _system(8, 4, 4, seed=13)
    s = type(sysm._sys).from_buffer_copy(sysm._sys)
    # only the size fields change: the entry point must refuse before any launch
    s.n_nodes = STATIC_SCRATCH_DOUBLES // (3 * s.n_smp) + 1
    t = torch.zeros(3, dtype=torch.float64, device=sysm.device)
    ls = sysm.loads(np.array([1.0])).struct(0)
    with pytest.raises(CvfesError, match="static scratch"):
        check(sysm.lib.cvfes_initial_state_static(sysm.stream, ctypes.byref(s), 0, ptr(t), ptr(t),
                                                  ctypes.byref(ls), 1e-6, 0.0, 0.0, ptr(t), ptr(t)))
