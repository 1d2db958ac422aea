// Kernels on the per-sample geometry X + u_s: element stress recovery and the matrix-free step of
// the updated-geometry mode.
//
// Stress (optimizedSolidAssemble.pyx:1116-1223; physicsSolverGPUs.cl:443-589): per element and
// sample, the local strain B T u_e of the constant-strain triangle on the geometry X + ug_s, the
// local stress E D B T u_e (D the 5x5 membrane + transverse-shear matrix / (1 - nu^2)) as the
// symmetric tensor [[s0, s2, s3], [s2, s1, s4], [s3, s4, 0]], rotated to global axes T^T S T;
// output (xx, yy, xy, xz, yz) as stress[e][c][s].
//
// Updated-geometry step (physicsSolver.py:294-324 with optimizedSolidAssemble.pyx:710-819, the
// K(X+u) product of physicsSolverSDampGPUs.cl:232-391): the stiffness and the unit-traction load are
// those of the geometry X + ug_s, where ug is the displacement at the last re-formation.  Nothing
// is stored: each thread owns one (row node, sample), walks the elements incident to the node,
// forms each element's frame on the moved geometry and applies its three blocks in the local
// frame, T^T (k_ab (T x_b)), so the matrix is never formed.  The update of u is the fused
// central-difference update of the linear step.
#include "geometry.h"
#include "rowops.h"

namespace cvfes {
namespace {

struct GeomArgs {
    const double *nodes;     // [N][3] reference coordinates
    const int32_t *elems;    // [E][3]
    const double *ug;        // [N][3][S] geometry displacement (null: reference geometry)
    const int32_t *nptr;     // [N+1]
    const int32_t *nent;     // [3E]
    double nu, g, gk, inv_1mnu2;
};

__device__ __forceinline__ void moved_points(const GeomArgs &g, const int32_t (&ids)[3], int64_t S,
                                             int64_t s, double (&X)[3][3]) {
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            X[b][k] = g.nodes[3 * static_cast<int64_t>(ids[b]) + k];
            if (g.ug) X[b][k] += g.ug[(static_cast<int64_t>(ids[b]) * 3 + k) * S + s];
        }
}

// FAST: reciprocal-normalised frames and one reciprocal for the element scale (results within a few
// ulp of the divided forms).  GEO_U: the geometry displacement is the current u (interval 1), so
// each node's u is loaded once for the geometry and the product.  VW samples per lane.
template <bool UNIFORM, bool BETA, bool DAMP, bool FAST, bool GEO_U, int VW>
__global__ __launch_bounds__(kBlock) void k_step_updated(StepArgs a, GeomArgs g, int64_t step) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    const int64_t S = a.k.n_smp;
    const uint32_t lo = static_cast<uint32_t>(s) * 8u;   // the lane's byte offset in every row
    const uint32_t sb = static_cast<uint32_t>(S) * 8u;   // bytes of one entry across the samples
    double acc[VW][3] = {};
    double fl[VW][3] = {};    // unit-traction load of the row node on the moved geometry
    for (int32_t q = g.nptr[row]; q < g.nptr[row + 1]; ++q) {
        const int32_t ent = g.nent[q];
        const int32_t e = ent / 3;
        const int la = ent - 3 * e;
        const int32_t ids[3] = {g.elems[3 * e], g.elems[3 * e + 1], g.elems[3 * e + 2]};
        double ref[3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int k = 0; k < 3; ++k) ref[b][k] = g.nodes[3 * static_cast<int64_t>(ids[b]) + k];
        // per-sample state of the element's nodes: u (and up), the geometry displacement
        // (wave-uniform rows: buffer loads from scalar resources per node, one per-lane offset)
        double xu[VW][3][3], xg[VW][3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int64_t off = (static_cast<int64_t>(ids[b]) * 3 + k) * S + s;
                Vec<VW> uu, pp, gg;
                if constexpr (UNIFORM) {
                    const int64_t node = static_cast<int64_t>(ids[b]) * 3 * S;
                    uu = to_vec<VW>(ldb<VW, 0>(buffer_rsrc(a.u + node, 3 * sb), lo, k * sb));
                    if constexpr (BETA)
                        pp = to_vec<VW>(ldb<VW, 0>(buffer_rsrc(a.up + node, 3 * sb), lo, k * sb));
                } else {
                    uu = load_vec<VW>(a.u + off);
                    if constexpr (BETA) pp = load_vec<VW>(a.up + off);
                }
                if constexpr (!GEO_U) {
                    if (g.ug) gg = load_vec<VW>(g.ug + off);
                }
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    xu[v][b][k] = BETA ? uu.v[v] + a.bdt * (uu.v[v] - pp.v[v]) : uu.v[v];
                    if constexpr (GEO_U) xg[v][b][k] = uu.v[v];
                    else xg[v][b][k] = g.ug ? gg.v[v] : 0.0;
                }
            }
        Vec<VW> cf;
        if constexpr (UNIFORM) cf = to_vec<VW>(ldb<VW, 0>(buffer_rsrc(a.k.coef + static_cast<int64_t>(e) * S, sb), lo, 0));
        else cf = load_vec<VW>(a.k.coef + static_cast<int64_t>(e) * S + s);
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            double X[3][3];
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int k = 0; k < 3; ++k) X[b][k] = ref[b][k] + xg[v][b][k];
            const Frame f = frame_from_points<FAST>(X);
            const double bb[3] = {f.ly[1] - f.ly[2], f.ly[2] - f.ly[0], f.ly[0] - f.ly[1]};
            const double cc[3] = {f.lx[2] - f.lx[1], f.lx[0] - f.lx[2], f.lx[1] - f.lx[0]};
            const double scale = FAST ? cf.v[v] * g.inv_1mnu2 * (1.0 / (4.0 * f.area))
                                      : cf.v[v] * g.inv_1mnu2 / (4.0 * f.area);
            double L0 = 0.0, L1 = 0.0, L2 = 0.0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const double *x = xu[v][b];
                const double x0 = f.t[0][0] * x[0] + f.t[0][1] * x[1] + f.t[0][2] * x[2];
                const double x1 = f.t[1][0] * x[0] + f.t[1][1] * x[1] + f.t[1][2] * x[2];
                const double x2 = f.t[2][0] * x[0] + f.t[2][1] * x[1] + f.t[2][2] * x[2];
                const double ba = bb[la], ca = cc[la];
                L0 += (ba * bb[b] + g.g * ca * cc[b]) * x0 + (g.nu * ba * cc[b] + g.g * ca * bb[b]) * x1;
                L1 += (g.nu * ca * bb[b] + g.g * ba * cc[b]) * x0 + (ca * cc[b] + g.g * ba * bb[b]) * x1;
                L2 += g.gk * (ba * bb[b] + ca * cc[b]) * x2;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                acc[v][k] += scale * (f.t[0][k] * L0 + f.t[1][k] * L1 + f.t[2][k] * L2);
                fl[v][k] += (1.0 / 3.0) * f.area * f.t[2][k];
            }
        }
    }
    const double pn = a.load.p[step];
    const double qa = a.load.ga ? a.load.qa[step] : 0.0;
    const double qb = a.load.gb ? a.load.qb[step] : 0.0;
    const bool fixed = a.fixed != nullptr && a.fixed[row] != 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * S + s;
        const Vec<VW> uu = load_vec<VW>(a.u + off);
        const Vec<VW> pp = load_vec<VW>(a.up + off);
        const int64_t d = 3 * static_cast<int64_t>(row) + i;
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            const double m = a.lm_sstride ? a.lm[static_cast<int64_t>(row) * S + s + v] : a.lm[row];
            double force = pn * fl[v][i];
            if (a.load.ga) force += qa * a.load.ga[d];
            if (a.load.gb) force += qb * a.load.gb[d];
            if constexpr (DAMP) force -= a.ddt * (uu.v[v] - pp.v[v]);
            const double rhs = a.dt2 * (force - acc[v][i]) + m * (2.0 * uu.v[v] - pp.v[v]);
            const double un = m != 0.0 ? rhs / m : 0.0;
            o.v[v] = fixed ? a.bdy_u : un;
        }
        if (active) store_vec<VW>(a.out + off, o);
    }
}

__global__ __launch_bounds__(kBlock) void k_element_stress(int64_t n_elms, int64_t S, GeomArgs g,
                                                           const double *__restrict__ u,
                                                           const double *__restrict__ ave_E,
                                                           double *__restrict__ stress) {
    const int64_t t = xcd_block() * kBlock + threadIdx.x;
    if (t >= n_elms * S) return;
    const int64_t e = t / S;
    const int64_t s = t - e * S;
    const int32_t ids[3] = {g.elems[3 * e], g.elems[3 * e + 1], g.elems[3 * e + 2]};
    double X[3][3];
    moved_points(g, ids, S, s, X);
    const Frame f = frame_from_points(X);
    const double bb[3] = {f.ly[1] - f.ly[2], f.ly[2] - f.ly[0], f.ly[0] - f.ly[1]};
    const double cc[3] = {f.lx[2] - f.lx[1], f.lx[0] - f.lx[2], f.lx[1] - f.lx[0]};
    double eps[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        double x[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) x[k] = u[(static_cast<int64_t>(ids[b]) * 3 + k) * S + s];
        const double l0 = f.t[0][0] * x[0] + f.t[0][1] * x[1] + f.t[0][2] * x[2];
        const double l1 = f.t[1][0] * x[0] + f.t[1][1] * x[1] + f.t[1][2] * x[2];
        const double l2 = f.t[2][0] * x[0] + f.t[2][1] * x[1] + f.t[2][2] * x[2];
        eps[0] += bb[b] * l0;
        eps[1] += cc[b] * l1;
        eps[2] += cc[b] * l0 + bb[b] * l1;
        eps[3] += bb[b] * l2;
        eps[4] += cc[b] * l2;
    }
    const double c = ave_E[t] * g.inv_1mnu2 / (2.0 * f.area);
    double S3[3][3];
    S3[0][0] = c * (eps[0] + g.nu * eps[1]);
    S3[1][1] = c * (g.nu * eps[0] + eps[1]);
    S3[0][1] = S3[1][0] = c * g.g * eps[2];
    S3[0][2] = S3[2][0] = c * g.gk * eps[3];
    S3[1][2] = S3[2][1] = c * g.gk * eps[4];
    S3[2][2] = 0.0;
    // G = T^T S T
    double TS[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l)
            TS[i][l] = f.t[0][i] * S3[0][l] + f.t[1][i] * S3[1][l] + f.t[2][i] * S3[2][l];
    auto G = [&](int i, int j) { return TS[i][0] * f.t[0][j] + TS[i][1] * f.t[1][j] + TS[i][2] * f.t[2][j]; };
    double *out = stress + e * 5 * S + s;
    out[0] = G(0, 0);
    out[S] = G(1, 1);
    out[2 * S] = G(0, 1);
    out[3 * S] = G(0, 2);
    out[4 * S] = G(1, 2);
}

GeomArgs geom_args(const cvfes_geometry_t *geo, const cvfes_system_t *sys, const double *ug) {
    GeomArgs g;
    g.nodes = geo->nodes;
    g.elems = geo->elems;
    g.ug = ug;
    g.nptr = sys ? sys->nptr : nullptr;
    g.nent = sys ? sys->nent : nullptr;
    g.nu = geo->nu;
    g.g = 0.5 * (1.0 - geo->nu);
    g.gk = 0.5 * geo->kshear * (1.0 - geo->nu);
    g.inv_1mnu2 = 1.0 / (1.0 - geo->nu * geo->nu);
    return g;
}

template <bool UNIFORM, bool FAST, bool GEO_U, int VW>
void launch_updated_inst(hipStream_t st, const StepArgs &a, const GeomArgs &g, int64_t step, bool beta,
                         bool damp) {
    const int64_t work = a.k.n_nodes * lanes_per_row<UNIFORM, VW>(a.k.n_smp);
    const dim3 grid(grid_for(work)), block(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_step_updated<UNIFORM, true, true, FAST, GEO_U, VW>), grid, block, 0, st, a, g, step);
        else hipLaunchKernelGGL((k_step_updated<UNIFORM, true, false, FAST, GEO_U, VW>), grid, block, 0, st, a, g, step);
    } else {
        if (damp) hipLaunchKernelGGL((k_step_updated<UNIFORM, false, true, FAST, GEO_U, VW>), grid, block, 0, st, a, g, step);
        else hipLaunchKernelGGL((k_step_updated<UNIFORM, false, false, FAST, GEO_U, VW>), grid, block, 0, st, a, g, step);
    }
}

template <bool FAST, bool GEO_U, int VW>
void launch_updated_u(hipStream_t st, const StepArgs &a, const GeomArgs &g, int64_t step, bool beta,
                      bool damp) {
    if (use_uniform<VW>(a.k.n_smp)) launch_updated_inst<true, FAST, GEO_U, VW>(st, a, g, step, beta, damp);
    else launch_updated_inst<false, FAST, GEO_U, VW>(st, a, g, step, beta, damp);
}

// Updated-geometry kernel forms (sys->variant): GENERIC the divided frames with one sample per
// lane; ROWLOAD1 reciprocal frames, one sample per lane; ROWLOAD (and AUTO) reciprocal frames, two
// samples per lane for even sample counts.  geo_u: the geometry displacement is the current u.
void launch_updated(hipStream_t st, const StepArgs &a, const GeomArgs &g, int64_t step, bool beta,
                    bool damp, int64_t variant, bool geo_u) {
    const bool two = (variant == CVFES_VARIANT_AUTO || variant == CVFES_VARIANT_ROWLOAD) &&
                     a.k.n_smp % 2 == 0;
    if (variant == CVFES_VARIANT_GENERIC) {
        launch_updated_u<false, false, 1>(st, a, g, step, beta, damp);
    } else if (two) {
        if (geo_u) launch_updated_u<true, true, 2>(st, a, g, step, beta, damp);
        else launch_updated_u<true, false, 2>(st, a, g, step, beta, damp);
    } else {
        if (geo_u) launch_updated_u<true, true, 1>(st, a, g, step, beta, damp);
        else launch_updated_u<true, false, 1>(st, a, g, step, beta, damp);
    }
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int cvfes_element_stress(void *stream, int64_t n_elms, int64_t n_smp,
                                    const cvfes_geometry_t *geo, const double *u, const double *ug,
                                    const double *ave_E, double *stress) {
    CVFES_REQUIRE(n_elms >= 0 && n_smp >= 1, "cvfes_element_stress: bad sizes");
    CVFES_REQUIRE(geo && geo->nodes && geo->elems && u && ave_E && stress,
                  "cvfes_element_stress: null argument");
    CVFES_REQUIRE(geo->nu > -1.0 && geo->nu < 1.0, "cvfes_element_stress: nu out of range");
    if (n_elms == 0) return CVFES_OK;
    const GeomArgs g = geom_args(geo, nullptr, ug);
    hipLaunchKernelGGL(k_element_stress, dim3(grid_for(n_elms * n_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), n_elms, n_smp, g, u, ave_E, stress);
    return check_launch("k_element_stress");
}

extern "C" int cvfes_step_updated(void *stream, const cvfes_system_t *sys, const cvfes_geometry_t *geo,
                                  cvfes_update_t *upd, double **state, const cvfes_load_t *load,
                                  int64_t nsteps, double dt, double beta, double damp) {
    CVFES_REQUIRE(sys && geo && upd && state && load, "cvfes_step_updated: null argument");
    CVFES_REQUIRE(sys->n_nodes >= 1 && sys->n_smp >= 1 && nsteps >= 0 && dt > 0.0,
                  "cvfes_step_updated: bad sizes");
    CVFES_REQUIRE(sys->n_nodes * 3 < (int64_t(1) << 31), "cvfes_step_updated: too many nodes");
    // buffer resources of 3*S doubles (a node's state) must fit 31 bits
    CVFES_REQUIRE(sys->n_smp < (int64_t(1) << 31) / 24, "cvfes_step_updated: too many samples");
    CVFES_REQUIRE(sys->nptr && sys->nent && sys->coef && sys->lm && geo->nodes && geo->elems,
                  "cvfes_step_updated: the system needs nptr, nent, coef, lm and the geometry");
    CVFES_REQUIRE(geo->nu > -1.0 && geo->nu < 1.0, "cvfes_step_updated: nu out of range");
    CVFES_REQUIRE(upd->interval >= 1, "cvfes_step_updated: interval must be >= 1");
    CVFES_REQUIRE(upd->interval == 1 || upd->ug, "cvfes_step_updated: interval > 1 needs the ug buffer");
    CVFES_REQUIRE(state[0] && state[1] && state[2], "cvfes_step_updated: null state buffer");
    CVFES_REQUIRE(load->p && (!load->ga || load->qa) && (!load->gb || load->qb),
                  "cvfes_step_updated: load tables");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t S = sys->n_smp;
    double *cur = state[0], *prev = state[1], *nxt = state[2];
    for (int64_t n = 0; n < nsteps; ++n) {
        if (upd->counter >= upd->interval) {
            if (upd->interval > 1) {
                const hipError_t err = hipMemcpyAsync(upd->ug, cur, sizeof(double) * 3 * sys->n_nodes * S,
                                                      hipMemcpyDeviceToDevice, st);
                if (err != hipSuccess)
                    return fail(CVFES_ERR_HIP, "cvfes_step_updated: %s", hipGetErrorString(err));
            }
            upd->has_geometry = 1;
            upd->counter = 1;
        } else {
            upd->counter += 1;
        }
        const double *geom = upd->has_geometry ? (upd->interval == 1 ? cur : upd->ug) : nullptr;
        StepArgs a;
        a.k = RowArgs{sys->n_nodes, S, sys->indptr, sys->indices, sys->values, sys->cptr, sys->celem,
                      sys->khat, sys->coef, sys->uvalues, sys->vblk};
        a.u = cur;
        a.up = prev;
        a.out = nxt;
        a.lm = sys->lm;
        a.lm_sstride = sys->lm_sstride;
        a.f1 = sys->f1;
        a.fixed = sys->fixed;
        a.load = *load;
        a.bdy_u = sys->bdy_u;
        a.dt2 = dt * dt;
        a.bdt = beta / dt;
        a.ddt = damp / dt;
        a.order = nullptr;
        const GeomArgs g = geom_args(geo, sys, geom);
        launch_updated(st, a, g, n, beta != 0.0, damp != 0.0, sys->variant, geom != nullptr && geom == cur);
        if (int rc = check_launch("k_step_updated")) return rc;
        double *t = prev;
        prev = cur;
        cur = nxt;
        nxt = t;
    }
    state[0] = cur;
    state[1] = prev;
    state[2] = nxt;
    return CVFES_OK;
}
