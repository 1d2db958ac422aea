// Element-level kernels: membrane stiffness blocks, stiffness coefficients, lumped mass and load.
//
// All element math is the linear constant-strain triangle of CVFES (CST membrane with the 5x5
// constitutive matrix including the k = 5/6 transverse-shear rows):
//   frame   e1 = (X2-X1)/|X2-X1|, e2 = normalised (X0-X2) minus its e1 component, e3 = e1 x e2
//           (optimizedSolidAssemble.pyx:82-114)
//   local   x_a = e1.X_a, y_a = e2.X_a, area from the 3x3 determinant (pyx:274-284)
//   edges   b = (y1-y2, y2-y0, y0-y1), c = (x2-x1, x0-x2, x1-x0) (pyx:292-297)
//   block   k_ab = [[b_a b_b + g c_a c_b, nu b_a c_b + g c_a b_b, 0],
//                   [nu c_a b_b + g b_a c_b, c_a c_b + g b_a b_b, 0],
//                   [0, 0, gk (b_a b_b + c_a c_b)]] / (4 A (1 - nu^2)),  g = (1-nu)/2, gk = k g
//           global block T^T k_ab T with T rows (e1, e2, e3)  (physicsSolverSDampGPUs.cl:340-357)
// which equals A * TT^T B^T D B TT of pyx:299-329 per unit stiffness coefficient.
#include "geometry.h"

namespace cvfes {
namespace {

__global__ __launch_bounds__(kBlock) void k_element_matrices(
    int64_t n_elms, const double *__restrict__ nodes, const int32_t *__restrict__ elems, double nu,
    double kshear, const int32_t *__restrict__ cpos, double *__restrict__ khat,
    double *__restrict__ area_out, double *__restrict__ normal_out) {
    const int64_t e = xcd_block() * kBlock + threadIdx.x;
    if (e >= n_elms) return;
    const int n0 = elems[3 * e], n1 = elems[3 * e + 1], n2 = elems[3 * e + 2];
    const Frame f = element_frame(nodes, n0, n1, n2);
    const double b[3] = {f.ly[1] - f.ly[2], f.ly[2] - f.ly[0], f.ly[0] - f.ly[1]};
    const double c[3] = {f.lx[2] - f.lx[1], f.lx[0] - f.lx[2], f.lx[1] - f.lx[0]};
    const double g = 0.5 * (1.0 - nu);
    const double gk = 0.5 * kshear * (1.0 - nu);
    const double scale = 1.0 / (4.0 * f.area * (1.0 - nu * nu));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
            const double k00 = b[a] * b[bb] + g * c[a] * c[bb];
            const double k01 = nu * b[a] * c[bb] + g * c[a] * b[bb];
            const double k10 = nu * c[a] * b[bb] + g * b[a] * c[bb];
            const double k11 = c[a] * c[bb] + g * b[a] * b[bb];
            const double k22 = gk * (b[a] * b[bb] + c[a] * c[bb]);
            double *out = khat + 9 * static_cast<int64_t>(cpos[9 * e + 3 * a + bb]);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double v = f.t[0][i] * (k00 * f.t[0][j] + k01 * f.t[1][j]) +
                                     f.t[1][i] * (k10 * f.t[0][j] + k11 * f.t[1][j]) +
                                     f.t[2][i] * (k22 * f.t[2][j]);
                    out[3 * i + j] = v * scale;
                }
            }
        }
    }
    area_out[e] = f.area;
#pragma unroll
    for (int k = 0; k < 3; ++k) normal_out[3 * e + k] = f.t[2][k];
}

// coef[e,s] = sum_q w_q h_q E_q, Gauss points at edge midpoints (0,1), (0,2), (1,2), w_q = 1/3
// (physicsSolver.py:149-165,200-202).
__global__ __launch_bounds__(kBlock) void k_element_coef(
    int64_t n_elms, int64_t n_smp, const int32_t *__restrict__ elems,
    const double *__restrict__ vthick, int64_t tss, const double *__restrict__ vE,
    const double *__restrict__ elm_E, int log_modulus, double *__restrict__ coef) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= n_elms * n_smp) return;
    const int64_t e = g / n_smp;
    const int64_t s = g - e * n_smp;
    const int64_t tns = tss ? n_smp : 1;
    const int64_t n[3] = {elems[3 * e], elems[3 * e + 1], elems[3 * e + 2]};
    const double h0 = vthick[n[0] * tns + s * tss], h1 = vthick[n[1] * tns + s * tss],
                 h2 = vthick[n[2] * tns + s * tss];
    const double hq0 = 0.5 * h0 + 0.5 * h1, hq1 = 0.5 * h0 + 0.5 * h2, hq2 = 0.5 * h1 + 0.5 * h2;
    const double w = 1.0 / 3.0;
    double out;
    if (elm_E) {
        const double E = log_modulus ? exp(elm_E[g]) : elm_E[g];
        out = (hq0 * E) * w + (hq1 * E) * w + (hq2 * E) * w;
    } else {
        const double E0 = vE[n[0] * n_smp + s], E1 = vE[n[1] * n_smp + s], E2 = vE[n[2] * n_smp + s];
        const double Eq0 = 0.5 * E0 + 0.5 * E1, Eq1 = 0.5 * E0 + 0.5 * E2, Eq2 = 0.5 * E1 + 0.5 * E2;
        out = (hq0 * Eq0) * w + (hq1 * Eq1) * w + (hq2 * Eq2) * w;
    }
    coef[g] = out;
}

// One thread per (node, sample): pull the lumped mass from the incident elements; the s == 0 thread
// also pulls the unit-pressure load.
__global__ __launch_bounds__(kBlock) void k_lumped_mass_load(
    int64_t n_nodes, int64_t n_lm_smp, int64_t n_smp, const int32_t *__restrict__ nptr,
    const int32_t *__restrict__ nent, const int32_t *__restrict__ elems,
    const double *__restrict__ area, const double *__restrict__ normal,
    const double *__restrict__ vthick, int64_t tss, double density, double mass_scale,
    double *__restrict__ lm, double *__restrict__ f1) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= n_nodes * n_lm_smp) return;
    const int64_t a = g / n_lm_smp;
    const int64_t s = g - a * n_lm_smp;
    const int64_t tns = tss ? n_smp : 1;
    double m = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int32_t q = nptr[a]; q < nptr[a + 1]; ++q) {
        const int64_t e = nent[q] / 3;
        const int la = nent[q] - 3 * static_cast<int32_t>(e);
        const int64_t na = elems[3 * e + la];
        const int64_t nb = elems[3 * e + (la + 1) % 3];
        const int64_t nc = elems[3 * e + (la + 2) % 3];
        const double ha = vthick[na * tns + s * tss];
        const double hb = vthick[nb * tns + s * tss];
        const double hc = vthick[nc * tns + s * tss];
        const double A = area[e];
        m += A * (0.5 * ha + 0.25 * hb + 0.25 * hc);
        const double w = (1.0 / 3.0) * A;
        fx += w * normal[3 * e];
        fy += w * normal[3 * e + 1];
        fz += w * normal[3 * e + 2];
    }
    lm[g] = density * mass_scale * m;
    if (s == 0) {
        f1[3 * a] = fx;
        f1[3 * a + 1] = fy;
        f1[3 * a + 2] = fz;
    }
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int cvfes_element_matrices(void *stream, int64_t n_elms, const double *nodes,
                                      const int32_t *elems, double nu, double kshear,
                                      const int32_t *cpos, double *khat, double *area,
                                      double *normal) {
    CVFES_REQUIRE(n_elms >= 0, "cvfes_element_matrices: n_elms < 0");
    if (n_elms == 0) return CVFES_OK;
    CVFES_REQUIRE(nodes && elems && cpos && khat && area && normal,
                  "cvfes_element_matrices: null pointer");
    CVFES_REQUIRE(nu > -1.0 && nu < 1.0, "cvfes_element_matrices: Poisson ratio out of (-1, 1)");
    hipLaunchKernelGGL(k_element_matrices, dim3(grid_for(n_elms)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), n_elms, nodes, elems, nu, kshear, cpos,
                       khat, area, normal);
    return check_launch("k_element_matrices");
}

extern "C" int cvfes_element_coef(void *stream, int64_t n_elms, int64_t n_smp, const int32_t *elems,
                                  const double *vthick, int64_t thick_sstride, const double *vE,
                                  const double *elm_E, int log_modulus, double *coef) {
    CVFES_REQUIRE(n_elms >= 0 && n_smp >= 1, "cvfes_element_coef: bad sizes");
    CVFES_REQUIRE(thick_sstride == 0 || thick_sstride == 1, "cvfes_element_coef: thick_sstride must be 0 or 1");
    if (n_elms == 0) return CVFES_OK;
    CVFES_REQUIRE(elems && vthick && coef && (vE || elm_E), "cvfes_element_coef: null pointer");
    CVFES_REQUIRE(!log_modulus || elm_E, "cvfes_element_coef: log_modulus needs elm_E");
    hipLaunchKernelGGL(k_element_coef, dim3(grid_for(n_elms * n_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), n_elms, n_smp, elems, vthick,
                       thick_sstride, vE, elm_E, log_modulus, coef);
    return check_launch("k_element_coef");
}

extern "C" int cvfes_lumped_mass_load(void *stream, int64_t n_nodes, int64_t n_smp,
                                      const int32_t *nptr, const int32_t *nent,
                                      const int32_t *elems, const double *area,
                                      const double *normal, const double *vthick,
                                      int64_t thick_sstride, double density, double mass_scale,
                                      double *lm, double *f1) {
    CVFES_REQUIRE(n_nodes >= 0 && n_smp >= 1, "cvfes_lumped_mass_load: bad sizes");
    CVFES_REQUIRE(thick_sstride == 0 || thick_sstride == 1, "cvfes_lumped_mass_load: thick_sstride must be 0 or 1");
    if (n_nodes == 0) return CVFES_OK;
    CVFES_REQUIRE(nptr && nent && elems && area && normal && vthick && lm && f1,
                  "cvfes_lumped_mass_load: null pointer");
    const int64_t n_lm_smp = thick_sstride ? n_smp : 1;
    hipLaunchKernelGGL(k_lumped_mass_load, dim3(grid_for(n_nodes * n_lm_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), n_nodes, n_lm_smp, n_smp, nptr, nent,
                       elems, area, normal, vthick, thick_sstride, density, mass_scale, lm, f1);
    return check_launch("k_lumped_mass_load");
}
