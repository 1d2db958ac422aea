// Adjoint of the explicit central-difference step, and the gradient with respect to the element
// log-modulus.
//
// Forward step (per sample, A = dt^2 M^-1 diagonal, P zeroes fixed dofs, b1 = 1 + beta/dt,
// b0 = beta/dt, gamma = dt damp / m):
//   u_{n+1} = P[(2-gamma) u_n - (1-gamma) u_{n-1} - A K (b1 u_n - b0 u_{n-1}) + A f_n] + (I-P) g.
// For L = sum_n l_n(u_n) the adjoint recurrence, with mu_n = P lambda_n and K symmetric, is
//   lambda_n = dl_n/du_n + (2-gamma) mu_{n+1} - (1-gamma) mu_{n+2} - K A (b1 mu_{n+1} - b0 mu_{n+2}),
// the same gathered block product as the forward step with a column scale dt^2/m.  With
// coef_e = h E_e and K = sum_e coef_e khat_e, the derivative with respect to theta_e = log E_e is
//   dL/dtheta_e = -coef_e sum_n (A mu_{n+1})_e^T khat_e (b1 u_n - b0 u_{n-1})_e.
// u_{-1} depends on K through the initial acceleration; its adjoint
//   lambda_{-1} = -(1-gamma) mu_1 + b0 K A mu_1   (u_{-1} is not projected)
// adds -coef_e (dt^2/2 M^-1 lambda_{-1})_e^T khat_e (u0 + beta du0)_e.
#include "common.h"

namespace cvfes {
namespace {

struct AdjArgs {
    int64_t n_nodes, n_smp;
    const int32_t *indptr, *indices;
    const double *values;
    const int32_t *cptr, *celem;
    const double *khat, *coef;
    const double *mu1;     // mu_{n+1}
    const double *mu2;     // mu_{n+2}
    double *out;           // mu_n (projected) or lambda_{-1} (unprojected)
    const double *lm;
    int64_t lm_sstride;
    const uint8_t *fixed;
    const double *obs_u;   // u_n      (observation term, may be NULL)
    const double *obs_ref; // u_obs    (NULL = 0)
    const double *obs_w;   // [S] weight of the observation term per sample (NULL = 1)
    double dt2, bdt, dtdamp;
    int project;
};

// acc[i][v] += K_p (i, :) . x for block p (K symmetric: the same block serves the transposed
// product).  Matrix-free: acc += coef_e (khat_c x) over the block's element contributions, without
// forming the block.
template <bool MATFREE, int VW>
__device__ __forceinline__ void adj_block_apply(const AdjArgs &a, int32_t p, int64_t s,
                                                const double (&x)[3][VW], double (&acc)[3][VW]) {
    if constexpr (!MATFREE) {
        const double *base = a.values + (static_cast<int64_t>(p) * 9) * a.n_smp + s;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const Vec<VW> kv = load_vec<VW>(base + (3 * i + j) * a.n_smp);
#pragma unroll
                for (int v = 0; v < VW; ++v) acc[i][v] = fma(kv.v[v], x[j][v], acc[i][v]);
            }
    } else {
        for (int32_t c = a.cptr[p]; c < a.cptr[p + 1]; ++c) {
            const Vec<VW> cf = load_vec<VW>(a.coef + static_cast<int64_t>(a.celem[c]) * a.n_smp + s);
            const double *kh = a.khat + 9 * static_cast<int64_t>(c);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double h0 = kh[3 * i], h1 = kh[3 * i + 1], h2 = kh[3 * i + 2];
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    const double t = fma(h2, x[2][v], fma(h1, x[1][v], h0 * x[0][v]));
                    acc[i][v] = fma(cf.v[v], t, acc[i][v]);
                }
            }
        }
    }
}

template <int VW>
__device__ __forceinline__ Vec<VW> mass_at(const AdjArgs &a, int64_t node, int64_t s) {
    Vec<VW> m;
    if (a.lm_sstride) {
        m = load_vec<VW>(a.lm + node * a.n_smp + s);
    } else {
        const double mv = a.lm[node];
#pragma unroll
        for (int v = 0; v < VW; ++v) m.v[v] = mv;
    }
    return m;
}

template <bool UNIFORM, int VW>
__host__ __device__ __forceinline__ int64_t adj_lanes_per_row(int64_t n_smp) {
    const int64_t l = n_smp / VW;
    return UNIFORM ? (l + 63) / 64 * 64 : l;
}

template <bool MATFREE, bool UNIFORM, bool BETA, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_adjoint_step(AdjArgs a) {
    const int64_t lpr = adj_lanes_per_row<UNIFORM, VW>(a.n_smp);
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    int64_t r = g / lpr;
    int64_t s = (g - r * lpr) * VW;
    const bool active = s < a.n_smp;
    if (!active) s = a.n_smp - VW;
    if constexpr (UNIFORM) r = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(r));
    if (r >= a.n_nodes) return;
    const int32_t row = static_cast<int32_t>(r);
    const int64_t S = a.n_smp;
    double acc[3][VW];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[i][v] = 0.0;
    for (int32_t p = a.indptr[row]; p < a.indptr[row + 1]; ++p) {
        const int64_t col = a.indices[p];
        const Vec<VW> mc = mass_at<VW>(a, col, s);
        double x[3][VW];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (col * 3 + j) * S + s;
            const Vec<VW> m1 = load_vec<VW>(a.mu1 + off);
            Vec<VW> m2;
            if constexpr (BETA) m2 = load_vec<VW>(a.mu2 + off);
#pragma unroll
            for (int v = 0; v < VW; ++v) {
                double comb = m1.v[v];
                if constexpr (BETA) comb = fma(a.bdt, m1.v[v] - m2.v[v], m1.v[v]);
                x[j][v] = mc.v[v] != 0.0 ? comb * (a.dt2 / mc.v[v]) : 0.0;
            }
        }
        adj_block_apply<MATFREE, VW>(a, p, s, x, acc);
    }
    const bool fixed = a.project && a.fixed != nullptr && a.fixed[row] != 0;
    const Vec<VW> m = mass_at<VW>(a, row, s);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * S + s;
        const Vec<VW> m1 = load_vec<VW>(a.mu1 + off);
        const Vec<VW> m2 = load_vec<VW>(a.mu2 + off);
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            double gam = 0.0;
            if constexpr (DAMP) gam = m.v[v] != 0.0 ? a.dtdamp / m.v[v] : 0.0;
            double lam = (2.0 - gam) * m1.v[v] - (1.0 - gam) * m2.v[v] - acc[i][v];
            if (a.obs_u) {
                const double ref = a.obs_ref ? a.obs_ref[off + v] : 0.0;
                const double w = a.obs_w ? a.obs_w[s + v] : 1.0;
                lam += w * (a.obs_u[off + v] - ref);
            }
            o.v[v] = fixed ? 0.0 : lam;
        }
        if (active) store_vec<VW>(a.out + off, o);
    }
}

struct GradArgs {
    int64_t n_elms, n_smp;
    const int32_t *elems, *cpos;
    const double *khat, *coef, *lm;
    int64_t lm_sstride;
    const double *mu, *xc, *xp;
    double bdt, nu_scale;
    double *grad;
};

// sum_{a,b} nu_a^T khat_e(a,b) x_b for the thread's (element e, sample s), with nu = nu_scale * mu / m
// and x = xc + bdt (xc - xp) (xp may be NULL).  `kh(ab)` gives the 9 entries of unit block ab.
template <class KH>
__device__ __forceinline__ double grad_sum(const GradArgs &a, int64_t e, int64_t s, KH kh) {
    int64_t nd[3];
    double nu[9], x[9];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        nd[q] = a.elems[3 * e + q];
        const double m = a.lm[a.lm_sstride ? nd[q] * a.n_smp + s : nd[q]];
        const double sc = m != 0.0 ? a.nu_scale / m : 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (nd[q] * 3 + j) * a.n_smp + s;
            nu[3 * q + j] = sc * a.mu[off];
            const double c = a.xc[off];
            x[3 * q + j] = a.xp ? fma(a.bdt, c - a.xp[off], c) : c;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const auto *h = kh(3 * p + b);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) t = fma(h[3 * i + j], x[3 * b + j], t);
                acc = fma(nu[3 * p + i], t, acc);
            }
        }
    }
    return acc;
}

// grad[e,s] += -coef[e,s] * grad_sum, one thread per (element, sample), unit blocks read from
// memory.  UNI (sample count a multiple of 64): every wave holds one element, so its node ids and
// unit blocks are scalar loads.  At most 80 registers (6 waves per SIMD or more): left alone, the
// compiler hoists all 81 unit-block loads and runs at 3 waves per SIMD.
template <bool UNI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_grad_accum(GradArgs a) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= a.n_elms * a.n_smp) return;
    int64_t e = g / a.n_smp;
    if constexpr (UNI) e = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(e));
    const int64_t s = g - e * a.n_smp;
    const double acc = grad_sum(a, e, s, [&](int ab) {
        return a.khat + 9 * static_cast<int64_t>(a.cpos[9 * e + ab]);
    });
    a.grad[g] -= a.coef[g] * acc;
}

// The same with the unit blocks of the workgroup's elements staged in LDS first (cooperative
// loads, about 3 per thread at 32 samples), for sample counts that put several elements in a
// wave: the lanes of one element then read its 81 entries from LDS instead of issuing 81 vector
// loads each.  Needs n_smp >= kGradLdsMinSmp (at most kGradLdsElems elements per workgroup).
constexpr int kGradLdsMinSmp = 8;
constexpr int kGradLdsElems = (kBlock - 1) / kGradLdsMinSmp + 2;   // 33

__global__ __launch_bounds__(kBlock) void k_grad_accum_lds(GradArgs a) {
    __shared__ double sk[kGradLdsElems * 81];
    const int64_t g0 = xcd_block() * kBlock;
    const int64_t total = a.n_elms * a.n_smp;
    const int64_t e0 = g0 / a.n_smp;
    int64_t e1 = (g0 + kBlock - 1) / a.n_smp;
    if (e1 > a.n_elms - 1) e1 = a.n_elms - 1;
    const int nel = e0 <= e1 ? static_cast<int>(e1 - e0 + 1) : 0;    // <= kGradLdsElems
    for (int q = threadIdx.x; q < nel * 81; q += kBlock) {
        const int le = q / 81, r = q - le * 81, ab = r / 9, ij = r - ab * 9;
        sk[q] = a.khat[9 * static_cast<int64_t>(a.cpos[9 * (e0 + le) + ab]) + ij];
    }
    __syncthreads();
    const int64_t g = g0 + threadIdx.x;
    if (g >= total) return;
    const int64_t e = g / a.n_smp;
    const int64_t s = g - e * a.n_smp;
    const double *ske = sk + (e - e0) * 81;
    const double acc = grad_sum(a, e, s, [&](int ab) { return ske + 9 * ab; });
    a.grad[g] -= a.coef[g] * acc;
}

template <bool MF, bool UNI, int VW>
void adj_launch(hipStream_t st, const AdjArgs &a, bool beta, bool damp) {
    const dim3 g(grid_for(a.n_nodes * adj_lanes_per_row<UNI, VW>(a.n_smp))), b(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_adjoint_step<MF, UNI, true, true, VW>), g, b, 0, st, a);
        else hipLaunchKernelGGL((k_adjoint_step<MF, UNI, true, false, VW>), g, b, 0, st, a);
    } else {
        if (damp) hipLaunchKernelGGL((k_adjoint_step<MF, UNI, false, true, VW>), g, b, 0, st, a);
        else hipLaunchKernelGGL((k_adjoint_step<MF, UNI, false, false, VW>), g, b, 0, st, a);
    }
}

template <int VW>
bool adj_uniform(int64_t n_smp) {
    const int64_t l = n_smp / VW, padded = (l + 63) / 64 * 64;
    return 4 * l >= 3 * padded;
}

template <bool MF>
void adj_dispatch(hipStream_t st, const AdjArgs &a, bool beta, bool damp) {
    if (a.n_smp % 2 == 0) {
        if (adj_uniform<2>(a.n_smp)) adj_launch<MF, true, 2>(st, a, beta, damp);
        else adj_launch<MF, false, 2>(st, a, beta, damp);
    } else {
        if (adj_uniform<1>(a.n_smp)) adj_launch<MF, true, 1>(st, a, beta, damp);
        else adj_launch<MF, false, 1>(st, a, beta, damp);
    }
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int cvfes_adjoint_step(void *stream, const cvfes_system_t *sys, int method,
                                  const double *mu1, const double *mu2, double *mu_out,
                                  const double *obs_u, const double *obs_ref, const double *obs_w,
                                  double dt, double beta, double damp, int project) {
    CVFES_REQUIRE(sys && sys->indptr && sys->indices && sys->n_nodes >= 1 && sys->n_smp >= 1,
                  "cvfes_adjoint_step: empty system");
    CVFES_REQUIRE(method == CVFES_METHOD_STORED || method == CVFES_METHOD_MATFREE,
                  "cvfes_adjoint_step: unknown method %d", method);
    CVFES_REQUIRE(method == CVFES_METHOD_MATFREE || sys->values, "cvfes_adjoint_step: no values");
    CVFES_REQUIRE(method == CVFES_METHOD_STORED || (sys->cptr && sys->celem && sys->khat && sys->coef),
                  "cvfes_adjoint_step: matrix-free needs cptr/celem/khat/coef");
    CVFES_REQUIRE(mu1 && mu2 && mu_out && sys->lm, "cvfes_adjoint_step: null pointer");
    CVFES_REQUIRE(mu_out != mu1 && mu_out != mu2, "cvfes_adjoint_step: output aliases an input");
    CVFES_REQUIRE(dt > 0.0, "cvfes_adjoint_step: dt must be positive");
    AdjArgs a{};
    a.n_nodes = sys->n_nodes;
    a.n_smp = sys->n_smp;
    a.indptr = sys->indptr;
    a.indices = sys->indices;
    a.values = sys->values;
    a.cptr = sys->cptr;
    a.celem = sys->celem;
    a.khat = sys->khat;
    a.coef = sys->coef;
    a.mu1 = mu1;
    a.mu2 = mu2;
    a.out = mu_out;
    a.lm = sys->lm;
    a.lm_sstride = sys->lm_sstride;
    a.fixed = sys->fixed;
    a.obs_u = obs_u;
    a.obs_ref = obs_ref;
    a.obs_w = obs_w;
    a.dt2 = dt * dt;
    a.bdt = beta / dt;
    a.dtdamp = dt * damp;
    a.project = project;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (method == CVFES_METHOD_MATFREE)
        adj_dispatch<true>(st, a, beta != 0.0, damp != 0.0);
    else
        adj_dispatch<false>(st, a, beta != 0.0, damp != 0.0);
    return check_launch("k_adjoint_step");
}

extern "C" int cvfes_grad_accum(void *stream, const cvfes_system_t *sys, const int32_t *elems,
                                const int32_t *cpos, const double *mu, const double *x_cur,
                                const double *x_prev, double dt, double beta, double nu_scale,
                                double *grad) {
    CVFES_REQUIRE(sys && sys->n_elms >= 1 && sys->n_smp >= 1, "cvfes_grad_accum: empty system");
    CVFES_REQUIRE(elems && cpos && sys->khat && sys->coef && sys->lm && mu && x_cur && grad,
                  "cvfes_grad_accum: null pointer");
    CVFES_REQUIRE(dt > 0.0, "cvfes_grad_accum: dt must be positive");
    const GradArgs a{sys->n_elms, sys->n_smp, elems, cpos, sys->khat, sys->coef, sys->lm,
                     sys->lm_sstride, mu, x_cur, x_prev, beta / dt, nu_scale, grad};
    const dim3 grid(grid_for(sys->n_elms * sys->n_smp)), block(kBlock);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (sys->n_smp % 64 == 0)
        hipLaunchKernelGGL(k_grad_accum<true>, grid, block, 0, st, a);
    else if (sys->n_smp >= kGradLdsMinSmp)
        hipLaunchKernelGGL(k_grad_accum_lds, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_grad_accum<false>, grid, block, 0, st, a);
    return check_launch("k_grad_accum");
}
