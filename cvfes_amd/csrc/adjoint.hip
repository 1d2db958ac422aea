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
#include "rowops.h"

namespace cvfes {
namespace {

struct AdjArgs {
    RowArgs k;             // structure and operator (stored values or element data)
    const double *mu1;     // mu_{n+1}
    const double *mu2;     // mu_{n+2}
    double *out;           // mu_n (projected) or lambda_{-1} (unprojected)
    const double *lm;
    int64_t lm_sstride;
    const double *minv;    // 1/lm (0 where lm = 0), same layout; NULL = divide by lm
    const uint8_t *fixed;
    const double *obs_u;   // u_n      (observation term, may be NULL)
    const double *obs_ref; // u_obs    (NULL = 0)
    const double *obs_w;   // [S] weight of the observation term per sample (NULL = 1)
    double dt2, bdt, dtdamp;
    int project;
    unsigned occ_lds;      // launch only: unused LDS per workgroup that caps the occupancy
    bool nt_out;           // write `out` with non-temporal stores (stream_out)
};

template <int VW>
__device__ __forceinline__ Vec<VW> mass_at(const AdjArgs &a, int64_t node, int64_t s) {
    Vec<VW> m;
    if (a.lm_sstride) {
        m = load_vec<VW>(a.lm + node * a.k.n_smp + s);
    } else {
        const double mv = a.lm[node];
#pragma unroll
        for (int v = 0; v < VW; ++v) m.v[v] = mv;
    }
    return m;
}

// Column term of the adjoint product at node col: x = dt^2/m_col (b1 mu1 - b0 mu2) (0 where the
// mass is 0); K is symmetric, so the transposed product is the same gathered row product.  With the
// inverse masses (minv, 0 where m = 0) the scale is dt^2 * minv_col, one multiply per column
// instead of an FP64 division per column and sample (the division sequence costs the kernel about
// a sixth of its time at 256 samples).
template <bool BETA, int VW>
__device__ __forceinline__ void adj_column(const AdjArgs &a, int64_t col, int64_t s, double (&x)[3][VW]) {
    Vec<VW> q;
    if (a.minv) {
        if (a.lm_sstride) {
            q = load_vec<VW>(a.minv + col * a.k.n_smp + s);
#pragma unroll
            for (int v = 0; v < VW; ++v) q.v[v] *= a.dt2;
        } else {
            const double qc = a.dt2 * a.minv[col];
#pragma unroll
            for (int v = 0; v < VW; ++v) q.v[v] = qc;
        }
    } else {
        const Vec<VW> mc = mass_at<VW>(a, col, s);
#pragma unroll
        for (int v = 0; v < VW; ++v) q.v[v] = mc.v[v] != 0.0 ? a.dt2 / mc.v[v] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int64_t off = (col * 3 + j) * a.k.n_smp + s;
        const Vec<VW> m1 = load_vec<VW>(a.mu1 + off);
        Vec<VW> m2;
        if constexpr (BETA) m2 = load_vec<VW>(a.mu2 + off);
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            double comb = m1.v[v];
            if constexpr (BETA) comb = fma(a.bdt, m1.v[v] - m2.v[v], m1.v[v]);
            x[j][v] = comb * q.v[v];
        }
    }
}

// lambda_n of one block row from acc = K A (b1 mu1 - b0 mu2) and the observation term; stores
// mu_n = P lambda_n (or lambda when not projecting) when `active`.  Explicit fma throughout, so
// the rounding is the same in the separate and the fused kernels.
template <bool DAMP, int VW>
__device__ __forceinline__ void adj_finish(const AdjArgs &a, int32_t row, int64_t s, bool active,
                                           const double (&acc)[3][VW]) {
    const int64_t S = a.k.n_smp;
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
            double lam = fma(2.0 - gam, m1.v[v], -fma(1.0 - gam, m2.v[v], acc[i][v]));
            if (a.obs_u) {
                const double ref = a.obs_ref ? a.obs_ref[off + v] : 0.0;
                const double w = a.obs_w ? a.obs_w[s + v] : 1.0;
                lam = fma(w, a.obs_u[off + v] - ref, lam);
            }
            o.v[v] = fixed ? 0.0 : lam;
        }
        if (active) {   // not read again in this launch: large outputs are streamed (stream_out)
            if (a.nt_out) store_vec_nt<VW>(a.out + off, o);
            else store_vec<VW>(a.out + off, o);
        }
    }
}

template <int OP, bool UNIFORM, bool BETA, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_adjoint_step(AdjArgs a) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    double acc[3][VW];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[i][v] = 0.0;
    for (int32_t p = a.k.indptr[row]; p < a.k.indptr[row + 1]; ++p) {
        const int64_t col = a.k.indices[p];
        double x[3][VW];
        adj_column<BETA, VW>(a, col, s, x);
        block_apply<OP, VW>(a.k, p, col < row, s, x, acc);
    }
    adj_finish<DAMP, VW>(a, row, s, active, acc);
}

// One forward step (the generic row kernel's operations) and one adjoint step in a single pass
// over the operator: each block (stored values, or element coefficients and unit blocks) is
// loaded once and applied to both the forward column u + bdt (u - up) and the adjoint column
// dt^2/m (b1 mu1 - b0 mu2).  The products and updates are the operations of k_step and
// k_adjoint_step in the same order, so both outputs equal those of the separate launches bit for
// bit.  The forward and the adjoint may be at different time steps (the reverse sweep of one
// checkpoint segment runs with the recomputation of the segment before it).  (A register target
// of 5 waves per SIMD, 94 VGPRs instead of 100 for the stored two-sample instance, measured 2-3 %
// slower: the loads in flight per wave count for more here.)
template <int OP, bool UNIFORM, bool BETA, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_fused_step(StepArgs f, AdjArgs a, int64_t step) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(f.k.n_nodes, f.k.n_smp, row, s, active)) return;
    double accf[3][VW], acca[3][VW];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int v = 0; v < VW; ++v) accf[i][v] = acca[i][v] = 0.0;
    for (int32_t p = f.k.indptr[row]; p < f.k.indptr[row + 1]; ++p) {
        const int64_t col = f.k.indices[p];
        double xf[3][VW], xa[3][VW];
        state_column<BETA, VW>(f.u, f.up, f.bdt, col, f.k.n_smp, s, xf);
        adj_column<BETA, VW>(a, col, s, xa);
        block_apply2<OP, VW>(f.k, p, col < row, s, xf, accf, xa, acca);
    }
    step_update<DAMP, VW>(f, row, s, active, step, accf);
    adj_finish<DAMP, VW>(a, row, s, active, acca);
}

struct GradArgs {
    int64_t n_elms, n_smp;
    const int32_t *elems, *cpos;
    const double *khat, *coef, *lm;
    int64_t lm_sstride;
    const double *minv;    // 1/lm (0 where lm = 0), same layout; NULL = divide by lm
    const double *mu, *xc, *xp;
    double bdt, nu_scale;
    double *grad;
};

// sum_{a,b} nu_a^T khat_e(a,b) x_b over the element's 9 x 9 unit block; `kh(ab)` gives the 9
// entries of unit block ab.
template <class KH>
__device__ __forceinline__ double grad_form(const double (&nu)[9], const double (&x)[9], KH kh) {
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

// grad[e,s] = fma(-coef[e,s], sum_{a,b} nu_a^T khat_e(a,b) x_b, grad[e,s]) for the thread's
// (element e, sample s), with nu = nu_scale * mu / m (nu_scale * minv with the inverse masses) and
// x = xc + bdt (xc - xp) (x = xc when xp is NULL).
template <class KH>
__device__ __forceinline__ void grad_update(const GradArgs &a, int64_t e, int64_t s, KH kh) {
    double nu[9], x[9];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int64_t nd = a.elems[3 * e + q];
        const int64_t mi = a.lm_sstride ? nd * a.n_smp + s : nd;
        double sc;
        if (a.minv) {
            sc = a.nu_scale * a.minv[mi];
        } else {
            const double m = a.lm[mi];
            sc = m != 0.0 ? a.nu_scale / m : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (nd * 3 + j) * a.n_smp + s;
            nu[3 * q + j] = sc * a.mu[off];
            const double c = a.xc[off];
            x[3 * q + j] = a.xp ? fma(a.bdt, c - a.xp[off], c) : c;
        }
    }
    const int64_t idx = e * a.n_smp + s;
    a.grad[idx] = fma(-a.coef[idx], grad_form(nu, x, kh), a.grad[idx]);
}

// The same for two consecutive samples per lane (16-byte loads; the element's node scales and
// unit blocks serve 128 samples per wave).
template <class KH>
__device__ __forceinline__ void grad_update2(const GradArgs &a, int64_t e, int64_t s, KH kh) {
    double nu[2][9], x[2][9];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int64_t nd = a.elems[3 * e + q];
        double sc[2];
        if (a.lm_sstride) {
            const Vec<2> m = load_vec<2>((a.minv ? a.minv : a.lm) + nd * a.n_smp + s);
#pragma unroll
            for (int v = 0; v < 2; ++v)
                sc[v] = a.minv ? a.nu_scale * m.v[v] : (m.v[v] != 0.0 ? a.nu_scale / m.v[v] : 0.0);
        } else {
            const double m = a.minv ? a.minv[nd] : a.lm[nd];
            sc[0] = sc[1] = a.minv ? a.nu_scale * m : (m != 0.0 ? a.nu_scale / m : 0.0);
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (nd * 3 + j) * a.n_smp + s;
            const Vec<2> mu = load_vec<2>(a.mu + off);
            const Vec<2> c = load_vec<2>(a.xc + off);
            Vec<2> p;
            if (a.xp) p = load_vec<2>(a.xp + off);
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                nu[v][3 * q + j] = sc[v] * mu.v[v];
                x[v][3 * q + j] = a.xp ? fma(a.bdt, c.v[v] - p.v[v], c.v[v]) : c.v[v];
            }
        }
    }
    const int64_t idx = e * a.n_smp + s;
    const Vec<2> cf = load_vec<2>(a.coef + idx);
    Vec<2> g = load_vec<2>(a.grad + idx);
#pragma unroll
    for (int v = 0; v < 2; ++v) g.v[v] = fma(-cf.v[v], grad_form(nu[v], x[v], kh), g.v[v]);
    store_vec<2>(a.grad + idx, g);
}

// Sample counts that are multiples of 128: each wave one element and 128 of its samples, two per
// lane, waves sample group first as below.  124 VGPRs (4 waves per SIMD; a 5-wave target spills):
// 5.02 ms against 6.24 ms for one sample per lane on the 512,000-node tube with 256 samples.
__global__ __launch_bounds__(kBlock) void k_grad_accum_vw2(GradArgs a) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= a.n_elms * a.n_smp / 2) return;
    const int64_t w = g >> 6, group = w / a.n_elms;
    const int64_t e = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(w - group * a.n_elms));
    const int64_t s = (group * 64 + (g & 63)) * 2;
    grad_update2(a, e, s, [&](int ab) {
        return a.khat + 9 * static_cast<int64_t>(a.cpos[9 * e + ab]);
    });
}

// One thread per (element, sample), unit blocks read from memory.  UNI (sample count a multiple
// of 64): every wave holds one element and 64 of its samples, so its node ids and unit blocks are
// scalar loads, and the waves run sample group first (all elements for samples 0-63, then
// 64-127, ...), which keeps the gathered node data of one group in L2 between the elements that
// share it.  At most 80 registers (6 waves per SIMD or more): left alone, the compiler hoists all
// 81 unit-block loads and runs at 3 waves per SIMD.  (A form that applies two consecutive steps'
// terms in one pass, reading u_{n-1}, coef and grad once, needs 123 registers and measured 10 %
// slower per step.)
template <bool UNI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_grad_accum(GradArgs a) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= a.n_elms * a.n_smp) return;
    int64_t e, s;
    if constexpr (UNI) {
        const int64_t w = g >> 6, group = w / a.n_elms;
        e = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(w - group * a.n_elms));
        s = group * 64 + (g & 63);
    } else {
        e = g / a.n_smp;
        s = g - e * a.n_smp;
    }
    grad_update(a, e, s, [&](int ab) {
        return a.khat + 9 * static_cast<int64_t>(a.cpos[9 * e + ab]);
    });
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
    grad_update(a, e, s, [&](int ab) { return ske + 9 * ab; });
}

template <int MF, bool UNI, int VW>
void adj_launch(hipStream_t st, const AdjArgs &a, bool beta, bool damp) {
    const dim3 g(grid_for(a.k.n_nodes * lanes_per_row<UNI, VW>(a.k.n_smp))), b(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_adjoint_step<MF, UNI, true, true, VW>), g, b, a.occ_lds, st, a);
        else hipLaunchKernelGGL((k_adjoint_step<MF, UNI, true, false, VW>), g, b, a.occ_lds, st, a);
    } else {
        if (damp) hipLaunchKernelGGL((k_adjoint_step<MF, UNI, false, true, VW>), g, b, a.occ_lds, st, a);
        else hipLaunchKernelGGL((k_adjoint_step<MF, UNI, false, false, VW>), g, b, a.occ_lds, st, a);
    }
}

// Lane width and row mapping as the generic forward kernel chooses them (generic_vw; wave-uniform
// rows when they keep 3/4 of the lanes busy).
template <int MF>
void adj_dispatch(hipStream_t st, const AdjArgs &a, bool beta, bool damp) {
    if (generic_vw(a.k.n_smp) == 2) {
        if (use_uniform<2>(a.k.n_smp)) adj_launch<MF, true, 2>(st, a, beta, damp);
        else adj_launch<MF, false, 2>(st, a, beta, damp);
    } else {
        if (use_uniform<1>(a.k.n_smp)) adj_launch<MF, true, 1>(st, a, beta, damp);
        else adj_launch<MF, false, 1>(st, a, beta, damp);
    }
}

template <int MF, bool UNI, int VW>
void fused_launch(hipStream_t st, const StepArgs &f, const AdjArgs &a, bool beta, bool damp) {
    const dim3 g(grid_for(f.k.n_nodes * lanes_per_row<UNI, VW>(f.k.n_smp))), b(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_fused_step<MF, UNI, true, true, VW>), g, b, f.occ_lds, st, f, a, int64_t(0));
        else hipLaunchKernelGGL((k_fused_step<MF, UNI, true, false, VW>), g, b, f.occ_lds, st, f, a, int64_t(0));
    } else {
        if (damp) hipLaunchKernelGGL((k_fused_step<MF, UNI, false, true, VW>), g, b, f.occ_lds, st, f, a, int64_t(0));
        else hipLaunchKernelGGL((k_fused_step<MF, UNI, false, false, VW>), g, b, f.occ_lds, st, f, a, int64_t(0));
    }
}

template <int MF>
void fused_dispatch(hipStream_t st, const StepArgs &f, const AdjArgs &a, bool beta, bool damp) {
    if (generic_vw(f.k.n_smp) == 2) {
        if (use_uniform<2>(f.k.n_smp)) fused_launch<MF, true, 2>(st, f, a, beta, damp);
        else fused_launch<MF, false, 2>(st, f, a, beta, damp);
    } else {
        if (use_uniform<1>(f.k.n_smp)) fused_launch<MF, true, 1>(st, f, a, beta, damp);
        else fused_launch<MF, false, 1>(st, f, a, beta, damp);
    }
}

// Operator form of the generic row kernels for this system and method (as the forward step's
// generic kernel chooses it: full values when present, else the symmetric upper blocks).
int operator_form(const cvfes_system_t *sys, int method) {
    return method == CVFES_METHOD_MATFREE ? kOpMatfree : (sys->values ? kOpStored : kOpSym);
}

int check_adjoint_args(const cvfes_system_t *sys, int method, const double *mu1, const double *mu2,
                       const double *mu_out, double dt, const char *who) {
    CVFES_REQUIRE(sys && sys->indptr && sys->indices && sys->n_nodes >= 1 && sys->n_smp >= 1,
                  "%s: empty system", who);
    CVFES_REQUIRE(method == CVFES_METHOD_STORED || method == CVFES_METHOD_MATFREE,
                  "%s: unknown method %d", who, method);
    CVFES_REQUIRE(method == CVFES_METHOD_MATFREE || sys->values || (sys->uvalues && sys->vblk),
                  "%s: no stored values (full or symmetric)", who);
    CVFES_REQUIRE(method == CVFES_METHOD_STORED || (sys->cptr && sys->celem && sys->khat && sys->coef),
                  "%s: matrix-free needs cptr/celem/khat/coef", who);
    CVFES_REQUIRE(mu1 && mu2 && mu_out && sys->lm, "%s: null pointer", who);
    CVFES_REQUIRE(mu_out != mu1 && mu_out != mu2, "%s: output aliases an input", who);
    CVFES_REQUIRE(dt > 0.0, "%s: dt must be positive", who);
    CVFES_REQUIRE(sys->waves_per_simd == 0 || (sys->waves_per_simd >= 2 && sys->waves_per_simd <= 7),
                  "%s: waves_per_simd must be 0 or 2..7", who);
    return CVFES_OK;
}

AdjArgs adjoint_args(const cvfes_system_t *sys, const double *mu1, const double *mu2, double *mu_out,
                     const double *obs_u, const double *obs_ref, const double *obs_w, double dt,
                     double beta, double damp, int project) {
    AdjArgs a{};
    a.k = RowArgs{sys->n_nodes, sys->n_smp, sys->indptr, sys->indices, sys->values,
                  sys->cptr, sys->celem, sys->khat, sys->coef, sys->uvalues, sys->vblk};
    a.mu1 = mu1;
    a.mu2 = mu2;
    a.out = mu_out;
    a.lm = sys->lm;
    a.lm_sstride = sys->lm_sstride;
    a.minv = sys->minv;
    a.fixed = sys->fixed;
    a.obs_u = obs_u;
    a.obs_ref = obs_ref;
    a.obs_w = obs_w;
    a.dt2 = dt * dt;
    a.bdt = beta / dt;
    a.dtdamp = dt * damp;
    a.project = project;
    a.occ_lds = occupancy_lds(sys->waves_per_simd);
    a.nt_out = stream_out(sys->n_nodes, sys->n_smp);
    return a;
}

// minv[i] = 1 / lm[i], 0 where lm[i] = 0.
__global__ __launch_bounds__(kBlock) void k_inverse_mass(int64_t n, const double *__restrict__ lm,
                                                         double *__restrict__ minv) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) minv[i] = lm[i] != 0.0 ? 1.0 / lm[i] : 0.0;
}
}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int cvfes_adjoint_step(void *stream, const cvfes_system_t *sys, int method,
                                  const double *mu1, const double *mu2, double *mu_out,
                                  const double *obs_u, const double *obs_ref, const double *obs_w,
                                  double dt, double beta, double damp, int project) {
    if (int rc = check_adjoint_args(sys, method, mu1, mu2, mu_out, dt, "cvfes_adjoint_step")) return rc;
    const AdjArgs a = adjoint_args(sys, mu1, mu2, mu_out, obs_u, obs_ref, obs_w, dt, beta, damp, project);
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (operator_form(sys, method)) {
    case kOpMatfree: adj_dispatch<kOpMatfree>(st, a, beta != 0.0, damp != 0.0); break;
    case kOpStored: adj_dispatch<kOpStored>(st, a, beta != 0.0, damp != 0.0); break;
    default: adj_dispatch<kOpSym>(st, a, beta != 0.0, damp != 0.0); break;
    }
    return check_launch("k_adjoint_step");
}

extern "C" int cvfes_adjoint_fused_step(void *stream, const cvfes_system_t *sys, int method,
                                        double *const *state, const cvfes_load_t *load,
                                        const double *mu1, const double *mu2, double *mu_out,
                                        const double *obs_u, const double *obs_ref,
                                        const double *obs_w, double dt, double beta, double damp) {
    if (int rc = check_adjoint_args(sys, method, mu1, mu2, mu_out, dt, "cvfes_adjoint_fused_step"))
        return rc;
    CVFES_REQUIRE(state && state[0] && state[1] && state[2], "cvfes_adjoint_fused_step: null state buffer");
    const double *bufs[6] = {state[0], state[1], state[2], mu1, mu2, mu_out};
    for (int i = 0; i < 6; ++i)
        for (int j = i + 1; j < 6; ++j)
            CVFES_REQUIRE(bufs[i] != bufs[j] || (i == 3 && j == 4),
                          "cvfes_adjoint_fused_step: state and adjoint buffers must be distinct");
    CVFES_REQUIRE(sys->f1 && load && load->p && (!load->ga || load->qa) && (!load->gb || load->qb),
                  "cvfes_adjoint_fused_step: load tables");
    CVFES_REQUIRE(sys->lm_sstride == 0 || sys->lm_sstride == 1, "cvfes_adjoint_fused_step: lm_sstride");
    const int64_t kern = cvfes_step_kernel(sys, method);
    CVFES_REQUIRE(kern == CVFES_VARIANT_GENERIC,
                  "cvfes_adjoint_fused_step: the forward step of this system runs kernel variant "
                  "%lld, not the generic one the fused kernel replays",
                  static_cast<long long>(kern));
    StepArgs f{};
    f.k = RowArgs{sys->n_nodes, sys->n_smp, sys->indptr, sys->indices, sys->values,
                  sys->cptr, sys->celem, sys->khat, sys->coef, sys->uvalues, sys->vblk};
    f.u = state[0];
    f.up = state[1];
    f.out = state[2];
    f.lm = sys->lm;
    f.lm_sstride = sys->lm_sstride;
    f.f1 = sys->f1;
    f.fixed = sys->fixed;
    f.load = *load;
    f.bdy_u = sys->bdy_u;
    f.dt2 = dt * dt;
    f.bdt = beta / dt;
    f.ddt = damp / dt;
    f.order = nullptr;
    f.occ_lds = occupancy_lds(sys->waves_per_simd);
    f.nt_out = stream_out(sys->n_nodes, sys->n_smp);
    const AdjArgs a = adjoint_args(sys, mu1, mu2, mu_out, obs_u, obs_ref, obs_w, dt, beta, damp, 1);
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (operator_form(sys, method)) {
    case kOpMatfree: fused_dispatch<kOpMatfree>(st, f, a, beta != 0.0, damp != 0.0); break;
    case kOpStored: fused_dispatch<kOpStored>(st, f, a, beta != 0.0, damp != 0.0); break;
    default: fused_dispatch<kOpSym>(st, f, a, beta != 0.0, damp != 0.0); break;
    }
    return check_launch("k_fused_step");
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
                     sys->lm_sstride, sys->minv, mu, x_cur, x_prev, beta / dt, nu_scale, grad};
    const dim3 grid(grid_for(sys->n_elms * sys->n_smp)), block(kBlock);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (sys->n_smp % 128 == 0)
        hipLaunchKernelGGL(k_grad_accum_vw2, dim3(grid_for(sys->n_elms * sys->n_smp / 2)), block, 0, st, a);
    else if (sys->n_smp % 64 == 0)
        hipLaunchKernelGGL(k_grad_accum<true>, grid, block, 0, st, a);
    else if (sys->n_smp >= kGradLdsMinSmp)
        hipLaunchKernelGGL(k_grad_accum_lds, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_grad_accum<false>, grid, block, 0, st, a);
    return check_launch("k_grad_accum");
}

extern "C" int cvfes_inverse_mass(void *stream, int64_t n, const double *lm, double *minv) {
    CVFES_REQUIRE(n >= 0, "cvfes_inverse_mass: negative size");
    if (n == 0) return CVFES_OK;
    CVFES_REQUIRE(lm && minv, "cvfes_inverse_mass: null pointer");
    hipLaunchKernelGGL(k_inverse_mass, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), n, lm, minv);
    return check_launch("k_inverse_mass");
}
