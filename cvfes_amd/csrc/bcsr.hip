// Block-CSR stiffness: assembly of sample-fastest values, the multi-sample product, and the fused
// explicit central-difference step.
//
// Thread mapping of the row kernels: the (block row, sample group) space is flattened, with
// VW consecutive samples per lane (VW = 2 -> 16-byte loads).  For S = 128 one 64-lane wavefront owns
// one block row and streams that row's values [p][9][S] (9 x 1 KiB per block) contiguously, while
// the column state u[col][3][S] is gathered as coalesced 1 KiB rows that hit L2 (the XCD-contiguous
// block order keeps a row's neighbours on the same XCD).  When S/VW is a multiple of 64 the row
// index is wave-uniform and is moved to a scalar register, so row pointers, column indices and the
// matrix-free element blocks come in through scalar loads.
//
// The stiffness operator is either
//   METHOD_STORED : values[p][ij][s] read from HBM, or
//   METHOD_MATFREE: values rebuilt per block as sum_c coef[celem[c], s] * khat[c][ij]
//                   (sample-independent khat, scalar loads; coef [E][S] is cached across the 3
//                   rows that touch an element).
#include "common.h"

namespace cvfes {
namespace {

struct RowArgs {
    int64_t n_nodes;
    int64_t n_smp;
    const int32_t *indptr;
    const int32_t *indices;
    const double *values;
    const int32_t *cptr;
    const int32_t *celem;
    const double *khat;
    const double *coef;
};

struct StepArgs {
    RowArgs k;
    const double *u;      // u_n   [N][3][S]
    const double *up;     // u_n-1 [N][3][S]
    double *out;          // u_n+1 (step) or y = K u (spmv)
    const double *lm;
    int64_t lm_sstride;
    const double *f1;
    const uint8_t *fixed;
    cvfes_load_t load;
    double bdy_u;
    double dt2;
    double bdt;           // beta / dt
    double ddt;           // damp / dt
};

template <bool MATFREE, int VW>
__device__ __forceinline__ void block_values(const RowArgs &k, int32_t p, int64_t s,
                                             double (&kv)[9][VW]) {
    if constexpr (!MATFREE) {
        const double *base = k.values + (static_cast<int64_t>(p) * 9) * k.n_smp + s;
#pragma unroll
        for (int ij = 0; ij < 9; ++ij) {
            const Vec<VW> t = load_vec<VW>(base + ij * k.n_smp);
#pragma unroll
            for (int v = 0; v < VW; ++v) kv[ij][v] = t.v[v];
        }
    } else {
#pragma unroll
        for (int ij = 0; ij < 9; ++ij)
#pragma unroll
            for (int v = 0; v < VW; ++v) kv[ij][v] = 0.0;
        const int32_t c0 = k.cptr[p], c1 = k.cptr[p + 1];
        for (int32_t c = c0; c < c1; ++c) {
            const int64_t e = k.celem[c];
            const Vec<VW> cf = load_vec<VW>(k.coef + e * k.n_smp + s);
            const double *kh = k.khat + 9 * static_cast<int64_t>(c);
#pragma unroll
            for (int ij = 0; ij < 9; ++ij) {
                const double h = kh[ij];
#pragma unroll
                for (int v = 0; v < VW; ++v) kv[ij][v] = fma(cf.v[v], h, kv[ij][v]);
            }
        }
    }
}

// acc[i][v] = sum over the row's blocks of K_p (i, :) . x_col, with x = u + bdt (u - up) when BETA.
template <bool MATFREE, bool BETA, int VW>
__device__ __forceinline__ void row_product(const RowArgs &k, const double *__restrict__ u,
                                            const double *__restrict__ up, double bdt, int32_t row,
                                            int64_t s, double (&acc)[3][VW]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[i][v] = 0.0;
    const int32_t p0 = k.indptr[row], p1 = k.indptr[row + 1];
    for (int32_t p = p0; p < p1; ++p) {
        const int64_t col = k.indices[p];
        double x[3][VW];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (col * 3 + j) * k.n_smp + s;
            const Vec<VW> xu = load_vec<VW>(u + off);
            if constexpr (BETA) {
                const Vec<VW> xp = load_vec<VW>(up + off);
#pragma unroll
                for (int v = 0; v < VW; ++v) x[j][v] = fma(bdt, xu.v[v] - xp.v[v], xu.v[v]);
            } else {
#pragma unroll
                for (int v = 0; v < VW; ++v) x[j][v] = xu.v[v];
            }
        }
        double kv[9][VW];
        block_values<MATFREE, VW>(k, p, s, kv);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int v = 0; v < VW; ++v) acc[i][v] = fma(kv[3 * i + j][v], x[j][v], acc[i][v]);
    }
}

template <bool UNIFORM, int VW>
__device__ __forceinline__ bool row_and_sample(int64_t n_nodes, int64_t n_smp, int32_t &row,
                                               int64_t &s) {
    const int64_t lpr = n_smp / VW;   // lanes per block row
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    int64_t r = g / lpr;
    s = (g - r * lpr) * VW;
    if constexpr (UNIFORM) r = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(r));
    row = static_cast<int32_t>(r);
    return r < n_nodes;
}

template <bool MATFREE, bool UNIFORM, int VW>
__global__ __launch_bounds__(kBlock) void k_spmv(StepArgs a) {
    int32_t row;
    int64_t s;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s)) return;
    double acc[3][VW];
    row_product<MATFREE, false, VW>(a.k, a.u, nullptr, 0.0, row, s, acc);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) o.v[v] = acc[i][v];
        store_vec<VW>(a.out + (static_cast<int64_t>(row) * 3 + i) * a.k.n_smp + s, o);
    }
}

template <bool MATFREE, bool UNIFORM, bool BETA, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_step(StepArgs a, int64_t step) {
    int32_t row;
    int64_t s;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s)) return;
    double acc[3][VW];
    row_product<MATFREE, BETA, VW>(a.k, a.u, a.up, a.bdt, row, s, acc);
    const double pn = a.load.p[step];
    const double qa = a.load.ga ? a.load.qa[step] : 0.0;
    const double qb = a.load.gb ? a.load.qb[step] : 0.0;
    const bool fixed = a.fixed != nullptr && a.fixed[row] != 0;
    const int64_t nsm = a.k.n_smp;
    const double *lmp = a.lm + (a.lm_sstride ? static_cast<int64_t>(row) * nsm + s : row);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * nsm + s;
        const Vec<VW> uu = load_vec<VW>(a.u + off);
        const Vec<VW> pp = load_vec<VW>(a.up + off);
        const int64_t d = 3 * static_cast<int64_t>(row) + i;
        double fi = pn * a.f1[d];
        if (a.load.ga) fi += qa * a.load.ga[d];
        if (a.load.gb) fi += qb * a.load.gb[d];
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            const double m = a.lm_sstride ? lmp[v] : lmp[0];
            double force = fi;
            if constexpr (DAMP) force -= a.ddt * (uu.v[v] - pp.v[v]);
            const double rhs = a.dt2 * (force - acc[i][v]) + m * (2.0 * uu.v[v] - pp.v[v]);
            double un = m != 0.0 ? rhs / m : 0.0;
            o.v[v] = fixed ? a.bdy_u : un;
        }
        store_vec<VW>(a.out + off, o);
    }
}

// values[p][ij][s] = sum_c coef[celem[c]][s] * khat[c][ij]; one thread per (block, sample).
__global__ __launch_bounds__(kBlock) void k_assemble(int64_t nnzb, int64_t n_smp,
                                                      const int32_t *__restrict__ cptr,
                                                      const int32_t *__restrict__ celem,
                                                      const double *__restrict__ khat,
                                                      const double *__restrict__ coef,
                                                      double *__restrict__ values) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= nnzb * n_smp) return;
    const int64_t p = g / n_smp;
    const int64_t s = g - p * n_smp;
    double kv[9];
#pragma unroll
    for (int ij = 0; ij < 9; ++ij) kv[ij] = 0.0;
    for (int32_t c = cptr[p]; c < cptr[p + 1]; ++c) {
        const double cf = coef[static_cast<int64_t>(celem[c]) * n_smp + s];
        const double *kh = khat + 9 * static_cast<int64_t>(c);
#pragma unroll
        for (int ij = 0; ij < 9; ++ij) kv[ij] = fma(cf, kh[ij], kv[ij]);
    }
    double *out = values + p * 9 * n_smp + s;
#pragma unroll
    for (int ij = 0; ij < 9; ++ij) out[ij * n_smp] = kv[ij];
}

// x[n][j][s] = u0[3n+j] + beta * du0[3n+j]  (tile a sample-independent vector over samples)
__global__ __launch_bounds__(kBlock) void k_tile_combo(int64_t ndof, int64_t n_smp,
                                                        const double *__restrict__ u0,
                                                        const double *__restrict__ du0, double beta,
                                                        double *__restrict__ x) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= ndof * n_smp) return;
    const int64_t d = g / n_smp;
    x[g] = u0[d] + beta * du0[d];
}

// u = u0; ddu = (p0 f1 - damp du0 - Ku) / lm; up = u0 - dt du0 + dt^2/2 ddu  (physicsSolver.py:209-229)
__global__ __launch_bounds__(kBlock) void k_init_finish(int64_t ndof, int64_t n_smp,
                                                         const double *__restrict__ u0,
                                                         const double *__restrict__ du0,
                                                         const double *__restrict__ f1,
                                                         cvfes_load_t load, double damp, double dt,
                                                         const double *__restrict__ lm,
                                                         int64_t lm_sstride,
                                                         const double *__restrict__ Ku,
                                                         double *__restrict__ u,
                                                         double *__restrict__ up) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= ndof * n_smp) return;
    const int64_t d = g / n_smp;
    const int64_t s = g - d * n_smp;
    const int64_t node = d / 3;
    const double m = lm[lm_sstride ? node * n_smp + s : node];
    double f0 = load.p[0] * f1[d];
    if (load.ga) f0 += load.qa[0] * load.ga[d];
    if (load.gb) f0 += load.qb[0] * load.gb[d];
    const double rhs = f0 - damp * du0[d] - Ku[g];
    const double ddu = m != 0.0 ? rhs / m : 0.0;
    u[g] = u0[d];
    up[g] = u0[d] - dt * du0[d] + 0.5 * (dt * dt) * ddu;
}

int validate(const cvfes_system_t *sys, int method, const char *who) {
    CVFES_REQUIRE(sys != nullptr, "%s: null system", who);
    CVFES_REQUIRE(sys->n_nodes >= 1 && sys->n_smp >= 1 && sys->nnzb >= 1, "%s: empty system", who);
    CVFES_REQUIRE(sys->n_nodes * 3 < (int64_t(1) << 31), "%s: too many nodes for 32-bit rows", who);
    CVFES_REQUIRE(sys->indptr && sys->indices, "%s: null block structure", who);
    CVFES_REQUIRE(method == CVFES_METHOD_STORED || method == CVFES_METHOD_MATFREE,
                  "%s: unknown method %d", who, method);
    if (method == CVFES_METHOD_STORED)
        CVFES_REQUIRE(sys->values, "%s: stored method needs values", who);
    else
        CVFES_REQUIRE(sys->cptr && sys->celem && sys->khat && sys->coef,
                      "%s: matrix-free method needs cptr/celem/khat/coef", who);
    return CVFES_OK;
}

RowArgs row_args(const cvfes_system_t *sys) {
    return RowArgs{sys->n_nodes, sys->n_smp, sys->indptr, sys->indices, sys->values,
                   sys->cptr,    sys->celem, sys->khat,   sys->coef};
}

// Dispatch helpers ------------------------------------------------------------------------------
template <bool MF, bool UNI, int VW>
void launch_spmv(hipStream_t st, const StepArgs &a) {
    const int64_t work = a.k.n_nodes * (a.k.n_smp / VW);
    hipLaunchKernelGGL((k_spmv<MF, UNI, VW>), dim3(grid_for(work)), dim3(kBlock), 0, st, a);
}

template <bool MF, int VW>
void launch_spmv_u(hipStream_t st, const StepArgs &a) {
    if ((a.k.n_smp / VW) % 64 == 0)
        launch_spmv<MF, true, VW>(st, a);
    else
        launch_spmv<MF, false, VW>(st, a);
}

void spmv_dispatch(hipStream_t st, bool mf, const StepArgs &a) {
    const bool v2 = a.k.n_smp % 2 == 0;
    if (mf) {
        if (v2) launch_spmv_u<true, 2>(st, a); else launch_spmv_u<true, 1>(st, a);
    } else {
        if (v2) launch_spmv_u<false, 2>(st, a); else launch_spmv_u<false, 1>(st, a);
    }
}

template <bool MF, bool UNI, bool BETA, bool DAMP, int VW>
void launch_step(hipStream_t st, const StepArgs &a, int64_t step) {
    const int64_t work = a.k.n_nodes * (a.k.n_smp / VW);
    hipLaunchKernelGGL((k_step<MF, UNI, BETA, DAMP, VW>), dim3(grid_for(work)), dim3(kBlock), 0,
                       st, a, step);
}

template <bool MF, bool UNI, int VW>
void step_bd(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    if (beta) {
        if (damp) launch_step<MF, UNI, true, true, VW>(st, a, step);
        else launch_step<MF, UNI, true, false, VW>(st, a, step);
    } else {
        if (damp) launch_step<MF, UNI, false, true, VW>(st, a, step);
        else launch_step<MF, UNI, false, false, VW>(st, a, step);
    }
}

template <bool MF, int VW>
void step_u(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    if ((a.k.n_smp / VW) % 64 == 0)
        step_bd<MF, true, VW>(st, a, step, beta, damp);
    else
        step_bd<MF, false, VW>(st, a, step, beta, damp);
}

void step_dispatch(hipStream_t st, bool mf, const StepArgs &a, int64_t step, bool beta, bool damp) {
    const bool v2 = a.k.n_smp % 2 == 0;
    if (mf) {
        if (v2) step_u<true, 2>(st, a, step, beta, damp); else step_u<true, 1>(st, a, step, beta, damp);
    } else {
        if (v2) step_u<false, 2>(st, a, step, beta, damp); else step_u<false, 1>(st, a, step, beta, damp);
    }
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int cvfes_assemble_bcsr(void *stream, int64_t nnzb, int64_t n_smp, const int32_t *cptr,
                                   const int32_t *celem, const double *khat, const double *coef,
                                   double *values) {
    CVFES_REQUIRE(nnzb >= 0 && n_smp >= 1, "cvfes_assemble_bcsr: bad sizes");
    if (nnzb == 0) return CVFES_OK;
    CVFES_REQUIRE(cptr && celem && khat && coef && values, "cvfes_assemble_bcsr: null pointer");
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(nnzb * n_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), nnzb, n_smp, cptr, celem, khat, coef,
                       values);
    return check_launch("k_assemble");
}

extern "C" int cvfes_bcsr_spmv(void *stream, const cvfes_system_t *sys, int method,
                               const double *x, double *y) {
    if (int rc = validate(sys, method, "cvfes_bcsr_spmv")) return rc;
    CVFES_REQUIRE(x && y && x != y, "cvfes_bcsr_spmv: x and y must be distinct device buffers");
    StepArgs a{};
    a.k = row_args(sys);
    a.u = x;
    a.out = y;
    spmv_dispatch(static_cast<hipStream_t>(stream), method == CVFES_METHOD_MATFREE, a);
    return check_launch("k_spmv");
}

int validate_load(const cvfes_load_t *load, const char *who) {
    CVFES_REQUIRE(load && load->p, "%s: null load / pressure table", who);
    CVFES_REQUIRE(!load->ga || load->qa, "%s: traction field A without its table", who);
    CVFES_REQUIRE(!load->gb || load->qb, "%s: traction field B without its table", who);
    return CVFES_OK;
}

extern "C" int cvfes_initial_state(void *stream, const cvfes_system_t *sys, int method,
                                   const double *u0, const double *du0, const cvfes_load_t *load,
                                   double dt, double beta, double damp, double *u, double *up,
                                   double *work) {
    if (int rc = validate(sys, method, "cvfes_initial_state")) return rc;
    if (int rc = validate_load(load, "cvfes_initial_state")) return rc;
    CVFES_REQUIRE(u0 && du0 && u && up && work && sys->lm && sys->f1,
                  "cvfes_initial_state: null pointer");
    CVFES_REQUIRE(dt > 0.0, "cvfes_initial_state: dt must be positive");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t ndof = 3 * sys->n_nodes, S = sys->n_smp;
    hipLaunchKernelGGL(k_tile_combo, dim3(grid_for(ndof * S)), dim3(kBlock), 0, st, ndof, S, u0,
                       du0, beta, u);
    if (int rc = check_launch("k_tile_combo")) return rc;
    StepArgs a{};
    a.k = row_args(sys);
    a.u = u;
    a.out = work;
    spmv_dispatch(st, method == CVFES_METHOD_MATFREE, a);
    if (int rc = check_launch("k_spmv")) return rc;
    hipLaunchKernelGGL(k_init_finish, dim3(grid_for(ndof * S)), dim3(kBlock), 0, st, ndof, S, u0,
                       du0, sys->f1, *load, damp, dt, sys->lm, sys->lm_sstride, work, u, up);
    return check_launch("k_init_finish");
}

extern "C" int cvfes_step(void *stream, const cvfes_system_t *sys, int method, double **state,
                          const cvfes_load_t *load, int64_t nsteps, double dt, double beta,
                          double damp) {
    if (int rc = validate(sys, method, "cvfes_step")) return rc;
    if (int rc = validate_load(load, "cvfes_step")) return rc;
    CVFES_REQUIRE(state && state[0] && state[1] && state[2], "cvfes_step: null state buffer");
    CVFES_REQUIRE(state[0] != state[1] && state[1] != state[2] && state[0] != state[2],
                  "cvfes_step: state buffers must be distinct");
    CVFES_REQUIRE(sys->lm && sys->f1, "cvfes_step: null lm/f1");
    CVFES_REQUIRE(nsteps >= 0, "cvfes_step: nsteps < 0");
    CVFES_REQUIRE(dt > 0.0, "cvfes_step: dt must be positive");
    CVFES_REQUIRE(sys->lm_sstride == 0 || sys->lm_sstride == 1, "cvfes_step: lm_sstride must be 0 or 1");
    hipStream_t st = static_cast<hipStream_t>(stream);
    StepArgs a{};
    a.k = row_args(sys);
    a.lm = sys->lm;
    a.lm_sstride = sys->lm_sstride;
    a.f1 = sys->f1;
    a.fixed = sys->fixed;
    a.load = *load;
    a.bdy_u = sys->bdy_u;
    a.dt2 = dt * dt;
    a.bdt = beta / dt;
    a.ddt = damp / dt;
    const bool mf = method == CVFES_METHOD_MATFREE;
    double *cur = state[0], *prev = state[1], *nxt = state[2];
    for (int64_t n = 0; n < nsteps; ++n) {
        a.u = cur;
        a.up = prev;
        a.out = nxt;
        step_dispatch(st, mf, a, n, beta != 0.0, damp != 0.0);
        if (int rc = check_launch("k_step")) return rc;
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
