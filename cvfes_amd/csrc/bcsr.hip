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
//   METHOD_MATFREE: each block applied as sum_c coef[celem[c], s] * (khat[c] x) over its element
//                   contributions (sample-independent khat, scalar loads; coef [E][S] is cached
//                   across the 3 rows that touch an element), or, in the EBE variant, element by
//                   element over element colours.
#include <mutex>

#include "rowops.h"

namespace cvfes {
namespace {

template <int OP, bool UNIFORM, int VW>
__global__ __launch_bounds__(kBlock) void k_spmv(StepArgs a) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    double acc[3][VW];
    row_product<OP, false, VW>(a.k, a.u, static_cast<const double *>(nullptr), 0.0, row, s, acc);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) o.v[v] = acc[i][v];
        if (active) store_vec<VW>(a.out + (static_cast<int64_t>(row) * 3 + i) * a.k.n_smp + s, o);
    }
}

template <int OP, bool UNIFORM, bool BETA, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_step(StepArgs a, int64_t step) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    if constexpr (UNIFORM) {
        if (a.order) row = a.order[row];
    }
    step_row<OP, BETA, DAMP, VW>(a, row, s, active, step);
}

// ---------------------------------------------------------------------------------------------
// Row step with every load of a block row issued up front (wave-uniform rows, two samples per lane).
//
// The generic k_step walks a row block by block, and for the matrix-free operator each element
// contribution is a dependent chain (scalar load of the element id -> vector load of its
// coefficient -> wait), so a wave has only a few loads in flight.  Here all addresses of a row are
// formed first and every load is issued before any arithmetic, without branches between them:
// block slots b < MAXB are clamped to the row's last block and masked, element slots are clamped
// to the block's last contribution and masked.  Clamping keeps every access inside the row's own
// blocks and contributions.  Limits (exact maxima over the mesh, computed on the host):
//   blocks per row <= MAXB, contributions per off-diagonal block <= 2,
//   contributions per diagonal block <= MAXB - 1 (a node has at least as many neighbours as
//   incident triangles in a manifold triangulation).
template <bool MATFREE, bool BETA, bool DAMP, bool NT, int MAXB, int VW>
__global__ __launch_bounds__(kBlock) void k_step_rowload(StepArgs a, int64_t step) {
    constexpr int MAXD = MAXB - 1;
    typedef typename VecT<VW>::type V;
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<true, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    if (a.order) row = a.order[row];
    const RowArgs &k = a.k;
    const int64_t S = k.n_smp;
    const int32_t p0 = k.indptr[row];
    const int32_t nb = k.indptr[row + 1] - p0;

    // own state (also needed by the update)
    V uo[3], po[3], wo[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * S + s;
        uo[i] = ldv<VW, false>(a.u + off);
        po[i] = ldv<VW, false>(a.up + off);
        wo[i] = BETA ? uo[i] + a.bdt * (uo[i] - po[i]) : uo[i];
    }
    V acc[3] = {splat<VW>(0.0), splat<VW>(0.0), splat<VW>(0.0)};
    if (nb > 0) {
        // neighbour states for every block slot
        int32_t col[MAXB];
        V w[MAXB][3];
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int32_t pb = p0 + (b < nb ? b : nb - 1);
            col[b] = k.indices[pb];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int64_t off = (static_cast<int64_t>(col[b]) * 3 + j) * S + s;
                const V xu = ldv<VW, false>(a.u + off);
                if constexpr (BETA) {
                    const V xp = ldv<VW, false>(a.up + off);
                    w[b][j] = xu + a.bdt * (xu - xp);
                } else {
                    w[b][j] = xu;
                }
            }
        }
        if constexpr (!MATFREE) {
#pragma unroll
            for (int b = 0; b < MAXB; ++b) {
                const bool valid = b < nb;
                const int32_t pb = p0 + (valid ? b : nb - 1);
                const double *base = k.values + static_cast<int64_t>(pb) * 9 * S + s;
                V kv[9];
#pragma unroll
                for (int ij = 0; ij < 9; ++ij) kv[ij] = ldv<VW, NT>(base + ij * S);
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) acc[i] += sel<VW>(valid, kv[3 * i + j]) * w[b][j];
            }
        } else {
            // diagonal block: the contributions of every element incident to the row node
            int32_t pd = p0;
#pragma unroll
            for (int b = 0; b < MAXB; ++b) pd += (b < nb && col[b] < row) ? 1 : 0;
            const int32_t cd0 = k.cptr[pd];
            const int32_t nd = k.cptr[pd + 1] - cd0;
            V cfd[MAXD];
#pragma unroll
            for (int d = 0; d < MAXD; ++d) {
                const int32_t c = cd0 + (d < nd ? d : nd - 1);
                cfd[d] = sel<VW>(d < nd, ldv<VW, false>(k.coef + static_cast<int64_t>(k.celem[c]) * S + s));
            }
            // off-diagonal blocks: one or two contributions each
            V cfo[MAXB][2];
            int32_t co[MAXB], no[MAXB];
#pragma unroll
            for (int b = 0; b < MAXB; ++b) {
                const bool valid = b < nb && col[b] != row;
                const int32_t pb = p0 + (b < nb ? b : nb - 1);
                co[b] = k.cptr[pb];
                no[b] = valid ? k.cptr[pb + 1] - co[b] : 0;
                const int32_t c1 = co[b] + (no[b] > 1 ? 1 : 0);
                cfo[b][0] = sel<VW>(no[b] > 0, ldv<VW, false>(k.coef + static_cast<int64_t>(k.celem[co[b]]) * S + s));
                cfo[b][1] = sel<VW>(no[b] > 1, ldv<VW, false>(k.coef + static_cast<int64_t>(k.celem[c1]) * S + s));
            }
            {
                V kv[9];
#pragma unroll
                for (int ij = 0; ij < 9; ++ij) kv[ij] = splat<VW>(0.0);
#pragma unroll
                for (int d = 0; d < MAXD; ++d) {
                    const double *kh = k.khat + 9 * static_cast<int64_t>(cd0 + (d < nd ? d : nd - 1));
#pragma unroll
                    for (int ij = 0; ij < 9; ++ij) kv[ij] += cfd[d] * kh[ij];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) acc[i] += kv[3 * i + j] * wo[j];
            }
#pragma unroll
            for (int b = 0; b < MAXB; ++b) {
                const double *kh0 = k.khat + 9 * static_cast<int64_t>(co[b]);
                const double *kh1 = k.khat + 9 * static_cast<int64_t>(co[b] + (no[b] > 1 ? 1 : 0));
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        acc[i] += (cfo[b][0] * kh0[3 * i + j] + cfo[b][1] * kh1[3 * i + j]) * w[b][j];
            }
        }
    }
    finish_row<DAMP, VW, NT>(a, row, s, active, step, uo, po, acc);   // NT variants stream u_{n+1}
}

// Stored step on the symmetric storage: the row's blocks are read from the upper entries
// (uvalues[vblk[p]]), transposed for the lower ones (col < row).  Every upper block is read twice per
// step, once by its own row and once, transposed, by its column's row, 1 and ntheta rows later in
// the XCD-ordered sweep (or one strip width later in strip order), so the second read is served by
// L2 and the HBM value stream is about halved.  All loads are buffer loads (one per-lane offset
// register for the whole row, resources in scalar registers): 124 VGPRs, four waves per SIMD.
// Value loads: NT = 0 plain (the fastest), 1 all non-temporal, 2 non-temporal for the mirrored
// (second) reads only.  NT = 2 has two load paths per block, which the compiler hoists into one
// wave per SIMD holding the whole row's loads.
// REC: the row's structure comes from its record (`rec`, see cvfes_system_t.row_records): one
// load chain shorter than row order -> indptr -> indices/vblk.
// Resources: states 3*S*8 bytes per node, values 9*S*8 bytes per block (resolve_variant keeps
// 72*S below 2^31).
template <bool BETA, bool DAMP, int NT, int MAXB, int VW, bool REC>
__global__ __launch_bounds__(kBlock) void k_step_sym(StepArgs a, const double *__restrict__ uvalues,
                                                     const int32_t *__restrict__ vblk,
                                                     const int32_t *__restrict__ rec, int64_t step) {
    typedef typename VecT<VW>::type V;
    constexpr int RECW = 2 + 2 * MAXB;
    constexpr int AUX_FIRST = NT == 1 ? kAuxNT : 0;
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<true, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    const int32_t *r = nullptr;
    if constexpr (REC) {
        r = rec + static_cast<int64_t>(row) * RECW;
        row = r[0];
    } else if (a.order) {
        row = a.order[row];
    }
    const RowArgs &k = a.k;
    const int64_t S = k.n_smp;
    const int32_t p0 = REC ? 0 : k.indptr[row];
    const int32_t nb = REC ? r[1] : k.indptr[row + 1] - p0;
    const uint32_t lo = static_cast<uint32_t>(s) * 8u;   // the lane's byte offset in every row
    const uint32_t sb = static_cast<uint32_t>(S) * 8u;   // bytes of one entry across the samples
    V uo[3], po[3];
    {
        const auto ru = buffer_rsrc(a.u + static_cast<int64_t>(row) * 3 * S, 3 * sb);
        const auto rp = buffer_rsrc(a.up + static_cast<int64_t>(row) * 3 * S, 3 * sb);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            uo[i] = ldb<VW, 0>(ru, lo, i * sb);
            po[i] = ldb<VW, 0>(rp, lo, i * sb);
        }
    }
    V acc[3] = {splat<VW>(0.0), splat<VW>(0.0), splat<VW>(0.0)};
    if (nb > 0) {
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const bool valid = b < nb;
            const int32_t pb = p0 + (valid ? b : nb - 1);
            const int32_t col = REC ? r[2 + b] : k.indices[pb];
            const bool tr = col < row;
            V w[3];
            const auto cu = buffer_rsrc(a.u + static_cast<int64_t>(col) * 3 * S, 3 * sb);
            const auto cp = buffer_rsrc(a.up + static_cast<int64_t>(col) * 3 * S, 3 * sb);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const V xu = ldb<VW, 0>(cu, lo, j * sb);
                if constexpr (BETA) {
                    const V xp = ldb<VW, 0>(cp, lo, j * sb);
                    w[j] = xu + a.bdt * (xu - xp);
                } else {
                    w[j] = xu;
                }
            }
            const int64_t ub = REC ? r[2 + MAXB + b] : vblk[pb];
            const auto rv = buffer_rsrc(uvalues + ub * 9 * S, 9 * sb);
            V kv[9];
            if (NT == 2 && tr) {
#pragma unroll
                for (int ij = 0; ij < 9; ++ij) kv[ij] = ldb<VW, kAuxNT>(rv, lo, (3 * (ij % 3) + ij / 3) * sb);
            } else {
#pragma unroll
                for (int ij = 0; ij < 9; ++ij)
                    kv[ij] = ldb<VW, AUX_FIRST>(rv, lo, (tr ? 3 * (ij % 3) + ij / 3 : ij) * sb);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc[i] += sel<VW>(valid, kv[3 * i + j]) * w[j];
        }
    }
    finish_row<DAMP, VW, NT != 0>(a, row, s, active, step, uo, po, acc);
}

// Matrix-free step with the row's coefficients loaded once per row (step_row_cached); one
// wavefront per (row, 64*VW samples), so rows must be wave-uniform.
template <bool BETA, bool DAMP, int VW, int MAXE>
__global__ __launch_bounds__(kBlock) void k_step_cached(StepArgs a, LdsArgs l, int64_t step) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<true, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    if (a.order) row = a.order[row];
    step_row_cached<BETA, DAMP, VW, MAXE>(a, l, row, s, active, step);
}

// Matrix-free step with each row's element data staged in LDS (step_row_lds); one wavefront per
// (row, 64*VW samples), so rows must be wave-uniform.
template <bool BETA, bool DAMP, int VW, int MAXB, int MAXE>
__global__ __launch_bounds__(kBlock) void k_step_lds(StepArgs a, LdsArgs l, int64_t step) {
    __shared__ LdsRow<VW, MAXE> s_row[kBlock / 64];
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<true, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    step_row_lds<BETA, DAMP, VW, MAXB, MAXE>(a, l, s_row[threadIdx.x / 64], row, s, active, step);
}

// values[p][ij][s] = sum_c coef[celem[c]][s] * khat[c][ij]; one thread per (block, sample).
__global__ __launch_bounds__(kBlock) void k_assemble(int64_t nnzb, int64_t n_smp,
                                                      const int32_t *__restrict__ bmap,
                                                      const int32_t *__restrict__ cptr,
                                                      const int32_t *__restrict__ celem,
                                                      const double *__restrict__ khat,
                                                      const double *__restrict__ coef,
                                                      double *__restrict__ values) {
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= nnzb * n_smp) return;
    const int64_t p = g / n_smp;                  // output block
    const int64_t s = g - p * n_smp;
    const int64_t b = bmap ? bmap[p] : p;         // block of the full structure
    double kv[9];
#pragma unroll
    for (int ij = 0; ij < 9; ++ij) kv[ij] = 0.0;
    for (int32_t c = cptr[b]; c < cptr[b + 1]; ++c) {
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
        CVFES_REQUIRE(sys->values || (sys->uvalues && sys->vblk),
                      "%s: stored method needs values (full, or symmetric: uvalues and vblk)", who);
    else
        CVFES_REQUIRE(sys->cptr && sys->celem && sys->khat && sys->coef,
                      "%s: matrix-free method needs cptr/celem/khat/coef", who);
    CVFES_REQUIRE(sys->waves_per_simd == 0 || (sys->waves_per_simd >= 2 && sys->waves_per_simd <= 7),
                  "%s: waves_per_simd must be 0 or 2..7", who);
    return CVFES_OK;
}

RowArgs row_args(const cvfes_system_t *sys) {
    return RowArgs{sys->n_nodes, sys->n_smp, sys->indptr, sys->indices, sys->values,
                   sys->cptr,    sys->celem, sys->khat,   sys->coef, sys->uvalues, sys->vblk};
}

// Dispatch helpers ------------------------------------------------------------------------------
// Wave-uniform rows unless padding to whole wavefronts would idle more than a quarter of the lanes.

template <int OP, bool UNI, int VW>
void launch_spmv(hipStream_t st, const StepArgs &a) {
    const int64_t work = a.k.n_nodes * lanes_per_row<UNI, VW>(a.k.n_smp);
    hipLaunchKernelGGL((k_spmv<OP, UNI, VW>), dim3(grid_for(work)), dim3(kBlock), a.occ_lds, st, a);
}

template <int OP, int VW>
void launch_spmv_u(hipStream_t st, const StepArgs &a) {
    if (use_uniform<VW>(a.k.n_smp))
        launch_spmv<OP, true, VW>(st, a);
    else
        launch_spmv<OP, false, VW>(st, a);
}

template <int OP>
void spmv_op(hipStream_t st, const StepArgs &a) {
    if (generic_vw(a.k.n_smp) == 2) launch_spmv_u<OP, 2>(st, a);
    else launch_spmv_u<OP, 1>(st, a);
}

// Operator form of the generic row kernels: matrix-free, or the stored values in the storage
// the system holds (full values when present, else the symmetric upper blocks).
int row_op(const RowArgs &k, bool mf) {
    return mf ? kOpMatfree : (k.values ? kOpStored : kOpSym);
}

void spmv_dispatch(hipStream_t st, bool mf, const StepArgs &a) {
    switch (row_op(a.k, mf)) {
    case kOpMatfree: spmv_op<kOpMatfree>(st, a); break;
    case kOpStored: spmv_op<kOpStored>(st, a); break;
    default: spmv_op<kOpSym>(st, a); break;
    }
}

template <int OP, bool UNI, bool BETA, bool DAMP, int VW>
void launch_step(hipStream_t st, const StepArgs &a, int64_t step) {
    const int64_t work = a.k.n_nodes * lanes_per_row<UNI, VW>(a.k.n_smp);
    hipLaunchKernelGGL((k_step<OP, UNI, BETA, DAMP, VW>), dim3(grid_for(work)), dim3(kBlock), a.occ_lds,
                       st, a, step);
}

template <int OP, bool UNI, int VW>
void step_bd(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    if (beta) {
        if (damp) launch_step<OP, UNI, true, true, VW>(st, a, step);
        else launch_step<OP, UNI, true, false, VW>(st, a, step);
    } else {
        if (damp) launch_step<OP, UNI, false, true, VW>(st, a, step);
        else launch_step<OP, UNI, false, false, VW>(st, a, step);
    }
}

template <int OP, int VW>
void step_u(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    if (use_uniform<VW>(a.k.n_smp))
        step_bd<OP, true, VW>(st, a, step, beta, damp);
    else
        step_bd<OP, false, VW>(st, a, step, beta, damp);
}

template <int OP>
void step_generic(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    if (generic_vw(a.k.n_smp) == 2) step_u<OP, 2>(st, a, step, beta, damp);
    else step_u<OP, 1>(st, a, step, beta, damp);
}

template <bool MF, bool NT, int MAXB, int VW>
void launch_rowload(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp) {
    const int64_t work = a.k.n_nodes * lanes_per_row<true, VW>(a.k.n_smp);
    const dim3 g(grid_for(work)), b(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_step_rowload<MF, true, true, NT, MAXB, VW>), g, b, a.occ_lds, st, a, step);
        else hipLaunchKernelGGL((k_step_rowload<MF, true, false, NT, MAXB, VW>), g, b, a.occ_lds, st, a, step);
    } else {
        if (damp) hipLaunchKernelGGL((k_step_rowload<MF, false, true, NT, MAXB, VW>), g, b, a.occ_lds, st, a, step);
        else hipLaunchKernelGGL((k_step_rowload<MF, false, false, NT, MAXB, VW>), g, b, a.occ_lds, st, a, step);
    }
}

template <bool MF, bool NT, int VW>
void rowload_maxb(hipStream_t st, const StepArgs &a, int64_t step, bool beta, bool damp, int maxb) {
    if (maxb == 7) launch_rowload<MF, NT, 7, VW>(st, a, step, beta, damp);
    else launch_rowload<MF, NT, 10, VW>(st, a, step, beta, damp);
}

template <bool BETA, bool DAMP, int VW, int MAXB, int MAXE>
void launch_lds_inst(hipStream_t st, const StepArgs &a, const LdsArgs &l, int64_t step) {
    const int64_t work = a.k.n_nodes * lanes_per_row<true, VW>(a.k.n_smp);
    hipLaunchKernelGGL((k_step_lds<BETA, DAMP, VW, MAXB, MAXE>), dim3(grid_for(work)), dim3(kBlock),
                       0, st, a, l, step);
}

template <int VW, int MAXB, int MAXE>
void launch_lds_bd(hipStream_t st, const StepArgs &a, const LdsArgs &l, int64_t step, bool beta,
                   bool damp) {
    if (beta) {
        if (damp) launch_lds_inst<true, true, VW, MAXB, MAXE>(st, a, l, step);
        else launch_lds_inst<true, false, VW, MAXB, MAXE>(st, a, l, step);
    } else {
        if (damp) launch_lds_inst<false, true, VW, MAXB, MAXE>(st, a, l, step);
        else launch_lds_inst<false, false, VW, MAXB, MAXE>(st, a, l, step);
    }
}

template <int NT, int MAXB, bool REC>
void launch_sym_rec(hipStream_t st, const cvfes_system_t *sys, const StepArgs &a, int64_t step,
                    bool beta, bool damp) {
    const int64_t work = a.k.n_nodes * lanes_per_row<true, 2>(a.k.n_smp);
    const dim3 grid(grid_for(work)), block(kBlock);
    const double *uv = sys->uvalues;
    const int32_t *vb = sys->vblk;
    const int32_t *rc = sys->row_records;
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_step_sym<true, true, NT, MAXB, 2, REC>), grid, block, a.occ_lds, st, a, uv, vb, rc, step);
        else hipLaunchKernelGGL((k_step_sym<true, false, NT, MAXB, 2, REC>), grid, block, a.occ_lds, st, a, uv, vb, rc, step);
    } else {
        if (damp) hipLaunchKernelGGL((k_step_sym<false, true, NT, MAXB, 2, REC>), grid, block, a.occ_lds, st, a, uv, vb, rc, step);
        else hipLaunchKernelGGL((k_step_sym<false, false, NT, MAXB, 2, REC>), grid, block, a.occ_lds, st, a, uv, vb, rc, step);
    }
}

template <int NT, int MAXB>
void launch_sym(hipStream_t st, const cvfes_system_t *sys, const StepArgs &a, int64_t step, bool beta,
                bool damp) {
    if (sys->row_records && sys->record_width == 2 + 2 * MAXB)
        launch_sym_rec<NT, MAXB, true>(st, sys, a, step, beta, damp);
    else
        launch_sym_rec<NT, MAXB, false>(st, sys, a, step, beta, damp);
}

template <int VW, int MAXE>
void launch_cached(hipStream_t st, const StepArgs &a, const LdsArgs &l, int64_t step, bool beta,
                   bool damp) {
    const int64_t work = a.k.n_nodes * lanes_per_row<true, VW>(a.k.n_smp);
    const dim3 grid(grid_for(work)), block(kBlock);
    if (beta) {
        if (damp) hipLaunchKernelGGL((k_step_cached<true, true, VW, MAXE>), grid, block, 0, st, a, l, step);
        else hipLaunchKernelGGL((k_step_cached<true, false, VW, MAXE>), grid, block, 0, st, a, l, step);
    } else {
        if (damp) hipLaunchKernelGGL((k_step_cached<false, true, VW, MAXE>), grid, block, 0, st, a, l, step);
        else hipLaunchKernelGGL((k_step_cached<false, false, VW, MAXE>), grid, block, 0, st, a, l, step);
    }
}

// LDS slots needed by the structure: 1 = (MAXB 7, MAXE 6), 2 = (MAXB 10, MAXE 9), 0 = too large.
int lds_slots(const cvfes_system_t *sys) {
    if (!sys->nptr || !sys->nent || !sys->cslot || sys->max_node_elems < 1 || sys->max_row_blocks < 1)
        return 0;
    if (sys->max_row_blocks <= 7 && sys->max_node_elems <= 6) return 1;
    if (sys->max_row_blocks <= 10 && sys->max_node_elems <= 9) return 2;
    return 0;
}

void launch_lds(hipStream_t st, const cvfes_system_t *sys, const StepArgs &a, int64_t step,
                bool beta, bool damp, bool two) {
    const LdsArgs l{sys->nptr, sys->nent, sys->cslot};
    const int slots = lds_slots(sys);
    if (two) {
        if (slots == 1) launch_lds_bd<2, 7, 6>(st, a, l, step, beta, damp);
        else launch_lds_bd<2, 10, 9>(st, a, l, step, beta, damp);
    } else {
        if (slots == 1) launch_lds_bd<1, 7, 6>(st, a, l, step, beta, damp);
        else launch_lds_bd<1, 10, 9>(st, a, l, step, beta, damp);
    }
}

// Row-load kernels apply when rows are wave-uniform (S % 128 == 0) and the structure fits their
// slots; returns the MAXB instance to use, or 0.
int rowload_slots(const cvfes_system_t *sys) {
    if (sys->n_smp % 128 != 0) return 0;
    if (sys->max_row_blocks < 1 || sys->max_offdiag_contrib < 1 || sys->max_diag_contrib < 1) return 0;
    if (sys->max_offdiag_contrib > 2) return 0;
    for (int maxb : {7, 10})
        if (sys->max_row_blocks <= maxb && sys->max_diag_contrib <= maxb - 1) return maxb;
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Element-by-element matrix-free variant (EBE): K x as a sum of element products,
// y_a += coef_e sum_b khat_e(a,b) x_b, one launch per element colour (no two elements of a colour
// share a node, so the read-modify-writes of y do not race), then the per-row update from y.  y
// lives in the output buffer: it is zeroed first, and the update reads it and overwrites it row
// by row.  Every node's y is read and written once per incident element, so this moves several
// times the bytes of the row gather; it is kept as the measured alternative.

__global__ __launch_bounds__(kBlock) void k_ebe_zero(double *__restrict__ y, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

// One thread per (element of colour c, VW samples); lanes map to samples as for block rows.
template <bool UNIFORM, bool BETA, int VW>
__global__ __launch_bounds__(kBlock) void k_ebe_colour(StepArgs a, const int32_t *__restrict__ color_ptr,
                                                      const int32_t *__restrict__ color_elems,
                                                      const int32_t *__restrict__ elems,
                                                      const int32_t *__restrict__ cpos, int c) {
    const int32_t c0 = color_ptr[c], c1 = color_ptr[c + 1];
    int32_t i;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(c1 - c0, a.k.n_smp, i, s, active)) return;
    const int64_t S = a.k.n_smp;
    const int64_t e = color_elems[c0 + i];
    int64_t nd[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) nd[q] = elems[3 * e + q];
    double x[3][3][VW];
#pragma unroll
    for (int b = 0; b < 3; ++b) state_column<BETA, VW>(a.u, a.up, a.bdt, nd[b], S, s, x[b]);
    const Vec<VW> cf = load_vec<VW>(a.k.coef + e * S + s);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double t[3][VW];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int v = 0; v < VW; ++v) t[r][v] = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double *kh = a.k.khat + 9 * static_cast<int64_t>(cpos[9 * e + 3 * q + b]);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double h0 = kh[3 * r], h1 = kh[3 * r + 1], h2 = kh[3 * r + 2];
#pragma unroll
                for (int v = 0; v < VW; ++v)
                    t[r][v] = fma(h2, x[b][2][v], fma(h1, x[b][1][v], fma(h0, x[b][0][v], t[r][v])));
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int64_t off = (nd[q] * 3 + r) * S + s;
            Vec<VW> y = load_vec<VW>(a.out + off);
#pragma unroll
            for (int v = 0; v < VW; ++v) y.v[v] = fma(cf.v[v], t[r][v], y.v[v]);
            if (active) store_vec<VW>(a.out + off, y);
        }
    }
}

template <bool UNIFORM, bool DAMP, int VW>
__global__ __launch_bounds__(kBlock) void k_ebe_finish(StepArgs a, int64_t step) {
    int32_t row;
    int64_t s;
    bool active;
    if (!row_and_sample<UNIFORM, VW>(a.k.n_nodes, a.k.n_smp, row, s, active)) return;
    double acc[3][VW];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const Vec<VW> y = load_vec<VW>(a.out + (static_cast<int64_t>(row) * 3 + i) * a.k.n_smp + s);
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[i][v] = y.v[v];
    }
    step_update<DAMP, VW>(a, row, s, active, step, acc);
}

template <bool UNI, int VW>
void ebe_inst(hipStream_t st, const cvfes_system_t *sys, const StepArgs &a, int64_t step, bool beta,
              bool damp) {
    const int64_t S = a.k.n_smp;
    const int64_t n = a.k.n_nodes * 3 * S;
    hipLaunchKernelGGL(k_ebe_zero, dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, st, a.out, n);
    const dim3 gc(grid_for(sys->max_color_elems * lanes_per_row<UNI, VW>(S))), b(kBlock);
    for (int c = 0; c < static_cast<int>(sys->n_colors); ++c) {
        if (beta)
            hipLaunchKernelGGL((k_ebe_colour<UNI, true, VW>), gc, b, 0, st, a, sys->color_ptr,
                               sys->color_elems, sys->elems, sys->cpos, c);
        else
            hipLaunchKernelGGL((k_ebe_colour<UNI, false, VW>), gc, b, 0, st, a, sys->color_ptr,
                               sys->color_elems, sys->elems, sys->cpos, c);
    }
    const dim3 gf(grid_for(a.k.n_nodes * lanes_per_row<UNI, VW>(S)));
    if (damp) hipLaunchKernelGGL((k_ebe_finish<UNI, true, VW>), gf, b, 0, st, a, step);
    else hipLaunchKernelGGL((k_ebe_finish<UNI, false, VW>), gf, b, 0, st, a, step);
}

void launch_ebe(hipStream_t st, const cvfes_system_t *sys, const StepArgs &a, int64_t step, bool beta,
                bool damp) {
    const int64_t S = a.k.n_smp;
    if (S % 2 == 0) {
        if (use_uniform<2>(S)) ebe_inst<true, 2>(st, sys, a, step, beta, damp);
        else ebe_inst<false, 2>(st, sys, a, step, beta, damp);
    } else {
        if (use_uniform<1>(S)) ebe_inst<true, 1>(st, sys, a, step, beta, damp);
        else ebe_inst<false, 1>(st, sys, a, step, beta, damp);
    }
}

bool ebe_ready(const cvfes_system_t *sys) {
    return sys->color_ptr && sys->color_elems && sys->elems && sys->cpos && sys->n_colors > 0 &&
           sys->max_color_elems > 0 && sys->max_color_elems <= sys->n_elms;
}

// The kernel cvfes_step launches for this system and operator: AUTO, and the variants that the
// structure, the sample count or the operator does not support, resolved to what actually runs.
int64_t resolve_variant(const cvfes_system_t *sys, bool mf) {
    const int maxb = rowload_slots(sys);
    const int64_t S = sys->n_smp;
    int64_t v = sys->variant;
    // Stored values in symmetric storage only (no full copy): the symmetric kernel, or the generic
    // one, which reads either storage.
    const bool sym_only = !mf && !sys->values && sys->uvalues && sys->vblk;
    // Measured on MI355X (102,400-node tube, 128 samples): the stored operator is fastest with the
    // symmetric kernel and plain loads (the mirrored reads must stay in L2), or, in full storage,
    // with the row-load kernel and non-temporal value loads; the matrix-free operator with the
    // generic kernel (the row-load form needs too many registers for enough waves in flight).
    // The symmetric kernel's buffer resources hold 72*S bytes (one block across the samples).
    const bool sym_ok = maxb && sys->uvalues && sys->vblk && 72 * S <= INT32_MAX;
    const int64_t fallback = mf ? CVFES_VARIANT_GENERIC
                                : (sym_only ? (sym_ok ? CVFES_VARIANT_SYM : CVFES_VARIANT_GENERIC)
                                            : CVFES_VARIANT_ROWLOAD_NT);
    if (v == CVFES_VARIANT_AUTO) v = fallback;
    if (sym_only && v != CVFES_VARIANT_SYM && v != CVFES_VARIANT_SYM_NT && v != CVFES_VARIANT_SYM_NTL)
        v = v == CVFES_VARIANT_GENERIC ? v : fallback;   // the full-storage kernels need `values`
    if (v == CVFES_VARIANT_LDS || v == CVFES_VARIANT_LDS2) {
        // LDS staging only exists for the matrix-free operator (and structures within its limits)
        if (mf && lds_slots(sys))
            return (v == CVFES_VARIANT_LDS2 && S % 2 == 0) ? CVFES_VARIANT_LDS2 : CVFES_VARIANT_LDS;
        v = fallback;
    }
    if (v == CVFES_VARIANT_CACHED || v == CVFES_VARIANT_CACHED1) {
        if (mf && lds_slots(sys) && use_uniform<2>(S) && S % 2 == 0) return v;
        v = fallback;
    }
    if (v == CVFES_VARIANT_EBE) {
        if (mf && ebe_ready(sys)) return v;
        v = fallback;
    }
    if (v == CVFES_VARIANT_SYM || v == CVFES_VARIANT_SYM_NT || v == CVFES_VARIANT_SYM_NTL) {
        if (!mf && sym_ok) return v;
        v = fallback;
    }
    if (maxb && v != CVFES_VARIANT_GENERIC) {
        const bool one = v == CVFES_VARIANT_ROWLOAD1 || v == CVFES_VARIANT_ROWLOAD1_NT;
        if (mf) return one ? CVFES_VARIANT_ROWLOAD1 : CVFES_VARIANT_ROWLOAD;   // no non-temporal form
        return v;
    }
    return CVFES_VARIANT_GENERIC;
}

void step_dispatch(hipStream_t st, const cvfes_system_t *sys, bool mf, const StepArgs &a,
                   int64_t step, bool beta, bool damp) {
    const int maxb = rowload_slots(sys);
    const int64_t v = resolve_variant(sys, mf);
    switch (v) {
    case CVFES_VARIANT_EBE:
        launch_ebe(st, sys, a, step, beta, damp);
        return;
    case CVFES_VARIANT_LDS:
    case CVFES_VARIANT_LDS2:
        launch_lds(st, sys, a, step, beta, damp, v == CVFES_VARIANT_LDS2);
        return;
    case CVFES_VARIANT_CACHED:
    case CVFES_VARIANT_CACHED1: {
        const LdsArgs l{sys->nptr, sys->nent, sys->cslot};
        const bool six = lds_slots(sys) == 1;
        if (v == CVFES_VARIANT_CACHED) {
            if (six) launch_cached<2, 6>(st, a, l, step, beta, damp);
            else launch_cached<2, 9>(st, a, l, step, beta, damp);
        } else {
            if (six) launch_cached<1, 6>(st, a, l, step, beta, damp);
            else launch_cached<1, 9>(st, a, l, step, beta, damp);
        }
        return;
    }
    case CVFES_VARIANT_SYM:
    case CVFES_VARIANT_SYM_NT:
    case CVFES_VARIANT_SYM_NTL: {
        const int nt = v == CVFES_VARIANT_SYM ? 0 : (v == CVFES_VARIANT_SYM_NT ? 1 : 2);
        if (maxb == 7) {
            if (nt == 0) launch_sym<0, 7>(st, sys, a, step, beta, damp);
            else if (nt == 1) launch_sym<1, 7>(st, sys, a, step, beta, damp);
            else launch_sym<2, 7>(st, sys, a, step, beta, damp);
        } else {
            if (nt == 0) launch_sym<0, 10>(st, sys, a, step, beta, damp);
            else if (nt == 1) launch_sym<1, 10>(st, sys, a, step, beta, damp);
            else launch_sym<2, 10>(st, sys, a, step, beta, damp);
        }
        return;
    }
    case CVFES_VARIANT_ROWLOAD:
    case CVFES_VARIANT_ROWLOAD_NT:
    case CVFES_VARIANT_ROWLOAD1:
    case CVFES_VARIANT_ROWLOAD1_NT: {
        const bool one = v == CVFES_VARIANT_ROWLOAD1 || v == CVFES_VARIANT_ROWLOAD1_NT;
        const bool nt = v == CVFES_VARIANT_ROWLOAD_NT || v == CVFES_VARIANT_ROWLOAD1_NT;
        if (mf) {
            if (one) rowload_maxb<true, false, 1>(st, a, step, beta, damp, maxb);
            else rowload_maxb<true, false, 2>(st, a, step, beta, damp, maxb);
        } else if (one) {
            if (nt) rowload_maxb<false, true, 1>(st, a, step, beta, damp, maxb);
            else rowload_maxb<false, false, 1>(st, a, step, beta, damp, maxb);
        } else {
            if (nt) rowload_maxb<false, true, 2>(st, a, step, beta, damp, maxb);
            else rowload_maxb<false, false, 2>(st, a, step, beta, damp, maxb);
        }
        return;
    }
    default:
        break;
    }
    switch (row_op(a.k, mf)) {
    case kOpMatfree: step_generic<kOpMatfree>(st, a, step, beta, damp); break;
    case kOpStored: step_generic<kOpStored>(st, a, step, beta, damp); break;
    default: step_generic<kOpSym>(st, a, step, beta, damp); break;
    }
}

}  // namespace
}  // namespace cvfes

using namespace cvfes;

extern "C" int64_t cvfes_step_kernel(const cvfes_system_t *sys, int method) {
    if (!sys || (method != CVFES_METHOD_STORED && method != CVFES_METHOD_MATFREE)) return -1;
    if (sys->variant < CVFES_VARIANT_AUTO || sys->variant > CVFES_VARIANT_EBE) return -1;
    return resolve_variant(sys, method == CVFES_METHOD_MATFREE);
}

extern "C" int cvfes_assemble_bcsr(void *stream, int64_t nnzb, int64_t n_smp, const int32_t *cptr,
                                   const int32_t *celem, const double *khat, const double *coef,
                                   double *values) {
    CVFES_REQUIRE(nnzb >= 0 && n_smp >= 1, "cvfes_assemble_bcsr: bad sizes");
    if (nnzb == 0) return CVFES_OK;
    CVFES_REQUIRE(cptr && celem && khat && coef && values, "cvfes_assemble_bcsr: null pointer");
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(nnzb * n_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), nnzb, n_smp, nullptr, cptr, celem, khat,
                       coef, values);
    return check_launch("k_assemble");
}

extern "C" int cvfes_assemble_upper(void *stream, int64_t nnzu, int64_t n_smp, const int32_t *ublk,
                                    const int32_t *cptr, const int32_t *celem, const double *khat,
                                    const double *coef, double *uvalues) {
    CVFES_REQUIRE(nnzu >= 0 && n_smp >= 1, "cvfes_assemble_upper: bad sizes");
    if (nnzu == 0) return CVFES_OK;
    CVFES_REQUIRE(ublk && cptr && celem && khat && coef && uvalues, "cvfes_assemble_upper: null pointer");
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(nnzu * n_smp)), dim3(kBlock), 0,
                       static_cast<hipStream_t>(stream), nnzu, n_smp, ublk, cptr, celem, khat,
                       coef, uvalues);
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
    a.occ_lds = occupancy_lds(sys->waves_per_simd);
    spmv_dispatch(static_cast<hipStream_t>(stream), method == CVFES_METHOD_MATFREE, a);
    return check_launch("k_spmv");
}

int validate_load(const cvfes_load_t *load, const char *who) {
    CVFES_REQUIRE(load && load->p, "%s: null load / pressure table", who);
    CVFES_REQUIRE(!load->ga || load->qa, "%s: traction field A without its table", who);
    CVFES_REQUIRE(!load->gb || load->qb, "%s: traction field B without its table", who);
    return CVFES_OK;
}

namespace {

// The two launches of cvfes_initial_state around the product K (u0 + beta du0), which `work`
// ([N][3][S]) holds between them.
int initial_state_impl(hipStream_t st, const cvfes_system_t *sys, int method, const double *u0,
                       const double *du0, const cvfes_load_t *load, double dt, double beta,
                       double damp, double *u, double *up, double *work) {
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

int validate_initial_state(const cvfes_system_t *sys, int method, const cvfes_load_t *load,
                           const double *u0, const double *du0, double dt, const double *u,
                           const double *up, const char *who) {
    if (int rc = validate(sys, method, who)) return rc;
    if (int rc = validate_load(load, who)) return rc;
    CVFES_REQUIRE(u0 && du0 && u && up && sys->lm && sys->f1, "%s: null pointer", who);
    CVFES_REQUIRE(dt > 0.0, "%s: dt must be positive", who);
    return CVFES_OK;
}

}  // namespace

extern "C" int cvfes_initial_state(void *stream, const cvfes_system_t *sys, int method,
                                   const double *u0, const double *du0, const cvfes_load_t *load,
                                   double dt, double beta, double damp, double *u, double *up,
                                   double *work) {
    if (int rc = validate_initial_state(sys, method, load, u0, du0, dt, u, up,
                                        "cvfes_initial_state"))
        return rc;
    CVFES_REQUIRE(work, "cvfes_initial_state: null work buffer");
    return initial_state_impl(static_cast<hipStream_t>(stream), sys, method, u0, du0, load, dt,
                              beta, damp, u, up, work);
}

// Device scratch of cvfes_initial_state_static: the product K (u0 + beta du0) between its two
// launches.  Every GPU the library runs on holds its own copy of this array until the process
// ends.  In host code `scratch` names a host-side stand-in, not device memory: kernels get the
// per-device address from hipGetSymbolAddress (scratch_device_ptr) and never the name itself.
static __device__ __align__(256) double scratch[CVFES_STATIC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double *scratch_ptr[kMaxDevices] = {};

// Address of the current device's copy of `scratch`, looked up once per device (with that device
// current) and kept.
int scratch_device_ptr(double **out, const char *who) {
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return fail(CVFES_ERR_HIP, "%s: hipGetDevice: %s", who, hipGetErrorString(err));
    CVFES_REQUIRE(dev >= 0 && dev < kMaxDevices, "%s: device %d beyond the %d supported", who, dev,
                  kMaxDevices);
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(CVFES_ERR_HIP, "%s: hipGetSymbolAddress(scratch) on device %d: %s", who, dev,
                        hipGetErrorString(err));
        scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = scratch_ptr[dev];
    return CVFES_OK;
}

}  // namespace

extern "C" int cvfes_initial_state_static(void *stream, const cvfes_system_t *sys, int method,
                                          const double *u0, const double *du0,
                                          const cvfes_load_t *load, double dt, double beta,
                                          double damp, double *u, double *up) {
    const char *who = "cvfes_initial_state_static";
    if (int rc = validate_initial_state(sys, method, load, u0, du0, dt, u, up, who)) return rc;
    const int64_t need = 3 * sys->n_nodes * sys->n_smp;
    CVFES_REQUIRE(need <= CVFES_STATIC_SCRATCH_DOUBLES,
                  "%s: 3 x %lld nodes x %lld samples = %lld values exceed the static scratch of %lld "
                  "(use cvfes_initial_state with a work buffer)",
                  who, static_cast<long long>(sys->n_nodes), static_cast<long long>(sys->n_smp),
                  static_cast<long long>(need), static_cast<long long>(CVFES_STATIC_SCRATCH_DOUBLES));
    double *work = nullptr;
    if (int rc = scratch_device_ptr(&work, who)) return rc;
    return initial_state_impl(static_cast<hipStream_t>(stream), sys, method, u0, du0, load, dt,
                              beta, damp, u, up, work);
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
    CVFES_REQUIRE(sys->variant >= CVFES_VARIANT_AUTO && sys->variant <= CVFES_VARIANT_EBE,
                  "cvfes_step: unknown variant %lld", static_cast<long long>(sys->variant));
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
    a.order = sys->row_order;
    a.occ_lds = occupancy_lds(sys->waves_per_simd);
    a.nt_out = stream_out(sys->n_nodes, sys->n_smp);
    const bool mf = method == CVFES_METHOD_MATFREE;
    double *cur = state[0], *prev = state[1], *nxt = state[2];
    for (int64_t n = 0; n < nsteps; ++n) {
        a.u = cur;
        a.up = prev;
        a.out = nxt;
        step_dispatch(st, sys, mf, a, n, beta != 0.0, damp != 0.0);
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
