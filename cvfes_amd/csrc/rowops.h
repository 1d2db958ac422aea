// Row operations shared by the single-mesh and the batched step kernels: the gathered block-row
// product (stored blocks, or element unit blocks scaled by the element coefficients) and the fused
// central-difference update of one block row.
#pragma once

#include "common.h"

namespace cvfes {

// Lanes per block row.  Packed (UNIFORM = false): S/VW lanes, so a wave may span several rows.
// Wave-uniform (UNIFORM = true): S/VW rounded up to whole wavefronts; lanes past the last sample
// are inactive (they read the last sample group and do not store), which keeps every row on
// whole wavefronts for any S.
template <bool UNIFORM, int VW>
__host__ __device__ __forceinline__ int64_t lanes_per_row(int64_t n_smp) {
    const int64_t l = n_smp / VW;
    return UNIFORM ? (l + 63) / 64 * 64 : l;
}

template <bool UNIFORM, int VW>
__device__ __forceinline__ bool row_and_sample(int64_t n_nodes, int64_t n_smp, int32_t &row,
                                               int64_t &s, bool &active) {
    const int64_t lpr = lanes_per_row<UNIFORM, VW>(n_smp);
    const int64_t g = xcd_block() * kBlock + threadIdx.x;
    int64_t r = g / lpr;
    s = (g - r * lpr) * VW;
    active = s < n_smp;
    if (!active) s = n_smp - VW;
    if constexpr (UNIFORM) r = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(r));
    row = static_cast<int32_t>(r);
    return r < n_nodes;
}

// Wave-uniform rows pay off when they keep at least 3/4 of the lanes busy.
template <int VW>
inline bool use_uniform(int64_t n_smp) {
    const int64_t l = n_smp / VW, padded = (l + 63) / 64 * 64;
    return 4 * l >= 3 * padded;
}

// Samples per lane of the generic row kernels (step, product, adjoint, fused): two for even
// counts, except where only one keeps the rows wave-uniform (64 samples: 32 two-sample lanes
// would put two rows in a wave, and the element blocks would come in per lane instead of through
// scalar loads).  The per-sample arithmetic is the same either way.
inline int generic_vw(int64_t n_smp) {
    if (n_smp % 2 != 0) return 1;
    return (!use_uniform<2>(n_smp) && use_uniform<1>(n_smp)) ? 1 : 2;
}

// Operator forms of the row functions (template argument OP): stored values in full storage,
// matrix-free (element coefficients x unit element blocks), stored values in symmetric storage.
constexpr int kOpStored = 0;
constexpr int kOpMatfree = 1;
constexpr int kOpSym = 2;

// Stored operator: full storage `values` [nnzb][9][S], or, when `values` is null, symmetric
// storage: the upper blocks `uvalues` [nnzu][9][S], block p of the row structure being
// uvalues[vblk[p]], read transposed when its column precedes its row.
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
    const double *uvalues;
    const int32_t *vblk;
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
    const int32_t *order; // visiting order of the rows (wave-uniform kernels; null = natural)
    unsigned occ_lds;     // launch only: unused LDS per workgroup that caps the occupancy
    bool nt_out;          // write `out` with non-temporal stores (stream_out)
};

// The same arguments with global-address-space pointers (see CVFES_G); the generic row functions
// below take either form.
// Read-only per-row data (structure, unit blocks, masses, loads) is in the constant address space
// (CVFES_K); the per-sample streams (values, coefficients, states) are global.
struct GRowArgs {
    int64_t n_nodes;
    int64_t n_smp;
    const CVFES_K int32_t *indptr;
    const CVFES_K int32_t *indices;
    const CVFES_G double *values;
    const CVFES_K int32_t *cptr;
    const CVFES_K int32_t *celem;
    const CVFES_K double *khat;
    const CVFES_G double *coef;
    const CVFES_G double *uvalues;
    const CVFES_K int32_t *vblk;
};

struct GLoad {
    const CVFES_K double *p, *ga, *qa, *gb, *qb;
};

struct GStepArgs {
    GRowArgs k;
    const CVFES_G double *u;
    const CVFES_G double *up;
    CVFES_G double *out;
    const CVFES_K double *lm;
    int64_t lm_sstride;
    const CVFES_K double *f1;
    const CVFES_K uint8_t *fixed;
    GLoad load;
    double bdy_u;
    double dt2;
    double bdt;
    double ddt;
    bool nt_out;
};

// Stored block p of the row: kv[ij][v] = K_p[ij] of sample s + v.  SYM: symmetric storage, the
// block is the upper block vblk[p], transposed when `tr` (its column precedes its row).
template <bool SYM, int VW, class RA>
__device__ __forceinline__ void stored_block(const RA &k, int32_t p, bool tr, int64_t s,
                                             double (&kv)[9][VW]) {
    if constexpr (!SYM) {
        const auto *base = k.values + (static_cast<int64_t>(p) * 9) * k.n_smp + s;
#pragma unroll
        for (int ij = 0; ij < 9; ++ij) {
            const Vec<VW> t = load_vec<VW>(base + ij * k.n_smp);
#pragma unroll
            for (int v = 0; v < VW; ++v) kv[ij][v] = t.v[v];
        }
    } else {
        const auto *base = k.uvalues + (static_cast<int64_t>(k.vblk[p]) * 9) * k.n_smp + s;
#pragma unroll
        for (int ij = 0; ij < 9; ++ij) {
            const int src = tr ? 3 * (ij % 3) + ij / 3 : ij;
            const Vec<VW> t = load_vec<VW>(base + src * k.n_smp);
#pragma unroll
            for (int v = 0; v < VW; ++v) kv[ij][v] = t.v[v];
        }
    }
}

// acc[i][v] += K_p (i, :) . x for block p of the row.  Stored: the 3x3 block is read.  Matrix-free:
// K_p = sum over the block's element contributions c of coef_e(c) * khat_c, applied as
// acc += coef_e (khat_c x), so the block itself is never formed and only the three sums per
// sample are live (khat_c is wave-uniform for wave-uniform rows: scalar loads).
// `tr`: the block's column precedes its row (symmetric storage reads it transposed).
template <int OP, int VW, class RA>
__device__ __forceinline__ void block_apply(const RA &k, int32_t p, bool tr, int64_t s,
                                            const double (&x)[3][VW], double (&acc)[3][VW]) {
    if constexpr (OP != kOpMatfree) {
        double kv[9][VW];
        stored_block<OP == kOpSym, VW>(k, p, tr, s, kv);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int v = 0; v < VW; ++v) acc[i][v] = fma(kv[3 * i + j][v], x[j][v], acc[i][v]);
    } else {
        const int32_t c0 = k.cptr[p], c1 = k.cptr[p + 1];
        for (int32_t c = c0; c < c1; ++c) {
            const int64_t e = k.celem[c];
            const Vec<VW> cf = load_vec<VW>(k.coef + e * k.n_smp + s);
            const auto *kh = k.khat + 9 * static_cast<int64_t>(c);
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

// Column term of the step's product at node col: x = u + bdt (u - up) when BETA, else u.
template <bool BETA, int VW, class DP>
__device__ __forceinline__ void state_column(DP u, DP up, double bdt, int64_t col, int64_t n_smp,
                                             int64_t s, double (&x)[3][VW]) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int64_t off = (col * 3 + j) * n_smp + s;
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
}

// block_apply for two columns at once (the block is loaded once): acc += K_p x and acc2 += K_p x2,
// each with exactly the operations block_apply performs.
template <int OP, int VW, class RA>
__device__ __forceinline__ void block_apply2(const RA &k, int32_t p, bool tr, int64_t s,
                                             const double (&x)[3][VW], double (&acc)[3][VW],
                                             const double (&x2)[3][VW], double (&acc2)[3][VW]) {
    if constexpr (OP != kOpMatfree) {
        double kv[9][VW];
        stored_block<OP == kOpSym, VW>(k, p, tr, s, kv);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    acc[i][v] = fma(kv[3 * i + j][v], x[j][v], acc[i][v]);
                    acc2[i][v] = fma(kv[3 * i + j][v], x2[j][v], acc2[i][v]);
                }
    } else {
        const int32_t c0 = k.cptr[p], c1 = k.cptr[p + 1];
        for (int32_t c = c0; c < c1; ++c) {
            const int64_t e = k.celem[c];
            const Vec<VW> cf = load_vec<VW>(k.coef + e * k.n_smp + s);
            const auto *kh = k.khat + 9 * static_cast<int64_t>(c);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double h0 = kh[3 * i], h1 = kh[3 * i + 1], h2 = kh[3 * i + 2];
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    const double t = fma(h2, x[2][v], fma(h1, x[1][v], h0 * x[0][v]));
                    acc[i][v] = fma(cf.v[v], t, acc[i][v]);
                    const double t2 = fma(h2, x2[2][v], fma(h1, x2[1][v], h0 * x2[0][v]));
                    acc2[i][v] = fma(cf.v[v], t2, acc2[i][v]);
                }
            }
        }
    }
}

// acc[i][v] = sum over the row's blocks of K_p (i, :) . x_col, with x = u + bdt (u - up) when BETA.
template <int OP, bool BETA, int VW, class RA, class DP>
__device__ __forceinline__ void row_product(const RA &k, DP u, DP up, double bdt, int32_t row,
                                            int64_t s, double (&acc)[3][VW]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int v = 0; v < VW; ++v) acc[i][v] = 0.0;
    const int32_t p0 = k.indptr[row], p1 = k.indptr[row + 1];
    for (int32_t p = p0; p < p1; ++p) {
        const int64_t col = k.indices[p];
        double x[3][VW];
        state_column<BETA, VW>(u, up, bdt, col, k.n_smp, s, x);
        block_apply<OP, VW>(k, p, col < row, s, x, acc);
    }
}

// Central-difference update of one block row and VW consecutive samples from its product
// acc = K x: load, damping, lumped mass and Dirichlet value; stores u_{n+1} when `active`.  Every
// product and sum that could be contracted is an explicit fma, so the rounding is the same in
// every kernel that uses it (single-mesh, batched, and the step fused with the adjoint).
template <bool DAMP, int VW, class SA>
__device__ __forceinline__ void step_update(const SA &a, int32_t row, int64_t s, bool active,
                                            int64_t step, const double (&acc)[3][VW]) {
    const double pn = a.load.p[step];
    const double qa = a.load.ga ? a.load.qa[step] : 0.0;
    const double qb = a.load.gb ? a.load.qb[step] : 0.0;
    const bool fixed = a.fixed != nullptr && a.fixed[row] != 0;
    const int64_t nsm = a.k.n_smp;
    const auto *lmp = a.lm + (a.lm_sstride ? static_cast<int64_t>(row) * nsm + s : row);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * nsm + s;
        const Vec<VW> uu = load_vec<VW>(a.u + off);
        const Vec<VW> pp = load_vec<VW>(a.up + off);
        const int64_t d = 3 * static_cast<int64_t>(row) + i;
        double fi = pn * a.f1[d];
        if (a.load.ga) fi = fma(qa, a.load.ga[d], fi);
        if (a.load.gb) fi = fma(qb, a.load.gb[d], fi);
        Vec<VW> o;
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            const double m = a.lm_sstride ? lmp[v] : lmp[0];
            double force = fi;
            if constexpr (DAMP) force = fma(-a.ddt, uu.v[v] - pp.v[v], force);
            const double rhs = fma(a.dt2, force - acc[i][v], m * fma(2.0, uu.v[v], -pp.v[v]));
            double un = m != 0.0 ? rhs / m : 0.0;
            o.v[v] = fixed ? a.bdy_u : un;
        }
        // u_{n+1} is not read again in this launch: large outputs are streamed past L2 (stream_out)
        if (active) {
            if (a.nt_out) store_vec_nt<VW>(a.out + off, o);
            else store_vec<VW>(a.out + off, o);
        }
    }
}

// Block-row product and central-difference update of one block row and VW consecutive samples.
template <int OP, bool BETA, bool DAMP, int VW, class SA>
__device__ __forceinline__ void step_row(const SA &a, int32_t row, int64_t s, bool active,
                                         int64_t step) {
    double acc[3][VW];
    row_product<OP, BETA, VW>(a.k, a.u, a.up, a.bdt, row, s, acc);
    step_update<DAMP, VW>(a, row, s, active, step, acc);
}

// ---------------------------------------------------------------------------------------------
// Ext-vector forms (one value or two consecutive samples per lane) used by the row-load and
// LDS-staged kernels.
typedef double dbl2 __attribute__((ext_vector_type(2)));

template <int VW> struct VecT;
template <> struct VecT<1> { typedef double type; };
template <> struct VecT<2> { typedef dbl2 type; };

template <int VW, bool NT>
__device__ __forceinline__ typename VecT<VW>::type ldv(const double *p) {
    typedef typename VecT<VW>::type V;
    if constexpr (NT)
        return __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
    else
        return *reinterpret_cast<const V *>(p);
}

// Buffer loads: a resource (base address and size in bytes, built in scalar registers from
// wave-uniform values) and a 32-bit per-lane byte offset.  Every load of a row then shares one
// per-lane address register (the sample offset) instead of holding a 64-bit address per load,
// which is what lets the symmetric step kernel keep four waves per SIMD.  Word 3 of the resource
// is the gfx9 raw-access form (32-bit data format, no swizzle).  AUX is the load's cache policy:
// 0 plain, kAuxNT non-temporal.  Every offset the kernels form lies inside its resource, so the
// hardware's range check (which would return zeros) never applies.
constexpr int kAuxNT = 2;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const double *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), static_cast<short>(0),
                                             static_cast<int>(bytes), 0x00020000);
}

template <int VW, int AUX>
__device__ __forceinline__ typename VecT<VW>::type ldb(__amdgpu_buffer_rsrc_t r, uint32_t lane_off,
                                                       uint32_t off) {
    if constexpr (VW == 2)
        return __builtin_bit_cast(dbl2, __builtin_amdgcn_raw_buffer_load_b128(
                                            r, static_cast<int>(lane_off), static_cast<int>(off), AUX));
    else
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
                                              r, static_cast<int>(lane_off), static_cast<int>(off), AUX));
}

template <int VW>
__device__ __forceinline__ Vec<VW> to_vec(typename VecT<VW>::type t) {
    Vec<VW> r;
    if constexpr (VW == 2) {
        r.v[0] = t[0];
        r.v[1] = t[1];
    } else {
        r.v[0] = t;
    }
    return r;
}

template <int VW>
__device__ __forceinline__ typename VecT<VW>::type splat(double x) {
    if constexpr (VW == 2) return dbl2{x, x};
    else return x;
}

template <int VW>
__device__ __forceinline__ typename VecT<VW>::type sel(bool keep, typename VecT<VW>::type v) {
    return keep ? v : splat<VW>(0.0);
}

// Central-difference update of one block row from its product acc = K (u + bdt (u - up)), with
// the row's own u/up already loaded; stores u_{n+1} when `active`, non-temporally when NTO (a
// compile-time choice: the row-load and symmetric kernels with non-temporal value loads, which
// are the forms for large problems, stream their output too).
template <bool DAMP, int VW, bool NTO = false>
__device__ __forceinline__ void finish_row(const StepArgs &a, int32_t row, int64_t s, bool active,
                                           int64_t step, const typename VecT<VW>::type (&uo)[3],
                                           const typename VecT<VW>::type (&po)[3],
                                           const typename VecT<VW>::type (&acc)[3]) {
    typedef typename VecT<VW>::type V;
    const int64_t S = a.k.n_smp;
    const double pn = a.load.p[step];
    const double qa = a.load.ga ? a.load.qa[step] : 0.0;
    const double qb = a.load.gb ? a.load.qb[step] : 0.0;
    const bool fixed = a.fixed != nullptr && a.fixed[row] != 0;
    const V m = a.lm_sstride ? ldv<VW, false>(a.lm + static_cast<int64_t>(row) * S + s)
                             : splat<VW>(a.lm[row]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t d = 3 * static_cast<int64_t>(row) + i;
        double fi = pn * a.f1[d];
        if (a.load.ga) fi += qa * a.load.ga[d];
        if (a.load.gb) fi += qb * a.load.gb[d];
        V force = splat<VW>(fi);
        if constexpr (DAMP) force -= a.ddt * (uo[i] - po[i]);
        const V rhs = a.dt2 * (force - acc[i]) + m * (2.0 * uo[i] - po[i]);
        V un;
        if constexpr (VW == 2) {
            un.x = m.x != 0.0 ? rhs.x / m.x : 0.0;
            un.y = m.y != 0.0 ? rhs.y / m.y : 0.0;
        } else {
            un = m != 0.0 ? rhs / m : 0.0;
        }
        if (fixed) un = splat<VW>(a.bdy_u);
        if (active) {
            V *dst = reinterpret_cast<V *>(a.out + (static_cast<int64_t>(row) * 3 + i) * S + s);
            if constexpr (NTO) __builtin_nontemporal_store(un, dst);
            else *dst = un;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Matrix-free row step with the row's element data staged in LDS.
//
// The generic matrix-free row walks blocks and, inside each block, element contributions, and
// every contribution is a dependent chain (element id -> coefficient load -> FMAs), about 25 memory
// round trips per row.  Here one wavefront owns one row (wave-uniform, 64*VW samples) and:
//   1. loads, all at once, the coefficient of every element incident to the row node (one
//      coalesced load per element), the unit element blocks and element slots of the row's
//      contributions (contiguous: one lane per value), and the neighbour states;
//   2. parks coefficients, element blocks and slots in its LDS slice;
//   3. rebuilds each block as sum_c coef[slot_c] * khat_c from LDS (element blocks and slots are
//      read at wave-uniform addresses, i.e. broadcast) and accumulates the product with the
//      neighbour state.
// A row node with ne incident elements has exactly 3*ne contributions (each element adds one to
// the blocks (row, b) of its three nodes b).  Limits (exact maxima, checked on the host): blocks
// per row <= MAXB, incident elements per node <= MAXE.
struct LdsArgs {
    const int32_t *nptr;    // [N+1] node -> incident (element, local) list
    const int32_t *nent;    // [3E]  e*3 + local
    const int32_t *cslot;   // [9E]  slot of contribution c's element in its row node's list
};

template <int VW, int MAXE>
struct LdsRow {
    typename VecT<VW>::type coef[MAXE][64];
    double khat[3 * MAXE * 9];
    int32_t slot[3 * MAXE];
};

// Orders this wavefront's LDS writes before its later reads (and earlier reads before later
// writes when a wavefront reuses its slice); only the own wavefront shares the slice.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool BETA, bool DAMP, int VW, int MAXB, int MAXE>
__device__ __forceinline__ void step_row_lds(const StepArgs &a, const LdsArgs &l, LdsRow<VW, MAXE> &sl,
                                             int32_t row, int64_t s, bool active, int64_t step) {
    constexpr int MAXC = 3 * MAXE;
    constexpr int KH = (MAXC * 9 + 63) / 64;     // element-block values per lane
    typedef typename VecT<VW>::type V;
    const int lane = threadIdx.x & 63;
    const RowArgs &k = a.k;
    const int64_t S = k.n_smp;
    const int32_t p0 = k.indptr[row];
    const int32_t nb = k.indptr[row + 1] - p0;
    const int32_t q0 = l.nptr[row];
    const int32_t ne = l.nptr[row + 1] - q0;
    V uo[3], po[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * S + s;
        uo[i] = ldv<VW, false>(a.u + off);
        po[i] = ldv<VW, false>(a.up + off);
    }
    V acc[3] = {splat<VW>(0.0), splat<VW>(0.0), splat<VW>(0.0)};
    if (nb > 0 && ne > 0) {
        const int32_t c0 = k.cptr[p0];
        const int32_t nc = k.cptr[p0 + nb] - c0;
        // 1. every load of the row, issued together
        V cf[MAXE];
#pragma unroll
        for (int e = 0; e < MAXE; ++e) {
            const int64_t el = l.nent[q0 + (e < ne ? e : ne - 1)] / 3;
            cf[e] = ldv<VW, false>(k.coef + el * S + s);
        }
        double kh[KH];
#pragma unroll
        for (int j = 0; j < KH; ++j) {
            const int idx = lane + 64 * j;
            kh[j] = idx < 9 * nc ? k.khat[9 * static_cast<int64_t>(c0) + idx] : 0.0;
        }
        const int32_t slot = lane < nc ? l.cslot[c0 + lane] : 0;
        V w[MAXB][3];
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int64_t col = k.indices[p0 + (b < nb ? b : nb - 1)];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int64_t off = (col * 3 + j) * S + s;
                const V xu = ldv<VW, false>(a.u + off);
                if constexpr (BETA) {
                    const V xp = ldv<VW, false>(a.up + off);
                    w[b][j] = xu + a.bdt * (xu - xp);
                } else {
                    w[b][j] = xu;
                }
            }
        }
        // 2. stage coefficients, element blocks and slots in this wavefront's LDS slice
        wave_lds_sync();
#pragma unroll
        for (int e = 0; e < MAXE; ++e) sl.coef[e][lane] = cf[e];
#pragma unroll
        for (int j = 0; j < KH; ++j) {
            const int idx = lane + 64 * j;
            if (idx < MAXC * 9) sl.khat[idx] = kh[j];
        }
        if (lane < MAXC) sl.slot[lane] = slot;
        wave_lds_sync();
        // 3. rebuild each block from LDS and accumulate its product
        int32_t cb = 0;
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            if (b < nb) {
                const int32_t ce = k.cptr[p0 + b + 1] - c0;
                V kv[9];
#pragma unroll
                for (int ij = 0; ij < 9; ++ij) kv[ij] = splat<VW>(0.0);
                for (int32_t j = cb; j < ce; ++j) {
                    const V c = sl.coef[sl.slot[j]][lane];
                    const double *kb = &sl.khat[9 * j];
#pragma unroll
                    for (int ij = 0; ij < 9; ++ij) kv[ij] += c * kb[ij];
                }
                cb = ce;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) acc[i] += kv[3 * i + j] * w[b][j];
            }
        }
    }
    finish_row<DAMP, VW>(a, row, s, active, step, uo, po, acc);
}

// ---------------------------------------------------------------------------------------------
// Matrix-free row step with the row node's element coefficients loaded once per row.
//
// In the generic matrix-free row every element contribution loads its coefficient again (a row
// has 3 contributions per incident element) and each load sits in a dependent chain.  Here the
// wavefront (one row, wave-uniform) first loads the coefficients of the node's incident elements
// (at most MAXE, all issued together), then walks the blocks; each contribution picks its
// element's coefficient from registers by its slot (cslot, uniform) and its unit block from the
// scalar cache, so no vector load remains inside the contribution loop.
template <bool BETA, bool DAMP, int VW, int MAXE>
__device__ __forceinline__ void step_row_cached(const StepArgs &a, const LdsArgs &l, int32_t row,
                                                int64_t s, bool active, int64_t step) {
    typedef typename VecT<VW>::type V;
    const RowArgs &k = a.k;
    const int64_t S = k.n_smp;
    const int32_t q0 = l.nptr[row];
    const int32_t ne = l.nptr[row + 1] - q0;
    V cf[MAXE];
#pragma unroll
    for (int e = 0; e < MAXE; ++e) cf[e] = splat<VW>(0.0);
    if (ne > 0) {                   // an isolated node has no elements and no blocks
#pragma unroll
        for (int e = 0; e < MAXE; ++e) {
            const int64_t el = l.nent[q0 + (e < ne ? e : ne - 1)] / 3;
            cf[e] = ldv<VW, false>(k.coef + el * S + s);
        }
    }
    V uo[3], po[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int64_t off = (static_cast<int64_t>(row) * 3 + i) * S + s;
        uo[i] = ldv<VW, false>(a.u + off);
        po[i] = ldv<VW, false>(a.up + off);
    }
    V acc[3] = {splat<VW>(0.0), splat<VW>(0.0), splat<VW>(0.0)};
    const int32_t p0 = k.indptr[row], p1 = k.indptr[row + 1];
    for (int32_t p = p0; p < p1; ++p) {
        const int64_t col = k.indices[p];
        V x[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int64_t off = (col * 3 + j) * S + s;
            const V xu = ldv<VW, false>(a.u + off);
            if constexpr (BETA) {
                const V xp = ldv<VW, false>(a.up + off);
                x[j] = xu + a.bdt * (xu - xp);
            } else {
                x[j] = xu;
            }
        }
        V kv[9];
#pragma unroll
        for (int ij = 0; ij < 9; ++ij) kv[ij] = splat<VW>(0.0);
        for (int32_t c = k.cptr[p]; c < k.cptr[p + 1]; ++c) {
            const int32_t slot = l.cslot[c];
            V cfc = cf[0];
#pragma unroll
            for (int e = 1; e < MAXE; ++e)
                if (slot == e) cfc = cf[e];
            const double *kh = k.khat + 9 * static_cast<int64_t>(c);
#pragma unroll
            for (int ij = 0; ij < 9; ++ij) kv[ij] += cfc * kh[ij];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) acc[i] += kv[3 * i + j] * x[j];
    }
    finish_row<DAMP, VW>(a, row, s, active, step, uo, po, acc);
}

}  // namespace cvfes
