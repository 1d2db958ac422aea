// Row operations shared by the single-mesh and the batched step kernels: block values (stored or
// rebuilt from element coefficients), the gathered block-row product and the fused
// central-difference update of one block row.
#pragma once

#include "common.h"

namespace cvfes {

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

// Block-row product, central-difference update, load and Dirichlet value for one block row and
// VW consecutive samples; stores u_{n+1} when `active`.
template <bool MATFREE, bool BETA, bool DAMP, int VW>
__device__ __forceinline__ void step_row(const StepArgs &a, int32_t row, int64_t s, bool active,
                                         int64_t step) {
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
        if (active) store_vec<VW>(a.out + off, o);
    }
}

}  // namespace cvfes
