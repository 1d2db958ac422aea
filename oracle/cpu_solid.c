/*
 * C/OpenMP restatement of CVFES's CPU membrane solid path (linear mode) - TEST INFRASTRUCTURE.
 *
 * Used by the tests (as a checker at sizes the NumPy oracle is too slow for) and by bench.py's
 * cpu_baseline leg (timed on the host cores).  Never part of the product path.
 *
 * Layouts follow the reference: K[nnzb][S][3][3] (physicsSolver.py:130-132), u[ndof][S].
 *
 *   oracle_assemble_K  : OptimizedSolidAssemble's stiffness half (optimizedSolidAssemble.pyx:82-127,
 *                        133-188,265-346) - frame T, local coordinates, area, B (5x9), B^T D, B^T D B,
 *                        lK = BTDB * coefK * area, gK = TT^T lK TT, binary-search insertion.  The
 *                        element loop is serial in the reference; here threads split the samples.
 *   oracle_step        : SolidSolver.Solve (physicsSolver.py:263-292) with LC = 0: the (dt^2) K
 *                        temporary, MultiplyByVector (pyx:1250-1274, parallel over block rows) and
 *                        the elementwise passes of the NumPy expression, then the boundary values.
 *                        EXTENSIONS (same as the NumPy oracle): load p * f1, Rayleigh beta, damp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static void frame(const double *nodes, const int64_t *e, double T[3][3], double *area,
                  double lx[3], double ly[3]) {
    double e0[3], e1[3], nrm, proj;
    for (int k = 0; k < 3; ++k) {
        e0[k] = nodes[3 * e[2] + k] - nodes[3 * e[1] + k];
        e1[k] = nodes[3 * e[0] + k] - nodes[3 * e[2] + k];
    }
    nrm = sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
    for (int k = 0; k < 3; ++k) T[0][k] = e0[k] / nrm;
    proj = e1[0] * T[0][0] + e1[1] * T[0][1] + e1[2] * T[0][2];
    for (int k = 0; k < 3; ++k) T[1][k] = e1[k] - proj * T[0][k];
    nrm = sqrt(T[1][0] * T[1][0] + T[1][1] * T[1][1] + T[1][2] * T[1][2]);
    for (int k = 0; k < 3; ++k) T[1][k] = T[1][k] / nrm;
    T[2][0] = T[0][1] * T[1][2] - T[0][2] * T[1][1];
    T[2][1] = T[0][2] * T[1][0] - T[0][0] * T[1][2];
    T[2][2] = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    for (int a = 0; a < 3; ++a) {
        const double *X = nodes + 3 * e[a];
        lx[a] = X[0] * T[0][0] + X[1] * T[0][1] + X[2] * T[0][2];
        ly[a] = X[0] * T[1][0] + X[1] * T[1][1] + X[2] * T[1][2];
    }
    *area = ((lx[1] * ly[2] - lx[2] * ly[1]) - (lx[0] * ly[2] - lx[2] * ly[0]) +
             (lx[0] * ly[1] - lx[1] * ly[0])) * 0.5;
}

static int64_t find_block(const int64_t *indptr, const int64_t *indices, int64_t row, int64_t col) {
    int64_t left = indptr[row], right = indptr[row + 1];
    int64_t ptr = (left + right) / 2;
    while (indices[ptr] != col) {
        if (indices[ptr] > col)
            right = ptr;
        else
            left = ptr;
        ptr = (left + right) / 2;
    }
    return ptr;
}

/* K must be zeroed by the caller. D is the 5x5 matrix of physicsSolver.py:184-188. */
int oracle_assemble_K(int64_t n_elms, int64_t n_smp, const double *nodes, const int64_t *elems,
                      const double *D, const double *coefK, const int64_t *indptr,
                      const int64_t *indices, double *K, int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
        const int64_t s0 = n_smp * tid / nt, s1 = n_smp * (tid + 1) / nt;
        double T[3][3], lx[3], ly[3], area, B[5][9], BTD[9][5], BTDB[9][9], TT[9][9];
        double lK[9][9], tmp[9][9], gK[9][9];
        int64_t pos[3][3];
        for (int64_t ie = 0; ie < n_elms && s0 < s1; ++ie) {
            const int64_t *e = elems + 3 * ie;
            frame(nodes, e, T, &area, lx, ly);
            memset(TT, 0, sizeof(TT));
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) TT[3 * a + i][3 * a + j] = T[i][j];
            const double y23 = ly[1] - ly[2], y31 = ly[2] - ly[0], y12 = ly[0] - ly[1];
            const double x32 = lx[2] - lx[1], x13 = lx[0] - lx[2], x21 = lx[1] - lx[0];
            const double c1 = 2.0 * area;
            memset(B, 0, sizeof(B));
            B[0][0] = B[2][1] = B[3][2] = y23 / c1;
            B[0][3] = B[2][4] = B[3][5] = y31 / c1;
            B[0][6] = B[2][7] = B[3][8] = y12 / c1;
            B[1][1] = B[2][0] = B[4][2] = x32 / c1;
            B[1][4] = B[2][3] = B[4][5] = x13 / c1;
            B[1][7] = B[2][6] = B[4][8] = x21 / c1;
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 5; ++j) {
                    BTD[i][j] = 0.0;
                    for (int k = 0; k < 5; ++k) BTD[i][j] += B[k][i] * D[5 * k + j];
                }
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 9; ++j) {
                    BTDB[i][j] = 0.0;
                    for (int k = 0; k < 5; ++k) BTDB[i][j] += BTD[i][k] * B[k][j];
                }
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) pos[a][b] = find_block(indptr, indices, e[a], e[b]);
            for (int64_t s = s0; s < s1; ++s) {
                const double cf = coefK[ie * n_smp + s];
                for (int i = 0; i < 9; ++i)
                    for (int j = 0; j < 9; ++j) lK[i][j] = BTDB[i][j] * cf * area;
                for (int i = 0; i < 9; ++i)      /* tmp = TT^T lK */
                    for (int j = 0; j < 9; ++j) {
                        tmp[i][j] = 0.0;
                        for (int k = 0; k < 9; ++k) tmp[i][j] += TT[k][i] * lK[k][j];
                    }
                for (int i = 0; i < 9; ++i)      /* gK = tmp TT */
                    for (int j = 0; j < 9; ++j) {
                        gK[i][j] = 0.0;
                        for (int k = 0; k < 9; ++k) gK[i][j] += tmp[i][k] * TT[k][j];
                    }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double *blk = K + ((pos[a][b] * n_smp) + s) * 9;
                        for (int i = 0; i < 3; ++i)
                            for (int j = 0; j < 3; ++j) blk[3 * i + j] += gK[3 * a + i][3 * b + j];
                    }
            }
        }
    }
    return 0;
}

/* res[3a+i][s] += sum_p sum_j M[p][s][i][j] v[3 col + j][s]  (MultiplyByVector). */
static void multiply_by_vector(int64_t n_nodes, int64_t n_smp, const int64_t *indptr,
                               const int64_t *indices, const double *M, const double *v,
                               double *res) {
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < n_nodes; ++a) {
        const int64_t r = 3 * a;
        for (int64_t ia = indptr[a]; ia < indptr[a + 1]; ++ia) {
            const int64_t c = 3 * indices[ia];
            const double *blk = M + ia * n_smp * 9;
            for (int64_t s = 0; s < n_smp; ++s) {
                const double *m = blk + 9 * s;
                res[r * n_smp + s] += m[0] * v[c * n_smp + s] + m[1] * v[(c + 1) * n_smp + s] +
                                      m[2] * v[(c + 2) * n_smp + s];
                res[(r + 1) * n_smp + s] += m[3] * v[c * n_smp + s] + m[4] * v[(c + 1) * n_smp + s] +
                                            m[5] * v[(c + 2) * n_smp + s];
                res[(r + 2) * n_smp + s] += m[6] * v[c * n_smp + s] + m[7] * v[(c + 1) * n_smp + s] +
                                            m[8] * v[(c + 2) * n_smp + s];
            }
        }
    }
}

int oracle_spmv(int64_t n_nodes, int64_t n_smp, const int64_t *indptr, const int64_t *indices,
                const double *K, const double *v, double *res, int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    memset(res, 0, sizeof(double) * 3 * n_nodes * n_smp);
    multiply_by_vector(n_nodes, n_smp, indptr, indices, K, v, res);
    return 0;
}

/*
 * One step.  u, up, LM, LHS [ndof][S]; f1 [ndof]; unew [ndof][S] out.  Follows
 * physicsSolver.py:272-285: A1 = dt^2 f; Ku = MultiplyByVector((dt^2) K, x); A2 = 2 LM u - Ku;
 * A3 = (0.5 dt LC - LM) up; RHS = A1 + A2 + A3; u = RHS / LHS where LHS != 0; u[bdy] = bdy_u.
 * Each NumPy operation is one pass over full-size arrays (serial, like NumPy); the (dt^2) K product
 * allocates a fresh temporary as the Python expression does.
 */
int oracle_step(int64_t n_nodes, int64_t n_smp, int64_t nnzb, const int64_t *indptr,
                const int64_t *indices, const double *K, const double *LM, const double *LHS,
                const double *f1, double p, const double *u, const double *up, double dt,
                double beta, double damp, const int64_t *bdy_dofs, int64_t n_bdy, double bdy_u,
                double *unew, int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    const int64_t ndof = 3 * n_nodes, n = ndof * n_smp, nk = nnzb * n_smp * 9;
    const double dt2 = dt * dt;
    double *Kdt = (double *)malloc(sizeof(double) * nk);
    double *x = (double *)malloc(sizeof(double) * n);
    double *Ku = (double *)malloc(sizeof(double) * n);
    double *A1 = (double *)malloc(sizeof(double) * n);
    double *A2 = (double *)malloc(sizeof(double) * n);
    double *A3 = (double *)malloc(sizeof(double) * n);
    if (!Kdt || !x || !Ku || !A1 || !A2 || !A3) {
        free(Kdt); free(x); free(Ku); free(A1); free(A2); free(A3);
        return -1;
    }
    for (int64_t k = 0; k < nk; ++k) Kdt[k] = dt2 * K[k];                        /* (dt**2)*K   */
    if (beta != 0.0) {
        const double bdt = beta / dt;
        for (int64_t k = 0; k < n; ++k) x[k] = u[k] + bdt * (u[k] - up[k]);
    } else {
        memcpy(x, u, sizeof(double) * n);
    }
    for (int64_t d = 0; d < ndof; ++d)                                          /* A1          */
        for (int64_t s = 0; s < n_smp; ++s) {
            const int64_t k = d * n_smp + s;
            double f = p * f1[d];
            if (damp != 0.0) f -= damp * (u[k] - up[k]) / dt;
            A1[k] = dt2 * f;
        }
    memset(Ku, 0, sizeof(double) * n);                                          /* Ku = zeros  */
    multiply_by_vector(n_nodes, n_smp, indptr, indices, Kdt, x, Ku);
    for (int64_t k = 0; k < n; ++k) A2[k] = 2.0 * LM[k];                        /* 2.0*LM      */
    for (int64_t k = 0; k < n; ++k) A2[k] = A2[k] * u[k];                       /* *u          */
    for (int64_t k = 0; k < n; ++k) A2[k] = A2[k] - Ku[k];                      /* - Ku        */
    for (int64_t k = 0; k < n; ++k) A3[k] = 0.0 - LM[k];                        /* 0.5dt LC-LM */
    for (int64_t k = 0; k < n; ++k) A3[k] = A3[k] * up[k];                      /* *up         */
    for (int64_t k = 0; k < n; ++k) A1[k] = A1[k] + A2[k];                      /* A1 + A2     */
    for (int64_t k = 0; k < n; ++k) A1[k] = A1[k] + A3[k];                      /* + A3        */
    for (int64_t k = 0; k < n; ++k) unew[k] = LHS[k] != 0.0 ? A1[k] / LHS[k] : 0.0;
    for (int64_t b = 0; b < n_bdy; ++b)
        for (int64_t s = 0; s < n_smp; ++s) unew[bdy_dofs[b] * n_smp + s] = bdy_u;
    free(Kdt); free(x); free(Ku); free(A1); free(A2); free(A3);
    return 0;
}

/*
 * n_steps steps of oracle_step with the pressure table p[n_steps], for long runs.  (dt^2) K is
 * formed once instead of every step: it is the same deterministic product each time, and the
 * vector passes of oracle_step are done per entry in the same order, so every step's result equals
 * oracle_step's bit for bit.  On entry u = u_0, up = u_{-1}; on return u = u_n, up = u_{n-1}.
 * After step snap_steps[i] (1-based count of completed steps) u is copied to snaps[i][ndof][S].
 */
int oracle_run(int64_t n_nodes, int64_t n_smp, int64_t nnzb, const int64_t *indptr,
               const int64_t *indices, const double *K, const double *LM, const double *LHS,
               const double *f1, const double *p, int64_t n_steps, double *u, double *up,
               double dt, double beta, double damp, const int64_t *bdy_dofs, int64_t n_bdy,
               double bdy_u, const int64_t *snap_steps, int64_t n_snap, double *snaps,
               int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    const int64_t ndof = 3 * n_nodes, n = ndof * n_smp, nk = nnzb * n_smp * 9;
    const double dt2 = dt * dt;
    double *Kdt = (double *)malloc(sizeof(double) * nk);
    double *x = (double *)malloc(sizeof(double) * n);
    double *Ku = (double *)malloc(sizeof(double) * n);
    double *unew = (double *)malloc(sizeof(double) * n);
    double *cur = (double *)malloc(sizeof(double) * n);
    double *prev = (double *)malloc(sizeof(double) * n);
    if (!Kdt || !x || !Ku || !unew || !cur || !prev) {
        free(Kdt); free(x); free(Ku); free(unew); free(cur); free(prev);
        return -1;
    }
    memcpy(cur, u, sizeof(double) * n);
    memcpy(prev, up, sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nk; ++k) Kdt[k] = dt2 * K[k];
    const double bdt = beta / dt;
    for (int64_t step = 0; step < n_steps; ++step) {
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            x[k] = beta != 0.0 ? cur[k] + bdt * (cur[k] - prev[k]) : cur[k];
            Ku[k] = 0.0;
        }
        multiply_by_vector(n_nodes, n_smp, indptr, indices, Kdt, x, Ku);
        const double pn = p[step];
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < ndof; ++d)
            for (int64_t s = 0; s < n_smp; ++s) {
                const int64_t k = d * n_smp + s;
                double f = pn * f1[d];
                if (damp != 0.0) f -= damp * (cur[k] - prev[k]) / dt;
                double a1 = dt2 * f;
                double a2 = 2.0 * LM[k];
                a2 = a2 * cur[k];
                a2 = a2 - Ku[k];
                double a3 = 0.0 - LM[k];
                a3 = a3 * prev[k];
                a1 = a1 + a2;
                a1 = a1 + a3;
                unew[k] = LHS[k] != 0.0 ? a1 / LHS[k] : 0.0;
            }
        for (int64_t b = 0; b < n_bdy; ++b)
            for (int64_t s = 0; s < n_smp; ++s) unew[bdy_dofs[b] * n_smp + s] = bdy_u;
        double *t = prev;
        prev = cur;
        cur = unew;
        unew = t;
        for (int64_t i = 0; i < n_snap; ++i)
            if (snap_steps[i] == step + 1) memcpy(snaps + i * n, cur, sizeof(double) * n);
    }
    memcpy(u, cur, sizeof(double) * n);
    memcpy(up, prev, sizeof(double) * n);
    free(Kdt); free(x); free(Ku); free(unew); free(cur); free(prev);
    return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
