/*
 * cvfes_mi355x.h - C ABI of the MI355X (gfx950) ensemble membrane solid step.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream (passed as void*; NULL is
 * the legacy default stream) and launches asynchronously on that stream.  Nothing here allocates
 * device memory: the caller owns every buffer.  All functions return CVFES_OK (0) or a negative
 * error code; cvfes_last_error() then describes the failure (thread-local string).
 *
 * Layouts (FP64 unless stated, "sample-fastest" = the sample index s is the fastest axis):
 *   nodes      [N][3]          reference coordinates
 *   elems      [E][3]  int32   triangle node ids
 *   u, up, ... [N][3][S]       = the reference's [ndof][nSmp] state layout
 *   coef       [E][S]          per element-sample stiffness coefficient (thickness x modulus)
 *   values     [nnzb][9][S]    stored 3x3 block-CSR values, block entry (i,j) at 3*i+j
 *   khat       [9E][9]         unit-coefficient element blocks, one per contribution, in block order
 *   lm         [N][S] or [N]   lumped mass per node (equal for its 3 dofs)
 *   f1         [N][3]          nodal load of a unit pressure
 */
#ifndef CVFES_MI355X_H
#define CVFES_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CVFES_ABI_VERSION 1

#define CVFES_OK 0
#define CVFES_ERR_ARG (-1)   /* invalid argument (null pointer, bad size)        */
#define CVFES_ERR_HIP (-2)   /* HIP runtime error (launch failure, bad pointer)  */

/* Method selector for the stiffness operator. */
#define CVFES_METHOD_STORED 0   /* stored block-CSR values [nnzb][9][S]                     */
#define CVFES_METHOD_MATFREE 1  /* block values rebuilt each step from coef x khat          */

/* Description of one assembled membrane system.  Pointers are device pointers; fields a method
 * does not use may be NULL (values for METHOD_MATFREE; cptr/celem/khat/coef for METHOD_STORED). */
typedef struct cvfes_system {
    int64_t n_nodes;          /* N                                                            */
    int64_t n_smp;            /* S (samples)                                                   */
    int64_t nnzb;             /* number of 3x3 blocks                                          */
    int64_t n_elms;           /* E                                                             */
    const int32_t *indptr;    /* [N+1] block-row pointers (node adjacency, sorted columns)     */
    const int32_t *indices;   /* [nnzb] block column (node) ids                                */
    const double *values;     /* [nnzb][9][S] stored values                                    */
    const int32_t *cptr;      /* [nnzb+1] contribution ranges per block                        */
    const int32_t *celem;     /* [9E] element of each contribution                             */
    const double *khat;       /* [9E][9] unit-coefficient element block of each contribution   */
    const double *coef;       /* [E][S] stiffness coefficient per element and sample           */
    const double *lm;         /* lumped mass, [N][S] (lm_sstride = 1) or [N] (lm_sstride = 0)  */
    int64_t lm_sstride;
    const double *f1;         /* [N][3] unit-pressure nodal load                               */
    const uint8_t *fixed;     /* [N] 1 = Dirichlet node (all 3 dofs), may be NULL              */
    double bdy_u;             /* prescribed displacement of fixed nodes                        */
    /* Exact structure maxima (0 = unknown): blocks per row, element contributions per
     * off-diagonal block, element contributions per diagonal block.  They enable the row kernels
     * that issue all loads of a row up front; wrong values give wrong results. */
    int64_t max_row_blocks;
    int64_t max_offdiag_contrib;
    int64_t max_diag_contrib;
    int64_t variant;          /* CVFES_VARIANT_*                                               */
} cvfes_system_t;

/* Step kernel variants (cvfes_system_t.variant). */
#define CVFES_VARIANT_AUTO 0       /* fastest measured variant the structure allows               */
#define CVFES_VARIANT_GENERIC 1    /* block-by-block row walk, any structure and sample count      */
#define CVFES_VARIANT_ROWLOAD 2    /* all loads of a row issued up front, two samples per lane     */
#define CVFES_VARIANT_ROWLOAD_NT 3 /* as ROWLOAD, stored values streamed with non-temporal loads   */
#define CVFES_VARIANT_ROWLOAD1 4   /* as ROWLOAD with one sample per lane (two waves per row)       */
#define CVFES_VARIANT_ROWLOAD1_NT 5 /* as ROWLOAD1 with non-temporal value loads                   */

/* Time-dependent nodal load: f_n = p[n] * f1 + qa[n] * ga + qb[n] * gb  (ga/gb may be NULL).
 * ga, gb are sample-independent [N][3] traction fields (e.g. two consecutive fluid wall-stress
 * snapshots interpolated in time, physicsSolverSDampGPUs.py:287-314), zero on fixed nodes.
 * p, qa, qb are device arrays with one entry per step of the call. */
typedef struct cvfes_load {
    const double *p;
    const double *ga;
    const double *qa;
    const double *gb;
    const double *qb;
} cvfes_load_t;

/* ABI version compiled into the library (CVFES_ABI_VERSION). */
int cvfes_abi_version(void);

/* Description of the last error raised on the calling thread ("" if none). */
const char *cvfes_last_error(void);

/* Wait for all work on `stream`. */
int cvfes_stream_sync(void *stream);

/*
 * Unit-coefficient element stiffness blocks, areas and normals.
 * Replaces: CoordinateTransformation + the B / B^T D B / TT^T.TT products of
 *   optimizedSolidAssemble.pyx:82-127,265-329 and the frame/area/B of
 *   physicsSolverSDampGPUs.cl:66-118,134-150 (factorised block formula, cl:340-357).
 * Block (a,b) of element e (K_e = coef[e,s] * khat_e) is written to khat[cpos[9e+3a+b]][0..8].
 */
int cvfes_element_matrices(void *stream, int64_t n_elms, const double *nodes, const int32_t *elems,
                           double nu, double kshear, const int32_t *cpos,
                           double *khat, double *area, double *normal);

/*
 * Stiffness coefficient per element and sample: coef[e,s] = sum_q w_q h_q E_q with node fields
 * interpolated at the edge-midpoint Gauss points, or E[e,s] * sum_q w_q h_q when elm_E != NULL.
 * Replaces: physicsSolver.py:154-165,200-202 (elmGE, elmGThick, coefK).
 *   vthick  [N][S] (thick_sstride = 1) or [N] (0);  vE [N][S] (ignored when elm_E given);
 *   elm_E   [E][S] or NULL.
 */
int cvfes_element_coef(void *stream, int64_t n_elms, int64_t n_smp, const int32_t *elems,
                       const double *vthick, int64_t thick_sstride, const double *vE,
                       const double *elm_E, double *coef);

/*
 * Lumped mass and unit-pressure nodal load, pulled per node over incident elements.
 * lm[a,s] = mass_scale * density * sum_e area_e (h_a/2 + h_b/4 + h_c/4);  mass_scale = 1/9 gives the
 * CPU path's lumping (optimizedSolidAssemble.pyx:194-207,313-323 with physicsSolver.py:191-197),
 * 1/3 the consistent-mass row sum of physicsSolverSDampGPUs.cl:172-221.
 * f1[a] = sum_e area_e/3 * normal_e  (optimizedSolidAssemble.pyx:263,338-346).
 *   nptr [N+1], nent [3E] = e*3 + local index; lm is [N][S] when thick_sstride = 1 else [N].
 */
int cvfes_lumped_mass_load(void *stream, int64_t n_nodes, int64_t n_smp, const int32_t *nptr,
                           const int32_t *nent, const int32_t *elems, const double *area,
                           const double *normal, const double *vthick, int64_t thick_sstride,
                           double density, double mass_scale, double *lm, double *f1);

/*
 * Stored block-CSR values: values[p][ij][s] = sum_{c in block p} coef[celem[c], s] * khat[c][ij].
 * Replaces: the K half of assembling() in OptimizedSolidAssemble (optimizedSolidAssemble.pyx:133-188,
 *   313-346), with the sample-fastest layout instead of the reference's [nnzb][S][3][3].
 */
int cvfes_assemble_bcsr(void *stream, int64_t nnzb, int64_t n_smp, const int32_t *cptr,
                        const int32_t *celem, const double *khat, const double *coef,
                        double *values);

/*
 * y = K x for all samples, x and y [N][3][S].
 * Replaces: MultiplyByVector (optimizedSolidAssemble.pyx:1250-1274); MultiplyBy1DVector
 *   (pyx:1280-1299) is the same call with x tiled over samples.
 */
int cvfes_bcsr_spmv(void *stream, const cvfes_system_t *sys, int method, const double *x,
                    double *y);

/*
 * Start of the central-difference recurrence from u0, du0 ([N][3], sample-independent):
 *   u = u0,  ddu = (f_0 - damp du0 - K (u0 + beta du0)) / lm,  up = u0 - dt du0 + dt^2/2 ddu,
 * with f_0 the first entry of `load`.
 * Replaces: physicsSolver.py:209-229 and calc_ddu + the host u_{-1} of
 *   physicsSolverSDampGPUs.py:232-277.  work is a [N][3][S] scratch buffer.
 */
int cvfes_initial_state(void *stream, const cvfes_system_t *sys, int method, const double *u0,
                        const double *du0, const cvfes_load_t *load, double dt, double beta,
                        double damp, double *u, double *up, double *work);

/*
 * nsteps explicit central-difference steps with the stiffness product, Rayleigh stiffness
 * damping, mass-independent damping, load and Dirichlet values fused into one kernel per step:
 *   x   = u + (beta/dt)(u - up)
 *   u+  = [dt^2 (f_n - damp (u - up)/dt - K x) + lm (2u - up)] / lm     (0 where lm = 0)
 *   u+  = bdy_u on fixed nodes
 * state[3] holds {u_n, u_{n-1}, work} device pointers; on return the array is rotated so that
 * state[0] is the newest displacement and state[1] the one before.  Step n of the call uses
 * entry n of the load tables.
 * Replaces: SolidSolver.Solve (physicsSolver.py:263-292) and GPUSolidSolver.Solve
 *   (physicsSolverSDampGPUs.py:325-399: assemble_K_P + calc_u + host BC + calc_u_appTrac).
 */
int cvfes_step(void *stream, const cvfes_system_t *sys, int method, double **state,
               const cvfes_load_t *load, int64_t nsteps, double dt, double beta, double damp);

#ifdef __cplusplus
}
#endif

#endif /* CVFES_MI355X_H */
