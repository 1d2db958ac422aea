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

#define CVFES_ABI_VERSION 12

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
    /* Node -> incident element lists and, per contribution, the slot of its element in its row
     * node's list (LDS-staged matrix-free kernel; may be NULL otherwise). */
    const int32_t *nptr;      /* [N+1]                                                          */
    const int32_t *nent;      /* [3E] e*3 + local index                                         */
    const int32_t *cslot;     /* [9E]                                                           */
    int64_t max_node_elems;   /* exact maximum of nptr[a+1] - nptr[a] (0 = unknown)             */
    /* Symmetric storage of the stored operator (SYM variants; may be NULL otherwise): only the
     * blocks (a, b) with b >= a are stored, uvalues[q][ij][s]; vblk[p] is the upper entry holding
     * block p of the full structure, itself when col(p) >= row(p), else the entry of the
     * mirrored block (col, row), read transposed. */
    const double *uvalues;    /* [nnzu][9][S]                                                   */
    const int32_t *vblk;      /* [nnzb]                                                         */
    int64_t nnzu;
    /* Order in which the wave-uniform step kernels visit the block rows (a permutation of
     * 0..N-1; NULL = natural).  Any permutation gives the same result; a compact one (neighbouring
     * nodes close together) lets the second read of a mirrored block hit in L2. */
    const int32_t *row_order;
    /* Element colouring for the element-by-element variant (EBE; may be NULL otherwise): the
     * elements of colour c are color_elems[color_ptr[c] .. color_ptr[c+1]), and no two of them
     * share a node (meshColoring.py:7-63 greedy colouring).  elems is [E][3] node ids. */
    const int32_t *color_ptr;   /* [n_colors+1]                                                  */
    const int32_t *color_elems; /* [E]                                                           */
    const int32_t *elems;       /* [E][3]                                                        */
    const int32_t *cpos;        /* [9E] contribution index of element e's block (a, b)          */
    int64_t n_colors;
    int64_t max_color_elems;    /* largest colour (launch bound)                                  */
    /* Cap on the wavefronts per SIMD of the step, adjoint and product kernels that stage nothing
     * in LDS (generic, row-load, symmetric), enforced with unused LDS per workgroup: 0 = as many
     * as the registers allow, else 2..7.  Fewer rows in flight keep a larger share of their
     * gathered neighbour data in the XCD's L2 when the rows are large (many samples, long rings);
     * the results do not depend on it. */
    int64_t waves_per_simd;
    /* Row records of the symmetric kernels (may be NULL): for every visited position v (in the
     * order of row_order, or natural), record_width int32 values: the row id, its block count,
     * then the columns of its blocks and their upper-storage entries (vblk), MAXB of each, padded
     * with the row's last block, where MAXB = (record_width - 2) / 2 is the kernel's block-slot
     * count (7 or 10, as the row-load kernels choose it).  One scalar load then gives a row its
     * whole structure.  The kernels fall back to indptr/indices/vblk when the width does not
     * match; the results are the same either way. */
    const int32_t *row_records;
    int64_t record_width;
    /* Inverse lumped masses 1/lm (0 where lm = 0), same layout as lm (may be NULL).  The adjoint
     * and gradient kernels scale by them (dt^2 * minv, nu_scale * minv) instead of dividing by lm
     * per column and sample; with NULL they divide.  The forward step kernels do not read it. */
    const double *minv;
} cvfes_system_t;

/* Step kernel variants (cvfes_system_t.variant). */
#define CVFES_VARIANT_AUTO 0       /* fastest measured variant the structure allows               */
#define CVFES_VARIANT_GENERIC 1    /* block-by-block row walk, any structure and sample count      */
#define CVFES_VARIANT_ROWLOAD 2    /* all loads of a row issued up front, two samples per lane     */
#define CVFES_VARIANT_ROWLOAD_NT 3 /* as ROWLOAD, stored values streamed with non-temporal loads   */
#define CVFES_VARIANT_ROWLOAD1 4   /* as ROWLOAD with one sample per lane (two waves per row)       */
#define CVFES_VARIANT_ROWLOAD1_NT 5 /* as ROWLOAD1 with non-temporal value loads                   */
#define CVFES_VARIANT_LDS 6        /* matrix-free: row coefficients and element blocks staged in LDS */
#define CVFES_VARIANT_LDS2 7       /* as LDS with two samples per lane                               */
#define CVFES_VARIANT_SYM 8        /* stored, symmetric: upper blocks only, lower ones read
                                      transposed from their mirror (about half the value bytes)   */
#define CVFES_VARIANT_SYM_NT 9     /* as SYM, every value load non-temporal                         */
#define CVFES_VARIANT_SYM_NTL 10   /* as SYM, the mirrored (second) reads non-temporal              */
#define CVFES_VARIANT_CACHED 11    /* matrix-free: the row node's element coefficients loaded once
                                      per row (two samples per lane)                             */
#define CVFES_VARIANT_CACHED1 12   /* as CACHED with one sample per lane                            */
#define CVFES_VARIANT_EBE 13       /* matrix-free element by element: one pass per element colour
                                    * adds each element's K_e x_e into the nodes' sums (no write
                                    * races inside a colour), then a per-row update (needs the
                                    * colouring fields)                                             */

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

/* sizeof of the ABI structs as compiled (binding checks): 0 cvfes_system_t, 1 cvfes_load_t,
 * 2 cvfes_batch_mesh_t, 3 cvfes_geometry_t, 4 cvfes_update_t; -1 otherwise.  Host only. */
int64_t cvfes_struct_size(int which);

/* Description of the last error raised on the calling thread ("" if none). */
const char *cvfes_last_error(void);

/* Wait for all work on `stream`. */
int cvfes_stream_sync(void *stream);

/*
 * Streaming calibration (no counterpart in the reference): the HBM rate reached with the step
 * kernels' own load and store forms, used as the achievable ceiling next to the 8 TB/s peak and as
 * the known byte count that scales the traffic counters.  Each lane moves 16-byte items (two
 * doubles), four independent loads in flight; `nontemporal` selects non-temporal loads (the
 * stored values' form) or ordinary ones; stores are 16-byte non-temporal (the step outputs' form).
 *   cvfes_stream_read: partial[b] = sum of the workgroup's items, for b < cvfes_stream_blocks(n)
 *     (reads 8 n bytes, writes 8 bytes per workgroup);
 *   cvfes_stream_copy: dst = src (reads and writes 8 n bytes).
 * n_doubles must be even and the buffers 16-byte aligned.
 */
int64_t cvfes_stream_blocks(int64_t n_doubles);
int cvfes_stream_read(void *stream, const double *src, int64_t n_doubles, int nontemporal,
                      double *partial);
int cvfes_stream_copy(void *stream, const double *src, int64_t n_doubles, int nontemporal,
                      double *dst);

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
 *   elm_E   [E][S] or NULL; with log_modulus = 1 it holds log E (the parameter of the gradient).
 */
int cvfes_element_coef(void *stream, int64_t n_elms, int64_t n_smp, const int32_t *elems,
                       const double *vthick, int64_t thick_sstride, const double *vE,
                       const double *elm_E, int log_modulus, double *coef);

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
 * minv[i] = 1 / lm[i] (0 where lm[i] = 0) for i < n: cvfes_system_t.minv of the adjoint and
 * gradient kernels (n = N, or N*S when the masses are per sample).
 * No counterpart in the reference (the adjoint is an addition of this library).
 */
int cvfes_inverse_mass(void *stream, int64_t n, const double *lm, double *minv);

/*
 * Symmetric stored values: uvalues[q][ij][s] = values[ublk[q]][ij][s] for the nnzu upper blocks
 * (col >= row) listed by ublk.  K is symmetric (K[a][b] = K[b][a]^T), so the step reads block p of
 * the full structure as uvalues[vblk[p]], transposed when col(p) < row(p).
 * Replaces: the K half of assembling() (optimizedSolidAssemble.pyx:133-188) for the symmetric
 *   storage of the same matrix.
 */
int cvfes_assemble_upper(void *stream, int64_t nnzu, int64_t n_smp, const int32_t *ublk,
                         const int32_t *cptr, const int32_t *celem, const double *khat,
                         const double *coef, double *uvalues);

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

/* Size of the static device scratch of cvfes_initial_state_static, in doubles (64 MiB). */
#define CVFES_STATIC_SCRATCH_DOUBLES (INT64_C(8) << 20)

/*
 * cvfes_initial_state without the work buffer, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It allocates nothing: the product K (u0 + beta du0) goes to a
 * static device array of CVFES_STATIC_SCRATCH_DOUBLES doubles that every GPU the library runs on
 * holds until the process ends (its address is looked up once per device on the first call, with
 * that device current; the stream must belong to the current device).  Systems with
 * 3 * n_nodes * n_smp above that limit are rejected with CVFES_ERR_ARG.  All calls on one GPU share
 * the array, so calls that overlap on one GPU (concurrent streams or threads) race; calls on one
 * stream do not.  Nothing in the array needs to start at zero (the product writes every entry
 * before it is read).  Results equal cvfes_initial_state's bit for bit (same kernels).
 */
int cvfes_initial_state_static(void *stream, const cvfes_system_t *sys, int method,
                               const double *u0, const double *du0, const cvfes_load_t *load,
                               double dt, double beta, double damp, double *u, double *up);

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

/*
 * The step kernel cvfes_step launches for this system and method (CVFES_VARIANT_*, with AUTO and
 * variants the structure or sample count does not support resolved), or -1 for bad arguments.
 * Host only.  No counterpart in the reference (one kernel there).
 */
int64_t cvfes_step_kernel(const cvfes_system_t *sys, int method);

/*
 * One step of the adjoint recurrence of cvfes_step (no counterpart in the reference, which has no
 * gradient; it is the transpose of the step above, K being symmetric):
 *   lambda = (2-gamma) mu1 - (1-gamma) mu2 - K A (mu1 + (beta/dt)(mu1 - mu2)) + w (obs_u - obs_ref)
 *   mu_out = lambda with fixed dofs zeroed when project = 1
 * A = dt^2/lm (column scale), gamma = dt damp / lm.  obs_u/obs_ref/obs_w (weights [S]) add the
 * observation term of L = sum w/2 |u_n - u_obs|^2 and may be NULL.  Arrays are [N][3][S].
 */
int cvfes_adjoint_step(void *stream, const cvfes_system_t *sys, int method, const double *mu1,
                       const double *mu2, double *mu_out, const double *obs_u,
                       const double *obs_ref, const double *obs_w, double dt, double beta,
                       double damp, int project);

/*
 * One forward step (as cvfes_step with nsteps = 1: state = {u_n, u_{n-1}, out}, load entry 0 of
 * `load`) and one projected adjoint step (as cvfes_adjoint_step with project = 1) in one kernel
 * that reads the operator once for both; the two may be at different time steps.  Both outputs
 * equal those of the separate calls bit for bit.  Needs cvfes_step_kernel(sys, method) to be
 * CVFES_VARIANT_GENERIC (the kernel whose operations it replays) and six distinct buffers
 * (mu1 may equal mu2).  The state array is not rotated.  No counterpart in the reference (used by
 * the checkpointed reverse sweep: the recomputation of one segment runs with the adjoint steps of
 * the next).
 */
int cvfes_adjoint_fused_step(void *stream, const cvfes_system_t *sys, int method,
                             double *const *state, const cvfes_load_t *load, const double *mu1,
                             const double *mu2, double *mu_out, const double *obs_u,
                             const double *obs_ref, const double *obs_w, double dt, double beta,
                             double damp);

/*
 * Gradient with respect to theta = log E per element and sample (no counterpart in the reference):
 *   grad[e,s] = fma(-coef[e,s], sum_ab (nu_scale mu/lm)_a^T khat_e(a,b) x_b, grad[e,s]),
 *   x = x_cur + (beta/dt)(x_cur - x_prev)  (x_prev may be NULL: x = x_cur).
 * elems [E][3], cpos [9E] (block slot of element e's (a,b) in khat), grad [E][S] accumulated.
 */
int cvfes_grad_accum(void *stream, const cvfes_system_t *sys, const int32_t *elems,
                     const int32_t *cpos, const double *mu, const double *x_cur,
                     const double *x_prev, double dt, double beta, double nu_scale, double *grad);


/*
 * Reference geometry of a mesh for the kernels that work on the moved geometry X + u.
 */
typedef struct cvfes_geometry {
    const double *nodes;    /* [N][3] reference coordinates                                      */
    const int32_t *elems;   /* [E][3] element node ids                                           */
    double nu;              /* Poisson ratio                                                     */
    double kshear;          /* transverse-shear factor (5/6)                                     */
} cvfes_geometry_t;

/*
 * Per-element stress of the displacement u ([N][3][S]) on the geometry X + ug (ug [N][3][S], or
 * NULL for the reference geometry), with the element-average Young's modulus ave_E [E][S]:
 * stress[e][c][s], c = (xx, yy, xy, xz, yz) in global axes.
 * Replaces: OptimizedCalculateStress (optimizedSolidAssemble.pyx:1116-1223) as called by
 *   SolidSolver.OptimizedCalculateStress (physicsSolver.py:416-424, ug = mesh.updateNodes - X);
 *   calc_stress (physicsSolverGPUs.cl:443-589).
 */
int cvfes_element_stress(void *stream, int64_t n_elms, int64_t n_smp, const cvfes_geometry_t *geo,
                         const double *u, const double *ug, const double *ave_E, double *stress);

/*
 * Re-formation schedule of the updated-geometry mode (host memory, updated by each call):
 * before every step, if counter >= interval the geometry becomes X + (current u) and counter is
 * set to 1, otherwise counter is incremented.  counter starts at 0, so the first re-formation
 * comes before step `interval` (0-based), as in the reference.  ug: device [N][3][S] buffer for
 * the geometry displacement (needed when interval > 1; with interval 1 the current u is used).
 */
typedef struct cvfes_update {
    int64_t interval;
    int64_t counter;
    int64_t has_geometry;   /* 0 until the first re-formation (reference geometry until then)   */
    double *ug;
} cvfes_update_t;

/*
 * nsteps steps of the updated-geometry mode: as cvfes_step, with the stiffness and the
 * unit-traction load those of the geometry X + ug (never stored: every step forms each element's
 * frame on the moved geometry in registers); lumped mass, coefficients and traction fields stay
 * as assembled.  Needs sys->nptr/nent/coef/lm.  sys->variant selects the kernel form:
 * CVFES_VARIANT_GENERIC divided frame normalisation, one sample per lane; CVFES_VARIANT_ROWLOAD1
 * reciprocal normalisation (within an ulp), one sample per lane; CVFES_VARIANT_ROWLOAD and AUTO
 * reciprocal normalisation, two samples per lane for even sample counts.
 * Replaces: SolidSolver.Solve with OptimizedAssembleUpdate (physicsSolver.py:263-324;
 *   optimizedSolidAssemble.pyx:710-819) for stiffness_update_interval = interval; the K(X+u) and
 *   follower-load products of assemble_K_P (physicsSolverSDampGPUs.cl:232-391) for interval 1.
 */
int cvfes_step_updated(void *stream, const cvfes_system_t *sys, const cvfes_geometry_t *geo,
                       cvfes_update_t *upd, double **state, const cvfes_load_t *load,
                       int64_t nsteps, double dt, double beta, double damp);

/*
 * Batched ensemble (no counterpart in the reference, which advances one mesh per process): many
 * meshes of uneven size, each with its own samples, operator, time step and loads, advance together
 * in one persistent launch per step.  Host description of one mesh:
 */
typedef struct cvfes_batch_mesh {
    cvfes_system_t sys;    /* device arrays of the mesh                                      */
    int64_t method;        /* CVFES_METHOD_*                                                  */
    double *state[3];      /* state buffers; with phase p: u = state[p], u_prev = state[p+2 mod 3],
                              output = state[p+1 mod 3]                                       */
    cvfes_load_t load;     /* tables indexed by the absolute step                             */
    double dt, beta, damp;
    int64_t task0;         /* set by cvfes_batch_plan: first wave task of the mesh             */
    int64_t lanes;         /* set by cvfes_batch_plan: lanes per block row                     */
    int64_t vw;            /* samples per lane.  On input: 0 = two for even counts, 1 = one
                              everywhere, 2 = two only where the rows are wave-uniform with two;
                              set by cvfes_batch_plan to the width used (1 or 2)             */
    int64_t packed;        /* set by cvfes_batch_plan: 1 = rows packed back to back in lanes,
                              0 = each row padded to whole wavefronts                         */
    int64_t n_mtasks;      /* set by cvfes_batch_plan: wave tasks of the mesh                   */
    int64_t slice;         /* set by cvfes_batch_plan: tasks per XCD slice, ceil(n_mtasks / 8)  */
    int64_t xtask0;        /* set by cvfes_batch_plan: first item of the mesh in every XCD's list */
} cvfes_batch_mesh_t;

/* Device bytes cvfes_batch_plan writes for n_mesh meshes: the descriptors followed by a compact
 * table of their xtask0 (-1 if n_mesh < 1). */
int64_t cvfes_batch_plan_bytes(int n_mesh);

/* Number the wave tasks of `meshes` (host array, updated in place) and copy the array, followed by
 * a compact table of the meshes' xtask0, to `dev_meshes` (device, cvfes_batch_plan_bytes(n_mesh)
 * bytes).  A wave task is 64 lanes of
 * one mesh, each lane vw consecutive samples of one block row.  A mesh whose sample count fills at
 * least 3/4 of whole wavefronts gets wave-uniform rows (lanes = S/vw rounded up to 64: row data
 * through scalar loads); otherwise its rows are packed back to back (lanes = S/vw).  Every mesh is
 * cut into 8 contiguous slices, one per XCD, and each XCD walks its slice of every mesh in turn, so
 * the XCDs get equal shares of every mesh (balanced) and contiguous rows (neighbour states shared in
 * their L2).  n_tasks receives 8 x the slice items of all meshes (the size of the task space; one
 * wavefront per task is n_waves = n_tasks).  kinds receives the paths the ensemble needs (bit 0:
 * wave-uniform rows with two samples per lane, bit 1: one sample per lane, bit 2: a stored mesh,
 * bit 3: beta or damp, bit 4: packed rows with two samples per lane); pass it to cvfes_batch_step
 * and cvfes_batch_graph_create, which then launch a kernel instance with only those paths (fewer
 * registers, more waves in flight; n_waves >= n_tasks selects the instance without the persistent
 * loop).  Blocks until the copy is done. */
int cvfes_batch_plan(void *stream, cvfes_batch_mesh_t *meshes, int n_mesh, void *dev_meshes,
                     int64_t *n_tasks, int64_t *kinds);

/* One step of every mesh (one launch).  Step index for the load tables: counter[0] + k_local
 * (k_local alone when counter is NULL).  n_waves: persistent grid size in wavefronts (0 = 16 per
 * compute unit).  kinds: the value cvfes_batch_plan returned, plus the caller's launch choices:
 * bit 5 caps the kernel at 64 registers (8 waves per SIMD, spilling where the paths need more),
 * bits 6..8 hold a cap on the waves per SIMD (2..7, through unused LDS per workgroup; 0 = none),
 * bit 9 orders the waves of a mesh whose rows take more than one wavefront sample group first
 * (every row for the first 64 lanes' samples, then the next group) instead of row by row; the
 * results do not depend on it. */
int cvfes_batch_step(void *stream, const void *dev_meshes, int n_mesh, int64_t n_tasks, int phase,
                     const int64_t *counter, int64_t k_local, int64_t n_waves, int64_t kinds);

/* counter[0] += by on the stream (a kernel, so that it can sit inside a captured graph). */
int cvfes_counter_add(void *stream, int64_t *counter, int64_t by);

/* Capture block_steps (a multiple of 3) batched steps starting at phase0, followed by
 * counter += block_steps, into an executable HIP graph returned in *graph. */
int cvfes_batch_graph_create(const void *dev_meshes, int n_mesh, int64_t n_tasks, int64_t *counter,
                             int64_t block_steps, int phase0, int64_t n_waves, int64_t kinds,
                             void **graph);
int cvfes_graph_launch(void *graph, void *stream);
int cvfes_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif

#endif /* CVFES_MI355X_H */
