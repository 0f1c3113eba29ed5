/*
 * spral_ssids_hip.h -- C-ABI of the MI355X (gfx950) numeric factorize/solve
 * backend for SSIDS.
 *
 * The entry points replace the GPU numeric subtree of the reference:
 * the Fortran host (src/ssids/fkeep.F90 inner_factor/inner_solve) reaches a
 * subtree only through the type-bound procedures of
 * symbolic_subtree_base/numeric_subtree_base (src/ssids/subtree.f90:26-119).
 * Each function below is what one of those procedures binds to through
 * iso_c_binding.  Plain pointers and sizes only: no HIP or torch types.
 *
 * Conventions (identical to the reference):
 *  - sptr, sparent, rptr, rlist, nptr, nlist are 1-based Fortran values;
 *  - nlist(2,k) is the 1-based column-major position of A-entry nlist(1,k)
 *    within the un-delayed front (src/ssids/anal.F90:1220,1232);
 *  - x arrives permuted into elimination order, column-major x(ldx,nrhs)
 *    (src/ssids/fkeep.F90:252-266);
 *  - flags follow src/ssids/datatypes.f90:25-59 (negative = error).
 *    Device failures use SSIDS_ERROR_CUDA_UNKNOWN (-51) with the HIP error
 *    code in stats.hip_error.
 */
#ifndef SPRAL_SSIDS_HIP_H
#define SPRAL_SSIDS_HIP_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interoperable subset of ssids_options (src/ssids/datatypes.f90:188-284),
 * the fields the numeric path honours.  Mirrors cpu_factor_options
 * (src/ssids/cpu/cpu_iface.f90:21-31) plus the fields the reference GPU path
 * ignored (u, small, pivot_method, failed_pivot_method are honoured here). */
struct spral_ssids_hip_options {
  int print_level;
  int action;              /* .true. = continue past singularity (+7)       */
  double small;            /* zero-pivot threshold (default 1e-20)          */
  double u;                /* relative pivot threshold (default 0.01)       */
  double multiplier;       /* pool slack for delays (default 1.1)           */
  int pivot_method;        /* 1 app_aggressive, 2 app_block, 3 tpp          */
  int failed_pivot_method; /* 1 tpp, 2 pass                                 */
  int schur_precision;     /* 0 fp64, 1 fp32, 2 fp16 trailing updates       */
  int refine_max;          /* max FP64 refinement steps in solve_refine     */
  double refine_tol;       /* target scaled backward error                  */
  int timing;              /* 1: time the gather/Schur launches with events */
  int level_schedule;      /* tree levels of the factorization (no reference *
                            * counterpart): 0 the cheapest of 1-3 by steps   *
                            * in sequence, 1 height from the leaves, 2       *
                            * greedy by urgency, 3 layered heavy fronts      */
};

/* Interoperable subset of ssids_inform, as cpu_factor_stats
 * (src/ssids/cpu/cpu_iface.f90:39-51) plus device diagnostics. */
struct spral_ssids_hip_stats {
  int flag;
  int num_delay;
  int64_t num_factor;
  int64_t num_flops;
  int num_neg;
  int num_two;
  int num_zero;
  int maxfront;
  int maxsupernode;
  int not_first_pass;
  int not_second_pass;
  int hip_error;           /* hipError_t of the first failing call, else 0  */
  int num_levels;          /* tree levels walked                            */
  double factor_ms;        /* device time of the numeric factorization      */
  double kernel_ms;        /* options.timing: summed device time of the     *
                            * gather launches (static plan: the rest-of-    *
                            * level gathers, tiles outside the first 64     *
                            * columns, timed in the production sequence;    *
                            * level-by-level path: whole-level gathers)     */
  double kernel_flops;     /* algorithmic flops of those launches           */
  double kernel_bytes;     /* algorithmic HBM bytes of those launches       */
  int kernel_launches;     /* number of gather launches timed               */
  int64_t pool_peak_bytes; /* peak device pool use                          */
  int static_plan;         /* 1: ran on the static plan (no per-level host
                              round trip), 0: level by level              */
};

/* Fill *opt with the reference defaults (src/ssids/datatypes.f90:188-284). */
void spral_ssids_hip_default_options(struct spral_ssids_hip_options *opt);

/* ---------------------------------------------------------------------------
 * Host analyse helpers.  The reference analyse is upstream of the numeric
 * path and is reused unchanged by a Fortran host; these restate it so the
 * numeric path can be driven standalone (tests, benchmark, service).
 * ------------------------------------------------------------------------- */

/* Restates basic_analyse (src/core_analyse.f90:38-151) followed by build_map
 * (src/ssids/anal.F90:1137-1239) and the unused-variable marking of
 * analyse_phase (src/ssids/anal.F90:1006-1014).  Input: lower-triangle CSC
 * (ptr 1-based, n+1 entries; row 1-based), order(i) = pivot position of i.
 * On exit order holds the final pivot order (0 for unused variables).
 * Returns an opaque handle (NULL on failure, *flag set) that owns the symbolic
 * arrays; query them with spral_ssids_hip_analyse_info/_arrays. */
void *spral_ssids_hip_analyse(int n, const int64_t *ptr, const int *row,
                              int *order, int nemin, int *flag);
void spral_ssids_hip_analyse_info(const void *akeep, int *nnodes,
                                  int64_t *rlist_len, int64_t *nlist_len,
                                  int64_t *num_factor, int64_t *num_flops,
                                  int *maxfront, int *maxsupernode);
void spral_ssids_hip_analyse_arrays(const void *akeep, int *sptr, int *sparent,
                                    int64_t *rptr, int *rlist, int64_t *nptr,
                                    int64_t *nlist);
void spral_ssids_hip_analyse_free(void *akeep);

/* Orderings standing in for METIS (src/metis5_wrapper.F90:102-182). */
void spral_ssids_hip_order_grid_nd(int nx, int ny, int nz, int *order);
int spral_ssids_hip_order_min_degree(int n, const int64_t *ptr, const int *row,
                                     int *order);
/* Approximate minimum degree (Amestoy, Davis, Duff 1996) on the quotient
 * graph, for large general patterns: lower CSC (1-based ptr, row) in,
 * order(i) = pivot position of i (1-based) out.  Returns 0. */
int spral_ssids_hip_order_amd(int n, const int64_t *ptr, const int *row, int *order);

/* ---------------------------------------------------------------------------
 * Symbolic subtree: replaces construct_gpu_symbolic_subtree
 * (src/ssids/gpu/subtree.f90:76-160).  Nodes sa..en-1 (1-based) of the
 * assembly tree; the whole tree is sa=1, en=nnodes+1.  Uploads rlist, the
 * child->parent row maps and nlist to the device.  Returns NULL on failure.
 * ------------------------------------------------------------------------- */
void *spral_ssids_hip_create_symbolic(int device, int n, int sa, int en,
                                      const int *sptr, const int *sparent,
                                      const int64_t *rptr, const int *rlist,
                                      const int64_t *nptr, const int64_t *nlist,
                                      const struct spral_ssids_hip_options *options,
                                      int *flag);
/* Replaces symbolic cleanup (src/ssids/gpu/subtree.f90:236-302). */
void spral_ssids_hip_destroy_symbolic(void *symbolic);

/* ---------------------------------------------------------------------------
 * Numeric factorization: replaces gpu_symbolic_subtree%factor
 * (src/ssids/gpu/subtree.f90:304-460).  aval is the host array of lower
 * triangle values (as passed to ssids_factor), scaling may be NULL.
 * Returns the numeric handle (factors resident in HBM) or NULL with
 * stats->flag < 0.  posdef = true factors A = L L^T as factor_node_posdef
 * (src/ssids/cpu/factor.hxx:139-171): pivots in order, all 1x1; a pivot
 * that is not positive fails with stats->flag = -6 (SSIDS_ERROR_NOT_POS_DEF).
 * The factors are held as L D L^T with L unit lower (L_chol = L D^{1/2}).
 * ------------------------------------------------------------------------- */
void *spral_ssids_hip_factor(bool posdef, void *symbolic, const double *aval,
                             int64_t naval, const double *scaling,
                             const struct spral_ssids_hip_options *options,
                             struct spral_ssids_hip_stats *stats);
/* As spral_ssids_hip_factor with aval/scaling already in device memory
 * (no PCIe traffic; the resident-factor path). */
void *spral_ssids_hip_factor_device(bool posdef, void *symbolic,
                                    const double *d_aval, int64_t naval,
                                    const double *d_scaling,
                                    const struct spral_ssids_hip_options *options,
                                    struct spral_ssids_hip_stats *stats);
/* Re-factor new values on an existing numeric handle of the same pattern,
 * reusing its device pool (serving refactor).  Returns flag. */
int spral_ssids_hip_refactor_device(void *numeric, const double *d_aval,
                                    const double *d_scaling,
                                    const struct spral_ssids_hip_options *options,
                                    struct spral_ssids_hip_stats *stats);
/* Replaces numeric cleanup (src/ssids/gpu/subtree.f90:462-520). */
void spral_ssids_hip_destroy_numeric(void *numeric);

/* ---------------------------------------------------------------------------
 * Solves with host x (replace gpu_numeric_subtree%solve_fwd/_diag/_diag_bwd/
 * _bwd, src/ssids/gpu/subtree.f90:555-738).  All RHS are solved together.
 * Return a flag (0 success).  For a positive-definite factorization fwd
 * solves L_chol y = x, bwd and diag_bwd L_chol^T y = x, and diag does
 * nothing, as the CPU subtree (src/ssids/cpu/NumericSubtree.hxx:342-343);
 * ssids_solve itself rejects jobs 2 and 4 then (src/ssids/ssids.f90:1237).
 * ------------------------------------------------------------------------- */
int spral_ssids_hip_solve_fwd(void *numeric, int nrhs, double *x, int ldx);
int spral_ssids_hip_solve_diag(void *numeric, int nrhs, double *x, int ldx);
int spral_ssids_hip_solve_diag_bwd(void *numeric, int nrhs, double *x, int ldx);
int spral_ssids_hip_solve_bwd(void *numeric, int nrhs, double *x, int ldx);

/* Device-resident solve: d_x in HBM, job as SSIDS_SOLVE_JOB_* (0 all,
 * 1 fwd, 2 diag, 3 bwd, 4 diag_bwd).  stream is a hipStream_t passed as void*
 * (NULL = the handle's own stream).  Asynchronous with respect to the host. */
int spral_ssids_hip_solve_device(void *numeric, int job, int nrhs, double *d_x,
                                 int ldx, void *stream);

/* Allocation-free variant of spral_ssids_hip_solve_device (no counterpart in
 * the reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; the solve workspace is a fixed 64 MiB device
 * array inside the library (one per GPU, held until the process ends)
 * instead of the handle's own buffer, and no allocating call is made.
 * Returns SPRAL_SSIDS_HIP_ERROR_SCRATCH_TOO_SMALL, before any launch, when
 * the solve needs more than SPRAL_SSIDS_HIP_SCRATCH_DOUBLES doubles
 * (spral_ssids_hip_solve_workspace_doubles(numeric, nrhs) gives the need).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (several streams or host threads) race. */
#define SPRAL_SSIDS_HIP_SCRATCH_DOUBLES (8 << 20)
#define SPRAL_SSIDS_HIP_ERROR_SCRATCH_TOO_SMALL (-61)
int spral_ssids_hip_solve_device_noalloc(void *numeric, int job, int nrhs, double *d_x,
                                         int ldx, void *stream);
int64_t spral_ssids_hip_solve_workspace_doubles(const void *numeric, int nrhs);

/* FP64 iterative refinement (no reference counterpart: the reference solve has
 * none).  x = solve(b), then x += solve(b - A x) with A the factorized values
 * (scaled system when a scaling was given) until every column's scaled
 * backward error max|b-Ax| / (||A||inf ||x||inf + ||b||inf)
 * (tests/ssids/ssids.f90:1762-1810) is <= tol, or maxit corrections.
 * bwd (host, nrhs) and *iters receive the final state.  b, x in elimination
 * order.  _device: b and x are distinct buffers in HBM.
 * Returns 0 when every column reached tol, SPRAL_SSIDS_HIP_WARNING_NOT_CONVERGED
 * when some column is still above tol after maxit corrections (x holds the
 * last iterate), SPRAL_SSIDS_HIP_ERROR_NONFINITE when a residual, solution or
 * right-hand side holds a NaN or infinity (bwd of that column is NaN), and a
 * negative SSIDS flag on other errors. */
#define SPRAL_SSIDS_HIP_WARNING_NOT_CONVERGED 60
#define SPRAL_SSIDS_HIP_ERROR_NONFINITE (-60)
int spral_ssids_hip_solve_refine(void *numeric, int nrhs, const double *b, double *x,
                                 int ldx, int maxit, double tol, double *bwd, int *iters);
int spral_ssids_hip_solve_refine_device(void *numeric, int nrhs, const double *d_b,
                                        double *d_x, int ldx, int maxit, double tol,
                                        double *bwd, int *iters, void *stream);

/* ---------------------------------------------------------------------------
 * Enquiry / alteration: replace gpu_numeric_subtree%enquire_indef and %alter
 * (src/ssids/gpu/subtree.f90:740-844, 909-981).  piv_order is 0-based as
 * produced by the subtree (the Fortran API adds 1, src/ssids/ssids.f90:1343);
 * d is 2 x n in pivot order.  Either pointer may be NULL.
 * ------------------------------------------------------------------------- */
int spral_ssids_hip_enquire_indef(const void *numeric, int *piv_order,
                                  double *d);
int spral_ssids_hip_alter(void *numeric, const double *d);
/* Positive-definite factorization: d[i] = diagonal of the Cholesky factor in
 * pivot order; replaces gpu_numeric_subtree%enquire_posdef
 * (src/ssids/gpu/subtree.f90:702-738, values as the CPU subtree's
 * src/ssids/cpu/NumericSubtree.hxx:424-433).  Returns -13
 * (SSIDS_ERROR_NOT_LLT) for an indefinite factorization; enquire_indef and
 * alter return -14 (SSIDS_ERROR_NOT_LDLT) for a positive-definite one. */
int spral_ssids_hip_enquire_posdef(const void *numeric, double *d);

/* Serialised factors (no reference counterpart: the reference's only on-disk
 * state is rb_dump of the matrix, src/ssids/ssids.f90:883-896).  save writes
 * the resident factors of `numeric` (L, D^{-1}, pivot order, the factorized
 * values and scaling kept for refinement, stats) to `path`; load rebuilds a
 * numeric handle on `symbolic`, which must be an analyse of the same pattern
 * with the same ordering (checked by a fingerprint of the row lists), whose
 * solves, enquiries and refinements then equal the original's bit for bit.
 * save returns 0 or -1 (I/O); load returns NULL with *flag -1 (I/O or a
 * malformed file), -2 (another analyse), -51 (device) or -50 (host memory). */
int spral_ssids_hip_save_numeric(const void *numeric, const char *path);
void *spral_ssids_hip_load_numeric(void *symbolic, const char *path,
                                   struct spral_ssids_hip_stats *stats, int *flag);
/* The scaling a numeric handle factorized with (n, elimination order) into s:
 * returns 1, or 0 when it has none (-51 on a device error). */
int spral_ssids_hip_numeric_scaling(const void *numeric, double *s);

/* Contribution block of the subtree root for a parent subtree held elsewhere;
 * same out-parameters as spral_ssids_cpu_subtree_get_contrib_dbl
 * (src/ssids/cpu/NumericSubtree.cxx:233-261).  Host memory owned by the
 * numeric handle until spral_ssids_hip_free_contrib.  Returns 0, or -51
 * (device error) / -50 (host allocation) with every out-parameter empty. */
int spral_ssids_hip_get_contrib(void *numeric, int *n, const double **val,
                                int *ldval, const int **rlist, int *ndelay,
                                const int **delay_perm,
                                const double **delay_val, int *lddelay);
void spral_ssids_hip_free_contrib(void *numeric);

/* Diagnostic accessor (no reference counterpart; used by parity tests):
 * factors of node `node` (0-based postorder index).  info[6] receives
 * m, n, ndelay_in, nelim, ldl, symbolic ncol.  L (ldl x n), D (2n) and
 * perm (n, 0-based elimination indices) may be NULL. */
int spral_ssids_hip_node_factors(const void *numeric, int node, int *info, double *L,
                                 double *D, int *perm);

/* Diagnostic accessor (no reference counterpart; used by parity tests): per
 * node (0-based postorder, arrays of nnodes), the pivots the a posteriori
 * pass accepted (nelim_app; the rest of nelim came from threshold partial
 * pivoting on the failed columns) and the delayed columns that arrived from
 * the children.  Returns 0. */
int spral_ssids_hip_node_pivots(const void *numeric, int *nelim_app, int *ndelay_in);

/* Planning helper (no reference counterpart; the factorization runs the
 * fronts of one tree level per launch sequence): the level of each of nnodes
 * fronts (0-based postorder; parent[i] > i, or -1 for a root; ncols = the
 * front's fully-summed columns) under schedule mode 0-3, as the
 * level_schedule option (0: the assignment the analyse uses by default).
 * Returns the number of levels, or -1 when parent is not a postorder forest
 * or mode is out of range. */
int spral_ssids_hip_tree_levels(int nnodes, const int *parent, const int *ncols, int mode,
                                int *level);

/* Number of visible HIP devices (replaces the cudaGetDeviceCount probe of
 * src/hw_topology/guess_topology.cxx:65). */
int spral_ssids_hip_device_count(void);

/* ---------------------------------------------------------------------------
 * Scaling.  Replaces equilib_scale_sym (src/scaling.f90:381-416; algorithm
 * inf_norm_equilib_sym, :480-521), which ssids_factor runs for
 * options%scaling >= 4 (src/ssids/ssids.f90:1003-1028): symmetric
 * infinity-norm equilibration of the lower-triangle CSC (ptr 1-based, n+1
 * entries; row 1-based), computed on the device.  scaling (n, original
 * variable order) receives the result; max_iterations and tol are
 * equilib_options (defaults 10 and 1e-8, tol a single-precision value as in
 * the reference); *iterations receives equilib_inform%iterations.  Returns 0,
 * -1 for an invalid argument or pattern (checked on the host before any
 * launch), -2 when device memory runs out, -3 on a HIP error.  The result is
 * bit-identical to the reference's.
 * ------------------------------------------------------------------------- */
int spral_ssids_hip_equilib_scale_sym(int n, const int64_t *ptr, const int *row,
                                      const double *val, double *scaling, int max_iterations,
                                      float tol, int *iterations);
/* As spral_ssids_hip_equilib_scale_sym with ptr, row, val and scaling in
 * device memory (the resident-pattern path).  The pattern is not checked on
 * the host here; row indices outside [1, n] are skipped on the device. */
int spral_ssids_hip_equilib_scale_sym_device(int n, const int64_t *d_ptr, const int *d_row,
                                             const double *d_val, double *d_scaling,
                                             int max_iterations, float tol, int *iterations);

#ifdef __cplusplus
}
#endif

#endif /* SPRAL_SSIDS_HIP_H */
