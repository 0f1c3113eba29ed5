/*
 * CPU oracle for the SSIDS numeric factorize/solve path -- TEST
 * INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline; the
 * product library never links or calls it.
 *
 * A plain-C++ restatement of the reference CPU numeric subtree with
 * pivot_method = TPP (src/ssids/cpu/factor.hxx:58-100 with PivotMethod::tpp):
 * multifrontal assembly (src/ssids/cpu/kernels/assemble.hxx:140-443),
 * threshold partial pivoting LDL^T (src/ssids/cpu/kernels/ldlt_tpp.cxx:166-270),
 * the contribution update (src/ssids/cpu/factor.hxx:82-96), inertia counting
 * (src/ssids/cpu/NumericSubtree.hxx:246-280), solves
 * (src/ssids/cpu/NumericSubtree.hxx:286-406, ldlt_tpp.cxx:272-317) and
 * enquire (src/ssids/cpu/NumericSubtree.hxx:424-468).
 *
 * oracle_factor_app restates the device factorization's pivoting instead:
 * the front layout the device assembles (delayed columns first) and a
 * posteriori pivoting in 64-column blocks (complete pivoting inside a block,
 * src/ssids/cpu/kernels/block_ldlt.hxx:289-414; the |l| <= 1/u test of
 * src/ssids/cpu/kernels/ldlt_app.cxx:304-322), then threshold partial
 * pivoting on the failed columns.
 */
#ifndef SSIDS_ORACLE_H
#define SSIDS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stats out (int64): [0] flag [1] num_delay [2] num_neg [3] num_two
 * [4] num_zero [5] maxfront [6] not_first_pass [7] num_factor
 * [8] num_flops [9] matrix_rank */
void *oracle_factor(int n, int nnodes, const int *sptr, const int *sparent,
                    const int64_t *rptr, const int *rlist, const int64_t *nptr,
                    const int64_t *nlist, const double *aval,
                    const double *scaling, int action, double u, double small,
                    int64_t *stats);
/* The same on the first max_nodes nodes in postorder only (every child
 * precedes its parent, so this is a prefix of the same factorization): a
 * bounded CPU sample of a large workload.  stats as above over those nodes;
 * the handle is for timing only (no solves). */
void *oracle_factor_prefix(int n, int nnodes, const int *sptr, const int *sparent,
                           const int64_t *rptr, const int *rlist, const int64_t *nptr,
                           const int64_t *nlist, const double *aval,
                           const double *scaling, int action, double u, double small,
                           int max_nodes, int64_t *stats);
/* The same factorization with the device's pivoting (APP, then TPP on the
 * failed columns) and front layout; stats as oracle_factor, not_first_pass
 * counting the columns APP failed. */
void *oracle_factor_app(int n, int nnodes, const int *sptr, const int *sparent,
                        const int64_t *rptr, const int *rlist, const int64_t *nptr,
                        const int64_t *nlist, const double *aval, const double *scaling,
                        int action, double u, double small, int64_t *stats);
/* per-node pivots accepted by APP (nnodes; equal to nelim without APP) */
void oracle_node_app(void *h, int *nelim_app);
/* job: 0 all, 1 fwd, 2 diag, 3 bwd, 4 diag_bwd; x (ldx,nrhs) in elimination
 * order (as the subtree sees it). */
void oracle_solve(void *h, int job, int nrhs, double *x, int ldx);
/* piv_order(n) 0-based signed as the subtree produces; d(2,n). */
void oracle_enquire(void *h, int *piv_order, double *d);
/* per-node nelim and ndelay_in (arrays of nnodes) */
void oracle_node_info(void *h, int *nelim, int *ndelay_in);
/* factors of one node: info = m, n, ndelay_in, nelim, ld, sn; L (ld x n),
 * D (2n), perm (n, 0-based); arrays may be NULL */
int oracle_node_factors(void *h, int node, int *info, double *L, double *D, int *perm);
void oracle_free(void *h);

#ifdef __cplusplus
}
#endif
#endif
