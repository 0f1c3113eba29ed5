/*
 * Matching-based scaling and ordering of symmetric matrices (host code in
 * libspral_ssids_hip.so).  Lower-triangle CSC input with 1-based ptr (n+1,
 * int64) and row, as the reference's src/scaling.f90 and
 * src/match_order.f90 take it.
 */
#ifndef SPRAL_SCALING_H
#define SPRAL_SCALING_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces hungarian_scale_sym (src/scaling.f90:134-170; algorithm
 * hungarian_wrapper :597-801, hungarian_match :938-1171): maximum-product
 * matching of |A| and the symmetric scaling from its duals, so that the
 * matched entries of S A S are 1 in modulus and no entry exceeds 1.
 * scaling (n) receives the scaling; match (n, may be NULL) the matched
 * column of each row (1-based, 0 unmatched); *matched the cardinality.
 * scale_if_singular as hungarian_options%scale_if_singular.  Returns 0,
 * 1 (WARNING_SINGULAR, structurally singular and scale_if_singular),
 * -2 (ERROR_SINGULAR, structurally singular otherwise; the scaling is still
 * written, as the reference does), -1 (allocation) or -3 (n < 0). */
int spral_hungarian_scale_sym(int n, const int64_t *ptr, const int *row, const double *val,
                              double *scaling, int *match, int scale_if_singular,
                              int *matched);

/* Replaces auction_scale_sym (src/scaling.f90:269-309; auction_match
 * :1481-1586, auction_match_core :1328-1466): the approximate matching of
 * the auction algorithm, one column bidding at a time, and the symmetric
 * scaling from its prices.  max_iterations, max_unchanged (3),
 * min_proportion (3, single precision) and eps_initial (single precision)
 * are auction_options (defaults 30000, {10, 100, 100}, {0.90, 0.0, 0.0},
 * 0.01).  match (n, may be NULL): matched column of each row, 1-based, 0
 * unmatched; *matched, *iterations and *unmatchable as auction_inform.
 * Returns 0, -1 (allocation) or -3 (n < 0). */
int spral_auction_scale_sym(int n, const int64_t *ptr, const int *row, const double *val,
                            double *scaling, int *match, int max_iterations,
                            const int *max_unchanged, const float *min_proportion,
                            float eps_initial, int *matched, int *iterations,
                            int *unmatchable);

/* Matching-based elimination order, replacing match_order_metis
 * (src/match_order.f90:135-208; mo_match :495-629, mo_split :220-396) as
 * ssids_analyse calls it for options%ordering = 2
 * (src/ssids/ssids.f90:318-340): the maximum-product matching of |A| split
 * into 1- and 2-cycles, matched pairs merged, the compressed graph ordered
 * by approximate minimum degree (standing in for METIS), each pair given
 * consecutive positions.  order (n) receives the pivot position of each
 * variable (1-based); scale (n) the matching scaling (what
 * options%scaling = 3 applies; 0 for an unmatched index of a structurally
 * singular matrix, as the reference).  Returns 0, 1 (structurally singular:
 * SSIDS_WARNING_ANAL_SINGULAR), -1 (allocation) or -2 (n < 0). */
int spral_match_order(int n, const int64_t *ptr, const int *row, const double *val,
                      int *order, double *scale);

#ifdef __cplusplus
}
#endif
#endif
