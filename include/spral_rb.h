/*
 * spral_rb.h -- Rutherford-Boeing sparse matrix files (host code), the
 * format of the reference's regression matrices.  Replaces rb_peek, rb_write
 * and rb_read of src/rutherford_boeing.f90 for assembled real, integer and
 * pattern matrices; return codes as that module (:27-38): 0 success, -1 bad
 * file, -2 not RB, -3 I/O, -4 complex type, -5 elemental file, -6 bad
 * matrix_type, -11 bad lwr_upr_full, -13 values not built, +1 auxiliary
 * values file.  Indices are 1-based as in the reference.
 */
#ifndef SPRAL_RB_H
#define SPRAL_RB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rb_peek (src/rutherford_boeing.f90:75-233).  Any output may be NULL;
 * type_code receives 3 characters, title 72, identifier 8 (no NUL). */
int spral_rb_peek(const char *filename, int *m, int *n, int64_t *nelt, int64_t *nvar,
                  int64_t *nval, int *matrix_type, char *type_code, char *title,
                  char *identifier);

/* rb_write (src/rutherford_boeing.f90:635-732), byte for byte.  val NULL
 * writes a pattern file; val_format NULL is the reference default
 * "(3e24.16)"; title / identifier NULL are "Matrix" / "0". */
int spral_rb_write(const char *filename, int matrix_type, int m, int n, const int64_t *ptr,
                   const int *row, const double *val, const char *val_format,
                   const char *title, const char *identifier);

/* rb_read (src/rutherford_boeing.f90:274-575) with values 0 (as in the
 * file) or 1 (pattern only), lwr_upr_full 1 (lower, as stored), 2 (upper)
 * or 3 (both), add_diagonal 0/1.  Returns a handle (NULL with *info < 0)
 * holding the matrix; spral_rb_fetch copies it out (ptr n+1, row and val
 * nz) and frees the handle. */
void *spral_rb_read(const char *filename, int lwr_upr_full, int values, int add_diagonal,
                    int *m, int *n, int64_t *nz, int *has_val, int *matrix_type, int *info);
void spral_rb_fetch(void *handle, int64_t *ptr, int *row, double *val);

#ifdef __cplusplus
}
#endif

#endif /* SPRAL_RB_H */
