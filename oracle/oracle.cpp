// CPU oracle: restatement of the reference CPU numeric subtree with
// threshold partial pivoting.  TEST INFRASTRUCTURE ONLY -- see oracle.h.
//
// Front layout follows the reference CPU (src/ssids/cpu/kernels/assemble.hxx:
// 227-270): the node's own fully-summed columns, then delayed columns from the
// children in child order, then the remaining rows.  Each child's whole
// trailing matrix (failed columns + contribution block) is added into the
// parent front before the parent is factorized.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

namespace {

struct SingularError {};

struct Node {
  int sn = 0, sm = 0;   // symbolic ncol / nrow
  int nd = 0;           // ndelay_in
  int m = 0, n = 0;     // front rows / fully-summed columns
  int nelim = 0;
  int nelim_app = 0;      // pivots accepted by APP (APP mode), else nelim
  std::vector<double> f;  // m x m column-major (lower used)
  std::vector<double> d;  // 2*n, D^{-1} in reference format
  std::vector<int> perm;  // n fully-summed variables (1-based pivot indices)
  const int *rows = nullptr;  // symbolic row list (1-based)
};

struct Oracle {
  int n = 0, nnodes = 0;
  std::vector<Node> nodes;
};

// ---------------- TPP kernel (src/ssids/cpu/kernels/ldlt_tpp.cxx) ----------

inline double &at(double *a, int lda, int r, int c) { return a[(size_t)c * lda + r]; }

bool col_small(int idx, int from, int to, const double *a, int lda, double small) {
  for (int c = from; c < idx; ++c)
    if (!(std::fabs(a[(size_t)c * lda + idx]) < small)) return false;
  for (int r = idx; r < to; ++r)
    if (!(std::fabs(a[(size_t)idx * lda + r]) < small)) return false;
  return true;
}

// column index of the largest |a(row, from..to-1)| (row fixed)
int row_argmax(int row, int from, int to, const double *a, int lda) {
  if (from >= to) return -1;
  int best = from;
  double bv = std::fabs(a[(size_t)from * lda + row]);
  for (int c = from + 1; c < to; ++c) {
    double v = std::fabs(a[(size_t)c * lda + row]);
    if (v > bv) { bv = v; best = c; }
  }
  return best;
}

// symmetric interchange of variables c1 and c2 in the lower triangle of an
// m-row front; perm swapped alongside.
void sym_swap(int c1, int c2, int m, int *perm, double *a, int lda) {
  if (c1 == c2) return;
  if (c2 < c1) std::swap(c1, c2);
  std::swap(perm[c1], perm[c2]);
  for (int c = 0; c < c1; ++c) std::swap(at(a, lda, c1, c), at(a, lda, c2, c));
  for (int i = c1 + 1; i < c2; ++i) std::swap(at(a, lda, i, c1), at(a, lda, c2, i));
  for (int r = c2 + 1; r < m; ++r) std::swap(at(a, lda, r, c1), at(a, lda, r, c2));
  std::swap(at(a, lda, c1, c1), at(a, lda, c2, c2));
}

// max |entry| in row/col `col` of the uneliminated part, skipping the
// diagonal and position `skip`
double rc_max_skip(int col, int nelim, int m, const double *a, int lda, int skip) {
  double best = 0.0;
  for (int c = nelim; c < col; ++c)
    if (c != skip) best = std::max(best, std::fabs(a[(size_t)c * lda + col]));
  for (int r = col + 1; r < m; ++r)
    if (r != skip) best = std::max(best, std::fabs(a[(size_t)col * lda + r]));
  return best;
}

// 2x2 stability and threshold test; fills d[0..3] with the inverse.
bool good_2x2(int t, int p, double maxt, double maxp, const double *a, int lda,
              double u, double small, double *d) {
  double a11 = a[(size_t)t * lda + t];
  double a21 = a[(size_t)t * lda + p];
  double a22 = a[(size_t)p * lda + p];
  double maxpiv = std::max(std::fabs(a11), std::max(std::fabs(a21), std::fabs(a22)));
  if (maxpiv < small) return false;
  double detscale = 1 / maxpiv;
  double det0 = (a11 * detscale) * a22;
  double det1 = (a21 * detscale) * a21;
  double det = det0 - det1;
  if (std::fabs(det) < std::max(small, std::max(std::fabs(det0 / 2), std::fabs(det1 / 2))))
    return false;
  d[0] = (a22 * detscale) / det;
  d[1] = (-a21 * detscale) / det;
  d[2] = std::numeric_limits<double>::infinity();
  d[3] = (a11 * detscale) / det;
  if (std::max(maxp, maxt) < small) return true;
  double x1 = std::fabs(d[0]) * maxt + std::fabs(d[1]) * maxp;
  double x2 = std::fabs(d[1]) * maxt + std::fabs(d[3]) * maxp;
  return u * std::max(x1, x2) < 1.0;
}

// eliminate 1x1 pivot at position k: scale column, rank-1 update of the
// fully-summed columns k+1..n-1 (rows k+1..m-1)
void elim_1x1(int k, int m, int n, double *a, int lda, double *d, double *ld) {
  double d11 = d[2 * k];
  double *a1 = &a[(size_t)k * lda];
  a1[k] = 1.0;
  for (int r = k + 1; r < m; ++r) { ld[r] = a1[r]; a1[r] *= d11; }
  for (int c = k + 1; c < n; ++c) {
    double s = ld[c];
    double *ac = &a[(size_t)c * lda];
    for (int r = c; r < m; ++r) ac[r] -= a1[r] * s;
  }
}

void elim_2x2(int k, int m, int n, double *a, int lda, double *d, double *ld,
              double *ld2) {
  double *a1 = &a[(size_t)k * lda];
  double *a2 = &a[(size_t)(k + 1) * lda];
  a1[k] = 1.0;
  a1[k + 1] = 0.0;
  a2[k + 1] = 1.0;
  double d11 = d[2 * k], d21 = d[2 * k + 1], d22 = d[2 * k + 3];
  for (int r = k + 2; r < m; ++r) {
    ld[r] = a1[r];
    ld2[r] = a2[r];
    a1[r] = d11 * ld[r] + d21 * ld2[r];
    a2[r] = d21 * ld[r] + d22 * ld2[r];
  }
  for (int c = k + 2; c < n; ++c) {
    double s1 = ld[c], s2 = ld2[c];
    double *ac = &a[(size_t)c * lda];
    for (int r = c; r < m; ++r) ac[r] -= a1[r] * s1 + a2[r] * s2;
  }
}

void zero_col(int col, int m, double *a, int lda) {
  for (int r = col; r < m; ++r) a[(size_t)col * lda + r] = 0.0;
}

// ldlt_tpp_factor restated: returns number eliminated among the first n cols
int tpp_factor(int m, int n, int *perm, double *a, int lda, double *d, bool action,
               double u, double small, int start = 0) {
  std::vector<double> ld(m), ld2(m);
  int nelim = start;
  while (nelim < n) {
    if (col_small(nelim, nelim, m, a, lda, small)) {
      if (!action) throw SingularError();
      zero_col(nelim, m, a, lda);
      d[2 * nelim] = 0.0;
      d[2 * nelim + 1] = 0.0;
      ++nelim;
      continue;
    }
    int p;
    for (p = nelim + 1; p < n; ++p) {
      if (col_small(p, nelim, m, a, lda, small)) {
        if (!action) throw SingularError();
        sym_swap(p, nelim, m, perm, a, lda);
        zero_col(nelim, m, a, lda);
        d[2 * nelim] = 0.0;
        d[2 * nelim + 1] = 0.0;
        ++nelim;
        break;
      }
      int t = row_argmax(p, nelim, p, a, lda);
      double maxt = rc_max_skip(t, nelim, m, a, lda, p);
      double maxp = rc_max_skip(p, nelim, m, a, lda, t);
      if (good_2x2(t, p, maxt, maxp, a, lda, u, small, &d[2 * nelim])) {
        sym_swap(t, nelim, m, perm, a, lda);
        sym_swap(p, nelim + 1, m, perm, a, lda);
        elim_2x2(nelim, m, n, a, lda, d, ld.data(), ld2.data());
        nelim += 2;
        break;
      }
      maxp = std::max(maxp, std::fabs(a[(size_t)t * lda + p]));
      if (std::fabs(a[(size_t)p * lda + p]) >= u * maxp) {
        sym_swap(p, nelim, m, perm, a, lda);
        d[2 * nelim] = 1 / a[(size_t)nelim * lda + nelim];
        d[2 * nelim + 1] = 0.0;
        elim_1x1(nelim, m, n, a, lda, d, ld.data());
        nelim += 1;
        break;
      }
    }
    if (p >= n) {
      p = nelim;
      double maxp = rc_max_skip(p, nelim, m, a, lda, -1);
      if (std::fabs(a[(size_t)p * lda + p]) >= u * maxp) {
        d[2 * nelim] = 1 / a[(size_t)nelim * lda + nelim];
        d[2 * nelim + 1] = 0.0;
        elim_1x1(nelim, m, n, a, lda, d, ld.data());
        nelim += 1;
      } else {
        break;
      }
    }
  }
  return nelim;
}

// ---------------- APP as the device runs it -------------------------------
// A restatement of the device factorization's a posteriori pivoting (the
// kernels k_diag / k_apply / k_update / k_swap): 64-column blocks; in each,
// complete pivoting on the symmetric w x w block with the rules of
// src/ssids/cpu/kernels/block_ldlt.hxx:289-414 (largest |entry| of the live
// block's lower triangle, ties to the smallest column then row, so an
// off-diagonal maximum gives the pair (column, row), column < row, as
// block_ldlt.hxx:323-330 orders a11 and a22; diagonal -> 1x1, else the 2x2
// test, else the larger diagonal as 1x1; a maximum below `small` makes every
// remaining index a zero pivot); the rows below take L21 = A21 L_bb^-T D^-1
// and the block passes up to the first column whose max |l| exceeds 1/u
// (src/ssids/cpu/kernels/ldlt_app.cxx:304-322), never splitting a 2x2; the
// failed columns are restored, updated by the accepted pivots and swapped
// behind the block's unprocessed candidates.  The block keeps both triangles
// and updates entry (r, j) with column bc's values, as the device does.
constexpr int kAppBlk = 64;

double sym_at(const double *a, int lda, int r, int c) {
  return r >= c ? a[(size_t)c * lda + r] : a[(size_t)r * lda + c];
}

// D (not its inverse) of pivot k from the stored D^{-1} (two: 0 1x1, 1 first
// of a 2x2, 2 second)
void pivot_d_host(const double *D, int k, int kend, double &e1, double &e2, int &two) {
  const double d11 = D[2 * k];
  if (!std::isfinite(d11)) {
    const double a11 = D[2 * (k - 1)], a21 = D[2 * (k - 1) + 1], a22 = D[2 * k + 1];
    const double det = a11 * a22 - a21 * a21;
    e1 = -a21 / det;
    e2 = a11 / det;
    two = 2;
  } else if (k + 1 < kend && !std::isfinite(D[2 * k + 2])) {
    const double a11 = d11, a21 = D[2 * k + 1], a22 = D[2 * k + 3];
    const double det = a11 * a22 - a21 * a21;
    e1 = a22 / det;
    e2 = -a21 / det;
    two = 1;
  } else {
    e1 = (d11 != 0.0) ? 1.0 / d11 : 0.0;
    e2 = 0.0;
    two = 0;
  }
}

// symmetric interchange of positions x and y in the panel: rows of the
// eliminated columns (< e) and the symmetric trailing entries
void panel_swap(int x, int y, int e, int m, int *perm, double *a, int lda) {
  if (x == y) return;
  if (y < x) std::swap(x, y);
  std::swap(perm[x], perm[y]);
  for (int c = 0; c < e; ++c) std::swap(at(a, lda, x, c), at(a, lda, y, c));
  for (int c = e; c < x; ++c) std::swap(at(a, lda, x, c), at(a, lda, y, c));
  for (int i = x + 1; i < y; ++i) std::swap(at(a, lda, i, x), at(a, lda, y, i));
  for (int r = y + 1; r < m; ++r) std::swap(at(a, lda, r, x), at(a, lda, r, y));
  std::swap(at(a, lda, x, x), at(a, lda, y, y));
}

// Returns nelim after APP (the failed columns sit at [nelim, n)); *nzero
// counts zero pivots; throws SingularError for a zero pivot without action.
int app_factor(int m, int n, int *perm, double *a, int lda, double *d, bool action,
               double u, double small, int *nzero) {
  int p = 0, q = n;
  const int nsteps = (n + kAppBlk - 1) / kAppBlk;
  const double lim = 1.0 / u;
  for (int s = 0; s < nsteps; ++s) {
    const int w = std::min(kAppBlk, q - p);
    if (w <= 0) break;
    // the block (both triangles) and backups in original index order
    std::vector<double> B((size_t)w * w), dbak, bak((size_t)w * (m - p - w));
    for (int c = 0; c < w; ++c)
      for (int r = 0; r < w; ++r) B[(size_t)c * w + r] = sym_at(a, lda, p + r, p + c);
    dbak = B;
    for (int k = 0; k < w; ++k)
      for (int i = p + w; i < m; ++i) bak[(size_t)k * (m - p - w) + (i - p - w)] = at(a, lda, i, p + k);
    // ---- complete pivoting on the block
    auto Bat = [&](int r, int c) -> double & { return B[(size_t)c * w + r]; };
    std::vector<int> lp(w), kind(w);
    std::vector<double> t((size_t)w * w, 0.0), dinv(2 * w + 2, 0.0);
    std::vector<char> elim(w, 0);
    int k = 0;
    while (k < w) {
      double vmax = 0.0;
      int bc = -1, brow = -1;
      for (int c = 0; c < w; ++c) {
        if (elim[c]) continue;
        for (int r = c; r < w; ++r) {  // the lower triangle (block_ldlt.hxx:77-101)
          if (elim[r]) continue;
          const double v = std::fabs(Bat(r, c));
          if (v > vmax) {  // NaN never wins; columns scanned in order, rows in order
            vmax = v;
            bc = c;
            brow = r;
          }
        }
      }
      if (!(vmax >= small) || bc < 0) {
        // every remaining index a zero pivot, in increasing index order
        for (int i = 0; i < w; ++i) {
          if (elim[i]) continue;
          for (int r = 0; r < w; ++r) t[(size_t)k * w + r] = 0.0;
          lp[k] = i;
          kind[k] = 0;
          dinv[2 * k] = 0.0;
          dinv[2 * k + 1] = 0.0;
          elim[i] = 1;
          ++k;
          ++*nzero;
        }
        if (!action) throw SingularError();
        break;
      }
      std::vector<double> c1(w), c2(w);
      for (int r = 0; r < w; ++r) c1[r] = Bat(r, bc);
      for (int r = 0; r < w; ++r) c2[r] = Bat(r, brow);
      int size = 1, p1 = bc, p2 = -1;
      double d11 = 0, d21 = 0, d22 = 0;
      const std::vector<double> *c1p = &c1, *c2p = &c2;
      if (brow != bc) {
        const double a11 = c1[bc], a22 = c2[brow], a21 = c1[brow];
        const double detscale = 1.0 / std::fabs(a21);
        const double detpiv = std::fma(a11 * detscale, a22, -std::fabs(a21));
        if (std::fabs(detpiv) >= std::fabs(a21) / 2) {
          size = 2;
          p2 = brow;
          d11 = (a22 * detscale) / detpiv;
          d21 = (-a21 * detscale) / detpiv;
          d22 = (a11 * detscale) / detpiv;
        } else {
          const bool first = std::fabs(a11) > std::fabs(a22);
          p1 = first ? bc : brow;
          c1p = first ? &c1 : &c2;
        }
      }
      if (size == 1) {
        const double di = 1.0 / (*c1p)[p1];
        elim[p1] = 1;
        for (int j = 0; j < w; ++j) {
          const double lj = (*c1p)[j] * di;
          for (int r = 0; r < w; ++r) Bat(r, j) = std::fma(-(*c1p)[r], lj, Bat(r, j));
          t[(size_t)k * w + j] = lj;
        }
        dinv[2 * k] = di;
        dinv[2 * k + 1] = 0.0;
        lp[k] = p1;
        kind[k] = 1;
      } else {
        elim[p1] = elim[p2] = 1;
        for (int j = 0; j < w; ++j) {
          const double x = c1[j], y = c2[j];
          const double l1 = std::fma(d11, x, d21 * y), l2 = std::fma(d21, x, d22 * y);
          for (int r = 0; r < w; ++r)
            Bat(r, j) = std::fma(-c1[r], l1, std::fma(-c2[r], l2, Bat(r, j)));
          t[(size_t)k * w + j] = l1;
          t[(size_t)(k + 1) * w + j] = l2;
        }
        dinv[2 * k] = d11;
        dinv[2 * k + 1] = d21;
        dinv[2 * k + 2] = std::numeric_limits<double>::infinity();
        dinv[2 * k + 3] = d22;
        lp[k] = p1;
        lp[k + 1] = p2;
        kind[k] = 2;
        kind[k + 1] = 3;
      }
      k += size;
    }
    std::vector<int> inv_lp(w);
    for (int i = 0; i < w; ++i) inv_lp[lp[i]] = i;
    // rows [p, p+w) of the eliminated columns left of p follow the order
    for (int c = 0; c < p; ++c) {
      std::vector<double> tmp(w);
      for (int r = 0; r < w; ++r) tmp[r] = at(a, lda, p + lp[r], c);
      for (int r = 0; r < w; ++r) at(a, lda, p + r, c) = tmp[r];
    }
    // L_bb in elimination order, D^{-1}, the variable list
    for (int kk = 0; kk < w; ++kk)
      for (int nr = kk; nr < w; ++nr) {
        const int r = lp[nr];
        double v;
        if (nr == kk) v = (kind[kk] == 0) ? 0.0 : 1.0;
        else if (nr == kk + 1 && kind[kk] == 2) v = 0.0;
        else v = t[(size_t)kk * w + r];
        at(a, lda, p + nr, p + kk) = v;
      }
    for (int j = 0; j < w; ++j) {
      d[2 * (p + j)] = dinv[2 * j];
      d[2 * (p + j) + 1] = dinv[2 * j + 1];
    }
    {
      std::vector<int> pv(w);
      for (int j = 0; j < w; ++j) pv[j] = perm[p + lp[j]];
      for (int j = 0; j < w; ++j) perm[p + j] = pv[j];
    }
    // M = L_bb^{-1} (unit lower, elimination order; a 2x2 pair's partner
    // entry is 0 and a zero pivot's column is 0 below the diagonal)
    std::vector<double> M((size_t)w * w, 0.0);  // M[row][col]
    for (int i = 0; i < w; ++i) M[(size_t)i * w + i] = 1.0;
    for (int kk = 0; kk < w; ++kk)
      for (int i = kk + 1; i < w; ++i) {
        const double lik = (kind[kk] == 2 && i == kk + 1) ? 0.0 : at(a, lda, p + i, p + kk);
        if (lik == 0.0) continue;
        for (int c = 0; c <= kk; ++c) M[(size_t)i * w + c] -= lik * M[(size_t)kk * w + c];
      }
    // rows below: y = (A21[:, lp] M^T) D^{-1}, column maxima of |y|
    std::vector<double> colmax(w, 0.0);
    for (int i = p + w; i < m; ++i) {
      std::vector<double> x(w, 0.0), y(w, 0.0);
      for (int j = 0; j < w; ++j) {
        double acc = 0.0;
        for (int kk = 0; kk <= j; ++kk) acc += bak[(size_t)lp[kk] * (m - p - w) + (i - p - w)] * M[(size_t)j * w + kk];
        x[j] = acc;
      }
      for (int j = 0; j < w; ++j) {
        const bool two_first = (j + 1 < w) && std::isinf(dinv[2 * j + 2]);
        const bool two_second = (j > 0) && std::isinf(dinv[2 * j]);
        if (two_first) y[j] = dinv[2 * j] * x[j] + dinv[2 * j + 1] * x[j + 1];
        else if (two_second) y[j] = dinv[2 * j - 1] * x[j - 1] + dinv[2 * j + 1] * x[j];
        else if (dinv[2 * j] == 0.0) y[j] = (std::fabs(x[j]) < small) ? 0.0 : INFINITY * x[j];
        else y[j] = x[j] * dinv[2 * j];
        at(a, lda, i, p + j) = y[j];
        const double v = std::fabs(y[j]);
        if (std::isnan(v) || v > colmax[j] || std::isnan(colmax[j]))
          colmax[j] = std::isnan(colmax[j]) ? colmax[j] : v;
      }
    }
    int npass = w;
    for (int j = 0; j < w; ++j)
      if (std::isnan(colmax[j]) || colmax[j] > lim) {
        npass = j;
        break;
      }
    if (npass > 0 && npass < w) {
      const double e11 = d[2 * (p + npass - 1)], e21 = d[2 * (p + npass - 1) + 1];
      if (std::isfinite(e11) && e21 != 0.0) --npass;
    }
    const int t0 = p + npass;
    // restore the failed block columns (original values, elimination order)
    for (int gj = t0; gj < p + w; ++gj) {
      const int jj = gj - p;
      for (int gi = gj; gi < p + w; ++gi)
        at(a, lda, gi, gj) = dbak[(size_t)lp[jj] * w + lp[gi - p]];
      for (int gi = p + w; gi < m; ++gi)
        at(a, lda, gi, gj) = bak[(size_t)lp[jj] * (m - p - w) + (gi - p - w)];
    }
    // the accepted pivots' update of the trailing panel [t0, n) x [t0, m):
    // entry (r, c) -= sum_k L(r, k) (L D)(c, k), the sum formed first
    std::vector<double> bv(std::max(t0 - p, 1));
    for (int c = t0; c < n; ++c) {
      for (int kk = p; kk < t0; ++kk) {
        double e1, e2;
        int two;
        pivot_d_host(d, kk, t0, e1, e2, two);
        const int ofs = (two == 1) ? 1 : (two == 2) ? -1 : 0;
        const double lk = at(a, lda, c, kk), lo = at(a, lda, c, kk + ofs);
        bv[kk - p] = (two == 0) ? lk * e1 : (two == 1) ? lk * e1 + lo * e2 : lo * e1 + lk * e2;
      }
      for (int r = c; r < m; ++r) {
        double acc = 0.0;
        for (int kk = p; kk < t0; ++kk) acc += at(a, lda, r, kk) * bv[kk - p];
        at(a, lda, r, c) -= acc;
      }
    }
    // failed columns behind the unprocessed candidates
    const int nf = w - npass, nu = q - (p + w);
    const int ns = (nf > 0 && nu > 0) ? std::min(nf, nu) : 0;
    for (int i = 0; i < ns; ++i) panel_swap(t0 + i, q - ns + i, t0, m, perm, a, lda);
    p = t0;
    q -= nf;
  }
  return p;
}

bool is_1x1(const double *d, int i, int nelim) {
  return i + 1 == nelim || std::isfinite(d[2 * i + 2]);
}

// contribution update F[n:m,n:m] -= L[n:m,0:ne] D L[n:m,0:ne]^T (lower)
void contrib_update(Node &nd) {
  int m = nd.m, n = nd.n, ne = nd.nelim, lda = m;
  if (m == n || ne == 0) return;
  double *a = nd.f.data();
  // LD = L * D (rows n..m-1)
  std::vector<double> ldm((size_t)(m - n) * ne);
  for (int i = 0; i < ne;) {
    if (is_1x1(nd.d.data(), i, ne)) {
      double d11 = nd.d[2 * i];
      double dd = (d11 != 0.0) ? 1 / d11 : 0.0;
      for (int r = n; r < m; ++r) ldm[(size_t)i * (m - n) + r - n] = dd * a[(size_t)i * lda + r];
      i += 1;
    } else {
      // D block = inverse of the stored 2x2 D^{-1}
      double d11 = nd.d[2 * i], d21 = nd.d[2 * i + 1], d22 = nd.d[2 * i + 3];
      double det = d11 * d22 - d21 * d21;
      double e11 = d22 / det, e21 = -d21 / det, e22 = d11 / det;
      for (int r = n; r < m; ++r) {
        double l1 = a[(size_t)i * lda + r], l2 = a[(size_t)(i + 1) * lda + r];
        ldm[(size_t)i * (m - n) + r - n] = e11 * l1 + e21 * l2;
        ldm[(size_t)(i + 1) * (m - n) + r - n] = e21 * l1 + e22 * l2;
      }
      i += 2;
    }
  }
  for (int c = n; c < m; ++c) {
    double *ac = &a[(size_t)c * lda];
    for (int k = 0; k < ne; ++k) {
      double s = ldm[(size_t)k * (m - n) + c - n];
      const double *lk = &a[(size_t)k * lda];
      for (int r = c; r < m; ++r) ac[r] -= lk[r] * s;
    }
  }
}

}  // namespace

extern "C" {

void *oracle_factor(int n, int nnodes, const int *sptr, const int *sparent,
                    const int64_t *rptr, const int *rlist, const int64_t *nptr,
                    const int64_t *nlist, const double *aval,
                    const double *scaling, int action, double u, double small,
                    int64_t *stats) {
  return oracle_factor_prefix(n, nnodes, sptr, sparent, rptr, rlist, nptr, nlist, aval,
                              scaling, action, u, small, nnodes, stats);
}

// app: the device's front layout (children's delayed columns first, in child
// order, then the node's own columns, then the rest) and APP followed by TPP
// on the failed columns; else the reference CPU layout and TPP throughout
static void *factor_impl(int n, int nnodes, const int *sptr, const int *sparent,
                         const int64_t *rptr, const int *rlist, const int64_t *nptr,
                         const int64_t *nlist, const double *aval, const double *scaling,
                         int action, double u, double small, int max_nodes, bool app,
                         int64_t *stats) {
  auto *o = new Oracle();
  o->n = n;
  o->nnodes = nnodes;
  o->nodes.resize(nnodes);
  for (int k = 0; k < 10; ++k) stats[k] = 0;
  // children lists in increasing order
  std::vector<std::vector<int>> children(nnodes + 1);
  for (int s = 0; s < nnodes; ++s) children[sparent[s] - 1].push_back(s);
  std::vector<int> map(n + 1, -1);
  int64_t num_delay = 0, not_first = 0, nfact = 0, nflops = 0;
  int maxfront = 0;
  const int nfac = std::min(nnodes, std::max(0, max_nodes));
  try {
    for (int s = 0; s < nfac; ++s) {
      Node &nd = o->nodes[s];
      nd.sn = sptr[s + 1] - sptr[s];
      nd.sm = (int)(rptr[s + 1] - rptr[s]);
      nd.rows = &rlist[rptr[s] - 1];
      nd.nd = 0;
      for (int c : children[s]) nd.nd += o->nodes[c].n - o->nodes[c].nelim;
      nd.m = nd.sm + nd.nd;
      nd.n = nd.sn + nd.nd;
      maxfront = std::max(maxfront, nd.m);
      int m = nd.m, lda = m;
      nd.f.assign((size_t)m * m, 0.0);
      nd.d.assign(2 * (size_t)nd.n, 0.0);
      nd.perm.assign(nd.n, 0);
      const int own0 = app ? nd.nd : 0;  // first own column
      for (int i = 0; i < nd.sn; ++i) nd.perm[own0 + i] = nd.rows[i];
      for (int i = 0; i < nd.sn; ++i) map[nd.rows[i]] = own0 + i;
      for (int i = nd.sn; i < nd.sm; ++i) map[nd.rows[i]] = i + nd.nd;
      double *a = nd.f.data();
      // A entries (src/ssids/cpu/kernels/assemble.hxx:49-80)
      for (int64_t k = nptr[s] - 1; k < nptr[s + 1] - 1; ++k) {
        int64_t src = nlist[2 * k] - 1, dst = nlist[2 * k + 1] - 1;
        int c = (int)(dst / nd.sm), r = (int)(dst % nd.sm);
        int rr = (r >= nd.sn) ? r + nd.nd : r + own0;
        double v = aval[src];
        if (scaling) v *= scaling[nd.rows[r] - 1] * scaling[nd.rows[c] - 1];
        at(a, lda, rr, c + own0) = v;
      }
      // children's trailing matrices
      int delay_col = app ? 0 : nd.sn;
      for (int c : children[s]) {
        Node &ch = o->nodes[c];
        int cm = ch.m, cne = ch.nelim, cn = ch.n;
        std::vector<int> pos(cm - cne);
        for (int k = cne; k < cn; ++k) {
          nd.perm[delay_col] = ch.perm[k];
          pos[k - cne] = delay_col++;
        }
        for (int k = cn; k < cm; ++k) pos[k - cne] = map[ch.rows[k - ch.nd]];
        const double *ca = ch.f.data();
        for (int kc = cne; kc < cm; ++kc) {
          int pc = pos[kc - cne];
          for (int kr = kc; kr < cm; ++kr) {
            int pr = pos[kr - cne];
            int hi = std::max(pr, pc), lo = std::min(pr, pc);
            at(a, lda, hi, lo) += ca[(size_t)kc * cm + kr];
          }
        }
      }
      // factor
      if (app) {
        int nz = 0;
        nd.nelim_app = app_factor(m, nd.n, nd.perm.data(), a, lda, nd.d.data(), action != 0,
                                  u, small, &nz);
        nd.nelim = tpp_factor(m, nd.n, nd.perm.data(), a, lda, nd.d.data(), action != 0, u,
                              small, nd.nelim_app);
      } else {
        nd.nelim = tpp_factor(m, nd.n, nd.perm.data(), a, lda, nd.d.data(), action != 0, u,
                              small);
        nd.nelim_app = nd.nelim;
      }
      not_first += nd.n - nd.nelim_app;
      contrib_update(nd);
      num_delay += nd.n - nd.nelim;
      for (int64_t j = m; j >= m - nd.nelim + 1; --j) {
        nfact += j;
        nflops += j * j;
      }
    }
  } catch (SingularError &) {
    stats[0] = -5;
    return o;
  }
  int64_t neg = 0, two = 0, zero = 0;
  for (int s = 0; s < nfac; ++s) {
    Node &nd = o->nodes[s];
    const double *d = nd.d.data();
    for (int i = 0; i < nd.nelim;) {
      double a11 = d[2 * i], a21 = d[2 * i + 1];
      if (is_1x1(d, i, nd.nelim)) {
        if (a11 == 0.0) zero++;
        if (a11 < 0.0) neg++;
        i += 1;
      } else {
        double a22 = d[2 * i + 3];
        two++;
        double det = a11 * a22 - a21 * a21, tr = a11 + a22;
        if (det < 0) neg++;
        else if (tr < 0) neg += 2;
        i += 2;
      }
    }
  }
  stats[0] = zero ? 7 : 0;
  stats[1] = num_delay;
  stats[2] = neg;
  stats[3] = two;
  stats[4] = zero;
  stats[5] = maxfront;
  stats[6] = not_first;
  stats[7] = nfact;
  stats[8] = nflops;
  stats[9] = (sptr[nnodes] - 1) - zero;
  return o;
}

void *oracle_factor_prefix(int n, int nnodes, const int *sptr, const int *sparent,
                           const int64_t *rptr, const int *rlist, const int64_t *nptr,
                           const int64_t *nlist, const double *aval,
                           const double *scaling, int action, double u, double small,
                           int max_nodes, int64_t *stats) {
  return factor_impl(n, nnodes, sptr, sparent, rptr, rlist, nptr, nlist, aval, scaling, action,
                     u, small, max_nodes, false, stats);
}

void *oracle_factor_app(int n, int nnodes, const int *sptr, const int *sparent,
                        const int64_t *rptr, const int *rlist, const int64_t *nptr,
                        const int64_t *nlist, const double *aval, const double *scaling,
                        int action, double u, double small, int64_t *stats) {
  return factor_impl(n, nnodes, sptr, sparent, rptr, rlist, nptr, nlist, aval, scaling, action,
                     u, small, nnodes, true, stats);
}

void oracle_node_app(void *h, int *nelim_app) {
  auto *o = static_cast<Oracle *>(h);
  for (int s = 0; s < o->nnodes; ++s) nelim_app[s] = o->nodes[s].nelim_app;
}

void oracle_node_info(void *h, int *nelim, int *ndelay_in) {
  auto *o = static_cast<Oracle *>(h);
  for (int s = 0; s < o->nnodes; ++s) {
    nelim[s] = o->nodes[s].nelim;
    ndelay_in[s] = o->nodes[s].nd;
  }
}

void oracle_solve(void *h, int job, int nrhs, double *x, int ldx) {
  auto *o = static_cast<Oracle *>(h);
  std::vector<double> xl;
  std::vector<int> map;
  bool fwd = (job == 0 || job == 1), bwd = (job == 0 || job == 3 || job == 4);
  bool diag = (job == 0 || job == 2 || job == 4);
  if (fwd) {
    for (int s = 0; s < o->nnodes; ++s) {
      Node &nd = o->nodes[s];
      int m = nd.m, ne = nd.nelim;
      map.resize(m);
      for (int i = 0; i < nd.n; ++i) map[i] = nd.perm[i] - 1;
      for (int i = nd.sn; i < nd.sm; ++i) map[i + nd.nd] = nd.rows[i] - 1;
      const double *l = nd.f.data();
      for (int r = 0; r < nrhs; ++r) {
        xl.assign(m, 0.0);
        for (int i = 0; i < m; ++i) xl[i] = x[(size_t)r * ldx + map[i]];
        for (int j = 0; j < ne; ++j) {
          double v = xl[j];
          for (int i = j + 1; i < m; ++i) xl[i] -= l[(size_t)j * m + i] * v;
        }
        for (int i = 0; i < m; ++i) x[(size_t)r * ldx + map[i]] = xl[i];
      }
    }
  }
  if (diag || bwd) {
    for (int s = o->nnodes - 1; s >= 0; --s) {
      Node &nd = o->nodes[s];
      int m = nd.m, ne = nd.nelim;
      map.resize(m);
      for (int i = 0; i < nd.n; ++i) map[i] = nd.perm[i] - 1;
      for (int i = nd.sn; i < nd.sm; ++i) map[i + nd.nd] = nd.rows[i] - 1;
      const double *l = nd.f.data();
      const double *d = nd.d.data();
      for (int r = 0; r < nrhs; ++r) {
        xl.assign(m, 0.0);
        int blkm = bwd ? m : ne;
        for (int i = 0; i < blkm; ++i) xl[i] = x[(size_t)r * ldx + map[i]];
        if (diag) {
          for (int i = 0; i < ne;) {
            if (is_1x1(d, i, ne)) {
              xl[i] *= d[2 * i];
              i += 1;
            } else {
              double x1 = xl[i], x2 = xl[i + 1];
              xl[i] = d[2 * i] * x1 + d[2 * i + 1] * x2;
              xl[i + 1] = d[2 * i + 1] * x1 + d[2 * i + 3] * x2;
              i += 2;
            }
          }
        }
        if (bwd) {
          for (int j = ne - 1; j >= 0; --j) {
            double acc = xl[j];
            for (int i = j + 1; i < m; ++i) acc -= l[(size_t)j * m + i] * xl[i];
            xl[j] = acc;
          }
        }
        for (int i = 0; i < ne; ++i) x[(size_t)r * ldx + map[i]] = xl[i];
      }
    }
  }
}

void oracle_enquire(void *h, int *piv_order, double *d) {
  auto *o = static_cast<Oracle *>(h);
  if (piv_order)
    for (int i = 0; i < o->n; ++i) piv_order[i] = 0;
  if (d)
    for (int i = 0; i < 2 * o->n; ++i) d[i] = 0.0;
  int piv = 0;
  double *dp = d;
  for (int s = 0; s < o->nnodes; ++s) {
    Node &nd = o->nodes[s];
    const double *dd = nd.d.data();
    for (int i = 0; i < nd.nelim;) {
      if (is_1x1(dd, i, nd.nelim)) {
        if (piv_order) piv_order[nd.perm[i] - 1] = piv;
        ++piv;
        if (dp) { *dp++ = dd[2 * i]; *dp++ = 0.0; }
        i += 1;
      } else {
        if (piv_order) {
          piv_order[nd.perm[i] - 1] = -(piv++);
          piv_order[nd.perm[i + 1] - 1] = -(piv++);
        } else {
          piv += 2;
        }
        if (dp) {
          *dp++ = dd[2 * i];
          *dp++ = dd[2 * i + 1];
          *dp++ = dd[2 * i + 3];
          *dp++ = 0.0;
        }
        i += 2;
      }
    }
  }
}

int oracle_node_factors(void *h, int node, int *info, double *L, double *D, int *perm) {
  auto *o = static_cast<Oracle *>(h);
  if (node < 0 || node >= o->nnodes) return -1;
  Node &nd = o->nodes[node];
  info[0] = nd.m;
  info[1] = nd.n;
  info[2] = nd.nd;
  info[3] = nd.nelim;
  info[4] = nd.m;
  info[5] = nd.sn;
  if (L)
    for (int j = 0; j < nd.n; ++j)
      for (int i = 0; i < nd.m; ++i) L[(size_t)j * nd.m + i] = nd.f[(size_t)j * nd.m + i];
  if (D)
    for (int i = 0; i < 2 * nd.n; ++i) D[i] = nd.d[i];
  if (perm)
    for (int i = 0; i < nd.n; ++i) perm[i] = nd.perm[i] - 1;
  return 0;
}

void oracle_free(void *h) { delete static_cast<Oracle *>(h); }

}  // extern "C"
