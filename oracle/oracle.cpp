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
               double u, double small) {
  std::vector<double> ld(m), ld2(m);
  int nelim = 0;
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

void *oracle_factor_prefix(int n, int nnodes, const int *sptr, const int *sparent,
                           const int64_t *rptr, const int *rlist, const int64_t *nptr,
                           const int64_t *nlist, const double *aval,
                           const double *scaling, int action, double u, double small,
                           int max_nodes, int64_t *stats) {
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
      for (int i = 0; i < nd.sn; ++i) nd.perm[i] = nd.rows[i];
      for (int i = 0; i < nd.sn; ++i) map[nd.rows[i]] = i;
      for (int i = nd.sn; i < nd.sm; ++i) map[nd.rows[i]] = i + nd.nd;
      double *a = nd.f.data();
      // A entries (src/ssids/cpu/kernels/assemble.hxx:49-80)
      for (int64_t k = nptr[s] - 1; k < nptr[s + 1] - 1; ++k) {
        int64_t src = nlist[2 * k] - 1, dst = nlist[2 * k + 1] - 1;
        int c = (int)(dst / nd.sm), r = (int)(dst % nd.sm);
        int rr = (r >= nd.sn) ? r + nd.nd : r;
        double v = aval[src];
        if (scaling) v *= scaling[nd.rows[r] - 1] * scaling[nd.rows[c] - 1];
        at(a, lda, rr, c) = v;
      }
      // children's trailing matrices
      int delay_col = nd.sn;
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
      nd.nelim = tpp_factor(m, nd.n, nd.perm.data(), a, lda, nd.d.data(),
                            action != 0, u, small);
      not_first += nd.n - nd.nelim;
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
