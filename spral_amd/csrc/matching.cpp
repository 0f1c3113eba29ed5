// Maximum-product matching scaling of a symmetric matrix (host C++ in the
// library): the weighted bipartite matching of Duff and Koster (shortest
// augmenting paths on reduced costs, as in MC64), its dual variables turned
// into a symmetric scaling.  Replaces hungarian_scale_sym
// (src/scaling.f90:134-170) with the pre/post-processing of
// hungarian_wrapper (:597-801), the start heuristic (:810-929), the
// assignment itself (hungarian_match, :938-1171) and the square case of
// match_postproc (:1588-1616).  The decisions (which row a column takes on a
// tie, the order in which equal-distance rows leave the queue) are the
// reference's, so the matching and the scaling are the same bit for bit.
//
// The auction scaling (auction_scale_sym, :269-309; auction_match
// :1481-1586, auction_match_core :1328-1466) is the same restated for the
// approximate matching of Bertsekas's auction, as the reference runs it.
//
// The same matching drives the matching-based elimination order of
// src/match_order.f90 (match_order_metis :135-208, mo_match :495-629,
// mo_split :220-396): matched pairs of the 1- and 2-cycle split are merged,
// the compressed graph is ordered (approximate minimum degree here: METIS is
// not available) and each pair takes consecutive pivot positions, so a
// saddle-point system's constraint rows meet their partners in one front.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <new>
#include <vector>

#include "../../include/spral_scaling.h"

namespace {

constexpr double kInf = std::numeric_limits<double>::max();

// Column-compressed pattern with costs, 0-based.
struct Graph {
  int n = 0;
  std::vector<int64_t> ptr;
  std::vector<int> row;
  std::vector<double> val;
};

// Both triangles of a lower-triangle CSC (explicit zeros dropped first).
// Column i of the result lists the entries (i, j), j < i, taken from the
// columns j that hold row i, in increasing j, then the column's own lower
// entries in input order: the order src/matrix_util.f90:3167-3302 produces.
// `logabs` stores log|a| instead of a.
Graph expand_lower(int n, const int64_t *ptr, const int *row, const double *val, bool logabs) {
  Graph g;
  g.n = n;
  std::vector<int64_t> lo_ptr(n + 1, 0);
  std::vector<int> lo_row;
  std::vector<double> lo_val;
  lo_row.reserve(ptr[n] - 1);
  lo_val.reserve(ptr[n] - 1);
  for (int j = 0; j < n; ++j) {
    lo_ptr[j] = (int64_t)lo_row.size();
    for (int64_t k = ptr[j] - 1; k < ptr[j + 1] - 1; ++k) {
      if (val[k] == 0.0) continue;
      lo_row.push_back(row[k] - 1);
      lo_val.push_back(logabs ? std::log(std::fabs(val[k])) : val[k]);
    }
  }
  lo_ptr[n] = (int64_t)lo_row.size();
  std::vector<int64_t> cnt(n, 0);
  for (int j = 0; j < n; ++j)
    for (int64_t k = lo_ptr[j]; k < lo_ptr[j + 1]; ++k) {
      ++cnt[j];
      if (lo_row[k] != j) ++cnt[lo_row[k]];
    }
  g.ptr.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) g.ptr[j + 1] = g.ptr[j] + cnt[j];
  g.row.resize(g.ptr[n]);
  g.val.resize(g.ptr[n]);
  std::vector<int64_t> up_fill(n);
  for (int j = 0; j < n; ++j) up_fill[j] = g.ptr[j];
  // upper part of column i: entries (i, j) of the lower columns j < i, in
  // increasing j
  for (int j = 0; j < n; ++j)
    for (int64_t k = lo_ptr[j]; k < lo_ptr[j + 1]; ++k) {
      const int i = lo_row[k];
      if (i == j) continue;
      g.row[up_fill[i]] = j;
      g.val[up_fill[i]] = lo_val[k];
      ++up_fill[i];
    }
  for (int j = 0; j < n; ++j) {
    int64_t dst = up_fill[j];
    for (int64_t k = lo_ptr[j]; k < lo_ptr[j + 1]; ++k, ++dst) {
      g.row[dst] = lo_row[k];
      g.val[dst] = lo_val[k];
    }
  }
  return g;
}

// Binary heap of rows keyed by d (positions 1..qlen of q), with pos[i] the
// position of row i.  Ties keep the element already in place.
struct RowHeap {
  std::vector<int> &q, &pos;
  const std::vector<double> &d;
  // d[idx] decreased (or idx just appended at pos[idx]): move it up
  void decreased(int idx) {
    int p = pos[idx];
    if (p <= 1) {
      q[p] = idx;
      return;
    }
    const double v = d[idx];
    while (p > 1) {
      const int pp = p / 2, pidx = q[pp];
      if (v >= d[pidx]) break;
      q[p] = pidx;
      pos[pidx] = p;
      p = pp;
    }
    q[p] = idx;
    pos[idx] = p;
  }
  // remove the element at position p0 (the last element fills the hole)
  void remove_at(int p0, int &qlen) {
    if (qlen == p0) {
      --qlen;
      return;
    }
    const int idx = q[qlen];
    const double v = d[idx];
    --qlen;
    int p = p0;
    if (p > 1) {
      for (;;) {
        const int pp = p / 2, qk = q[pp];
        if (v >= d[qk]) break;
        q[p] = qk;
        pos[qk] = p;
        p = pp;
        if (p <= 1) break;
      }
    }
    q[p] = idx;
    pos[idx] = p;
    if (p != p0) return;
    for (;;) {
      int c = 2 * p;
      if (c > qlen) break;
      double dk = d[q[c]];
      if (c < qlen) {
        const double dr = d[q[c + 1]];
        if (dk > dr) {
          ++c;
          dk = dr;
        }
      }
      if (v <= dk) break;
      const int qk = q[c];
      q[p] = qk;
      pos[qk] = p;
      p = c;
    }
    q[p] = idx;
    pos[idx] = p;
  }
  int pop(int &qlen) {
    const int top = q[1];
    remove_at(1, qlen);
    return top;
  }
};

// Minimum-cost assignment on a square n x n cost matrix (costs >= 0), with
// row duals u and column duals v.  rowmatch[i] = column matched to row i or
// -1; returns the cardinality.
int min_cost_assignment(const Graph &g, std::vector<int> &rowmatch, std::vector<double> &u,
                        std::vector<double> &v) {
  const int n = g.n, m = n;
  const auto &ptr = g.ptr;
  const auto &row = g.row;
  const auto &val = g.val;
  std::vector<int64_t> colentry(n, -1);  // entry of column j in the matching
  rowmatch.assign(m, -1);
  u.assign(m, kInf);
  v.assign(n, 0.0);
  int num = 0;
  // ---- start: each row on its cheapest column (the last one on ties),
  // where that column is free and not dense
  {
    std::vector<int> cheap(m, -1);
    std::vector<int64_t> cheap_at(m, -1);
    for (int j = 0; j < n; ++j)
      for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k) {
        const int i = row[k];
        if (val[k] > u[i]) continue;
        u[i] = val[k];
        cheap[i] = j;
        cheap_at[i] = k;
      }
    for (int i = 0; i < m; ++i) {
      const int j = cheap[i];
      if (j < 0 || colentry[j] >= 0) continue;
      if (ptr[j + 1] - ptr[j] > m / 10 && m > 50) continue;
      ++num;
      rowmatch[i] = j;
      colentry[j] = cheap_at[i];
    }
    if (num < std::min(m, n)) {
      // free columns: take the cheapest reduced entry (an unmatched row
      // first on ties), or free a matched row by moving its column to an
      // unmatched row at no extra reduced cost (a path of length 2)
      std::vector<double> dcol(n, 0.0);
      std::vector<int64_t> from(n);
      for (int j = 0; j < n; ++j) from[j] = ptr[j];
      for (int j = 0; j < n; ++j) {
        if (colentry[j] >= 0 || ptr[j] >= ptr[j + 1]) continue;
        int i0 = row[ptr[j]];
        double vj = val[ptr[j]] - u[i0];
        int64_t k0 = ptr[j];
        for (int64_t k = ptr[j] + 1; k < ptr[j + 1]; ++k) {
          const int i = row[k];
          const double di = val[k] - u[i];
          if (di > vj) continue;
          if (di == vj && di != kInf && (rowmatch[i] >= 0 || rowmatch[i0] < 0)) continue;
          vj = di;
          i0 = i;
          k0 = k;
        }
        dcol[j] = vj;
        if (rowmatch[i0] < 0) {
          ++num;
          colentry[j] = k0;
          rowmatch[i0] = j;
          from[j] = k0 + 1;
          continue;
        }
        bool moved = false;
        for (int64_t k = k0; k < ptr[j + 1] && !moved; ++k) {
          const int i = row[k];
          if (val[k] - u[i] > vj) continue;
          const int jj = rowmatch[i];
          for (int64_t kk = from[jj]; kk < ptr[jj + 1]; ++kk) {
            const int ii = row[kk];
            if (rowmatch[ii] >= 0) continue;
            if (val[kk] - u[ii] <= dcol[jj]) {
              colentry[jj] = kk;
              rowmatch[ii] = jj;
              from[jj] = kk + 1;
              ++num;
              colentry[j] = k;
              rowmatch[i] = j;
              from[j] = k + 1;
              moved = true;
              break;
            }
          }
          if (!moved) from[jj] = ptr[jj + 1];
        }
      }
    }
  }
  // ---- shortest augmenting paths (Dijkstra over reduced costs) from every
  // still-free column
  if (num < std::min(m, n)) {
    std::vector<double> d(std::max(m, n), kInf);
    std::vector<int> pos(m, 0), q(m + 2, 0), pred(n, -1);
    std::vector<int64_t> via(n, -1), touched(std::max(m, 1));
    RowHeap heap{q, pos, d};
    int64_t isp = -1;
    int jsp = -1;
    for (int jord = 0; jord < n; ++jord) {
      if (colentry[jord] >= 0) continue;
      double dmin = kInf, csp = kInf;
      int qlen = 0, low = m + 1, up = m + 1;
      int j = jord;
      pred[j] = -2;  // root
      int nt = 0;
      for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k) {
        const int i = row[k];
        const double dnew = val[k] - u[i];
        if (dnew >= csp) continue;
        if (rowmatch[i] < 0) {
          csp = dnew;
          isp = k;
          jsp = j;
        } else {
          if (dnew < dmin) dmin = dnew;
          d[i] = dnew;
          touched[nt++] = k;
        }
      }
      for (int kk = 0; kk < nt; ++kk) {
        const int64_t k = touched[kk];
        const int i = row[k];
        if (csp <= d[i]) {
          d[i] = kInf;
          continue;
        }
        if (d[i] <= dmin) {
          q[--low] = i;
          pos[i] = low;
        } else {
          pos[i] = ++qlen;
          heap.decreased(i);
        }
        const int jj = rowmatch[i];
        via[jj] = k;
        pred[jj] = j;
      }
      for (int it = 0; it < num; ++it) {
        if (low == up) {
          if (qlen == 0) break;
          if (d[q[1]] >= csp) break;
          dmin = d[q[1]];
          while (qlen > 0) {
            if (d[q[1]] > dmin) break;
            const int i = heap.pop(qlen);
            q[--low] = i;
            pos[i] = low;
          }
        }
        const int q0 = q[up - 1];
        const double dq0 = d[q0];
        if (dq0 >= csp) break;
        --up;
        j = rowmatch[q0];
        const double vj = dq0 - val[colentry[j]] + u[q0];
        for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k) {
          const int i = row[k];
          if (pos[i] >= up) continue;
          const double dnew = vj + val[k] - u[i];
          if (dnew >= csp) continue;
          if (rowmatch[i] < 0) {
            csp = dnew;
            isp = k;
            jsp = j;
          } else {
            if (d[i] <= dnew) continue;
            if (pos[i] >= low) continue;
            d[i] = dnew;
            if (dnew <= dmin) {
              if (pos[i] != 0) heap.remove_at(pos[i], qlen);
              q[--low] = i;
              pos[i] = low;
            } else {
              if (pos[i] == 0) pos[i] = ++qlen;
              heap.decreased(i);
            }
            const int jj = rowmatch[i];
            via[jj] = k;
            pred[jj] = j;
          }
        }
      }
      if (csp != kInf) {
        // augment along the path back to the root, then lower the duals of
        // the rows whose distance became final
        ++num;
        int i = row[isp];
        rowmatch[i] = jsp;
        colentry[jsp] = isp;
        j = jsp;
        for (int it = 0; it < num; ++it) {
          const int jj = pred[j];
          if (jj == -2) break;
          const int64_t k = via[j];
          i = row[k];
          rowmatch[i] = jj;
          colentry[jj] = k;
          j = jj;
        }
        for (int kk = up; kk <= m; ++kk) {
          const int r = q[kk];
          u[r] = u[r] + d[r] - csp;
        }
      }
      for (int kk = low; kk <= m; ++kk) {
        const int r = q[kk];
        d[r] = kInf;
        pos[r] = 0;
      }
      for (int kk = 1; kk <= qlen; ++kk) {
        const int r = q[kk];
        d[r] = kInf;
        pos[r] = 0;
      }
    }
  }
  for (int j = 0; j < n; ++j) {
    const int64_t k = colentry[j];
    v[j] = (k >= 0) ? val[k] - u[row[k]] : 0.0;
  }
  for (int i = 0; i < m; ++i)
    if (rowmatch[i] < 0) u[i] = 0.0;
  return num;
}

// matching order helpers --------------------------------------------------

// mo_match: costs log(max_k |a_kj|) - log|a_ij| (column maximum clamped at 0
// before its log), the assignment, and the log scaling (u + v - cmax) / 2;
// a structurally singular matrix is matched again on its matched index set
// and its unmatched indices get -huge (scale 0 after the exponential).
// match[i]: partner of i (0-based) or -1.  Returns 0, or 1 when singular.
int match_and_log_scale(const Graph &full_abs, std::vector<double> &scale,
                        std::vector<int> &match) {
  const int n = full_abs.n;
  Graph g = full_abs;
  std::vector<double> cmax(n);
  for (int j = 0; j < n; ++j) {
    double mx = 0.0;
    for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) mx = std::max(mx, g.val[k]);
    if (mx != 0.0) mx = std::log(mx);
    cmax[j] = mx;
    for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) g.val[k] = mx - std::log(g.val[k]);
  }
  std::vector<int> rm;
  std::vector<double> u, v;
  const int rank = min_cost_assignment(g, rm, u, v);
  scale.assign(n, 0.0);
  if (rank == n) {
    for (int i = 0; i < n; ++i) scale[i] = (u[i] + v[i] - cmax[i]) / 2;
    match = rm;
    return 0;
  }
  std::vector<int> o2n(n, -1), n2o;
  for (int i = 0; i < n; ++i)
    if (rm[i] >= 0) {
      o2n[i] = (int)n2o.size();
      n2o.push_back(i);
    }
  Graph r;
  r.n = (int)n2o.size();
  r.ptr.assign(r.n + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (rm[i] < 0) continue;
    for (int64_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
      const int jj = g.row[k];
      if (rm[jj] < 0) continue;
      r.row.push_back(o2n[jj]);
      r.val.push_back(g.val[k]);
    }
    r.ptr[o2n[i] + 1] = (int64_t)r.row.size();
  }
  std::vector<int> rm2;
  min_cost_assignment(r, rm2, u, v);
  for (int i = 0; i < n; ++i) {
    const int j = o2n[i];
    scale[i] = (j < 0) ? -kInf : (u[j] + v[j] - cmax[i]) / 2;
  }
  match.assign(n, -1);
  for (int i = 0; i < r.n; ++i)
    if (rm2[i] >= 0) match[n2o[i]] = n2o[rm2[i]];
  return 1;
}

}  // namespace

namespace ssids_hip {
int order_amd(int n, const int64_t *ptr, const int *row, int *order);
}

extern "C" {

int spral_hungarian_scale_sym(int n, const int64_t *ptr, const int *row, const double *val,
                              double *scaling, int *match, int scale_if_singular,
                              int *matched) {
  if (n < 0) return -3;
  *matched = 0;
  if (n == 0) return 0;
  try {
    // costs: column maximum of log|a| minus log|a| (maximum product ->
    // minimum sum, every cost >= 0)
    Graph g = expand_lower(n, ptr, row, val, true);
    std::vector<double> cmax(n);
    for (int j = 0; j < n; ++j) {
      double mx = -kInf;
      for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) mx = std::max(mx, g.val[k]);
      cmax[j] = mx;
      for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) g.val[k] = mx - g.val[k];
    }
    std::vector<int> rm;
    std::vector<double> u, v;
    const int num = min_cost_assignment(g, rm, u, v);
    *matched = num;
    int flag = 0;
    std::vector<double> rs(n), cs(n);
    if (num == n) {
      // duals -> log scalings, balanced so both sets have the same mean
      double ravg = 0.0, cavg = 0.0;
      for (int i = 0; i < n; ++i) {
        rs[i] = u[i];
        cs[i] = v[i] - cmax[i];
      }
      for (int i = 0; i < n; ++i) ravg += rs[i];
      for (int i = 0; i < n; ++i) cavg += cs[i];
      ravg /= n;
      cavg /= n;
      const double adjust = (ravg - cavg) / 2;
      for (int i = 0; i < n; ++i) {
        rs[i] -= adjust;
        cs[i] += adjust;
      }
    } else {
      // structurally singular: as the reference, the mean of the row and
      // column duals less the column maximum is the symmetric log scaling,
      // for every index (hungarian_wrapper's reduced problem keeps all rows:
      // its filter tests for a negative match, and an unmatched row holds
      // 0, so the second assignment is this one again)
      flag = scale_if_singular ? 1 : -2;
      for (int i = 0; i < n; ++i) rs[i] = (u[i] + v[i] - cmax[i]) / 2;
      cs = rs;
    }
    for (int i = 0; i < n; ++i) scaling[i] = std::exp((rs[i] + cs[i]) / 2);
    if (match)
      for (int i = 0; i < n; ++i) match[i] = rm[i] >= 0 ? rm[i] + 1 : 0;
    return flag;
  } catch (std::bad_alloc &) {
    return -1;
  }
}

int spral_auction_scale_sym(int n, const int64_t *ptr, const int *row, const double *val,
                            double *scaling, int *match, int max_iterations,
                            const int *max_unchanged, const float *min_proportion,
                            float eps_initial, int *matched, int *iterations,
                            int *unmatchable) {
  if (n < 0) return -3;
  *matched = *iterations = *unmatchable = 0;
  if (n == 0) return 0;
  try {
    // weights: 2 max + 1 - (column max of log|a| - log|a|), so that a
    // larger matching always pays (auction_match, src/scaling.f90:1481-1586)
    Graph g = expand_lower(n, ptr, row, val, true);
    std::vector<double> cmax(n, 0.0);
    for (int j = 0; j < n; ++j) {
      if (g.ptr[j + 1] <= g.ptr[j]) continue;
      double mx = -kInf;
      for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) mx = std::max(mx, g.val[k]);
      cmax[j] = mx;
      for (int64_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k) g.val[k] = mx - g.val[k];
    }
    double maxentry = -kInf;
    for (double x : g.val) maxentry = std::max(maxentry, x);
    maxentry = 2 * maxentry + 1;
    for (double &x : g.val) x = maxentry - x;
    // auction (src/scaling.f90:1328-1466): one column bids at a time for the
    // row of its best net value, paying the margin over its second best
    // plus eps; outbid columns rejoin the next pass
    std::vector<double> u(n, 0.0), v(n);
    for (int j = 0; j < n; ++j) v[j] = -cmax[j];
    std::vector<int> colmatch(n, 0), owner(n, -1), next(n);
    for (int j = 0; j < n; ++j) next[j] = j;
    int tail = n, unmatched = n, prev = -1, nunchanged = 0, itr;
    double eps = (double)eps_initial;
    int unmatchable_count = 0;
    for (itr = 1; itr <= max_iterations; ++itr) {
      if (unmatched == 0) break;
      if (unmatched != prev) nunchanged = 0;
      prev = unmatched;
      ++nunchanged;
      const float prop = (float)(n - unmatched) / (float)n;  // default-real arithmetic
      bool stop = false;
      for (int t = 0; t < 3; ++t)
        if (nunchanged >= max_unchanged[t] && prop >= min_proportion[t]) stop = true;
      if (stop) break;
      eps = std::min(1.0, eps + 1.0 / (n + 1));
      int insert = 0;
      for (int cp = 0; cp < tail; ++cp) {
        const int col = next[cp];
        if (colmatch[col] != 0) continue;  // matched (row + 1) or ineligible (-1)
        if (g.ptr[col] == g.ptr[col + 1]) continue;
        int64_t k = g.ptr[col];
        int bestr = g.row[k];
        double bestu = g.val[k] - u[bestr], bestv = -kInf;
        for (++k; k < g.ptr[col + 1]; ++k) {
          const double x = g.val[k] - u[g.row[k]];
          if (x > bestu) {
            bestv = bestu;
            bestr = g.row[k];
            bestu = x;
          } else if (x > bestv) {
            bestv = x;
          }
        }
        if (bestv == -kInf) bestv = 0.0;
        if (bestu > 0) {
          u[bestr] = u[bestr] + bestu - bestv + eps;
          v[col] = bestv - eps;
          colmatch[col] = bestr + 1;
          --unmatched;
          const int kk = owner[bestr];
          owner[bestr] = col;
          if (kk >= 0) {
            colmatch[kk] = 0;
            ++unmatched;
            next[insert++] = kk;
          }
        } else {
          colmatch[col] = -1;
          --unmatched;
          ++unmatchable_count;
        }
      }
      tail = insert;
    }
    *iterations = itr - 1;
    *unmatchable = unmatchable_count;
    int nm = 0;
    std::vector<int> rowmatch(n, 0);
    for (int j = 0; j < n; ++j)
      if (colmatch[j] > 0) {
        ++nm;
        rowmatch[colmatch[j] - 1] = j + 1;
      }
    *matched = nm;
    std::vector<double> rs(n), cs(n);
    for (int i = 0; i < n; ++i) rs[i] = -u[i] + maxentry;
    for (int j = 0; j < n; ++j) cs[j] = -v[j] - cmax[j];
    // square post-processing (match_postproc): equal mean row and column logs
    double ravg = 0.0, cavg = 0.0;
    for (int i = 0; i < n; ++i) ravg += rs[i];
    for (int i = 0; i < n; ++i) cavg += cs[i];
    ravg /= n;
    cavg /= n;
    const double adjust = (ravg - cavg) / 2;
    for (int i = 0; i < n; ++i) {
      rs[i] -= adjust;
      cs[i] += adjust;
    }
    for (int i = 0; i < n; ++i) scaling[i] = std::exp((rs[i] + cs[i]) / 2);
    if (match)
      for (int i = 0; i < n; ++i) match[i] = rowmatch[i];
    return 0;
  } catch (std::bad_alloc &) {
    return -1;
  }
}

int spral_match_order(int n, const int64_t *ptr, const int *row, const double *val,
                      int *order, double *scale) {
  if (n < 0) return -2;
  if (n == 0) return 0;
  try {
    Graph full = expand_lower(n, ptr, row, val, false);
    for (double &x : full.val) x = std::fabs(x);
    std::vector<double> ls;
    std::vector<int> cperm;
    const int flag = match_and_log_scale(full, ls, cperm);
    // split the matching into 1- and 2-cycles: a cycle is walked from its
    // smallest index, consecutive indices paired, an odd cycle leaving a
    // singleton; pair[i] = partner, -1 singleton, -2 unmatched
    std::vector<int> pair(n, 0);
    for (int i = 0; i < n; ++i) {
      if (pair[i] != 0) continue;
      int j = i;
      for (;;) {
        if (cperm[j] < 0) {
          pair[j] = -2;
          break;
        }
        if (cperm[j] == i) {
          pair[j] = -1;
          break;
        }
        const int jj = cperm[j];
        pair[j] = jj + 1;  // stored 1-based while walking: 0 means unseen
        pair[jj] = j + 1;
        j = cperm[jj];
        if (j == i) break;
      }
    }
    for (int i = 0; i < n; ++i)
      if (pair[i] > 0) pair[i] -= 1;
    // compressed numbering: a pair's first (smaller) index names it
    std::vector<int> o2n(n), n2o;
    for (int i = 0; i < n; ++i) {
      const int j = pair[i];
      if (j >= 0 && j < i) continue;
      o2n[i] = (int)n2o.size();
      n2o.push_back(i);
      if (j >= 0) o2n[j] = o2n[i];
    }
    const int nc = (int)n2o.size();
    // compressed lower pattern (diagonal kept), 1-based for the ordering
    std::vector<int64_t> cp(nc + 1, 1);
    std::vector<int> cr;
    std::vector<int> seen(nc, -1);
    for (int c = 0; c < nc; ++c) {
      const int i = n2o[c];
      const int cols[2] = {i, pair[i]};
      for (int t = 0; t < 2; ++t) {
        const int col = cols[t];
        if (col < 0) continue;
        for (int64_t k = full.ptr[col]; k < full.ptr[col + 1]; ++k) {
          const int kr = o2n[full.row[k]];
          if (seen[kr] == c || kr < c) continue;
          seen[kr] = c;
          cr.push_back(kr + 1);
        }
      }
      cp[c + 1] = (int64_t)cr.size() + 1;
    }
    std::vector<int> corder(nc);
    ssids_hip::order_amd(nc, cp.data(), cr.data(), corder.data());
    std::vector<int> seq(nc);
    for (int c = 0; c < nc; ++c) seq[corder[c] - 1] = c;
    int k = 1;
    for (int p = 0; p < nc; ++p) {
      const int j = n2o[seq[p]];
      order[j] = k++;
      if (pair[j] >= 0) order[pair[j]] = k++;
    }
    for (int i = 0; i < n; ++i) scale[i] = std::exp(ls[i]);
    return flag;
  } catch (std::bad_alloc &) {
    return -1;
  }
}

}  // extern "C"
