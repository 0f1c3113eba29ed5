// Host-side symbolic analysis for the SSIDS factorize path.
//
// Produces exactly the symbolic data SSIDS's analyse phase hands to a numeric
// subtree: the final pivot order, supernode partition (sptr), assembly tree
// (sparent), row lists (rptr/rlist, each list ascending) and the A->front
// scatter map (nptr/nlist).  Every routine below follows the semantics of the
// reference analyse so the output is identical integer-for-integer:
//   elimination tree      src/core_analyse.f90:173-219   (find_etree, Liu)
//   postorder             src/core_analyse.f90:228-331   (find_postorder)
//   column counts         src/core_analyse.f90:367-484   (find_col_counts, GNP)
//   supernode amalgamation src/core_analyse.f90:496-679  (find_supernodes)
//   child ordering        src/core_analyse.f90:685-777   (sort_by_val[_ms])
//   merge test / merge    src/core_analyse.f90:782-829   (do_merge, merge_nodes)
//   statistics            src/core_analyse.f90:838-880   (calc_stats)
//   row lists             src/core_analyse.f90:887-967   (find_row_lists)
//   row-list sort         src/core_analyse.f90:972-1027  (dbl_tr_sort)
//   A->L map              src/ssids/anal.F90:1137-1239   (build_map)
//   pattern expansion     src/ssids/anal.F90:42-85       (expand_pattern)
// All index arrays use the reference's 1-based convention at the interface
// (the Fortran host passes them through unchanged).  Internally vectors are
// sized n+2 and indexed 1-based so each step reads like the definition.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host.hpp"

namespace ssids_hip {

namespace {

using i64 = int64_t;

// Elimination tree of P A P^T (Liu's algorithm with path compression).
// aptr/arow: full (both triangles) pattern, 1-based.  perm(i) = pivot
// position of column i; invp its inverse.  parent(piv) in 1..n+1.
void etree(int n, const std::vector<i64>& aptr, const std::vector<int>& arow,
           const std::vector<int>& perm, const std::vector<int>& invp,
           std::vector<int>& parent) {
  std::vector<int> vforest(n + 2, n + 1);
  for (int piv = 1; piv <= n; ++piv) {
    int col = invp[piv];
    for (i64 k = aptr[col]; k < aptr[col + 1]; ++k) {
      int j = perm[arow[k]];
      if (j >= piv) continue;
      int a = j;
      while (vforest[a] < piv) {
        int nxt = vforest[a];
        vforest[a] = piv;
        a = nxt;
      }
      if (vforest[a] == piv) continue;
      parent[a] = piv;
      vforest[a] = piv;
    }
    parent[piv] = n + 1;
  }
}

// Relabel the tree in depth-first order.  Children are visited so that the
// first child (in original numbering) is numbered lowest; roots with an empty
// column are numbered last and reduce realn.  Updates perm, invp, parent.
void postorder(int n, int& realn, const std::vector<i64>& aptr,
               std::vector<int>& perm, std::vector<int>& invp,
               std::vector<int>& parent) {
  realn = n;
  std::vector<int> chead(n + 2, -1), cnext(n + 2, -1);
  for (int i = n; i >= 1; --i) {
    int p = parent[i];
    cnext[i] = chead[p];
    chead[p] = i;
  }
  std::vector<int> label(n + 2, 0), stack;
  stack.reserve(n + 1);
  stack.push_back(n + 1);
  int next_id = n + 1;
  auto empty_col = [&](int v) { return aptr[invp[v] + 1] - aptr[invp[v]] == 0; };
  while (!stack.empty()) {
    int v = stack.back();
    stack.pop_back();
    label[v] = next_id--;
    if (v == n + 1) {
      for (int c = chead[v]; c != -1; c = cnext[c])
        if (!empty_col(c)) stack.push_back(c);
      for (int c = chead[v]; c != -1; c = cnext[c])
        if (empty_col(c)) { --realn; stack.push_back(c); }
    } else {
      for (int c = chead[v]; c != -1; c = cnext[c]) stack.push_back(c);
    }
  }
  std::vector<int> tmp(invp.begin(), invp.end());
  for (int i = 1; i <= n; ++i) invp[label[i]] = tmp[i];
  for (int i = 1; i <= n; ++i) perm[invp[i]] = i;
  for (int i = 1; i <= n; ++i) tmp[i] = label[parent[i]];
  for (int i = 1; i <= n; ++i) parent[label[i]] = tmp[i];
}

// Top of the partial elimination tree containing u (path halving).
int forest_find(std::vector<int>& vf, int u) {
  int prev = -1, cur = u;
  while (vf[cur] != 0) {
    prev = cur;
    cur = vf[cur];
    if (vf[cur] != 0) vf[prev] = vf[cur];
  }
  return cur;
}

// Column counts of L (diagonal included) by net-weight accumulation
// (Gilbert, Ng & Peyton).  cc has n+2 entries; cc(n+1) is a sentinel.
void col_counts(int n, const std::vector<i64>& aptr, const std::vector<int>& arow,
                const std::vector<int>& perm, const std::vector<int>& invp,
                const std::vector<int>& parent, std::vector<int>& cc) {
  std::vector<int> first(n + 2);
  for (int i = 1; i <= n + 1; ++i) first[i] = i;
  for (int i = 1; i <= n; ++i) {
    int p = parent[i];
    first[p] = std::min(first[i], first[p]);
    cc[i] = (first[i] == i) ? 1 : 0;
  }
  cc[n + 1] = n + 1;
  std::vector<int> vforest(n + 2, 0), last_p(n + 2, 0), last_nbr(n + 2, 0);
  for (int piv = 1; piv <= n; ++piv) {
    int col = invp[piv];
    for (i64 k = aptr[col]; k < aptr[col + 1]; ++k) {
      int u = perm[arow[k]];
      if (u <= piv) continue;
      if (first[piv] > last_nbr[u]) {
        cc[piv] += 1;
        int pp = last_p[u];
        if (pp != 0) {
          int lca = forest_find(vforest, pp);
          cc[lca] -= 1;
        }
        last_p[u] = piv;
      }
      last_nbr[u] = piv;
    }
    int par = parent[piv];
    cc[par] += cc[piv] - 1;
    vforest[piv] = par;
  }
}

// Stable-by-construction ordering of a child list into decreasing val[];
// insertion variant below 16 entries, top-down merge above (the two give
// different orders among equal keys, and the tree depends on that order).
void sort_desc(int cnt, int* idx, const std::vector<int>& val);

void sort_desc_merge(int cnt, int* idx, const std::vector<int>& val) {
  if (cnt <= 1) return;
  if (cnt < 16) { sort_desc(cnt, idx, val); return; }
  int mid = (cnt - 1) / 2 + 1;
  sort_desc_merge(mid, idx, val);
  sort_desc_merge(cnt - mid, idx + mid, val);
  std::vector<int> left(idx, idx + mid);
  int j = 0, k = mid, out = 0;
  int jj = left[0], kk = idx[k];
  for (out = 0; out < cnt; ++out) {
    if (val[jj] >= val[kk]) {
      idx[out] = jj;
      if (++j >= mid) break;
      jj = left[j];
    } else {
      idx[out] = kk;
      if (++k >= cnt) break;
      kk = idx[k];
    }
  }
  for (int t = out + 1; j < mid; ++t, ++j) idx[t] = left[j];
}

void sort_desc(int cnt, int* idx, const std::vector<int>& val) {
  if (cnt >= 16) { sort_desc_merge(cnt, idx, val); return; }
  // Insert idx[kor-1] into the already ordered tail idx[kor..cnt-1].
  for (int kor = cnt - 1; kor >= 1; --kor) {
    int item = idx[kor - 1];
    int key = val[item];
    int k = kor;
    for (; k < cnt; ++k) {
      if (key >= val[idx[k]]) break;
      idx[k - 1] = idx[k];
    }
    idx[k - 1] = item;
  }
}

}  // namespace

int basic_analyse(int n, const i64* ptr_full, const int* row_full, int* order,
                  int nemin, SymbolicData& out) {
  // 1-based copies of the full pattern
  std::vector<i64> aptr(n + 2);
  for (int i = 0; i <= n; ++i) aptr[i + 1] = ptr_full[i];
  i64 nnz = (n > 0) ? aptr[n + 1] - 1 : 0;
  std::vector<int> arow(nnz + 1);
  for (i64 k = 0; k < nnz; ++k) arow[k + 1] = row_full[k];

  std::vector<int> perm(n + 2), invp(n + 2);
  for (int i = 1; i <= n; ++i) perm[i] = order[i - 1];
  for (int i = 1; i <= n; ++i) invp[perm[i]] = i;

  int info = 0;
  int realn = n;
  std::vector<int> parent(n + 2, 0);
  etree(n, aptr, arow, perm, invp, parent);
  postorder(n, realn, aptr, perm, invp, parent);
  if (realn != n) info = 1;  // structurally singular (empty columns)

  std::vector<int> cc(n + 2, 0);
  col_counts(n, aptr, arow, perm, invp, parent, cc);

  // ---- supernode amalgamation ----
  std::vector<int> nelim(n + 2, 1), nvert(n + 2, 1), vhead(n + 2, -1),
      vnext(n + 2, -1);
  std::vector<i64> ezero(n + 2, 0);
  const i64 kHuge = INT64_MAX;
  ezero[n + 1] = kHuge;
  nelim[n + 1] = n + 1 + nemin;
  std::vector<int> chead(n + 2, -1), cnext(n + 2, -1);
  for (int i = realn; i >= 1; --i) {
    int p = parent[i];
    cnext[i] = chead[p];
    chead[p] = i;
  }
  std::vector<char> mark(n + 2, 0);
  std::vector<int> child(n + 1);
  for (int par = 1; par <= n + 1; ++par) {
    int nch = 0;
    for (int c = chead[par]; c != -1; c = cnext[c]) child[nch++] = c;
    sort_desc(nch, child.data(), cc);
    for (int j = 0; j < nch; ++j) {
      int node = child[j];
      bool merge;
      if (ezero[par] == kHuge) {
        merge = false;
      } else {
        merge = (cc[par] == cc[node] - 1 && nelim[par] == 1) ||
                (nelim[par] < nemin && nelim[node] < nemin);
      }
      if (merge) {
        vnext[node] = vhead[par];
        vhead[par] = node;
        ezero[par] = ezero[par] + ezero[node] +
                     (i64)(cc[par] - 1 + nelim[par] - cc[node] + 1) * nelim[par];
        nelim[par] += nelim[node];
        nvert[par] += nvert[node];
        mark[node] = 0;
      } else {
        mark[node] = 1;
      }
    }
  }
  std::vector<int> sptr(n + 2), npar(n + 2), scc(n + 2), sperm(n + 2),
      vmap(n + 2);
  int nnodes = 0, v = 1;
  std::vector<int> stack;
  for (int node = 1; node <= realn; ++node) {
    if (!mark[node]) continue;
    ++nnodes;
    sptr[nnodes] = v;
    npar[nnodes] = parent[node];
    scc[nnodes] = cc[node] + nelim[node] - 1;
    v += nvert[node];
    int k = v;
    stack.clear();
    stack.push_back(node);
    while (!stack.empty()) {
      int i = stack.back();
      stack.pop_back();
      sperm[i] = --k;
      vmap[i] = nnodes;
      if (vnext[i] != -1) stack.push_back(vnext[i]);
      if (vhead[i] != -1) stack.push_back(vhead[i]);
    }
  }
  sptr[nnodes + 1] = v;
  vmap[n + 1] = nnodes + 1;
  for (int i = realn + 1; i <= n; ++i) sperm[i] = i;
  std::vector<int> sparent(nnodes + 2);
  for (int s = 1; s <= nnodes; ++s) sparent[s] = vmap[npar[s]];

  // ---- apply supernodal permutation to perm, invp, cc ----
  {
    std::vector<int> tmp(n + 2);
    for (int i = 1; i <= n; ++i) tmp[i] = cc[i];
    for (int i = 1; i <= n; ++i) cc[sperm[i]] = tmp[i];
    for (int i = 1; i <= n; ++i) tmp[i] = invp[i];
    for (int i = 1; i <= n; ++i) invp[sperm[i]] = tmp[i];
    for (int i = 1; i <= n; ++i) perm[invp[i]] = i;
  }

  // ---- row lists ----
  std::vector<i64> rptr(nnodes + 2);
  i64 total = 0;
  for (int s = 1; s <= nnodes; ++s) total += scc[s];
  std::vector<int> rlist(total + 1);
  {
    std::vector<int> seen(n + 2, 0), ch(nnodes + 2, -1), cn(nnodes + 2, -1);
    for (int s = nnodes; s >= 1; --s) {
      int p = sparent[s];
      cn[s] = ch[p];
      ch[p] = s;
    }
    rptr[1] = 1;
    for (int s = 1; s <= nnodes; ++s) {
      rptr[s + 1] = rptr[s] + scc[s];
      i64 idx = rptr[s];
      for (int piv = sptr[s]; piv < sptr[s + 1]; ++piv) {
        seen[piv] = s;
        rlist[idx++] = piv;
      }
      for (int c = ch[s]; c != -1; c = cn[c]) {
        for (i64 k = rptr[c]; k < rptr[c + 1]; ++k) {
          int j = rlist[k];
          if (j < sptr[s] || seen[j] == s) continue;
          seen[j] = s;
          rlist[idx++] = j;
        }
      }
      for (int piv = sptr[s]; piv < sptr[s + 1]; ++piv) {
        int col = invp[piv];
        for (i64 k = aptr[col]; k < aptr[col + 1]; ++k) {
          int j = perm[arow[k]];
          if (j < piv || seen[j] == s) continue;
          seen[j] = s;
          rlist[idx++] = j;
        }
      }
    }
    // ascending order within each list (the double-transpose sort yields this)
    for (int s = 1; s <= nnodes; ++s)
      std::sort(rlist.begin() + rptr[s], rlist.begin() + rptr[s + 1]);
  }

  // ---- statistics (64-bit throughout) ----
  i64 nfact = 0, nflops = 0;
  for (int s = 1; s <= nnodes; ++s) {
    i64 ne = sptr[s + 1] - sptr[s];
    i64 m = scc[s] - ne;
    nfact += ne * (ne + 1) / 2 + ne * m;
    for (i64 j = 1; j <= ne; ++j) nflops += (m + j) * (m + j);
  }

  // ---- outputs (1-based values, 0-based storage) ----
  out.n = n;
  out.nnodes = nnodes;
  out.sptr.assign(sptr.begin() + 1, sptr.begin() + nnodes + 2);
  out.sparent.assign(sparent.begin() + 1, sparent.begin() + nnodes + 1);
  out.rptr.assign(rptr.begin() + 1, rptr.begin() + nnodes + 2);
  out.rlist.assign(rlist.begin() + 1, rlist.end());
  out.num_factor = nfact;
  out.num_flops = nflops;
  for (int i = 1; i <= n; ++i) order[i - 1] = perm[i];
  return info;
}

// Build nptr/nlist for the lower-triangle CSC (ptr,row) (1-based) given the
// final order.  lcol(nlist(2,k)) = val(nlist(1,k)), nlist(2,k) being the
// 1-based column-major position within the node's blkm x blkn front.
void build_map(int n, const i64* ptr, const int* row, const int* order,
               SymbolicData& sd) {
  i64 nz = ptr[n] - 1;
  std::vector<int> perm(n + 1), invp(n + 1);
  for (int i = 1; i <= n; ++i) perm[i] = std::abs(order[i - 1]);
  for (int i = 1; i <= n; ++i)
    if (perm[i] >= 1 && perm[i] <= n) invp[perm[i]] = i;
  // transpose of the strictly-lower part, remembering origin positions
  std::vector<i64> tptr(n + 3, 0);
  for (int i = 1; i <= n; ++i)
    for (i64 jj = ptr[i - 1]; jj < ptr[i]; ++jj) {
      int k = row[jj - 1];
      if (k == i) continue;
      tptr[k + 2]++;
    }
  tptr[1] = tptr[2] = 1;
  for (int i = 1; i <= n; ++i) tptr[i + 2] += tptr[i + 1];
  std::vector<int> trow(nz + 1);
  std::vector<i64> origin(nz + 1);
  for (int i = 1; i <= n; ++i)
    for (i64 jj = ptr[i - 1]; jj < ptr[i]; ++jj) {
      int k = row[jj - 1];
      if (k == i) continue;
      trow[tptr[k + 1]] = i;
      origin[tptr[k + 1]] = jj;
      tptr[k + 1]++;
    }
  int nnodes = sd.nnodes;
  sd.nptr.assign(nnodes + 1, 0);
  sd.nlist.assign(2 * (size_t)nz, 0);
  std::vector<int> map(n + 1, 0);
  i64 pp = 1;
  for (int s = 1; s <= nnodes; ++s) {
    i64 r0 = sd.rptr[s - 1], r1 = sd.rptr[s];
    int blkm = (int)(r1 - r0);
    sd.nptr[s - 1] = pp;
    for (i64 jj = r0; jj < r1; ++jj) map[sd.rlist[jj - 1]] = (int)(jj - r0 + 1);
    int s0 = sd.sptr[s - 1], s1 = sd.sptr[s];
    for (int j = s0; j < s1; ++j) {
      int col = invp[j];
      for (i64 k = tptr[col]; k < tptr[col + 1]; ++k) {
        int r = perm[trow[k]];
        if (r < j) continue;
        sd.nlist[2 * (pp - 1) + 1] = (i64)(j - s0) * blkm + map[r];
        sd.nlist[2 * (pp - 1) + 0] = origin[k];
        ++pp;
      }
    }
    for (int j = s0; j < s1; ++j) {
      int col = invp[j];
      for (i64 k = ptr[col - 1]; k < ptr[col]; ++k) {
        int r = perm[row[k - 1]];
        if (r < j) continue;
        sd.nlist[2 * (pp - 1) + 1] = (i64)(j - s0) * blkm + map[r];
        sd.nlist[2 * (pp - 1) + 0] = k;
        ++pp;
      }
    }
  }
  sd.nptr[nnodes] = pp;
  sd.nlist.resize(2 * (size_t)(pp - 1));
}

// Full pattern (both triangles) from lower CSC, 1-based in and out.
void expand_pattern(int n, const i64* ptr, const int* row, std::vector<i64>& aptr,
                    std::vector<int>& arow) {
  i64 nz = ptr[n] - 1;
  aptr.assign(n, 0);  // aptr[j-1] for column j (1-based values)
  std::vector<i64> cnt(n + 2, 0);
  for (int j = 1; j <= n; ++j)
    for (i64 kk = ptr[j - 1]; kk < ptr[j]; ++kk) {
      int i = row[kk - 1];
      cnt[i]++;
      if (i != j) cnt[j]++;
    }
  // Fill from the top of each column downwards, as the reference does, so the
  // per-column entry order is the same.
  std::vector<i64> end(n + 2, 0);
  end[1] = cnt[1];
  for (int j = 2; j <= n; ++j) end[j] = end[j - 1] + cnt[j];
  arow.assign(2 * nz + 1, 0);
  std::vector<i64> pos(end);
  for (int j = 1; j <= n; ++j)
    for (i64 kk = ptr[j - 1]; kk < ptr[j]; ++kk) {
      int i = row[kk - 1];
      arow[pos[i]] = j;
      pos[i]--;
      if (i == j) continue;
      arow[pos[j]] = i;
      pos[j]--;
    }
  for (int j = 1; j <= n; ++j) aptr[j - 1] = pos[j] + 1;
  aptr.push_back(end[n] + 1);
  // compact: arow is 1-based (index 0 unused); shift to 0-based storage
  std::vector<int> r(arow.begin() + 1, arow.begin() + 1 + (end[n]));
  arow.swap(r);
}

}  // namespace ssids_hip
