// Fill-reducing orderings used in place of METIS (absent from this image).
//
// The reference reaches METIS_NodeND through src/metis5_wrapper.F90:102-182 or
// takes a user order (options%ordering=0, checked by src/ssids/anal.F90:152-202).
// Both orderings here are deterministic so the same order can be fed to the
// numeric path and to any CPU baseline:
//  * order_grid_nd: geometric nested dissection of an nx*ny*nz grid whose
//    variable (i,j,k) is i + nx*(j + ny*k) (0-based), separators are middle
//    planes of the longest box dimension, ordered after both halves, each
//    with the vertices next to an enclosing separator last;
//  * order_min_degree: exact minimum degree on the explicit elimination graph
//    (ties to the lowest index), for moderate n;
//  * order_amd: approximate minimum degree on the quotient graph, for large
//    general patterns (the explicit graph above grows with the fill).
#include <algorithm>
#include <cstdint>
#include <set>
#include <vector>

#include "host.hpp"

namespace ssids_hip {

namespace {

struct Box {
  int lo[3], hi[3];  // half-open
};

// Separator vertices that touch the box's outside (an ancestor separator,
// never the domain boundary) come last, after the rest of the separator.
// Each such vertex brings ancestor rows of its own into the fill; with them
// last, the separator's interior shares one structure and is a single
// supernode (instead of one per grid line, as in plain row-major order), and
// the boundary vertices follow as small fronts of their own.  Ordered first
// instead, they fill into the enclosing separators and amalgamate into the
// root: on 64^3 a 6,112-column root against 4,128, and 241 pivot steps in
// sequence over the tree levels against 215.
void nd_box(const Box& b, const int dims[3], std::vector<int>& seq) {
  const int nx = dims[0], ny = dims[1];
  int ext[3] = {b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]};
  if (ext[0] <= 0 || ext[1] <= 0 || ext[2] <= 0) return;
  int vol = ext[0] * ext[1] * ext[2];
  int d = 0;
  if (ext[1] > ext[d]) d = 1;
  if (ext[2] > ext[d]) d = 2;
  if (vol <= 8 || ext[d] < 3) {
    for (int k = b.lo[2]; k < b.hi[2]; ++k)
      for (int j = b.lo[1]; j < b.hi[1]; ++j)
        for (int i = b.lo[0]; i < b.hi[0]; ++i) seq.push_back(i + nx * (j + ny * k));
    return;
  }
  int mid = b.lo[d] + ext[d] / 2;
  Box left = b, right = b, sep = b;
  left.hi[d] = mid;
  right.lo[d] = mid + 1;
  sep.lo[d] = mid;
  sep.hi[d] = mid + 1;
  nd_box(left, dims, seq);
  nd_box(right, dims, seq);
  auto touches_outside = [&](const int p[3]) {
    for (int e = 0; e < 3; ++e) {
      if (p[e] == b.lo[e] && b.lo[e] > 0) return true;
      if (p[e] == b.hi[e] - 1 && b.hi[e] < dims[e]) return true;
    }
    return false;
  };
  for (int pass = 0; pass < 2; ++pass)
    for (int k = sep.lo[2]; k < sep.hi[2]; ++k)
      for (int j = sep.lo[1]; j < sep.hi[1]; ++j)
        for (int i = sep.lo[0]; i < sep.hi[0]; ++i) {
          const int p[3] = {i, j, k};
          if (touches_outside(p) == (pass == 1)) seq.push_back(i + nx * (j + ny * k));
        }
}

}  // namespace

void order_grid_nd(int nx, int ny, int nz, int* order) {
  std::vector<int> seq;
  seq.reserve((size_t)nx * ny * nz);
  Box b{{0, 0, 0}, {nx, ny, nz}};
  const int dims[3] = {nx, ny, nz};
  nd_box(b, dims, seq);
  for (size_t p = 0; p < seq.size(); ++p) order[seq[p]] = (int)p + 1;
}

// ptr/row: lower-triangle CSC, 1-based.  Returns 0 on success.
int order_min_degree(int n, const int64_t* ptr, const int* row, int* order) {
  std::vector<std::vector<int>> adj(n);
  for (int j = 0; j < n; ++j)
    for (int64_t k = ptr[j] - 1; k < ptr[j + 1] - 1; ++k) {
      int i = row[k] - 1;
      if (i == j || i < 0 || i >= n) continue;
      adj[i].push_back(j);
      adj[j].push_back(i);
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::set<std::pair<int, int>> pq;  // (degree, index)
  std::vector<char> done(n, 0);
  for (int v = 0; v < n; ++v) pq.insert({(int)adj[v].size(), v});
  std::vector<int> merged;
  int pos = 0;
  while (!pq.empty()) {
    int v = pq.begin()->second;
    pq.erase(pq.begin());
    done[v] = 1;
    order[v] = ++pos;
    const std::vector<int> nb = adj[v];
    for (int u : nb) {
      pq.erase({(int)adj[u].size(), u});
      merged.clear();
      std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(),
                     std::back_inserter(merged));
      std::vector<int> nu;
      nu.reserve(merged.size());
      for (int w : merged)
        if (w != u && w != v && !done[w]) nu.push_back(w);
      adj[u].swap(nu);
      pq.insert({(int)adj[u].size(), u});
    }
    std::vector<int>().swap(adj[v]);
  }
  return 0;
}


// Approximate minimum degree: Amestoy, Davis and Duff, "An approximate
// minimum degree ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4),
// 886-905 (1996).  The elimination graph is held as a quotient graph: each
// principal variable i keeps its remaining variable neighbours A_i and the
// elements E_i it belongs to; each element e (a pivot already eliminated)
// keeps its variable list L_e.  Eliminating the pivot p of least approximate
// degree makes p an element with L_p = A_p u (union of L_e, e in E_p) and
// absorbs every element of E_p.  For i in L_p the external degree is bounded
// as in the paper:
//   d_i <- min(nleft - |i|, d_i + |L_p \ i|,
//              |A_i \ i| + |L_p \ i| + sum_{e in E_i, e != p} |L_e \ L_p|),
// with |L_e \ L_p| counted once per element from the members of L_p.
// Elements wholly inside L_p are absorbed into p, and variables of L_p with
// the same A_i and E_i (indistinguishable) merge into one supervariable,
// found by hashing those lists.  Sizes are weighted by supervariable size.
// Degrees sit in bucket lists; the pivot is the least degree, ties to the
// most recently inserted.  Deterministic.
int order_amd(int n, const int64_t* ptr, const int* row, int* order) {
  if (n <= 0) return 0;
  std::vector<std::vector<int>> A(n), E(n), Le(n), members(n);
  for (int j = 0; j < n; ++j)
    for (int64_t k = ptr[j] - 1; k < ptr[j + 1] - 1; ++k) {
      const int i = row[k] - 1;
      if (i == j || i < 0 || i >= n) continue;
      A[i].push_back(j);
      A[j].push_back(i);
    }
  for (auto& a : A) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  enum : char { VAR = 0, ELEM = 1, DEAD = 2, MERGED = 3 };
  std::vector<char> st(n, VAR);
  std::vector<int> nv(n, 1), deg(n), edeg(n, 0);
  // degree buckets (doubly linked)
  std::vector<int> head(n + 1, -1), nxt(n, -1), prv(n, -1);
  auto bucket_in = [&](int i) {
    const int d = std::min(std::max(deg[i], 0), n);
    prv[i] = -1;
    nxt[i] = head[d];
    if (head[d] >= 0) prv[head[d]] = i;
    head[d] = i;
  };
  auto bucket_out = [&](int i) {
    const int d = std::min(std::max(deg[i], 0), n);
    if (prv[i] >= 0) nxt[prv[i]] = nxt[i];
    else head[d] = nxt[i];
    if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
    nxt[i] = prv[i] = -1;
  };
  for (int i = 0; i < n; ++i) {
    deg[i] = (int)A[i].size();
    bucket_in(i);
  }
  std::vector<int> mark(n, -1), wmark(n, -1), w(n, 0), dpart(n, 0), dold(n, 0);
  std::vector<char> inbucket(n, 1);
  int stage = 0, mindeg = 0, pos = 0, nleft = n;
  std::vector<int> Lp;
  std::vector<std::pair<unsigned long long, int>> hv;
  while (nleft > 0) {
    // ---- pivot: least approximate degree
    while (mindeg <= n && head[mindeg] < 0) ++mindeg;
    if (mindeg > n) return -1;  // cannot happen: every live variable is bucketed
    const int p = head[mindeg];
    bucket_out(p);
    inbucket[p] = 0;
    ++stage;
    // ---- L_p = A_p u L_e (e in E_p), principal variables other than p
    Lp.clear();
    mark[p] = stage;
    for (int v : A[p])
      if (st[v] == VAR && mark[v] != stage) {
        mark[v] = stage;
        Lp.push_back(v);
      }
    for (int e : E[p]) {
      if (st[e] != ELEM) continue;
      for (int v : Le[e])
        if (st[v] == VAR && mark[v] != stage) {
          mark[v] = stage;
          Lp.push_back(v);
        }
      st[e] = DEAD;  // absorbed into p
      std::vector<int>().swap(Le[e]);
    }
    std::vector<int>().swap(A[p]);
    std::vector<int>().swap(E[p]);
    // p and its merged members take the next positions
    order[p] = ++pos;
    for (int v : members[p]) order[v] = ++pos;
    nleft -= nv[p];
    st[p] = ELEM;
    int degme = 0;
    for (int i : Lp) degme += nv[i];
    edeg[p] = degme;
    // ---- |L_e \ L_p| for the elements next to L_p
    for (int i : Lp) {
      if (inbucket[i]) {
        bucket_out(i);
        inbucket[i] = 0;
      }
      for (int e : E[i]) {
        if (st[e] != ELEM || e == p) continue;
        if (wmark[e] != stage) {
          wmark[e] = stage;
          w[e] = edeg[e];
        }
        w[e] -= nv[i];
      }
    }
    // ---- update each i in L_p: lists, element absorption, degree bound
    for (int i : Lp) {
      auto& Ei = E[i];
      size_t q = 0;
      int esum = 0;
      for (int e : Ei) {
        if (st[e] != ELEM || e == p) continue;
        if (w[e] <= 0 && wmark[e] == stage) {  // L_e inside L_p: absorb into p
          st[e] = DEAD;
          std::vector<int>().swap(Le[e]);
          continue;
        }
        Ei[q++] = e;
        esum += (wmark[e] == stage) ? w[e] : edeg[e];
      }
      Ei.resize(q);
      Ei.push_back(p);
      auto& Ai = A[i];
      q = 0;
      int asum = 0;
      for (int v : Ai) {
        if (st[v] != VAR || mark[v] == stage) continue;  // p, elements, or in L_p
        Ai[q++] = v;
        asum += nv[v];
      }
      Ai.resize(q);
      // the degree is bounded once supervariables are merged (their weight
      // is then internal, not external)
      dpart[i] = asum + esum;
      dold[i] = deg[i];
    }
    // ---- indistinguishable variables of L_p: same A_i and E_i
    hv.clear();
    for (int i : Lp) {
      auto& Ai = A[i];
      auto& Ei = E[i];
      std::sort(Ai.begin(), Ai.end());
      std::sort(Ei.begin(), Ei.end());
      unsigned long long h = 1469598103934665603ull;
      for (int v : Ai) h = (h ^ (unsigned)v) * 1099511628211ull;
      h = (h ^ 0x9e3779b9u) * 1099511628211ull;
      for (int e : Ei) h = (h ^ (unsigned)e) * 1099511628211ull;
      hv.push_back({h, i});
    }
    std::sort(hv.begin(), hv.end());
    for (size_t a = 0; a < hv.size(); ++a) {
      const int i = hv[a].second;
      if (st[i] != VAR) continue;
      for (size_t b = a + 1; b < hv.size() && hv[b].first == hv[a].first; ++b) {
        const int j = hv[b].second;
        if (st[j] != VAR || A[i] != A[j] || E[i] != E[j]) continue;
        // merge j into i (j's own entries in other lists are skipped as MERGED)
        nv[i] += nv[j];
        members[i].push_back(j);
        for (int v : members[j]) members[i].push_back(v);
        std::vector<int>().swap(members[j]);
        std::vector<int>().swap(A[j]);
        std::vector<int>().swap(E[j]);
        st[j] = MERGED;
        nv[j] = 0;
      }
    }
    for (int i : Lp)
      if (st[i] == VAR) {
        const long long ext = (long long)degme - nv[i];
        long long d = dpart[i] + ext;
        d = std::min<long long>(d, (long long)dold[i] + ext);
        d = std::min<long long>(d, (long long)nleft - nv[i]);
        deg[i] = (int)std::max<long long>(d, 0);
        bucket_in(i);
        inbucket[i] = 1;
        mindeg = std::min(mindeg, deg[i]);
      }
    // L_p keeps the principal variables only
    auto& L = Le[p];
    for (int i : Lp)
      if (st[i] == VAR) L.push_back(i);
  }
  return 0;
}

}  // namespace ssids_hip
