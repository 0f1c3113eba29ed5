// Fill-reducing orderings used in place of METIS (absent from this image).
//
// The reference reaches METIS_NodeND through src/metis5_wrapper.F90:102-182 or
// takes a user order (options%ordering=0, checked by src/ssids/anal.F90:152-202).
// Both orderings here are deterministic so the same order can be fed to the
// numeric path and to any CPU baseline:
//  * order_grid_nd: geometric nested dissection of an nx*ny*nz grid whose
//    variable (i,j,k) is i + nx*(j + ny*k) (0-based), separators are middle
//    planes of the longest box dimension, ordered after both halves;
//  * order_min_degree: exact minimum degree on the explicit elimination graph
//    (ties to the lowest index), for moderate n.
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

void nd_box(const Box& b, int nx, int ny, std::vector<int>& seq) {
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
  nd_box(left, nx, ny, seq);
  nd_box(right, nx, ny, seq);
  for (int k = sep.lo[2]; k < sep.hi[2]; ++k)
    for (int j = sep.lo[1]; j < sep.hi[1]; ++j)
      for (int i = sep.lo[0]; i < sep.hi[0]; ++i) seq.push_back(i + nx * (j + ny * k));
}

}  // namespace

void order_grid_nd(int nx, int ny, int nz, int* order) {
  std::vector<int> seq;
  seq.reserve((size_t)nx * ny * nz);
  Box b{{0, 0, 0}, {nx, ny, nz}};
  nd_box(b, nx, ny, seq);
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

}  // namespace ssids_hip
