// Host-side data shared by the analyse, ordering and factor drivers.
#pragma once
#include <cstdint>
#include <vector>

namespace ssids_hip {

// Symbolic output of the analyse phase (reference akeep fields
// src/ssids/akeep.f90:25-80).  Values are 1-based as in the reference; the
// arrays themselves are 0-based C storage.
struct SymbolicData {
  int n = 0;
  int nnodes = 0;
  std::vector<int> sptr;        // nnodes+1
  std::vector<int> sparent;     // nnodes (root's parent = nnodes+1)
  std::vector<int64_t> rptr;    // nnodes+1
  std::vector<int> rlist;       // rptr[nnodes]-1
  std::vector<int64_t> nptr;    // nnodes+1
  std::vector<int64_t> nlist;   // 2 x (nptr[nnodes]-1), column-major pairs
  int64_t num_factor = 0;
  int64_t num_flops = 0;
};

int basic_analyse(int n, const int64_t* ptr_full, const int* row_full, int* order,
                  int nemin, SymbolicData& out);
void build_map(int n, const int64_t* ptr, const int* row, const int* order,
               SymbolicData& sd);
void expand_pattern(int n, const int64_t* ptr, const int* row,
                    std::vector<int64_t>& aptr, std::vector<int>& arow);

// Levels of the assembly tree for the numeric factorization (parent[i] > i
// or -1; ncols = fully-summed columns per front): level[child] < level[parent].
// mode: 0 the cheapest by the cost model, 1 height, 2 greedy, 3 layered.
std::vector<int> tree_levels(const std::vector<int>& parent, const std::vector<int>& ncols,
                             int mode = 0);

// Orderings (1-based order(i) = pivot position of variable i).
void order_grid_nd(int nx, int ny, int nz, int* order);
int order_min_degree(int n, const int64_t* ptr, const int* row, int* order);
int order_amd(int n, const int64_t* ptr, const int* row, int* order);

}  // namespace ssids_hip
