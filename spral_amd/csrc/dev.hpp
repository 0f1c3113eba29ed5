// Device-side data of the numeric factorization (shared by factor and solve).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ssids_hip {

constexpr int kBlk = 64;        // APP pivot block width (columns per step)
constexpr int kThreads = 256;   // threads per workgroup in most kernels

// Per-node descriptor, stored contiguously in tree-level order ("slots").
// Front layout (rows and columns): [delays in (nd) | own columns (sn) |
// remaining rows]; the first n = nd+sn columns are fully summed.
//   L : ldl x n column-major, rows [0,m) -- the fully-summed panel; after the
//       factorization columns [0,nelim) hold unit-lower L, [nelim,n) the
//       failed (delayed) columns updated by all eliminated pivots.
//   C : ldc x (m-n) column-major -- contribution block (rows/cols [n,m)).
// The child->parent map of a trailing index k in [nelim, m) is monotone:
//   k <  n : doff + (k - nelim)          (into the parent's delay slots)
//   k >= n : parent.nd + rmap[k - n]      (parent symbolic row position)
struct NodeDev {
  double *L;
  double *C;
  double *D;        // 2*n, D^{-1} as [d11, d21, (inf for 2x2 second), d22]
  double *Dd;       // 3*n: per pivot (D_kk, D_k,partner, partner offset 0/+1/-1)
  int *perm;        // n fully-summed variables (0-based elimination index)
  int *perm0;       // the same list before pivoting (as assembled)
  int *ipos;        // pre-pivot -> post-pivot position (filled after factor)
  const int *rows;  // sm symbolic rows (0-based elimination index)
  const int *rmap;  // sm-sn parent symbolic positions of contribution rows
  const int64_t *acol;  // sn+1 offsets into the A-entry arrays
  double *bak;      // (m x kBlk) backup of the panel rows below the block
  double *dbak;     // kBlk x kBlk backup of the diagonal block
  double *swp;      // 2*kBlk x m gather buffer for the failed-column swap
  int *pmap;        // (m - nelim) parent positions of the trailing indices
  int *ctab;        // per child (stride T+1, T = 64-row tiles of this front):
                    // tile t -> first child trailing index mapping to >= 64t
  int sm, sn, nd, m, n, ldl, ldc, ldbak;
  int parent;       // slot of parent or -1
  int cbeg, cend;   // children slots range in the child list
  int doff;         // offset of this node's delayed columns in parent
  int nsteps;       // APP steps this node needs (ceil(n/kBlk))
  int is_root;
  // factorization state (device-written)
  int p, q, w, npass;
  int nelim;
  int nelim_app;
  int flag;         // 0 ok, -5 singular (action false)
  int nzero;
  int lp[kBlk];
  unsigned long long colmax[kBlk];
};

// Per-node view used by the solve kernels.
struct SolveNode {
  const double *L;
  const double *D;
  const int *perm;   // post-pivot variable of fully-summed positions
  const int *ipos;   // pre-pivot -> post-pivot position (n entries)
  const int *rows;   // symbolic rows
  const int *rmap;
  double *X;         // m x nrhs workspace (ld m)
  double *P;         // backward GEMV partial sums (gemv workgroups x nrhs x nelim)
  int m, n, nd, sn, nelim, ldl, doff, cbeg, cend;
};

struct FactorParams {
  double u, small;
  int action;
  int tpp_failed;   // failed_pivot_method == TPP
  int pivot_tpp;    // pivot_method == TPP (whole panel with TPP)
};

}  // namespace ssids_hip
