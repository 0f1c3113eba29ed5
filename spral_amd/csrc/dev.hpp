// Device-side data of the numeric factorization (shared by factor and solve).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ssids_hip {

constexpr int kBlk = 64;        // APP pivot block width (columns per step)
constexpr int kThreads = 256;   // threads per workgroup in most kernels
// least workgroups of a large front's APP trailing update (two per CU)
constexpr int kUpdMinWgs = 512;

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
  double *LD;       // ldl x n: rows [n, m) of L21 D (the parent's gather operand);
                    // allocated with C, absent for roots
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
  double *W;        // kBlk x kBlk row-major: inverse of the step's unit lower
                    // diagonal block (elimination order), for k_apply
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
  int ndcap;        // delayed columns the node's buffers can take (static plan)
  int skip;         // 1: delays exceeded ndcap here or below; node not factored
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
  // The solve workspace of a call with nrhs right-hand sides is one array ws
  // holding, per front, X (m x nrhs, ld m) and then the backward partial
  // sums P (bwd row workgroups x nrhs x nelim): X = ws + nrhs * woff and
  // P = X + nrhs * m, so the descriptors do not depend on nrhs.
  int64_t woff;
  const int *ctab;   // per child, per 64-row tile of this front: first child
                     // trailing index whose position reaches the tile (stride T+1)
  int m, n, nd, sn, nelim, ldl, doff, cbeg, cend;
};

__host__ __device__ inline double *solve_X(double *ws, const SolveNode &Q, int nrhs) {
  return ws + (size_t)nrhs * Q.woff;
}
__host__ __device__ inline double *solve_P(double *ws, const SolveNode &Q, int nrhs) {
  return ws + (size_t)nrhs * Q.woff + (size_t)nrhs * Q.m;
}

// Global-address-space views of device pointers read from descriptors: loads
// and stores through them compile to global (not flat) memory instructions.
typedef const double __attribute__((address_space(1))) gcdouble;
typedef double __attribute__((address_space(1))) gdouble;
typedef const int __attribute__((address_space(1))) gcint;

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ inline int idiv_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Per-level launch tables.  A level whose fronts need at most S APP steps has
// 2S+5 tables, each an exclusive prefix over the level's fronts of the
// workgroups one launch gives each front (last entry = total):
//   0          gather: lower 64x64 tiles of the front
//   1..S       apply, step t-1: t-1 chunks of earlier rows + 64-row blocks below
//   S+1        failed-column swap
//   S+2..2S+1  trailing panel update, step t-S-2
//   2S+2       contribution update of a root's TPP pivots
//   2S+3       gather of the diagonal tile (0, 0) fused with step 0's pivot
//              block: one workgroup per front
//   2S+4       gather, the tiles outside the first tile column
//   2S+5       gather, the first tile column's off-diagonal tiles
__host__ __device__ inline int level_tables(int S) { return 2 * S + 6; }
__host__ __device__ inline int level_table_count(int t, int S, int m, int n, int nsteps,
                                                 int is_root) {
  if (t == 0) {
    const int T = idiv_up(m, kBlk);
    return T * (T + 1) / 2;
  }
  if (t <= S) {
    const int stp = t - 1;
    return stp < nsteps ? stp + imax(1, idiv_up(m, kBlk)) : 0;
  }
  // failed-column swap: up to 32 elements a thread (most steps swap nothing,
  // and a small grid dispatches faster)
  if (t == S + 1) return nsteps > 1 ? imax(1, idiv_up(2LL * kBlk * m, kThreads * 32)) : 0;
  if (t <= 2 * S + 1) {
    const int stp = t - S - 2;
    if (stp >= nsteps) return 0;
    const int t0 = imin(n, (stp + 1) * kBlk);
    const int TI = idiv_up(m - t0, kBlk), TJ = idiv_up(n - t0, kBlk);
    const int nt = TJ * TI - TJ * (TJ - 1) / 2;
    // the count above assumes every earlier pivot passed; failed columns
    // leave a larger trailing panel (down to the whole front), which the
    // workgroups stride over, so a large front keeps at least kUpdMinWgs
    const int TM = idiv_up(m, kBlk), TN = idiv_up(n, kBlk);
    const int nt_all = TN * TM - TN * (TN - 1) / 2;
    return imax(1, imax(imin(nt, 4096), imin(nt_all, kUpdMinWgs)));
  }
  if (t == 2 * S + 2) {
    if (!is_root || m <= n) return 0;
    const int T = idiv_up(m - n, kBlk);
    return imax(1, imin(T * (T + 1) / 2, 4096));
  }
  // gather split in three launches: the diagonal tile (with step 0's pivot
  // block), the rest of the first tile column (what step 0's apply needs),
  // then the remaining tiles
  const int T = idiv_up(m, kBlk);
  if (t == 2 * S + 3) return T > 0 ? 1 : 0;
  if (t == 2 * S + 4) return T * (T + 1) / 2 - T;
  return imax(T - 1, 0);
}

// Index li of the table segment holding workgroup b: off[li] <= b < off[li+1]
// for a non-decreasing prefix table off[0..n]; -1 when b >= off[n].  Each
// wave searches with 64 samples per round (ballot of off[sample] <= b), so a
// table of up to 64^k entries costs k dependent loads.  Every lane of the
// wave must be active; the result is wave-uniform.
__device__ inline int table_find(const int *off, int n, int b) {
  const int lane = threadIdx.x & 63;
  if (b >= off[n]) return -1;
  int lo = 0, hi = n;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int step = (hi - lo + 63) >> 6;
    const int idx = lo + lane * step;
    const bool le = idx < hi && off[idx] <= b;
    const unsigned long long mask = __ballot(le);  // a prefix of the lanes; lane 0 set
    lo += (63 - __clzll((long long)mask)) * step;
    hi = (lo + step < hi) ? lo + step : hi;
  }
  return lo;
}

// v if keep, else +0.0.  The value passes through an empty asm first, so the
// selection is not turned into a branch around the load that produced it
// (which would serialize a batch of loads behind their uses).
__device__ inline double masked(double v, bool keep) {
  asm("" : "+v"(v));
  return keep ? v : 0.0;
}

// value held by lane 0 (wave-uniform by construction), as a scalar
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Wave-wide maxima with DPP lane permutations inside each row of 16 lanes
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror),
// then the four row results read as scalars.  Every lane of the wave must be
// active; the result is wave-uniform.
__device__ inline unsigned dpp_max_u32(unsigned v, int ctrl) {
  unsigned t;
  switch (ctrl) {
    case 0: t = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false); break;
    case 1: t = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false); break;
    case 2: t = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false); break;
    default: t = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false); break;
  }
  return t > v ? t : v;
}
__device__ inline unsigned wave_max_u32(unsigned v) {
  v = dpp_max_u32(v, 0);
  v = dpp_max_u32(v, 1);
  v = dpp_max_u32(v, 2);
  v = dpp_max_u32(v, 3);
  const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
  const unsigned r1 = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
  const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
  const unsigned r3 = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
  const unsigned a = r0 > r1 ? r0 : r1, b = r2 > r3 ? r2 : r3;
  return a > b ? a : b;
}
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
  const unsigned hi = wave_max_u32((unsigned)(v >> 32));
  const unsigned lo = wave_max_u32(((unsigned)(v >> 32) == hi) ? (unsigned)v : 0u);
  return ((unsigned long long)hi << 32) | lo;
}

struct FactorParams {
  double u, small;
  int action;
  int tpp_failed;   // failed_pivot_method == TPP
  int pivot_tpp;    // pivot_method == TPP (whole panel with TPP)
  int posdef;       // A = L L^T: pivots in order, every one 1x1 and > 0
  int c32;          // contribution blocks held in FP32, the children's Schur
                    // updates on FP32 MFMA (mixed precision; the panel, the
                    // pivot blocks and their tests stay FP64)
};

}  // namespace ssids_hip
