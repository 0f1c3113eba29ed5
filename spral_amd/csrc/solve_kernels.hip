// Forward / diagonal / backward solve kernels over the resident factors.
//
// Forward (levels leaves -> root), per front s with nelim = ne:
//   k_fwd_tri : one workgroup per front: X = own b + children's updates
//               (children in order: deterministic), then the unit-lower
//               ne x ne triangle (64-column blocks staged in LDS); the
//               eliminated values y go to x by variable.
//   k_fwd_gemv: many workgroups per front: X[ne:m] -= L21 y (rows split
//               across workgroups, up to 8 right-hand sides per pass so L21
//               is read once per 8 RHS).  X[ne:m] is the update the parent
//               gathers.
// Backward (root -> leaves):
//   k_bwd_gemv: many workgroups per front: partial sums of L21^T x[rows].
//   k_bwd_tri : one workgroup per front: reduce the partials, L11^T solve.
// Fronts with more than kBigTri eliminated columns (the separators at the
// top of the tree) solve their triangle with a launch sequence instead of one
// workgroup: per 64-column block a one-workgroup diagonal solve and a
// many-workgroup update of (forward) or partial sums over (backward) the
// triangle's rows below the block.
// Reference semantics: src/ssids/cpu/NumericSubtree.hxx:286-406 and
// src/ssids/cpu/kernels/ldlt_app.cxx:2537-2589.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dev.hpp"

namespace ssids_hip {

constexpr int kRhsChunk = 8;
constexpr int kBigTri = 256;

__device__ inline int find_node_s(const int *wg_off, int nlvl, int b) {
  return table_find(wg_off, nlvl, b);
}

// post-pivot parent position of child trailing index k
__device__ inline int post_pos(const SolveNode &P, const SolveNode &c, int k) {
  int pre = (k < c.n) ? c.doff + (k - c.nelim) : P.nd + c.rmap[k - c.n];
  return (pre < P.n) ? P.ipos[pre] : pre;
}

__global__ void k_ipos(NodeDev *nodes, const int *slots, int *scratch) {
  // one workgroup per front of the level; fully-summed variable sets of one
  // level are disjoint, so the variable-indexed scratch is conflict free.
  const NodeDev &N = nodes[slots[blockIdx.x]];
  for (int i = threadIdx.x; i < N.n; i += blockDim.x) scratch[N.perm[i]] = i;
  __syncthreads();
  for (int j = threadIdx.x; j < N.n; j += blockDim.x) N.ipos[j] = scratch[N.perm0[j]];
}

// unit-lower triangle solve of X[0:ne] (one workgroup), 64-column blocks:
// the diagonal block is staged in LDS and solved by one wave, the rows below
// inside the triangle are updated by the whole workgroup.
__device__ void tri_lower(const SolveNode &P, double *X, double (*lb)[kBlk + 1], double *yb) {
  const int tid = threadIdx.x, ne = P.nelim;
  for (int j0 = 0; j0 < ne; j0 += kBlk) {
    const int j1 = min(j0 + kBlk, ne), wb = j1 - j0;
    for (int idx = tid; idx < wb * wb; idx += kThreads) {
      int r = idx % wb, c = idx / wb;
      lb[r][c] = (r > c) ? P.L[(size_t)(j0 + c) * P.ldl + j0 + r] : 0.0;
    }
    __syncthreads();
    if (tid < 64) {
      double v = (tid < wb) ? X[j0 + tid] : 0.0;
      for (int j = 0; j < wb; ++j) {
        double yj = __shfl(v, j);
        if (tid > j && tid < wb) v -= lb[tid][j] * yj;
      }
      if (tid < wb) {
        X[j0 + tid] = v;
        yb[tid] = v;
      }
    }
    __syncthreads();
    for (int i = j1 + tid; i < ne; i += kThreads) {
      double acc = X[i];
      const double *lc = P.L + (size_t)j0 * P.ldl + i;
      for (int j = 0; j < wb; ++j) acc -= lc[(size_t)j * P.ldl] * yb[j];
      X[i] = acc;
    }
    __syncthreads();
  }
}

// X = own b + the children's updates.  Workgroup q of a front owns the
// front positions (before pivoting) [256q, 256q+256): it sets them from b,
// then adds each child's update in child order (the child's trailing indices
// that land in the range come from the assembly's tile table), so every
// entry is summed in a fixed order by one workgroup.
__global__ __launch_bounds__(kThreads) void k_fwd_gather(const SolveNode *sn, const int *slots,
                                                         const int *child_list, double *x,
                                                         int ldx, int nrhs, const int *wg_off,
                                                         int nlvl) {
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  const SolveNode &P = sn[slots[li]];
  const int tid = threadIdx.x, m = P.m, n = P.n, nd = P.nd;
  const int q0 = (blockIdx.x - wg_off[li]) * kThreads, q1 = min(q0 + kThreads, m);
  const int T = (m + 63) / 64;
  const int I0 = q0 / 64, I1 = min(T, (q1 + 63) / 64);
  const int pre = q0 + tid;
  const int post = (pre < n) ? ((pre < q1) ? P.ipos[pre] : 0) : pre;
  const int vrow = (pre >= nd && pre < n) ? P.rows[pre - nd] : -1;
  for (int r = 0; r < nrhs; ++r) {
    double *X = P.X + (size_t)r * m;
    const double *xr = x + (size_t)r * ldx;
    if (pre < q1) X[post] = (vrow >= 0) ? xr[vrow] : 0.0;
    __syncthreads();
    for (int ci = P.cbeg; ci < P.cend; ++ci) {
      const SolveNode &c = sn[child_list[ci]];
      const int *tab = P.ctab + (size_t)(ci - P.cbeg) * (T + 1);
      const int kb = tab[I0], ke = tab[I1];
      const int k = kb + tid;  // at most 256 child indices land in the range
      if (k < ke) X[post_pos(P, c, k)] += c.X[(size_t)r * c.m + k];
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_fwd_tri(const SolveNode *sn, const int *slots,
                                                      double *x, int ldx, int nrhs) {
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double yb[kBlk];
  const SolveNode &P = sn[slots[blockIdx.x]];
  const int tid = threadIdx.x, m = P.m, ne = P.nelim;
  if (ne > kBigTri) return;  // triangle solved by the big-front launches
  for (int r = 0; r < nrhs; ++r) {
    double *X = P.X + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    tri_lower(P, X, lb, yb);
    for (int i = tid; i < ne; i += kThreads) xr[P.perm[i]] = X[i];
    __syncthreads();
  }
}

// X[ne:m, r] -= L[ne:m, 0:ne] X[0:ne, r] for rows of this workgroup
__global__ __launch_bounds__(kThreads) void k_fwd_gemv(const SolveNode *sn, const int *slots,
                                                       const int *wg_off, int nlvl, int nrhs) {
  __shared__ double ys[kRhsChunk][kBlk];
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  const SolveNode &P = sn[slots[li]];
  const int local = blockIdx.x - wg_off[li], count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, m = P.m, ne = P.nelim;
  for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
    const int nr = min(kRhsChunk, nrhs - r0);
    for (int i0 = ne + local * kThreads; i0 < m; i0 += count * kThreads) {
      const int i = i0 + tid;
      double acc[kRhsChunk];
#pragma unroll
      for (int q = 0; q < kRhsChunk; ++q) acc[q] = 0.0;
      for (int j0 = 0; j0 < ne; j0 += kBlk) {
        const int wb = min(kBlk, ne - j0);
        __syncthreads();
        for (int idx = tid; idx < kRhsChunk * kBlk; idx += kThreads) {
          int q = idx / kBlk, j = idx % kBlk;
          ys[q][j] = (q < nr && j < wb) ? P.X[(size_t)(r0 + q) * m + j0 + j] : 0.0;
        }
        __syncthreads();
        if (i < m) {
          const double *lc = P.L + (size_t)j0 * P.ldl + i;
          for (int j = 0; j < wb; ++j) {
            double l = lc[(size_t)j * P.ldl];
#pragma unroll
            for (int q = 0; q < kRhsChunk; ++q) acc[q] += l * ys[q][j];
          }
        }
      }
      if (i < m)
        for (int q = 0; q < nr; ++q) P.X[(size_t)(r0 + q) * m + i] -= acc[q];
    }
  }
}

__global__ void k_dsolve(const SolveNode *sn, int nslots, double *x, int ldx, int nrhs) {
  const int s = blockIdx.x;
  if (s >= nslots) return;
  const SolveNode &P = sn[s];
  const int ne = P.nelim;
  const double *d = P.D;
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    bool first2 = (i + 1 < ne) && !isfinite(d[2 * i + 2]);
    bool second2 = !isfinite(d[2 * i]);
    for (int r = 0; r < nrhs; ++r) {
      double *xr = x + (size_t)r * ldx;
      if (first2) {
        double x1 = xr[P.perm[i]], x2 = xr[P.perm[i + 1]];
        xr[P.perm[i]] = d[2 * i] * x1 + d[2 * i + 1] * x2;
        xr[P.perm[i + 1]] = d[2 * i + 1] * x1 + d[2 * i + 3] * x2;
      } else if (!second2) {
        xr[P.perm[i]] *= d[2 * i];
      }
    }
  }
}

// partial[wg][r][j] = sum over this workgroup's rows i >= ne of L[i,j] x[var(i), r]
__global__ __launch_bounds__(kThreads) void k_bwd_gemv(const SolveNode *sn, const int *slots,
                                                       const int *wg_off, int nlvl,
                                                       const double *x, int ldx, int nrhs) {
  __shared__ double red[kThreads / 64][kBlk];
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  const SolveNode &P = sn[slots[li]];
  const int local = blockIdx.x - wg_off[li], count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = P.m, n = P.n, ne = P.nelim;
  double *part = P.P;  // count x nrhs x ne partial sums
  for (int r = 0; r < nrhs; ++r) {
    const double *xr = x + (size_t)r * ldx;
    for (int j0 = 0; j0 < ne; j0 += kBlk) {
      const int wb = min(kBlk, ne - j0);
      double acc[kBlk];
#pragma unroll
      for (int t = 0; t < kBlk; ++t) acc[t] = 0.0;
      for (int i = ne + local * kThreads + tid; i < m; i += count * kThreads) {
        double xi = (i < n) ? xr[P.perm[i]] : xr[P.rows[i - P.nd]];
        const double *lr = P.L + (size_t)j0 * P.ldl + i;
#pragma unroll
        for (int t = 0; t < kBlk; ++t)
          if (t < wb) acc[t] += lr[(size_t)t * P.ldl] * xi;
      }
#pragma unroll
      for (int t = 0; t < kBlk; ++t) {
        double v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave][t] = v;
      }
      __syncthreads();
      if (tid < wb) {
        double v = 0.0;
        for (int q = 0; q < kThreads / 64; ++q) v += red[q][tid];
        part[((size_t)local * nrhs + r) * ne + j0 + tid] = v;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_bwd_tri(const SolveNode *sn, const int *slots,
                                                      const int *wg_off, double *x, int ldx,
                                                      int nrhs) {
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double part[kThreads / 64][kBlk];
  __shared__ double xb[kBlk];
  const int li = blockIdx.x;
  const SolveNode &P = sn[slots[li]];
  const int count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = P.m, ne = P.nelim;
  const double *partial = P.P;
  for (int r = 0; r < nrhs; ++r) {
    double *X = P.X + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    // X[j] = x[perm[j]] - sum of the gemv partials (rows >= ne)
    for (int j = tid; j < ne; j += kThreads) {
      double v = xr[P.perm[j]];
#pragma unroll 8
      for (int g = 0; g < count; ++g) v -= partial[((size_t)g * nrhs + r) * ne + j];
      X[j] = v;
    }
    __syncthreads();
    if (ne > kBigTri) continue;  // triangle solved by the big-front launches
    // L11^T solve, last 64-column block first
    const int nb = (ne + kBlk - 1) / kBlk;
    for (int b = nb - 1; b >= 0; --b) {
      const int j0 = b * kBlk, j1 = min(j0 + kBlk, ne), wb = j1 - j0;
      // X[j] -= L[j1:ne, j]^T X[j1:ne] for j in the block
      double acc[kBlk];
#pragma unroll
      for (int t = 0; t < kBlk; ++t) acc[t] = 0.0;
      for (int i = j1 + tid; i < ne; i += kThreads) {
        double xi = X[i];
        const double *lr = P.L + (size_t)j0 * P.ldl + i;
#pragma unroll
        for (int t = 0; t < kBlk; ++t)
          if (t < wb) acc[t] += lr[(size_t)t * P.ldl] * xi;
      }
#pragma unroll
      for (int t = 0; t < kBlk; ++t) {
        double v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) part[wave][t] = v;
      }
      for (int idx = tid; idx < wb * wb; idx += kThreads) {
        int rr = idx % wb, cc = idx / wb;
        lb[rr][cc] = (rr > cc) ? P.L[(size_t)(j0 + cc) * P.ldl + j0 + rr] : 0.0;
      }
      __syncthreads();
      if (tid < wb) {
        double v = X[j0 + tid];
        for (int q = 0; q < kThreads / 64; ++q) v -= part[q][tid];
        xb[tid] = v;
      }
      __syncthreads();
      if (tid < 64) {
        double v = (tid < wb) ? xb[tid] : 0.0;
        for (int j = wb - 1; j >= 0; --j) {
          double yj = __shfl(v, j);
          if (tid < j) v -= lb[j][tid] * yj;
        }
        if (tid < wb) X[j0 + tid] = v;
      }
      __syncthreads();
    }
    for (int i = tid; i < ne; i += kThreads) xr[P.perm[i]] = X[i];
    __syncthreads();
  }
}

// ---- big fronts: forward -------------------------------------------------
// X[j0:j1] <- L_bb^{-1} X[j0:j1] for every right-hand side (one workgroup,
// one wave per right-hand side at a time)
__global__ __launch_bounds__(kThreads) void k_tri_fdiag(const SolveNode *sn, int s, int j0,
                                                        int nrhs) {
  __shared__ double lb[kBlk][kBlk + 1];
  const SolveNode &P = sn[s];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ne = P.nelim, m = P.m, j1 = min(j0 + kBlk, ne), wb = j1 - j0;
  for (int idx = tid; idx < wb * wb; idx += kThreads) {
    const int r = idx % wb, c = idx / wb;
    lb[r][c] = (r > c) ? P.L[(size_t)(j0 + c) * P.ldl + j0 + r] : 0.0;
  }
  __syncthreads();
  for (int r = wave; r < nrhs; r += kThreads / 64) {
    double *X = P.X + (size_t)r * m;
    double v = (lane < wb) ? X[j0 + lane] : 0.0;
    for (int j = 0; j < wb; ++j) {
      const double yj = __shfl(v, j);
      if (lane > j && lane < wb) v -= lb[lane][j] * yj;
    }
    if (lane < wb) X[j0 + lane] = v;
  }
}

// X[i] -= L[i, j0:j1] X[j0:j1] for the triangle's rows i in [j1, ne)
__global__ __launch_bounds__(kThreads) void k_tri_frows(const SolveNode *sn, int s, int j0,
                                                        int nrhs) {
  __shared__ double ys[kRhsChunk][kBlk];
  const SolveNode &P = sn[s];
  const int tid = threadIdx.x, ne = P.nelim, m = P.m;
  const int j1 = min(j0 + kBlk, ne), wb = j1 - j0;
  const int i = j1 + blockIdx.x * kThreads + tid;
  for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk) {
    const int nr = min(kRhsChunk, nrhs - r0);
    __syncthreads();
    for (int idx = tid; idx < kRhsChunk * kBlk; idx += kThreads) {
      const int q = idx / kBlk, j = idx % kBlk;
      ys[q][j] = (q < nr && j < wb) ? P.X[(size_t)(r0 + q) * m + j0 + j] : 0.0;
    }
    __syncthreads();
    if (i < ne) {
      double acc[kRhsChunk];
#pragma unroll
      for (int q = 0; q < kRhsChunk; ++q) acc[q] = 0.0;
      const double *lc = P.L + (size_t)j0 * P.ldl + i;
      for (int j = 0; j < wb; ++j) {
        const double l = lc[(size_t)j * P.ldl];
#pragma unroll
        for (int q = 0; q < kRhsChunk; ++q) acc[q] += l * ys[q][j];
      }
      for (int q = 0; q < nr; ++q) P.X[(size_t)(r0 + q) * m + i] -= acc[q];
    }
  }
}

// ---- big fronts: backward ------------------------------------------------
// part[g][r][t] = sum over workgroup g's rows i in [j1, ne) of L[i, j0+t] X[i, r]
__global__ __launch_bounds__(kThreads) void k_tri_brows(const SolveNode *sn, int s, int j0,
                                                        int nrhs, double *part) {
  __shared__ double red[kThreads / 64][kBlk];
  const SolveNode &P = sn[s];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ne = P.nelim, m = P.m, j1 = min(j0 + kBlk, ne), wb = j1 - j0;
  const int i = j1 + blockIdx.x * kThreads + tid;
  for (int r = 0; r < nrhs; ++r) {
    double acc[kBlk];
#pragma unroll
    for (int t = 0; t < kBlk; ++t) acc[t] = 0.0;
    if (i < ne) {
      const double xi = P.X[(size_t)r * m + i];
      const double *lr = P.L + (size_t)j0 * P.ldl + i;
#pragma unroll
      for (int t = 0; t < kBlk; ++t)
        if (t < wb) acc[t] = lr[(size_t)t * P.ldl] * xi;
    }
#pragma unroll
    for (int t = 0; t < kBlk; ++t) {
      double v = acc[t];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (lane == 0) red[wave][t] = v;
    }
    __syncthreads();
    if (tid < kBlk) {
      double v = 0.0;
      for (int q = 0; q < kThreads / 64; ++q) v += red[q][tid];
      part[((size_t)blockIdx.x * nrhs + r) * kBlk + tid] = v;
    }
    __syncthreads();
  }
}

// X[j0:j1] -= sum of the partials, then X[j0:j1] <- L_bb^{-T} X[j0:j1]
__global__ __launch_bounds__(kThreads) void k_tri_bdiag(const SolveNode *sn, int s, int j0,
                                                        int ngroups, int nrhs,
                                                        const double *part) {
  __shared__ double lb[kBlk][kBlk + 1];
  const SolveNode &P = sn[s];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ne = P.nelim, m = P.m, j1 = min(j0 + kBlk, ne), wb = j1 - j0;
  for (int idx = tid; idx < wb * wb; idx += kThreads) {
    const int r = idx % wb, c = idx / wb;
    lb[r][c] = (r > c) ? P.L[(size_t)(j0 + c) * P.ldl + j0 + r] : 0.0;
  }
  __syncthreads();
  for (int r = wave; r < nrhs; r += kThreads / 64) {
    double *X = P.X + (size_t)r * m;
    double v = 0.0;
    if (lane < wb) {
      v = X[j0 + lane];
      for (int g = 0; g < ngroups; ++g) v -= part[((size_t)g * nrhs + r) * kBlk + lane];
    }
    for (int j = wb - 1; j >= 0; --j) {
      const double yj = __shfl(v, j);
      if (lane < j) v -= lb[j][lane] * yj;
    }
    if (lane < wb) X[j0 + lane] = v;
  }
}

// x[perm[i]] = X[i] for the front's eliminated variables
__global__ void k_tri_scatter(const SolveNode *sn, int s, double *x, int ldx, int nrhs) {
  const SolveNode &P = sn[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nelim) return;
  for (int r = 0; r < nrhs; ++r) x[(size_t)r * ldx + P.perm[i]] = P.X[(size_t)r * P.m + i];
}

void launch_big_fwd(hipStream_t st, const SolveNode *sn, int s, int ne, double *x, int ldx,
                    int nrhs) {
  for (int j0 = 0; j0 < ne; j0 += kBlk) {
    hipLaunchKernelGGL(k_tri_fdiag, dim3(1), dim3(kThreads), 0, st, sn, s, j0, nrhs);
    const int rows = ne - std::min(j0 + kBlk, ne);
    if (rows > 0)
      hipLaunchKernelGGL(k_tri_frows, dim3((rows + kThreads - 1) / kThreads), dim3(kThreads), 0,
                         st, sn, s, j0, nrhs);
  }
  hipLaunchKernelGGL(k_tri_scatter, dim3((ne + 255) / 256), dim3(256), 0, st, sn, s, x, ldx,
                     nrhs);
}
// part: room for (ne / kThreads + 1) * nrhs * kBlk doubles
void launch_big_bwd(hipStream_t st, const SolveNode *sn, int s, int ne, double *x, int ldx,
                    int nrhs, double *part) {
  const int nb = (ne + kBlk - 1) / kBlk;
  for (int b = nb - 1; b >= 0; --b) {
    const int j0 = b * kBlk, rows = ne - std::min(j0 + kBlk, ne);
    const int ng = (rows + kThreads - 1) / kThreads;
    if (ng > 0)
      hipLaunchKernelGGL(k_tri_brows, dim3(ng), dim3(kThreads), 0, st, sn, s, j0, nrhs, part);
    hipLaunchKernelGGL(k_tri_bdiag, dim3(1), dim3(kThreads), 0, st, sn, s, j0, ng, nrhs,
                       (const double *)part);
  }
  hipLaunchKernelGGL(k_tri_scatter, dim3((ne + 255) / 256), dim3(256), 0, st, sn, s, x, ldx,
                     nrhs);
}
int big_tri_threshold() { return kBigTri; }

void launch_ipos(hipStream_t st, int nlvl, NodeDev *nodes, const int *slots, int *scratch) {
  if (nlvl > 0)
    hipLaunchKernelGGL(k_ipos, dim3(nlvl), dim3(256), 0, st, nodes, slots, scratch);
}
void launch_fwd_gather(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                       const int *child_list, double *x, int ldx, int nrhs, const int *wg_off,
                       int nwg) {
  if (nlvl > 0 && nwg > 0)
    hipLaunchKernelGGL(k_fwd_gather, dim3(nwg), dim3(kThreads), 0, st, sn, slots, child_list,
                       x, ldx, nrhs, wg_off, nlvl);
}
void launch_fwd_tri(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                    double *x, int ldx, int nrhs) {
  if (nlvl <= 0) return;
  hipLaunchKernelGGL(k_fwd_tri, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, x, ldx, nrhs);
}
void launch_fwd_gemv(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                     int nrhs, const int *gemv_off, int gemv_wg) {
  if (nlvl > 0 && gemv_wg > 0)
    hipLaunchKernelGGL(k_fwd_gemv, dim3(gemv_wg), dim3(kThreads), 0, st, sn, slots, gemv_off,
                       nlvl, nrhs);
}
void launch_dsolve(hipStream_t st, int nslots, const SolveNode *sn, double *x, int ldx,
                   int nrhs) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_dsolve, dim3(nslots), dim3(64), 0, st, sn, nslots, x, ldx, nrhs);
}
void launch_bwd(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots, double *x,
                int ldx, int nrhs, const int *gemv_off, int gemv_wg) {
  if (nlvl <= 0) return;
  if (gemv_wg > 0)
    hipLaunchKernelGGL(k_bwd_gemv, dim3(gemv_wg), dim3(kThreads), 0, st, sn, slots, gemv_off,
                       nlvl, x, ldx, nrhs);
  hipLaunchKernelGGL(k_bwd_tri, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, gemv_off, x, ldx,
                     nrhs);
}

}  // namespace ssids_hip
