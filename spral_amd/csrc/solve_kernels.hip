// Forward / diagonal / backward solve kernels over the resident factors.
//
// The forward solve is a multifrontal extend-add of vectors: each front keeps
// its trailing part X_s[nelim:m] (the update passed to the parent), the
// parent gathers its children's updates in child order (deterministic), and
// eliminated values are written to x by variable.  The backward solve reads x
// by variable, so fronts of one level never write the same entry.
// Dense parts are swept in 64-column blocks so L is read once per RHS
// (reference semantics: src/ssids/cpu/NumericSubtree.hxx:286-406,
// src/ssids/cpu/kernels/ldlt_app.cxx:2537-2589).
#include <hip/hip_runtime.h>

#include <cmath>

#include "dev.hpp"

namespace ssids_hip {

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

__global__ __launch_bounds__(kThreads) void k_fwd(const SolveNode *sn, const int *slots,
                                                  const int *child_list, double *x,
                                                  int ldx, int nrhs) {
  __shared__ double yb[kBlk];
  const SolveNode &P = sn[slots[blockIdx.x]];
  const int tid = threadIdx.x, m = P.m, n = P.n, ne = P.nelim;
  for (int r = 0; r < nrhs; ++r) {
    double *X = P.X + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    for (int i = tid; i < m; i += kThreads) X[i] = 0.0;
    __syncthreads();
    for (int j = P.nd + tid; j < n; j += kThreads) X[P.ipos[j]] = xr[P.rows[j - P.nd]];
    __syncthreads();
    for (int ci = P.cbeg; ci < P.cend; ++ci) {
      const SolveNode &c = sn[child_list[ci]];
      const double *Xc = c.X + (size_t)r * c.m;
      for (int k = c.nelim + tid; k < c.m; k += kThreads) X[post_pos(P, c, k)] += Xc[k];
      __syncthreads();
    }
    // blocked forward substitution with unit lower L (columns [0, ne))
    for (int j0 = 0; j0 < ne; j0 += kBlk) {
      const int j1 = min(j0 + kBlk, ne);
      if (tid < 64) {
        // one wave solves the diagonal block
        double v = (j0 + tid < j1) ? X[j0 + tid] : 0.0;
        for (int j = j0; j < j1; ++j) {
          double yj = __shfl(v, j - j0);
          if (tid > j - j0 && j0 + tid < j1) v -= P.L[(size_t)j * P.ldl + j0 + tid] * yj;
        }
        if (j0 + tid < j1) {
          X[j0 + tid] = v;
          yb[tid] = v;
        }
      }
      __syncthreads();
      for (int i = j1 + tid; i < m; i += kThreads) {
        double acc = X[i];
        for (int j = j0; j < j1; ++j) acc -= P.L[(size_t)j * P.ldl + i] * yb[j - j0];
        X[i] = acc;
      }
      __syncthreads();
    }
    for (int i = tid; i < ne; i += kThreads) xr[P.perm[i]] = X[i];
    __syncthreads();
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

__global__ __launch_bounds__(kThreads) void k_bwd(const SolveNode *sn, const int *slots,
                                                  double *x, int ldx, int nrhs) {
  __shared__ double part[kThreads / 64][kBlk];
  __shared__ double xb[kBlk];
  const SolveNode &P = sn[slots[blockIdx.x]];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = P.m, n = P.n, ne = P.nelim;
  for (int r = 0; r < nrhs; ++r) {
    double *X = P.X + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    for (int i = tid; i < m; i += kThreads)
      X[i] = (i < n) ? xr[P.perm[i]] : xr[P.rows[i - P.nd]];
    __syncthreads();
    // backward sweep over 64-column blocks, last block first
    const int nb = (ne + kBlk - 1) / kBlk;
    for (int b = nb - 1; b >= 0; --b) {
      const int j0 = b * kBlk, j1 = min(j0 + kBlk, ne);
      const int w = j1 - j0;
      // X[j] -= L[j1:m, j]^T X[j1:m] for j in block
      double acc[kBlk];
#pragma unroll
      for (int t = 0; t < kBlk; ++t) acc[t] = 0.0;
      for (int i = j1 + tid; i < m; i += kThreads) {
        double xi = X[i];
#pragma unroll
        for (int t = 0; t < kBlk; ++t)
          if (t < w) acc[t] += P.L[(size_t)(j0 + t) * P.ldl + i] * xi;
      }
#pragma unroll
      for (int t = 0; t < kBlk; ++t) {
        double v = acc[t];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) part[wave][t] = v;
      }
      __syncthreads();
      if (tid < w) {
        double v = X[j0 + tid];
        for (int q = 0; q < kThreads / 64; ++q) v -= part[q][tid];
        xb[tid] = v;
      }
      __syncthreads();
      // unit upper solve L_bb^T within the block (one wave)
      if (tid < 64) {
        double v = (tid < w) ? xb[tid] : 0.0;
        for (int j = w - 1; j >= 0; --j) {
          double yj = __shfl(v, j);
          if (tid < j) v -= P.L[(size_t)(j0 + tid) * P.ldl + j0 + j] * yj;
        }
        if (tid < w) X[j0 + tid] = v;
      }
      __syncthreads();
    }
    for (int i = tid; i < ne; i += kThreads) xr[P.perm[i]] = X[i];
    __syncthreads();
  }
}

void launch_ipos(hipStream_t st, int nlvl, NodeDev *nodes, const int *slots, int *scratch) {
  if (nlvl > 0)
    hipLaunchKernelGGL(k_ipos, dim3(nlvl), dim3(256), 0, st, nodes, slots, scratch);
}
void launch_fwd(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                const int *child_list, double *x, int ldx, int nrhs) {
  if (nlvl > 0)
    hipLaunchKernelGGL(k_fwd, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, child_list, x,
                       ldx, nrhs);
}
void launch_dsolve(hipStream_t st, int nslots, const SolveNode *sn, double *x, int ldx,
                   int nrhs) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_dsolve, dim3(nslots), dim3(64), 0, st, sn, nslots, x, ldx, nrhs);
}
void launch_bwd(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots, double *x,
                int ldx, int nrhs) {
  if (nlvl > 0)
    hipLaunchKernelGGL(k_bwd, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, x, ldx, nrhs);
}

}  // namespace ssids_hip
