// Forward / diagonal / backward solve kernels over the resident factors.
//
// Positions of a front (after pivoting): [0, ne) eliminated, [ne, n) failed
// fully-summed columns passed on to the parent, [n, m) the rows below.  X
// (m x nrhs per front, ld m) holds the forward values of the front's positions;
// X and the backward partial sums live in the call's workspace ws (solve_X,
// solve_P).  Every kernel writes what it reads of ws earlier in the same
// solve, so ws needs no clearing between calls.
//
// Forward (levels leaves -> root), two launches per level:
//   k_fwd_tri : one workgroup per front: X[0:n] = own b + the children's
//               updates that land on fully-summed positions, then the
//               unit-lower triangle on X[0:ne] (fronts with ne <= kBigTri);
//               the eliminated values go to x by variable.
//   k_fwd_rows: 64-row workgroups over [ne, m): the children's updates that
//               land on the rows (summed in LDS in a fixed order) minus
//               L[rows, 0:ne] X[0:ne].  X[ne:m] is what the parent gathers.
// Backward (root -> leaves), one launch per level:
//   k_bwd     : 64-row workgroups over [ne, m) write partial sums of
//               L[rows, 0:ne]^T x[rows]; the last workgroup of a front to
//               finish (device-scope counter) reduces them in workgroup order
//               and solves the triangle with L11^T.
// Fronts with ne > kBigTri (separators at the top of the tree) solve their
// triangle with one launch per 64-column block (k_big_fwd / k_big_bwd).
//
// Loads of L are issued in batches of 16 independent addresses (column indices
// clamped into the block instead of guarded, the surplus products multiplied
// by zero or discarded), and the 64-step diagonal-block substitutions keep the
// block in registers and broadcast each solved value with a lane read.
// Reference semantics: src/ssids/cpu/NumericSubtree.hxx:286-406 and
// src/ssids/cpu/kernels/ldlt_app.cxx:2537-2589.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dev.hpp"

namespace ssids_hip {

constexpr int kRows = 64;                  // rows per workgroup of the row kernels
constexpr int kGroups = kThreads / kRows;  // column groups (one wave each)
constexpr int kYChunk = 256;               // columns of X[0:ne] staged per pass
constexpr int kRhsChunk = 8;
constexpr int kBigTri = 256;
constexpr int kTriCap = 1024;  // fully-summed positions gathered in LDS

__device__ inline int find_node_s(const int *wg_off, int nlvl, int b) {
  return table_find(wg_off, nlvl, b);
}

// value of v in lane j (j uniform)
__device__ inline double bcast_lane(double v, int j) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, j);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes of a[t] for every t, spread over the lanes: the
// result in a lane is the sum for t = lane >> (6 - log2 T).  Halving
// exchanges (T - 1 shuffles) then butterflies over the remaining lane bits.
template <int T>
__device__ inline double wave_reduce_scatter(double (&a)[T], int lane) {
  constexpr int k = (T >= 64) ? 6 : (T >= 32) ? 5 : (T >= 16) ? 4 : (T >= 8) ? 3
                  : (T >= 4) ? 2 : (T >= 2) ? 1 : 0;
#pragma unroll
  for (int s = 0; s < k; ++s) {
    const int mask = 32 >> s;
    const bool up = (lane & mask) != 0;
#pragma unroll
    for (int i = 0; i < (T >> (s + 1)); ++i) {
      const double send = up ? a[i] : a[i + (T >> (s + 1))];
      const double keep = up ? a[i + (T >> (s + 1))] : a[i];
      a[i] = keep + __shfl_xor(send, mask);
    }
  }
  double v = a[0];
#pragma unroll
  for (int mask = 32 >> k; mask > 0; mask >>= 1) v += __shfl_xor(v, mask);
  return v;
}

// One wave: acc[j0:j0+wb] <- L_bb^{-1} acc[j0:j0+wb] (unit lower, forward
// substitution); lane i holds row i of the block.  Returns the lane's value
// (0 for lanes >= wb).
__device__ inline double fwd_diag_wave(gcdouble *PL, int ldl, int j0, int wb, double v,
                                       int lane) {
  double l[kBlk];
  const int ri = j0 + imin(lane, wb - 1);
#pragma unroll
  for (int j = 0; j < kBlk; ++j) {
    l[j] = masked(PL[(size_t)(j0 + imin(j, wb - 1)) * ldl + ri], j < lane && lane < wb);
  }
  __builtin_amdgcn_sched_barrier(0);  // all loads issued before the chain
  if (lane >= wb) v = 0.0;
#pragma unroll
  for (int j = 0; j < kBlk; ++j) v -= l[j] * bcast_lane(v, j);
  return v;
}

// One wave: x_b <- L_bb^{-T} x_b (backward substitution); lane j reads
// column j of the block from lb (lb[i][j] = L[j0+i, j0+j] for i > j, else 0).
__device__ inline double bwd_diag_wave(const double (*lb)[kBlk + 1], double v, int lane) {
  double c[kBlk];
#pragma unroll
  for (int i = 0; i < kBlk; ++i) c[i] = lb[i][lane];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = kBlk - 1; i >= 0; --i) v -= c[i] * bcast_lane(v, i);
  return v;
}

// lb[i][j] = L[j0+i, j0+j] for j < i < wb, else 0 (whole workgroup)
__device__ inline void stage_lower(gcdouble *PL, int ldl, int j0, int wb,
                                   double (*lb)[kBlk + 1]) {
#pragma unroll
  for (int it = 0; it < kBlk * kBlk / kThreads; ++it) {
    const int idx = threadIdx.x + it * kThreads, i = idx & (kBlk - 1), j = idx / kBlk;
    lb[i][j] = masked(PL[(size_t)(j0 + imin(j, wb - 1)) * ldl + j0 + imin(i, wb - 1)],
                      j < i && i < wb);
  }
}

// ---- forward --------------------------------------------------------------
// One workgroup per front.  X[0:n] = own b + the children's updates on
// fully-summed positions (children taken four at a time, a wave each; their
// adds applied wave by wave, so every position is summed in a fixed order),
// then the triangle X[0:ne] <- L11^{-1} X[0:ne] when ne <= kBigTri.
__global__ __launch_bounds__(kThreads) void k_fwd_tri(const SolveNode *sn, const int *slots,
                                                      const int *child_list, double *x,
                                                      int ldx, int nrhs, double *ws) {
  __shared__ double xs[kTriCap];
  __shared__ double yb[kBlk];
  __shared__ int nchunk[kGroups];
  const SolveNode &P = sn[slots[blockIdx.x]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gcint *Pipos = (gcint *)P.ipos;
  gcint *Prows = (gcint *)P.rows;
  gcint *Pctab = (gcint *)P.ctab;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, n = P.n, nd = P.nd, ne = P.nelim;
  const int cbeg = P.cbeg, cend = P.cend;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int T = idiv_up(m, 64), Tn = idiv_up(n, 64);
  for (int r = 0; r < nrhs; ++r) {
    gdouble *X = PX + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    double *acc = (n <= kTriCap) ? xs : (double *)X;
    for (int p = tid; p < n; p += kThreads) {
      const int vrow = (p >= nd) ? Prows[p - nd] : -1;
      acc[Pipos[p]] = (vrow >= 0) ? xr[vrow] : 0.0;
    }
    for (int c0 = cbeg; c0 < cend; c0 += kGroups) {
      const int ci = c0 + wave;
      int kb = 0, ke = 0, cn = 0, cne = 0, cdoff = 0, cm = 0;
      gcint *crmap = nullptr;
      gcdouble *cX = nullptr;
      if (ci < cend) {
        const SolveNode &c = sn[child_list[ci]];
        cn = c.n;
        cne = c.nelim;
        cdoff = c.doff;
        cm = c.m;
        crmap = (gcint *)c.rmap;
        cX = (gcdouble *)solve_X(ws, c, nrhs);
        gcint *tab = Pctab + (size_t)(ci - cbeg) * (T + 1);
        kb = tab[0];
        ke = tab[Tn];
      }
      if (lane == 0) nchunk[wave] = idiv_up(ke - kb, 64);
      __syncthreads();
      const int nch = imax(imax(nchunk[0], nchunk[1]), imax(nchunk[2], nchunk[3]));
      __syncthreads();  // nchunk is rewritten by the next group
      for (int it = 0; it < nch; ++it) {
        const int k = kb + it * 64 + lane;
        int pos = -1;
        double val = 0.0;
        if (k < ke) {
          const int pre = (k < cn) ? cdoff + (k - cne) : nd + crmap[k - cn];
          if (pre < n) {
            pos = Pipos[pre];
            val = cX[(size_t)r * cm + k];
          }
        }
        for (int w = 0; w < kGroups; ++w) {
          if (wave == w && pos >= 0) acc[pos] += val;
          __syncthreads();
        }
      }
    }
    __syncthreads();
    if (ne <= kBigTri) {
      for (int j0 = 0; j0 < ne; j0 += kBlk) {
        const int j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
        if (wave == 0) {
          const double v =
              fwd_diag_wave(PL, ldl, j0, wb, (lane < wb) ? acc[j0 + lane] : 0.0, lane);
          if (lane < wb) acc[j0 + lane] = v;
          yb[lane] = v;
        }
        __syncthreads();
        const int rows = ne - j1;  // rows of the triangle below the block
        if (tid < rows) {
          const int i = j1 + tid;
          double s = 0.0;
#pragma unroll
          for (int u0 = 0; u0 < kBlk; u0 += 16) {
            double lv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
              lv[u] = PL[(size_t)(j0 + imin(u0 + u, wb - 1)) * ldl + i];
#pragma unroll
            for (int u = 0; u < 16; ++u) s += lv[u] * yb[u0 + u];
          }
          acc[i] -= s;
        }
        __syncthreads();
      }
      for (int p = tid; p < ne; p += kThreads) xr[Pperm[p]] = acc[p];
    }
    if (acc != (double *)X)
      for (int p = tid; p < n; p += kThreads) X[p] = acc[p];
    __syncthreads();
  }
}

// k_fwd_tri for nrhs >= 2: each wave takes its own right-hand sides
// (r = wave, wave + 4, ...), so no workgroup barrier falls inside a
// right-hand side.  Every sum keeps k_fwd_tri's order (the children's
// values arrive on k_fwd_tri's schedule; a row's triangle products are
// summed in column order), so each column is bitwise what k_fwd_tri gives.
// Wave-scope fences order one lane's LDS / global writes before another
// lane's reads.
__device__ inline void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__global__ __launch_bounds__(kThreads) void k_fwd_tri_mr(const SolveNode *sn, const int *slots,
                                                         const int *child_list, double *x,
                                                         int ldx, int nrhs, double *ws) {
  __shared__ double xs[kGroups][kTriCap];
  __shared__ double ybs[kGroups][kBlk];
  const SolveNode &P = sn[slots[blockIdx.x]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gcint *Pipos = (gcint *)P.ipos;
  gcint *Prows = (gcint *)P.rows;
  gcint *Pctab = (gcint *)P.ctab;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, n = P.n, nd = P.nd, ne = P.nelim;
  const int cbeg = P.cbeg, cend = P.cend;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int T = idiv_up(m, 64), Tn = idiv_up(n, 64);
  double *yb = ybs[wave];
  for (int r = wave; r < nrhs; r += kGroups) {
    gdouble *X = PX + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    double *acc = (n <= kTriCap) ? xs[wave] : (double *)X;
    for (int p = lane; p < n; p += 64) {
      const int vrow = (p >= nd) ? Prows[p - nd] : -1;
      acc[Pipos[p]] = (vrow >= 0) ? xr[vrow] : 0.0;
    }
    wave_fence();
    // k_fwd_tri's schedule: children four at a time, 64-entry chunk by
    // chunk, the four children's chunks in turn (a position fed by several
    // children receives their values in that order)
    for (int c0 = cbeg; c0 < cend; c0 += kGroups) {
      int nch = 0;
      for (int g = 0; g < kGroups && c0 + g < cend; ++g) {
        gcint *tab = Pctab + (size_t)(c0 + g - cbeg) * (T + 1);
        nch = imax(nch, idiv_up(tab[Tn] - tab[0], 64));
      }
      for (int it = 0; it < nch; ++it) {
        for (int g = 0; g < kGroups && c0 + g < cend; ++g) {
          const int ci = c0 + g;
          const SolveNode &c = sn[child_list[ci]];
          gcint *tab = Pctab + (size_t)(ci - cbeg) * (T + 1);
          const int kb = tab[0], ke = tab[Tn];
          const int k = kb + it * 64 + lane;
          if (k < ke) {
            const int pre = (k < c.n) ? c.doff + (k - c.nelim) : nd + c.rmap[k - c.n];
            if (pre < n) acc[Pipos[pre]] += solve_X(ws, c, nrhs)[(size_t)r * c.m + k];
          }
          wave_fence();
        }
      }
    }
    if (ne <= kBigTri) {
      for (int j0 = 0; j0 < ne; j0 += kBlk) {
        const int j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
        const double v =
            fwd_diag_wave(PL, ldl, j0, wb, (lane < wb) ? acc[j0 + lane] : 0.0, lane);
        if (lane < wb) acc[j0 + lane] = v;
        yb[lane] = v;
        wave_fence();
        for (int i = j1 + lane; i < ne; i += 64) {
          double sum = 0.0;
#pragma unroll
          for (int u0 = 0; u0 < kBlk; u0 += 16) {
            double lv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
              lv[u] = PL[(size_t)(j0 + imin(u0 + u, wb - 1)) * ldl + i];
#pragma unroll
            for (int u = 0; u < 16; ++u) sum += lv[u] * yb[u0 + u];
          }
          acc[i] -= sum;
        }
        wave_fence();
      }
      for (int p = lane; p < ne; p += 64) xr[Pperm[p]] = acc[p];
    }
    if (acc != (double *)X)
      for (int p = lane; p < n; p += 64) X[p] = acc[p];
    wave_fence();
  }
}

// 64-row workgroups over [ne, m) of every front of the level (RC right-hand
// sides per pass):
//   X[i] = (gathered fully-summed value, i < n) or (sum of the children's
//          updates on row i, i >= n)  -  L[i, 0:ne] X[0:ne]
// Thread (row rr, group g) takes columns g, g+4, ...; the four groups' sums
// are added in group order.
template <int RC>
__global__ __launch_bounds__(kThreads) void k_fwd_rows(const SolveNode *sn, const int *slots,
                                                       const int *child_list,
                                                       const int *wg_off, int nlvl, int nrhs,
                                                       double *ws) {
  __shared__ double xs[RC][kRows];
  __shared__ double ys[RC][kYChunk];
  __shared__ double red[kGroups][RC][kRows];
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  const SolveNode &P = sn[slots[li]];
  const int ne = P.nelim, m = P.m;
  const int r0 = ne + (blockIdx.x - wg_off[li]) * kRows, r1 = imin(m, r0 + kRows);
  if (r0 >= r1) return;  // front without rows below
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pctab = (gcint *)P.ctab;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  const int ldl = P.ldl, n = P.n, nd = P.nd, cbeg = P.cbeg, cend = P.cend;
  const int tid = threadIdx.x, lane = tid & 63, rr = lane, g = tid >> 6;
  const int row = imin(r0 + rr, r1 - 1);
  const int T = idiv_up(m, 64);
  const int lo = imax(r0, n), I0 = lo / 64, I1 = idiv_up(r1, 64);
  for (int rc = 0; rc < nrhs; rc += RC) {
    const int nr = imin(RC, nrhs - rc);
    for (int idx = tid; idx < RC * kRows; idx += kThreads) {
      const int q = idx / kRows, i = idx % kRows, rw = r0 + i;
      xs[q][i] = (q < nr && rw < n && rw < r1) ? PX[(size_t)(rc + q) * m + rw] : 0.0;
    }
    __syncthreads();
    if (lo < r1) {
      // children four at a time (a wave each); a 64-row range holds at most
      // 128 positions, so at most two indices per lane
      for (int c0 = cbeg; c0 < cend; c0 += kGroups) {
        const int ci = c0 + g;
        int pos[2] = {-1, -1};
        double val[2][RC];
        if (ci < cend) {
          const SolveNode &c = sn[child_list[ci]];
          const int cn = c.n, cne = c.nelim, cdoff = c.doff, cm = c.m;
          gcint *crmap = (gcint *)c.rmap;
          gcdouble *cX = (gcdouble *)solve_X(ws, c, nrhs);
          gcint *tab = Pctab + (size_t)(ci - cbeg) * (T + 1);
          const int kb = tab[I0], ke = tab[I1];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int k = kb + lane + 64 * h;
            if (k < ke) {
              const int pre = (k < cn) ? cdoff + (k - cne) : nd + crmap[k - cn];
              if (pre >= lo && pre < r1) {
                pos[h] = pre - r0;
#pragma unroll
                for (int q = 0; q < RC; ++q)
                  val[h][q] = (q < nr) ? cX[(size_t)(rc + q) * cm + k] : 0.0;
              }
            }
          }
        }
        for (int w = 0; w < kGroups; ++w) {
          if (g == w) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
              if (pos[h] >= 0) {
#pragma unroll
                for (int q = 0; q < RC; ++q) xs[q][pos[h]] += val[h][q];
              }
          }
          __syncthreads();
        }
      }
    }
    double acc[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[q] = 0.0;
    for (int j0 = 0; j0 < ne; j0 += kYChunk) {
      const int jw = imin(kYChunk, ne - j0);
      for (int idx = tid; idx < RC * kYChunk; idx += kThreads) {
        const int q = idx / kYChunk, j = idx % kYChunk;
        ys[q][j] = (q < nr && j < jw) ? PX[(size_t)(rc + q) * m + j0 + j] : 0.0;
      }
      __syncthreads();
      for (int t0 = 0; t0 < kYChunk / kGroups && kGroups * t0 < jw; t0 += 16) {
        double lv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int jl = g + kGroups * (t0 + u);
          lv[u] = PL[(size_t)(j0 + imin(jl, jw - 1)) * ldl + row];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int jl = g + kGroups * (t0 + u);
#pragma unroll
          for (int q = 0; q < RC; ++q) acc[q] += lv[u] * ys[q][jl];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < RC; ++q) red[g][q][rr] = acc[q];
    __syncthreads();
    for (int idx = tid; idx < RC * kRows; idx += kThreads) {
      const int q = idx / kRows, i = idx % kRows;
      if (q < nr && r0 + i < r1)
        PX[(size_t)(rc + q) * m + r0 + i] =
            xs[q][i] - (((red[0][q][i] + red[1][q][i]) + red[2][q][i]) + red[3][q][i]);
    }
    __syncthreads();
  }
}

// x = D^{-1} x over every front; half = 1 (positive definite, 1x1 pivots
// only): x = D^{-1/2} x
__global__ void k_dsolve(const SolveNode *sn, int nslots, double *x, int ldx, int nrhs,
                         int half) {
  const int s = blockIdx.x;
  if (s >= nslots) return;
  const SolveNode &P = sn[s];
  gcint *Pperm = (gcint *)P.perm;
  gcdouble *d = (gcdouble *)P.D;
  const int ne = P.nelim;
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    bool first2 = (i + 1 < ne) && !isfinite(d[2 * i + 2]);
    bool second2 = !isfinite(d[2 * i]);
    for (int r = 0; r < nrhs; ++r) {
      double *xr = x + (size_t)r * ldx;
      if (half) {
        xr[Pperm[i]] *= sqrt(d[2 * i]);
      } else if (first2) {
        double x1 = xr[Pperm[i]], x2 = xr[Pperm[i + 1]];
        xr[Pperm[i]] = d[2 * i] * x1 + d[2 * i + 1] * x2;
        xr[Pperm[i + 1]] = d[2 * i + 1] * x1 + d[2 * i + 3] * x2;
      } else if (!second2) {
        xr[Pperm[i]] *= d[2 * i];
      }
    }
  }
}

// ---- backward -------------------------------------------------------------
// 256-row workgroups over [ne, m) of every front of the level (a row per
// thread):  part[wg][r][j] = sum over the workgroup's rows i of
// L[i, j] x[var(i), r].  Columns go 16 at a time: the wave's 64 row products
// are summed by a reduce-scatter, then the four waves' sums in wave order.
template <int RC>
__global__ __launch_bounds__(kThreads) void k_bwd_rows(const SolveNode *sn, const int *slots,
                                                       const int *wg_off, int nlvl,
                                                       const double *x, int ldx, int nrhs,
                                                       double *ws) {
  __shared__ double red[kGroups][RC][kBlk];
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  const SolveNode &P = sn[slots[li]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gcint *Prows = (gcint *)P.rows;
  gdouble *part = (gdouble *)solve_P(ws, P, nrhs);  // count x nrhs x ne
  const int ldl = P.ldl, m = P.m, n = P.n, nd = P.nd, ne = P.nelim;
  const int local = blockIdx.x - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = ne + local * kThreads, r1 = imin(m, r0 + kThreads);
  const int i = r0 + tid, row = imin(i, r1 - 1);
  const int var = (i < r1) ? ((i < n) ? Pperm[i] : Prows[i - nd]) : -1;
  for (int rc = 0; rc < nrhs; rc += RC) {
    const int nr = imin(RC, nrhs - rc);
    double xv[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) xv[q] = (q < nr && var >= 0) ? x[(size_t)(rc + q) * ldx + var] : 0.0;
    for (int j0 = 0; j0 < ne; j0 += kBlk) {
      for (int u0 = 0; u0 < kBlk && j0 + u0 < ne; u0 += 16) {
        double lv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) lv[u] = PL[(size_t)imin(j0 + u0 + u, ne - 1) * ldl + row];
#pragma unroll
        for (int q = 0; q < RC; ++q) {
          double a[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) a[u] = lv[u] * xv[q];
          const double s = wave_reduce_scatter<16>(a, lane);
          if ((lane & 3) == 0) red[wave][q][u0 + (lane >> 2)] = s;
        }
      }
      __syncthreads();
      for (int idx = tid; idx < RC * kBlk; idx += kThreads) {
        const int q = idx / kBlk, c = idx % kBlk;
        if (q < nr && j0 + c < ne)
          part[((size_t)local * nrhs + rc + q) * ne + j0 + c] =
              ((red[0][q][c] + red[1][q][c]) + red[2][q][c]) + red[3][q][c];
      }
      __syncthreads();
    }
  }
}

// One workgroup per front of the level: X[j] = x[perm[j]] - sum of the row
// workgroups' partials (wave w sums workgroups w, w+4, ...; the waves' sums
// are subtracted in wave order), then for ne <= kBigTri the triangle
// L11^T X = X, 64-column blocks from the last, and x[perm[j]] = X[j].
__global__ __launch_bounds__(kThreads) void k_bwd_tri(const SolveNode *sn, const int *slots,
                                                      const int *wg_off, double *x, int ldx,
                                                      int nrhs, double *ws) {
  __shared__ double xt[kBigTri];
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double red[kGroups][kBlk];
  __shared__ double pr[kGroups][kYChunk];
  const int li = blockIdx.x;
  const SolveNode &P = sn[slots[li]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gcdouble *part = (gcdouble *)solve_P(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim;
  const int count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int r = 0; r < nrhs; ++r) {
    gdouble *X = PX + (size_t)r * m;
    double *xr = x + (size_t)r * ldx;
    double *acc = (ne <= kBigTri) ? xt : (double *)X;
    for (int jc = 0; jc < ne; jc += kYChunk) {
      double s[kYChunk / 64];
#pragma unroll
      for (int k = 0; k < kYChunk / 64; ++k) s[k] = 0.0;
#pragma unroll 4
      for (int w = wave; w < count; w += kGroups) {
        gcdouble *pw = part + ((size_t)w * nrhs + r) * ne;
#pragma unroll
        for (int k = 0; k < kYChunk / 64; ++k) {
          const int j = jc + lane + 64 * k;
          if (j < ne) s[k] += pw[j];
        }
      }
#pragma unroll
      for (int k = 0; k < kYChunk / 64; ++k) pr[wave][lane + 64 * k] = s[k];
      __syncthreads();
      const int j = jc + tid;
      if (j < ne)
        acc[j] = xr[Pperm[j]] - (((pr[0][tid] + pr[1][tid]) + pr[2][tid]) + pr[3][tid]);
      __syncthreads();
    }
    if (ne > kBigTri) continue;  // the triangle is solved by the big-front launches
    for (int b = idiv_up(ne, kBlk) - 1; b >= 0; --b) {
      const int j0 = b * kBlk, j1 = imin(j0 + kBlk, ne), wb = j1 - j0, rows = ne - j1;
      stage_lower(PL, ldl, j0, wb, lb);
      // X[j] -= sum over the triangle's rows i in [j1, ne) of L[i, j] X[i]
      double s = 0.0;
      if (wave * 64 < rows) {
        const int i = j1 + imin(tid, rows - 1);
        const double xi = (tid < rows) ? xt[i] : 0.0;
        double a[kBlk];
#pragma unroll
        for (int u = 0; u < kBlk; ++u)
          a[u] = PL[(size_t)(j0 + imin(u, wb - 1)) * ldl + i] * xi;
        s = wave_reduce_scatter<kBlk>(a, lane);
      }
      red[wave][lane] = s;
      __syncthreads();
      if (wave == 0) {
        double v = 0.0;
        if (lane < wb)
          v = xt[j0 + lane] - (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
        v = bwd_diag_wave(lb, v, lane);
        if (lane < wb) xt[j0 + lane] = v;
      }
      __syncthreads();
    }
    for (int j = tid; j < ne; j += kThreads) xr[Pperm[j]] = xt[j];
    __syncthreads();
  }
}

// k_bwd_tri for nrhs >= 2: the right-hand sides four at a time, wave w on
// right-hand side r0 + w; the staged diagonal block is shared by the four.
// Each wave reproduces the four row groups of k_bwd_tri (its partial sums
// over workgroups g, g+4, ..., and its triangle rows [j1 + 64g, ...)) and
// combines them in k_bwd_tri's order, so each column is bitwise what
// k_bwd_tri gives.
__global__ __launch_bounds__(kThreads) void k_bwd_tri_mr(const SolveNode *sn, const int *slots,
                                                         const int *wg_off, double *x,
                                                         int ldx, int nrhs, double *ws) {
  __shared__ double xts[kGroups][kBigTri];
  __shared__ double lb[kBlk][kBlk + 1];
  const int li = blockIdx.x;
  const SolveNode &P = sn[slots[li]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gcdouble *part = (gcdouble *)solve_P(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim;
  const int count = wg_off[li + 1] - wg_off[li];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int r0 = 0; r0 < nrhs; r0 += kGroups) {
    const int r = r0 + wave;
    const bool mine = r < nrhs;
    double *xt = xts[wave];
    if (mine) {
      gdouble *X = PX + (size_t)r * m;
      double *xr = x + (size_t)r * ldx;
      double *acc = (ne <= kBigTri) ? xt : (double *)X;
      for (int jc = 0; jc < ne; jc += kYChunk) {
#pragma unroll
        for (int k = 0; k < kYChunk / 64; ++k) {
          const int j = jc + lane + 64 * k;
          double sg[kGroups];
#pragma unroll
          for (int g = 0; g < kGroups; ++g) {
            double sum = 0.0;
            for (int w = g; w < count; w += kGroups)
              if (j < ne) sum += part[((size_t)w * nrhs + r) * ne + j];
            sg[g] = sum;
          }
          if (j < ne) acc[j] = xr[Pperm[j]] - (((sg[0] + sg[1]) + sg[2]) + sg[3]);
        }
      }
    }
    if (ne > kBigTri) continue;  // the triangle is solved by the big-front launches
    __syncthreads();
    for (int b = idiv_up(ne, kBlk) - 1; b >= 0; --b) {
      const int j0 = b * kBlk, j1 = imin(j0 + kBlk, ne), wb = j1 - j0, rows = ne - j1;
      stage_lower(PL, ldl, j0, wb, lb);
      __syncthreads();  // the staged block is read by every wave
      if (mine) {
        double sg[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
          double sw = 0.0;
          if (g * 64 < rows) {
            const int t = 64 * g + lane;
            const int i = j1 + imin(t, rows - 1);
            const double xi = (t < rows) ? xt[i] : 0.0;
            double a[kBlk];
#pragma unroll
            for (int u = 0; u < kBlk; ++u)
              a[u] = PL[(size_t)(j0 + imin(u, wb - 1)) * ldl + i] * xi;
            sw = wave_reduce_scatter<kBlk>(a, lane);
          }
          sg[g] = sw;
        }
        double v = 0.0;
        if (lane < wb) v = xt[j0 + lane] - (((sg[0] + sg[1]) + sg[2]) + sg[3]);
        v = bwd_diag_wave(lb, v, lane);
        if (lane < wb) xt[j0 + lane] = v;
      }
      __syncthreads();
    }
    if (mine) {
      double *xr = x + (size_t)r * ldx;
      for (int j = lane; j < ne; j += 64) xr[Pperm[j]] = xt[j];
    }
    __syncthreads();
  }
}

// ---- big fronts -------------------------------------------------------------
// Forward, block [j0, j1): every workgroup solves the diagonal block from
// X[j0:j1] (as left by the previous block's launch); workgroup q then updates
// the triangle's rows [j1 + 64q, ...) with the solution, and workgroup 0 keeps
// the solution in Y (the front's partial-sum area, nrhs x ne).
__global__ __launch_bounds__(kThreads) void k_big_fwd(const SolveNode *sn, int s, int j0,
                                                      int nrhs, double *ws) {
  __shared__ double ys[kRhsChunk][kBlk];
  __shared__ double red[kGroups][kRhsChunk][kRows];
  const SolveNode &P = sn[s];
  gcdouble *PL = (gcdouble *)P.L;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gdouble *Y = (gdouble *)solve_P(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim, j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
  const int tid = threadIdx.x, lane = tid & 63, rr = lane, g = tid >> 6, wave = g;
  const int r0 = j1 + blockIdx.x * kRows, r1 = imin(ne, r0 + kRows);
  const int row = (r0 < r1) ? imin(r0 + rr, r1 - 1) : 0;
  for (int rc = 0; rc < nrhs; rc += kRhsChunk) {
    const int nr = imin(kRhsChunk, nrhs - rc);
    for (int q = wave; q < kRhsChunk; q += kGroups) {
      double v = 0.0;
      if (q < nr) {
        v = fwd_diag_wave(PL, ldl, j0, wb,
                          (lane < wb) ? PX[(size_t)(rc + q) * m + j0 + lane] : 0.0, lane);
        if (blockIdx.x == 0 && lane < wb) Y[(size_t)(rc + q) * ne + j0 + lane] = v;
      }
      ys[q][lane] = v;
    }
    __syncthreads();
    if (r0 < r1) {
      double acc[kRhsChunk];
#pragma unroll
      for (int q = 0; q < kRhsChunk; ++q) acc[q] = 0.0;
      double lv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        lv[u] = PL[(size_t)(j0 + imin(g + kGroups * u, wb - 1)) * ldl + row];
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int q = 0; q < kRhsChunk; ++q) acc[q] += lv[u] * ys[q][g + kGroups * u];
#pragma unroll
      for (int q = 0; q < kRhsChunk; ++q) red[g][q][rr] = acc[q];
    }
    __syncthreads();
    if (r0 < r1)
      for (int idx = tid; idx < kRhsChunk * kRows; idx += kThreads) {
        const int q = idx / kRows, i = idx % kRows;
        if (q < nr && r0 + i < r1)
          PX[(size_t)(rc + q) * m + r0 + i] -=
              ((red[0][q][i] + red[1][q][i]) + red[2][q][i]) + red[3][q][i];
      }
    __syncthreads();
  }
}

// X[0:ne] = Y and x[perm[i]] = Y[i] after the last forward block
__global__ void k_big_fwd_fin(const SolveNode *sn, int s, double *x, int ldx, int nrhs,
                              double *ws) {
  const SolveNode &P = sn[s];
  gcint *Pperm = (gcint *)P.perm;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gcdouble *Y = (gcdouble *)solve_P(ws, P, nrhs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, ne = P.nelim, m = P.m;
  if (i >= ne) return;
  const int v = Pperm[i];
  for (int r = 0; r < nrhs; ++r) {
    const double y = Y[(size_t)r * ne + i];
    PX[(size_t)r * m + i] = y;
    x[(size_t)r * ldx + v] = y;
  }
}

// Backward, block [j0, j1) (blocks from the last): every workgroup solves the
// diagonal block with L_bb^T from X[j0:j1] (complete: every later block has
// already pushed its contribution); workgroup q then pushes the solution into
// the columns [64q, 64q + 64) of X[0:j0]:  X[c] -= L[j0:j1, c]^T x_b.
// Workgroup 0 keeps the solution in Y (the front's partial-sum area).
__global__ __launch_bounds__(kThreads) void k_big_bwd(const SolveNode *sn, int s, int j0,
                                                      int nrhs, double *ws) {
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double xb[kRhsChunk][kBlk];
  const SolveNode &P = sn[s];
  gcdouble *PL = (gcdouble *)P.L;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gdouble *Y = (gdouble *)solve_P(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim, j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6, wave = g;
  const int c0 = blockIdx.x * kBlk, c1 = imin(j0, c0 + kBlk);  // columns of X to update
  stage_lower(PL, ldl, j0, wb, lb);
  __syncthreads();
  for (int rc = 0; rc < nrhs; rc += kRhsChunk) {
    const int nr = imin(kRhsChunk, nrhs - rc);
    for (int q = wave; q < kRhsChunk; q += kGroups) {
      double v = 0.0;
      if (q < nr) {
        v = bwd_diag_wave(lb, (lane < wb) ? PX[(size_t)(rc + q) * m + j0 + lane] : 0.0, lane);
        if (lane >= wb) v = 0.0;
        if (blockIdx.x == 0 && lane < wb) Y[(size_t)(rc + q) * ne + j0 + lane] = v;
      }
      xb[q][lane] = v;
    }
    __syncthreads();
    if (c0 < c1) {
      // lane = row t of the block, wave g = columns c0 + g + 4u
      double lv[16];
      const int t = j0 + imin(lane, wb - 1);
#pragma unroll
      for (int u = 0; u < 16; ++u) lv[u] = PL[(size_t)imin(c0 + g + kGroups * u, c1 - 1) * ldl + t];
      const int c = c0 + g + kGroups * (lane >> 2);
      for (int q = 0; q < nr; ++q) {
        double a[16];
        const double xt = xb[q][lane];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = lv[u] * xt;
        const double sum = wave_reduce_scatter<16>(a, lane);
        if ((lane & 3) == 0 && c < c1) PX[(size_t)(rc + q) * m + c] -= sum;
      }
    }
    __syncthreads();
  }
}

// x[perm[i]] = Y[i] after the first block
__global__ void k_big_bwd_fin(const SolveNode *sn, int s, double *x, int ldx, int nrhs,
                              double *ws) {
  const SolveNode &P = sn[s];
  gcint *Pperm = (gcint *)P.perm;
  gcdouble *Y = (gcdouble *)solve_P(ws, P, nrhs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x, ne = P.nelim;
  if (i >= ne) return;
  const int v = Pperm[i];
  for (int r = 0; r < nrhs; ++r) x[(size_t)r * ldx + v] = Y[(size_t)r * ne + i];
}

// ---- launchers ----------------------------------------------------------------
int big_tri_threshold() { return kBigTri; }
int fwd_rows_per_wg() { return kRows; }
int bwd_rows_per_wg() { return kThreads; }

void launch_big_fwd(hipStream_t st, const SolveNode *sn, int s, int ne, double *x, int ldx,
                    int nrhs, double *ws) {
  for (int j0 = 0; j0 < ne; j0 += kBlk) {
    const int rows = ne - std::min(j0 + kBlk, ne);
    hipLaunchKernelGGL(k_big_fwd, dim3(std::max(1, (rows + kRows - 1) / kRows)),
                       dim3(kThreads), 0, st, sn, s, j0, nrhs, ws);
  }
  hipLaunchKernelGGL(k_big_fwd_fin, dim3((ne + 255) / 256), dim3(256), 0, st, sn, s, x, ldx,
                     nrhs, ws);
}
void launch_big_bwd(hipStream_t st, const SolveNode *sn, int s, int ne, double *x, int ldx,
                    int nrhs, double *ws) {
  for (int j0 = (ne - 1) / kBlk * kBlk; j0 >= 0; j0 -= kBlk)
    hipLaunchKernelGGL(k_big_bwd, dim3(std::max(1, (j0 + kBlk - 1) / kBlk)), dim3(kThreads), 0,
                       st, sn, s, j0, nrhs, ws);
  hipLaunchKernelGGL(k_big_bwd_fin, dim3((ne + 255) / 256), dim3(256), 0, st, sn, s, x, ldx,
                     nrhs, ws);
}

void launch_fwd_tri(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                    const int *child_list, double *x, int ldx, int nrhs, double *ws) {
  if (nlvl <= 0) return;
  if (nrhs == 1)
    hipLaunchKernelGGL(k_fwd_tri, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, child_list, x,
                       ldx, nrhs, ws);
  else  // a wave per right-hand side: the same sums, no barrier per right-hand side
    hipLaunchKernelGGL(k_fwd_tri_mr, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, child_list,
                       x, ldx, nrhs, ws);
}
void launch_fwd_rows(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                     const int *child_list, int nrhs, const int *wg_off, int nwg, double *ws) {
  if (nlvl <= 0 || nwg <= 0) return;
  if (nrhs == 1)
    hipLaunchKernelGGL(k_fwd_rows<1>, dim3(nwg), dim3(kThreads), 0, st, sn, slots, child_list,
                       wg_off, nlvl, nrhs, ws);
  else
    hipLaunchKernelGGL(k_fwd_rows<kRhsChunk>, dim3(nwg), dim3(kThreads), 0, st, sn, slots,
                       child_list, wg_off, nlvl, nrhs, ws);
}
void launch_dsolve(hipStream_t st, int nslots, const SolveNode *sn, double *x, int ldx,
                   int nrhs, int half) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_dsolve, dim3(nslots), dim3(64), 0, st, sn, nslots, x, ldx, nrhs,
                       half);
}
// rows_off / nrows: the level's 256-row workgroup table and its total
void launch_bwd(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots, double *x,
                int ldx, int nrhs, const int *rows_off, int nrows, double *ws) {
  if (nlvl <= 0) return;
  if (nrows > 0) {
    if (nrhs == 1)
      hipLaunchKernelGGL(k_bwd_rows<1>, dim3(nrows), dim3(kThreads), 0, st, sn, slots, rows_off,
                         nlvl, x, ldx, nrhs, ws);
    else
      hipLaunchKernelGGL(k_bwd_rows<kRhsChunk>, dim3(nrows), dim3(kThreads), 0, st, sn, slots,
                         rows_off, nlvl, x, ldx, nrhs, ws);
  }
  if (nrhs == 1)
    hipLaunchKernelGGL(k_bwd_tri, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, rows_off, x,
                       ldx, nrhs, ws);
  else  // four right-hand sides at a time, one a wave
    hipLaunchKernelGGL(k_bwd_tri_mr, dim3(nlvl), dim3(kThreads), 0, st, sn, slots, rows_off, x,
                       ldx, nrhs, ws);
}

}  // namespace ssids_hip
