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
// Backward (root -> leaves), two launches per level:
//   k_bwd_rows_mr: 256-row workgroups over [ne, m) write partial sums of
//               L[rows, 0:ne]^T x[rows] (FP64 MFMA, up to 64 right-hand sides
//               a pass);
//   k_bwd_tri : one workgroup per front: the partials reduced in workgroup
//               order, then the triangle with L11^T.
// Fronts with ne > kBigTri (separators at the top of the tree) solve their
// triangle in one launch per level over 64-column blocks, the blocks of a
// front chained by flags (k_big_fwd_lvl / k_big_bwd_lvl).
// Several right-hand sides: the front kernels take one right-hand side per
// workgroup (gridDim.y of them in parallel) and the row kernels' passes run
// in parallel the same way, so every column is summed exactly as a
// single-right-hand-side solve sums it.
//
// Loads of L are issued in batches of 16 independent addresses (column indices
// clamped into the block instead of guarded, the surplus products multiplied
// by zero or discarded), and the 64-step diagonal-block substitutions keep the
// block in registers and broadcast each solved value with a lane read.
// Reference semantics: src/ssids/cpu/NumericSubtree.hxx:286-406 and
// src/ssids/cpu/kernels/ldlt_app.cxx:2537-2589.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>

#include "dev.hpp"

namespace ssids_hip {

constexpr int kRows = 64;                  // rows per workgroup of the row kernels
constexpr int kGroups = kThreads / kRows;  // column groups (one wave each)
constexpr int kYChunk = 256;               // columns of X[0:ne] staged per pass
constexpr int kRhsChunk = 8;
constexpr int kTriRhsPar = 64;  // right-hand sides solved in parallel by the front kernels
constexpr int kBigTri = 256;
constexpr int kTriCap = 1024;  // fully-summed positions gathered in LDS

__device__ inline int find_node_s(const int *wg_off, int nlvl, int b) {
  return table_find(wg_off, nlvl, b);
}

// FP64 MFMA accumulators of the multi-right-hand-side products
typedef double mr_acc_t __attribute__((ext_vector_type(4)));
constexpr int kMrChunks = 4;  // 16-column MFMA outputs per wave: 64 right-hand sides a pass
// independent accumulators per output (4-position step u goes to u % kMrSplit,
// the partial sums added in order at the end): 16 dependent MFMAs become
// 4 chains of 4
constexpr int kMrSplit = 4;
// right-hand sides a wave substitutes through a diagonal block together when
// there are more than four (independent chains of lane reads and FMAs in
// one instruction stream; each right-hand side keeps its own sequence)
constexpr int kDiagGroup = 4;

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
  for (int r = blockIdx.y; r < nrhs; r += gridDim.y) {
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
  for (int rc = blockIdx.y * RC; rc < nrhs; rc += gridDim.y * RC) {
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
// Workgroups over 256-row blocks of [ne, m) times kBwdCols-column chunks of
// [0, ne), for every front of the level (a row per thread):
// part[rb][r][j] = sum over row block rb's rows i of L[i, j] x[var(i), r].
// Columns go 16 at a time: the wave's 64 row products are summed by a
// reduce-scatter, then the four waves' sums in wave order.  Each (rb, j) is
// written by exactly one workgroup, so the sums do not depend on the split.
constexpr int kBwdCols = 64;

// k_bwd_rows on FP64 MFMA for any number of right-hand sides (the launch
// used for every nrhs, so each column is the same whatever nrhs): wave w
// takes columns 16w..16w+15 of the workgroup's 64-column chunk, the
// right-hand sides are the 16-wide output columns (64 a pass), and the
// workgroup's 256 rows the K dimension in steps of 4, in row order, dealt to
// kMrSplit accumulators whose sums are added in order.
// NCH: 16-column output chunks per pass (1 for up to 16 right-hand sides,
// which keeps the registers for several workgroups per CU; a column's
// arithmetic is the same in both instantiations)
template <int NCH>
__global__ __launch_bounds__(kThreads) void k_bwd_rows_mr(const SolveNode *sn, const int *slots,
                                                          const int *wg_off, int nlvl,
                                                          const double *x, int ldx, int nrhs,
                                                          double *ws) {
  constexpr int kSteps = 8;  // 4-row steps loaded per batch
  __shared__ int svar[kThreads];
  const int li = find_node_s(wg_off, nlvl, blockIdx.x);
  if (li < 0) return;
  const SolveNode &P = sn[slots[li]];
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  gcint *Prows = (gcint *)P.rows;
  gdouble *part = (gdouble *)solve_P(ws, P, nrhs);  // count x nrhs x ne
  const int ldl = P.ldl, m = P.m, n = P.n, nd = P.nd, ne = P.nelim;
  const int nchunk = idiv_up(imax(ne, 1), kBwdCols);
  const int local_all = blockIdx.x - wg_off[li];
  const int local = local_all / nchunk, chunk = local_all - local * nchunk;
  const int jbeg = chunk * kBwdCols, jend = imin(ne, jbeg + kBwdCols);
  const int r0 = ne + local * kThreads, r1 = imin(m, r0 + kThreads);
  if (jend <= jbeg || r1 <= r0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ri = lane & 15, kq = lane >> 4;
  {
    const int i = r0 + tid;
    svar[tid] = (i < r1) ? ((i < n) ? Pperm[i] : Prows[i - nd]) : 0;
  }
  __syncthreads();
  const int nrow = r1 - r0;
  // A operand: column acol of L (the lane's column of the wave's 16, clamped)
  const int acol = imin(jbeg + 16 * wave + ri, jend - 1);
  gcdouble *La = PL + (size_t)acol * ldl + r0;
  for (int rc = 0; rc < nrhs; rc += 16 * NCH) {
    const int nr = imin(16 * NCH, nrhs - rc);
    mr_acc_t acc[NCH][kMrSplit];
#pragma unroll
    for (int h = 0; h < NCH; ++h)
#pragma unroll
      for (int t = 0; t < kMrSplit; ++t) acc[h][t] = (mr_acc_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < nrow; k0 += 4 * kSteps) {
      double av[kSteps];
      int vv[kSteps];
#pragma unroll
      for (int u = 0; u < kSteps; ++u) {
        const int kk = k0 + 4 * u + kq;
        av[u] = masked(La[imin(kk, nrow - 1)], kk < nrow);
        vv[u] = svar[imin(kk, nrow - 1)];
      }
#pragma unroll
      for (int h = 0; h < NCH; ++h) {
        if (16 * h >= nr) break;
        const int q = 16 * h + ri;
        const double *xq = x + (size_t)(rc + imin(q, nr - 1)) * ldx;
        double bv[kSteps];
#pragma unroll
        for (int u = 0; u < kSteps; ++u)
          bv[u] = masked(xq[vv[u]], q < nr && k0 + 4 * u + kq < nrow);
#pragma unroll
        for (int u = 0; u < kSteps; ++u)
          acc[h][u % kMrSplit] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              av[u], bv[u], acc[h][u % kMrSplit], 0, 0, 0);
      }
    }
    // element (column kq + 4 rg of the wave's 16, right-hand side ri) of chunk h
#pragma unroll
    for (int h = 0; h < NCH; ++h)
#pragma unroll
      for (int rg = 0; rg < 4; ++rg) {
        const int c = 16 * wave + kq + 4 * rg, q = 16 * h + ri;
        if (jbeg + c < jend && q < nr)
          part[((size_t)local * nrhs + rc + q) * ne + jbeg + c] =
              ((acc[h][0][rg] + acc[h][1][rg]) + acc[h][2][rg]) + acc[h][3][rg];
      }
  }
}

// Value of position j after x = D^{-1} x (the expressions of k_dsolve, so
// the result is the same whether the diagonal solve runs on its own or
// inside the backward solve)
__device__ inline double dinv_at(gcdouble *d, gcint *perm, const double *xr, int j, int ne) {
  const bool first2 = (j + 1 < ne) && !isfinite(d[2 * j + 2]);
  const bool second2 = !isfinite(d[2 * j]);
  if (first2) return d[2 * j] * xr[perm[j]] + d[2 * j + 1] * xr[perm[j + 1]];
  if (second2) return d[2 * j - 1] * xr[perm[j - 1]] + d[2 * j + 1] * xr[perm[j]];
  return xr[perm[j]] * d[2 * j];
}

// One workgroup per front of the level: X[j] = x[perm[j]] - sum of the row
// workgroups' partials (wave w sums workgroups w, w+4, ...; the waves' sums
// are subtracted in wave order), then for ne <= kBigTri the triangle
// L11^T X = X, 64-column blocks from the last, and x[perm[j]] = X[j].
__global__ __launch_bounds__(kThreads) void k_bwd_tri(const SolveNode *sn, const int *slots,
                                                      const int *wg_off, double *x, int ldx,
                                                      int nrhs, double *ws, int dfuse) {
  __shared__ double xt[kBigTri];
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double red[kGroups][kBlk];
  __shared__ double pr[kGroups][kYChunk];
  const int li = blockIdx.x;
  const SolveNode &P = sn[slots[li]];
  gcdouble *PL = (gcdouble *)P.L;
  gcdouble *Pd = (gcdouble *)P.D;
  gcint *Pperm = (gcint *)P.perm;
  gdouble *PX = (gdouble *)solve_X(ws, P, nrhs);
  gcdouble *part = (gcdouble *)solve_P(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim;
  const int count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int r = blockIdx.y; r < nrhs; r += gridDim.y) {
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
        acc[j] = (dfuse ? dinv_at(Pd, Pperm, xr, j, ne) : xr[Pperm[j]]) -
                 (((pr[0][tid] + pr[1][tid]) + pr[2][tid]) + pr[3][tid]);
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

// ---- big fronts: one launch per tree level ---------------------------------
// Fronts with ne > kBigTri solve their triangle, L11 y = X[0:ne] (forward) or
// L11^T x = X[0:ne] (backward), in one launch per tree level over 64-column
// blocks: a workgroup per block of every big front of the level.  Within a
// front, workgroups take their blocks by ticket (an atomic counter; forward
// in increasing, backward in decreasing block order), so a workgroup waits
// only for blocks whose workgroups took their tickets earlier and are
// running: every wait ends.  Block b's workgroup accumulates the products of
// the blocks it depends on in block order (not arrival order), solves its
// diagonal block on one wave per right-hand side, stores the block's
// solution into X and x, and raises the block's flag.  The flag is a
// device-scope release store after a workgroup barrier; readers poll it with
// relaxed device-scope loads and read the published solutions with
// device-scope loads (coherent across the XCDs' L2 caches).  The sync words
// (per big front: forward ticket and one flag per block, then the same for
// the backward solve) live after the call's workspace and are cleared by
// k_big_sync_clear at the start of every solve.
__device__ inline double load_dev(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void wait_flag(const int *f) {
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    __builtin_amdgcn_s_sleep(1);
  // the solution loads that follow stay behind the poll (a compiler barrier;
  // the hardware issues them only after the poll's value has returned)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ inline void raise_flag(int *f) {
  __hip_atomic_store(f, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_big_sync_clear(int *sync, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sync[i] = 0;
}

// The products of the other blocks go to FP64 MFMA (v_mfma_f64_16x16x4f64)
// for up to 64 right-hand sides at once: wave w takes 16 rows (forward) or
// 16 columns (backward) of block b, the right-hand sides are the 16-wide
// output columns, and the 64 positions of each other block are the K
// dimension in steps of 4, in block order.  An output column depends only
// on its own right-hand side, so every column is the same whatever nrhs.
// The diagonal block is then solved once per right-hand side (lane = row,
// the right-hand sides dealt to the waves in turn).

// forward: block b's rows minus L[b rows, earlier columns] Y[earlier]
__global__ __launch_bounds__(kThreads) void k_big_fwd_lvl(const SolveNode *sn, const int *boff,
                                                          const int *bslot, const int *bsync,
                                                          int nbig, double *x, int ldx,
                                                          int nrhs, double *ws, int *sync) {
  __shared__ double red[16 * kMrChunks][kBlk];
  __shared__ int s_b;
  const int li = table_find(boff, nbig, blockIdx.x);
  if (li < 0) return;
  const SolveNode &P = sn[bslot[li]];
  int *tick = sync + bsync[li];
  int *flag = tick + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ri = lane & 15, kq = lane >> 4;
  if (tid == 0) s_b = atomicAdd(tick, 1);
  __syncthreads();
  const int b = s_b;
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  double *PX = solve_X(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m, ne = P.nelim;
  const int j0 = b * kBlk, j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
  // A operand row of this lane (the wave's 16 rows of block b, clamped)
  const int arow = j0 + imin(16 * wave + ri, wb - 1);
  // L of the diagonal block for the substitutions (lane = row), loaded once
  double l[kBlk];
  {
    const int rr = j0 + imin(lane, wb - 1);
#pragma unroll
    for (int j = 0; j < kBlk; ++j)
      l[j] = masked(PL[(size_t)(j0 + imin(j, wb - 1)) * ldl + rr], j < lane && lane < wb);
  }
  for (int rc = 0; rc < nrhs; rc += 16 * kMrChunks) {
    const int nr = imin(16 * kMrChunks, nrhs - rc);
    mr_acc_t acc[kMrChunks][kMrSplit];
#pragma unroll
    for (int h = 0; h < kMrChunks; ++h)
#pragma unroll
      for (int t = 0; t < kMrSplit; ++t) acc[h][t] = (mr_acc_t){0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < b; ++j) {  // every block before b is a full block
      const int c0 = j * kBlk;
      double av[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) av[u] = PL[(size_t)(c0 + 4 * u + kq) * ldl + arow];
      if (rc == 0) wait_flag(flag + j);
#pragma unroll
      for (int h = 0; h < kMrChunks; ++h) {
        if (16 * h >= nr) break;
        const int q = 16 * h + ri;
        const double *y = PX + (size_t)(rc + imin(q, nr - 1)) * m + c0 + kq;
        double bv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) bv[u] = masked(load_dev(y + 4 * u), q < nr);
#pragma unroll
        for (int u = 0; u < 16; ++u)
          acc[h][u % kMrSplit] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              av[u], bv[u], acc[h][u % kMrSplit], 0, 0, 0);
      }
    }
    // element (row kq + 4 rg of the wave's rows, right-hand side ri) of chunk h
#pragma unroll
    for (int h = 0; h < kMrChunks; ++h)
#pragma unroll
      for (int rg = 0; rg < 4; ++rg)
        red[16 * h + ri][16 * wave + kq + 4 * rg] =
            ((acc[h][0][rg] + acc[h][1][rg]) + acc[h][2][rg]) + acc[h][3][rg];
    __syncthreads();
    auto diag = [&](auto grp) {
      constexpr int G = decltype(grp)::value;
      for (int q0 = wave * G; q0 < nr; q0 += kGroups * G) {
        double v[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
          const int q = q0 + r;
          v[r] = (q < nr && lane < wb) ? PX[(size_t)(rc + q) * m + j0 + lane] - red[q][lane] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kBlk; ++j)
#pragma unroll
          for (int r = 0; r < G; ++r) v[r] = fma(-l[j], bcast_lane(v[r], j), v[r]);
#pragma unroll
        for (int r = 0; r < G; ++r) {
          const int q = q0 + r;
          if (q < nr && lane < wb) {
            PX[(size_t)(rc + q) * m + j0 + lane] = v[r];
            x[(size_t)(rc + q) * ldx + Pperm[j0 + lane]] = v[r];
          }
        }
      }
    };
    if (nr > kGroups) diag(std::integral_constant<int, kDiagGroup>());
    else diag(std::integral_constant<int, 1>());
    __syncthreads();
  }
  if (tid == 0) raise_flag(flag + b);
}

// backward: block b's values minus L[later rows, b columns]^T X[later]
__global__ __launch_bounds__(kThreads) void k_big_bwd_lvl(const SolveNode *sn, const int *boff,
                                                          const int *bslot, const int *bsync,
                                                          int nbig, double *x, int ldx,
                                                          int nrhs, double *ws, int *sync) {
  __shared__ double lb[kBlk][kBlk + 1];
  __shared__ double red[16 * kMrChunks][kBlk];
  __shared__ int s_b;
  const int li = table_find(boff, nbig, blockIdx.x);
  if (li < 0) return;
  const SolveNode &P = sn[bslot[li]];
  const int ne = P.nelim, nb = idiv_up(ne, kBlk);
  int *tick = sync + bsync[li] + nb + 1;
  int *flag = tick + 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ri = lane & 15, kq = lane >> 4;
  if (tid == 0) s_b = nb - 1 - atomicAdd(tick, 1);
  __syncthreads();
  const int b = s_b;
  gcdouble *PL = (gcdouble *)P.L;
  gcint *Pperm = (gcint *)P.perm;
  double *PX = solve_X(ws, P, nrhs);
  const int ldl = P.ldl, m = P.m;
  const int j0 = b * kBlk, j1 = imin(j0 + kBlk, ne), wb = j1 - j0;
  stage_lower(PL, ldl, j0, wb, lb);
  // A operand column of this lane (the wave's 16 columns of block b, clamped)
  const int acol = j0 + imin(16 * wave + ri, wb - 1);
  for (int rc = 0; rc < nrhs; rc += 16 * kMrChunks) {
    const int nr = imin(16 * kMrChunks, nrhs - rc);
    mr_acc_t acc[kMrChunks][kMrSplit];
#pragma unroll
    for (int h = 0; h < kMrChunks; ++h)
#pragma unroll
      for (int t = 0; t < kMrSplit; ++t) acc[h][t] = (mr_acc_t){0.0, 0.0, 0.0, 0.0};
    for (int j = nb - 1; j > b; --j) {
      const int r0 = j * kBlk, rows = imin(kBlk, ne - r0);
      double av[16];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        av[u] = masked(PL[(size_t)acol * ldl + r0 + imin(4 * u + kq, rows - 1)],
                       4 * u + kq < rows);
      if (rc == 0) wait_flag(flag + j);
#pragma unroll
      for (int h = 0; h < kMrChunks; ++h) {
        if (16 * h >= nr) break;
        const int q = 16 * h + ri;
        const double *xq = PX + (size_t)(rc + imin(q, nr - 1)) * m + r0;
        double bv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
          bv[u] = masked(load_dev(xq + imin(4 * u + kq, rows - 1)), q < nr && 4 * u + kq < rows);
#pragma unroll
        for (int u = 0; u < 16; ++u)
          acc[h][u % kMrSplit] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              av[u], bv[u], acc[h][u % kMrSplit], 0, 0, 0);
      }
    }
    // element (column kq + 4 rg of the wave's columns, right-hand side ri)
#pragma unroll
    for (int h = 0; h < kMrChunks; ++h)
#pragma unroll
      for (int rg = 0; rg < 4; ++rg)
        red[16 * h + ri][16 * wave + kq + 4 * rg] =
            ((acc[h][0][rg] + acc[h][1][rg]) + acc[h][2][rg]) + acc[h][3][rg];
    __syncthreads();
    double c[kBlk];
#pragma unroll
    for (int i = 0; i < kBlk; ++i) c[i] = lb[i][lane];
    auto diag = [&](auto grp) {
      constexpr int G = decltype(grp)::value;
      for (int q0 = wave * G; q0 < nr; q0 += kGroups * G) {
        double v[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
          const int q = q0 + r;
          v[r] = (q < nr && lane < wb) ? PX[(size_t)(rc + q) * m + j0 + lane] - red[q][lane] : 0.0;
        }
#pragma unroll
        for (int i = kBlk - 1; i >= 0; --i)
#pragma unroll
          for (int r = 0; r < G; ++r) v[r] = fma(-c[i], bcast_lane(v[r], i), v[r]);
#pragma unroll
        for (int r = 0; r < G; ++r) {
          const int q = q0 + r;
          if (q < nr && lane < wb) {
            PX[(size_t)(rc + q) * m + j0 + lane] = v[r];
            x[(size_t)(rc + q) * ldx + Pperm[j0 + lane]] = v[r];
          }
        }
      }
    };
    if (nr > kGroups) diag(std::integral_constant<int, kDiagGroup>());
    else diag(std::integral_constant<int, 1>());
    __syncthreads();
  }
  if (tid == 0) raise_flag(flag + b);
}

// ---- launchers ----------------------------------------------------------------
int big_tri_threshold() { return kBigTri; }
int fwd_rows_per_wg() { return kRows; }
int bwd_rows_per_wg() { return kThreads; }
int bwd_cols_per_wg() { return kBwdCols; }

// tab: the level's big fronts, [boff (nbig+1) | slots (nbig) | sync offsets
// (nbig)]; nwg = boff[nbig], the level's 64-column blocks
void launch_big_fwd_lvl(hipStream_t st, const SolveNode *sn, const int *tab, int nbig, int nwg,
                        double *x, int ldx, int nrhs, double *ws, int *sync) {
  if (nbig <= 0) return;
  hipLaunchKernelGGL(k_big_fwd_lvl, dim3(nwg), dim3(kThreads), 0, st, sn, tab, tab + nbig + 1,
                     tab + 2 * nbig + 1, nbig, x, ldx, nrhs, ws, sync);
}
void launch_big_bwd_lvl(hipStream_t st, const SolveNode *sn, const int *tab, int nbig, int nwg,
                        double *x, int ldx, int nrhs, double *ws, int *sync) {
  if (nbig <= 0) return;
  hipLaunchKernelGGL(k_big_bwd_lvl, dim3(nwg), dim3(kThreads), 0, st, sn, tab, tab + nbig + 1,
                     tab + 2 * nbig + 1, nbig, x, ldx, nrhs, ws, sync);
}
void launch_big_sync_clear(hipStream_t st, int *sync, int n) {
  if (n > 0)
    hipLaunchKernelGGL(k_big_sync_clear, dim3((n + 255) / 256), dim3(256), 0, st, sync, n);
}

void launch_fwd_tri(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                    const int *child_list, double *x, int ldx, int nrhs, double *ws) {
  if (nlvl <= 0) return;
  // a workgroup per front and right-hand side: each right-hand side gets the
  // single-right-hand-side sums, in parallel
  hipLaunchKernelGGL(k_fwd_tri, dim3(nlvl, imin(nrhs, kTriRhsPar)), dim3(kThreads), 0, st, sn,
                     slots, child_list, x, ldx, nrhs, ws);
}
void launch_fwd_rows(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots,
                     const int *child_list, int nrhs, const int *wg_off, int nwg, double *ws) {
  if (nlvl <= 0 || nwg <= 0) return;
  if (nrhs == 1)
    hipLaunchKernelGGL(k_fwd_rows<1>, dim3(nwg), dim3(kThreads), 0, st, sn, slots, child_list,
                       wg_off, nlvl, nrhs, ws);
  else  // the passes of kRhsChunk right-hand sides in parallel
    hipLaunchKernelGGL(k_fwd_rows<kRhsChunk>, dim3(nwg, idiv_up(nrhs, kRhsChunk)),
                       dim3(kThreads), 0, st, sn, slots, child_list, wg_off, nlvl, nrhs, ws);
}
void launch_dsolve(hipStream_t st, int nslots, const SolveNode *sn, double *x, int ldx,
                   int nrhs, int half) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_dsolve, dim3(nslots), dim3(64), 0, st, sn, nslots, x, ldx, nrhs,
                       half);
}
// rows_off: the level's 256-row workgroup table (per front); chunk_off /
// nwg: the row workgroups x column chunks table of k_bwd_rows and its total
void launch_bwd(hipStream_t st, int nlvl, const SolveNode *sn, const int *slots, double *x,
                int ldx, int nrhs, const int *rows_off, const int *chunk_off, int nwg,
                double *ws, int dfuse) {
  if (nlvl <= 0) return;
  if (nwg > 0) {
    if (nrhs <= 16)
      hipLaunchKernelGGL(k_bwd_rows_mr<1>, dim3(nwg), dim3(kThreads), 0, st, sn, slots,
                         chunk_off, nlvl, x, ldx, nrhs, ws);
    else
      hipLaunchKernelGGL(k_bwd_rows_mr<kMrChunks>, dim3(nwg), dim3(kThreads), 0, st, sn, slots,
                         chunk_off, nlvl, x, ldx, nrhs, ws);
  }
  hipLaunchKernelGGL(k_bwd_tri, dim3(nlvl, imin(nrhs, kTriRhsPar)), dim3(kThreads), 0, st, sn,
                     slots, rows_off, x, ldx, nrhs, ws, dfuse);
}

}  // namespace ssids_hip
