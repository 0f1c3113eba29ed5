// Numeric factorization kernels for gfx950 (CDNA4, wave64).
//
// One tree level is processed per launch sequence; a launch covers every
// front of the level.  Workgroups find their front by binary search in a
// per-level prefix array of workgroup counts (wg_off).
//
//  k_gather     gather assembly: a workgroup builds a 64x64 tile of one parent
//               front in LDS from its children (children in order =>
//               deterministic sums) through LDS-staged child->parent maps,
//               computing the children's Schur complements on MFMA on the way.
//  k_diag       one workgroup per front: factor the w<=64 column diagonal
//               block with complete pivoting (1x1/2x2) in registers
//               (pivot rules of src/ssids/cpu/kernels/block_ldlt.hxx:213-414).
//  k_apply      L = A Lbb^{-T} D^{-1} for the rows below the block, keeping a
//               backup and the column max |l| (a posteriori test,
//               src/ssids/cpu/kernels/ldlt_app.cxx:304-414).
//  k_update     restore failed columns and apply the accepted pivots to the
//               trailing panel columns with v_mfma_f64_16x16x4_f64 on 64x64
//               tiles (the contribution block is updated lazily by k_gather).
//  k_swap_*     move failed columns behind the unprocessed candidates.
//  k_tail       end of a level, one workgroup per front: APP result,
//               threshold partial pivoting on the failed columns
//               (src/ssids/cpu/kernels/ldlt_tpp.cxx:166-270), D and L21 D
//               operands of the parent's gather, solve position index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "dev.hpp"

namespace ssids_hip {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// front of workgroup b in a level table, or -1 past the table's total (grids
// sized for the plan's capacities may be larger than the level needs)
__device__ inline int find_node(const int *wg_off, int nlvl, int b) {
  return table_find(wg_off, nlvl, b);
}

// element (i,j), i>=j, of the front (panel for j<n, contribution block else)
__device__ inline double *front_at(const NodeDev &N, int i, int j) {
  if (j < N.n) return N.L + (size_t)j * N.ldl + i;
  return N.C + (size_t)(j - N.n) * N.ldc + (i - N.n);
}

// parent position of trailing index k of child c (monotone in k)
__device__ inline int child_pos(const NodeDev &c, int doff, int pnd, int k) {
  if (k < c.n) return doff + (k - c.nelim);
  return pnd + c.rmap[k - c.n];
}

__device__ inline const double *child_src(const NodeDev &c, int kr, int k) {
  if (k < c.n) return c.L + (size_t)k * c.ldl + kr;
  return c.C + (size_t)(k - c.n) * c.ldc + (kr - c.n);
}

// ---------------------------------------------------------------------------
// Diagonal block factorization, NW waves per front.  Lane j of wave v keeps
// rows [vR, vR+R) (R = 64/NW) of column j of the symmetric w x w block in
// registers; pivots are tracked in a uniform bit mask (no physical swaps).
// Per pivot step:
//  * search: every lane scans its live rows i >= j for max |a(i,j)| (the
//    lower triangle, as block_ldlt.hxx:77-101 searches it: each symmetric
//    pair once, in one orientation, so a maximum off the diagonal gives the
//    pair (column, row) with column < row whatever the rounding of the two
//    mirrored entries; rows ascending, first maximum kept); each wave
//    reduces to its (value, smallest column, that column's row) with DPP
//    and a ballot, and the waves' candidates are merged through LDS in the
//    same order;
//  * the owners of the pivot column(s) broadcast them through LDS;
//  * rank-1 / rank-2 update of every column's rows in registers;
//  * the L column(s) of the step go to LDS, indexed by elimination step.
// Pivot rules of src/ssids/cpu/kernels/block_ldlt.hxx:289-414: maximum on the
// diagonal -> 1x1; else the 2x2 test of block_ldlt.hxx:213-218, else the
// larger diagonal as 1x1.  A maximum below `small` makes every remaining
// index a zero pivot.  The a posteriori test of k_apply then checks
// |l| <= 1/u over the full column.  The elimination order lp[] permutes the
// block on write back.
// ---------------------------------------------------------------------------
constexpr int kDiagWaves = 4;

constexpr int kGT = 64;  // gather tile edge

template <int PART, typename CT>
__device__ __forceinline__ int gather_tile(
    NodeDev *nodes, const int *lvl_slots, const int *wg_off, int nlvl,
    const int *child_list, const double *aval, const int64_t *asrc, const int *arow,
    const int *acix, const double *scaling, double (&acc)[kGT][kGT + 1], int (&mr)[kGT],
    int (&mc)[kGT]);

struct GatherArgs {  // the level's gather inputs (a pivot launch that assembles)
  const int *wg_off;
  int nlvl;
  const int *child_list;
  const double *aval;
  const int64_t *asrc;
  const int *arow;
  const int *acix;
  const double *scaling;
};

// FIRST: one workgroup per front gathers the front's diagonal tile (0, 0)
// (the gather's part 1) and runs step 0's pivot block on it, as soon as its
// block is assembled and in a workgroup that already holds its CU, instead of
// queueing behind the rest of the level's gather.  The first tile column's
// other tiles (part 3) are gathered by k_gather, whose smaller LDS fits
// twice as many workgroups on a CU.
// C32 (FIRST only): the children's contribution blocks are FP32
template <int NW, bool FIRST, bool C32>
__global__ __launch_bounds__(64 * NW) void k_diag(NodeDev *nodes, const int *lvl_slots,
                                                  int step, FactorParams fp, GatherArgs ga) {
  constexpr int R = kBlk / NW;
  __shared__ double t[kBlk][kBlk + 1];  // load staging, then L by step: t[k][row]
  __shared__ double cs2[2][2][kBlk];    // broadcast pivot columns (by step parity)
  __shared__ unsigned long long ckey[2][NW];
  __shared__ int ccr[2][NW];  // candidate column << 7 | row
  __shared__ double dinv[2 * kBlk + 2];
  __shared__ int lp[kBlk], inv_lp[kBlk], kind[kBlk];
  __shared__ double mi[kBlk][kBlk + 1];  // inverse of L_bb (elimination order)
  int s;
  if constexpr (FIRST) {
    static_assert(64 * NW == kThreads, "one workgroup shape for both phases");
    static_assert(kGT == kBlk, "the gather tile is the pivot block's staging array");
    // the tile's LDS is the pivot block's own (t, lp, inv_lp), used in turn
    typedef typename std::conditional<C32, float, double>::type CT;
    s = gather_tile<1, CT>(nodes, lvl_slots, ga.wg_off, ga.nlvl, ga.child_list, ga.aval,
                           ga.asrc, ga.arow, ga.acix, ga.scaling, t, lp, inv_lp);
    if (s < 0) return;
    __syncthreads();  // the block's stores are visible to the whole workgroup
  } else {
    s = lvl_slots[blockIdx.x];
  }
  NodeDev &N = nodes[s];
  if (step >= N.nsteps) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r0 = wv * R;
  int p, q;
  if (step == 0) {
    p = 0;
    q = N.n;
  } else {
    p = N.p + N.npass;
    q = N.q - (N.w - N.npass);
  }
  const int w = max(0, min(kBlk, q - p));
  __syncthreads();  // every thread has read N.p/N.npass before they change
  if (tid == 0) {
    N.p = p;
    N.q = q;
    N.w = w;
    N.npass = 0;
  }
  if (w == 0) return;
  const int ldl = N.ldl;
  double *L = N.L;
  {
    // the w x w block (lower triangle mirrored), zero outside; every thread's
    // loads are unconditional (indices clamped into the block) so they are
    // all in flight together
    gcdouble *Lg = (gcdouble *)L;
    gdouble *dbak = (gdouble *)N.dbak;
    double v[kBlk * kBlk / (64 * NW)];
#pragma unroll
    for (int it = 0; it < kBlk * kBlk / (64 * NW); ++it) {
      const int idx = tid + it * 64 * NW, r = idx % kBlk, c = idx / kBlk;
      v[it] = masked(Lg[(size_t)(p + imin(c, w - 1)) * ldl + p + imin(r, w - 1)],
                     r >= c && r < w);
    }
#pragma unroll
    for (int it = 0; it < kBlk * kBlk / (64 * NW); ++it) {
      const int idx = tid + it * 64 * NW, r = idx % kBlk, c = idx / kBlk;
      if (r >= c) {
        t[r][c] = v[it];
        t[c][r] = v[it];
        if (r < w) {
          dbak[r + c * kBlk] = v[it];
          dbak[c + r * kBlk] = v[it];
        }
      }
    }
  }
  if (tid < kBlk) N.colmax[tid] = 0ull;
  __syncthreads();
  double a[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = t[r0 + r][lane];
  __syncthreads();
  unsigned long long elim = (w < 64) ? (~0ull << w) : 0ull;  // eliminated / absent
  int k = 0, par = 0;
  // every remaining index becomes a zero pivot, in increasing index order;
  // code -6: a non-positive pivot of the positive-definite factorization
  auto zero_rest = [&](int code) {
    int kk = k;
    for (int i = 0; i < w; ++i) {
      if ((elim >> i) & 1ull) continue;
      if (wv == 0) t[kk][lane] = 0.0;
      if (tid == 0) {
        lp[kk] = i;
        kind[kk] = 0;
        dinv[2 * kk] = 0.0;
        dinv[2 * kk + 1] = 0.0;
      }
      ++kk;
    }
    if (tid == 0) {
      if (code == -6) {
        N.flag = -6;
      } else {
        N.nzero += w - k;
        if (!fp.action) N.flag = -5;
      }
    }
    k = w;
  };
  int cpar = 0;
  while (k < w) {
    // parity-buffered: a step with one barrier (positive definite) must not
    // overwrite the columns a slower wave still reads from the step before
    double (*cs)[kBlk] = cs2[cpar];
    cpar ^= 1;
    int bc, brow;
    if (fp.posdef) {
      // Cholesky order: the lowest remaining index, always 1x1
      bc = brow = __ffsll((long long)~elim) - 1;
    } else {
      const bool live = !((elim >> lane) & 1ull);
      // ---- search this lane's live rows
      double bv = 0.0;
      int br = 64;
#pragma unroll
      for (int r = 0; r < R; ++r) {  // NaN never wins; all-zero keeps br = 64
        const double av = fabs(a[r]);
        const bool take = r0 + r >= lane && !((elim >> (r0 + r)) & 1ull) && av > bv;
        bv = take ? av : bv;
        br = take ? r0 + r : br;
      }
      // ---- wave candidate: value, then smallest column, then its row
      const unsigned long long key = live ? (unsigned long long)__double_as_longlong(bv) : 0ull;
      const unsigned long long kw = wave_max_u64(key);
      const unsigned long long hit = __ballot(live && key == kw);
      const int cw = hit ? __ffsll((long long)hit) - 1 : 64;
      const int rw = hit ? __builtin_amdgcn_readlane(br, cw) : 64;
      if (lane == 0) {
        ckey[par][wv] = kw;
        ccr[par][wv] = (cw << 7) | rw;
      }
      __syncthreads();
      // largest value; among the waves holding it, smallest column, then row
      unsigned long long kv[NW];
      int cr[NW];
#pragma unroll
      for (int v = 0; v < NW; ++v) {
        kv[v] = ckey[par][v];
        cr[v] = ccr[par][v];
      }
      unsigned long long kmax = kv[0];
#pragma unroll
      for (int v = 1; v < NW; ++v) kmax = (kv[v] > kmax) ? kv[v] : kmax;
      int best = 1 << 30;  // above every encoding; kmax is some wave's key
#pragma unroll
      for (int v = 0; v < NW; ++v) best = (kv[v] == kmax && cr[v] < best) ? cr[v] : best;
      kmax = ((unsigned long long)(unsigned)uni((int)(kmax >> 32)) << 32) |
             (unsigned)uni((int)(kmax & 0xffffffffull));
      best = uni(best);
      bc = best >> 7;
      brow = best & 127;
      par ^= 1;
      const double vmax = __longlong_as_double((long long)kmax);
      if (!(vmax >= fp.small) || bc >= 64 || brow >= 64) {
        zero_rest(-5);
        break;
      }
    }
    // ---- owners broadcast their columns: cs[0] <- column bc, cs[1] <- column brow
    if (lane == bc || lane == brow) {
      double *dst = cs[lane == bc ? 0 : 1] + r0;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r] = a[r];
    }
    __syncthreads();
    if (fp.posdef && !(uni(cs[0][bc]) > 0.0)) {
      zero_rest(-6);
      break;
    }
    int size, p1, p2 = -1;
    const double *c1p, *c2p = cs[1];
    double d11 = 0.0, d21 = 0.0, d22 = 0.0;
    if (brow == bc) {
      size = 1;
      p1 = bc;
      c1p = cs[0];
    } else {
      const double a11 = uni(cs[0][bc]), a22 = uni(cs[1][brow]), a21 = uni(cs[0][brow]);
      // explicit fma throughout the block (no contraction left to the
      // compiler), so the oracle's restatement can round the same way
      const double detscale = 1.0 / fabs(a21);
      const double detpiv = fma(a11 * detscale, a22, -fabs(a21));
      if (fabs(detpiv) >= fabs(a21) / 2) {
        size = 2;
        p1 = bc;
        p2 = brow;
        c1p = cs[0];
        d11 = (a22 * detscale) / detpiv;
        d21 = (-a21 * detscale) / detpiv;
        d22 = (a11 * detscale) / detpiv;
      } else {
        size = 1;
        const bool first = fabs(a11) > fabs(a22);
        p1 = first ? bc : brow;
        c1p = first ? cs[0] : cs[1];
      }
    }
    if (size == 1) {
      const double di = 1.0 / c1p[p1];
      elim |= 1ull << p1;
      const double lj = c1p[lane] * di;  // L(lane, k)
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = fma(-c1p[r0 + r], lj, a[r]);
      if (wv == 0) t[k][lane] = lj;
      if (tid == 0) {
        dinv[2 * k] = di;
        dinv[2 * k + 1] = 0.0;
        lp[k] = p1;
        kind[k] = 1;
      }
    } else {
      elim |= (1ull << p1) | (1ull << p2);
      const double x = c1p[lane], y = c2p[lane];
      const double l1 = fma(d11, x, d21 * y), l2 = fma(d21, x, d22 * y);
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = fma(-c1p[r0 + r], l1, fma(-c2p[r0 + r], l2, a[r]));
      if (wv == 0) {
        t[k][lane] = l1;
        t[k + 1][lane] = l2;
      }
      if (tid == 0) {
        dinv[2 * k] = d11;
        dinv[2 * k + 1] = d21;
        dinv[2 * k + 2] = INFINITY;
        dinv[2 * k + 3] = d22;
        lp[k] = p1;
        lp[k + 1] = p2;
        kind[k] = 2;
        kind[k + 1] = 3;
      }
    }
    k += size;
  }
  __syncthreads();
  // ---- write back L_bb (rows permuted into elimination order), D^{-1}, lp
  for (int i = tid; i < w; i += 64 * NW) inv_lp[lp[i]] = i;
  __syncthreads();
  for (int idx = tid; idx < w * w; idx += 64 * NW) {
    const int r = idx % w, kk = idx / w;  // original row r, elimination column kk
    const int nr = inv_lp[r];
    if (nr < kk) continue;
    double v;
    if (nr == kk) v = (kind[kk] == 0) ? 0.0 : 1.0;
    else if (nr == kk + 1 && kind[kk] == 2) v = 0.0;
    else v = t[kk][r];
    L[(size_t)(p + kk) * ldl + p + nr] = v;
  }
  for (int j = tid; j < w; j += 64 * NW) {
    N.D[2 * (p + j)] = dinv[2 * j];
    N.D[2 * (p + j) + 1] = dinv[2 * j + 1];
    N.lp[j] = lp[j];
  }
  // ---- M = L_bb^{-1} in elimination order for k_apply
  // (L21 = A21[:, lp] M^T D^{-1}).  mi[k][i] = L_bb[i][k] (zero outside the
  // strict lower w x w part), then forward substitution by columns: once
  // m[k] is final it updates every later m[i]
  for (int idx = tid; idx < kBlk * kBlk; idx += 64 * NW) {
    const int i = idx & (kBlk - 1), k = idx / kBlk;
    double v = 0.0;
    if (i > k && i < w && !(kind[k] == 2 && i == k + 1)) v = t[k][lp[i]];
    mi[k][i] = v;
  }
  __syncthreads();
  {
    // wave v: columns [Cv, Cv+C) (C = 64/NW); NW lanes per column: lane =
    // NW * (column in wave) + phase, rows phase, phase+NW, ... in registers
    constexpr int C = kBlk / NW;
    static_assert((NW & (NW - 1)) == 0 && C * NW == kBlk, "NW: a power of two");
    const int col = wv * C + lane / NW, ph = lane % NW;
    double mm[C];
#pragma unroll
    for (int q = 0; q < C; ++q) mm[q] = (ph + NW * q == col) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < kBlk - 1; ++k) {
      const double mk = __shfl(mm[k / NW], (lane & ~(NW - 1)) | (k % NW));
#pragma unroll
      for (int q = 0; q < C; ++q)
        if (ph + NW * q > k) mm[q] -= mi[k][ph + NW * q] * mk;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) N.W[(ph + NW * q) * kBlk + col] = mm[q];  // row-major
  }
  // permute the variable list (rows of earlier eliminated columns: k_apply)
  const int pv = (tid < w) ? N.perm[p + lp[tid]] : 0;
  __syncthreads();
  if (tid < w) N.perm[p + tid] = pv;
}

// ---------------------------------------------------------------------------
// Apply the block pivots to the rows below and record column maxima
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_apply(NodeDev *nodes, const int *lvl_slots,
                                                    const int *wg_off, int nlvl, int step,
                                                    FactorParams fp) {
  // M = L_bb^{-1} (or the left-rows chunk), then, once the MFMAs have read
  // it, X = A21[:, lp] M^T for this workgroup's 64 rows (one tile of LDS:
  // four workgroups fit a CU)
  __shared__ double ms[kBlk][kBlk + 1];
  double (*xs)[kBlk + 1] = ms;
  __shared__ double dv[2 * kBlk + 2];
  __shared__ int lps[kBlk];
  const int li = find_node(wg_off, nlvl, blockIdx.x);
  if (li < 0) return;
  NodeDev &N = nodes[lvl_slots[li]];
  if (step >= N.nsteps) return;
  const int p = N.p, w = N.w;
  if (w == 0) return;
  int local = blockIdx.x - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ldl = N.ldl;
  double *L = N.L;
  // the first nleft workgroups permute rows p..p+w-1 of the eliminated
  // columns left of p into the block's elimination order (64-column chunks)
  const int nleft = step;  // upper bound on ceil(p / 64)
  if (local < nleft) {
    const int cb = local * kBlk;
    if (cb >= p) return;
    for (int idx = tid; idx < kBlk * kBlk; idx += kThreads) {
      int r = idx % kBlk, cc = idx / kBlk;
      if (r < w && cb + cc < p) ms[r][cc] = L[(size_t)(cb + cc) * ldl + p + r];
    }
    __syncthreads();
    for (int idx = tid; idx < kBlk * kBlk; idx += kThreads) {
      int r = idx % kBlk, cc = idx / kBlk;
      if (r < w && cb + cc < p) L[(size_t)(cb + cc) * ldl + p + r] = ms[N.lp[r]][cc];
    }
    return;
  }
  local -= nleft;
  // ---- rows [rb, rb+64) below the block: X = A21[:, lp] M^T on MFMA
  const int r0 = p + w, m = N.m;
  const int rb = r0 + local * kBlk;
  for (int idx = tid; idx < kBlk * kBlk; idx += kThreads)
    ms[idx / kBlk][idx & (kBlk - 1)] = N.W[idx];
  for (int j = tid; j < 2 * w; j += kThreads) dv[j] = N.D[2 * p + j];
  if (tid == 0) {
    dv[2 * w] = 0.0;
    dv[2 * w + 1] = 0.0;
  }
  for (int j = tid; j < kBlk; j += kThreads) lps[j] = (j < w) ? N.lp[j] : 0;
  __syncthreads();
  const int ri = lane & 15, kq = lane >> 4;
  const int r = rb + wave * 16 + ri;  // A row of this lane (fragment layout)
  const bool rok = r < m;
  double av[kBlk / 4];
  {  // unconditional loads (row and pivot clamped into range), unused masked
    const int rc = imin(r, m - 1);
    gcdouble *Lg = (gcdouble *)L;
#pragma unroll
    for (int u = 0; u < kBlk / 4; ++u) {
      const int k = 4 * u + kq;
      av[u] = masked(Lg[(size_t)(p + lps[imin(k, w - 1)]) * ldl + rc], rok && k < w);
    }
  }
  // descriptor fields once: the global stores below could otherwise alias
  // them and force a reload per iteration
  gdouble *bak = (gdouble *)N.bak;
  gdouble *LD = (gdouble *)N.LD;
  gdouble *Lw = (gdouble *)L;
  unsigned long long *colmax = N.colmax;
  const int ldbak = N.ldbak, nfs = N.n;
#pragma unroll
  for (int u = 0; u < kBlk / 4; ++u) {
    const int k = 4 * u + kq;
    if (rok && k < w) bak[(size_t)k * ldbak + (r - r0)] = av[u];
  }
  double4_t acc[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) acc[g] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < kBlk / 4; ++u) {
#pragma unroll
    for (int g = 0; g < 4; ++g)  // B[k][j] = M[j][k]
      acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], ms[g * 16 + ri][4 * u + kq], acc[g],
                                                    0, 0, 0);
  }
  __syncthreads();  // every wave has read M before X overwrites it
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) xs[wave * 16 + 4 * q + kq][g * 16 + ri] = acc[g][q];
  __syncthreads();
  // ---- x D^{-1}, store, column maxima (thread: row `row`, columns cb0 + 4u)
  const int row = lane, cb0 = wave;
  const int rr = rb + row;
  const bool ok = rr < m;
  const double small = fp.small;
#pragma unroll
  for (int u = 0; u < kBlk / 4; ++u) {
    const int j = cb0 + 4 * u;
    double y = 0.0;
    if (j < w) {
      const double xj = xs[row][j];
      const bool two_first = (j + 1 < w) && isinf(dv[2 * j + 2]);
      const bool two_second = (j > 0) && isinf(dv[2 * j]);
      bool zero_piv = false;
      if (two_first) {
        y = dv[2 * j] * xj + dv[2 * j + 1] * xs[row][j + 1];
      } else if (two_second) {
        y = dv[2 * j - 1] * xs[row][j - 1] + dv[2 * j + 1] * xj;
      } else {
        const double d11 = dv[2 * j];
        zero_piv = (d11 == 0.0);
        if (zero_piv) y = (fabs(xj) < small) ? 0.0 : INFINITY * xj;
        else y = xj * d11;
      }
      if (ok) {
        Lw[(size_t)(p + j) * ldl + rr] = y;
        // L21 D is the product before D^{-1} (zero for a zero pivot): the
        // parent's gather operand, for the contribution rows
        if (LD && rr >= nfs) LD[(size_t)(p + j) * ldl + rr] = zero_piv ? 0.0 : xj;
      }
    }
    // column maximum over this wave's 64 rows (NaN propagates as a failure)
    double v = (ok && j < w) ? fabs(y) : 0.0;
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off);
      v = (o > v || isnan(o)) ? o : v;
    }
    if (lane == 0 && j < w) {
      const unsigned long long bits = isnan(v) ? 0x7ff0000000000001ull
                                               : (unsigned long long)__double_as_longlong(v);
      atomicMax(&colmax[j], bits);
    }
  }
}

// number of block pivots that pass the threshold test (split 2x2 avoided):
// the first column whose max |l| exceeds 1/u.  Lane j reads column j's
// maximum, so the w loads are in flight together; every lane of the wave
// must be active, and the result is wave-uniform.
__device__ inline int block_npass(const NodeDev &N, double u) {
  const int w = N.w, p = N.p, lane = threadIdx.x & 63;
  const double lim = 1.0 / u;
  const unsigned long long lb = (unsigned long long)__double_as_longlong(lim);
  const unsigned long long b =
      (lane < w) ? __atomic_load_n(&N.colmax[lane], __ATOMIC_RELAXED) : 0ull;
  const unsigned long long fail = __ballot(lane < w && b > lb);
  int np = fail ? __ffsll((long long)fail) - 1 : w;
  if (np > 0 && np < w) {
    double d11 = N.D[2 * (p + np - 1)], d21 = N.D[2 * (p + np - 1) + 1];
    if (isfinite(d11) && d21 != 0.0) np--;
  }
  return np;
}

// D for pivot k (from the stored D^{-1}); two: 0 = 1x1, 1 = first of 2x2,
// 2 = second of 2x2.
__device__ inline void pivot_d(const double *D, int k, int kend, double &e1, double &e2,
                               int &two) {
  double d11 = D[2 * k];
  if (!isfinite(d11)) {  // second of a 2x2 block
    double a11 = D[2 * (k - 1)], a21 = D[2 * (k - 1) + 1], a22 = D[2 * k + 1];
    double det = a11 * a22 - a21 * a21;
    e1 = -a21 / det;  // off-diagonal of D block
    e2 = a11 / det;   // D22
    two = 2;
  } else if (k + 1 < kend && !isfinite(D[2 * k + 2])) {
    double a11 = d11, a21 = D[2 * k + 1], a22 = D[2 * k + 3];
    double det = a11 * a22 - a21 * a21;
    e1 = a22 / det;   // D11
    e2 = -a21 / det;  // off-diagonal
    two = 1;
  } else {
    e1 = (d11 != 0.0) ? 1.0 / d11 : 0.0;
    e2 = 0.0;
    two = 0;
  }
}

// ---------------------------------------------------------------------------
// Panel update with MFMA:  F[i,j] -= sum_k L[i,a0+k] (L D)[j,a0+k]
// mode 0 (APP step): accepted pivots [p, p+npass) applied to the panel's
//   trailing columns [t0, n), rows [t0, m); failed block columns are restored
//   from the backup first.  The contribution block is NOT touched here: its
//   Schur update is applied lazily when the parent gathers it (k_gather).
// mode 1: subtree roots with a contribution block: all nelim pivots applied
//   to [n, m)^2 so the block can be exported.
// ---------------------------------------------------------------------------
constexpr int kTile = 64;

__device__ inline void tile_of(int tile, int TI, int TJ, int &I, int &J) {
  // lower tiles (I >= J), column-major over J
  int t = tile;
  J = 0;
  while (J < TJ && t >= TI - J) {
    t -= TI - J;
    ++J;
  }
  I = J + t;
}

__global__ __launch_bounds__(kThreads, 4) void k_update(NodeDev *nodes, const int *lvl_slots,
                                                     const int *wg_off, int nlvl, int step,
                                                     int mode, FactorParams fp) {
  // 32-pivot K chunks: the two operand tiles take 34 KB, four workgroups a CU
  constexpr int kUpdK = 32;
  __shared__ double As[kTile][kUpdK + 1];
  __shared__ double Bs[kTile][kUpdK + 1];
  const int li = find_node(wg_off, nlvl, blockIdx.x);
  if (li < 0) return;
  NodeDev &N = nodes[lvl_slots[li]];
  const int local = blockIdx.x - wg_off[li], count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int a0, K, t0, cend, fa, fb, p = 0, w = 0;
  if (mode == 0) {
    if (step >= N.nsteps) return;
    p = N.p;
    w = N.w;
    if (w == 0) return;
    int np = fp.posdef ? N.w : block_npass(N, fp.u);
    if (local == 0 && tid == 0) N.npass = np;
    a0 = p;
    K = np;
    t0 = p + np;
    cend = N.n;
    fa = p + np;
    fb = p + w;
  } else {
    if (!N.is_root || N.m <= N.n) return;
    a0 = 0;
    K = N.nelim;
    t0 = N.n;
    cend = N.m;
    fa = fb = 0;
    if (K <= 0) return;
  }
  const int m = N.m;
  if (m - t0 <= 0 || cend - t0 <= 0) return;
  const int TI = (m - t0 + kTile - 1) / kTile, TJ = (cend - t0 + kTile - 1) / kTile;
  const int ntiles = TJ * TI - TJ * (TJ - 1) / 2;
  const double *D = N.D;
  const int ldl = N.ldl;
  const int Kc = kUpdK;  // K chunk
  // descriptor fields once (the stores below could alias them)
  gdouble *NL = (gdouble *)N.L;
  gdouble *NC = (gdouble *)N.C;
  float __attribute__((address_space(1))) *NCf = (float __attribute__((address_space(1))) *)N.C;
  const bool c32 = fp.c32;  // the contribution block (mode 1 only) is FP32
  gcdouble *dbak = (gcdouble *)N.dbak;
  gcdouble *bak = (gcdouble *)N.bak;
  const int *lpv = N.lp;
  const int nfs = N.n, ldc = N.ldc, ldbak = N.ldbak;
  const int gmax = imin(m, cend) - 1;  // last row a B operand reads
  for (int tile = local; tile < ntiles; tile += count) {
    int I, J;
    tile_of(tile, TI, TJ, I, J);
    const int r0 = t0 + I * kTile, c0 = t0 + J * kTile;
    const bool has_fail = (fb > fa) && (c0 < fb) && (c0 + kTile > fa);
    if (K == 0 && !has_fail) continue;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    double4_t acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += Kc) {
      const int kn = min(Kc, K - k0);
      const int Kp = (kn + 3) & ~3;
      __syncthreads();
      // every thread's loads unconditional (pivot and rows clamped, a 2x2
      // partner always inside the pivots), the entries past the chunk or the
      // front masked to zero.  B = L D keeps the expression of each pivot
      // kind as it is (its rounding is part of the result).
#pragma unroll
      for (int it = 0; it < kTile * kUpdK / kThreads; ++it) {
        const int idx = tid + it * kThreads, i = idx % kTile, kk = idx / kTile;
        const int kc = imin(kk, kn - 1), k = a0 + k0 + kc;
        const int gi = r0 + i, gj = c0 + i, gjc = imin(gj, gmax);
        const double av = masked(NL[(size_t)k * ldl + imin(gi, m - 1)], kk < kn && gi < m);
        double e1, e2;
        int two;
        pivot_d(D, k, a0 + K, e1, e2, two);
        const int ofs = (two == 1) ? 1 : (two == 2) ? -1 : 0;
        const double lk = NL[(size_t)k * ldl + gjc];
        const double lo = NL[(size_t)(k + ofs) * ldl + gjc];
        double bv = (two == 0) ? lk * e1 : (two == 1) ? lk * e1 + lo * e2 : lo * e1 + lk * e2;
        As[i][kk] = av;
        Bs[i][kk] = masked(bv, kk < kn && gj <= gmax);
      }
      __syncthreads();
      for (int kk = 0; kk < Kp; kk += 4) {
        const int kl = kk + (lane >> 4);
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
          double av = As[wr + bi * 16 + (lane & 15)][kl];
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            double bv = Bs[wc + bj * 16 + (lane & 15)][kl];
            acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[bi][bj], 0, 0, 0);
          }
        }
      }
    }
    // element (i,j) of a 16x16 block lives in lane j + 16*(i%4), register i/4
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rg = 0; rg < 4; ++rg) {
          int i = wr + bi * 16 + (lane >> 4) + 4 * rg;
          int j = wc + bj * 16 + (lane & 15);
          int gi = r0 + i, gj = c0 + j;
          if (gi < m && gj < cend && gi >= gj && gj >= nfs && c32) {
            float __attribute__((address_space(1))) *dc =
                NCf + (size_t)(gj - nfs) * ldc + (gi - nfs);
            *dc = (float)((double)*dc - acc[bi][bj][rg]);
          } else if (gi < m && gj < cend && gi >= gj) {
            gdouble *dst = (gj < nfs) ? NL + (size_t)gj * ldl + gi
                                      : NC + (size_t)(gj - nfs) * ldc + (gi - nfs);
            double old;
            if (gj >= fa && gj < fb) {
              int jj = gj - p;
              if (gi < p + w) old = dbak[lpv[gi - p] + lpv[jj] * kBlk];
              else old = bak[(size_t)jj * ldbak + (gi - p - w)];
            } else {
              old = *dst;
            }
            *dst = old - acc[bi][bj][rg];
          }
        }
  }
}

// ---------------------------------------------------------------------------
// The APP step's trailing panel update (k_update mode 0) with the MFMA
// operands loaded straight into the 16x16x4 fragment layout: lane i + 16 kq
// holds row i of A (= L, pivot kq of the 4-pivot step) and column i of
// B = L D for the same pivot; two register buffers of two steps each
// alternate, so the loads of the next 8 pivots are in flight during the
// MFMAs.  The per-pivot coefficients of L D come from pivot_d once per
// workgroup.  Each element gets the same products in the same order as in
// k_update.  Three workgroups per SIMD fit (162 VGPRs): loading the tile's
// current values before the products as well needs 198 registers and two
// workgroups, and measured slower (64^3 factor 47.4 against 46.7 ms, the
// LDS-staged k_update 48.1 ms).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads, 3) void k_update_panel(NodeDev *nodes,
                                                           const int *lvl_slots,
                                                           const int *wg_off, int nlvl,
                                                           int step, FactorParams fp) {
  __shared__ double ce1[kBlk], ce2[kBlk];
  __shared__ int ctwo[kBlk];
  const int li = find_node(wg_off, nlvl, blockIdx.x);
  if (li < 0) return;
  NodeDev &N = nodes[lvl_slots[li]];
  const int local = blockIdx.x - wg_off[li], count = wg_off[li + 1] - wg_off[li];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (step >= N.nsteps) return;
  const int p = N.p, w = N.w;
  if (w == 0) return;
  const int np = fp.posdef ? w : block_npass(N, fp.u);
  if (local == 0 && tid == 0) N.npass = np;
  const int a0 = p, K = np, t0 = p + np, cend = N.n, fa = p + np, fb = p + w;
  const int m = N.m;
  if (m - t0 <= 0 || cend - t0 <= 0) return;
  const int TI = (m - t0 + kTile - 1) / kTile, TJ = (cend - t0 + kTile - 1) / kTile;
  const int ntiles = TJ * TI - TJ * (TJ - 1) / 2;
  {
    const double *D = N.D;
    for (int t = tid; t < K; t += kThreads) {
      double e1, e2;
      int two;
      pivot_d(D, a0 + t, a0 + K, e1, e2, two);
      ce1[t] = e1;
      ce2[t] = e2;
      ctwo[t] = two;
    }
  }
  __syncthreads();
  const int ldl = N.ldl;
  gdouble *NL = (gdouble *)N.L;
  gcdouble *dbak = (gcdouble *)N.dbak;
  gcdouble *bak = (gcdouble *)N.bak;
  const int *lpv = N.lp;
  const int ldbak = N.ldbak;
  const int gmax = imin(m, cend) - 1;  // last row a B operand reads
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ri = lane & 15, kq = lane >> 4;
  constexpr int KC = 2;  // 4-pivot steps per register buffer
  for (int tile = local; tile < ntiles; tile += count) {
    int I, J;
    tile_of(tile, TI, TJ, I, J);
    const int r0 = t0 + I * kTile, c0 = t0 + J * kTile;
    const bool has_fail = (fb > fa) && (c0 < fb) && (c0 + kTile > fa);
    if (K == 0 && !has_fail) continue;
    double4_t acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = (double4_t){0.0, 0.0, 0.0, 0.0};
    const int ra0 = r0 + wr + ri, ra1 = ra0 + 16, cb0 = c0 + wc + ri, cb1 = cb0 + 16;
    const int ra0c = imin(ra0, m - 1), ra1c = imin(ra1, m - 1);
    const int cb0c = imin(cb0, gmax), cb1c = imin(cb1, gmax);
    // raw operands of KC steps: A rows 0/1, L of B columns 0/1 (pivot and
    // its 2x2 partner)
    double xa0[KC], xa1[KC], xk0[KC], xo0[KC], xk1[KC], xo1[KC];
    double ya0[KC], ya1[KC], yk0[KC], yo0[KC], yk1[KC], yo1[KC];
    auto load = [&](int T0, double *A0, double *A1, double *K0, double *O0, double *K1,
                    double *O1) {
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int kc = imin(T0 + 4 * u + kq, K - 1);
        const int two = ctwo[kc];
        const int ofs = (two == 1) ? 1 : (two == 2) ? -1 : 0;
        const size_t ok = (size_t)(a0 + kc) * ldl, oo = (size_t)(a0 + kc + ofs) * ldl;
        A0[u] = NL[ok + ra0c];
        A1[u] = NL[ok + ra1c];
        K0[u] = NL[ok + cb0c];
        O0[u] = NL[oo + cb0c];
        K1[u] = NL[ok + cb1c];
        O1[u] = NL[oo + cb1c];
      }
    };
    auto mfma = [&](int T0, const double *A0, const double *A1, const double *K0,
                    const double *O0, const double *K1, const double *O1) {
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        if (T0 + 4 * u >= K) break;
        const int kk = T0 + 4 * u + kq;
        const bool kok = kk < K;
        const int kc = imin(kk, K - 1);
        const double e1 = ce1[kc], e2 = ce2[kc];
        const int two = ctwo[kc];
        const double b0 = (two == 0) ? K0[u] * e1
                          : (two == 1) ? K0[u] * e1 + O0[u] * e2 : O0[u] * e1 + K0[u] * e2;
        const double b1 = (two == 0) ? K1[u] * e1
                          : (two == 1) ? K1[u] * e1 + O1[u] * e2 : O1[u] * e1 + K1[u] * e2;
        const double av0 = masked(A0[u], kok && ra0 < m), av1 = masked(A1[u], kok && ra1 < m);
        const double bv0 = masked(b0, kok && cb0 <= gmax), bv1 = masked(b1, kok && cb1 <= gmax);
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv1, acc[1][1], 0, 0, 0);
      }
    };
    if (K > 0) {
      load(0, xa0, xa1, xk0, xo0, xk1, xo1);
      load(4 * KC, ya0, ya1, yk0, yo0, yk1, yo1);
      for (int T = 0; T < K; T += 8 * KC) {
        if (T > 0) load(T + 4 * KC, ya0, ya1, yk0, yo0, yk1, yo1);
        mfma(T, xa0, xa1, xk0, xo0, xk1, xo1);
        if (T + 4 * KC >= K) break;
        load(T + 8 * KC, xa0, xa1, xk0, xo0, xk1, xo1);
        mfma(T + 4 * KC, ya0, ya1, yk0, yo0, yk1, yo1);
      }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int rg = 0; rg < 4; ++rg) {
          const int gi = r0 + wr + bi * 16 + kq + 4 * rg, gj = c0 + wc + bj * 16 + ri;
          if (gi < m && gj < cend && gi >= gj) {
            double old;
            if (gj >= fa && gj < fb) {
              const int jj = gj - p;
              old = (gi < p + w) ? dbak[lpv[gi - p] + lpv[jj] * kBlk]
                                 : bak[(size_t)jj * ldbak + (gi - p - w)];
            } else {
              old = NL[(size_t)gj * ldl + gi];
            }
            NL[(size_t)gj * ldl + gi] = old - acc[bi][bj][rg];
          }
        }
  }
}

// D (not its inverse) per pivot for the lazy contribution update:
// Dd[3k] = D_kk, Dd[3k+1] = D_k,partner, Dd[3k+2] = partner offset.
// ---------------------------------------------------------------------------
// Gather assembly of one level (replaces zero + scatter + extend-add + the
// Schur update of the children's contribution blocks).  A workgroup owns a
// 64x64 tile of one parent front and builds it in LDS:
//   for each child in order, the child's trailing index ranges that map into
//   the tile come from the k_maptab tables and their parent positions are
//   staged in LDS; failed (delayed) columns are copied from the child panel;
//   contribution columns are added as C_c (read down its columns) minus
//   L21 (D L21)^T on MFMA (K = the child's nelim; operands straight from
//   global memory in the fragment layout), both through the maps;
//   then the tile's A entries are added and the tile is stored once.
// ---------------------------------------------------------------------------
// Per level, for every child of the level's fronts: the parent position of
// each trailing index (pmap) and, for each 64-row tile of the parent, the
// first trailing index whose parent position reaches the tile (ptab).
constexpr int kMapParts = 8;  // workgroups per front (trailing indices strided over them)
__global__ __launch_bounds__(kThreads) void k_maptab(NodeDev *nodes, const int *lvl_slots,
                                                     const int *child_list) {
  const NodeDev &P = nodes[lvl_slots[blockIdx.x]];
  if (P.skip) return;
  const int part = blockIdx.y;
  const int T = (P.m + kGT - 1) / kGT;
  int doff = 0;
  for (int ci = P.cbeg; ci < P.cend; ++ci) {
    NodeDev &c = nodes[child_list[ci]];
    int *tab = P.ctab + (size_t)(ci - P.cbeg) * (T + 1);
    const int cdoff = doff;
    doff += c.n - c.nelim;
    const int nt = c.m - c.nelim;
    for (int t = part * kThreads + threadIdx.x; t < nt; t += kMapParts * kThreads) {
      const int k = c.nelim + t;
      const int pk = child_pos(c, cdoff, P.nd, k);
      c.pmap[t] = pk;
      const int tile = pk / kGT;
      const int prev = (t == 0) ? -1 : child_pos(c, cdoff, P.nd, k - 1) / kGT;
      for (int I = prev + 1; I <= tile; ++I) tab[I] = k;
      if (t == nt - 1)
        for (int I = tile + 1; I <= T; ++I) tab[I] = c.m;
    }
    if (nt == 0 && part == 0)
      for (int I = threadIdx.x; I <= T; I += kThreads) tab[I] = c.m;
  }
}

// PART 0: every tile of the level's fronts; 1: the first tile column only
// (the pivot block and the rows below it); 2: the remaining tiles
// Returns the front's slot when this workgroup built the front's first
// diagonal tile (0, 0), else -1.

// CT: element type of the contribution blocks (double, or float for the
// mixed-precision factorization, whose Schur terms then run on FP32 MFMA)
template <int PART, typename CT>
__device__ __forceinline__ int gather_tile(
    NodeDev *nodes, const int *lvl_slots, const int *wg_off, int nlvl,
    const int *child_list, const double *aval, const int64_t *asrc, const int *arow,
    const int *acix, const double *scaling, double (&acc)[kGT][kGT + 1], int (&mr)[kGT],
    int (&mc)[kGT]) {
  constexpr bool F32 = sizeof(CT) == 4;
  const int li = find_node(wg_off, nlvl, blockIdx.x);
  if (li < 0) return -1;
  const int s = lvl_slots[li];
  NodeDev &P = nodes[s];
  // descriptor fields once, into registers (loads below never wait on them)
  const int m = P.m, n = P.n, nd = P.nd, pldl = P.ldl, pldc = P.ldc;
  const int cbeg = P.cbeg, cend = P.cend;
  gdouble *PL = (gdouble *)P.L;
  CT __attribute__((address_space(1))) *PC = (CT __attribute__((address_space(1))) *)P.C;
  gcint *pctab = (gcint *)P.ctab;
  gcint *prows = (gcint *)P.rows;
  const int64_t *pacol = P.acol;
  const int T = (m + kGT - 1) / kGT;
  // PART 1: the diagonal tile (0, 0); 3: the first tile column's other
  // tiles; 2: the tiles outside the first column; 0: every tile
  const int local = blockIdx.x - wg_off[li] + (PART == 2 ? T : PART == 3 ? 1 : 0);
  if (local >= (PART == 1 ? 1 : PART == 3 ? T : T * (T + 1) / 2)) return -1;
  int I, J;
  tile_of(local, T, T, I, J);
  const int r0 = I * kGT, c0 = J * kGT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int idx = tid; idx < kGT * kGT; idx += kThreads) acc[idx % kGT][idx / kGT] = 0.0;
  // fully-summed variable list (diagonal tiles own their columns)
  if (I == J) {
    for (int j = c0 + tid; j < min(c0 + kGT, n); j += kThreads) {
      int v;
      if (j >= nd) {
        v = prows[j - nd];
      } else {
        int doff = 0;
        v = -1;
        for (int ci = cbeg; ci < cend; ++ci) {
          const NodeDev &c = nodes[child_list[ci]];
          int nf = c.n - c.nelim;
          if (j >= doff && j < doff + nf) {
            v = c.perm[c.nelim + (j - doff)];
            break;
          }
          doff += nf;
        }
      }
      P.perm[j] = v;
      P.perm0[j] = v;
    }
  }
  // this tile's A entries (entry -> column precomputed), loaded up front
  constexpr int kAE = 1;  // entries per thread held in registers
  const int cs0 = max(c0, nd) - nd, cs1 = min(c0 + kGT, n) - nd;
  int64_t ae0 = 0, ae1 = 0;
  if (cs1 > cs0) {
    ae0 = pacol[cs0];
    ae1 = pacol[cs1];
  }
  double av_[kAE];
  int ar_[kAE], ac_[kAE];
#pragma unroll
  for (int u = 0; u < kAE; ++u) {
    const int64_t e = ae0 + tid + (int64_t)u * kThreads;
    ar_[u] = -1;
    av_[u] = 0.0;
    ac_[u] = 0;
    if (e < ae1) {
      ar_[u] = arow[e] + nd;
      ac_[u] = acix[e];
      av_[u] = aval[asrc[e]];
    }
  }
  __syncthreads();
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int ri = lane & 15, kq = lane >> 4;
  for (int ci = cbeg; ci < cend; ++ci) {
    const NodeDev &c = nodes[child_list[ci]];
    const int cn = c.n, cm = c.m, kb = c.nelim, cldl = c.ldl, cldc = c.ldc;
    gcdouble *cL = (gcdouble *)c.L;
    gcdouble *cLD = (gcdouble *)c.LD;
    const CT __attribute__((address_space(1))) *cC =
        (const CT __attribute__((address_space(1))) *)c.C;
    gcint *cpmap = (gcint *)c.pmap;
    if (cm <= kb) continue;
    gcint *tab = pctab + (size_t)(ci - cbeg) * (T + 1);
    const int kr0 = tab[I], kr1 = tab[I + 1];
    const int k0 = tab[J], k1 = tab[J + 1];
    if (kr1 <= kr0 || k1 <= k0 || kr1 - 1 < k0) continue;
    for (int t = tid; t < kr1 - kr0; t += kThreads) mr[t] = cpmap[kr0 + t - kb] - r0;
    for (int t = tid; t < k1 - k0; t += kThreads) mc[t] = cpmap[k0 + t - kb] - c0;
    // thread (a, b0): row a of the tile's row range, columns b0, b0+4, ...
    // (no index division; a wave reads 64 consecutive rows of one column)
    const int ta = tid & 63, tb = tid >> 6;
    const int kc0 = max(k0, cn);
    const int rr0 = max(kr0, cn);
    const int nrow = kr1 - rr0, ncol = k1 - kc0;
    const bool contrib = (k1 > kc0 && kr1 > rr0);
    // C_c, lower part, down its columns: all loads issued before any use
    constexpr int NU = kGT / (kThreads / 64);
    double cv[NU];
    const bool rowok = contrib && ta < nrow;
    const int kr = rr0 + ta;
    const CT __attribute__((address_space(1))) *Cr = cC + (kr - cn);
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int b = tb + u * (kThreads / 64), k = kc0 + b;
      cv[u] = (rowok && b < ncol && kr >= k) ? (double)Cr[(size_t)(k - cn) * cldc] : 0.0;
    }
    // L21 (L21 D)^T on MFMA: each wave a 32x32 quadrant, operands in the
    // 16x16x4 fragment layout (lane i+16k holds row i, pivot k) straight
    // from L21 and the child's L21 D; two register buffers alternate so
    // the loads of the next 16 pivots are in flight during the MFMAs
    const int K = kb;
    const int ra0 = wr + ri, ra1 = ra0 + 16, cb0 = wc + ri, cb1 = cb0 + 16;
    const bool va0 = ra0 < nrow, va1 = ra1 < nrow, vb0 = cb0 < ncol, vb1 = cb1 < ncol;
    gcdouble *La = cL + rr0;
    gcdouble *Lb = cLD + kc0;
    // FP64: v_mfma_f64_16x16x4_f64, D[i][j] in lane j + 16 (i % 4), register
    // i / 4.  Mixed precision: operands rounded to FP32, v_mfma_f32_16x16x4_f32
    // with FP32 accumulation, D[i][j] in lane j + 16 (i / 4), register i % 4
    typedef typename std::conditional<F32, float4_t, double4_t>::type acc4_t;
    acc4_t g[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) g[bi][bj] = (acc4_t){0, 0, 0, 0};
    constexpr int KC = 2;  // k-steps (of 4 pivots) per buffer
    double xa0[KC], xa1[KC], xb0[KC], xb1[KC], ya0[KC], ya1[KC], yb0[KC], yb1[KC];
    auto load = [&](int T0, double *A0, double *A1, double *B0, double *B1) {
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        const int t = T0 + 4 * u + kq;
        const bool ok = t < K;
        const size_t o = (size_t)t * cldl;
        A0[u] = (ok && va0) ? La[o + ra0] : 0.0;
        A1[u] = (ok && va1) ? La[o + ra1] : 0.0;
        B0[u] = (ok && vb0) ? Lb[o + cb0] : 0.0;
        B1[u] = (ok && vb1) ? Lb[o + cb1] : 0.0;
      }
    };
    auto mfma = [&](int T0, const double *A0, const double *A1, const double *B0,
                    const double *B1) {
#pragma unroll
      for (int u = 0; u < KC; ++u) {
        if (T0 + 4 * u >= K) break;
        if constexpr (F32) {
          const float a0 = (float)A0[u], a1 = (float)A1[u], b0 = (float)B0[u],
                      b1 = (float)B1[u];
          g[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, g[0][0], 0, 0, 0);
          g[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, g[0][1], 0, 0, 0);
          g[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, g[1][0], 0, 0, 0);
          g[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, g[1][1], 0, 0, 0);
        } else {
          g[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(A0[u], B0[u], g[0][0], 0, 0, 0);
          g[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(A0[u], B1[u], g[0][1], 0, 0, 0);
          g[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(A1[u], B0[u], g[1][0], 0, 0, 0);
          g[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(A1[u], B1[u], g[1][1], 0, 0, 0);
        }
      }
    };
    // the first two operand buffers go out with the C_c loads, before the
    // barrier (they depend on the tables only)
    if (contrib) {
      load(0, xa0, xa1, xb0, xb1);
      load(4 * KC, ya0, ya1, yb0, yb1);
    }
    __syncthreads();  // maps ready
    // failed (delayed) columns of the child: copy from its panel
    const int kf1 = min(k1, cn);
    if (kf1 > k0) {
      const int nr = kr1 - kr0, ncf = kf1 - k0;
      if (ta < nr) {
        const int mra = mr[ta], krr = kr0 + ta;
        for (int b = tb; b < ncf; b += kThreads / 64) {
          const int k = k0 + b;
          if (krr >= k) acc[mra][mc[b]] += cL[(size_t)k * cldl + krr];
        }
      }
    }
    if (contrib) {
      if (rowok) {
        const int mra = mr[rr0 - kr0 + ta];
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const int b = tb + u * (kThreads / 64);
          if (b < ncol && kr >= kc0 + b) acc[mra][mc[kc0 - k0 + b]] += cv[u];
        }
      }
      for (int t0 = 0; t0 < K; t0 += 8 * KC) {
        if (t0 > 0) load(t0 + 4 * KC, ya0, ya1, yb0, yb1);
        mfma(t0, xa0, xa1, xb0, xb1);
        if (t0 + 4 * KC >= K) break;
        load(t0 + 8 * KC, xa0, xa1, xb0, xb1);
        mfma(t0 + 4 * KC, ya0, ya1, yb0, yb1);
      }
      __syncthreads();  // the C_c pass has finished with acc
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int rg = 0; rg < 4; ++rg) {
            const int i = wr + bi * 16 + (F32 ? 4 * kq + rg : kq + 4 * rg);
            const int j = wc + bj * 16 + ri;
            if (i < nrow && j < ncol) {
              const int krr = rr0 + i, k = kc0 + j;
              if (krr >= k) acc[mr[krr - kr0]][mc[k - k0]] -= (double)g[bi][bj][rg];
            }
          }
    }
    __syncthreads();
  }
  // A entries
#pragma unroll
  for (int u = 0; u < kAE; ++u) {
    const int r = ar_[u];
    if (r < r0 || r >= r0 + kGT) continue;
    const int cc = ac_[u];
    double v = av_[u];
    if (scaling) v *= scaling[prows[r - nd]] * scaling[prows[cc]];
    acc[r - r0][cc + nd - c0] += v;
  }
  for (int64_t e = ae0 + tid + (int64_t)kAE * kThreads; e < ae1; e += kThreads) {
    const int r = arow[e] + nd;
    if (r < r0 || r >= r0 + kGT) continue;
    const int cc = acix[e];
    double v = aval[asrc[e]];
    if (scaling) v *= scaling[prows[r - nd]] * scaling[prows[cc]];
    acc[r - r0][cc + nd - c0] += v;
  }
  __syncthreads();
  // store the tile once: panel columns to L, contribution columns to C
  for (int idx = tid; idx < kGT * kGT; idx += kThreads) {
    const int i = idx % kGT, j = idx / kGT;
    const int gr = r0 + i, gc = c0 + j;
    if (gr < m && gc < m && gr >= gc) {
      if (gc < n) PL[(size_t)gc * pldl + gr] = acc[i][j];
      else PC[(size_t)(gc - n) * pldc + (gr - n)] = (CT)acc[i][j];
    }
  }
  return (I == 0 && J == 0) ? s : -1;
}

template <int PART, typename CT>
__global__ __launch_bounds__(kThreads, 4) void k_gather(
    NodeDev *nodes, const int *lvl_slots, const int *wg_off, int nlvl,
    const int *child_list, const double *aval, const int64_t *asrc, const int *arow,
    const int *acix, const double *scaling) {
  __shared__ double acc[kGT][kGT + 1];
  __shared__ int mr[kGT], mc[kGT];
  gather_tile<PART, CT>(nodes, lvl_slots, wg_off, nlvl, child_list, aval, asrc, arow, acix,
                        scaling, acc, mr, mc);
}


// ---------------------------------------------------------------------------
// Move failed columns of this step behind the unprocessed candidates:
// disjoint interchanges (p+npass+i) <-> (q-ns+i), i < ns = min(nf, nu).
// Two passes through a gather buffer so the permutation is applied in place.
// ---------------------------------------------------------------------------
__device__ inline int swp_count(const NodeDev &N) {
  int nf = N.w - N.npass;
  int nu = N.q - (N.p + N.w);
  if (nf <= 0 || nu <= 0) return 0;
  return min(nf, nu);
}

__device__ inline int swp_pi(const NodeDev &N, int ns, int x) {
  int f0 = N.p + N.npass, u0 = N.q - ns;
  if (x >= f0 && x < f0 + ns) return u0 + (x - f0);
  if (x >= u0 && x < u0 + ns) return f0 + (x - u0);
  return x;
}

__device__ inline double *swp_at(const NodeDev &N, int a, int b) {
  int hi = max(a, b), lo = min(a, b);
  return N.L + (size_t)lo * N.ldl + hi;
}

__global__ __launch_bounds__(kThreads) void k_swap(NodeDev *nodes, const int *lvl_slots,
                                                   const int *wg_off, int nlvl, int step,
                                                   int phase) {
  const int li = find_node(wg_off, nlvl, blockIdx.x);
  if (li < 0) return;
  NodeDev &N = nodes[lvl_slots[li]];
  if (step >= N.nsteps || N.w == 0) return;
  const int ns = swp_count(N);
  if (ns == 0) return;
  const int local = blockIdx.x - wg_off[li], count = wg_off[li + 1] - wg_off[li];
  const int m = N.m, e = N.p + N.npass;
  const int f0 = e, u0 = N.q - ns;
  const int total = 2 * ns * m;
  for (int idx = local * kThreads + threadIdx.x; idx < total; idx += count * kThreads) {
    int si = idx / m, x = idx % m;
    int sv = (si < ns) ? f0 + si : u0 + (si - ns);
    if (phase == 0) {
      int ps = swp_pi(N, ns, sv);
      double v;
      if (x < e) v = N.L[(size_t)x * N.ldl + ps];
      else v = *swp_at(N, ps, swp_pi(N, ns, x));
      N.swp[(size_t)si * m + x] = v;
    } else {
      double v = N.swp[(size_t)si * m + x];
      if (x < e) N.L[(size_t)x * N.ldl + sv] = v;
      else *swp_at(N, sv, x) = v;
    }
  }
  if (phase == 1 && local == 0 && threadIdx.x < ns) {
    int i = threadIdx.x;
    int t = N.perm[f0 + i];
    N.perm[f0 + i] = N.perm[u0 + i];
    N.perm[u0 + i] = t;
  }
}

// ---------------------------------------------------------------------------
// Threshold partial pivoting on the remaining fully-summed columns
// [nelim, n) of the panel, one workgroup per front.  Updates the panel
// columns only; the contribution block is updated afterwards (k_update mode 1).
// ---------------------------------------------------------------------------
// Workgroup size of k_tail.  A level with fewer fronts than CUs gets 16
// waves per front: the TPP trailing update streams the panel from L2 and HBM,
// and more waves keep more of its loads in flight (on a delay-heavy KKT root
// 1024 threads beat 512 by 7% and 256 by 28%; the 216 bytes a lane spills at
// 1024 sit outside the update loops).  Wider levels keep 4 waves per front,
// so that many small fronts share each CU.
constexpr int kTailWide = 1024, kTailNarrow = 256, kTailWideMaxFronts = 256;

template <int NT>
struct TppShared {
  double red[NT / 64];
  int ired[NT / 64];
  int flag;
  // one candidate per wave (batched scan): result code, partner, 2x2 D^-1
  int code[NT / 64];
  int tt[NT / 64];
  double dd[3][NT / 64];
};

template <int NT>
__device__ double wg_max(double v, TppShared<NT> &sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh.red[0];
  for (int i = 1; i < NT / 64; ++i) r = fmax(r, sh.red[i]);
  __syncthreads();
  return r;
}

template <int NT>
__device__ int wg_and(int v, TppShared<NT> &sh) {
  for (int off = 32; off > 0; off >>= 1) v &= __shfl_xor(v, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh.ired[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = sh.ired[0];
  for (int i = 1; i < NT / 64; ++i) r &= sh.ired[i];
  __syncthreads();
  return r;
}

__device__ inline double &Lat(NodeDev &N, int r, int c) {
  return N.L[(size_t)c * N.ldl + r];
}

// all |entries| of row/col idx within [from, to) below small?
template <int NT>
__device__ int tpp_col_small(NodeDev &N, int idx, int from, int to, double small,
                             TppShared<NT> &sh) {
  int ok = 1;
  for (int c = from + threadIdx.x; c < idx; c += NT)
    if (!(fabs(Lat(N, idx, c)) < small)) ok = 0;
  for (int r = idx + threadIdx.x; r < to; r += NT)
    if (!(fabs(Lat(N, r, idx)) < small)) ok = 0;
  return wg_and(ok, sh);
}

template <int NT>
__device__ double tpp_rc_max(NodeDev &N, int col, int nelim, int m, int skip,
                             TppShared<NT> &sh) {
  double best = 0.0;
  for (int c = nelim + threadIdx.x; c < col; c += NT)
    if (c != skip) best = fmax(best, fabs(Lat(N, col, c)));
  for (int r = col + 1 + threadIdx.x; r < m; r += NT)
    if (r != skip) best = fmax(best, fabs(Lat(N, r, col)));
  return wg_max(best, sh);
}

template <int NT>
__device__ void tpp_swap(NodeDev &N, int c1, int c2, int m) {
  if (c1 == c2) return;
  if (c2 < c1) { int t = c1; c1 = c2; c2 = t; }
  const int tid = threadIdx.x;
  for (int c = tid; c < c1; c += NT) {
    double t = Lat(N, c1, c); Lat(N, c1, c) = Lat(N, c2, c); Lat(N, c2, c) = t;
  }
  for (int i = c1 + 1 + tid; i < c2; i += NT) {
    double t = Lat(N, i, c1); Lat(N, i, c1) = Lat(N, c2, i); Lat(N, c2, i) = t;
  }
  for (int r = c2 + 1 + tid; r < m; r += NT) {
    double t = Lat(N, r, c1); Lat(N, r, c1) = Lat(N, r, c2); Lat(N, r, c2) = t;
  }
  if (tid == 0) {
    double t = Lat(N, c1, c1); Lat(N, c1, c1) = Lat(N, c2, c2); Lat(N, c2, c2) = t;
    int pt = N.perm[c1]; N.perm[c1] = N.perm[c2]; N.perm[c2] = pt;
  }
  __syncthreads();
}

constexpr int kTppCols = 8;  // trailing columns per pass of the TPP update

template <int NT>
__device__ void tpp_elim1(NodeDev &N, int k, int m, int n, double *ld) {
  const double d11 = N.D[2 * k];
  for (int r = k + 1 + threadIdx.x; r < m; r += NT) {
    ld[r] = Lat(N, r, k);
    Lat(N, r, k) = ld[r] * d11;
  }
  __syncthreads();
  // trailing update, kTppCols columns per pass: a thread keeps its row's
  // pivot entry in a register and issues the column loads together
  for (int c0 = k + 1; c0 < n; c0 += kTppCols) {
    double sc[kTppCols];
#pragma unroll
    for (int j = 0; j < kTppCols; ++j) sc[j] = (c0 + j < n) ? ld[c0 + j] : 0.0;
    for (int r = c0 + threadIdx.x; r < m; r += NT) {
      const double lk = Lat(N, r, k);
      double v[kTppCols];
#pragma unroll
      for (int j = 0; j < kTppCols; ++j)
        v[j] = (c0 + j < n && r >= c0 + j) ? Lat(N, r, c0 + j) : 0.0;
#pragma unroll
      for (int j = 0; j < kTppCols; ++j)
        if (c0 + j < n && r >= c0 + j) Lat(N, r, c0 + j) = v[j] - lk * sc[j];
    }
  }
  if (threadIdx.x == 0) Lat(N, k, k) = 1.0;
  __syncthreads();
}

template <int NT>
__device__ void tpp_elim2(NodeDev &N, int k, int m, int n, double *ld, double *ld2) {
  const double d11 = N.D[2 * k], d21 = N.D[2 * k + 1], d22 = N.D[2 * k + 3];
  for (int r = k + 2 + threadIdx.x; r < m; r += NT) {
    double x1 = Lat(N, r, k), x2 = Lat(N, r, k + 1);
    ld[r] = x1;
    ld2[r] = x2;
    Lat(N, r, k) = d11 * x1 + d21 * x2;
    Lat(N, r, k + 1) = d21 * x1 + d22 * x2;
  }
  __syncthreads();
  for (int c0 = k + 2; c0 < n; c0 += kTppCols) {
    double s1[kTppCols], s2[kTppCols];
#pragma unroll
    for (int j = 0; j < kTppCols; ++j) {
      s1[j] = (c0 + j < n) ? ld[c0 + j] : 0.0;
      s2[j] = (c0 + j < n) ? ld2[c0 + j] : 0.0;
    }
    for (int r = c0 + threadIdx.x; r < m; r += NT) {
      const double l1 = Lat(N, r, k), l2 = Lat(N, r, k + 1);
      double v[kTppCols];
#pragma unroll
      for (int j = 0; j < kTppCols; ++j)
        v[j] = (c0 + j < n && r >= c0 + j) ? Lat(N, r, c0 + j) : 0.0;
#pragma unroll
      for (int j = 0; j < kTppCols; ++j)
        if (c0 + j < n && r >= c0 + j) Lat(N, r, c0 + j) = v[j] - (l1 * s1[j] + l2 * s2[j]);
    }
  }
  if (threadIdx.x == 0) {
    Lat(N, k, k) = 1.0;
    Lat(N, k + 1, k) = 0.0;
    Lat(N, k + 1, k + 1) = 1.0;
  }
  __syncthreads();
}

// One TPP candidate p (> nelim) tested by one wave, reading the panel only
// (the tests of a scan are read-only, so the waves of a workgroup test
// consecutive candidates at once and the first that passes is taken, as a
// one-at-a-time scan would).  The decisions are those of the scan's steps
// (ldlt_tpp.cxx:192-259): a zero row/column; else t = the first largest
// |a(p, c)|, c in [nelim, p), the 2x2 test on (t, p) against the largest
// other entries of rows/columns t and p, else the 1x1 test on p.  Two passes:
// row p and column p give the zero test, t, row p's largest entry other
// than t (its second largest: t holds the first largest) and column p's
// largest; row t and column t give t's largest other than p.  Every maximum
// is exact, so the decision equals the one-at-a-time scan's.  Returns the
// code (0 fail, 1 zero column, 2 2x2, 3 1x1), wave-uniform; t and the 2x2
// D^-1 in the out arguments.
__device__ inline int tpp_wave_test(NodeDev &N, int p, int nelim, int m, double u,
                                    double small, int &t_out, double &d0, double &d1,
                                    double &d3) {
  const int lane = threadIdx.x & 63;
  // pass A: row p over [nelim, p) and column p over [p, m)
  int big = 0;                        // some |entry| not below small (NaN counts)
  double m1 = -1.0, m2 = -1.0;        // row p: largest, second largest (-1: none)
  int a1 = INT_MAX;                   // first column holding m1
  for (int c = nelim + lane; c < p; c += 64) {
    const double v = fabs(Lat(N, p, c));
    if (!(v < small)) big = 1;
    if (v > m1) {
      m2 = m1;
      m1 = v;
      a1 = c;
    } else if (v > m2) {
      m2 = v;
    }
  }
  double cp = 0.0;                    // column p below the diagonal
  for (int r = p + lane; r < m; r += 64) {
    const double v = fabs(Lat(N, r, p));
    if (!(v < small)) big = 1;
    if (r > p) cp = fmax(cp, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double om1 = __shfl_xor(m1, off), om2 = __shfl_xor(m2, off);
    const int oa1 = __shfl_xor(a1, off);
    if (om1 > m1 || (om1 == m1 && oa1 < a1)) {
      m2 = fmax(om2, m1);
      m1 = om1;
      a1 = oa1;
    } else {
      m2 = fmax(m2, om1);
    }
    cp = fmax(cp, __shfl_xor(cp, off));
  }
  if (!__ballot(big)) return 1;
  // t: the first largest of row p (nelim when every entry is NaN)
  const int t = (a1 == INT_MAX) ? nelim : __builtin_amdgcn_readfirstlane(a1);
  t_out = t;
  // maxp = rc_max(p, skip t): row p without column t, column p below
  double maxp = fmax(fmax(0.0, (a1 == INT_MAX) ? -1.0 : m2), cp);
  // pass B: maxt = rc_max(t, skip p): row t over [nelim, t), column t over
  // (t, m) without row p
  double mt = 0.0;
  for (int c = nelim + lane; c < t; c += 64) mt = fmax(mt, fabs(Lat(N, t, c)));
  for (int r = t + 1 + lane; r < m; r += 64)
    if (r != p) mt = fmax(mt, fabs(Lat(N, r, t)));
  for (int off = 32; off > 0; off >>= 1) mt = fmax(mt, __shfl_xor(mt, off));
  const double maxt = uni(mt);
  maxp = uni(maxp);
  // 2x2 test (t < p)
  const double a11 = Lat(N, t, t), a21 = Lat(N, p, t), a22 = Lat(N, p, p);
  const double maxpiv = fmax(fabs(a11), fmax(fabs(a21), fabs(a22)));
  if (!(maxpiv < small)) {
    const double detscale = 1 / maxpiv;
    const double det0 = (a11 * detscale) * a22;
    const double det1 = (a21 * detscale) * a21;
    const double det = det0 - det1;
    if (!(fabs(det) < fmax(small, fmax(fabs(det0 / 2), fabs(det1 / 2))))) {
      const double e0 = (a22 * detscale) / det;
      const double e1 = (-a21 * detscale) / det;
      const double e3 = (a11 * detscale) / det;
      bool ok2;
      if (fmax(maxp, maxt) < small) {
        ok2 = true;
      } else {
        const double x1 = fabs(e0) * maxt + fabs(e1) * maxp;
        const double x2 = fabs(e1) * maxt + fabs(e3) * maxp;
        ok2 = u * fmax(x1, x2) < 1.0;
      }
      if (ok2) {
        d0 = e0;
        d1 = e1;
        d3 = e3;
        return 2;
      }
    }
  }
  const double maxp2 = fmax(maxp, fabs(a21));
  return (fabs(a22) >= u * maxp2) ? 3 : 0;
}

// TPP on the columns [nelim, n) that APP left, by the whole workgroup;
// returns the new nelim (uniform) and stores it in N.nelim.
template <int NT>
__device__ int tpp_front(NodeDev &N, const FactorParams &fp, int nelim, TppShared<NT> &sh,
                         int &s_t) {
  const int n = N.n, m = N.m;
  if (nelim >= n) return nelim;
  // as factor_node_indef (src/ssids/cpu/factor.hxx:77-80): TPP finishes a
  // front with no contribution block (m == n), or every front when the pivot
  // method or the failed-pivot method is TPP
  if (!(fp.pivot_tpp || m == n || fp.tpp_failed)) return nelim;
  const double u = fp.u, small = fp.small;
  double *ld = N.bak, *ld2 = N.bak + N.ldbak;
  const int tid = threadIdx.x;
  while (nelim < n) {
    if (tpp_col_small(N, nelim, nelim, m, small, sh)) {
      if (!fp.action) { if (tid == 0) N.flag = -5; return nelim; }
      for (int r = nelim + tid; r < m; r += NT) Lat(N, r, nelim) = 0.0;
      if (tid == 0) { N.D[2 * nelim] = 0.0; N.D[2 * nelim + 1] = 0.0; N.nzero++; }
      __syncthreads();
      ++nelim;
      continue;
    }
    int done = 0;
    // candidates nelim+1, nelim+2, ... in batches of one per wave; the first
    // that passes in order is taken
    constexpr int NW = NT / 64;
    const int wv = threadIdx.x >> 6;
    for (int base = nelim + 1; base < n && !done; base += NW) {
      const int pc = base + wv;
      int code = 0, tw = 0;
      double e0 = 0.0, e1 = 0.0, e3 = 0.0;
      if (pc < n) code = tpp_wave_test(N, pc, nelim, m, u, small, tw, e0, e1, e3);
      if ((tid & 63) == 0) {
        sh.code[wv] = code;
        sh.tt[wv] = tw;
        sh.dd[0][wv] = e0;
        sh.dd[1][wv] = e1;
        sh.dd[2][wv] = e3;
      }
      __syncthreads();
      int w0 = -1;
      for (int v = 0; v < NW; ++v)
        if (sh.code[v] != 0) {
          w0 = v;
          break;
        }
      int c = 0, t = 0;
      double d0 = 0.0, d1 = 0.0, d3 = 0.0;
      if (w0 >= 0) {
        c = sh.code[w0];
        t = sh.tt[w0];
        d0 = sh.dd[0][w0];
        d1 = sh.dd[1][w0];
        d3 = sh.dd[2][w0];
      }
      __syncthreads();  // every thread has read the batch before it is reused
      if (w0 < 0) continue;
      const int p = base + w0;
      done = 1;
      if (c == 1) {  // a zero row/column: a zero pivot at nelim
        if (!fp.action) { if (tid == 0) N.flag = -5; return nelim; }
        tpp_swap<NT>(N, p, nelim, m);
        for (int r = nelim + tid; r < m; r += NT) Lat(N, r, nelim) = 0.0;
        if (tid == 0) { N.D[2 * nelim] = 0.0; N.D[2 * nelim + 1] = 0.0; N.nzero++; }
        __syncthreads();
        ++nelim;
      } else if (c == 2) {
        tpp_swap<NT>(N, t, nelim, m);
        tpp_swap<NT>(N, p, nelim + 1, m);
        if (tid == 0) {
          N.D[2 * nelim] = d0;
          N.D[2 * nelim + 1] = d1;
          N.D[2 * nelim + 2] = INFINITY;
          N.D[2 * nelim + 3] = d3;
        }
        __syncthreads();
        tpp_elim2<NT>(N, nelim, m, n, ld, ld2);
        nelim += 2;
      } else {
        tpp_swap<NT>(N, p, nelim, m);
        if (tid == 0) {
          N.D[2 * nelim] = 1.0 / Lat(N, nelim, nelim);
          N.D[2 * nelim + 1] = 0.0;
        }
        __syncthreads();
        tpp_elim1<NT>(N, nelim, m, n, ld);
        nelim += 1;
      }
    }
    if (!done) {
      const int p = nelim;
      double maxp = tpp_rc_max(N, p, nelim, m, -1, sh);
      if (fabs(Lat(N, p, p)) >= u * maxp) {
        if (tid == 0) {
          N.D[2 * nelim] = 1.0 / Lat(N, nelim, nelim);
          N.D[2 * nelim + 1] = 0.0;
        }
        __syncthreads();
        tpp_elim1<NT>(N, nelim, m, n, ld);
        nelim += 1;
      } else {
        break;
      }
    }
  }
  if (tid == 0) N.nelim = nelim;
  return nelim;
}

// ---------------------------------------------------------------------------
// End of a level, one workgroup per front, in order:
//  * the APP result: nelim = p + npass of the last step (nelim_app);
//  * TPP on the columns APP failed (tpp_front);
//  * Dd: D per pivot in the (value, partner value, partner offset) form the
//    gather's L21 D operand uses;
//  * L21 D for the rows [n, m) and the TPP pivots (k_apply wrote it for the
//    APP pivots);
//  * ipos: pre-pivot -> post-pivot position of the fully-summed variables
//    (the variable-indexed scratch is conflict free: the fully-summed
//    variable sets of one level are disjoint).
// Values passed between the phases stay in registers; descriptor fields
// written here are re-read only after a barrier.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(NT) void k_tail(NodeDev *nodes, const int *lvl_slots,
                                                   FactorParams fp, int *scratch) {
  __shared__ TppShared<NT> sh;
  __shared__ int s_t;
  NodeDev &N = nodes[lvl_slots[blockIdx.x]];
  const int tid = threadIdx.x;
  int nelim;
  {
    const int nsteps = N.nsteps, p = N.p, npass = N.npass, q = N.q, w = N.w;
    nelim = (nsteps > 0) ? p + npass : 0;
    __syncthreads();  // every thread has read the step state
    if (tid == 0) {
      if (nsteps > 0) {
        N.p = nelim;
        N.q = q - (w - npass);
      }
      N.nelim = nelim;
      N.nelim_app = nelim;
    }
    __syncthreads();
  }
  const int t0 = nelim;  // APP pivots end here
  nelim = tpp_front(N, fp, nelim, sh, s_t);
  __syncthreads();
  // Dd
  for (int k = tid; k < nelim; k += NT) {
    double e1, e2;
    int two;
    pivot_d(N.D, k, nelim, e1, e2, two);
    if (two == 0) {
      N.Dd[3 * k] = e1;
      N.Dd[3 * k + 1] = 0.0;
      N.Dd[3 * k + 2] = 0.0;
    } else if (two == 1) {
      N.Dd[3 * k] = e1;
      N.Dd[3 * k + 1] = e2;
      N.Dd[3 * k + 2] = 1.0;
    } else {
      N.Dd[3 * k] = e2;
      N.Dd[3 * k + 1] = e1;
      N.Dd[3 * k + 2] = -1.0;
    }
  }
  __syncthreads();
  // L21 D for the TPP pivots
  if (!N.skip && N.LD && t0 < nelim) {
    const int n = N.n, m = N.m, ldl = N.ldl;
    const double *L = N.L, *Dd = N.Dd;
    double *LD = N.LD;
    for (int i = n + tid; i < m; i += NT) {
      for (int t = t0; t < nelim; ++t) {
        const double e = Dd[3 * t], f = Dd[3 * t + 1];
        const int off = (int)Dd[3 * t + 2];
        double v = L[(size_t)t * ldl + i] * e;
        if (off != 0) v += L[(size_t)(t + off) * ldl + i] * f;
        LD[(size_t)t * ldl + i] = v;
      }
    }
  }
  // ipos
  const int n = N.n;
  for (int i = tid; i < n; i += NT) scratch[N.perm[i]] = i;
  __syncthreads();
  for (int j = tid; j < n; j += NT) N.ipos[j] = scratch[N.perm0[j]];
}

// inertia and delay counts over all fronts (integer atomics: deterministic)
constexpr int kStatsWaves = 16, kStatsGroups = 64;
__global__ __launch_bounds__(64 * kStatsWaves) void k_stats(const NodeDev *nodes, int nslots,
                                                            int *counters, int pivot_tpp,
                                                            int tpp_failed) {
  // a wave per front (fronts strided over the grid's waves), a lane per
  // pivot: pivot i is the second member of a 2x2 when its D^{-1} slot holds
  // infinity ([d11, d21 | inf, d22]), the first member when the next one is,
  // a 1x1 otherwise.  Sums go wave -> workgroup (LDS) -> 7 atomics per
  // workgroup (integer: deterministic).
  constexpr int kC = 8;
  __shared__ int red[kStatsWaves][kC];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int c[kC] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int s = blockIdx.x * kStatsWaves + wv; s < nslots; s += gridDim.x * kStatsWaves) {
    const NodeDev &N = nodes[s];
    const int ne = N.nelim;
    gcdouble *d = (gcdouble *)N.D;
    for (int i = lane; i < ne; i += 64) {
      const double a11 = d[2 * i];
      if (!isfinite(a11)) continue;
      if (i + 1 < ne && !isfinite(d[2 * i + 2])) {
        const double a21 = d[2 * i + 1], a22 = d[2 * i + 3];
        c[1]++;
        const double det = a11 * a22 - a21 * a21, tr = a11 + a22;
        if (det < 0) c[0]++;
        else if (tr < 0) c[0] += 2;
      } else {
        if (a11 == 0.0) c[2]++;
        if (a11 < 0.0) c[0]++;
      }
    }
    if (lane == 0) {
      c[3] += N.n - N.nelim;
      // not_first_pass / not_second_pass as src/ssids/cpu/factor.hxx:74-105:
      // the columns APP left (TPP pivoting: the columns TPP left), and the
      // columns a TPP pass over APP's failures left
      if (pivot_tpp) {
        c[4] += N.n - N.nelim;
      } else if (N.nelim_app < N.n) {
        c[4] += N.n - N.nelim_app;
        if (N.m == N.n || tpp_failed) c[7] += N.n - N.nelim;
      }
      c[5] += (N.flag == -5);
      c[6] += (N.flag == -6);
    }
  }
#pragma unroll
  for (int j = 0; j < kC; ++j) {
    int v = c[j];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < kC) {
    int v = 0;
    for (int q = 0; q < kStatsWaves; ++q) v += red[q][threadIdx.x];
    if (v) atomicAdd(&counters[threadIdx.x], v);
  }
}

// ---------------------------------------------------------------------------
// Level set-up for the static plan (one workgroup): sizes each front of the
// level from its children's results -- nd = sum of the children's failed
// columns, m = sm + nd, n = sn + nd -- and writes the level's launch tables
// (exclusive prefix sums of level_table_count).  A front whose delays exceed
// the capacity its buffers were planned with, or with a skipped child, is
// skipped (m = n = 0, no work) and counted in *overflow when it is the first
// on its path; the host then refactors with the per-level schedule.
// Thread tid sizes fronts tid, tid+1024, ...; after a workgroup barrier each
// wave scans its own share of the tables.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_level_setup(NodeDev *nodes, int s0, int nl,
                                                      const int *child_list, int pivot_tpp,
                                                      int S, int *tab, int *overflow) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int li = tid; li < nl; li += 1024) {
    NodeDev &D = nodes[s0 + li];
    int nd = 0, poison = 0;
    for (int k = D.cbeg; k < D.cend; ++k) {
      const NodeDev &c = nodes[child_list[k]];
      nd += c.n - c.nelim;
      poison |= c.skip;
    }
    if (poison || nd > D.ndcap) {
      D.skip = 1;
      D.nd = 0;
      D.m = 0;
      D.n = 0;
      D.nsteps = 0;
      if (!poison) atomicAdd(overflow, 1);
    } else {
      D.skip = 0;
      D.nd = nd;
      D.m = D.sm + nd;
      D.n = D.sn + nd;
      D.nsteps = pivot_tpp ? 0 : (D.n + kBlk - 1) / kBlk;
    }
  }
  __syncthreads();  // the sizes above are visible to every wave
  // wave wv scans tables wv, wv + 16, ... over 64-front chunks, the running
  // sum in a register (no workgroup barrier per table)
  const int ntab = level_tables(S);
  for (int t = wv; t < ntab; t += 16) {
    int *out = tab + (size_t)t * (nl + 1);
    int carry = 0;
    for (int base = 0; base < nl; base += 64) {
      const int li = base + lane;
      int v = 0;
      if (li < nl) {
        const NodeDev &D = nodes[s0 + li];
        v = level_table_count(t, S, D.m, D.n, D.nsteps, D.is_root);
      }
      int x = v;  // inclusive scan within the wave
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
      }
      if (li < nl) out[li] = carry + x - v;
      carry += __shfl(x, 63);
    }
    if (lane == 0) out[nl] = carry;
  }
}

// the per-front results the host needs after a statically planned factor
__global__ void k_pack_nodes(const NodeDev *nodes, int nn, int *out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nn) return;
  const NodeDev &D = nodes[s];
  int *o = out + 8 * (size_t)s;
  o[0] = D.nd;
  o[1] = D.m;
  o[2] = D.n;
  o[3] = D.nelim;
  o[4] = D.nelim_app;
  o[5] = D.flag;
  o[6] = D.skip;
  o[7] = D.nsteps;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void launch_level_setup(hipStream_t st, NodeDev *nodes, int s0, int nl, const int *child_list,
                        int pivot_tpp, int S, int *tab, int *overflow) {
  if (nl > 0)
    hipLaunchKernelGGL(k_level_setup, dim3(1), dim3(1024), 0, st, nodes, s0, nl, child_list,
                       pivot_tpp, S, tab, overflow);
}
// D^{-1} (two per pivot) and the eliminated variables of every front, into
// contiguous arrays in node order: node i's pivots at [off[i], off[i+1]),
// with off from the host's read-back of each front's nelim
__global__ void k_pack_dperm(const NodeDev *nodes, const int *slot_of, const int64_t *off,
                             int nn, double *out_d, int *out_p) {
  for (int i = blockIdx.x; i < nn; i += gridDim.x) {
    const NodeDev &D = nodes[slot_of[i]];
    const int64_t o = off[i];
    const int ne = (int)(off[i + 1] - o);
    for (int k = threadIdx.x; k < 2 * ne; k += blockDim.x) out_d[2 * o + k] = D.D[k];
    for (int k = threadIdx.x; k < ne; k += blockDim.x) out_p[o + k] = D.perm[k];
  }
}
void launch_pack_dperm(hipStream_t st, const NodeDev *nodes, const int *slot_of,
                       const int64_t *off, int nn, double *out_d, int *out_p) {
  if (nn > 0)
    hipLaunchKernelGGL(k_pack_dperm, dim3(std::min(nn, 4096)), dim3(256), 0, st, nodes,
                       slot_of, off, nn, out_d, out_p);
}
void launch_pack_nodes(hipStream_t st, const NodeDev *nodes, int nn, int *out) {
  if (nn > 0)
    hipLaunchKernelGGL(k_pack_nodes, dim3((nn + 255) / 256), dim3(256), 0, st, nodes, nn, out);
}
void launch_gather_first(hipStream_t st, int nwg, NodeDev *nodes, const int *slots,
                         const int *wg_off, int nlvl, const int *child_list,
                         const double *aval, const int64_t *asrc, const int *arow,
                         const int *acix, const double *scaling, FactorParams fp) {
  const GatherArgs ga{wg_off, nlvl, child_list, aval, asrc, arow, acix, scaling};
  if (nwg <= 0) return;
  if (fp.c32)
    hipLaunchKernelGGL((k_diag<kDiagWaves, true, true>), dim3(nwg), dim3(kThreads), 0, st,
                       nodes, slots, 0, fp, ga);
  else
    hipLaunchKernelGGL((k_diag<kDiagWaves, true, false>), dim3(nwg), dim3(kThreads), 0, st,
                       nodes, slots, 0, fp, ga);
}
void launch_gather(hipStream_t st, int part, int nwg, NodeDev *nodes, const int *slots,
                   const int *wg_off, int nlvl, const int *child_list, const double *aval,
                   const int64_t *asrc, const int *arow, const int *acix,
                   const double *scaling, int c32) {
  if (nwg <= 0) return;
  auto kern = c32 ? ((part == 3)   ? k_gather<3, float>
                     : (part == 2) ? k_gather<2, float>
                                   : k_gather<0, float>)
                  : ((part == 3)   ? k_gather<3, double>
                     : (part == 2) ? k_gather<2, double>
                                   : k_gather<0, double>);
  hipLaunchKernelGGL(kern, dim3(nwg), dim3(kThreads), 0, st, nodes, slots, wg_off, nlvl,
                     child_list, aval, asrc, arow, acix, scaling);
}
void launch_maptab(hipStream_t st, int nlvl, NodeDev *nodes, const int *slots,
                   const int *child_list) {
  if (nlvl > 0)
    hipLaunchKernelGGL(k_maptab, dim3(nlvl, kMapParts), dim3(kThreads), 0, st, nodes, slots,
                       child_list);
}
void launch_tail(hipStream_t st, int nlvl, NodeDev *nodes, const int *slots, FactorParams fp,
                 int *scratch) {
  if (nlvl <= 0) return;
  if (nlvl < kTailWideMaxFronts)
    hipLaunchKernelGGL(k_tail<kTailWide>, dim3(nlvl), dim3(kTailWide), 0, st, nodes, slots, fp,
                       scratch);
  else
    hipLaunchKernelGGL(k_tail<kTailNarrow>, dim3(nlvl), dim3(kTailNarrow), 0, st, nodes, slots,
                       fp, scratch);
}
void launch_diag(hipStream_t st, int nlvl, NodeDev *nodes, const int *slots, int step,
                 FactorParams fp) {
  if (nlvl > 0)
    hipLaunchKernelGGL((k_diag<kDiagWaves, false, false>), dim3(nlvl), dim3(64 * kDiagWaves),
                       0, st,
                       nodes, slots, step, fp, GatherArgs{});
}
void launch_apply(hipStream_t st, int nwg, NodeDev *nodes, const int *slots,
                  const int *wg_off, int nlvl, int step, FactorParams fp) {
  if (nwg > 0)
    hipLaunchKernelGGL(k_apply, dim3(nwg), dim3(kThreads), 0, st, nodes, slots, wg_off,
                       nlvl, step, fp);
}
void launch_update(hipStream_t st, int nwg, NodeDev *nodes, const int *slots,
                   const int *wg_off, int nlvl, int step, int mode, FactorParams fp) {
  if (nwg <= 0) return;
  if (mode == 0)
    hipLaunchKernelGGL(k_update_panel, dim3(nwg), dim3(kThreads), 0, st, nodes, slots, wg_off,
                       nlvl, step, fp);
  else
    hipLaunchKernelGGL(k_update, dim3(nwg), dim3(kThreads), 0, st, nodes, slots, wg_off,
                       nlvl, step, mode, fp);
}
void launch_swap(hipStream_t st, int nwg, NodeDev *nodes, const int *slots,
                 const int *wg_off, int nlvl, int step, int phase) {
  if (nwg > 0)
    hipLaunchKernelGGL(k_swap, dim3(nwg), dim3(kThreads), 0, st, nodes, slots, wg_off,
                       nlvl, step, phase);
}
void launch_stats(hipStream_t st, const NodeDev *nodes, int nslots, int *counters,
                  const FactorParams &fp) {
  if (nslots > 0)
    hipLaunchKernelGGL(k_stats,
                       dim3(std::min(kStatsGroups, (nslots + kStatsWaves - 1) / kStatsWaves)),
                       dim3(64 * kStatsWaves), 0, st, nodes, nslots, counters, fp.pivot_tpp,
                       fp.tpp_failed);
}

}  // namespace ssids_hip
