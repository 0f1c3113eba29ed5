// Host driver of the HIP numeric path: symbolic upload, the tree-level walk
// over an HBM pool, the solves, enquiry and alteration.  Replaces the GPU
// numeric subtree host code of the reference (src/ssids/gpu/subtree.f90,
// factor.f90, solve.f90) with a level schedule whose only host round trip is
// one small read-back per level (delayed-pivot counts size the parents).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <new>
#include <stdexcept>
#include <vector>

#include "../../include/spral_ssids_hip.h"
#include "api_lock.hpp"
#include "dev.hpp"
#include "host.hpp"

namespace ssids_hip {

// kernel launchers (factor_kernels.hip / solve_kernels.hip)
void launch_gather(hipStream_t, int, int, NodeDev *, const int *, const int *, int, const int *,
                   const double *, const int64_t *, const int *, const int *, const double *,
                   int);
void launch_gather_first(hipStream_t, int, NodeDev *, const int *, const int *, int,
                         const int *, const double *, const int64_t *, const int *,
                         const int *, const double *, FactorParams);
void launch_maptab(hipStream_t, int, NodeDev *, const int *, const int *);
void launch_diag(hipStream_t, int, NodeDev *, const int *, int, FactorParams);
void launch_apply(hipStream_t, int, NodeDev *, const int *, const int *, int, int,
                  FactorParams);
void launch_update(hipStream_t, int, NodeDev *, const int *, const int *, int, int, int,
                   FactorParams);
void launch_swap(hipStream_t, int, NodeDev *, const int *, const int *, int, int, int);
void launch_tail(hipStream_t, int, NodeDev *, const int *, FactorParams, int *);
void launch_stats(hipStream_t, const NodeDev *, int, int *, const FactorParams &);
void launch_level_setup(hipStream_t, NodeDev *, int, int, const int *, int, int, int *, int *);
void launch_pack_nodes(hipStream_t, const NodeDev *, int, int *);
void launch_pack_dperm(hipStream_t, const NodeDev *, const int *, const int64_t *, int, double *,
                       int *);
void launch_fwd_tri(hipStream_t, int, const SolveNode *, const int *, const int *, double *,
                    int, int, double *);
void launch_fwd_rows(hipStream_t, int, const SolveNode *, const int *, const int *, int,
                     const int *, int, double *);
void launch_big_fwd_lvl(hipStream_t, const SolveNode *, const int *, int, int, double *, int,
                        int, double *, int *);
void launch_big_bwd_lvl(hipStream_t, const SolveNode *, const int *, int, int, double *, int,
                        int, double *, int *);
void launch_big_sync_clear(hipStream_t, int *, int);
int big_tri_threshold();
int fwd_rows_per_wg();
int bwd_rows_per_wg();
int bwd_cols_per_wg();
void launch_dsolve(hipStream_t, int, const SolveNode *, double *, int, int, int);
void launch_bwd(hipStream_t, int, const SolveNode *, const int *, double *, int, int,
                const int *, const int *, int, double *, int);

void launch_residual(hipStream_t, int, const int64_t *, const int *, const int64_t *,
                     const double *, const double *, const double *, double *, int, int);
void launch_anorm(hipStream_t, int, const int64_t *, const int64_t *, const double *,
                  unsigned long long *);
void launch_colmax3(hipStream_t, int, const double *, const double *, const double *, int, int,
                    unsigned long long *);
void launch_axpy(hipStream_t, int, int, int, const double *, double *, const int *);
void launch_scale_rows(hipStream_t, int, int, int, const double *, double *);

struct HipError : std::runtime_error {
  hipError_t code;
  explicit HipError(hipError_t c) : std::runtime_error(hipGetErrorString(c)), code(c) {}
};

#define HIPCHK(x)                        \
  do {                                   \
    hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) throw HipError(e_); \
  } while (0)

// Copies the library makes are ordered on a stream and have completed on
// return.  None goes through the legacy null stream: a copy there waits for
// every blocking stream of the process, so a solve on one handle would wait
// for a refactor running on another (a service runs both at once).
static void copy_now(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                     hipStream_t st) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
  HIPCHK(hipStreamSynchronize(st));
}
static void copy2d_now(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                       size_t height, hipMemcpyKind kind, hipStream_t st) {
  HIPCHK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, st));
  HIPCHK(hipStreamSynchronize(st));
}
// a non-blocking stream of this host thread, for copies outside any handle
// (kept until the process ends)
static hipStream_t thread_stream() {
  thread_local std::map<int, hipStream_t> streams;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipStream_t &st = streams[dev];
  if (!st) HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return st;
}

// ---------------------------------------------------------------------------
// Device memory pool: large chunks, best-fit free blocks (indexed by size, so
// an allocation is O(log blocks) however many fronts a batch places) with
// coalescing on release.
// ---------------------------------------------------------------------------
class Pool {
 public:
  ~Pool() {
    for (auto &c : chunks_) (void)hipFree(c.base);
  }
  void *alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    for (auto &c : chunks_) {
      auto it = c.bysize.lower_bound({bytes, 0});  // smallest block that fits
      if (it == c.bysize.end()) continue;
      const size_t sz = it->first, off = it->second;
      c.bysize.erase(it);
      c.free.erase(off);
      if (sz > bytes) c.put(off + bytes, sz - bytes);
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return (char *)c.base + off;
    }
    size_t csz = std::max(bytes, chunk_min_);
    Chunk c;
    HIPCHK(hipMalloc(&c.base, csz));
    c.size = csz;
    if (csz > bytes) c.put(bytes, csz - bytes);
    chunks_.push_back(std::move(c));
    used_ += bytes;
    peak_ = std::max(peak_, used_);
    return chunks_.back().base;
  }
  void release(void *p, size_t bytes) {
    if (!p) return;
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    for (auto &c : chunks_) {
      char *b = (char *)c.base;
      if ((char *)p >= b && (char *)p < b + c.size) {
        size_t off = (char *)p - b, sz = bytes;
        // merge with the free neighbours on either side
        auto nx = c.free.lower_bound(off);
        if (nx != c.free.end() && off + sz == nx->first) {
          sz += nx->second;
          c.bysize.erase({nx->second, nx->first});
          nx = c.free.erase(nx);
        }
        if (nx != c.free.begin()) {
          auto pv = std::prev(nx);
          if (pv->first + pv->second == off) {
            off = pv->first;
            sz += pv->second;
            c.bysize.erase({pv->second, pv->first});
            c.free.erase(pv);
          }
        }
        c.put(off, sz);
        used_ -= bytes;
        return;
      }
    }
  }
  void set_chunk_min(size_t b) { chunk_min_ = b; }
  size_t peak() const { return peak_; }

 private:
  struct Chunk {
    void *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;                 // offset -> size
    std::set<std::pair<size_t, size_t>> bysize;    // (size, offset)
    void put(size_t off, size_t sz) {
      free[off] = sz;
      bysize.insert({sz, off});
    }
  };
  std::vector<Chunk> chunks_;
  size_t chunk_min_ = size_t(1) << 30;
  size_t used_ = 0, peak_ = 0;
};

template <class T>
T *dev_upload(const std::vector<T> &v, std::vector<void *> &owned) {
  if (v.empty()) return nullptr;
  void *p;
  HIPCHK(hipMalloc(&p, v.size() * sizeof(T)));
  copy_now(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, thread_stream());
  owned.push_back(p);
  return (T *)p;
}

// ---------------------------------------------------------------------------
// Symbolic subtree
// ---------------------------------------------------------------------------
struct Symbolic {
  int device = 0, n = 0, nnodes = 0, sa = 1;
  std::vector<int> sn, sm, parent;        // per node (local 0-based)
  std::vector<std::vector<int>> children;
  std::vector<int> slot;                  // node -> slot
  std::vector<int> node_of;               // slot -> node
  std::vector<int> lvl_ptr;               // slots of level L: [lvl_ptr[L], lvl_ptr[L+1])
  std::vector<int> child_list;            // children slots, per node contiguous
  std::vector<int> cbeg, cend;            // per node (node-indexed)
  std::vector<int64_t> row_off, rmap_off, acol_off;
  std::vector<int> rows_h;                // 0-based elimination indices
  std::vector<int> rmap_h;                // host copy of the child -> parent row map
  int64_t naval_min = 0;                  // max A index referenced + 1
  int64_t num_flops = 0;
  // device
  int *d_rows = nullptr, *d_rmap = nullptr, *d_arow = nullptr, *d_acix = nullptr;
  int *d_child = nullptr;
  int64_t *d_asrc = nullptr, *d_acol = nullptr;
  // full symmetric matrix in elimination order (CSR) for refinement residuals
  int64_t *d_fptr = nullptr, *d_fsrc = nullptr;
  int *d_fcol = nullptr;
  std::vector<void *> owned;
  ~Symbolic() {
    for (void *p : owned) (void)hipFree(p);
  }
};

// Static factor plan: every front's buffers placed in the pool up front for a
// delay capacity ndcap per front, so a factorization runs as one stream of
// launches with no host round trip (device kernels size each level from the
// children's results).  Temporaries (contribution blocks, workspaces) share
// pool memory by lifetime: a child's block is reused by later levels once its
// parent has been assembled, which stream order makes safe.
struct Plan {
  bool valid = false;
  int gen = 0;                             // bumped by every build (graphs key on it)
  int pivot_tpp = -1;
  int c32 = -1;                            // contribution-block precision planned for
  std::vector<NodeDev> tmpl;               // slot-indexed static fields
  NodeDev *d_tmpl = nullptr;
  struct Level {
    int maxsteps = 0;
    bool multi = false;
    size_t tab_off = 0;
    std::vector<int> cap;                   // capacity workgroups per table
  };
  std::vector<Level> lv;
  int *d_tabs = nullptr;
  size_t tabs_cap = 0;
  int *d_pack = nullptr;                   // 8 ints per front
  std::vector<std::pair<void *, size_t>> allocs;
  // the large transient blocks (contribution blocks, L21 D, workspaces of
  // kArenaMin bytes or more), at offsets packed by lifetime
  void *arena = nullptr;
  size_t arena_bytes = 0;
  ~Plan() {
    if (arena) (void)hipFree(arena);
    if (d_tmpl) (void)hipFree(d_tmpl);
    if (d_tabs) (void)hipFree(d_tabs);
    if (d_pack) (void)hipFree(d_pack);
  }
};

static std::shared_mutex api_mu;
static thread_local int api_depth = 0;

ApiShared::ApiShared() {
  if (api_depth++ == 0) api_mu.lock_shared();
}
ApiShared::~ApiShared() {
  if (--api_depth == 0) api_mu.unlock_shared();
}
CaptureExclusive::CaptureExclusive() : nested_(api_depth > 0) {
  if (nested_) api_mu.unlock_shared();
  api_mu.lock();
}
CaptureExclusive::~CaptureExclusive() {
  api_mu.unlock();
  if (nested_) api_mu.lock_shared();
}

struct Numeric {
  Symbolic *symb = nullptr;
  int c32 = 0;  // contribution blocks in FP32 (schur_precision 1, mixed precision)
  hipStream_t stream = nullptr;
  // second stream of the static plan: a level's gather beyond its first tile
  // column runs there while the first pivot step runs on `stream`
  hipStream_t stream2 = nullptr;
  // graph capture records on this stream only.  It and `stream2` are
  // non-blocking: an operation on the legacy null stream from another host
  // thread (a synchronous copy) would otherwise depend on the capturing
  // stream, which invalidates the capture
  hipStream_t stream_cap = nullptr;
  hipEvent_t ev_level = nullptr, ev_rest = nullptr, ev_first = nullptr;
  // the static-plan factorization captured as a graph, replayed by later
  // refactors with the same plan, parameters and buffers
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<double> gkey;
  Pool pool;
  Plan plan;
  std::vector<int> nd_seen;               // per slot: most delays seen in a factor
  std::vector<NodeDev> nodes;             // slot-indexed host mirror
  NodeDev *d_nodes = nullptr;
  int *d_lvl = nullptr;                   // per-level scratch (slots + offsets)
  size_t d_lvl_cap = 0;
  int *d_counters = nullptr;
  int *d_scratch_n = nullptr;             // n ints (ipos)
  std::vector<std::pair<void *, size_t>> persistent;
  // solve
  SolveNode *d_sn = nullptr;
  int *d_slots_all = nullptr;             // 0..nnodes-1 (slots in level order)
  // per level: nlvl+1 workgroup offsets of the solve's row kernels (forward:
  // 64-row workgroups, backward: 256-row workgroups), one array
  int *d_gemv_tab = nullptr;
  std::vector<int> gemv_tab_off, gemv_total;  // forward: offset into d_gemv_tab, total
  std::vector<int> bwd_tab_off, bwd_total;    // backward: row workgroups per front
  std::vector<int> bwdc_tab_off, bwdc_total;  // backward: row workgroups x column chunks
  std::vector<int> bwd_count;                 // per slot: backward row workgroups
  // per level: the fronts whose triangle is solved by the big-front level
  // kernels, as a device table [block offsets | slots | sync offsets]
  std::vector<std::vector<int>> big_of_level;
  std::vector<int> bigtab_off, big_blocks;     // per level: table offset, blocks
  int *d_bigtab = nullptr;
  int64_t solve_sync_ints = 0;                // sync words after the workspace
  std::vector<SolveNode> sn_h;
  double *d_X = nullptr;                  // solve workspace (solve_ws_per_rhs x nrhs)
  size_t d_X_cap = 0;
  int64_t solve_ws_per_rhs = 0;           // doubles of workspace per right-hand side
  double *d_x = nullptr;
  size_t d_x_cap = 0;
  bool posdef = false;                    // A = L L^T (pivots in order, all > 0)
  spral_ssids_hip_stats stats{};
  double *d_aval = nullptr;               // copy of the factorized values
  size_t d_aval_cap = 0;
  double *d_scal = nullptr;               // scaling in elimination order (or null)
  double *d_r = nullptr;                  // refinement workspace (2 x n x nrhs)
  size_t d_r_cap = 0;
  unsigned long long *d_norms = nullptr;  // 1 + 3*nrhs ordered-bit maxima
  int *d_active = nullptr;                // nrhs flags
  int d_norms_nrhs = 0;                   // columns d_norms/d_active hold
  double anorm = -1.0;
  // contribution export
  std::vector<double> contrib_val, contrib_delay_val;
  std::vector<int> contrib_rlist, contrib_delay_perm;
  ~Numeric() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &p : persistent) pool.release(p.first, p.second);
    if (d_nodes) (void)hipFree(d_nodes);
    if (d_lvl) (void)hipFree(d_lvl);
    if (d_counters) (void)hipFree(d_counters);
    if (d_scratch_n) (void)hipFree(d_scratch_n);
    if (d_sn) (void)hipFree(d_sn);
    if (d_slots_all) (void)hipFree(d_slots_all);
    if (d_gemv_tab) (void)hipFree(d_gemv_tab);
    if (d_bigtab) (void)hipFree(d_bigtab);
    if (d_X) (void)hipFree(d_X);
    if (d_x) (void)hipFree(d_x);
    if (d_aval) (void)hipFree(d_aval);
    if (d_scal) (void)hipFree(d_scal);
    if (d_r) (void)hipFree(d_r);
    if (d_norms) (void)hipFree(d_norms);
    if (d_active) (void)hipFree(d_active);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_level) (void)hipEventDestroy(ev_level);
    if (ev_rest) (void)hipEventDestroy(ev_rest);
    if (ev_first) (void)hipEventDestroy(ev_first);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream_cap) (void)hipStreamDestroy(stream_cap);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static Symbolic *build_symbolic(int device, int n, int sa, int en, const int *sptr,
                                const int *sparent, const int64_t *rptr, const int *rlist,
                                const int64_t *nptr, const int64_t *nlist, int schedule) {
  auto S = std::make_unique<Symbolic>();
  S->device = device;
  S->n = n;
  S->sa = sa;
  const int nn = en - sa;
  S->nnodes = nn;
  S->sn.resize(nn);
  S->sm.resize(nn);
  S->parent.assign(nn, -1);
  S->children.assign(nn, {});
  S->row_off.resize(nn + 1);
  S->rmap_off.resize(nn + 1);
  S->acol_off.resize(nn + 1);
  for (int i = 0; i < nn; ++i) {
    int s = sa - 1 + i;
    S->sn[i] = sptr[s + 1] - sptr[s];
    S->sm[i] = (int)(rptr[s + 1] - rptr[s]);
    int par = sparent[s] - 1 - (sa - 1);
    if (par >= 0 && par < nn) S->parent[i] = par;
  }
  for (int i = 0; i < nn; ++i)
    if (S->parent[i] >= 0) S->children[S->parent[i]].push_back(i);
  // rows (0-based) and child->parent row maps
  int64_t tot = 0;
  for (int i = 0; i < nn; ++i) {
    S->row_off[i] = tot;
    tot += S->sm[i];
  }
  S->row_off[nn] = tot;
  S->rows_h.resize(tot);
  for (int i = 0; i < nn; ++i) {
    int s = sa - 1 + i;
    for (int k = 0; k < S->sm[i]; ++k) S->rows_h[S->row_off[i] + k] = rlist[rptr[s] - 1 + k] - 1;
  }
  std::vector<int> rmap;
  std::vector<int> pos(n, -1);
  int64_t rtot = 0;
  for (int i = 0; i < nn; ++i) {
    S->rmap_off[i] = rtot;
    rtot += S->sm[i] - S->sn[i];
  }
  S->rmap_off[nn] = rtot;
  rmap.assign(rtot, 0);
  for (int p = 0; p < nn; ++p) {
    for (int k = 0; k < S->sm[p]; ++k) pos[S->rows_h[S->row_off[p] + k]] = k;
    for (int c : S->children[p])
      for (int k = S->sn[c]; k < S->sm[c]; ++k)
        rmap[S->rmap_off[c] + k - S->sn[c]] = pos[S->rows_h[S->row_off[c] + k]];
  }
  // A entries sorted by column per node
  std::vector<int64_t> asrc;
  std::vector<int> arow, acix;  // front row and own-column index of each entry
  std::vector<int64_t> acol;
  int64_t amax = 0;
  for (int i = 0; i < nn; ++i) {
    int s = sa - 1 + i;
    int sm = S->sm[i], snn = S->sn[i];
    int64_t e0 = nptr[s] - 1, e1 = nptr[s + 1] - 1;
    std::vector<std::vector<std::pair<int64_t, int>>> bycol(snn);
    for (int64_t e = e0; e < e1; ++e) {
      int64_t src = nlist[2 * e] - 1, dst = nlist[2 * e + 1] - 1;
      int c = (int)(dst / sm), r = (int)(dst % sm);
      bycol[c].push_back({src, r});
      amax = std::max(amax, src + 1);
    }
    S->acol_off[i] = (int64_t)acol.size();
    for (int c = 0; c < snn; ++c) {
      acol.push_back((int64_t)asrc.size());
      for (auto &pr : bycol[c]) {
        asrc.push_back(pr.first);
        arow.push_back(pr.second);
        acix.push_back(c);
      }
    }
    acol.push_back((int64_t)asrc.size());
  }
  S->acol_off[nn] = (int64_t)acol.size();
  S->naval_min = amax;
  // tree levels (children precede parents: postorder)
  const std::vector<int> lev = tree_levels(S->parent, S->sn, schedule);
  int nlev = 0;
  for (int i = 0; i < nn; ++i) nlev = std::max(nlev, lev[i] + 1);
  S->lvl_ptr.assign(nlev + 1, 0);
  for (int i = 0; i < nn; ++i) S->lvl_ptr[lev[i] + 1]++;
  for (int L = 0; L < nlev; ++L) S->lvl_ptr[L + 1] += S->lvl_ptr[L];
  S->slot.assign(nn, 0);
  S->node_of.assign(nn, 0);
  {
    std::vector<int> fill(S->lvl_ptr.begin(), S->lvl_ptr.end() - 1);
    for (int i = 0; i < nn; ++i) {
      int sl = fill[lev[i]]++;
      S->slot[i] = sl;
      S->node_of[sl] = i;
    }
  }
  S->cbeg.assign(nn, 0);
  S->cend.assign(nn, 0);
  for (int i = 0; i < nn; ++i) {
    S->cbeg[i] = (int)S->child_list.size();
    for (int c : S->children[i]) S->child_list.push_back(S->slot[c]);
    S->cend[i] = (int)S->child_list.size();
  }
  int64_t fl = 0;
  for (int i = 0; i < nn; ++i) {
    int64_t ne = S->sn[i], mm = S->sm[i] - ne;
    for (int64_t j = 1; j <= ne; ++j) fl += (mm + j) * (mm + j);
  }
  S->num_flops = fl;
  HIPCHK(hipSetDevice(device));
  S->d_rows = dev_upload(S->rows_h, S->owned);
  S->d_rmap = dev_upload(rmap, S->owned);
  S->rmap_h = rmap;
  S->d_asrc = dev_upload(asrc, S->owned);
  S->d_arow = dev_upload(arow, S->owned);
  S->d_acix = dev_upload(acix, S->owned);
  S->d_acol = dev_upload(acol, S->owned);
  S->d_child = dev_upload(S->child_list, S->owned);
  // full CSR (rows sorted by column) of the matrix in elimination order
  {
    std::vector<int64_t> cnt(n + 1, 0);
    auto each = [&](auto fn) {
      for (int i = 0; i < nn; ++i) {
        const int *rows = &S->rows_h[S->row_off[i]];
        int64_t a0 = acol[S->acol_off[i]];
        for (int c = 0; c < S->sn[i]; ++c)
          for (int64_t e = acol[S->acol_off[i] + c]; e < acol[S->acol_off[i] + c + 1]; ++e)
            fn(rows[arow[e]], rows[c], asrc[e]);
        (void)a0;
      }
    };
    each([&](int r, int c, int64_t) {
      cnt[r + 1]++;
      if (r != c) cnt[c + 1]++;
    });
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int64_t> fsrc(cnt[n]);
    std::vector<int> fcol(cnt[n]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    each([&](int r, int c, int64_t src) {
      fcol[pos[r]] = c;
      fsrc[pos[r]++] = src;
      if (r != c) {
        fcol[pos[c]] = r;
        fsrc[pos[c]++] = src;
      }
    });
    for (int i = 0; i < n; ++i) {
      std::vector<std::pair<int, int64_t>> rowv;
      for (int64_t e = cnt[i]; e < cnt[i + 1]; ++e) rowv.push_back({fcol[e], fsrc[e]});
      std::sort(rowv.begin(), rowv.end());
      for (int64_t e = cnt[i]; e < cnt[i + 1]; ++e) {
        fcol[e] = rowv[e - cnt[i]].first;
        fsrc[e] = rowv[e - cnt[i]].second;
      }
    }
    S->d_fptr = dev_upload(cnt, S->owned);
    S->d_fcol = dev_upload(fcol, S->owned);
    S->d_fsrc = dev_upload(fsrc, S->owned);
  }
  return S.release();
}

// ---------------------------------------------------------------------------
// Numeric factorization
// ---------------------------------------------------------------------------
static void ensure_lvl(Numeric &N, size_t ints) {
  if (ints <= N.d_lvl_cap) return;
  if (N.d_lvl) HIPCHK(hipFree(N.d_lvl));
  size_t cap = std::max(ints, size_t(1) << 16);
  HIPCHK(hipMalloc(&N.d_lvl, cap * sizeof(int)));
  N.d_lvl_cap = cap;
}

static inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

static hipEvent_t timed_event(hipStream_t s) {
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(e, s));
  return e;
}

// Static fields of a front (everything but the delay-dependent sizes).
static void node_static(const Symbolic &S, int sl, NodeDev &D) {
  const int i = S.node_of[sl];
  D.sm = S.sm[i];
  D.sn = S.sn[i];
  D.parent = S.parent[i] >= 0 ? S.slot[S.parent[i]] : -1;
  D.cbeg = S.cbeg[i];
  D.cend = S.cend[i];
  D.is_root = (S.parent[i] < 0) ? 1 : 0;
  D.rows = S.d_rows + S.row_off[i];
  D.rmap = S.d_rmap + S.rmap_off[i];
  D.acol = S.d_acol + S.acol_off[i];
}

// The contribution block and L21 D share one allocation (same lifetime)
static void set_contrib(NodeDev &D, char *base, size_t cbytes, size_t ldbytes) {
  D.C = (D.ldc > 0) ? (double *)base : nullptr;
  D.LD = ldbytes ? (double *)(base + cbytes) : nullptr;
}
// The APP workspace: backups, the block inverse, the swap buffer
static void set_work(NodeDev &D, char *w, int nsteps) {
  D.bak = (double *)w;
  D.dbak = D.bak + (size_t)D.ldbak * kBlk;
  D.W = D.dbak + kBlk * kBlk;
  D.swp = (nsteps > 1) ? D.W + kBlk * kBlk : nullptr;
}

// Bytes of a front's transient blocks (D.ldl, D.ldc, D.ldbak set): the
// contribution block (FP32 under mixed precision), L21 D (not for roots) and
// the APP workspace
static void trans_sizes(const Numeric &N, const NodeDev &D, int n, int nsteps, int m,
                        size_t &cbytes, size_t &ldbytes, size_t &wb) {
  const size_t cesz = N.c32 ? sizeof(float) : sizeof(double);
  cbytes = ((size_t)D.ldc * (D.sm - D.sn) * cesz + 255) & ~size_t(255);
  ldbytes = D.is_root ? 0 : (size_t)D.ldl * n * sizeof(double);
  wb = ((size_t)D.ldbak * kBlk + 2 * kBlk * kBlk) * sizeof(double);
  if (nsteps > 1) wb += (size_t)2 * kBlk * m * sizeof(double);
}

// Buffers of a front with room for m rows / n fully-summed columns.
// Persistent factors come from `keep`; the contribution block and the
// workspace come from `trans(bytes, kind)` (kind 0: contribution block,
// 1: workspace) and are returned through cb/wk for the
// caller to release.  `trans` may return null for a block placed later
// (set_contrib / set_work then set its pointers).
template <class Keep, class Trans>
static void node_buffers(Numeric &N, const Symbolic &S, int sl, NodeDev &D, int m, int n,
                         int nsteps, Keep keep, Trans trans, std::pair<void *, size_t> &cb,
                         std::pair<void *, size_t> &wk) {
  const int i = S.node_of[sl];
  // columns start on 128-byte lines (16 doubles), so a 64-row tile column is
  // four whole lines for the kernels that read and write fronts by tiles
  auto lines = [](int r) { return (r + 15) & ~15; };
  D.ldl = lines(std::max(m, 1));
  D.ldc = (D.sm > D.sn) ? lines(D.sm - D.sn) : 0;
  D.ldbak = std::max(m, 1);
  D.L = (double *)keep((size_t)D.ldl * n * sizeof(double));
  D.pmap = (int *)keep(std::max(m, 1) * sizeof(int));
  D.D = (double *)keep(2 * (size_t)std::max(n, 1) * sizeof(double));
  D.Dd = (double *)keep(3 * (size_t)std::max(n, 1) * sizeof(double));
  D.perm = (int *)keep(std::max(n, 1) * sizeof(int));
  D.perm0 = (int *)keep(std::max(n, 1) * sizeof(int));
  D.ipos = (int *)keep(std::max(n, 1) * sizeof(int));
  size_t cbytes, ldbytes, wb;
  trans_sizes(N, D, n, nsteps, m, cbytes, ldbytes, wb);
  char *cbase = (cbytes + ldbytes) ? (char *)trans(cbytes + ldbytes, 0) : nullptr;
  if (cbase) set_contrib(D, cbase, cbytes, ldbytes);
  cb = {cbase, cbytes + ldbytes};
  // child tile tables: built by the assembly, kept for the forward solve's gather
  const size_t tabn = (size_t)(cdiv(m, 64) + 1) * std::max<size_t>(S.children[i].size(), 1);
  D.ctab = (int *)keep(tabn * sizeof(int));
  char *w = (char *)trans(wb, 1);
  if (w) set_work(D, w, nsteps);
  wk = {w, wb};
}

// Called with no launch of this handle in flight (a factorization starts, or
// a static factorization has ended and synchronised).
static void plan_release(Numeric &N) {
  for (auto &a : N.plan.allocs) N.pool.release(a.first, a.second);
  N.plan.allocs.clear();
  if (N.plan.arena) (void)hipFree(N.plan.arena);
  N.plan.arena = nullptr;
  N.plan.arena_bytes = 0;
  N.plan.valid = false;
}

// Offsets for blocks (bytes, first level, last level) with closed lifetimes:
// largest first, each at the lowest offset free of every placed block whose
// lifetime meets its own.  Returns the arena size.
struct ArenaBlock {
  size_t bytes;
  int t0, t1;
  int slot, kind;  // kind 0: contribution block + L21 D, 1: workspace
  size_t off;
};
static size_t arena_place(std::vector<ArenaBlock> &b) {
  std::vector<int> order(b.size());
  for (size_t k = 0; k < b.size(); ++k) order[k] = (int)k;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    if (b[x].bytes != b[y].bytes) return b[x].bytes > b[y].bytes;
    return x < y;
  });
  std::vector<int> placed;
  std::vector<std::pair<size_t, size_t>> busy;
  size_t arena = 0;
  for (int k : order) {
    busy.clear();
    for (int j : placed)
      if (!(b[j].t1 < b[k].t0 || b[k].t1 < b[j].t0)) busy.push_back({b[j].off, b[j].bytes});
    std::sort(busy.begin(), busy.end());
    size_t cur = 0;
    for (auto &o : busy) {
      if (o.first >= cur + b[k].bytes) break;
      cur = std::max(cur, o.first + o.second);
    }
    b[k].off = cur;
    placed.push_back(k);
    arena = std::max(arena, cur + b[k].bytes);
  }
  return arena;
}

// Place every front for a delay capacity of
//   ndcap = seen + ceil((multiplier - 1) (sn + seen)),
// at most the children's fully-summed columns and at most what fits in the
// APP steps the front needs anyway (seen = most delays observed at this
// front in an earlier factorization, else 0).
static void plan_build(Numeric &N, int pivot_tpp, double multiplier) {
  Symbolic &S = *N.symb;
  Plan &P = N.plan;
  plan_release(N);
  for (auto &p : N.persistent) N.pool.release(p.first, p.second);
  N.persistent.clear();
  const int nn = S.nnodes, nlev = (int)S.lvl_ptr.size() - 1;
  if ((int)N.nd_seen.size() != nn) N.nd_seen.assign(nn, 0);
  const double slack = std::max(0.0, multiplier - 1.0);
  P.pivot_tpp = pivot_tpp;
  P.c32 = N.c32;
  ++P.gen;
  P.tmpl.assign(nn, NodeDev{});
  P.lv.assign(nlev, Plan::Level{});
  std::vector<std::pair<void *, size_t>> cof(nn), wof(nn);
  auto keep = [&](size_t b) {
    void *p = N.pool.alloc(b);
    P.allocs.push_back({p, b});
    return p;
  };
  // Transient blocks of kArenaMin bytes or more go to one arena, placed by
  // lifetime once every front is sized: a contribution block lives from its
  // front's level to its parent's, a workspace for its own level.  In the
  // pool, the space such a block frees is split by the smaller kept blocks
  // placed after it while the blocks grow toward the root, so each level near
  // the root would take a new chunk (a 27-point KKT tree at 56^3: 334 GiB of
  // chunks for 97 GiB of live blocks; with the arena 145 GiB).
  constexpr size_t kArenaMin = size_t(64) << 20;
  std::vector<int> lev_of(nn);
  for (int L = 0; L < nlev; ++L)
    for (int sl = S.lvl_ptr[L]; sl < S.lvl_ptr[L + 1]; ++sl) lev_of[sl] = L;
  std::vector<ArenaBlock> big;
  int cur_sl = 0, cur_lev = 0;
  auto trans = [&](size_t b, int kind) -> void * {
    if (b < kArenaMin) return N.pool.alloc(b);
    const NodeDev &D = P.tmpl[cur_sl];
    const int t1 = (kind == 1) ? cur_lev : (D.parent >= 0 ? lev_of[D.parent] : nlev);
    big.push_back({(b + 255) & ~size_t(255), cur_lev, t1, cur_sl, kind, 0});
    return nullptr;
  };
  size_t tab_total = 0;
  for (int L = 0; L < nlev; ++L) {
    const int s0 = S.lvl_ptr[L], s1 = S.lvl_ptr[L + 1], nl = s1 - s0;
    cur_lev = L;
    Plan::Level &lv = P.lv[L];
    std::vector<int> mc(nl), nc(nl), sc(nl);
    for (int sl = s0; sl < s1; ++sl) {
      NodeDev &D = P.tmpl[sl];
      node_static(S, sl, D);
      const int i = S.node_of[sl];
      int room = 0;  // most delays the children can hand up
      for (int c : S.children[i]) {
        const NodeDev &C = P.tmpl[S.slot[c]];
        room += C.sn + C.ndcap;
      }
      const int seen = N.nd_seen[sl];
      int cap = seen + (int)std::ceil(slack * (D.sn + seen));
      if (seen > 0) {
        // a front that has received delays receives a varying number when
        // the values change (refactors of one pattern): room for half as
        // many again as the most seen, past the last APP step's padding (a
        // factorization that exceeds the plan repeats itself level by level)
        cap = std::max(cap, seen + (seen + 1) / 2);
      } else if (!pivot_tpp) {
        // no slack beyond the last APP step the front needs anyway: an extra
        // step would cost every factorization a round of empty launches
        cap = std::min(cap, cdiv(D.sn, kBlk) * kBlk - D.sn);
      }
      D.ndcap = std::max(seen, std::min(cap, room));
      const int m = D.sm + D.ndcap, n = D.sn + D.ndcap;
      const int nsteps = pivot_tpp ? 0 : cdiv(n, kBlk);
      mc[sl - s0] = m;
      nc[sl - s0] = n;
      sc[sl - s0] = nsteps;
      lv.maxsteps = std::max(lv.maxsteps, nsteps);
      if (nsteps > 1) lv.multi = true;
      cur_sl = sl;
      node_buffers(N, S, sl, D, m, n, nsteps, keep, trans, cof[sl], wof[sl]);
    }
    const int ntab = level_tables(lv.maxsteps);
    lv.cap.assign(ntab, 0);
    for (int t = 0; t < ntab; ++t)
      for (int li = 0; li < nl; ++li)
        lv.cap[t] += level_table_count(t, lv.maxsteps, mc[li], nc[li], sc[li],
                                       P.tmpl[s0 + li].is_root);
    lv.tab_off = tab_total;
    tab_total += (size_t)ntab * (nl + 1);
    // this level's fronts have consumed their children's blocks
    for (int sl = s0; sl < s1; ++sl) {
      for (int c : S.children[S.node_of[sl]]) {
        const int cs = S.slot[c];
        N.pool.release(cof[cs].first, cof[cs].second);
        cof[cs] = {nullptr, 0};
      }
      N.pool.release(wof[sl].first, wof[sl].second);
    }
  }
  for (int sl = 0; sl < nn; ++sl)  // roots keep their contribution block
    if (cof[sl].first) P.allocs.push_back(cof[sl]);
  if (!big.empty()) {
    P.arena_bytes = arena_place(big);
    HIPCHK(hipMalloc(&P.arena, P.arena_bytes));
    for (const ArenaBlock &a : big) {
      NodeDev &D = P.tmpl[a.slot];
      const int n = D.sn + D.ndcap, m = D.sm + D.ndcap;
      const int nsteps = pivot_tpp ? 0 : cdiv(n, kBlk);
      char *base = (char *)P.arena + a.off;
      if (a.kind == 0) {
        size_t cbytes, ldbytes, wb;
        trans_sizes(N, D, n, nsteps, m, cbytes, ldbytes, wb);
        set_contrib(D, base, cbytes, ldbytes);
      } else {
        set_work(D, base, nsteps);
      }
    }
  }
  if (!P.d_tmpl) HIPCHK(hipMalloc(&P.d_tmpl, std::max(nn, 1) * sizeof(NodeDev)));
  copy_now(P.d_tmpl, P.tmpl.data(), nn * sizeof(NodeDev), hipMemcpyHostToDevice, N.stream);
  if (tab_total > P.tabs_cap) {
    if (P.d_tabs) HIPCHK(hipFree(P.d_tabs));
    HIPCHK(hipMalloc(&P.d_tabs, std::max<size_t>(tab_total, 1) * sizeof(int)));
    P.tabs_cap = tab_total;
  }
  if (!P.d_pack) HIPCHK(hipMalloc(&P.d_pack, 8 * (size_t)std::max(nn, 1) * sizeof(int)));
  P.valid = true;
}

static bool graphs_disabled() {
  const char *e = std::getenv("SPRAL_SSIDS_HIP_NO_GRAPH");
  return e && *e && *e != '0';
}

// One factorization on the static plan: all levels launched back to back, one
// read-back at the end.  Returns false when some front's delays exceeded its
// capacity (its results and its ancestors' are then void).
static bool factor_static(Numeric &N, const FactorParams &fp, const spral_ssids_hip_options &opt,
                          std::vector<hipEvent_t> &tev) {
  Symbolic &S = *N.symb;
  Plan &P = N.plan;
  const int nn = S.nnodes, nlev = (int)S.lvl_ptr.size() - 1;
  hipStream_t st = N.stream;
  int *overflow = N.d_counters + 8;
  // The gather of a level is split: its first tile column (the first pivot
  // block and the rows below it) on `st`, the other tiles on `st2`.  The
  // first step's pivot block and apply touch only the first 64 columns, so
  // they overlap the rest of the gather; `st` joins `st2` before the first
  // trailing update.  A timing run (opt.timing) launches the same sequence
  // eagerly, with events on `st2` around each level's rest-of-level gather.
  if (!N.stream2) HIPCHK(hipStreamCreateWithFlags(&N.stream2, hipStreamNonBlocking));
  if (!N.ev_level) HIPCHK(hipEventCreateWithFlags(&N.ev_level, hipEventDisableTiming));
  if (!N.ev_rest) HIPCHK(hipEventCreateWithFlags(&N.ev_rest, hipEventDisableTiming));
  if (!N.ev_first) HIPCHK(hipEventCreateWithFlags(&N.ev_first, hipEventDisableTiming));
  hipStream_t st2 = N.stream2;
  // the whole launch sequence (no host round trip, no memset: the counters
  // are cleared before it), captured once as a graph when not timing
  auto enqueue = [&](hipStream_t st) {
  HIPCHK(hipMemcpyAsync(N.d_nodes, P.d_tmpl, nn * sizeof(NodeDev), hipMemcpyDeviceToDevice,
                        st));
  for (int L = 0; L < nlev; ++L) {
    const int s0 = S.lvl_ptr[L], nl = S.lvl_ptr[L + 1] - s0;
    const Plan::Level &lv = P.lv[L];
    const int ms = lv.maxsteps;
    int *tab = P.d_tabs + lv.tab_off;
    auto toff = [&](int t) { return (const int *)(tab + (size_t)t * (nl + 1)); };
    const int *d_slots = N.d_slots_all + s0;
    launch_level_setup(st, N.d_nodes, s0, nl, S.d_child, fp.pivot_tpp, ms, tab, overflow);
    launch_maptab(st, nl, N.d_nodes, d_slots, S.d_child);
    HIPCHK(hipEventRecord(N.ev_level, st));
    HIPCHK(hipStreamWaitEvent(st2, N.ev_level, 0));
    // the first tile column's other tiles on st2, ahead of the rest
    launch_gather(st2, 3, lv.cap[2 * ms + 5], N.d_nodes, d_slots, toff(2 * ms + 5), nl,
                  S.d_child, N.d_aval, S.d_asrc, S.d_arow, S.d_acix, N.d_scal, fp.c32);
    HIPCHK(hipEventRecord(N.ev_first, st2));
    // the diagonal tiles, each with step 0's pivot block
    launch_gather_first(st, lv.cap[2 * ms + 3], N.d_nodes, d_slots, toff(2 * ms + 3), nl,
                        S.d_child, N.d_aval, S.d_asrc, S.d_arow, S.d_acix, N.d_scal, fp);
    if (opt.timing) tev.emplace_back(timed_event(st2));
    launch_gather(st2, 2, lv.cap[2 * ms + 4], N.d_nodes, d_slots, toff(2 * ms + 4), nl,
                  S.d_child, N.d_aval, S.d_asrc, S.d_arow, S.d_acix, N.d_scal, fp.c32);
    if (opt.timing) tev.emplace_back(timed_event(st2));
    HIPCHK(hipEventRecord(N.ev_rest, st2));
    bool joined = false;
    auto join = [&]() {
      if (!joined) HIPCHK(hipStreamWaitEvent(st, N.ev_rest, 0));
      joined = true;
    };
    for (int stp = 0; stp < ms; ++stp) {
      if (stp > 0) launch_diag(st, nl, N.d_nodes, d_slots, stp, fp);
      // step 0's apply reads the whole first tile column
      if (stp == 0) HIPCHK(hipStreamWaitEvent(st, N.ev_first, 0));
      launch_apply(st, lv.cap[1 + stp], N.d_nodes, d_slots, toff(1 + stp), nl, stp, fp);
      // one step (every n <= 64): the update stays inside the first tile
      // column too; with more steps it reaches the later columns
      if (ms > 1) join();
      launch_update(st, lv.cap[ms + 2 + stp], N.d_nodes, d_slots, toff(ms + 2 + stp), nl, stp,
                    0, fp);
      if (lv.multi) {
        launch_swap(st, lv.cap[ms + 1], N.d_nodes, d_slots, toff(ms + 1), nl, stp, 0);
        launch_swap(st, lv.cap[ms + 1], N.d_nodes, d_slots, toff(ms + 1), nl, stp, 1);
      }
    }
    join();
    launch_tail(st, nl, N.d_nodes, d_slots, fp, N.d_scratch_n);
    if (lv.cap[2 * ms + 2] > 0)
      launch_update(st, lv.cap[2 * ms + 2], N.d_nodes, d_slots, toff(2 * ms + 2), nl, 0, 1, fp);
  }
  launch_pack_nodes(st, N.d_nodes, nn, P.d_pack);
  };
  if (!opt.timing && !graphs_disabled()) {
    // a graph is valid for one plan, one set of pivot parameters and the
    // same value and scaling buffers
    const std::vector<double> key = {(double)P.gen, fp.u, fp.small, (double)fp.action,
                                     (double)fp.tpp_failed, (double)fp.pivot_tpp,
                                     (double)fp.posdef, (double)fp.c32,
                                     (double)(uintptr_t)N.d_aval, (double)(uintptr_t)N.d_scal};
    bool captured = true;
    if (!N.gexec || N.gkey != key) {
      if (N.gexec) HIPCHK(hipGraphExecDestroy(N.gexec));
      if (N.graph) HIPCHK(hipGraphDestroy(N.graph));
      N.gexec = nullptr;
      N.graph = nullptr;
      // thread-local, on the private non-blocking stream, while no other
      // host thread is inside the library; the graph is launched on `st`
      if (!N.stream_cap) HIPCHK(hipStreamCreateWithFlags(&N.stream_cap, hipStreamNonBlocking));
      hipStream_t cs = N.stream_cap;
      CaptureExclusive only_caller;
      HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      try {
        enqueue(cs);
      } catch (HipError &e) {
        ce = e.code;
      } catch (...) {
        hipGraph_t g = nullptr;  // leave capture mode before reporting
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hipGraph_t g = nullptr;
      const hipError_t ee = hipStreamEndCapture(cs, &g);
      if (ce == hipSuccess) ce = ee;
      if (ce == hipSuccess) {
        N.graph = g;
        HIPCHK(hipGraphInstantiate(&N.gexec, N.graph, nullptr, nullptr, 0));
        N.gkey = key;
      } else {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        // a HIP call made outside the library by another host thread can
        // invalidate a capture: nothing captured has run, so this call
        // launches the sequence eagerly and the next one captures again
        if (ce != hipErrorStreamCaptureInvalidated && ce != hipErrorStreamCaptureImplicit &&
            ce != hipErrorStreamCaptureUnsupported)
          throw HipError(ce);
        captured = false;
      }
    }
    if (captured) HIPCHK(hipGraphLaunch(N.gexec, st));
    else enqueue(st);
  } else {
    enqueue(st);
  }
  HIPCHK(hipGetLastError());
  std::vector<int> pack(8 * (size_t)nn);
  int ovf = 0;
  HIPCHK(hipMemcpyAsync(pack.data(), P.d_pack, pack.size() * sizeof(int),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&ovf, overflow, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (ovf > 0) return false;
  N.nodes = P.tmpl;
  for (int sl = 0; sl < nn; ++sl) {
    NodeDev &D = N.nodes[sl];
    const int *o = &pack[8 * (size_t)sl];
    D.nd = o[0];
    D.m = o[1];
    D.n = o[2];
    D.nelim = o[3];
    D.nelim_app = o[4];
    D.flag = o[5];
    D.skip = o[6];
    D.nsteps = o[7];
  }
  return true;
}

// One factorization level by level, sizing each level's fronts on the host
// from the read-back of the levels below (exact sizes, any number of delays).
static void factor_dynamic(Numeric &N, const FactorParams &fp,
                           const spral_ssids_hip_options &opt, std::vector<hipEvent_t> &tev) {
  Symbolic &S = *N.symb;
  const int nn = S.nnodes, nlev = (int)S.lvl_ptr.size() - 1;
  hipStream_t st = N.stream;
  for (auto &p : N.persistent) N.pool.release(p.first, p.second);
  N.persistent.clear();
  N.nodes.assign(nn, NodeDev{});
  std::vector<std::pair<void *, size_t>> contrib_of(nn), work_of(nn);
  auto keep = [&](size_t bytes) {
    void *p = N.pool.alloc(bytes);
    N.persistent.push_back({p, bytes});
    return p;
  };
  for (int L = 0; L < nlev; ++L) {
    const int s0 = S.lvl_ptr[L], s1 = S.lvl_ptr[L + 1], nl = s1 - s0;
    int maxsteps = 0;
    bool multi = false;
    for (int sl = s0; sl < s1; ++sl) {
      const int i = S.node_of[sl];
      NodeDev &D = N.nodes[sl];
      D = NodeDev{};
      node_static(S, sl, D);
      int nd = 0;
      for (int c : S.children[i]) {
        const NodeDev &C = N.nodes[S.slot[c]];
        nd += C.n - C.nelim;
      }
      D.nd = nd;
      D.ndcap = nd;
      D.m = D.sm + nd;
      D.n = D.sn + nd;
      D.nsteps = fp.pivot_tpp ? 0 : cdiv(D.n, kBlk);
      maxsteps = std::max(maxsteps, D.nsteps);
      if (D.nsteps > 1) multi = true;
      node_buffers(N, S, sl, D, D.m, D.n, D.nsteps, keep,
                   [&](size_t b, int) { return N.pool.alloc(b); }, contrib_of[sl],
                   work_of[sl]);
    }
    HIPCHK(hipMemcpyAsync(N.d_nodes + s0, N.nodes.data() + s0, nl * sizeof(NodeDev),
                          hipMemcpyHostToDevice, st));
    // per-level launch tables (exact sizes): slots | tables of level_table_count
    const int ntab = level_tables(maxsteps);
    std::vector<int> tab;
    tab.reserve(nl + (size_t)ntab * (nl + 1));
    for (int sl = s0; sl < s1; ++sl) tab.push_back(sl);
    std::vector<int> tot(ntab, 0);
    for (int t = 0; t < ntab; ++t) {
      int acc = 0;
      for (int li = 0; li < nl; ++li) {
        tab.push_back(acc);
        const NodeDev &D = N.nodes[s0 + li];
        acc += level_table_count(t, maxsteps, D.m, D.n, D.nsteps, D.is_root);
      }
      tab.push_back(acc);
      tot[t] = acc;
    }
    ensure_lvl(N, tab.size());
    HIPCHK(hipMemcpyAsync(N.d_lvl, tab.data(), tab.size() * sizeof(int),
                          hipMemcpyHostToDevice, st));
    const int *d_slots = N.d_lvl;
    auto toff = [&](int t) { return (const int *)(N.d_lvl + nl + (size_t)t * (nl + 1)); };
    const int ms = maxsteps;
    launch_maptab(st, nl, N.d_nodes, d_slots, S.d_child);
    if (opt.timing) tev.emplace_back(timed_event(st));
    launch_gather(st, 0, tot[0], N.d_nodes, d_slots, toff(0), nl, S.d_child, N.d_aval,
                  S.d_asrc, S.d_arow, S.d_acix, N.d_scal, fp.c32);
    if (opt.timing) tev.emplace_back(timed_event(st));
    for (int stp = 0; stp < ms; ++stp) {
      launch_diag(st, nl, N.d_nodes, d_slots, stp, fp);
      launch_apply(st, tot[1 + stp], N.d_nodes, d_slots, toff(1 + stp), nl, stp, fp);
      launch_update(st, tot[ms + 2 + stp], N.d_nodes, d_slots, toff(ms + 2 + stp), nl, stp, 0,
                    fp);
      if (multi) {
        launch_swap(st, tot[ms + 1], N.d_nodes, d_slots, toff(ms + 1), nl, stp, 0);
        launch_swap(st, tot[ms + 1], N.d_nodes, d_slots, toff(ms + 1), nl, stp, 1);
      }
    }
    launch_tail(st, nl, N.d_nodes, d_slots, fp, N.d_scratch_n);
    if (tot[2 * ms + 2] > 0)
      launch_update(st, tot[2 * ms + 2], N.d_nodes, d_slots, toff(2 * ms + 2), nl, 0, 1, fp);
    HIPCHK(hipGetLastError());
    // read back the level's results (nelim, flag)
    std::vector<NodeDev> back(nl);
    HIPCHK(hipMemcpyAsync(back.data(), N.d_nodes + s0, nl * sizeof(NodeDev),
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int bad = 0;
    for (int li = 0; li < nl; ++li) {
      NodeDev &D = N.nodes[s0 + li];
      D.nelim = back[li].nelim;
      D.nelim_app = back[li].nelim_app;
      D.flag = back[li].flag;
      if (back[li].flag < 0) bad = back[li].flag;
    }
    // children contributions have been assembled: release them and the
    // workspaces of this level
    for (int sl = s0; sl < s1; ++sl) {
      for (int c : S.children[S.node_of[sl]]) {
        const int cs = S.slot[c];
        N.pool.release(contrib_of[cs].first, contrib_of[cs].second);
        contrib_of[cs] = {nullptr, 0};
      }
      N.pool.release(work_of[sl].first, work_of[sl].second);
      work_of[sl] = {nullptr, 0};
    }
    if (bad < 0) {
      N.stats.flag = bad;
      break;
    }
  }
  // roots keep their contribution block (exported by get_contrib)
  for (int sl = 0; sl < nn; ++sl)
    if (contrib_of[sl].first) N.persistent.push_back(contrib_of[sl]);
  if ((int)N.nd_seen.size() != nn) N.nd_seen.assign(nn, 0);
  for (int sl = 0; sl < nn; ++sl) N.nd_seen[sl] = std::max(N.nd_seen[sl], N.nodes[sl].nd);
}

static bool plan_disabled() {
  const char *e = std::getenv("SPRAL_SSIDS_HIP_NO_PLAN");
  return e && *e && *e != '0';
}

// Solve descriptors and launch tables of a factorization (the factors in
// N.nodes): per-front solve views, row-workgroup tables, the workspace
// layout and the big-front level tables.
static void build_solve_tables(Numeric &N) {
  Symbolic &S = *N.symb;
  const int nn = S.nnodes;
  N.sn_h.assign(nn, SolveNode{});
  std::vector<int> doff(nn, 0);
  for (int i = 0; i < nn; ++i) {
    int acc = 0;
    for (int c : S.children[i]) {
      int cs = S.slot[c];
      doff[cs] = acc;
      acc += N.nodes[cs].n - N.nodes[cs].nelim;
    }
  }
  for (int sl = 0; sl < nn; ++sl) {
    const NodeDev &D = N.nodes[sl];
    SolveNode &Q = N.sn_h[sl];
    Q.L = D.L;
    Q.D = D.D;
    Q.perm = D.perm;
    Q.ipos = D.ipos;
    Q.rows = D.rows;
    Q.rmap = D.rmap;
    Q.m = D.m;
    Q.n = D.n;
    Q.nd = D.nd;
    Q.sn = D.sn;
    Q.nelim = D.nelim;
    Q.ldl = D.ldl;
    Q.doff = doff[sl];
    Q.cbeg = D.cbeg;
    Q.cend = D.cend;
    Q.ctab = D.ctab;
  }
  if (!N.d_sn) HIPCHK(hipMalloc(&N.d_sn, std::max(nn, 1) * sizeof(SolveNode)));
  if (!N.d_slots_all) {
    std::vector<int> iota(std::max(nn, 1));
    for (int i = 0; i < nn; ++i) iota[i] = i;
    HIPCHK(hipMalloc(&N.d_slots_all, iota.size() * sizeof(int)));
    copy_now(N.d_slots_all, iota.data(), iota.size() * sizeof(int),
                     hipMemcpyHostToDevice, N.stream);
  }
  // solve row-workgroup tables (depend on nelim): every front gets at least
  // one, which in the backward solve also reduces and solves the triangle
  {
    const int nlev2 = (int)S.lvl_ptr.size() - 1;
    std::vector<int> tab;
    N.gemv_tab_off.assign(nlev2, 0);
    N.gemv_total.assign(nlev2, 0);
    N.bwd_tab_off.assign(nlev2, 0);
    N.bwd_total.assign(nlev2, 0);
    N.bwd_count.assign(nn, 0);
    N.bwdc_tab_off.assign(nlev2, 0);
    N.bwdc_total.assign(nlev2, 0);
    for (int L = 0; L < nlev2; ++L) {
      N.gemv_tab_off[L] = (int)tab.size();
      int acc = 0;
      for (int sl = S.lvl_ptr[L]; sl < S.lvl_ptr[L + 1]; ++sl) {
        tab.push_back(acc);
        acc += cdiv(N.nodes[sl].m - N.nodes[sl].nelim, fwd_rows_per_wg());
      }
      tab.push_back(acc);
      N.gemv_total[L] = acc;
      N.bwd_tab_off[L] = (int)tab.size();
      acc = 0;
      for (int sl = S.lvl_ptr[L]; sl < S.lvl_ptr[L + 1]; ++sl) {
        tab.push_back(acc);
        const int c = cdiv(N.nodes[sl].m - N.nodes[sl].nelim, bwd_rows_per_wg());
        N.bwd_count[sl] = c;
        acc += c;
      }
      tab.push_back(acc);
      N.bwd_total[L] = acc;
      N.bwdc_tab_off[L] = (int)tab.size();
      acc = 0;
      for (int sl = S.lvl_ptr[L]; sl < S.lvl_ptr[L + 1]; ++sl) {
        tab.push_back(acc);
        acc += N.bwd_count[sl] * cdiv(std::max(N.nodes[sl].nelim, 1), bwd_cols_per_wg());
      }
      tab.push_back(acc);
      N.bwdc_total[L] = acc;
    }
    if (N.d_gemv_tab) HIPCHK(hipFree(N.d_gemv_tab));
    N.d_gemv_tab = nullptr;
    HIPCHK(hipMalloc(&N.d_gemv_tab, std::max<size_t>(tab.size(), 1) * sizeof(int)));
    copy_now(N.d_gemv_tab, tab.data(), tab.size() * sizeof(int),
                     hipMemcpyHostToDevice, N.stream);
  }
  // per front, per right-hand side: X (m) and the backward partial sums
  // (row workgroups x nelim, at least nelim: big fronts keep their
  // triangle's solution there)
  {
    int64_t off = 0;
    for (int sl = 0; sl < nn; ++sl) {
      N.sn_h[sl].woff = off;
      off += (int64_t)N.nodes[sl].m +
             (int64_t)std::max(N.bwd_count[sl], 1) * std::max(N.nodes[sl].nelim, 1);
    }
    N.solve_ws_per_rhs = std::max<int64_t>(off, 1);
  }
  copy_now(N.d_sn, N.sn_h.data(), nn * sizeof(SolveNode), hipMemcpyHostToDevice, N.stream);
  N.big_of_level.assign(S.lvl_ptr.size() - 1, {});
  for (int L = 0; L + 1 < (int)S.lvl_ptr.size(); ++L)
    for (int sl = S.lvl_ptr[L]; sl < S.lvl_ptr[L + 1]; ++sl)
      if (N.nodes[sl].nelim > big_tri_threshold()) N.big_of_level[L].push_back(sl);
  {
    // per level [block offsets (nbig+1) | slots (nbig) | sync offsets (nbig)];
    // sync words per big front: forward ticket + nb flags, backward the same
    std::vector<int> tab;
    const int nlev2 = (int)S.lvl_ptr.size() - 1;
    N.bigtab_off.assign(nlev2, 0);
    N.big_blocks.assign(nlev2, 0);
    int64_t sync = 0;
    for (int L = 0; L < nlev2; ++L) {
      const auto &bl = N.big_of_level[L];
      N.bigtab_off[L] = (int)tab.size();
      int acc = 0;
      for (int sl : bl) {
        tab.push_back(acc);
        acc += cdiv(N.nodes[sl].nelim, kBlk);
      }
      tab.push_back(acc);
      N.big_blocks[L] = acc;
      for (int sl : bl) tab.push_back(sl);
      for (int sl : bl) {
        tab.push_back((int)sync);
        sync += 2 * (cdiv(N.nodes[sl].nelim, kBlk) + 1);
      }
    }
    N.solve_sync_ints = sync;
    if (N.d_bigtab) HIPCHK(hipFree(N.d_bigtab));
    N.d_bigtab = nullptr;
    if (!tab.empty()) {
      HIPCHK(hipMalloc(&N.d_bigtab, tab.size() * sizeof(int)));
      copy_now(N.d_bigtab, tab.data(), tab.size() * sizeof(int),
                       hipMemcpyHostToDevice, N.stream);
    }
  }
}

static void factor_impl(Numeric &N, const double *d_aval, const double *d_scaling,
                        const spral_ssids_hip_options &opt, int64_t naval) {
  Symbolic &S = *N.symb;
  // Schur-update precision: 0 FP64, 1 FP32 (mixed); FP16 is not built
  if (opt.schur_precision < 0 || opt.schur_precision > 1) {
    N.stats.flag = -98;  // SSIDS_ERROR_UNIMPLEMENTED
    return;
  }
  N.c32 = (opt.schur_precision == 1);
  // keep the values (and scaling) for refinement residuals
  if ((size_t)naval > N.d_aval_cap) {
    if (N.d_aval) HIPCHK(hipFree(N.d_aval));
    HIPCHK(hipMalloc(&N.d_aval, std::max<int64_t>(naval, 1) * sizeof(double)));
    N.d_aval_cap = naval;
  }
  if (naval > 0 && d_aval != N.d_aval)
    HIPCHK(hipMemcpyAsync(N.d_aval, d_aval, naval * sizeof(double), hipMemcpyDeviceToDevice,
                          N.stream));
  if (d_scaling) {
    if (!N.d_scal) HIPCHK(hipMalloc(&N.d_scal, std::max(S.n, 1) * sizeof(double)));
    HIPCHK(hipMemcpyAsync(N.d_scal, d_scaling, S.n * sizeof(double), hipMemcpyDeviceToDevice,
                          N.stream));
  } else if (N.d_scal) {
    HIPCHK(hipFree(N.d_scal));
    N.d_scal = nullptr;
  }
  N.anorm = -1.0;
  const int nn = S.nnodes;
  hipStream_t st = N.stream;
  FactorParams fp;
  fp.u = opt.u;
  fp.small = opt.small;
  fp.action = opt.action;
  fp.tpp_failed = (opt.failed_pivot_method != 2);
  fp.pivot_tpp = (opt.pivot_method == 3) && !N.posdef;
  fp.posdef = N.posdef;
  fp.c32 = N.c32;
  if (!N.d_nodes) HIPCHK(hipMalloc(&N.d_nodes, std::max(nn, 1) * sizeof(NodeDev)));
  if (!N.d_counters) HIPCHK(hipMalloc(&N.d_counters, 16 * sizeof(int)));
  if (!N.d_scratch_n) HIPCHK(hipMalloc(&N.d_scratch_n, std::max(S.n, 1) * sizeof(int)));
  if (!N.d_slots_all) {
    std::vector<int> iota(std::max(nn, 1));
    for (int i = 0; i < nn; ++i) iota[i] = i;
    HIPCHK(hipMalloc(&N.d_slots_all, iota.size() * sizeof(int)));
    copy_now(N.d_slots_all, iota.data(), iota.size() * sizeof(int),
                     hipMemcpyHostToDevice, N.stream);
  }
  HIPCHK(hipMemsetAsync(N.d_counters, 0, 16 * sizeof(int), st));
  std::vector<hipEvent_t> tev;  // pairs around the gather launches
  hipEvent_t ev0, ev1;
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));
  HIPCHK(hipEventRecord(ev0, st));
  bool done = false;
  if (!plan_disabled()) {
    if (!N.plan.valid || N.plan.pivot_tpp != fp.pivot_tpp || N.plan.c32 != N.c32)
      plan_build(N, fp.pivot_tpp, opt.multiplier);
    done = factor_static(N, fp, opt, tev);
    if (done) {
      // the delays of every factorization inform the next plan
      for (int sl = 0; sl < nn; ++sl)
        N.nd_seen[sl] = std::max(N.nd_seen[sl], N.nodes[sl].nd);
    } else {
      for (auto e : tev) (void)hipEventDestroy(e);
      tev.clear();
      plan_release(N);  // rebuilt with the delays seen, on the next factorization
    }
  }
  if (!done) factor_dynamic(N, fp, opt, tev);
  const int nlev = (int)S.lvl_ptr.size() - 1;
  int maxfront = 0;
  for (int sl = 0; sl < nn; ++sl) maxfront = std::max(maxfront, N.nodes[sl].m);
  launch_stats(st, N.d_nodes, nn, N.d_counters, fp);
  int cnt[16];
  HIPCHK(hipMemcpyAsync(cnt, N.d_counters, 16 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(ev1, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
  HIPCHK(hipEventDestroy(ev0));
  HIPCHK(hipEventDestroy(ev1));
  spral_ssids_hip_stats &o = N.stats;
  if (o.flag >= 0) {
    o.num_neg = cnt[0];
    o.num_two = cnt[1];
    o.num_zero = cnt[2];
    o.num_delay = cnt[3];
    o.not_first_pass = cnt[4];
    o.not_second_pass = cnt[7];
    o.flag = (cnt[2] > 0) ? 7 : 0;  // SSIDS_WARNING_FACT_SINGULAR
    if (cnt[5] > 0) o.flag = -5;
    if (cnt[6] > 0) o.flag = -6;  // SSIDS_ERROR_NOT_POS_DEF
  }
  int64_t nf = 0, fl = 0;
  int maxsn = 0;
  for (int sl = 0; sl < nn; ++sl) {
    const NodeDev &D = N.nodes[sl];
    for (int64_t j = D.m; j >= D.m - D.nelim + 1; --j) {
      nf += j;
      fl += j * j;
    }
    maxsn = std::max(maxsn, S.sn[S.node_of[sl]]);
  }
  o.num_factor = nf;
  o.num_flops = fl;
  o.maxfront = maxfront;
  o.maxsupernode = maxsn;
  o.num_levels = nlev;
  o.factor_ms = ms;
  o.pool_peak_bytes = (int64_t)(N.pool.peak() + N.plan.arena_bytes);
  o.static_plan = done ? 1 : 0;
  o.kernel_ms = 0.0;
  o.kernel_launches = 0;
  for (size_t k = 0; k + 1 < tev.size(); k += 2) {
    float e = 0.f;
    HIPCHK(hipEventElapsedTime(&e, tev[k], tev[k + 1]));
    o.kernel_ms += e;
    o.kernel_launches++;
  }
  for (auto e : tev) (void)hipEventDestroy(e);
  // algorithmic work of the gather kernels: each child's Schur complement
  // C_c - L21 (L21 D)^T on the lower triangle (K*N*(N+1) flops, K = nelim,
  // N = m-n), and HBM bytes = parent fronts written once (lower part) +
  // children's contribution blocks (lower), L21 and L21 D panels and
  // delayed columns read once.
  // On the static plan the timed launches are the rest-of-level gathers
  // (k_gather part 2: every parent tile outside the first 64-column tile
  // column), so only that part is counted: a child's contribution entry
  // (i, j), i >= j, lands there when its parent column P(j) >= 64 (P is
  // monotone, so that is every j from the first such j*).
  double gf = 0.0, gb = 0.0;
  const bool part2 = done;
  for (int sl = 0; sl < nn; ++sl) {
    const NodeDev &D = N.nodes[sl];
    double m = D.m, n = D.n;
    if (!part2) {
      gb += 8.0 * (n * (n + 1) / 2 + (m - n) * n + (m - n) * (m - n + 1) / 2);  // written
    } else if (m > kBlk) {
      gb += 8.0 * (m - kBlk) * (m - kBlk + 1) / 2;  // written, columns >= 64
    }
    if (D.parent < 0) continue;
    const double K = D.nelim;
    if (!part2) {
      const double Nc = m - n;
      gf += K * Nc * (Nc + 1);
      gb += 8.0 * (Nc * (Nc + 1) / 2 + 2 * Nc * K + (n - D.nelim) * m);  // read
      continue;
    }
    const int node = S.node_of[sl], pn = N.nodes[D.parent].nd;
    const int Nc = D.m - D.n;
    int js = Nc;  // first contribution index with parent column >= 64
    for (int j = 0; j < Nc; ++j)
      if (pn + S.rmap_h[S.rmap_off[node] + j] >= kBlk) {
        js = j;
        break;
      }
    const double r = Nc - js;
    gf += K * r * (r + 1);
    gb += 8.0 * (r * (r + 1) / 2 + 2 * r * K);  // C and the L21 / L21 D rows read
    int ndl = 0;  // delayed columns landing at parent column >= 64 (copied)
    for (int k = D.nelim; k < D.n; ++k) ndl += (D.doff + (k - D.nelim) >= kBlk) ? 1 : 0;
    gb += 8.0 * ndl * m;
  }
  o.kernel_flops = gf;
  o.kernel_bytes = gb;
  build_solve_tables(N);
}

// Doubles of solve workspace for nrhs right-hand sides: the per-front X and
// partial sums, then the big fronts' sync words.
static size_t solve_ws_need(const Numeric &N, int nrhs) {
  return (size_t)N.solve_ws_per_rhs * nrhs + (size_t)(N.solve_sync_ints + 1) / 2;
}

// The handle's own solve workspace, grown to nrhs right-hand sides.
static double *solve_ws(Numeric &N, int nrhs) {
  const size_t need = solve_ws_need(N, nrhs);
  if (need > N.d_X_cap) {
    if (N.d_X) HIPCHK(hipFree(N.d_X));
    N.d_X = nullptr;
    N.d_X_cap = 0;
    HIPCHK(hipMalloc(&N.d_X, need * sizeof(double)));
    N.d_X_cap = need;
  }
  return N.d_X;
}

// One solve job on device-resident x with workspace ws (at least
// solve_ws_per_rhs x nrhs doubles).  Launches only: no allocation.
static void solve_launch(Numeric &N, int job, int nrhs, double *d_x, int ldx, hipStream_t st,
                         double *ws) {
  Symbolic &S = *N.symb;
  const int nlev = (int)S.lvl_ptr.size() - 1;
  // positive definite, as the CPU subtree (src/ssids/cpu/NumericSubtree.hxx:
  // 342-343): the diagonal solve alone is a no-op, diag + bwd is bwd
  if (N.posdef && job == 2) return;
  if (N.posdef && job == 4) job = 3;
  // the slot list of all levels is simply 0..nnodes-1 in level order
  const int *slots = N.d_slots_all;
  bool fwd = (job == 0 || job == 1), bwd = (job == 0 || job == 3 || job == 4);
  bool diag = (job == 0 || job == 2 || job == 4);
  // positive definite: L_chol = L D^{1/2}, so the forward (job 1) and
  // backward (job 3) halves of L L^T carry D^{-1/2} on their inner side
  const bool half_after_fwd = N.posdef && job == 1, half_before_bwd = N.posdef && job == 3;
  int *sync = reinterpret_cast<int *>(ws + (size_t)N.solve_ws_per_rhs * nrhs);
  if (fwd || bwd) launch_big_sync_clear(st, sync, (int)N.solve_sync_ints);
  if (fwd)
    for (int L = 0; L < nlev; ++L) {
      const int nl = S.lvl_ptr[L + 1] - S.lvl_ptr[L];
      launch_fwd_tri(st, nl, N.d_sn, slots + S.lvl_ptr[L], S.d_child, d_x, ldx, nrhs, ws);
      launch_big_fwd_lvl(st, N.d_sn, N.d_bigtab + N.bigtab_off[L], (int)N.big_of_level[L].size(),
                         N.big_blocks[L], d_x, ldx, nrhs, ws, sync);
      launch_fwd_rows(st, nl, N.d_sn, slots + S.lvl_ptr[L], S.d_child, nrhs,
                      N.d_gemv_tab + N.gemv_tab_off[L], N.gemv_total[L], ws);
    }
  // with a backward solve after it, D^{-1} is applied per front inside the
  // backward kernels (a front's pivots only mix its own positions)
  const int dfuse = (diag && bwd && !N.posdef) ? 1 : 0;
  if (diag && !dfuse) launch_dsolve(st, S.nnodes, N.d_sn, d_x, ldx, nrhs, 0);
  if (half_after_fwd || half_before_bwd)
    launch_dsolve(st, S.nnodes, N.d_sn, d_x, ldx, nrhs, 1);
  if (bwd)
    for (int L = nlev - 1; L >= 0; --L) {
      launch_bwd(st, S.lvl_ptr[L + 1] - S.lvl_ptr[L], N.d_sn, slots + S.lvl_ptr[L], d_x,
                 ldx, nrhs, N.d_gemv_tab + N.bwd_tab_off[L], N.d_gemv_tab + N.bwdc_tab_off[L],
                 N.bwdc_total[L], ws, dfuse);
      launch_big_bwd_lvl(st, N.d_sn, N.d_bigtab + N.bigtab_off[L], (int)N.big_of_level[L].size(),
                         N.big_blocks[L], d_x, ldx, nrhs, ws, sync);
    }
  HIPCHK(hipGetLastError());
}

// Workspace of the allocation-free solve entry point: a file-scope device
// array, one copy per GPU the library runs on, held until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress with
// that device current); host code passes only that address, never the
// array's name, to kernels.  The solve kernels write every workspace entry
// they read earlier in the same solve, so nothing needs clearing between
// calls.  Calls that overlap on one GPU share the array and race.
constexpr size_t kScratchDoubles = size_t(8) << 20;  // 64 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
static std::mutex scratch_mu;
static double *scratch_addr[kMaxScratchDevices];

static double *scratch_on(int device) {
  if (device < 0 || device >= kMaxScratchDevices) throw HipError(hipErrorInvalidDevice);
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_addr[device]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_addr[device] = static_cast<double *>(p);
  }
  return scratch_addr[device];
}

static void solve_device(Numeric &N, int job, int nrhs, double *d_x, int ldx,
                         hipStream_t st) {
  solve_launch(N, job, nrhs, d_x, ldx, st, solve_ws(N, nrhs));
}

static int solve_host(Numeric &N, int job, int nrhs, double *x, int ldx) {
  Symbolic &S = *N.symb;
  if (nrhs < 1 || ldx < S.n) return -10;
  HIPCHK(hipSetDevice(S.device));
  size_t need = (size_t)S.n * nrhs;
  if (need > N.d_x_cap) {
    if (N.d_x) HIPCHK(hipFree(N.d_x));
    HIPCHK(hipMalloc(&N.d_x, need * sizeof(double)));
    N.d_x_cap = need;
  }
  HIPCHK(hipMemcpy2DAsync(N.d_x, S.n * sizeof(double), x, ldx * sizeof(double),
                          S.n * sizeof(double), nrhs, hipMemcpyHostToDevice, N.stream));
  solve_device(N, job, nrhs, N.d_x, S.n, N.stream);
  HIPCHK(hipMemcpy2DAsync(x, ldx * sizeof(double), N.d_x, S.n * sizeof(double),
                          S.n * sizeof(double), nrhs, hipMemcpyDeviceToHost, N.stream));
  HIPCHK(hipStreamSynchronize(N.stream));
  return 0;
}

// Solve A x = b (scaled factorization S A S when scaling was given) with FP64
// iterative refinement against the original values: x0 = solve(b),
// x_{k+1} = x_k + solve(b - A x_k) until the scaled backward error of every
// column is <= tol or maxit corrections were made.  bwd[k] and *iters report
// the final state.  All vectors are device, in elimination order, ld >= n.
static void solve_scaled(Numeric &N, int nrhs, double *d_v, int ld, hipStream_t st) {
  Symbolic &S = *N.symb;
  if (N.d_scal) launch_scale_rows(st, S.n, nrhs, ld, N.d_scal, d_v);
  solve_device(N, 0, nrhs, d_v, ld, st);
  if (N.d_scal) launch_scale_rows(st, S.n, nrhs, ld, N.d_scal, d_v);
}

static int refine_device(Numeric &N, int nrhs, const double *d_b, double *d_x, int ld,
                         int maxit, double tol, double *bwd, int *iters, hipStream_t st) {
  Symbolic &S = *N.symb;
  const int n = S.n;
  size_t need = (size_t)ld * nrhs;
  if (need > N.d_r_cap) {
    if (N.d_r) HIPCHK(hipFree(N.d_r));
    N.d_r = nullptr;
    N.d_r_cap = 0;
    HIPCHK(hipMalloc(&N.d_r, std::max(need, size_t(1)) * sizeof(double)));
    N.d_r_cap = need;
  }
  // the per-column norms and flags are sized by nrhs alone (a call with a
  // larger ld and fewer columns can fill d_r without them growing)
  if (nrhs > N.d_norms_nrhs) {
    if (N.d_norms) HIPCHK(hipFree(N.d_norms));
    if (N.d_active) HIPCHK(hipFree(N.d_active));
    N.d_norms = nullptr;
    N.d_active = nullptr;
    N.d_norms_nrhs = 0;
    HIPCHK(hipMalloc(&N.d_norms, (1 + 3 * (size_t)nrhs) * sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&N.d_active, nrhs * sizeof(int)));
    N.d_norms_nrhs = nrhs;
  }
  std::vector<unsigned long long> nb(1 + 3 * (size_t)nrhs);
  if (N.anorm < 0) {
    HIPCHK(hipMemsetAsync(N.d_norms, 0, sizeof(unsigned long long), st));
    launch_anorm(st, n, S.d_fptr, S.d_fsrc, N.d_aval, N.d_norms);
    unsigned long long a = 0;
    HIPCHK(hipMemcpyAsync(&a, N.d_norms, sizeof(a), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double v;
    std::memcpy(&v, &a, sizeof(v));
    N.anorm = v;
  }
  if (d_x != d_b)
    HIPCHK(hipMemcpy2DAsync(d_x, ld * sizeof(double), d_b, ld * sizeof(double),
                            n * sizeof(double), nrhs, hipMemcpyDeviceToDevice, st));
  solve_scaled(N, nrhs, d_x, ld, st);
  std::vector<int> active(nrhs, 1);
  int it = 0;
  bool nonfinite = false, converged = false;
  for (;;) {
    launch_residual(st, n, S.d_fptr, S.d_fcol, S.d_fsrc, N.d_aval, d_x, d_b, N.d_r, ld, nrhs);
    HIPCHK(hipMemsetAsync(N.d_norms, 0, (1 + 3 * (size_t)nrhs) * sizeof(unsigned long long), st));
    launch_colmax3(st, n, N.d_r, d_x, d_b, ld, nrhs, N.d_norms + 1);
    HIPCHK(hipMemcpyAsync(nb.data(), N.d_norms, nb.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool all_ok = true;
    nonfinite = false;
    for (int k = 0; k < nrhs; ++k) {
      double rv, xv, bv;
      std::memcpy(&rv, &nb[1 + 3 * k], 8);
      std::memcpy(&xv, &nb[1 + 3 * k + 1], 8);
      std::memcpy(&bv, &nb[1 + 3 * k + 2], 8);
      const double den = N.anorm * xv + bv;
      // the maxima propagate NaN; a non-finite residual, solution or
      // right-hand side is never taken as converged
      if (!std::isfinite(rv) || !std::isfinite(den)) {
        bwd[k] = std::numeric_limits<double>::quiet_NaN();
        nonfinite = true;
      } else {
        bwd[k] = den > 0 ? rv / den : 0.0;
      }
      active[k] = (bwd[k] <= tol) ? 0 : 1;
      if (active[k]) all_ok = false;
    }
    if (all_ok || it >= maxit || nonfinite) break;
    HIPCHK(hipMemcpyAsync(N.d_active, active.data(), nrhs * sizeof(int),
                          hipMemcpyHostToDevice, st));
    solve_scaled(N, nrhs, N.d_r, ld, st);
    launch_axpy(st, n, nrhs, ld, N.d_r, d_x, N.d_active);
    ++it;
  }
  *iters = it;
  converged = true;
  for (int k = 0; k < nrhs; ++k) converged = converged && active[k] == 0;
  if (nonfinite) return SPRAL_SSIDS_HIP_ERROR_NONFINITE;
  return converged ? 0 : SPRAL_SSIDS_HIP_WARNING_NOT_CONVERGED;
}

// ---- serialised factors ------------------------------------------------------
// File layout (native endianness): magic "SSIDSHIP", version, the handle's
// shape (n, nodes, posdef, FP32 contribution blocks, scaling present, values
// count), a fingerprint of the symbolic structure (row lists), the stats,
// the factorized values and scaling (for refinement residuals), then per
// front in slot order its sizes and state, L (ldl x n), D^{-1} (2n), the
// pivoted variable list, the pre-pivot -> post-pivot positions, the forward
// solve's child tile table and, for a root that keeps one, its contribution
// block.
namespace {
constexpr char kMagic[8] = {'S', 'S', 'I', 'D', 'S', 'H', 'I', 'P'};
constexpr int kVersion = 1;

uint64_t symbolic_fingerprint(const Symbolic &S) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over the row lists and sizes
  auto mix = [&](uint64_t v) {
    for (int b = 0; b < 8; ++b) {
      h ^= (v >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  mix((uint64_t)S.n);
  mix((uint64_t)S.nnodes);
  for (int i = 0; i < S.nnodes; ++i) {
    mix((uint64_t)S.sn[i]);
    mix((uint64_t)S.sm[i]);
  }
  for (int v : S.rows_h) mix((uint64_t)(uint32_t)v);
  return h;
}

struct FileW {
  FILE *f;
  bool ok = true;
  template <class T> void put(const T *p, size_t n) {
    if (ok && n && fwrite(p, sizeof(T), n, f) != n) ok = false;
  }
  template <class T> void put1(const T &v) { put(&v, 1); }
};
struct FileR {
  FILE *f;
  bool ok = true;
  template <class T> void get(T *p, size_t n) {
    if (ok && n && fread(p, sizeof(T), n, f) != n) ok = false;
  }
  template <class T> T get1() {
    T v{};
    get(&v, 1);
    return v;
  }
};

size_t ctab_ints(const Symbolic &S, int sl, int m) {
  const int i = S.node_of[sl];
  return (size_t)(cdiv(m, 64) + 1) * std::max<size_t>(S.children[i].size(), 1);
}
}  // namespace

static int save_numeric(const Numeric &N, const char *path) {
  const Symbolic &S = *N.symb;
  HIPCHK(hipSetDevice(S.device));
  HIPCHK(hipStreamSynchronize(N.stream));
  FILE *f = std::fopen(path, "wb");
  if (!f) return -1;
  FileW w{f};
  w.put(kMagic, 8);
  w.put1(kVersion);
  const int hdr[5] = {S.n, S.nnodes, N.posdef ? 1 : 0, N.c32, N.d_scal ? 1 : 0};
  w.put(hdr, 5);
  const int64_t naval = (int64_t)N.d_aval_cap;
  w.put1(naval);
  w.put1(symbolic_fingerprint(S));
  w.put1(N.stats);
  std::vector<double> hv(std::max<int64_t>(naval, S.n));
  if (naval > 0) {
    copy_now(hv.data(), N.d_aval, naval * sizeof(double), hipMemcpyDeviceToHost, N.stream);
    w.put(hv.data(), naval);
  }
  if (N.d_scal) {
    copy_now(hv.data(), N.d_scal, S.n * sizeof(double), hipMemcpyDeviceToHost, N.stream);
    w.put(hv.data(), S.n);
  }
  std::vector<double> dbuf;
  std::vector<int> ibuf;
  std::vector<char> cbuf;
  for (int sl = 0; sl < S.nnodes && w.ok; ++sl) {
    const NodeDev &D = N.nodes[sl];
    const int st[10] = {D.m, D.n, D.nd, D.nelim, D.nelim_app, D.flag, D.nzero, D.nsteps, D.ldl,
                        D.ldc};
    w.put(st, 10);
    dbuf.resize((size_t)D.ldl * D.n + 2 * (size_t)D.n);
    if (D.n > 0) {
      copy_now(dbuf.data(), D.L, (size_t)D.ldl * D.n * sizeof(double),
                       hipMemcpyDeviceToHost, N.stream);
      copy_now(dbuf.data() + (size_t)D.ldl * D.n, D.D, 2 * (size_t)D.n * sizeof(double),
                       hipMemcpyDeviceToHost, N.stream);
      w.put(dbuf.data(), dbuf.size());
      ibuf.resize(2 * (size_t)D.n);
      copy_now(ibuf.data(), D.perm, D.n * sizeof(int), hipMemcpyDeviceToHost, N.stream);
      copy_now(ibuf.data() + D.n, D.ipos, D.n * sizeof(int), hipMemcpyDeviceToHost, N.stream);
      w.put(ibuf.data(), ibuf.size());
    }
    const size_t tn = ctab_ints(S, sl, D.m);
    ibuf.resize(tn);
    copy_now(ibuf.data(), D.ctab, tn * sizeof(int), hipMemcpyDeviceToHost, N.stream);
    w.put(ibuf.data(), tn);
    const int cm = D.m - D.n;
    const size_t cb = (D.is_root && cm > 0 && D.C)
                          ? (size_t)D.ldc * cm * (N.c32 ? sizeof(float) : sizeof(double))
                          : 0;
    w.put1((int64_t)cb);
    if (cb) {
      cbuf.resize(cb);
      copy_now(cbuf.data(), D.C, cb, hipMemcpyDeviceToHost, N.stream);
      w.put(cbuf.data(), cb);
    }
  }
  const bool ok = w.ok && std::fclose(f) == 0;
  return ok ? 0 : -1;
}

static Numeric *load_numeric(Symbolic *Sp, const char *path, int *flag) {
  Symbolic &S = *Sp;
  *flag = 0;
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    *flag = -1;
    return nullptr;
  }
  FileR r{f};
  char magic[8];
  r.get(magic, 8);
  const int ver = r.get1<int>();
  int hdr[5];
  r.get(hdr, 5);
  const int64_t naval = r.get1<int64_t>();
  const uint64_t fp = r.get1<uint64_t>();
  if (!r.ok || std::memcmp(magic, kMagic, 8) != 0 || ver != kVersion) {
    std::fclose(f);
    *flag = -1;
    return nullptr;
  }
  if (hdr[0] != S.n || hdr[1] != S.nnodes || fp != symbolic_fingerprint(S)) {
    std::fclose(f);
    *flag = -2;  // factors of another analyse
    return nullptr;
  }
  auto *N = new Numeric();
  N->symb = &S;
  N->posdef = hdr[2] != 0;
  N->c32 = hdr[3];
  try {
    HIPCHK(hipSetDevice(S.device));
    HIPCHK(hipStreamCreate(&N->stream));
    N->stats = r.get1<spral_ssids_hip_stats>();
    std::vector<double> hv(std::max<int64_t>(naval, S.n));
    HIPCHK(hipMalloc(&N->d_aval, std::max<int64_t>(naval, 1) * sizeof(double)));
    N->d_aval_cap = naval;
    if (naval > 0) {
      r.get(hv.data(), naval);
      copy_now(N->d_aval, hv.data(), naval * sizeof(double), hipMemcpyHostToDevice, N->stream);
    }
    if (hdr[4]) {
      r.get(hv.data(), S.n);
      HIPCHK(hipMalloc(&N->d_scal, std::max(S.n, 1) * sizeof(double)));
      copy_now(N->d_scal, hv.data(), S.n * sizeof(double), hipMemcpyHostToDevice, N->stream);
    }
    const int nn = S.nnodes;
    N->nodes.assign(nn, NodeDev{});
    N->nd_seen.assign(nn, 0);
    auto keep = [&](size_t bytes) {
      void *p = N->pool.alloc(std::max<size_t>(bytes, 1));
      N->persistent.push_back({p, std::max<size_t>(bytes, 1)});
      return p;
    };
    std::vector<double> dbuf;
    std::vector<int> ibuf;
    std::vector<char> cbuf;
    for (int sl = 0; sl < nn && r.ok; ++sl) {
      NodeDev &D = N->nodes[sl];
      node_static(S, sl, D);
      int st[10];
      r.get(st, 10);
      D.m = st[0];
      D.n = st[1];
      D.nd = st[2];
      D.nelim = st[3];
      D.nelim_app = st[4];
      D.flag = st[5];
      D.nzero = st[6];
      D.nsteps = st[7];
      D.ldl = st[8];
      D.ldc = st[9];
      if (D.m < D.n || D.n < 0 || D.ldl < D.m || D.nelim > D.n || D.n - D.nd != D.sn ||
          D.m - D.nd != D.sm) {
        r.ok = false;
        break;
      }
      N->nd_seen[sl] = D.nd;
      D.L = (double *)keep((size_t)D.ldl * D.n * sizeof(double));
      D.D = (double *)keep(2 * (size_t)std::max(D.n, 1) * sizeof(double));
      D.perm = (int *)keep(std::max(D.n, 1) * sizeof(int));
      D.ipos = (int *)keep(std::max(D.n, 1) * sizeof(int));
      if (D.n > 0) {
        dbuf.resize((size_t)D.ldl * D.n + 2 * (size_t)D.n);
        r.get(dbuf.data(), dbuf.size());
        ibuf.resize(2 * (size_t)D.n);
        r.get(ibuf.data(), ibuf.size());
        if (!r.ok) break;
        copy_now(D.L, dbuf.data(), (size_t)D.ldl * D.n * sizeof(double),
                         hipMemcpyHostToDevice, N->stream);
        copy_now(D.D, dbuf.data() + (size_t)D.ldl * D.n,
                         2 * (size_t)D.n * sizeof(double), hipMemcpyHostToDevice, N->stream);
        copy_now(D.perm, ibuf.data(), D.n * sizeof(int), hipMemcpyHostToDevice, N->stream);
        copy_now(D.ipos, ibuf.data() + D.n, D.n * sizeof(int), hipMemcpyHostToDevice, N->stream);
      }
      const size_t tn = ctab_ints(S, sl, D.m);
      D.ctab = (int *)keep(tn * sizeof(int));
      ibuf.resize(tn);
      r.get(ibuf.data(), tn);
      if (!r.ok) break;
      copy_now(D.ctab, ibuf.data(), tn * sizeof(int), hipMemcpyHostToDevice, N->stream);
      const int64_t cb = r.get1<int64_t>();
      const int cm = D.m - D.n;
      const size_t cbmax =
          (size_t)std::max(D.ldc, 0) * std::max(cm, 0) * (N->c32 ? sizeof(float) : sizeof(double));
      if (cb < 0 || (size_t)cb > cbmax) {
        r.ok = false;
        break;
      }
      if (cb > 0) {
        D.C = (double *)keep((size_t)cb);
        cbuf.resize(cb);
        r.get(cbuf.data(), cb);
        if (!r.ok) break;
        copy_now(D.C, cbuf.data(), cb, hipMemcpyHostToDevice, N->stream);
      }
    }
    std::fclose(f);
    f = nullptr;
    if (!r.ok) {
      delete N;
      *flag = -1;
      return nullptr;
    }
    if (!N->d_nodes) HIPCHK(hipMalloc(&N->d_nodes, std::max(nn, 1) * sizeof(NodeDev)));
    copy_now(N->d_nodes, N->nodes.data(), nn * sizeof(NodeDev), hipMemcpyHostToDevice, N->stream);
    if (!N->d_counters) HIPCHK(hipMalloc(&N->d_counters, 16 * sizeof(int)));
    if (!N->d_scratch_n) HIPCHK(hipMalloc(&N->d_scratch_n, std::max(S.n, 1) * sizeof(int)));
    build_solve_tables(*N);
    return N;
  } catch (HipError &) {
    if (f) std::fclose(f);
    delete N;
    *flag = -51;
  } catch (std::bad_alloc &) {
    if (f) std::fclose(f);
    delete N;
    *flag = -50;
  }
  return nullptr;
}

}  // namespace ssids_hip

using namespace ssids_hip;

static void set_error(spral_ssids_hip_stats *st, const HipError &e) {
  if (st) {
    st->flag = -51;  // SSIDS_ERROR_CUDA_UNKNOWN slot
    st->hip_error = (int)e.code;
  }
}

extern "C" {

void spral_ssids_hip_free_contrib(void *numeric);

int spral_ssids_hip_device_count(void) {
  ApiShared api_guard;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

void *spral_ssids_hip_create_symbolic(int device, int n, int sa, int en, const int *sptr,
                                      const int *sparent, const int64_t *rptr,
                                      const int *rlist, const int64_t *nptr,
                                      const int64_t *nlist,
                                      const struct spral_ssids_hip_options *options,
                                      int *flag) {
  ApiShared api_guard;
  *flag = 0;
  const int schedule = options ? options->level_schedule : 0;
  if (schedule < 0 || schedule > 3) {
    *flag = -1;  // no such level schedule
    return nullptr;
  }
  try {
    return build_symbolic(device, n, sa, en, sptr, sparent, rptr, rlist, nptr, nlist,
                          schedule);
  } catch (HipError &e) {
    *flag = -51;
  } catch (std::bad_alloc &) {
    *flag = -50;
  }
  return nullptr;
}

void spral_ssids_hip_destroy_symbolic(void *symbolic) {
  ApiShared api_guard;
  delete static_cast<Symbolic *>(symbolic);
}

void *spral_ssids_hip_factor_device(bool posdef, void *symbolic, const double *d_aval,
                                    int64_t naval, const double *d_scaling,
                                    const struct spral_ssids_hip_options *options,
                                    struct spral_ssids_hip_stats *stats) {
  ApiShared api_guard;
  Symbolic *S = static_cast<Symbolic *>(symbolic);
  spral_ssids_hip_options opt;
  if (options) opt = *options;
  else spral_ssids_hip_default_options(&opt);
  std::memset(stats, 0, sizeof(*stats));
  if (naval < S->naval_min) {
    stats->flag = -3;
    return nullptr;
  }
  auto *N = new Numeric();
  N->symb = S;
  N->posdef = posdef;
  try {
    HIPCHK(hipSetDevice(S->device));
    HIPCHK(hipStreamCreate(&N->stream));
    factor_impl(*N, d_aval, d_scaling, opt, naval);
    *stats = N->stats;
    if (stats->flag < 0) {
      delete N;
      return nullptr;
    }
    return N;
  } catch (HipError &e) {
    set_error(stats, e);
  } catch (std::bad_alloc &) {
    stats->flag = -50;
  }
  delete N;
  return nullptr;
}

void *spral_ssids_hip_factor(bool posdef, void *symbolic, const double *aval, int64_t naval,
                             const double *scaling,
                             const struct spral_ssids_hip_options *options,
                             struct spral_ssids_hip_stats *stats) {
  ApiShared api_guard;
  Symbolic *S = static_cast<Symbolic *>(symbolic);
  double *d_a = nullptr, *d_s = nullptr;
  void *res = nullptr;
  try {
    HIPCHK(hipSetDevice(S->device));
    HIPCHK(hipMalloc(&d_a, std::max<int64_t>(naval, 1) * sizeof(double)));
    copy_now(d_a, aval, naval * sizeof(double), hipMemcpyHostToDevice, thread_stream());
    if (scaling) {
      HIPCHK(hipMalloc(&d_s, std::max(S->n, 1) * sizeof(double)));
      copy_now(d_s, scaling, S->n * sizeof(double), hipMemcpyHostToDevice, thread_stream());
    }
    res = spral_ssids_hip_factor_device(posdef, symbolic, d_a, naval, d_s, options, stats);
  } catch (HipError &e) {
    std::memset(stats, 0, sizeof(*stats));
    set_error(stats, e);
  }
  if (d_a) (void)hipFree(d_a);
  if (d_s) (void)hipFree(d_s);
  return res;
}

int spral_ssids_hip_refactor_device(void *numeric, const double *d_aval,
                                    const double *d_scaling,
                                    const struct spral_ssids_hip_options *options,
                                    struct spral_ssids_hip_stats *stats) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  spral_ssids_hip_options opt;
  if (options) opt = *options;
  else spral_ssids_hip_default_options(&opt);
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    N->stats = spral_ssids_hip_stats{};
    factor_impl(*N, d_aval, d_scaling, opt, (int64_t)N->d_aval_cap);
    *stats = N->stats;
    return stats->flag;
  } catch (HipError &e) {
    std::memset(stats, 0, sizeof(*stats));
    set_error(stats, e);
    return -51;
  } catch (std::bad_alloc &) {
    stats->flag = -50;
    return -50;
  }
}

void spral_ssids_hip_destroy_numeric(void *numeric) {
  ApiShared api_guard;
  delete static_cast<Numeric *>(numeric);
}

static int solve_entry(void *numeric, int job, int nrhs, double *x, int ldx) {
  try {
    return solve_host(*static_cast<Numeric *>(numeric), job, nrhs, x, ldx);
  } catch (HipError &) {
    return -51;
  } catch (std::bad_alloc &) {
    return -50;
  }
}

int spral_ssids_hip_solve_fwd(void *numeric, int nrhs, double *x, int ldx) {
  ApiShared api_guard;
  return solve_entry(numeric, 1, nrhs, x, ldx);
}
int spral_ssids_hip_solve_diag(void *numeric, int nrhs, double *x, int ldx) {
  ApiShared api_guard;
  return solve_entry(numeric, 2, nrhs, x, ldx);
}
int spral_ssids_hip_solve_diag_bwd(void *numeric, int nrhs, double *x, int ldx) {
  ApiShared api_guard;
  return solve_entry(numeric, 4, nrhs, x, ldx);
}
int spral_ssids_hip_solve_bwd(void *numeric, int nrhs, double *x, int ldx) {
  ApiShared api_guard;
  return solve_entry(numeric, 3, nrhs, x, ldx);
}

int spral_ssids_hip_solve_device(void *numeric, int job, int nrhs, double *d_x, int ldx,
                                 void *stream) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  if (job < 0 || job > 4) return -11;
  if (nrhs < 1 || ldx < N->symb->n) return -10;
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    hipStream_t st = stream ? (hipStream_t)stream : N->stream;
    solve_device(*N, job, nrhs, d_x, ldx, st);
    return 0;
  } catch (HipError &) {
    return -51;
  }
}

int spral_ssids_hip_solve_device_noalloc(void *numeric, int job, int nrhs, double *d_x,
                                         int ldx, void *stream) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  if (job < 0 || job > 4) return -11;
  if (nrhs < 1 || ldx < N->symb->n) return -10;
  if (solve_ws_need(*N, nrhs) > kScratchDoubles)
    return SPRAL_SSIDS_HIP_ERROR_SCRATCH_TOO_SMALL;
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    double *ws = scratch_on(N->symb->device);
    hipStream_t st = stream ? (hipStream_t)stream : N->stream;
    solve_launch(*N, job, nrhs, d_x, ldx, st, ws);
    return 0;
  } catch (HipError &) {
    return -51;
  }
}

int64_t spral_ssids_hip_solve_workspace_doubles(const void *numeric, int nrhs) {
  ApiShared api_guard;
  return (int64_t)solve_ws_need(*static_cast<const Numeric *>(numeric), nrhs);
}

int spral_ssids_hip_solve_refine_device(void *numeric, int nrhs, const double *d_b,
                                        double *d_x, int ldx, int maxit, double tol,
                                        double *bwd, int *iters, void *stream) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  if (nrhs < 1 || ldx < N->symb->n) return -10;
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    hipStream_t st = stream ? (hipStream_t)stream : N->stream;
    return refine_device(*N, nrhs, d_b, d_x, ldx, maxit, tol, bwd, iters, st);
  } catch (HipError &) {
    return -51;
  }
}

int spral_ssids_hip_solve_refine(void *numeric, int nrhs, const double *b, double *x, int ldx,
                                 int maxit, double tol, double *bwd, int *iters) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  const int n = N->symb->n;
  if (nrhs < 1 || ldx < n) return -10;
  double *db = nullptr, *dx = nullptr;
  int rc = 0;
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    HIPCHK(hipMalloc(&db, (size_t)n * nrhs * sizeof(double)));
    HIPCHK(hipMalloc(&dx, (size_t)n * nrhs * sizeof(double)));
    copy2d_now(db, n * sizeof(double), b, ldx * sizeof(double), n * sizeof(double),
                       nrhs, hipMemcpyHostToDevice, N->stream);
    rc = refine_device(*N, nrhs, db, dx, n, maxit, tol, bwd, iters, N->stream);
    HIPCHK(hipMemcpy2DAsync(x, ldx * sizeof(double), dx, n * sizeof(double),
                            n * sizeof(double), nrhs, hipMemcpyDeviceToHost, N->stream));
    HIPCHK(hipStreamSynchronize(N->stream));
  } catch (HipError &) {
    rc = -51;
  }
  if (db) (void)hipFree(db);
  if (dx) (void)hipFree(dx);
  return rc;
}

int spral_ssids_hip_enquire_posdef(const void *numeric, double *d) {
  ApiShared api_guard;
  const Numeric *N = static_cast<const Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  if (!N->posdef) return -13;  // SSIDS_ERROR_NOT_LLT
  try {
    for (int i = 0; i < S.n; ++i) d[i] = 0.0;
    double *dp = d;
    // diagonal of L_chol = L D^{1/2}: sqrt(D_ii) = 1 / sqrt(D^{-1}_ii), in
    // the pivot order of enquire_indef (nodes in postorder)
    for (int i = 0; i < S.nnodes; ++i) {
      const NodeDev &D = N->nodes[S.slot[i]];
      const int ne = D.nelim;
      if (ne == 0) continue;
      std::vector<double> dh(2 * (size_t)ne);
      copy_now(dh.data(), D.D, 2 * ne * sizeof(double), hipMemcpyDeviceToHost, N->stream);
      for (int k = 0; k < ne; ++k) *dp++ = 1.0 / std::sqrt(dh[2 * k]);
    }
    return 0;
  } catch (HipError &) {
    return -51;
  }
}

int spral_ssids_hip_enquire_indef(const void *numeric, int *piv_order, double *d) {
  ApiShared api_guard;
  const Numeric *N = static_cast<const Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  if (N->posdef) return -14;  // SSIDS_ERROR_NOT_LDLT
  try {
    if (piv_order)
      for (int i = 0; i < S.n; ++i) piv_order[i] = 0;
    if (d)
      for (int i = 0; i < 2 * S.n; ++i) d[i] = 0.0;
    int piv = 0;
    double *dp = d;
    // every front's D^{-1} and eliminated variables, packed on the device and
    // copied back once (not two copies per front)
    const int nn = S.nnodes;
    std::vector<int64_t> off(nn + 1, 0);
    std::vector<int> slot_of(std::max(nn, 1));
    for (int i = 0; i < nn; ++i) {
      slot_of[i] = S.slot[i];
      off[i + 1] = off[i] + std::max(0, N->nodes[S.slot[i]].nelim);
    }
    const int64_t tot = off[nn];
    std::vector<double> dall(2 * (size_t)std::max<int64_t>(tot, 1));
    std::vector<int> pall(std::max<int64_t>(tot, 1));
    if (tot > 0) {
      int *d_slot = nullptr, *d_p = nullptr;
      int64_t *d_off = nullptr;
      double *d_d = nullptr;
      try {
        HIPCHK(hipMalloc(&d_slot, nn * sizeof(int)));
        HIPCHK(hipMalloc(&d_off, (nn + 1) * sizeof(int64_t)));
        HIPCHK(hipMalloc(&d_d, 2 * tot * sizeof(double)));
        HIPCHK(hipMalloc(&d_p, tot * sizeof(int)));
        copy_now(d_slot, slot_of.data(), nn * sizeof(int), hipMemcpyHostToDevice, N->stream);
        copy_now(d_off, off.data(), (nn + 1) * sizeof(int64_t), hipMemcpyHostToDevice, N->stream);
        launch_pack_dperm(N->stream, N->d_nodes, d_slot, d_off, nn, d_d, d_p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dall.data(), d_d, 2 * tot * sizeof(double),
                              hipMemcpyDeviceToHost, N->stream));
        HIPCHK(hipMemcpyAsync(pall.data(), d_p, tot * sizeof(int), hipMemcpyDeviceToHost,
                              N->stream));
        HIPCHK(hipStreamSynchronize(N->stream));
      } catch (HipError &) {
        (void)hipFree(d_slot);
        (void)hipFree(d_off);
        (void)hipFree(d_d);
        (void)hipFree(d_p);
        throw;
      }
      (void)hipFree(d_slot);
      (void)hipFree(d_off);
      (void)hipFree(d_d);
      (void)hipFree(d_p);
    }
    // pivot order follows the original node order (postorder), as the
    // reference walks nodes 0..nnodes-1
    for (int i = 0; i < nn; ++i) {
      const int ne = (int)(off[i + 1] - off[i]);
      const double *dh = dall.data() + 2 * off[i];
      const int *ph = pall.data() + off[i];
      for (int k = 0; k < ne;) {
        if (k + 1 == ne || std::isfinite(dh[2 * k + 2])) {
          if (piv_order) piv_order[ph[k]] = piv;
          ++piv;
          if (dp) { *dp++ = dh[2 * k]; *dp++ = 0.0; }
          k += 1;
        } else {
          if (piv_order) {
            piv_order[ph[k]] = -(piv++);
            piv_order[ph[k + 1]] = -(piv++);
          } else {
            piv += 2;
          }
          if (dp) {
            *dp++ = dh[2 * k];
            *dp++ = dh[2 * k + 1];
            *dp++ = dh[2 * k + 3];
            *dp++ = 0.0;
          }
          k += 2;
        }
      }
    }
    return 0;
  } catch (HipError &) {
    return -51;
  }
}

int spral_ssids_hip_alter(void *numeric, const double *d) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  if (N->posdef) return -14;  // SSIDS_ERROR_NOT_LDLT
  try {
    const double *dp = d;
    for (int i = 0; i < S.nnodes; ++i) {
      const NodeDev &D = N->nodes[S.slot[i]];
      int ne = D.nelim;
      if (ne == 0) continue;
      std::vector<double> dh(2 * (size_t)ne);
      copy_now(dh.data(), D.D, 2 * ne * sizeof(double), hipMemcpyDeviceToHost, N->stream);
      for (int k = 0; k < ne;) {
        if (k + 1 == ne || std::isfinite(dh[2 * k + 2])) {
          dh[2 * k] = dp[0];
          dh[2 * k + 1] = 0.0;
          dp += 2;
          k += 1;
        } else {
          dh[2 * k] = dp[0];
          dh[2 * k + 1] = dp[1];
          dh[2 * k + 3] = dp[2];
          dp += 4;
          k += 2;
        }
      }
      copy_now(D.D, dh.data(), 2 * ne * sizeof(double), hipMemcpyHostToDevice, N->stream);
    }
    return 0;
  } catch (HipError &) {
    return -51;
  }
}

int spral_ssids_hip_save_numeric(const void *numeric, const char *path) {
  ApiShared api_guard;
  try {
    return save_numeric(*static_cast<const Numeric *>(numeric), path);
  } catch (HipError &) {
    return -51;
  } catch (std::bad_alloc &) {
    return -50;
  }
}

void *spral_ssids_hip_load_numeric(void *symbolic, const char *path,
                                   struct spral_ssids_hip_stats *stats, int *flag) {
  ApiShared api_guard;
  Numeric *N = load_numeric(static_cast<Symbolic *>(symbolic), path, flag);
  if (stats) {
    if (N) *stats = N->stats;
    else std::memset(stats, 0, sizeof(*stats));
  }
  return N;
}

int spral_ssids_hip_numeric_scaling(const void *numeric, double *s) {
  ApiShared api_guard;
  const Numeric *N = static_cast<const Numeric *>(numeric);
  if (!N->d_scal) return 0;
  try {
    HIPCHK(hipSetDevice(N->symb->device));
    copy_now(s, N->d_scal, N->symb->n * sizeof(double), hipMemcpyDeviceToHost, N->stream);
  } catch (HipError &) {
    return -51;
  }
  return 1;
}

int spral_ssids_hip_get_contrib(void *numeric, int *n, const double **val, int *ldval,
                                const int **rlist, int *ndelay, const int **delay_perm,
                                const double **delay_val, int *lddelay) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  *n = 0;
  *val = nullptr;
  *ldval = 0;
  *rlist = nullptr;
  *ndelay = 0;
  *delay_perm = nullptr;
  *delay_val = nullptr;
  *lddelay = 0;
  // the subtree root is the last node in postorder
  if (S.nnodes == 0) return 0;
  try {
    HIPCHK(hipSetDevice(S.device));
    HIPCHK(hipStreamSynchronize(N->stream));
    const NodeDev &D = N->nodes[S.slot[S.nnodes - 1]];
    const int cm = D.m - D.n;
    N->contrib_val.assign((size_t)cm * cm, 0.0);
    if (cm > 0) {
      if (N->c32) {  // mixed precision: the block is held in FP32
        std::vector<float> f((size_t)cm * cm);
        copy2d_now(f.data(), cm * sizeof(float), D.C, D.ldc * sizeof(float),
                           cm * sizeof(float), cm, hipMemcpyDeviceToHost, N->stream);
        for (size_t k = 0; k < f.size(); ++k) N->contrib_val[k] = f[k];
      } else {
        copy2d_now(N->contrib_val.data(), cm * sizeof(double), D.C,
                           D.ldc * sizeof(double), cm * sizeof(double), cm,
                           hipMemcpyDeviceToHost, N->stream);
      }
      // symmetrise the lower triangle into the full square the reference passes
      for (int j = 0; j < cm; ++j)
        for (int i = 0; i < j; ++i)
          N->contrib_val[(size_t)j * cm + i] = N->contrib_val[(size_t)i * cm + j];
    }
    N->contrib_rlist.resize(cm);
    const int i_root = S.nnodes - 1;
    for (int k = 0; k < cm; ++k)
      N->contrib_rlist[k] = S.rows_h[S.row_off[i_root] + D.sn + k] + 1;
    const int nf = D.n - D.nelim;
    N->contrib_delay_perm.resize(nf);
    N->contrib_delay_val.assign((size_t)nf * D.m, 0.0);
    if (nf > 0) {
      std::vector<int> ph(D.n);
      copy_now(ph.data(), D.perm, D.n * sizeof(int), hipMemcpyDeviceToHost, N->stream);
      for (int k = 0; k < nf; ++k) N->contrib_delay_perm[k] = ph[D.nelim + k] + 1;
      copy2d_now(N->contrib_delay_val.data(), D.m * sizeof(double),
                         D.L + (size_t)D.nelim * D.ldl, D.ldl * sizeof(double),
                         D.m * sizeof(double), nf, hipMemcpyDeviceToHost, N->stream);
    }
    *n = cm;
    *val = N->contrib_val.data();
    *ldval = cm;
    *rlist = N->contrib_rlist.data();
    *ndelay = nf;
    *delay_perm = N->contrib_delay_perm.data();
    *delay_val = nf ? N->contrib_delay_val.data() + D.nelim : nullptr;
    *lddelay = D.m;
    return 0;
  } catch (HipError &) {
    spral_ssids_hip_free_contrib(numeric);
    return -51;
  } catch (std::bad_alloc &) {
    spral_ssids_hip_free_contrib(numeric);
    return -50;
  }
}

int spral_ssids_hip_node_factors(const void *numeric, int node, int *info, double *L,
                                 double *D, int *perm) {
  ApiShared api_guard;
  const Numeric *N = static_cast<const Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  if (node < 0 || node >= S.nnodes) return -1;
  const NodeDev &Dn = N->nodes[S.slot[node]];
  info[0] = Dn.m;
  info[1] = Dn.n;
  info[2] = Dn.nd;
  info[3] = Dn.nelim;
  info[4] = Dn.ldl;
  info[5] = Dn.sn;
  try {
    if (L && Dn.n > 0)
      copy_now(L, Dn.L, (size_t)Dn.ldl * Dn.n * sizeof(double), hipMemcpyDeviceToHost, N->stream);
    if (D && Dn.n > 0)
      copy_now(D, Dn.D, 2 * (size_t)Dn.n * sizeof(double), hipMemcpyDeviceToHost, N->stream);
    if (perm && Dn.n > 0)
      copy_now(perm, Dn.perm, Dn.n * sizeof(int), hipMemcpyDeviceToHost, N->stream);
  } catch (HipError &) {
    return -51;
  }
  return 0;
}

int spral_ssids_hip_node_pivots(const void *numeric, int *nelim_app, int *ndelay_in) {
  ApiShared api_guard;
  const Numeric *N = static_cast<const Numeric *>(numeric);
  const Symbolic &S = *N->symb;
  for (int node = 0; node < S.nnodes; ++node) {
    const NodeDev &Dn = N->nodes[S.slot[node]];
    if (nelim_app) nelim_app[node] = Dn.nelim_app;
    if (ndelay_in) ndelay_in[node] = Dn.nd;
  }
  return 0;
}

void spral_ssids_hip_free_contrib(void *numeric) {
  ApiShared api_guard;
  Numeric *N = static_cast<Numeric *>(numeric);
  std::vector<double>().swap(N->contrib_val);
  std::vector<double>().swap(N->contrib_delay_val);
  std::vector<int>().swap(N->contrib_rlist);
  std::vector<int>().swap(N->contrib_delay_perm);
}

}  // extern "C"
