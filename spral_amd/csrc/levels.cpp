// Tree levels of the numeric factorization.
//
// Every front of a level is factored by one launch sequence: the level's
// gathers, then its APP steps (pivot block, apply, update) in sequence for all
// of its fronts at once, so a level costs the most APP steps of its fronts
// (ceil(columns / 64)), plus a fixed part for its other launches.  A level
// assignment is valid when every front's level is above its children's.
//
// Height from the leaves (the level of a front is one more than its highest
// child) places large fronts of different subtrees on different levels when
// the subtrees differ in height: two halves of a nested dissection whose
// separators sit below chains of small fronts of different lengths, or the
// independent trees of a batch.  Two other assignments line them up:
//  * greedy: the most urgent ready front (the least latest start in the
//    schedule that runs every front as late as the root allows, with fronts
//    weighted by their steps) and every ready front no heavier than it form
//    the next level; heavier ready fronts wait for a level that needs them;
//  * layered: fronts of at least kHeavySteps steps with the same number of
//    such ancestors share a level, the gap between two such layers fitting
//    the longest chain of lighter fronts between them; the lighter fronts
//    sit as late as their parents allow.
// The cheapest of the three by the cost model is kept (mode 0); a new
// assignment must beat height from the leaves by kMinGain, since the placement
// of the small fronts also moves the lifetimes of their contribution blocks.
// Modes 1-3 force height, greedy or layered (the schedule autotuner times
// them on the device instead of trusting the model).
#include <algorithm>
#include <cstdint>
#include <set>
#include <tuple>
#include <vector>

#include "../../include/spral_ssids_hip.h"
#include "host.hpp"

namespace ssids_hip {

namespace {

constexpr int kStepCols = 64;          // APP block width
constexpr int kHeavySteps = 2;         // layered: fronts of at least this many steps
constexpr int kLevelCostQuarter = 1;   // a level's fixed part, in quarter steps
constexpr double kMinGain = 0.05;

int64_t level_cost(const std::vector<int> &lev, const std::vector<int> &w) {
  int nl = 0;
  for (int l : lev) nl = std::max(nl, l + 1);
  std::vector<int> mx(nl, 0);
  for (size_t i = 0; i < lev.size(); ++i) mx[lev[i]] = std::max(mx[lev[i]], w[i]);
  int64_t c = 0;
  for (int L = 0; L < nl; ++L) c += 4 * (int64_t)mx[L] + kLevelCostQuarter;
  return c;
}

// children before parents (postorder): parent[i] > i or -1
std::vector<int> levels_height(const std::vector<int> &parent) {
  const int nn = (int)parent.size();
  std::vector<int> lev(nn, 0);
  for (int i = 0; i < nn; ++i)
    if (parent[i] >= 0) lev[parent[i]] = std::max(lev[parent[i]], lev[i] + 1);
  return lev;
}

std::vector<int> levels_greedy(const std::vector<int> &parent, const std::vector<int> &w) {
  const int nn = (int)parent.size();
  // weighted distance to the root, the front itself excluded
  std::vector<int64_t> top(nn, 0);
  for (int i = nn - 1; i >= 0; --i)
    if (parent[i] >= 0) top[i] = top[parent[i]] + w[parent[i]];
  std::vector<int> pending(nn, 0), lev(nn, -1);
  for (int i = 0; i < nn; ++i)
    if (parent[i] >= 0) pending[parent[i]]++;
  // urgency: latest start = -(top + w), so larger top + w is more urgent;
  // ties to the heavier front, then the lower index
  auto key = [&](int i) { return std::make_tuple(-(top[i] + w[i]), -w[i], i); };
  std::set<std::tuple<int64_t, int, int>> ready;
  for (int i = 0; i < nn; ++i)
    if (pending[i] == 0) ready.insert(key(i));
  std::vector<int> batch, next;
  for (int L = 0; !ready.empty(); ++L) {
    const int W = w[std::get<2>(*ready.begin())];
    batch.clear();
    for (auto it = ready.begin(); it != ready.end();) {
      const int i = std::get<2>(*it);
      if (w[i] <= W) {
        batch.push_back(i);
        it = ready.erase(it);
      } else {
        ++it;
      }
    }
    for (int i : batch) {
      lev[i] = L;
      const int p = parent[i];
      if (p >= 0 && --pending[p] == 0) ready.insert(key(p));
    }
  }
  return lev;
}

std::vector<int> levels_layered(const std::vector<int> &parent, const std::vector<int> &w) {
  const int nn = (int)parent.size();
  std::vector<char> heavy(nn);
  for (int i = 0; i < nn; ++i) heavy[i] = w[i] >= kHeavySteps;
  // hd: heavy proper ancestors; k: lighter fronts between the front and its
  // nearest heavy ancestor
  std::vector<int> hd(nn, 0), k(nn, 0);
  int maxhd = 0;
  for (int i = nn - 1; i >= 0; --i) {
    const int p = parent[i];
    if (p < 0) continue;
    hd[i] = hd[p] + (heavy[p] ? 1 : 0);
    k[i] = heavy[p] ? 0 : k[p] + 1;
    maxhd = std::max(maxhd, hd[i]);
  }
  std::vector<int> gap(maxhd + 1, 0);
  for (int i = 0; i < nn; ++i)
    if (heavy[i]) gap[hd[i]] = std::max(gap[hd[i]], k[i] + (hd[i] > 0 ? 1 : 0));
  std::vector<int> layer(maxhd + 1, 0);
  layer[0] = gap[0];
  for (int d = 1; d <= maxhd; ++d) layer[d] = layer[d - 1] + gap[d];
  // distance from the top level
  std::vector<int> dist(nn, 0);
  for (int i = nn - 1; i >= 0; --i) {
    const int p = parent[i];
    if (heavy[i]) dist[i] = layer[hd[i]];
    else dist[i] = (p >= 0) ? dist[p] + 1 : 0;
    if (p >= 0 && dist[i] <= dist[p]) dist[i] = dist[p] + 1;
  }
  int dmax = 0;
  for (int d : dist) dmax = std::max(dmax, d);
  std::vector<int> lev(nn);
  for (int i = 0; i < nn; ++i) lev[i] = dmax - dist[i];
  return lev;
}

}  // namespace

std::vector<int> tree_levels(const std::vector<int> &parent, const std::vector<int> &ncols,
                             int mode) {
  const int nn = (int)parent.size();
  std::vector<int> w(nn);
  for (int i = 0; i < nn; ++i) w[i] = std::max(1, (ncols[i] + kStepCols - 1) / kStepCols);
  if (mode == 2) return levels_greedy(parent, w);
  if (mode == 3) return levels_layered(parent, w);
  std::vector<int> best = levels_height(parent);
  if (mode == 1) return best;
  const int64_t base = level_cost(best, w);
  int64_t bc = base;
  for (auto cand : {levels_greedy(parent, w), levels_layered(parent, w)}) {
    const int64_t c = level_cost(cand, w);
    if (c < bc && (double)c <= (1.0 - kMinGain) * (double)base) {
      bc = c;
      best = std::move(cand);
    }
  }
  return best;
}

}  // namespace ssids_hip

extern "C" int spral_ssids_hip_tree_levels(int nnodes, const int *parent, const int *ncols,
                                           int mode, int *level) {
  if (nnodes < 0 || (nnodes > 0 && (!parent || !ncols || !level)) || mode < 0 || mode > 3)
    return -1;
  std::vector<int> par(parent, parent + nnodes), nc(ncols, ncols + nnodes);
  for (int i = 0; i < nnodes; ++i)
    if (par[i] != -1 && (par[i] <= i || par[i] >= nnodes)) return -1;
  const std::vector<int> lev = ssids_hip::tree_levels(par, nc, mode);
  int nl = 0;
  for (int i = 0; i < nnodes; ++i) {
    level[i] = lev[i];
    nl = std::max(nl, lev[i] + 1);
  }
  return nl;
}
