// extern "C" entry points of the host analyse and orderings.
#include <algorithm>
#include <cstdint>
#include <new>
#include <vector>

#include "../../include/spral_ssids_hip.h"
#include "host.hpp"

using ssids_hip::SymbolicData;

extern "C" {

void spral_ssids_hip_default_options(struct spral_ssids_hip_options *opt) {
  // Defaults of src/ssids/datatypes.f90:188-284
  opt->print_level = 0;
  opt->action = 1;
  opt->small = 1e-20;
  // u and multiplier are default-real (single precision) literals in the
  // reference, widened to double: 0.01f and 1.1f, not 0.01 and 1.1
  opt->u = (double)0.01f;
  opt->multiplier = (double)1.1f;
  opt->pivot_method = 2;
  opt->failed_pivot_method = 1;
  opt->schur_precision = 0;
  opt->refine_max = 0;
  opt->refine_tol = 1e-14;
  opt->timing = 0;
  opt->level_schedule = 0;
}

void *spral_ssids_hip_analyse(int n, const int64_t *ptr, const int *row,
                              int *order, int nemin, int *flag) {
  *flag = 0;
  if (n < 0) { *flag = -2; return nullptr; }
  if (nemin < 1) nemin = 32;  // nemin_default, src/ssids/anal.F90:990
  try {
    // check the order is a permutation (src/ssids/anal.F90:152-202)
    std::vector<char> seen(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      int p = std::abs(order[i]);
      if (p < 1 || p > n || seen[p]) { *flag = -8; return nullptr; }
      seen[p] = 1;
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = std::abs(order[i]);
    std::vector<int64_t> aptr;
    std::vector<int> arow;
    ssids_hip::expand_pattern(n, ptr, row, aptr, arow);
    auto *sd = new SymbolicData();
    int info = ssids_hip::basic_analyse(n, aptr.data(), arow.data(), ord.data(),
                                        nemin, *sd);
    if (info == 1) *flag = 6;  // SSIDS_WARNING_ANAL_SINGULAR
    ssids_hip::build_map(n, ptr, row, ord.data(), *sd);
    // unused variables (beyond the last supernode) are marked 0
    std::vector<int> invp(n + 1);
    for (int i = 0; i < n; ++i) invp[ord[i]] = i;
    for (int j = sd->sptr[sd->nnodes]; j <= n; ++j) ord[invp[j]] = 0;
    for (int i = 0; i < n; ++i) order[i] = ord[i];
    return sd;
  } catch (std::bad_alloc &) {
    *flag = -50;
    return nullptr;
  }
}

void spral_ssids_hip_analyse_info(const void *akeep, int *nnodes,
                                  int64_t *rlist_len, int64_t *nlist_len,
                                  int64_t *num_factor, int64_t *num_flops,
                                  int *maxfront, int *maxsupernode) {
  auto *sd = static_cast<const SymbolicData *>(akeep);
  *nnodes = sd->nnodes;
  *rlist_len = (int64_t)sd->rlist.size();
  *nlist_len = (int64_t)sd->nlist.size() / 2;
  *num_factor = sd->num_factor;
  *num_flops = sd->num_flops;
  int mf = 0, ms = 0;
  for (int s = 0; s < sd->nnodes; ++s) {
    mf = std::max(mf, (int)(sd->rptr[s + 1] - sd->rptr[s]));
    ms = std::max(ms, sd->sptr[s + 1] - sd->sptr[s]);
  }
  *maxfront = mf;
  *maxsupernode = ms;
}

void spral_ssids_hip_analyse_arrays(const void *akeep, int *sptr, int *sparent,
                                    int64_t *rptr, int *rlist, int64_t *nptr,
                                    int64_t *nlist) {
  auto *sd = static_cast<const SymbolicData *>(akeep);
  if (sptr) std::copy(sd->sptr.begin(), sd->sptr.end(), sptr);
  if (sparent) std::copy(sd->sparent.begin(), sd->sparent.end(), sparent);
  if (rptr) std::copy(sd->rptr.begin(), sd->rptr.end(), rptr);
  if (rlist) std::copy(sd->rlist.begin(), sd->rlist.end(), rlist);
  if (nptr) std::copy(sd->nptr.begin(), sd->nptr.end(), nptr);
  if (nlist) std::copy(sd->nlist.begin(), sd->nlist.end(), nlist);
}

void spral_ssids_hip_analyse_free(void *akeep) {
  delete static_cast<SymbolicData *>(akeep);
}

void spral_ssids_hip_order_grid_nd(int nx, int ny, int nz, int *order) {
  ssids_hip::order_grid_nd(nx, ny, nz, order);
}

int spral_ssids_hip_order_min_degree(int n, const int64_t *ptr, const int *row,
                                     int *order) {
  return ssids_hip::order_min_degree(n, ptr, row, order);
}

int spral_ssids_hip_order_amd(int n, const int64_t *ptr, const int *row, int *order) {
  return ssids_hip::order_amd(n, ptr, row, order);
}

}  // extern "C"
