// Symmetric infinity-norm equilibration on the device, the scaling that
// ssids_factor computes for options%scaling >= 4 (src/ssids/ssids.f90:
// 1003-1028) with equilib_scale_sym (src/scaling.f90:403-416), whose body is
// inf_norm_equilib_sym (src/scaling.f90:480-521): Algorithm 1 of Knight, Ruiz
// and Ucar, "A symmetry preserving algorithm for matrix scaling" (INRIA RR
// 7552, 2012).  From s = 1, each iteration takes
//   m_i = max over the stored lower entries a_rc with r = i or c = i of
//         |(s_r a_rc) s_c|,
//   s_i <- s_i / sqrt(m_i) where m_i > 0,
// and stops once max_i |1 - m_i| < tol (tested after the update).
//
// Every step is exactly rounded (products in the reference's operand order,
// max, sqrt, division), so the result matches the reference bit for bit.
// The maxima are scattered with 64-bit atomicMax on the bit patterns of the
// non-negative values, which order them as the values do: the result does
// not depend on the order of the updates.  The iterations run without a host
// round trip: each launch first reads the previous iteration's convergence
// value and returns at once when it has converged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/spral_ssids_hip.h"
#include "api_lock.hpp"

namespace ssids_hip {

namespace {

constexpr int kEqThreads = 256;

__device__ inline bool eq_converged(const unsigned long long *conv, int it, double tol) {
  return it > 0 && __longlong_as_double((long long)conv[it - 1]) < tol;
}

__global__ void k_eq_init(int n, int maxit, double *s, unsigned long long *mx,
                          unsigned long long *conv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    s[i] = 1.0;
    mx[i] = 0ull;
  }
  if (i < maxit) conv[i] = 0ull;
}

// one wave per column c: the column's own maximum is reduced in the wave
// (one atomic), each row r gets its own atomic
__global__ void k_eq_max(int n, const int64_t *ptr, const int *row, const double *val,
                         const double *s, unsigned long long *mx,
                         const unsigned long long *conv, int it, double tol) {
  if (eq_converged(conv, it, tol)) return;
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (kEqThreads / 64) + (threadIdx.x >> 6);
  if (c >= n) return;
  const double sc = s[c];
  unsigned long long cm = 0ull;
  for (int64_t e = ptr[c] - 1 + lane; e < ptr[c + 1] - 1; e += 64) {
    const int r = row[e] - 1;
    if (r < 0 || r >= n) continue;  // never written outside the vector
    const double v = fabs((s[r] * val[e]) * sc);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    atomicMax(&mx[r], b);
    cm = (b > cm) ? b : cm;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(cm, off);
    cm = (o > cm) ? o : cm;
  }
  if (lane == 0 && cm) atomicMax(&mx[c], cm);
}

// s_i /= sqrt(m_i) where m_i > 0; conv[it] = max |1 - m_i|; m_i reset to 0
__global__ void k_eq_update(int n, double *s, unsigned long long *mx, unsigned long long *conv,
                            int it, double tol) {
  if (eq_converged(conv, it, tol)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < n) {
    const double m = __longlong_as_double((long long)mx[i]);
    if (m > 0.0) s[i] = s[i] / sqrt(m);
    d = fabs(1.0 - m);
    mx[i] = 0ull;
  }
  unsigned long long b = (unsigned long long)__double_as_longlong(d);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = (o > b) ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(&conv[it], b);
}

int equilib_run(int n, const int64_t *d_ptr, const int *d_row, const double *d_val,
                double *d_s, int max_iterations, float tol, int *iterations) {
  if (n < 0 || max_iterations < 0) return -1;
  if (iterations) *iterations = 0;
  if (n == 0) return 0;
  const int maxit = max_iterations;
  unsigned long long *mx = nullptr, *conv = nullptr;
  if (hipMalloc(&mx, sizeof(unsigned long long) * n) != hipSuccess) return -2;
  if (hipMalloc(&conv, sizeof(unsigned long long) * (maxit > 0 ? maxit : 1)) != hipSuccess) {
    (void)hipFree(mx);
    return -2;
  }
  // the reference's tolerance is a default-kind real: compare in double
  // against its single-precision value
  const double tolv = (double)tol;
  const int gi = (std::max(n, maxit) + kEqThreads - 1) / kEqThreads;
  hipLaunchKernelGGL(k_eq_init, dim3(gi), dim3(kEqThreads), 0, 0, n, maxit, d_s, mx, conv);
  const int gc = (n + kEqThreads / 64 - 1) / (kEqThreads / 64);
  const int gu = (n + kEqThreads - 1) / kEqThreads;
  for (int it = 0; it < maxit; ++it) {
    hipLaunchKernelGGL(k_eq_max, dim3(gc), dim3(kEqThreads), 0, 0, n, d_ptr, d_row, d_val, d_s,
                       mx, conv, it, tolv);
    hipLaunchKernelGGL(k_eq_update, dim3(gu), dim3(kEqThreads), 0, 0, n, d_s, mx, conv, it,
                       tolv);
  }
  int rc = (hipGetLastError() == hipSuccess) ? 0 : -3;
  std::vector<unsigned long long> hc(maxit > 0 ? maxit : 1, 0ull);
  if (rc == 0 && maxit > 0 &&
      hipMemcpy(hc.data(), conv, sizeof(unsigned long long) * maxit, hipMemcpyDeviceToHost) !=
          hipSuccess)
    rc = -3;
  if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = -3;
  if (rc == 0 && iterations) {
    // the reference reports itr - 1: the index of the converged iteration,
    // or max_iterations when none converged
    int its = maxit;
    for (int it = 0; it < maxit; ++it) {
      double v;
      static_assert(sizeof(v) == sizeof(hc[0]), "bit copy");
      std::memcpy(&v, &hc[it], sizeof(v));
      if (v < tolv) {
        its = it;
        break;
      }
    }
    *iterations = its;
  }
  (void)hipFree(mx);
  (void)hipFree(conv);
  return rc;
}

}  // namespace
}  // namespace ssids_hip

extern "C" {

int spral_ssids_hip_equilib_scale_sym_device(int n, const int64_t *d_ptr, const int *d_row,
                                             const double *d_val, double *d_scaling,
                                             int max_iterations, float tol, int *iterations) {
  ssids_hip::ApiShared api_guard;
  return ssids_hip::equilib_run(n, d_ptr, d_row, d_val, d_scaling, max_iterations, tol,
                                iterations);
}

int spral_ssids_hip_equilib_scale_sym(int n, const int64_t *ptr, const int *row,
                                      const double *val, double *scaling, int max_iterations,
                                      float tol, int *iterations) {
  if (n < 0 || !ptr) return -1;
  if (n == 0) {
    if (iterations) *iterations = 0;
    return 0;
  }
  // the pattern is checked here, on the host, before any launch: 1-based,
  // ptr non-decreasing from 1, every row index in [1, n]
  if (ptr[0] != 1) return -1;
  for (int c = 0; c < n; ++c)
    if (ptr[c + 1] < ptr[c]) return -1;
  const int64_t nz = ptr[n] - 1;
  for (int64_t e = 0; e < nz; ++e)
    if (row[e] < 1 || row[e] > n) return -1;
  ssids_hip::ApiShared api_guard;
  int64_t *d_ptr = nullptr;
  int *d_row = nullptr;
  double *d_val = nullptr, *d_s = nullptr;
  int rc = 0;
  if (hipMalloc(&d_ptr, sizeof(int64_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&d_row, sizeof(int) * (nz > 0 ? nz : 1)) != hipSuccess ||
      hipMalloc(&d_val, sizeof(double) * (nz > 0 ? nz : 1)) != hipSuccess ||
      hipMalloc(&d_s, sizeof(double) * n) != hipSuccess)
    rc = -2;
  if (rc == 0 &&
      (hipMemcpy(d_ptr, ptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
       (nz > 0 &&
        (hipMemcpy(d_row, row, sizeof(int) * nz, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_val, val, sizeof(double) * nz, hipMemcpyHostToDevice) != hipSuccess))))
    rc = -3;
  if (rc == 0)
    rc = ssids_hip::equilib_run(n, d_ptr, d_row, d_val, d_s, max_iterations, tol, iterations);
  if (rc == 0 &&
      hipMemcpy(scaling, d_s, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    rc = -3;
  if (d_ptr) (void)hipFree(d_ptr);
  if (d_row) (void)hipFree(d_row);
  if (d_val) (void)hipFree(d_val);
  if (d_s) (void)hipFree(d_s);
  return rc;
}

}  // extern "C"
