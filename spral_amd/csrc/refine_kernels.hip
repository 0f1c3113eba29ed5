// FP64 iterative refinement support: residual r = b - A x with the original
// matrix held as a full symmetric CSR in elimination order (each row summed
// in a fixed order: deterministic), infinity norms and the correction update.
// The scaled backward error follows tests/ssids/ssids.f90:1762-1810:
//   max_i |r_i| / (||A||_inf ||x||_inf + ||b||_inf), per right-hand side.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ssids_hip {

// kResLanes lanes per row (stencil and KKT rows are short; longer rows are
// strided over the group), the group's sums reduced by a fixed butterfly
constexpr int kResLanes = 8;
__global__ void k_residual(int n, const int64_t *rp, const int *ci, const int64_t *src,
                           const double *aval, const double *scale_unused, const double *x,
                           const double *b, double *r, int ld, int nrhs) {
  const int sub = threadIdx.x & (kResLanes - 1);
  const int row = blockIdx.x * (blockDim.x / kResLanes) + threadIdx.x / kResLanes;
  if (row >= n) return;  // whole groups of kResLanes lanes leave together
  const int64_t e0 = rp[row], e1 = rp[row + 1];
  for (int k = 0; k < nrhs; ++k) {
    const double *xk = x + (size_t)k * ld;
    double acc = 0.0;
    for (int64_t e = e0 + sub; e < e1; e += kResLanes) acc += aval[src[e]] * xk[ci[e]];
    for (int off = kResLanes / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kResLanes);
    if (sub == 0) r[(size_t)k * ld + row] = b[(size_t)k * ld + row] - acc;
  }
}

// ||A||_inf: absolute row sums, max over rows (result as ordered bits)
__global__ void k_anorm(int n, const int64_t *rp, const int64_t *src, const double *aval,
                        unsigned long long *out) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double s = 0.0;
  for (int64_t e = rp[row]; e < rp[row + 1]; ++e) s += fabs(aval[src[e]]);
  atomicMax(out, (unsigned long long)__double_as_longlong(s));
}

// max |r|, max |x| and max |b| per column in one pass, into out[3k + 0..2]
// (ordered bits; the maximum propagates NaN, and as an ordered bit pattern a
// positive NaN sorts above +inf, so the atomic keeps it too); one atomic per
// workgroup
__global__ __launch_bounds__(256) void k_colmax3(int n, const double *rv, const double *xv,
                                                 const double *bv, int ld, int nrhs,
                                                 unsigned long long *out) {
  __shared__ double red[3][4];
  const int k = blockIdx.y;
  auto nmax = [](double a, double b) { return (b > a || b != b) ? b : a; };
  double m[3] = {0.0, 0.0, 0.0};
  const double *v[3] = {rv + (size_t)k * ld, xv + (size_t)k * ld, bv + (size_t)k * ld};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = nmax(m[q], fabs(v[q][i]));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    for (int off = 32; off > 0; off >>= 1) m[q] = nmax(m[q], __shfl_xor(m[q], off));
    if (lane == 0) red[q][wave] = m[q];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    double t = nmax(nmax(red[q][0], red[q][1]), nmax(red[q][2], red[q][3]));
    t = fabs(t);  // a NaN with the sign bit clear
    atomicMax(&out[k * 3 + q], (unsigned long long)__double_as_longlong(t));
  }
}

__global__ void k_axpy(int n, int nrhs, int ld, const double *dx, double *x,
                       const int *active) {
  const int k = blockIdx.y;
  if (!active[k]) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    x[(size_t)k * ld + i] += dx[(size_t)k * ld + i];
}

__global__ void k_scale_rows(int n, int nrhs, int ld, const double *s, double *v) {
  const int k = blockIdx.y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v[(size_t)k * ld + i] *= s[i];
}

void launch_scale_rows(hipStream_t st, int n, int nrhs, int ld, const double *s, double *v) {
  if (n <= 0) return;
  int gx = (n + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_scale_rows, dim3(gx, nrhs), dim3(256), 0, st, n, nrhs, ld, s, v);
}

void launch_residual(hipStream_t st, int n, const int64_t *rp, const int *ci,
                     const int64_t *src, const double *aval, const double *x, const double *b,
                     double *r, int ld, int nrhs) {
  if (n <= 0) return;
  const int rows_per = 256 / kResLanes;
  hipLaunchKernelGGL(k_residual, dim3((n + rows_per - 1) / rows_per), dim3(256), 0, st, n, rp,
                     ci, src, aval, nullptr, x, b, r, ld, nrhs);
}
void launch_anorm(hipStream_t st, int n, const int64_t *rp, const int64_t *src,
                  const double *aval, unsigned long long *out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_anorm, dim3((n + 255) / 256), dim3(256), 0, st, n, rp, src, aval, out);
}
void launch_colmax3(hipStream_t st, int n, const double *r, const double *x, const double *b,
                    int ld, int nrhs, unsigned long long *out) {
  if (n <= 0) return;
  int gx = (n + 255) / 256;
  if (gx > 256) gx = 256;
  hipLaunchKernelGGL(k_colmax3, dim3(gx, nrhs), dim3(256), 0, st, n, r, x, b, ld, nrhs, out);
}
void launch_axpy(hipStream_t st, int n, int nrhs, int ld, const double *dx, double *x,
                 const int *active) {
  if (n <= 0) return;
  int gx = (n + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_axpy, dim3(gx, nrhs), dim3(256), 0, st, n, nrhs, ld, dx, x, active);
}

}  // namespace ssids_hip
