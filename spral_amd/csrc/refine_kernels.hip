// FP64 iterative refinement support: residual r = b - A x with the original
// matrix held as a full symmetric CSR in elimination order (each row summed
// in a fixed order: deterministic), infinity norms and the correction update.
// The scaled backward error follows tests/ssids/ssids.f90:1762-1810:
//   max_i |r_i| / (||A||_inf ||x||_inf + ||b||_inf), per right-hand side.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ssids_hip {

// one wave per row (rows are short: stencils and KKT rows)
__global__ void k_residual(int n, const int64_t *rp, const int *ci, const int64_t *src,
                           const double *aval, const double *scale_unused, const double *x,
                           const double *b, double *r, int ld, int nrhs) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;
  for (int k = 0; k < nrhs; ++k) {
    const double *xk = x + (size_t)k * ld;
    double acc = 0.0;
    for (int64_t e = rp[row] + lane; e < rp[row + 1]; e += 64) acc += aval[src[e]] * xk[ci[e]];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) r[(size_t)k * ld + row] = b[(size_t)k * ld + row] - acc;
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

// per-column max |v| for nv vectors of length n (ld), into out[col*3 + slot]
__global__ void k_colmax(int n, const double *v, int ld, int nrhs, int slot,
                         unsigned long long *out) {
  const int k = blockIdx.y;
  if (k >= nrhs) return;
  // NaN-propagating maximum (fmax would drop a NaN): as an ordered bit
  // pattern a positive NaN sorts above +inf, so the atomic keeps it too
  auto nmax = [](double a, double b) { return (b > a || b != b) ? b : a; };
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = nmax(m, fabs(v[(size_t)k * ld + i]));
  for (int off = 32; off > 0; off >>= 1) m = nmax(m, __shfl_xor(m, off));
  m = fabs(m);  // a NaN with the sign bit clear
  if ((threadIdx.x & 63) == 0)
    atomicMax(&out[k * 3 + slot], (unsigned long long)__double_as_longlong(m));
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
  int rows_per = 4;  // 4 waves per workgroup
  hipLaunchKernelGGL(k_residual, dim3((n + rows_per - 1) / rows_per), dim3(256), 0, st, n, rp,
                     ci, src, aval, nullptr, x, b, r, ld, nrhs);
}
void launch_anorm(hipStream_t st, int n, const int64_t *rp, const int64_t *src,
                  const double *aval, unsigned long long *out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_anorm, dim3((n + 255) / 256), dim3(256), 0, st, n, rp, src, aval, out);
}
void launch_colmax(hipStream_t st, int n, const double *v, int ld, int nrhs, int slot,
                   unsigned long long *out) {
  if (n <= 0) return;
  int gx = (n + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_colmax, dim3(gx, nrhs), dim3(256), 0, st, n, v, ld, nrhs, slot, out);
}
void launch_axpy(hipStream_t st, int n, int nrhs, int ld, const double *dx, double *x,
                 const int *active) {
  if (n <= 0) return;
  int gx = (n + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_axpy, dim3(gx, nrhs), dim3(256), 0, st, n, nrhs, ld, dx, x, active);
}

}  // namespace ssids_hip
