// Test-only probes of compiler builtins the kernels rely on (in-bounds: one
// wave of 64 lanes, fixed-size buffers): the v_mfma_f64_16x16x4_f64 and
// v_mfma_f32_16x16x4_f32 operand/result layouts (16x4 A, 4x16 B, 16x16 D)
// and the DPP wave maximum
// of the diagonal-block pivot search.  Built into tests/hip/libprobe.so;
// never linked by the product library.
#include <hip/hip_runtime.h>

#include "dev.hpp"
typedef double double4_t __attribute__((ext_vector_type(4)));

__global__ void k_probe(const double *A, const double *B, double *out) {
  const int lane = threadIdx.x;  // 64 lanes
  // assumed layout: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k
  double a = A[(lane & 15) * 4 + (lane >> 4)];
  double b = B[(lane >> 4) * 16 + (lane & 15)];
  double4_t acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = acc[r];  // raw lane/register dump
}

extern "C" int probe_mfma_f64(const double *hA, const double *hB, double *hout) {
  double *A, *B, *O;
  if (hipMalloc(&A, 64 * sizeof(double)) != hipSuccess) return -1;
  if (hipMalloc(&B, 64 * sizeof(double)) != hipSuccess) return -1;
  if (hipMalloc(&O, 256 * sizeof(double)) != hipSuccess) return -1;
  (void)hipMemcpy(A, hA, 64 * sizeof(double), hipMemcpyHostToDevice);
  (void)hipMemcpy(B, hB, 64 * sizeof(double), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, 0, A, B, O);
  hipError_t e = hipMemcpy(hout, O, 256 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(A);
  (void)hipFree(B);
  (void)hipFree(O);
  return e == hipSuccess ? 0 : -2;
}

typedef float float4_t __attribute__((ext_vector_type(4)));

__global__ void k_probe_f32(const float *A, const float *B, float *out) {
  const int lane = threadIdx.x;  // 64 lanes
  float a = A[(lane & 15) * 4 + (lane >> 4)];
  float b = B[(lane >> 4) * 16 + (lane & 15)];
  float4_t acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = acc[r];  // raw lane/register dump
}

extern "C" int probe_mfma_f32(const float *hA, const float *hB, float *hout) {
  float *A, *B, *O;
  if (hipMalloc(&A, 64 * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&B, 64 * sizeof(float)) != hipSuccess) return -1;
  if (hipMalloc(&O, 256 * sizeof(float)) != hipSuccess) return -1;
  (void)hipMemcpy(A, hA, 64 * sizeof(float), hipMemcpyHostToDevice);
  (void)hipMemcpy(B, hB, 64 * sizeof(float), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe_f32, dim3(1), dim3(64), 0, 0, A, B, O);
  hipError_t e = hipMemcpy(hout, O, 256 * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(A);
  (void)hipFree(B);
  (void)hipFree(O);
  return e == hipSuccess ? 0 : -2;
}

__global__ void k_probe_wave_max(const unsigned long long *in, unsigned long long *out) {
  out[threadIdx.x] = ssids_hip::wave_max_u64(in[threadIdx.x]);  // 64 lanes, 64 outputs
}

extern "C" int probe_wave_max_u64(const unsigned long long *hin, unsigned long long *hout) {
  unsigned long long *I, *O;
  if (hipMalloc(&I, 64 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMalloc(&O, 64 * sizeof(unsigned long long)) != hipSuccess) return -1;
  (void)hipMemcpy(I, hin, 64 * sizeof(unsigned long long), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe_wave_max, dim3(1), dim3(64), 0, 0, I, O);
  hipError_t e = hipMemcpy(hout, O, 64 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(I);
  (void)hipFree(O);
  return e == hipSuccess ? 0 : -2;
}

// table_find over a prefix table: every workgroup b of a 256-thread grid
// records the segment the search returns (in-bounds: n + 1 table entries,
// one output per workgroup)
__global__ void k_probe_table_find(const int *off, int n, int *out) {
  const int li = ssids_hip::table_find(off, n, blockIdx.x);
  if (threadIdx.x == 0) out[blockIdx.x] = li;
}

extern "C" int probe_table_find(const int *hoff, int n, int nb, int *hout) {
  int *O, *R;
  if (hipMalloc(&O, (n + 1) * sizeof(int)) != hipSuccess) return -1;
  if (hipMalloc(&R, nb * sizeof(int)) != hipSuccess) return -1;
  (void)hipMemcpy(O, hoff, (n + 1) * sizeof(int), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe_table_find, dim3(nb), dim3(256), 0, 0, O, n, R);
  hipError_t e = hipMemcpy(hout, R, nb * sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(O);
  (void)hipFree(R);
  return e == hipSuccess ? 0 : -2;
}
