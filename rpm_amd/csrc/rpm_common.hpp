// Shared helpers for the rpm_amd HIP library: complex FP64 arithmetic,
// error reporting, wave/block reductions.  gfx950 (CDNA4) only: wave64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <string>

#include "../../include/rpm_amd.h"

namespace rpm {

constexpr int kWave = 64;

struct __align__(16) c128 {
  double re, im;
};

__host__ __device__ inline c128 cmake(double r, double i) { return c128{r, i}; }
__host__ __device__ inline c128 cadd(c128 a, c128 b) { return c128{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline c128 csub(c128 a, c128 b) { return c128{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline c128 cscale(c128 a, double s) { return c128{a.re * s, a.im * s}; }
// acc + a * b
__host__ __device__ inline c128 cfma(c128 a, c128 b, c128 acc) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
  return acc;
}
// acc + conj(a) * b
__host__ __device__ inline c128 cfmaconj(c128 a, c128 b, c128 acc) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(-a.im, b.re, acc.im);
  return acc;
}
// acc - a * b
__host__ __device__ inline c128 cfms(c128 a, c128 b, c128 acc) {
  acc.re = fma(-a.re, b.re, acc.re);
  acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(-a.re, b.im, acc.im);
  acc.im = fma(-a.im, b.re, acc.im);
  return acc;
}

// ---------------------------------------------------------------- errors

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define RPM_HIP_CHECK(expr)                                                         \
  do {                                                                              \
    hipError_t rpm_err__ = (expr);                                                  \
    if (rpm_err__ != hipSuccess)                                                    \
      return ::rpm::fail(RPM_E_HIP, "%s failed: %s (%s:%d)", #expr,                 \
                         hipGetErrorString(rpm_err__), __FILE__, __LINE__);         \
  } while (0)

#define RPM_LAUNCH_CHECK() RPM_HIP_CHECK(hipGetLastError())

// ---------------------------------------------------------------- reductions

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ inline c128 wave_sum(c128 v) {
  v.re = wave_sum(v.re);
  v.im = wave_sum(v.im);
  return v;
}

// Deterministic block sum; every thread receives the result.  `smem` must
// hold blockDim.x / kWave doubles.  Ends with a barrier so smem can be reused.
__device__ inline double block_sum(double v, double* smem) {
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += smem[i];
  __syncthreads();
  return s;
}

__device__ inline c128 block_sum(c128 v, c128* smem) {
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave, nw = blockDim.x / kWave;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  c128 s{0.0, 0.0};
  for (int i = 0; i < nw; ++i) s = cadd(s, smem[i]);
  __syncthreads();
  return s;
}

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace rpm

// Opaque operator handle shared by the operator and Lanczos translation units.
struct rpm_operator {
  int kind;             // 1 = two-radical pair stencil, 2 = generic local terms
  long long n;          // state dimension
  void* impl;           // kind-specific device tables
};

namespace rpm {
// Apply H with an optional device scale s (out = s * H in) and an optional copy
// of the scaled input (copy_out[x] = s * in[x]) written for every x.
// `active` (device int, nullable): when it points to 0 the launch does nothing.
int operator_apply_ex(const rpm_operator* op, const c128* in, c128* out, int batch, long long ld,
                      const double* in_scale, c128* copy_out, const int* active,
                      hipStream_t stream);
}  // namespace rpm
