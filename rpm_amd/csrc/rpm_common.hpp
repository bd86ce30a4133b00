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

// Per-vector control block of a Krylov basis under construction.
struct BasisCtl {
  int active;       // 0 once the basis broke down (or the vector is idle)
  int size;         // basis size built so far / final
  int target;       // requested basis size
  int nfix;         // explicit corrections of the lagged projection applied in this basis
  double scale;     // 1/norm of the vector the next matvec normalises
  double beta_prev; // beta_{j-1} of the Lanczos three-term step (0 for j = 0)
  int begin;        // Lanczos steps begin <= j < stop run in the current call (a prefix of the
  int stop;         // basis now, the rest when a later call resumes it)
  int part;         // 1 if the trajectory takes part in the current call
  int npend;        // paired steps: vectors pending materialisation (1: v_j; 2: v_{j-1} and v_j)
  double zscale;    // paired steps: v_{j-1} = zscale * z - V c_A for the pending z
  int flags;        // paired steps: how the pair closes, and kBasisShort
  int pad_;
};
constexpr int kBasisShort = 1 << 4;   // the call stopped the basis early; a later call continues it

// Step j of the basis runs for this trajectory in the current call.
__host__ __device__ inline bool basis_step_live(const BasisCtl& c, int j) {
  return c.active && j < c.target && j >= c.begin && j < c.stop;
}

// Options of one batched H·psi launch (all nullable / off by default):
//  * ctl: per-vector control; vector b is skipped when !ctl[b].active or
//    j >= ctl[b].target, and its input is scaled by ctl[b].scale
//    (out = s * H in);
//  * copy: copy[b * copy_ld + x] = s * in[b][x] for every x (the normalised
//    Krylov vector, written as a side product of the matvec);
//  * pairs: per-vector 4x4 electron matrix (16 complex per vector, row
//    (eA,eB)) replacing the operator's own electron-pair/Zeeman matrix.
struct ApplyCtl {
  const BasisCtl* ctl = nullptr;
  int j = 0;
  c128* copy = nullptr;
  long long copy_ld = 0;
  const c128* pairs = nullptr;
  // Lanczos three-term step (with ctl): out -= ctl[b].beta_prev * prev[b * prev_ld + i], and
  // per-workgroup partial sums of conj(s in) (s H in) -- alpha_j before the subtraction -- at
  // dots[b * dots_ld + unit] (unit = tile group or block, apply_units() per vector).
  const c128* prev = nullptr;
  long long prev_ld = 0;
  c128* dots = nullptr;
  int dots_ld = 0;
  // affine epilogue (pair operators; Chebyshev / Clenshaw recurrences):
  //   out = aff_s * (H in) + aff_in * in + aff_c0 * x0 + aff_c1 * x1
  // with x0, x1 (nullable) per vector at b * ld like in and out.
  int affine = 0;
  double aff_s = 1.0;
  double aff_in = 0.0;
  c128 aff_c0 = {0.0, 0.0};
  double aff_c1 = 0.0;
  const c128* x0 = nullptr;
  const c128* x1 = nullptr;
};

// Output value of one amplitude: s * (H in) normally, the affine form when requested.
__device__ inline c128 apply_epilogue(const ApplyCtl& A, c128 hx, c128 xin, double s, long long b, long long ld,
                                      long long idx) {
  if (!A.affine) return cscale(hx, s);
  c128 y = cscale(hx, A.aff_s);
  y.re = fma(A.aff_in, xin.re, y.re);
  y.im = fma(A.aff_in, xin.im, y.im);
  if (A.x0) y = cfma(A.aff_c0, A.x0[b * ld + idx], y);
  if (A.x1) {
    const c128 v = A.x1[b * ld + idx];
    y.re = fma(A.aff_c1, v.re, y.re);
    y.im = fma(A.aff_c1, v.im, y.im);
  }
  return y;
}

// Electron-pair coefficients for vector b: diagonal, eA flip, eB flip, both flip.
struct PairCoef {
  c128 d[4], fa[4], fb[4], bo[4];
};

__device__ inline void load_pair_coef(const c128* __restrict__ P, PairCoef& C) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    C.d[e] = P[e * 4 + e];
    C.fa[e] = P[e * 4 + (e ^ 2)];
    C.fb[e] = P[e * 4 + (e ^ 1)];
    C.bo[e] = P[e * 4 + (e ^ 3)];
  }
}

int operator_apply_ex(const rpm_operator* op, const c128* in, c128* out, int batch, long long ld,
                      const ApplyCtl& A, hipStream_t stream);
// Number of partial sums per vector that operator_apply_ex writes to ApplyCtl::dots.
int apply_units(const rpm_operator* op);

// Block-wide sum of one complex value from every thread into dots[slot] (thread 0 writes).
// Every thread of the block must call it.  smem: blockDim.x / kWave entries.
__device__ inline void block_dot_store(c128 v, c128* smem, c128* dst) {
  const c128 t = block_sum(v, smem);
  if (threadIdx.x == 0) *dst = t;
}
}  // namespace rpm
