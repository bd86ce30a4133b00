// Device pieces of the adjoint (costate) sweep and of segmented propagation:
//   * electron cross-densities C[e][e'] = sum_n conj(a[e, n]) b[e', n], the
//     4x4 reduced objects every electron-only derivative dH/dp needs;
//   * the singlet kick g += c P psi of the costate at an output time;
//   * an instantaneous local unitary on one tensor axis of a state
//     (R/poster_reproduction/core.py:285-292).
// All batched over vectors (grid.y); reductions are two-stage in a fixed order.
#include <utility>
#include <vector>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kDensThreads = 256;
constexpr int kDensPer = 8;   // nuclear indices per thread

// partial[(b * nblk + blk) * 16 + e*4 + e'] = sum over the block's nuclear indices of conj(a_e) b_e'
__global__ void __launch_bounds__(kDensThreads) k_cross_density(const c128* __restrict__ A, const c128* __restrict__ Bv,
                                                                long long ld, long long nn, c128* __restrict__ partial) {
  __shared__ c128 sm[kDensThreads / kWave];
  const long long b = blockIdx.y;
  const c128* a = A + b * ld;
  const c128* v = Bv + b * ld;
  c128 acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = cmake(0.0, 0.0);
  const long long base = static_cast<long long>(blockIdx.x) * kDensThreads * kDensPer + threadIdx.x;
  for (int q = 0; q < kDensPer; ++q) {
    const long long n = base + static_cast<long long>(q) * kDensThreads;
    if (n >= nn) break;
    c128 x[4], y[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      x[e] = a[e * nn + n];
      y[e] = v[e * nn + n];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int f = 0; f < 4; ++f) acc[e * 4 + f] = cfmaconj(x[e], y[f], acc[e * 4 + f]);
  }
  c128* out = partial + (b * gridDim.x + blockIdx.x) * 16;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const c128 t = block_sum(acc[i], sm);
    if (threadIdx.x == 0) out[i] = t;
  }
}

// acc[b*16 + i] += weight * sum_blk partial[b][blk][i]; one 64-thread block per b, lane i < 16.
__global__ void k_density_finish(const c128* __restrict__ partial, int nblk, double weight, c128* __restrict__ acc) {
  const int b = blockIdx.x;
  const int i = threadIdx.x;
  if (i >= 16) return;
  c128 s{0.0, 0.0};
  for (int k = 0; k < nblk; ++k) s = cadd(s, partial[(static_cast<long long>(b) * nblk + k) * 16 + i]);
  c128& dst = acc[b * 16 + i];
  dst.re = fma(weight, s.re, dst.re);
  dst.im = fma(weight, s.im, dst.im);
}

// g[e=1] += c (psi_1 - psi_2)/2, g[e=2] -= c (psi_1 - psi_2)/2   (P = |S><S|, S = (|01> - |10>)/sqrt2)
__global__ void __launch_bounds__(256) k_singlet_kick(const c128* __restrict__ psi, c128* __restrict__ g, long long ld,
                                                      long long nn, double c) {
  const long long b = blockIdx.y;
  const long long n = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (n >= nn) return;
  const c128* x = psi + b * ld;
  c128* y = g + b * ld;
  const c128 p1 = x[nn + n], p2 = x[2 * nn + n];
  const double re = 0.5 * c * (p1.re - p2.re), im = 0.5 * c * (p1.im - p2.im);
  y[nn + n].re += re;
  y[nn + n].im += im;
  y[2 * nn + n].re -= re;
  y[2 * nn + n].im -= im;
}

// out = (1 (x) U (x) 1) in on the axis with dimension d and C-order stride `inner`.
__global__ void __launch_bounds__(256) k_axis_apply(const c128* __restrict__ in, c128* __restrict__ out, long long ld,
                                                    long long rows, long long inner, int d,
                                                    const c128* __restrict__ U) {
  const long long b = blockIdx.y;
  const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= rows) return;
  const long long o = t / inner, i = t - o * inner;
  const long long base = o * d * inner + i;
  const c128* x = in + b * ld;
  c128* y = out + b * ld;
  c128 v[16];
  for (int k = 0; k < d; ++k) v[k] = x[base + k * inner];
  for (int r = 0; r < d; ++r) {
    c128 s{0.0, 0.0};
    for (int k = 0; k < d; ++k) s = cfma(U[r * d + k], v[k], s);
    y[base + r * inner] = s;
  }
}

// out = alpha x (+ beta y when y is given), complex coefficients.
__global__ void __launch_bounds__(256) k_axpby(c128 alpha, const c128* __restrict__ x, c128 beta,
                                               const c128* __restrict__ y, c128* __restrict__ out, long long ld,
                                               long long n) {
  const long long b = blockIdx.y;
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  c128 v = cfma(alpha, x[b * ld + i], cmake(0.0, 0.0));
  if (y) v = cfma(beta, y[b * ld + i], v);
  out[b * ld + i] = v;
}

}  // namespace rpm

using namespace rpm;

extern "C" size_t rpm_cross_density_workspace_bytes(long long nn, int batch) {
  const long long nblk = (nn + kDensThreads * kDensPer - 1) / (kDensThreads * kDensPer);
  return static_cast<size_t>(nblk) * (batch > 0 ? batch : 1) * 16 * sizeof(c128) + 256;
}

extern "C" int rpm_cross_density(const void* d_a, const void* d_b, int batch, long long ld, long long nn,
                                 double weight, void* d_acc, void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_a || !d_b || !d_acc || !d_ws || batch < 1 || batch > 65535 || nn < 1 || ld < 4 * nn)
    return fail(RPM_E_INVALID, "rpm_cross_density: bad arguments");
  if (ws_bytes < rpm_cross_density_workspace_bytes(nn, batch))
    return fail(RPM_E_INVALID, "rpm_cross_density: workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nblk = ceil_div(nn, kDensThreads * kDensPer);
  c128* part = static_cast<c128*>(d_ws);
  hipLaunchKernelGGL(k_cross_density, dim3(nblk, batch), dim3(kDensThreads), 0, s, static_cast<const c128*>(d_a),
                     static_cast<const c128*>(d_b), ld, nn, part);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_density_finish, dim3(batch), dim3(64), 0, s, part, nblk, weight, static_cast<c128*>(d_acc));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_singlet_kick(const void* d_psi, void* d_g, int batch, long long ld, long long nn, double coef,
                                void* stream) {
  if (!d_psi || !d_g || d_psi == d_g || batch < 1 || batch > 65535 || nn < 1 || ld < 4 * nn)
    return fail(RPM_E_INVALID, "rpm_singlet_kick: bad arguments");
  hipLaunchKernelGGL(k_singlet_kick, dim3(ceil_div(nn, 256), batch), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const c128*>(d_psi), static_cast<c128*>(d_g), ld, nn, coef);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_axis_apply(int naxes, const int* dims, int axis, const void* d_U, const void* d_in, void* d_out,
                              int batch, long long ld, void* stream) {
  if (!dims || !d_U || !d_in || !d_out || d_in == d_out || naxes < 1 || naxes > 64 || axis < 0 || axis >= naxes ||
      batch < 1 || batch > 65535)
    return fail(RPM_E_INVALID, "rpm_axis_apply: bad arguments");
  long long n = 1, inner = 1;
  for (int i = 0; i < naxes; ++i) {
    if (dims[i] < 1 || dims[i] > 16) return fail(RPM_E_INVALID, "rpm_axis_apply: axis dimension %d", dims[i]);
    n *= dims[i];
    if (i > axis) inner *= dims[i];
  }
  if (ld < n) return fail(RPM_E_INVALID, "rpm_axis_apply: ld %lld below N %lld", ld, n);
  const int d = dims[axis];
  const long long rows = n / d;
  hipLaunchKernelGGL(k_axis_apply, dim3(ceil_div(rows, 256), batch), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const c128*>(d_in), static_cast<c128*>(d_out), ld, rows, inner, d,
                     static_cast<const c128*>(d_U));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

// ---------------------------------------------------------------- adjoint sweep

namespace {

struct SweepCtx {
  const rpm_operator* op;
  int batch;
  long long n, nn;
  double shift, hs;       // X = (H - shift) / hs
  hipStream_t s;
  c128* dens_ws;
  size_t dens_bytes;
};

int affine(const SweepCtx& C, const c128* in, c128* out, const c128* pairs, double s, double a_in, const c128* x0,
           c128 c0, const c128* x1, double c1) {
  ApplyCtl A;
  A.pairs = pairs;
  A.affine = 1;
  A.aff_s = s;
  A.aff_in = a_in;
  A.x0 = x0;
  A.aff_c0 = c0;
  A.x1 = x1;
  A.aff_c1 = c1;
  return operator_apply_ex(C.op, in, out, C.batch, C.n, A, C.s);
}

int scaled_copy(const SweepCtx& C, c128 alpha, const c128* x, c128* out) {
  hipLaunchKernelGGL(k_axpby, dim3(ceil_div(C.n, 256), C.batch), dim3(256), 0, C.s, alpha, x, cmake(0.0, 0.0),
                     static_cast<const c128*>(nullptr), out, C.n, C.n);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

int density(const SweepCtx& C, const c128* a, const c128* b, double weight, c128* acc) {
  return rpm_cross_density(a, b, C.batch, C.n, C.nn, weight, acc, C.dens_ws, C.dens_bytes, C.s);
}

inline c128 conjc(const double* c) { return cmake(c[0], -c[1]); }

// out = sum_n conj(c_n) T_n(X) seed by Clenshaw's recurrence (K >= 2 terms beyond c_0).
// With u != NULL, also acc += sum_n w_n / hs * C(nu_{n+1}, u_n) (w_0 = 1, w_n = 2) for the
// Clenshaw vectors nu.  ra, rb: scratch vectors.
int clenshaw(const SweepCtx& C, const c128* seed, const double* coef, int K, const c128* pairs, c128* const* u,
             c128* acc, c128* ra, c128* rb, c128* out) {
  const double s2 = 2.0 / C.hs, a2 = -2.0 * C.shift / C.hs;
  int rc = scaled_copy(C, conjc(coef + 2 * K), seed, ra);          // nu_K
  if (rc) return rc;
  if (u && (rc = density(C, ra, u[K - 1], (K - 1 >= 1 ? 2.0 : 1.0) / C.hs, acc))) return rc;
  rc = affine(C, ra, rb, pairs, s2, a2, seed, conjc(coef + 2 * (K - 1)), nullptr, 0.0);   // nu_{K-1}
  if (rc) return rc;
  if (u && (rc = density(C, rb, u[K - 2], (K - 2 >= 1 ? 2.0 : 1.0) / C.hs, acc))) return rc;
  c128* hi = ra;   // nu_{n+2}
  c128* lo = rb;   // nu_{n+1}
  for (int n = K - 2; n >= 1; --n) {
    rc = affine(C, lo, hi, pairs, s2, a2, seed, conjc(coef + 2 * n), hi, -1.0);            // nu_n over nu_{n+2}
    if (rc) return rc;
    if (u && (rc = density(C, hi, u[n - 1], (n - 1 >= 1 ? 2.0 : 1.0) / C.hs, acc))) return rc;
    c128* t = hi;
    hi = lo;
    lo = t;
  }
  // lo = nu_1, hi = nu_2
  return affine(C, lo, out, pairs, 1.0 / C.hs, -C.shift / C.hs, seed, conjc(coef), hi, -1.0);
}

}  // namespace

extern "C" size_t rpm_adjoint_workspace_bytes(const rpm_operator* op, int batch, int kmax) {
  if (!op || batch < 1 || kmax < 2) return 0;
  const size_t vec = static_cast<size_t>(op->n) * batch * sizeof(c128);
  const size_t vecs = 6 + static_cast<size_t>(kmax);   // psi', g', 2 ring, u_1..u_{K-1} (+1 spare)
  return vecs * ((vec + 255) & ~size_t(255)) + rpm_cross_density_workspace_bytes(op->n / 4, batch) + 256;
}

extern "C" int rpm_adjoint_sweep(const rpm_operator* op, int batch, void* d_psi, int nint, const double* h_int,
                                 const double* weights, double shift, double half_width, int ncls,
                                 const int* nterms, int kmax, const double* coef, const int* cls,
                                 const void* d_pairs, const int* seg, void* d_acc, const int* slot, void* d_ws,
                                 size_t ws_bytes, void* stream) {
  if (!op || op->kind != 1 || !d_psi || nint < 1 || !h_int || !weights || !nterms || !coef || !cls || !d_acc ||
      !slot || !d_ws || batch < 1 || batch > 65535 || ncls < 1 || kmax < 2 || !(half_width > 0.0))
    return fail(RPM_E_INVALID, "rpm_adjoint_sweep: bad arguments");
  if (d_pairs && !seg) return fail(RPM_E_INVALID, "rpm_adjoint_sweep: per-interval segments needed with pairs");
  for (int c = 0; c < ncls; ++c)
    if (nterms[c] < 2 || nterms[c] > kmax) return fail(RPM_E_INVALID, "Chebyshev length %d outside 2..%d", nterms[c], kmax);
  for (int k = 0; k < nint; ++k)
    if (cls[k] < 0 || cls[k] >= ncls || slot[k] < 0 || (seg && seg[k] < 0))
      return fail(RPM_E_INVALID, "rpm_adjoint_sweep: interval %d has a bad class, slot or segment", k);
  if (ws_bytes < rpm_adjoint_workspace_bytes(op, batch, kmax))
    return fail(RPM_E_INVALID, "rpm_adjoint_sweep: workspace too small");
  SweepCtx C;
  C.op = op;
  C.batch = batch;
  C.n = op->n;
  C.nn = op->n / 4;
  C.shift = shift;
  C.hs = half_width;
  C.s = static_cast<hipStream_t>(stream);
  const size_t vec = ((static_cast<size_t>(C.n) * batch * sizeof(c128)) + 255) & ~size_t(255);
  char* base = static_cast<char*>(d_ws);
  auto vecp = [&](int i) { return reinterpret_cast<c128*>(base + vec * i); };
  c128* psi_a = static_cast<c128*>(d_psi);
  c128* psi_b = vecp(0);
  c128* g_a = vecp(1);
  c128* g_b = vecp(2);
  c128* ra = vecp(3);
  c128* rb = vecp(4);
  std::vector<c128*> u(kmax);
  for (int i = 1; i < kmax; ++i) u[i] = vecp(4 + i);
  C.dens_ws = reinterpret_cast<c128*>(base + vec * (5 + kmax));
  C.dens_bytes = ws_bytes - vec * (5 + kmax);
  RPM_HIP_CHECK(hipMemsetAsync(g_a, 0, static_cast<size_t>(C.n) * batch * sizeof(c128), C.s));
  const c128* pairs0 = static_cast<const c128*>(d_pairs);
  c128* acc0 = static_cast<c128*>(d_acc);
  int rc = RPM_OK;
  for (int k = nint; k >= 1; --k) {
    const int iv = k - 1;                     // interval [t_{k-1}, t_k]
    const int K = nterms[cls[iv]];
    const double* cf = coef + static_cast<size_t>(cls[iv]) * (kmax + 1) * 2;
    const c128* pairs = pairs0 ? pairs0 + static_cast<long long>(seg[iv]) * batch * 16 : nullptr;
    c128* acc = acc0 + static_cast<long long>(slot[iv]) * batch * 16;
    // costate source of the output at t_k: g += 2 W_k P psi_k
    if ((rc = rpm_singlet_kick(psi_a, g_a, batch, C.n, C.nn, 2.0 * weights[k], C.s))) return rc;
    // psi_{k-1} = U^H psi_k
    if ((rc = clenshaw(C, psi_a, cf, K, pairs, nullptr, nullptr, ra, rb, psi_b))) return rc;
    // u_n = T_n(X) psi_{k-1}, n < K
    u[0] = psi_b;
    if ((rc = affine(C, u[0], u[1], pairs, 1.0 / C.hs, -C.shift / C.hs, nullptr, cmake(0.0, 0.0), nullptr, 0.0)))
      return rc;
    for (int n = 1; n + 1 < K; ++n)
      if ((rc = affine(C, u[n], u[n + 1], pairs, 2.0 / C.hs, -2.0 * C.shift / C.hs, nullptr, cmake(0.0, 0.0), u[n - 1],
                       -1.0)))
        return rc;
    // g_{k-1}' = U^H g_k with the interval's derivative densities
    if ((rc = clenshaw(C, g_a, cf, K, pairs, u.data(), acc, ra, rb, g_b))) return rc;
    std::swap(g_a, g_b);
    std::swap(psi_a, psi_b);                  // psi_{k-1} becomes the current state
  }
  if (psi_a != d_psi)
    RPM_HIP_CHECK(hipMemcpyAsync(d_psi, psi_a, static_cast<size_t>(C.n) * batch * sizeof(c128),
                                 hipMemcpyDeviceToDevice, C.s));
  return RPM_OK;
}
