// Device pieces of the adjoint (costate) sweep and of segmented propagation:
//   * electron cross-densities C[e][e'] = sum_n conj(a[e, n]) b[e', n], the
//     4x4 reduced objects every electron-only derivative dH/dp needs;
//   * the costate seeds of the Simpson-yield adjoint sweep;
//   * an instantaneous local unitary on one tensor axis of a state
//     (R/poster_reproduction/core.py:285-292).
// All batched over vectors (grid.y); reductions are two-stage in a fixed order.
#include <utility>
#include <vector>

#include "rpm_common.hpp"

namespace rpm {

__global__ void __launch_bounds__(256) k_zero(c128* __restrict__ x, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) x[i] = c128{0.0, 0.0};
}

constexpr int kDensThreads = 256;
constexpr int kDensPer = 8;   // nuclear indices per thread

// partial[(b * nblk + blk) * 16 + e*4 + e'] = weight * (sum over the block's nuclear indices of
// conj(a_e) b_e'), added to the partial already there when `accumulate` (the sweep sums all
// Chebyshev orders of a step in the block partials and reduces them once per step)
__global__ void __launch_bounds__(kDensThreads) k_cross_density(const c128* __restrict__ A, const c128* __restrict__ Bv,
                                                                long long ld, long long nn, c128* __restrict__ partial,
                                                                double weight, int accumulate) {
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
    if (threadIdx.x == 0) {
      c128 v = accumulate ? out[i] : cmake(0.0, 0.0);
      v.re = fma(weight, t.re, v.re);
      v.im = fma(weight, t.im, v.im);
      out[i] = v;
    }
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
                     static_cast<const c128*>(d_b), ld, nn, part, 1.0, 0);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_density_finish, dim3(batch), dim3(64), 0, s, part, nblk, weight, static_cast<c128*>(d_acc));
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

namespace rpm {

// chi[x] = u[e=1][x] - u[e=2][x] for x < nn (the singlet part of u, times sqrt2).
// chi of vector set z (grid.z) at u + z * ustride, into chi + z * cstride
__global__ void __launch_bounds__(256) k_singlet_diff(const c128* __restrict__ u, c128* __restrict__ chi, long long ld,
                                                      long long cld, long long nn, long long ustride,
                                                      long long cstride) {
  const long long b = blockIdx.y;
  const long long x = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (x >= nn) return;
  const c128* uz = u + blockIdx.z * ustride;
  const c128 p = uz[b * ld + nn + x], q = uz[b * ld + 2 * nn + x];
  chi[blockIdx.z * cstride + b * cld + x] = csub(p, q);
}

// Singlet sums of the costate seeds of every Chebyshev order of a step, in one pass over the
// chi_j = u_j[1] - u_j[2] (chi: nj blocks of cstride, trajectory b at b * cld):
//   S_m = sum_j G[m][j] chi_j,  m in [blockIdx.z * kSeedRows, ...) and m < nm,
// summed over j in ascending order for every m.
constexpr int kSeedRows = 8;
constexpr int kMaxTerms = 255;     // Chebyshev terms per sweep step: K + 1 <= 256 columns of G in LDS
__global__ void __launch_bounds__(256) k_seed_sums(const c128* __restrict__ chi, long long cstride, long long cld,
                                                   int nj, const c128* __restrict__ G, int gstride, int nm,
                                                   c128* __restrict__ S, long long nn) {
  const long long b = blockIdx.y;
  const long long x = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const int m0 = static_cast<int>(blockIdx.z) * kSeedRows;
  __shared__ c128 gs[kSeedRows][kMaxTerms + 1];
  for (int t = threadIdx.x; t < kSeedRows * nj; t += 256) {
    const int i = t / nj, j = t % nj;
    gs[i][j] = m0 + i < nm ? G[static_cast<long long>(m0 + i) * gstride + j] : cmake(0.0, 0.0);
  }
  __syncthreads();
  if (x >= nn) return;
  c128 acc[kSeedRows];
#pragma unroll
  for (int i = 0; i < kSeedRows; ++i) acc[i] = cmake(0.0, 0.0);
  for (int j = 0; j < nj; ++j) {
    const c128 c = chi[j * cstride + b * cld + x];
#pragma unroll
    for (int i = 0; i < kSeedRows; ++i) acc[i] = cfma(gs[i][j], c, acc[i]);
  }
#pragma unroll
  for (int i = 0; i < kSeedRows; ++i)
    if (m0 + i < nm) S[(m0 + i) * cstride + b * cld + x] = acc[i];
}

// Costate seed of one Chebyshev order m:
//   out = cg g + P sum_j G[m][j] u_j,  (P v)[1] = -(P v)[2] = (v[1] - v[2]) / 2,
// with the singlet sum S_m = sum_j G[m][j] chi_j from k_seed_sums.
__global__ void __launch_bounds__(256) k_seed(const c128* __restrict__ g, c128 cg, const c128* __restrict__ Sm,
                                              long long cld, double mult, c128* __restrict__ out, long long ld,
                                              long long nn) {
  const long long b = blockIdx.y;
  const long long x = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (x >= nn) return;
  const c128 s = cscale(Sm[b * cld + x], 0.5 * mult);
  const c128* gb = g + b * ld;
  c128* ob = out + b * ld;
  ob[x] = cfma(cg, gb[x], cmake(0.0, 0.0));
  ob[nn + x] = cfma(cg, gb[nn + x], s);
  ob[2 * nn + x] = cfma(cg, gb[2 * nn + x], cmake(-s.re, -s.im));
  ob[3 * nn + x] = cfma(cg, gb[3 * nn + x], cmake(0.0, 0.0));
}

}  // namespace rpm

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

// weight * C(a, b) added to the sweep step's density block partials (started by `first`); the
// step's partials are reduced into its accumulator once, by density_finish
int density(const SweepCtx& C, const c128* a, const c128* b, double weight, bool first) {
  const int nblk = ceil_div(C.nn, kDensThreads * kDensPer);
  hipLaunchKernelGGL(k_cross_density, dim3(nblk, C.batch), dim3(kDensThreads), 0, C.s, a, b, C.n, C.nn, C.dens_ws,
                     weight, first ? 0 : 1);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

int density_finish(const SweepCtx& C, c128* acc) {
  const int nblk = ceil_div(C.nn, kDensThreads * kDensPer);
  hipLaunchKernelGGL(k_density_finish, dim3(C.batch), dim3(64), 0, C.s, C.dens_ws, nblk, 1.0, acc);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

inline c128 conjc(const double* c) { return cmake(c[0], -c[1]); }

// out = sum_n conj(c_n) T_n(X) seed by Clenshaw's recurrence (K >= 2).  ra, rb: scratch.
int clenshaw_state(const SweepCtx& C, const c128* seed, const double* coef, int K, const c128* pairs, c128* ra,
                   c128* rb, c128* out) {
  const double s2 = 2.0 / C.hs, a2 = -2.0 * C.shift / C.hs;
  int rc = scaled_copy(C, conjc(coef + 2 * K), seed, ra);
  if (rc) return rc;
  if ((rc = affine(C, ra, rb, pairs, s2, a2, seed, conjc(coef + 2 * (K - 1)), nullptr, 0.0))) return rc;
  c128* hi = ra;
  c128* lo = rb;
  for (int n = K - 2; n >= 1; --n) {
    if ((rc = affine(C, lo, hi, pairs, s2, a2, seed, conjc(coef + 2 * n), hi, -1.0))) return rc;
    std::swap(hi, lo);
  }
  return affine(C, lo, out, pairs, 1.0 / C.hs, -C.shift / C.hs, seed, conjc(coef), hi, -1.0);
}

}  // namespace

extern "C" size_t rpm_adjoint_workspace_bytes(const rpm_operator* op, int batch, int kmax) {
  if (!op || batch < 1 || kmax < 2) return 0;
  const size_t vec = ((static_cast<size_t>(op->n) * batch * sizeof(c128)) + 255) & ~size_t(255);
  const size_t qvec = ((static_cast<size_t>(op->n / 4) * batch * sizeof(c128)) + 255) & ~size_t(255);
  // psi', g (2), ring a/b, seed z, u_1..u_K; then chi_0..chi_K and the seed sums S_0..S_K
  const size_t vecs = 6 + static_cast<size_t>(kmax);
  return vecs * vec + 2 * (kmax + 1) * qvec + rpm_cross_density_workspace_bytes(op->n / 4, batch) + 256;
}

extern "C" int rpm_adjoint_sweep(const rpm_operator* op, int batch, void* d_psi, int nmacro, double shift,
                                 double half_width, int ncls, const int* nterms, int kmax, const double* coef,
                                 const void* d_G, const int* cls, const double* mult, const void* d_pairs,
                                 const int* seg, int nseg, void* d_acc, const int* slot, int nslot, void* d_ws,
                                 size_t ws_bytes, void* stream) {
  if (!op || op->kind != 1 || !d_psi || nmacro < 1 || !nterms || !coef || !d_G || !cls || !mult || !d_acc ||
      !slot || !d_ws || batch < 1 || batch > 65535 || ncls < 1 || kmax < 2 || kmax > kMaxTerms || !(half_width > 0.0) ||
      nslot < 1)
    return fail(RPM_E_INVALID, "rpm_adjoint_sweep: bad arguments");
  if (d_pairs && (!seg || nseg < 1))
    return fail(RPM_E_INVALID, "rpm_adjoint_sweep: per-step segments and their count needed with pairs");
  for (int c = 0; c < ncls; ++c)
    if (nterms[c] < 2 || nterms[c] > kmax) return fail(RPM_E_INVALID, "Chebyshev length %d outside 2..%d", nterms[c], kmax);
  // every index the kernels use to address d_G, d_acc and d_pairs is checked here, on the host
  for (int k = 0; k < nmacro; ++k)
    if (cls[k] < 0 || cls[k] >= ncls || slot[k] < 0 || slot[k] >= nslot ||
        (d_pairs && (seg[k] < 0 || seg[k] >= nseg)))
      return fail(RPM_E_INVALID, "rpm_adjoint_sweep: step %d has a bad class, slot or segment", k);
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
  const size_t qvec = ((static_cast<size_t>(C.nn) * batch * sizeof(c128)) + 255) & ~size_t(255);
  char* base = static_cast<char*>(d_ws);
  auto vecp = [&](int i) { return reinterpret_cast<c128*>(base + vec * i); };
  c128* psi_a = static_cast<c128*>(d_psi);
  c128* psi_b = vecp(0);
  c128* g_a = vecp(1);
  c128* g_b = vecp(2);
  c128* ra = vecp(3);
  c128* rb = vecp(4);
  c128* z = vecp(5);
  std::vector<c128*> u(kmax + 1);
  for (int i = 1; i <= kmax; ++i) u[i] = vecp(5 + i);
  c128* chi = reinterpret_cast<c128*>(base + vec * (6 + kmax));
  c128* ssum = reinterpret_cast<c128*>(base + vec * (6 + kmax) + qvec * (kmax + 1));
  const long long cstride = static_cast<long long>(qvec / sizeof(c128));   // per order j
  C.dens_ws = reinterpret_cast<c128*>(base + vec * (6 + kmax) + 2 * qvec * (kmax + 1));
  C.dens_bytes = ws_bytes - (vec * (6 + kmax) + 2 * qvec * (kmax + 1));
  if (C.dens_bytes < rpm_cross_density_workspace_bytes(C.nn, batch))
    return fail(RPM_E_INVALID, "rpm_adjoint_sweep: workspace too small");
  {
    // a fill kernel, not hipMemsetAsync: memset nodes misbehave on replay when a caller captures
    // this entry point in a HIP graph
    const long long total = C.n * static_cast<long long>(batch);
    hipLaunchKernelGGL(k_zero, dim3(ceil_div(total, 256)), dim3(256), 0, C.s, g_a, total);
    RPM_LAUNCH_CHECK();
  }
  const c128* pairs0 = static_cast<const c128*>(d_pairs);
  const c128* G0 = static_cast<const c128*>(d_G);
  c128* acc0 = static_cast<c128*>(d_acc);
  const double s2 = 2.0 / C.hs, a2 = -2.0 * C.shift / C.hs;
  const int nqb = ceil_div(C.nn, 256);
  int rc = RPM_OK;
  for (int k = nmacro - 1; k >= 0; --k) {
    const int K = nterms[cls[k]];
    const double* cf = coef + static_cast<size_t>(cls[k]) * (kmax + 1) * 2;
    const c128* G = G0 + static_cast<size_t>(cls[k]) * (kmax + 1) * (kmax + 1);
    const c128* pairs = pairs0 ? pairs0 + static_cast<long long>(seg[k]) * batch * 16 : nullptr;
    c128* acc = acc0 + static_cast<long long>(slot[k]) * batch * 16;
    // psi at the start of the step: psi_a = U^H psi_b
    if ((rc = clenshaw_state(C, psi_a, cf, K, pairs, ra, rb, psi_b))) return rc;
    // u_j = T_j(X) psi_a, j <= K, and their singlet parts
    u[0] = psi_b;
    if ((rc = affine(C, u[0], u[1], pairs, 1.0 / C.hs, -C.shift / C.hs, nullptr, cmake(0.0, 0.0), nullptr, 0.0)))
      return rc;
    for (int n = 1; n < K; ++n)
      if ((rc = affine(C, u[n], u[n + 1], pairs, s2, a2, nullptr, cmake(0.0, 0.0), u[n - 1], -1.0))) return rc;
    // chi_0 from psi_a, chi_1..chi_K from the stored T_j(X) psi_a (evenly spaced in the workspace)
    hipLaunchKernelGGL(k_singlet_diff, dim3(nqb, batch, 1), dim3(256), 0, C.s, u[0], chi, C.n, C.nn, C.nn, 0LL, 0LL);
    RPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_singlet_diff, dim3(nqb, batch, K), dim3(256), 0, C.s, u[1], chi + cstride, C.n, C.nn, C.nn,
                       static_cast<long long>(vec / sizeof(c128)), cstride);
    RPM_LAUNCH_CHECK();
    // Clenshaw over the costate with vector seeds z_m = conj(c_m) g + mult P sum_j G[m][j] u_j; the
    // singlet sums of all orders in one pass over the chi_j
    hipLaunchKernelGGL(k_seed_sums, dim3(nqb, batch, ceil_div(K + 1, kSeedRows)), dim3(256), 0, C.s, chi, cstride,
                       C.nn, K + 1, G, kmax + 1, K + 1, ssum, C.nn);
    RPM_LAUNCH_CHECK();
    auto seed = [&](int m) {
      hipLaunchKernelGGL(k_seed, dim3(nqb, batch), dim3(256), 0, C.s, g_a, conjc(cf + 2 * m), ssum + m * cstride,
                         C.nn, mult[k], z, C.n, C.nn);
      return hipGetLastError() == hipSuccess ? RPM_OK : fail(RPM_E_HIP, "k_seed launch failed");
    };
    if ((rc = seed(K))) return rc;
    c128* hi = ra;   // nu_{n+2}
    c128* lo = rb;   // nu_{n+1}
    RPM_HIP_CHECK(hipMemcpyAsync(hi, z, static_cast<size_t>(C.n) * batch * sizeof(c128), hipMemcpyDeviceToDevice,
                                 C.s));                                                  // nu_K
    if ((rc = density(C, hi, u[K - 1], (K - 1 >= 1 ? 2.0 : 1.0) / C.hs, true))) return rc;
    if ((rc = seed(K - 1))) return rc;
    if ((rc = affine(C, hi, lo, pairs, s2, a2, z, cmake(1.0, 0.0), nullptr, 0.0))) return rc;   // nu_{K-1}
    if ((rc = density(C, lo, u[K - 2], (K - 2 >= 1 ? 2.0 : 1.0) / C.hs, false))) return rc;
    for (int n = K - 2; n >= 1; --n) {
      if ((rc = seed(n))) return rc;
      if ((rc = affine(C, lo, hi, pairs, s2, a2, z, cmake(1.0, 0.0), hi, -1.0))) return rc;     // nu_n
      if ((rc = density(C, hi, u[n - 1], (n - 1 >= 1 ? 2.0 : 1.0) / C.hs, false))) return rc;
      std::swap(hi, lo);
    }
    // lo = nu_1, hi = nu_2: costate carried back to the start of the step
    if ((rc = seed(0))) return rc;
    if ((rc = affine(C, lo, g_b, pairs, 1.0 / C.hs, -C.shift / C.hs, z, cmake(1.0, 0.0), hi, -1.0))) return rc;
    if ((rc = density_finish(C, acc))) return rc;
    std::swap(g_a, g_b);
    std::swap(psi_a, psi_b);
  }
  if (psi_a != d_psi)
    RPM_HIP_CHECK(hipMemcpyAsync(d_psi, psi_a, static_cast<size_t>(C.n) * batch * sizeof(c128),
                                 hipMemcpyDeviceToDevice, C.s));
  return RPM_OK;
}
