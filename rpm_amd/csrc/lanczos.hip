// Lanczos basis with classical Gram-Schmidt applied twice (CGS2), norms,
// singlet projection of a Krylov basis and the basis-combination update.
//
// Follows R/rpm_statevector_engine.py:651-688 (basis), 492-497 (singlet
// population), 521-529 (projected observable), 904 (endpoint update).
//
// All reductions are two-stage with a fixed summation order, so results are
// bitwise reproducible run to run.  Every basis kernel reads a device flag
// and returns immediately once the basis has broken down (beta <= tol), so
// the host never waits inside the basis loop.
#include <cmath>
#include <cstring>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kThreads = 256;
constexpr int kEPT = 4;                        // elements per thread in vector passes
constexpr int kChunk = kThreads * kEPT;        // elements per block
constexpr int kColGroup = 8;                   // basis columns per block in the dot pass

struct BasisCtl {
  int active;       // 0 once the basis broke down
  int size;         // basis size built so far / final
  double scale;     // 1/norm of the vector that the next matvec normalises
  double pad;
};

// ---------------------------------------------------------------- plain reductions

// partial[b * nblk + blk] = sum |x|^2 over the block's chunk of vector b
__global__ void __launch_bounds__(kThreads) k_sumsq(const c128* __restrict__ x, long long n, long long ld,
                                                    double* __restrict__ partial, const int* active) {
  if (active && *active == 0) return;
  __shared__ double sm[kThreads / kWave];
  const long long b = blockIdx.y;
  const c128* v = x + b * ld;
  const long long base = static_cast<long long>(blockIdx.x) * kChunk + threadIdx.x;
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const long long i = base + static_cast<long long>(e) * kThreads;
    if (i < n) {
      const c128 a = v[i];
      s = fma(a.re, a.re, s);
      s = fma(a.im, a.im, s);
    }
  }
  s = block_sum(s, sm);
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = s;
}

// singlet population partials: sum over nuclear n of |(x[nn+n] - x[2nn+n])|^2 / 2
__global__ void __launch_bounds__(kThreads) k_singlet_sumsq(const c128* __restrict__ x, long long nn, long long ld,
                                                            double* __restrict__ partial) {
  __shared__ double sm[kThreads / kWave];
  const long long b = blockIdx.y;
  const c128* v = x + b * ld;
  const long long base = static_cast<long long>(blockIdx.x) * kChunk + threadIdx.x;
  const double r2 = 1.0 / std::sqrt(2.0);
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const long long i = base + static_cast<long long>(e) * kThreads;
    if (i < nn) {
      const c128 a = v[nn + i], c = v[2 * nn + i];
      const double re = (a.re - c.re) * r2, im = (a.im - c.im) * r2;
      s = fma(re, re, s);
      s = fma(im, im, s);
    }
  }
  s = block_sum(s, sm);
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = s;
}

// out[b] = sqrt(sum partial) (take_sqrt) or the plain sum; one block per b.
__global__ void __launch_bounds__(kThreads) k_finish_sum(const double* __restrict__ partial, int nblk,
                                                         double* __restrict__ out, int take_sqrt) {
  __shared__ double sm[kThreads / kWave];
  const int b = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kThreads) s += partial[static_cast<long long>(b) * nblk + i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) out[b] = take_sqrt ? std::sqrt(s) : s;
}

// ---------------------------------------------------------------- basis kernels

// Norm of the starting state: norm, 1/norm into ctl, reset the control block.
__global__ void k_basis_start(const double* __restrict__ partial, int nblk, double* __restrict__ norm_out,
                              BasisCtl* ctl, int count) {
  __shared__ double sm[kThreads / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kThreads) s += partial[i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    const double nrm = std::sqrt(s);
    *norm_out = nrm;
    ctl->active = nrm > 0.0 ? 1 : 0;
    ctl->size = nrm > 0.0 ? count : 0;
    ctl->scale = nrm > 0.0 ? 1.0 / nrm : 0.0;
  }
}

// partial[c * nchunk + blk] = sum_x conj(V_c[x]) w[x] over the block's chunk, for c in the group.
__global__ void __launch_bounds__(kThreads) k_dots(const c128* __restrict__ V, long long n, int k,
                                                   const c128* __restrict__ w, c128* __restrict__ partial,
                                                   const BasisCtl* ctl) {
  if (ctl && ctl->active == 0) return;
  __shared__ c128 sm[kThreads / kWave];
  const int nchunk = gridDim.x;
  const long long base = static_cast<long long>(blockIdx.x) * kChunk + threadIdx.x;
  c128 wv[kEPT];
#pragma unroll
  for (int e = 0; e < kEPT; ++e) {
    const long long i = base + static_cast<long long>(e) * kThreads;
    wv[e] = i < n ? w[i] : cmake(0.0, 0.0);
  }
  const int c0 = blockIdx.y * kColGroup;
  const int c1 = min(k, c0 + kColGroup);
  for (int c = c0; c < c1; ++c) {
    const c128* v = V + static_cast<long long>(c) * n;
    c128 s{0.0, 0.0};
#pragma unroll
    for (int e = 0; e < kEPT; ++e) {
      const long long i = base + static_cast<long long>(e) * kThreads;
      if (i < n) s = cfmaconj(v[i], wv[e], s);
    }
    s = block_sum(s, sm);
    if (threadIdx.x == 0) partial[static_cast<long long>(c) * nchunk + blockIdx.x] = s;
  }
}

// h[c] = sum_blk partial[c][blk]; optionally alpha[alpha_slot] = Re h[alpha_col].
__global__ void __launch_bounds__(kThreads) k_reduce_cols(const c128* __restrict__ partial, int nchunk,
                                                          c128* __restrict__ h, double* __restrict__ alpha,
                                                          int alpha_col, int alpha_slot, const BasisCtl* ctl) {
  if (ctl && ctl->active == 0) return;
  __shared__ c128 sm[kThreads / kWave];
  const int c = blockIdx.x;
  c128 s{0.0, 0.0};
  for (int i = threadIdx.x; i < nchunk; i += kThreads) s = cadd(s, partial[static_cast<long long>(c) * nchunk + i]);
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    h[c] = s;
    if (alpha && c == alpha_col) alpha[alpha_slot] = s.re;
  }
}

// w -= sum_c V_c h_c over the chunk (one element per thread); optional partial sums of |w|^2.
constexpr int kUpdChunk = kThreads;
__global__ void __launch_bounds__(kThreads) k_update(const c128* __restrict__ V, long long n, int k,
                                                     const c128* __restrict__ h, c128* __restrict__ w,
                                                     double* __restrict__ norm_partial, const BasisCtl* ctl) {
  if (ctl && ctl->active == 0) return;
  __shared__ double sm[kThreads / kWave];
  const long long i = static_cast<long long>(blockIdx.x) * kUpdChunk + threadIdx.x;
  c128 wv = i < n ? w[i] : cmake(0.0, 0.0);
  if (i < n) {
    int c = 0;
    for (; c + 4 <= k; c += 4) {
      const c128 v0 = V[static_cast<long long>(c) * n + i], v1 = V[static_cast<long long>(c + 1) * n + i];
      const c128 v2 = V[static_cast<long long>(c + 2) * n + i], v3 = V[static_cast<long long>(c + 3) * n + i];
      wv = cfms(v0, h[c], wv);
      wv = cfms(v1, h[c + 1], wv);
      wv = cfms(v2, h[c + 2], wv);
      wv = cfms(v3, h[c + 3], wv);
    }
    for (; c < k; ++c) wv = cfms(V[static_cast<long long>(c) * n + i], h[c], wv);
    w[i] = wv;
  }
  if (norm_partial) {
    double s = fma(wv.re, wv.re, wv.im * wv.im);
    s = block_sum(s, sm);
    if (threadIdx.x == 0) norm_partial[blockIdx.x] = s;
  }
}

// Fused CGS pass over an LDS-staged chunk of the basis: optionally
// w <- w - V h (update), then partial[c][blk] = sum_x conj(V_c[x]) w[x] for
// all c < k.  Each block walks kGsPer chunks of kGsChunk elements; V is read
// from HBM once per pass.  k <= kGsMaxK.
constexpr int kGsChunk = 32;
constexpr int kGsPer = 4;
constexpr int kGsMaxK = 64;
__global__ void __launch_bounds__(256) k_gs_pass(const c128* __restrict__ V, long long n, int k,
                                                 const c128* __restrict__ h, c128* __restrict__ w,
                                                 c128* __restrict__ partial, const BasisCtl* ctl) {
  if (ctl && ctl->active == 0) return;
  __shared__ c128 vs[kGsMaxK][kGsChunk + 1];
  __shared__ c128 ws[kGsChunk];
  __shared__ c128 red[256];
  __shared__ c128 hs[kGsMaxK];
  const int tid = threadIdx.x;
  const int nblk = gridDim.x;
  if (h && tid < k) hs[tid] = h[tid];
  // dot mapping: column c = tid % 64, part q = tid / 64 (4 parts of 8 elements)
  const int dc = tid % 64, dq = tid / 64;
  c128 dot{0.0, 0.0};
  for (int it = 0; it < kGsPer; ++it) {
    const long long x0 = (static_cast<long long>(blockIdx.x) * kGsPer + it) * kGsChunk;
    if (x0 >= n) break;
    __syncthreads();
    for (int s = tid; s < k * kGsChunk; s += 256) {
      const int c = s / kGsChunk, i = s % kGsChunk;
      const long long x = x0 + i;
      vs[c][i] = x < n ? V[static_cast<long long>(c) * n + x] : cmake(0.0, 0.0);
    }
    if (tid < kGsChunk) {
      const long long x = x0 + tid;
      ws[tid] = x < n ? w[x] : cmake(0.0, 0.0);
    }
    __syncthreads();
    if (h) {
      // 8 threads per element, each over k/8 columns
      const int i = tid / 8, part = tid % 8;
      c128 s{0.0, 0.0};
      for (int c = part; c < k; c += 8) s = cfma(vs[c][i], hs[c], s);
      red[tid] = s;
      __syncthreads();
      if (part == 0) {
        c128 t = red[tid];
        for (int p = 1; p < 8; ++p) t = cadd(t, red[tid + p]);
        const c128 nw = csub(ws[i], t);
        ws[i] = nw;
        const long long x = x0 + i;
        if (x < n) w[x] = nw;
      }
      __syncthreads();
    }
    if (dc < k) {
#pragma unroll
      for (int i = dq * 8; i < dq * 8 + 8; ++i) dot = cfmaconj(vs[dc][i], ws[i], dot);
    }
  }
  __syncthreads();
  red[tid] = dot;
  __syncthreads();
  if (tid < 64 && tid < k) {
    const c128 t = cadd(cadd(red[tid], red[tid + 64]), cadd(red[tid + 128], red[tid + 192]));
    partial[static_cast<long long>(tid) * nblk + blockIdx.x] = t;
  }
}

// beta_j = ||w||; breakdown test; scale for the next matvec.
__global__ void __launch_bounds__(kThreads) k_finish_beta(const double* __restrict__ partial, int nchunk,
                                                          double* __restrict__ beta, int j, double tol,
                                                          BasisCtl* ctl) {
  if (ctl->active == 0) return;
  __shared__ double sm[kThreads / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < nchunk; i += kThreads) s += partial[i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    const double b = std::sqrt(s);
    beta[j] = b;
    if (b <= tol) {
      ctl->active = 0;
      ctl->size = j + 1;
    } else {
      ctl->scale = 1.0 / b;
    }
  }
}

__global__ void k_basis_info(const BasisCtl* ctl, int* info) {
  if (threadIdx.x == 0) {
    info[0] = ctl->size;
    info[1] = ctl->active;
  }
}

// ---------------------------------------------------------------- combine / gram

// out[x] = scale * sum_j V_j[x] c_j
__global__ void __launch_bounds__(kThreads) k_combine(const c128* __restrict__ V, int k, long long n,
                                                      const c128* __restrict__ c, const double* __restrict__ scale,
                                                      c128* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  c128 s{0.0, 0.0};
  for (int j = 0; j < k; ++j) s = cfma(V[static_cast<long long>(j) * n + i], c[j], s);
  const double sc = scale ? *scale : 1.0;
  out[i] = cscale(s, sc);
}

constexpr int kGramTile = 64;
constexpr int kGramChunk = 16;

// Partial S[a][b] = sum_n conj(A_a[n]) A_b[n] for a in [a0, a0+64), b in [b0, b0+64),
// A_a[n] = (V_a[nn + n] - V_a[2 nn + n]) / sqrt2.  One 4x4 register block per thread.
__global__ void __launch_bounds__(256) k_gram(const c128* __restrict__ V, int k, long long nn, int a0, int b0,
                                              long long n_per_block, c128* __restrict__ partial) {
  __shared__ c128 ra[kGramChunk][kGramTile + 1];
  __shared__ c128 rb[kGramChunk][kGramTile + 1];
  const int tid = threadIdx.x;
  const int ia = tid / 16, ib = tid % 16;
  const long long N = 4 * nn;
  const long long n0 = static_cast<long long>(blockIdx.x) * n_per_block;
  const long long n1 = min(nn, n0 + n_per_block);
  const double r2 = 1.0 / std::sqrt(2.0);
  c128 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = cmake(0.0, 0.0);
  for (long long c0 = n0; c0 < n1; c0 += kGramChunk) {
    // stage: 64 rows x kGramChunk columns for each side
    for (int s = tid; s < kGramTile * kGramChunk; s += 256) {
      const int row = s / kGramChunk, col = s % kGramChunk;
      const long long n = c0 + col;
      c128 va{0.0, 0.0}, vb{0.0, 0.0};
      if (n < n1) {
        if (a0 + row < k) {
          const c128* v = V + static_cast<long long>(a0 + row) * N;
          const c128 p = v[nn + n], q = v[2 * nn + n];
          va = cmake((p.re - q.re) * r2, (p.im - q.im) * r2);
        }
        if (b0 + row < k) {
          const c128* v = V + static_cast<long long>(b0 + row) * N;
          const c128 p = v[nn + n], q = v[2 * nn + n];
          vb = cmake((p.re - q.re) * r2, (p.im - q.im) * r2);
        }
      }
      ra[col][row] = va;
      rb[col][row] = vb;
    }
    __syncthreads();
    for (int col = 0; col < kGramChunk; ++col) {
      c128 xa[4], xb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xa[u] = ra[col][ia * 4 + u];
        xb[u] = rb[col][ib * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = cfmaconj(xa[u], xb[v], acc[u][v]);
    }
    __syncthreads();
  }
  c128* out = partial + static_cast<long long>(blockIdx.x) * kGramTile * kGramTile;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) out[(ia * 4 + u) * kGramTile + ib * 4 + v] = acc[u][v];
}

// S tile = sum over blocks of the partial tiles, written into S (k x k row-major)
__global__ void __launch_bounds__(256) k_gram_reduce(const c128* __restrict__ partial, int nblk, int k, int a0,
                                                     int b0, c128* __restrict__ S) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= kGramTile * kGramTile) return;
  const int a = a0 + e / kGramTile, b = b0 + e % kGramTile;
  if (a >= k || b >= k) return;
  c128 s{0.0, 0.0};
  for (int i = 0; i < nblk; ++i) s = cadd(s, partial[static_cast<long long>(i) * kGramTile * kGramTile + e]);
  S[a * k + b] = s;
}

// P = (S + S^H) / 2
__global__ void k_hermitize(const c128* __restrict__ S, int k, c128* __restrict__ P) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= k * k) return;
  const int a = e / k, b = e % k;
  const c128 x = S[a * k + b], y = S[b * k + a];
  P[e] = cmake(0.5 * (x.re + y.re), 0.5 * (x.im - y.im));
}

}  // namespace rpm

using namespace rpm;

// ---------------------------------------------------------------- C ABI

extern "C" size_t rpm_reduce_workspace_bytes(long long n, int batch) {
  const long long nblk = (n + kChunk - 1) / kChunk;
  return static_cast<size_t>(nblk) * (batch > 0 ? batch : 1) * sizeof(double) + 256;
}

extern "C" int rpm_norm(const void* d_x, long long n, int batch, long long ld, double* d_out, void* d_ws,
                        size_t ws_bytes, void* stream) {
  if (!d_x || !d_out || n < 1 || batch < 1 || batch > 65535 || ld < n)
    return fail(RPM_E_INVALID, "rpm_norm: bad arguments");
  if (ws_bytes < rpm_reduce_workspace_bytes(n, batch) || !d_ws)
    return fail(RPM_E_INVALID, "rpm_norm: workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nblk = ceil_div(n, kChunk);
  double* part = static_cast<double*>(d_ws);
  hipLaunchKernelGGL(k_sumsq, dim3(nblk, batch), dim3(kThreads), 0, s, static_cast<const c128*>(d_x), n, ld,
                     part, nullptr);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_finish_sum, dim3(batch), dim3(kThreads), 0, s, part, nblk, d_out, 1);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_singlet_population(const void* d_x, long long nn, int batch, long long ld, double* d_out,
                                      void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_x || !d_out || nn < 1 || batch < 1 || batch > 65535 || ld < 4 * nn)
    return fail(RPM_E_INVALID, "rpm_singlet_population: bad arguments");
  if (ws_bytes < rpm_reduce_workspace_bytes(nn, batch) || !d_ws)
    return fail(RPM_E_INVALID, "rpm_singlet_population: workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nblk = ceil_div(nn, kChunk);
  double* part = static_cast<double*>(d_ws);
  hipLaunchKernelGGL(k_singlet_sumsq, dim3(nblk, batch), dim3(kThreads), 0, s, static_cast<const c128*>(d_x), nn,
                     ld, part);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_finish_sum, dim3(batch), dim3(kThreads), 0, s, part, nblk, d_out, 0);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

namespace {
struct BasisWs {
  c128* w[2];
  c128* partial;     // [m][nchunk]
  c128* h;           // [m]
  double* npart;     // [nchunk]
  BasisCtl* ctl;
  size_t bytes;
};

BasisWs basis_layout(char* base, long long n, int m) {
  long long nchunk = (n + kChunk - 1) / kChunk;
  const long long gsb = (n + kGsChunk * kGsPer - 1) / (kGsChunk * kGsPer);
  const long long nupd = (n + kUpdChunk - 1) / kUpdChunk;
  if (gsb > nchunk) nchunk = gsb;
  if (nupd > nchunk) nchunk = nupd;
  size_t off = 0;
  auto take = [&](size_t b) { size_t o = off; off += (b + 255) & ~size_t(255); return o; };
  const size_t o_w0 = take(n * 16), o_w1 = take(n * 16), o_p = take(static_cast<size_t>(m) * nchunk * 16);
  const size_t o_h = take(static_cast<size_t>(m) * 16), o_np = take(nchunk * 8), o_ctl = take(sizeof(BasisCtl));
  BasisWs w;
  w.w[0] = reinterpret_cast<c128*>(base + o_w0);
  w.w[1] = reinterpret_cast<c128*>(base + o_w1);
  w.partial = reinterpret_cast<c128*>(base + o_p);
  w.h = reinterpret_cast<c128*>(base + o_h);
  w.npart = reinterpret_cast<double*>(base + o_np);
  w.ctl = reinterpret_cast<BasisCtl*>(base + o_ctl);
  w.bytes = off;
  return w;
}
}  // namespace

extern "C" size_t rpm_lanczos_workspace_bytes(const rpm_operator* op, int m) {
  if (!op || m < 1) return 0;
  return basis_layout(nullptr, op->n, m).bytes;
}

extern "C" int rpm_lanczos_basis(const rpm_operator* op, const void* d_state, int target, void* d_basis,
                                 double* d_alpha, double* d_beta, int* d_info, double* d_norm,
                                 double tol, void* d_ws, size_t ws_bytes, void* stream) {
  if (!op || !d_state || !d_basis || !d_alpha || !d_beta || !d_info || !d_norm || !d_ws || target < 1)
    return fail(RPM_E_INVALID, "rpm_lanczos_basis: bad arguments");
  const long long n = op->n;
  const int count = static_cast<int>(target < n ? target : n);
  BasisWs w = basis_layout(static_cast<char*>(d_ws), n, count);
  if (ws_bytes < w.bytes) return fail(RPM_E_INVALID, "rpm_lanczos_basis: workspace %zu < %zu", ws_bytes, w.bytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nchunk = ceil_div(n, kChunk);
  const int nchunk_upd = ceil_div(n, kUpdChunk);
  const int gs_blocks = ceil_div(n, kGsChunk * kGsPer);
  c128* V = static_cast<c128*>(d_basis);
  const c128* x0 = static_cast<const c128*>(d_state);

  hipLaunchKernelGGL(k_sumsq, dim3(nchunk, 1), dim3(kThreads), 0, s, x0, n, n, w.npart, nullptr);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_basis_start, dim3(1), dim3(kThreads), 0, s, w.npart, nchunk, d_norm, w.ctl, count);
  RPM_LAUNCH_CHECK();
  const int* active = &w.ctl->active;
  for (int j = 0; j < count; ++j) {
    const c128* in = j == 0 ? x0 : w.w[j % 2];
    c128* out = w.w[(j + 1) % 2];
    c128* vj = V + static_cast<long long>(j) * n;
    int rc = operator_apply_ex(op, in, out, 1, n, &w.ctl->scale, vj, active, s);
    if (rc) return rc;
    if (j == count - 1) {
      // last vector: only alpha_j = Re <v_j, H v_j> is used
      hipLaunchKernelGGL(k_dots, dim3(nchunk, 1), dim3(kThreads), 0, s, vj, n, 1, out, w.partial, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(kThreads), 0, s, w.partial, nchunk, w.h, d_alpha, 0, j,
                         w.ctl);
      RPM_LAUNCH_CHECK();
      break;
    }
    const int k = j + 1;
    if (k <= kGsMaxK) {
      // three passes over V: h1 = V^H w | w -= V h1, h2 = V^H w | w -= V h2, ||w||
      hipLaunchKernelGGL(k_gs_pass, dim3(gs_blocks), dim3(256), 0, s, V, n, k, nullptr, out, w.partial, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reduce_cols, dim3(k), dim3(kThreads), 0, s, w.partial, gs_blocks, w.h, d_alpha, j, j,
                         w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gs_pass, dim3(gs_blocks), dim3(256), 0, s, V, n, k, w.h, out, w.partial, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reduce_cols, dim3(k), dim3(kThreads), 0, s, w.partial, gs_blocks, w.h, nullptr, 0, 0,
                         w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_update, dim3(nchunk_upd), dim3(kThreads), 0, s, V, n, k, w.h, out, w.npart, w.ctl);
      RPM_LAUNCH_CHECK();
    } else {
      const dim3 gd(nchunk, ceil_div(k, kColGroup));
      // pass 1: h1 = V^H w (alpha_j = Re h1[j]); w -= V h1
      hipLaunchKernelGGL(k_dots, gd, dim3(kThreads), 0, s, V, n, k, out, w.partial, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reduce_cols, dim3(k), dim3(kThreads), 0, s, w.partial, nchunk, w.h, d_alpha, j, j, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_update, dim3(nchunk_upd), dim3(kThreads), 0, s, V, n, k, w.h, out, nullptr, w.ctl);
      RPM_LAUNCH_CHECK();
      // pass 2: h2 = V^H w; w -= V h2; beta = ||w||
      hipLaunchKernelGGL(k_dots, gd, dim3(kThreads), 0, s, V, n, k, out, w.partial, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reduce_cols, dim3(k), dim3(kThreads), 0, s, w.partial, nchunk, w.h, nullptr, 0, 0, w.ctl);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_update, dim3(nchunk_upd), dim3(kThreads), 0, s, V, n, k, w.h, out, w.npart, w.ctl);
      RPM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_finish_beta, dim3(1), dim3(kThreads), 0, s, w.npart, nchunk_upd, d_beta, j, tol, w.ctl);
    RPM_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_basis_info, dim3(1), dim3(64), 0, s, w.ctl, d_info);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_krylov_combine(const void* d_basis, int k, long long n, const void* d_c, const double* d_scale,
                                  void* d_out, void* stream) {
  if (!d_basis || !d_c || !d_out || k < 1 || n < 1) return fail(RPM_E_INVALID, "rpm_krylov_combine: bad arguments");
  hipLaunchKernelGGL(k_combine, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<const c128*>(d_basis), k, n, static_cast<const c128*>(d_c), d_scale,
                     static_cast<c128*>(d_out));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

namespace {
long long gram_blocks(long long nn) {
  // about 256 blocks, each an integer number of kGramChunk-wide chunks
  long long per = (nn + 255) / 256;
  per = ((per + kGramChunk - 1) / kGramChunk) * kGramChunk;
  if (per < kGramChunk) per = kGramChunk;
  return per;
}
}  // namespace

extern "C" size_t rpm_singlet_gram_workspace_bytes(int k, long long nn) {
  if (k < 1 || nn < 1) return 0;
  const long long per = gram_blocks(nn);
  const long long nblk = (nn + per - 1) / per;
  return static_cast<size_t>(nblk) * kGramTile * kGramTile * 16 + static_cast<size_t>(k) * k * 16 + 512;
}

extern "C" int rpm_singlet_gram(const void* d_basis, int k, long long nn, void* d_P, void* d_ws, size_t ws_bytes,
                                void* stream) {
  if (!d_basis || !d_P || !d_ws || k < 1 || nn < 1) return fail(RPM_E_INVALID, "rpm_singlet_gram: bad arguments");
  if (ws_bytes < rpm_singlet_gram_workspace_bytes(k, nn)) return fail(RPM_E_INVALID, "rpm_singlet_gram: workspace");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long per = gram_blocks(nn);
  const int nblk = static_cast<int>((nn + per - 1) / per);
  c128* partial = static_cast<c128*>(d_ws);
  c128* S = reinterpret_cast<c128*>(static_cast<char*>(d_ws) +
                                    ((static_cast<size_t>(nblk) * kGramTile * kGramTile * 16 + 255) & ~size_t(255)));
  for (int a0 = 0; a0 < k; a0 += kGramTile)
    for (int b0 = 0; b0 < k; b0 += kGramTile) {
      hipLaunchKernelGGL(k_gram, dim3(nblk), dim3(256), 0, s, static_cast<const c128*>(d_basis), k, nn, a0, b0, per,
                         partial);
      RPM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_gram_reduce, dim3(kGramTile * kGramTile / 256), dim3(256), 0, s, partial, nblk, k, a0, b0,
                         S);
      RPM_LAUNCH_CHECK();
    }
  hipLaunchKernelGGL(k_hermitize, dim3(ceil_div(static_cast<long long>(k) * k, 256)), dim3(256), 0, s, S, k,
                     static_cast<c128*>(d_P));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}
