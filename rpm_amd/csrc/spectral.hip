// Factorized spectral yields (no electron-electron coupling): transition-line
// binning of one radical's electron-spin correlation spectra on a uniform
// frequency grid, for a batch of field orientations.
//
// Follows R/rpm_quantum_needle_reproduction.py:145-172 (transition_lines and
// binned_spectrum). Line (i, j) of a radical with eigenvalues lambda and
// electron operators T_p in its eigenbasis has frequency lambda_i - lambda_j
// and weight T_p[i][j] T_q[j][i] / M. T_q is Hermitian, so the weight is
// computed as T_p[i][j] conj(T_q[i][j]) / M. The line sits at grid position
// pos = (lambda_i - lambda_j) / delta + half (divided, as the reference does,
// so every line gets the same position and fraction) and is split between
// bins floor(pos) and floor(pos) + 1 by linear interpolation.
//
// The (q, p) spectrum is the complex conjugate of the (p, q) one (same
// frequencies, conjugate weights), so only the six pairs p <= q are binned,
// in the order (0,0) (1,1) (2,2) (0,1) (0,2) (1,2).
//
// k_bin_tiles: one workgroup per (tile of kTileBins bins, orientation, pair).
// The eigenvalues are ascending, so for a fixed row i the computed pos is
// non-increasing in j (subtraction, division and addition all round
// monotonically): the lines of row i whose floor(pos) touches the tile form
// one contiguous j range, found by two binary searches on that same
// expression. Deposits go to the tile in LDS; the tile is then written once,
// coalesced, with zeros in the padding up to the grid stride (the FFT input
// needs no separate clear). Each bin belongs to exactly one tile, so every
// bin receives exactly the reference's terms; only their summation order
// (LDS atomics) varies.
#include <rocsolver/rocsolver.h>

#include <memory>
#include <mutex>
#include <vector>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kSpectralPairs = 6;
constexpr int kTileBins = 4032;        // 63 KB of LDS (re, im); with the row tables 80 KB: two workgroups per CU
constexpr int kBinThreads = 512;
constexpr int kFlatRows = 1024;        // radicals up to this dimension keep eigenvalues and row ranges in LDS
__constant__ int kPairP[kSpectralPairs] = {0, 1, 2, 0, 0, 1};
__constant__ int kPairQ[kSpectralPairs] = {0, 1, 2, 1, 2, 2};

__device__ __forceinline__ double line_pos(double li, double lj, double delta, double half) {
  return (li - lj) / delta + half;
}

// First j in [0, d) with line_pos(li, lam[j]) < x (line_pos is non-increasing in j); d if none.
__device__ __forceinline__ int first_below(const double* __restrict__ lam, int d, double li, double delta,
                                           double half, double x) {
  int lo = 0, hi = d;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (line_pos(li, lam[mid], delta, half) < x)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

__device__ __forceinline__ void deposit_line(double* __restrict__ t_re, double* __restrict__ t_im, bool diag,
                                             const c128* __restrict__ tp_row, const c128* __restrict__ tq_row,
                                             int d, int i, int j, double li, double lj, double m, double delta,
                                             double hf, long long b0, long long b_end) {
  const double pos = line_pos(li, lj, delta, hf);
  const double fl = floor(pos);
  const long long lower = static_cast<long long>(fl);
  const double f = pos - fl;
  const c128 a = tp_row[static_cast<long long>(i) * d + j];
  double wr, wi = 0.0;
  if (diag) {
    wr = (a.re * a.re + a.im * a.im) / m;
  } else {
    const c128 b = tq_row[static_cast<long long>(i) * d + j];
    wr = (a.re * b.re + a.im * b.im) / m;
    wi = (a.im * b.re - a.re * b.im) / m;
  }
  if (lower >= b0 && lower < b_end) {
    atomicAdd(&t_re[lower - b0], wr * (1.0 - f));
    if (!diag) atomicAdd(&t_im[lower - b0], wi * (1.0 - f));
  }
  if (lower + 1 >= b0 && lower + 1 < b_end) {
    atomicAdd(&t_re[lower + 1 - b0], wr * f);
    if (!diag) atomicAdd(&t_im[lower + 1 - b0], wi * f);
  }
}

// kFlat (d <= kFlatRows): eigenvalues, per-row j ranges and their prefix sums live in LDS, and the
// tile's lines are dealt to the threads in order (line L to thread L mod kBinThreads), so that
// neighbouring lanes read neighbouring T entries of one row and no lane walks a long range alone.
// Otherwise each thread walks whole rows, searching the eigenvalues in global memory.
template <bool kFlat>
__global__ void __launch_bounds__(kBinThreads) k_bin_tiles(const double* __restrict__ lam_all,
                                                           const c128* __restrict__ ops, int d, double m,
                                                           double delta, long long half, c128* __restrict__ grid,
                                                           long long gstride, int* __restrict__ overflow,
                                                           int accumulate) {
  __shared__ double t_re[kTileBins];
  __shared__ double t_im[kTileBins];
  __shared__ double s_lam[kFlat ? kFlatRows : 1];
  __shared__ int s_j0[kFlat ? kFlatRows : 1];
  __shared__ int s_pre[kFlat ? kFlatRows : 1];
  __shared__ int s_wave[kBinThreads / kWave];
  __shared__ int s_total;
  const long long o = blockIdx.y;
  const int pair = blockIdx.z;
  const int p = kPairP[pair], q = kPairQ[pair];
  const bool diag = (p == q);
  const long long size = 2 * half + 1;
  const long long b0 = static_cast<long long>(blockIdx.x) * kTileBins;
  const double* lam_g = lam_all + o * d;
  const double hf = static_cast<double>(half);
  c128* g = grid + (o * kSpectralPairs + pair) * gstride;

  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // The extreme lines bound every computed position (monotone rounding): the reference's
    // "grid too small" condition, lower < 0 or lower + 1 >= size, holds for some line iff it
    // holds for one of these two.
    const double lo_pos = floor(line_pos(lam_g[0], lam_g[d - 1], delta, hf));
    const double hi_pos = floor(line_pos(lam_g[d - 1], lam_g[0], delta, hf));
    if (!(lo_pos >= 0.0) || !(hi_pos + 1.0 < static_cast<double>(size))) atomicAdd(overflow, 1);
  }
  if (b0 >= size) {
    // padding of the FFT input: zeros (block-uniform branch); nothing to add when accumulating
    if (!accumulate)
      for (int k = threadIdx.x; k < kTileBins; k += kBinThreads) {
        const long long idx = b0 + k;
        if (idx < gstride) g[idx] = c128{0.0, 0.0};
      }
    return;
  }

  for (int k = threadIdx.x; k < kTileBins; k += kBinThreads) {
    t_re[k] = 0.0;
    t_im[k] = 0.0;
  }
  const long long dd = static_cast<long long>(d) * d;
  const c128* tp_row = ops + (o * 3 + p) * dd;
  const c128* tq_row = ops + (o * 3 + q) * dd;
  const double x_hi = static_cast<double>(b0 + kTileBins);   // pos < b0 + kTileBins
  const double x_lo = static_cast<double>(b0 - 1);           // pos >= b0 - 1
  const long long b_end = (b0 + kTileBins < size) ? b0 + kTileBins : size;   // bins of this tile in the grid

  if constexpr (kFlat) {
    for (int k = threadIdx.x; k < d; k += kBinThreads) s_lam[k] = lam_g[k];
    __syncthreads();
    // 1. the j range [j0, j1) of every row; the length goes to s_pre for the scan
    for (int i = threadIdx.x; i < d; i += kBinThreads) {
      const double li = s_lam[i];
      const int j0 = first_below(s_lam, d, li, delta, hf, x_hi);
      const int j1 = first_below(s_lam, d, li, delta, hf, x_lo);
      s_j0[i] = j0;
      s_pre[i] = j1 - j0;
    }
    __syncthreads();
    // 2. exclusive prefix sum of the lengths over rows (each thread owns `chunk` consecutive rows)
    const int chunk = (d + kBinThreads - 1) / kBinThreads;
    const int r0 = threadIdx.x * chunk;
    const int r1 = (r0 + chunk < d) ? r0 + chunk : d;
    int len[kFlatRows / kBinThreads];
    int local = 0;
#pragma unroll
    for (int c = 0; c < kFlatRows / kBinThreads; ++c) {
      len[c] = (r0 + c < r1) ? s_pre[r0 + c] : 0;
      local += len[c];
    }
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    int incl = local;
    for (int off = 1; off < kWave; off <<= 1) {
      const int t = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += t;
    }
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();   // every thread has read its lengths; s_pre may now be overwritten
    if (threadIdx.x == 0) {
      int run = 0;
      for (int w = 0; w < kBinThreads / kWave; ++w) {
        const int t = s_wave[w];
        s_wave[w] = run;
        run += t;
      }
      s_total = run;
    }
    __syncthreads();
    int run = s_wave[wave] + incl - local;
#pragma unroll
    for (int c = 0; c < kFlatRows / kBinThreads; ++c) {
      if (r0 + c < r1) s_pre[r0 + c] = run;
      run += len[c];
    }
    __syncthreads();
    // 3. line L = (row, j): row = last r with s_pre[r] <= L
    const int total = s_total;
    for (int L = threadIdx.x; L < total; L += kBinThreads) {
      int lo = 0, hi = d;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_pre[mid] > L)
          hi = mid;
        else
          lo = mid + 1;
      }
      const int i = lo - 1;
      const int j = s_j0[i] + (L - s_pre[i]);
      if (i < 0 || j >= d) continue;   // unreachable when the scan is consistent; keeps every read in bounds
      deposit_line(t_re, t_im, diag, tp_row, tq_row, d, i, j, s_lam[i], s_lam[j], m, delta, hf, b0, b_end);
    }
  } else {
    __syncthreads();
    for (int i = threadIdx.x; i < d; i += kBinThreads) {
      const double li = lam_g[i];
      const int j0 = first_below(lam_g, d, li, delta, hf, x_hi);
      const int j1 = first_below(lam_g, d, li, delta, hf, x_lo);
      for (int j = j0; j < j1; ++j)
        deposit_line(t_re, t_im, diag, tp_row, tq_row, d, i, j, li, lam_g[j], m, delta, hf, b0, b_end);
    }
  }
  __syncthreads();

  for (int k = threadIdx.x; k < kTileBins; k += kBinThreads) {
    const long long idx = b0 + k;
    if (idx < gstride) {
      if (accumulate) {
        const c128 old = g[idx];
        g[idx] = c128{old.re + t_re[k], old.im + t_im[k]};
      } else {
        g[idx] = c128{t_re[k], t_im[k]};
      }
    }
  }
}

// k_spectral_yields: the Parseval sums of the convolved spectra.
// For orientation b and lifetime t:
//   S[b][t] = sum_f Re( cd_b(f) LD_t(f) + co_b(f) LO_t(f) ),
//   cd_b = sum_{k<3} FA_b,k FB_b,k   (pairs p = q),   co_b = sum_{k>=3} FA_b,k FB_b,k   (p < q),
// where FA, FB are the FFTs of the two radicals' six binned spectra and LD_t, LO_t the inverse
// FFTs of K_t and 2 Re K_t on the output grid. A workgroup owns a frequency range for a group
// of kYieldGroup orientations: LD/LO are read once per group (registers), FA/FB once in total.
// Each workgroup writes its partial sums to partial[blockIdx.x][b][t] (no atomics: the result
// is deterministic); the host adds the partials.
constexpr int kYieldGroup = 8;
constexpr int kYieldMaxLifetimes = 8;
constexpr int kYieldThreads = 256;

__global__ void __launch_bounds__(kYieldThreads) k_spectral_yields(const c128* __restrict__ fa,
                                                                   const c128* __restrict__ fb,
                                                                   const c128* __restrict__ ld,
                                                                   const c128* __restrict__ lo, int nb,
                                                                   long long n, int nl,
                                                                   double* __restrict__ partial) {
  __shared__ double red[kYieldThreads / kWave][kYieldGroup * kYieldMaxLifetimes];
  const int b0 = blockIdx.y * kYieldGroup;
  double acc[kYieldGroup][kYieldMaxLifetimes];
#pragma unroll
  for (int g = 0; g < kYieldGroup; ++g)
#pragma unroll
    for (int t = 0; t < kYieldMaxLifetimes; ++t) acc[g][t] = 0.0;

  const long long step = static_cast<long long>(gridDim.x) * kYieldThreads;
  for (long long f = static_cast<long long>(blockIdx.x) * kYieldThreads + threadIdx.x; f < n; f += step) {
    c128 kd[kYieldMaxLifetimes], ko[kYieldMaxLifetimes];
#pragma unroll
    for (int t = 0; t < kYieldMaxLifetimes; ++t) {
      if (t < nl) {
        kd[t] = ld[static_cast<long long>(t) * n + f];
        ko[t] = lo[static_cast<long long>(t) * n + f];
      }
    }
#pragma unroll
    for (int g = 0; g < kYieldGroup; ++g) {
      const int b = b0 + g;
      if (b < nb) {
        const c128* pa = fa + static_cast<long long>(b) * kSpectralPairs * n + f;
        const c128* pb = fb + static_cast<long long>(b) * kSpectralPairs * n + f;
        c128 cd = cmake(0.0, 0.0), co = cmake(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 3; ++k) cd = cfma(pa[k * n], pb[k * n], cd);
#pragma unroll
        for (int k = 3; k < kSpectralPairs; ++k) co = cfma(pa[k * n], pb[k * n], co);
#pragma unroll
        for (int t = 0; t < kYieldMaxLifetimes; ++t)
          if (t < nl) acc[g][t] += (cd.re * kd[t].re - cd.im * kd[t].im) + (co.re * ko[t].re - co.im * ko[t].im);
      }
    }
  }

  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int g = 0; g < kYieldGroup; ++g)
#pragma unroll
    for (int t = 0; t < kYieldMaxLifetimes; ++t) {
      const double v = wave_sum(acc[g][t]);
      if (lane == 0) red[wave][g * kYieldMaxLifetimes + t] = v;
    }
  __syncthreads();
  if (threadIdx.x < kYieldGroup * kYieldMaxLifetimes) {
    const int g = threadIdx.x / kYieldMaxLifetimes, t = threadIdx.x % kYieldMaxLifetimes;
    const int b = b0 + g;
    if (b < nb && t < nl) {
      double v = 0.0;
      for (int w = 0; w < kYieldThreads / kWave; ++w) v += red[w][threadIdx.x];
      partial[(static_cast<long long>(blockIdx.x) * nb + b) * nl + t] = v;
    }
  }
}

}  // namespace rpm

using namespace rpm;

extern "C" int rpm_bin_lines(const double* d_eigenvalues, const void* d_ops, int d, int norient, double weight_divisor,
                             double delta, long long half, void* d_grid, long long grid_stride, int accumulate,
                             int* d_overflow, void* stream) {
  if (!d_eigenvalues || !d_ops || !d_grid || !d_overflow || d < 1 || d > 46340 || norient < 1 || norient > 65535 ||
      !(weight_divisor > 0.0) || !(delta > 0.0) || half < 1 || grid_stride < 2 * half + 1 ||
      (accumulate != 0 && accumulate != 1))
    return fail(RPM_E_INVALID, "rpm_bin_lines: bad arguments");
  const long long tiles = (grid_stride + kTileBins - 1) / kTileBins;
  if (tiles > 0x7fffffffLL) return fail(RPM_E_INVALID, "rpm_bin_lines: grid too long");
  const dim3 grid(static_cast<unsigned>(tiles), norient, kSpectralPairs);
  if (d <= kFlatRows)
    hipLaunchKernelGGL(k_bin_tiles<true>, grid, dim3(kBinThreads), 0, static_cast<hipStream_t>(stream), d_eigenvalues,
                       static_cast<const c128*>(d_ops), d, weight_divisor, delta, half,
                       static_cast<c128*>(d_grid), grid_stride, d_overflow, accumulate);
  else
    hipLaunchKernelGGL(k_bin_tiles<false>, grid, dim3(kBinThreads), 0, static_cast<hipStream_t>(stream),
                       d_eigenvalues, static_cast<const c128*>(d_ops), d, weight_divisor, delta,
                       half, static_cast<c128*>(d_grid), grid_stride, d_overflow, accumulate);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_spectral_yields(const void* d_fa, const void* d_fb, const void* d_ld, const void* d_lo, int norient,
                                   long long n, int nlifetimes, int nblocks, double* d_partial, void* stream) {
  if (!d_fa || !d_fb || !d_ld || !d_lo || !d_partial || norient < 1 || n < 1 || nlifetimes < 1 ||
      nlifetimes > kYieldMaxLifetimes || nblocks < 1 || (norient + kYieldGroup - 1) / kYieldGroup > 65535)
    return fail(RPM_E_INVALID, "rpm_spectral_yields: bad arguments");
  hipLaunchKernelGGL(k_spectral_yields, dim3(nblocks, (norient + kYieldGroup - 1) / kYieldGroup), dim3(kYieldThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const c128*>(d_fa), static_cast<const c128*>(d_fb),
                     static_cast<const c128*>(d_ld), static_cast<const c128*>(d_lo), norient, n, nlifetimes, d_partial);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

// ----------------------------------------------------------------------------
// Batched Hermitian eigen-decomposition (rocSOLVER strided-batched).
//
// The batch is handed to rocSOLVER as one call, so the matrices are reduced
// concurrently rather than one after another. rocSOLVER is column-major: a
// row-major Hermitian H read column-major is H^T = conj(H), whose eigenvalues
// are those of H and whose eigenvectors are conj(V). On return, each matrix
// buffer read row-major is therefore V^H (row k = conjugated eigenvector k).

namespace {

constexpr int kMaxDevices = 64;

// One rocBLAS handle per (device, stream), bound to its stream when created: a handle's internal
// device workspace is then only used by kernels of that one stream, which run in order. Its mutex is
// held for each whole solve, since a handle must not be used by two host threads at once. Handles
// live until the process ends.
struct SolverSlot {
  int dev;
  hipStream_t stream;
  rocblas_handle handle;
  std::mutex lock;
};
std::mutex g_solver_mutex;
std::vector<std::unique_ptr<SolverSlot>> g_solvers;

SolverSlot* solver_slot(int dev, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_solver_mutex);
  for (auto& s : g_solvers)
    if (s->dev == dev && s->stream == stream) return s.get();
  auto slot = std::make_unique<SolverSlot>();
  slot->dev = dev;
  slot->stream = stream;
  if (rocblas_create_handle(&slot->handle) != rocblas_status_success) return nullptr;
  if (rocblas_set_stream(slot->handle, stream) != rocblas_status_success) {
    rocblas_destroy_handle(slot->handle);
    return nullptr;
  }
  g_solvers.push_back(std::move(slot));
  return g_solvers.back().get();
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" size_t rpm_eigh_workspace_bytes(int n, int batch) {
  if (n < 1 || batch < 1) return 0;
  return align256(sizeof(int) * static_cast<size_t>(batch)) + align256(sizeof(double) * static_cast<size_t>(n) * batch);
}

extern "C" int rpm_eigh_batched(void* d_a, int n, int batch, double* d_w, void* d_ws, size_t ws_bytes, int method,
                                void* stream) {
  if (!d_a || !d_w || !d_ws || n < 1 || batch < 1 || (method != 0 && method != 1))
    return fail(RPM_E_INVALID, "rpm_eigh_batched: bad arguments");
  if (ws_bytes < rpm_eigh_workspace_bytes(n, batch)) return fail(RPM_E_INVALID, "rpm_eigh_batched: workspace too small");
  int dev = 0;
  RPM_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(RPM_E_UNSUPPORTED, "rpm_eigh_batched: device index out of range");
  SolverSlot* slot = solver_slot(dev, static_cast<hipStream_t>(stream));
  if (!slot) return fail(RPM_E_HIP, "rpm_eigh_batched: rocBLAS handle for this stream could not be created");
  std::lock_guard<std::mutex> hold(slot->lock);
  rocblas_handle h = slot->handle;
  auto* info = static_cast<rocblas_int*>(d_ws);
  auto* e = reinterpret_cast<double*>(static_cast<char*>(d_ws) + align256(sizeof(int) * static_cast<size_t>(batch)));
  auto* a = static_cast<rocblas_double_complex*>(d_a);
  const rocblas_stride sa = static_cast<rocblas_stride>(n) * n;
  rocblas_status st;
  if (method == 0)
    st = rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, n, a, n, sa, d_w, n, e, n,
                                          info, batch);
  else
    st = rocsolver_zheevdj_strided_batched(h, rocblas_evect_original, rocblas_fill_lower, n, a, n, sa, d_w, n, info,
                                           batch);
  if (st != rocblas_status_success) return fail(RPM_E_HIP, "rpm_eigh_batched: rocSOLVER call failed");
  return RPM_OK;
}
