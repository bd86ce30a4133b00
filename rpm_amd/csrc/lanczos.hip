// Lanczos basis with full reorthogonalisation, norms, singlet projection of a
// Krylov basis and the basis-combination update, for one trajectory or a batch
// of trajectories (grid.y = trajectory).
//
// Orthogonalisation of w = H v_j (two projections, as classical Gram-Schmidt
// applied twice): the three-term step w - beta_{j-1} v_{j-1} - alpha_j v_j
// (formed inside k_lcgs_pass) and then one full projection
// w <- w - V (V^H w), applied one step late so that the projection of step j
// and the materialisation of v_{j+1} share one read of V (k_lcgs_pass).  The
// reference subtracts the same three terms and then projects twice more; the
// bases agree to round-off.
//
// Follows R/rpm_statevector_engine.py:651-688 (basis), 492-497 (singlet
// population), 521-529 (projected observable), 904 (endpoint update).
//
// All reductions are two-stage with a fixed summation order, so results are
// bitwise reproducible run to run.  Each trajectory of a batch carries a
// BasisCtl block (active flag, basis size, target size, 1/norm): every basis
// kernel reads it and returns at once when the trajectory is idle at the
// current Lanczos index (breakdown, or a smaller target), so the host never
// waits inside the basis loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kThreads = 256;
constexpr int kEPT = 4;                        // elements per thread in the norm reductions
constexpr int kChunk = kThreads * kEPT;        // elements per block
constexpr int kMaxBasis = 64;                  // largest Krylov dimension on the device

// ---------------------------------------------------------------- plain reductions

// partial[b * nblk + blk] = sum |x|^2 over the block's chunk of vector b
__global__ void __launch_bounds__(kThreads) k_sumsq(const c128* __restrict__ x, long long n, long long ld,
                                                    double* __restrict__ partial, const int* __restrict__ targets) {
  const long long b = blockIdx.y;
  if (targets && targets[b] <= 0) return;   // vector not part of this basis call
  __shared__ double sm[kThreads / kWave];
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

// Trajectory b takes part in the step of Lanczos index j (inside its window of this call) that
// still needs `need` more vectors after j: j + need < target.
__device__ inline bool basis_live(const BasisCtl* ctl, int b, int j, int need) {
  const BasisCtl& c = ctl[b];
  return basis_step_live(c, j) && j + need < c.target;
}

// Norm of each starting state and the control blocks of this call. Trajectory b with targets[b] > 0
// (every trajectory without targets) starts a new basis of eventual size min(targets[b], uniform) and
// builds its first stops[b] vectors now (all of them without stops); with targets[b] == 0 and
// stops[b] > 0 it resumes the basis an earlier call left (same workspace) up to stops[b] vectors;
// otherwise it sits this call out and keeps everything. Resets the k_lcgs_finish counters of new
// bases.
__global__ void __launch_bounds__(kThreads) k_basis_start(const double* __restrict__ partial, int nblk,
                                                          double* __restrict__ norm_out, BasisCtl* ctl,
                                                          unsigned* __restrict__ counter,
                                                          const int* __restrict__ targets,
                                                          const int* __restrict__ stops, int uniform, long long n) {
  __shared__ double sm[kThreads / kWave];
  const int b = blockIdx.x;
  const int tg = targets ? targets[b] : uniform;
  const int st = stops ? stops[b] : 0;
  BasisCtl& C = ctl[b];
  if (tg <= 0) {
    if (threadIdx.x == 0) {
      if (targets && tg == 0 && st > 0) {
        C.part = 1;
        C.begin = C.stop;
        C.stop = max(C.stop, min(st, C.target));
        C.flags = 0;
      } else {
        C.part = 0;
        C.begin = C.stop;     // no step of this call is live for it
      }
    }
    return;
  }
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kThreads) s += partial[static_cast<long long>(b) * nblk + i];
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    const double nrm = std::sqrt(s);
    int count = min(tg, uniform);
    if (count > n) count = static_cast<int>(n);
    const int first = st > 0 ? min(st, count) : count;
    counter[b] = 0;
    C.nfix = 0;
    norm_out[b] = nrm;
    const bool live = nrm > 0.0;
    C.active = live ? 1 : 0;
    C.size = live ? count : 0;
    C.target = live ? count : 0;
    C.scale = live ? 1.0 / nrm : 0.0;
    C.beta_prev = 0.0;
    C.begin = 0;
    C.stop = live ? first : 0;
    C.part = 1;
    C.npend = 1;
    C.zscale = 0.0;
    C.flags = 0;
  }
}

// h[c] = sum_blk partial[c][blk]; optionally alpha[j] = Re h[alpha_col]. With dpart, the last block
// of each trajectory instead forms dsum[b] = sum_blk dpart[b][blk] (the |y|^2 of the lagged pass),
// in the order k_lcgs_finish used to.
__global__ void __launch_bounds__(kThreads) k_reduce_cols(const c128* __restrict__ partial, int nblk, long long pstride,
                                                          c128* __restrict__ h, int hstride, double* __restrict__ alpha,
                                                          int astride, int alpha_col, int j, int need,
                                                          const BasisCtl* ctl, const double* __restrict__ dpart,
                                                          double* __restrict__ dsum) {
  const int b = blockIdx.y;
  if (!basis_live(ctl, b, j, need)) return;
  __shared__ c128 sm[kThreads / kWave];
  if (dpart && blockIdx.x == gridDim.x - 1) {
    __shared__ double smd[kThreads / kWave];
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += kThreads) s += dpart[static_cast<long long>(b) * nblk + i];
    s = block_sum(s, smd);
    if (threadIdx.x == 0) dsum[b] = s;
    return;
  }
  const int c = blockIdx.x;
  const c128* p = partial + b * pstride + static_cast<long long>(c) * nblk;
  c128 s{0.0, 0.0};
  for (int i = threadIdx.x; i < nblk; i += kThreads) s = cadd(s, p[i]);
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    h[b * hstride + c] = s;
    if (alpha && c == alpha_col) alpha[b * astride + j] = s.re;
  }
}

// Re sum_u dots[b][u] (the matvec's per-unit partials of <v_j, H v_j>), fixed order; every thread
// of the (kThreads-wide) block receives it.
__device__ inline double alpha_sum(const c128* __restrict__ dots, int units, int b, c128* sm) {
  c128 s{0.0, 0.0};
  for (int i = threadIdx.x; i < units; i += kThreads) s = cadd(s, dots[static_cast<long long>(b) * units + i]);
  return block_sum(s, sm).re;
}

// atilde[b] = Re sum_u dots[b][u] (alpha~ of the lagged pass). A single trajectory's basis forms it
// inside k_lcgs_pass instead.
__global__ void __launch_bounds__(kThreads) k_alpha(const c128* __restrict__ dots, int units,
                                                    double* __restrict__ atilde, int j, const BasisCtl* ctl) {
  const int b = blockIdx.x;
  if (!basis_live(ctl, b, j, 0)) return;
  __shared__ c128 sm[kThreads / kWave];
  const double a = alpha_sum(dots, units, b, sm);
  if (threadIdx.x == 0) atilde[b] = a;
}

// ---------------------------------------------------------------- lagged one-pass orthogonalisation
//
// Step j works on the materialised basis v_0 .. v_{j-1} and the vector x_j that the matvec
// consumed, which is v_j plus a pending O(eps) combination of the basis, v_j = x_j - V_<j c_j
// (c_j from step j-1). The matvec writes H x_j; one pass over V_<j
//   * forms u = H x_j - beta_{j-1} v_{j-1} as it stages H x_j,
//   * materialises v_j = x_j - V_<j c_j in V[j] (x_j = s in, from the matvec's input),
//   * forms D = V_<j^H u, gamma = v_j^H u and y = u - alpha~ v_j (alpha~ = x_j^H H x_j from the
//     matvec's partials), with |y|^2.
// Then (k_lcgs_finish), with H V_<j = V_<=j T + O(eps):
//   alpha_j = Re gamma - beta_{j-1} Re c_j[j-1],
//   w''_j = (1 - V_<=j V_<=j^H) u = y - V_<=j g, g = [D; gamma - alpha~],
//   beta_j^2 = |y|^2 - |g|^2, x_{j+1} = y / beta_j, c_{j+1} = g / beta_j.
// w''_j is the three-term step followed by one full projection, as in two-pass classical
// Gram-Schmidt; the projection is only applied one step later, inside the next pass. The matvec
// acting on x_{j+1} instead of v_{j+1} adds H V c, which lies in the span of the basis and is
// removed by the next projection; |c| grows by about |H| / beta per step, so when |c| passes
// kFixThreshold, or when |g| is not small against |y| (near breakdown), k_lcgs_finish applies the
// correction to y explicitly and takes beta_j from the corrected vector. One read of V per step
// instead of two.
constexpr int kPassPer = 4;                // 64-element runs per k_lcgs_pass block (2 and 8 measured slower)
constexpr int kPassElems = kPassPer * kWave;
constexpr double kFixThreshold = 1e-6;     // |c_{j+1}| above which the correction is applied at once
constexpr double kCancel = 1e-2;           // beta^2 <= kCancel |y|^2: the norm is taken explicitly
constexpr int kFixBlocks = 256;            // blocks per trajectory of k_lcgs_finish
constexpr int kPassWaves = 4;
static_assert(kPassWaves * kWave == kThreads, "k_lcgs_pass shares the kThreads-wide reductions");

// One block per kPassElems elements of one trajectory's vectors; columns c < j are dealt to the waves,
// column c to wave c % 4, and each wave streams its columns one at a time with the next column's
// loads in flight (a round structure that waited for every load of a round before reducing it was
// latency-bound: 977 us per launch whatever j). kStream: the basis columns are read with
// non-temporal loads. That pays when the bases of a batch are far larger than the 256 MB Infinity
// Cache (+5% at 32 trajectories) and costs when one trajectory's basis fits in it (-10%).
// The four loads of one column for this lane: elements it * kWave + lane of the block (kFull: all in
// range, so every load is a constant offset from the column start; else clamped to the last element,
// which then weighs nothing).
template <bool kStream, bool kFull>
__device__ inline void pass_load(const c128* __restrict__ col, int lane, int cnt, c128 (&v)[kPassPer]) {
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) {
    const int e = kFull ? it * kWave + lane : min(it * kWave + lane, cnt - 1);
    if constexpr (kStream) {
      v[it].re = __builtin_nontemporal_load(&col[e].re);
      v[it].im = __builtin_nontemporal_load(&col[e].im);
    } else {
      v[it] = col[e];
    }
  }
}

// Body of k_lcgs_pass for one block (kThreads == kPassElems: thread t owns element x0 + t).
template <bool kStream, bool kFull>
__device__ inline void pass_body(c128* __restrict__ Vb, long long n, int j, long long x0, int cnt,
                                 const c128* __restrict__ xin, c128* __restrict__ u, const c128* __restrict__ cpend_b,
                                 double at, double bprev, double sc, c128* __restrict__ part_b,
                                 double* __restrict__ ypart_b, c128* s_u, c128 (*s_corr)[kPassElems], c128* s_c,
                                 c128* sm, double* smd, int blk, int nblk) {
  const int t = threadIdx.x, lane = t % kWave, q = t / kWave;
  const bool mine = t < cnt;
  const int et = mine ? t : cnt - 1;
  // This wave's columns q, q + 4, q + 8, ... < j are streamed one at a time with the next columns'
  // loads in flight. Each column's partial sums the wave's elements in the order it = 0, 1, ... and
  // the correction of each element adds the wave's columns in ascending order, then the waves in a
  // fixed order, so the basis is bitwise the same for every batch size.
  const int ncol = j > q ? (j - q + kPassWaves - 1) / kPassWaves : 0;
  c128 va[kPassPer], vb[kPassPer];
  // the wave's first column is requested before the block stages u (which it does not need)
  if (ncol > 0) pass_load<kStream, kFull>(Vb + static_cast<long long>(q) * n + x0, lane, cnt, va);
  {
    c128 ue = u[x0 + et];                       // u = H x_j - beta_{j-1} v_{j-1} (three-term step)
    if (j > 0) {
      const c128 p = Vb[static_cast<long long>(j - 1) * n + x0 + et];
      ue.re = fma(-bprev, p.re, ue.re);
      ue.im = fma(-bprev, p.im, ue.im);
    }
    s_u[t] = mine ? ue : cmake(0.0, 0.0);       // elements past cnt weigh nothing
  }
  if (t < j) s_c[t] = cpend_b[t];
  __syncthreads();
  c128 corr[kPassPer];
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) corr[it] = cmake(0.0, 0.0);
  auto load = [&](int k, c128(&v)[kPassPer]) {
    pass_load<kStream, kFull>(Vb + static_cast<long long>(q + kPassWaves * k) * n + x0, lane, cnt, v);
  };
  auto use = [&](int k, const c128(&v)[kPassPer]) {
    const int c = q + kPassWaves * k;
    const c128 cc = s_c[c];
    c128 acc = cmake(0.0, 0.0);
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) {
      acc = cfmaconj(v[it], s_u[it * kWave + lane], acc);
      corr[it] = cfma(v[it], cc, corr[it]);
    }
    acc = wave_sum(acc);
    if (lane == 0) part_b[static_cast<long long>(c) * nblk + blk] = acc;
  };
  // the loop body has no branch, so the next column's loads stay outstanding until their use (a load
  // under a branch would make the compiler wait for everything at the join); measured: a third
  // column in flight, or fewer registers for more waves, does not change the time per column
  int k = 0;
  for (; k + 2 < ncol; k += 2) {
    load(k + 1, vb);
    use(k, va);
    load(k + 2, va);
    use(k + 1, vb);
  }
  if (k + 1 < ncol) {
    load(k + 1, vb);
    use(k, va);
    use(k + 1, vb);
  } else if (k < ncol) {
    use(k, va);
  }
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) s_corr[q][it * kWave + lane] = corr[it];
  __syncthreads();
  // v_j = s x_j - V_<j c_j, y = u - alpha~ v_j, gamma = v_j^H u, |y|^2
  c128 g = cmake(0.0, 0.0);
  double yy = 0.0;
  if (mine) {
    const c128 ue = s_u[t];
    const c128 cs = cadd(cadd(cadd(s_corr[0][t], s_corr[1][t]), s_corr[2][t]), s_corr[3][t]);
    const c128 v = csub(cscale(xin[x0 + t], sc), cs);
    Vb[static_cast<long long>(j) * n + x0 + t] = v;
    const c128 y = cmake(fma(-at, v.re, ue.re), fma(-at, v.im, ue.im));
    u[x0 + t] = y;
    g = cfmaconj(v, ue, g);
    yy = fma(y.re, y.re, fma(y.im, y.im, yy));
  }
  g = block_sum(g, sm);
  yy = block_sum(yy, smd);
  if (t == 0) {
    part_b[static_cast<long long>(j) * nblk + blk] = g;
    ypart_b[blk] = yy;
  }
}

// Grid (blocks, batch): block blockIdx.x is element block blk0 + blockIdx.x of nblk. kFull: every
// element of those blocks is in range (all but a vector's last block when kPassElems does not divide
// N, which a second launch with kFull = false covers).
template <bool kStream, bool kFull>
__global__ void __launch_bounds__(kThreads) k_lcgs_pass(c128* __restrict__ V, long long vstride, long long n, int j,
                                                        const c128* __restrict__ in, long long ild,
                                                        const c128* __restrict__ cpend, int cstride,
                                                        double* __restrict__ atilde,
                                                        const c128* __restrict__ adots, int units, c128* __restrict__ U,
                                                        c128* __restrict__ partial, long long pstride,
                                                        double* __restrict__ ypart, const BasisCtl* ctl, int blk0,
                                                        int nblk) {
  static_assert(kPassElems == kThreads, "one element per thread outside the column stream");
  const int b = blockIdx.y;
  if (!basis_live(ctl, b, j, 0)) return;
  __shared__ c128 s_u[kPassElems];
  __shared__ c128 s_corr[kPassWaves][kPassElems];
  __shared__ c128 s_c[kMaxBasis];
  __shared__ c128 sm[kThreads / kWave];
  __shared__ double smd[kThreads / kWave];
  double at;
  if (adots) {
    // alpha~ from the matvec's partials, formed by every block in the same order as k_alpha;
    // block 0 stores it for k_lcgs_finish
    at = alpha_sum(adots, units, b, sm);
    if (blk0 + blockIdx.x == 0 && threadIdx.x == 0) atilde[b] = at;
  } else {
    at = atilde[b];
  }
  const int blk = blk0 + blockIdx.x;
  const long long x0 = static_cast<long long>(blk) * kPassElems;
  const int cnt = kFull ? kPassElems : static_cast<int>(n - x0 < kPassElems ? n - x0 : kPassElems);
  const BasisCtl& C = ctl[b];
  pass_body<kStream, kFull>(V + b * vstride, n, j, x0, cnt, in + b * ild, U + b * n, cpend + b * cstride, at,
                            C.beta_prev, C.scale, partial + b * pstride, ypart + static_cast<long long>(b) * nblk,
                            s_u, s_corr, s_c, sm, smd, blk, nblk);
}

// alpha_j, beta_j and c_{j+1} from the pass's sums h = [D; gamma] and |y|^2 (see above), and the
// explicit correction when it is due. Grid (kFixBlocks, batch): every block forms the same scalars
// in the same order; block 0 stores them. When the correction is due (|c_{j+1}| > kFixThreshold,
// or |g| not small against |y|), every block subtracts its share of V_<=j g from y, and the last
// block of the trajectory to finish (device-scope counter; the fences run on this rare path only)
// takes beta_j from the corrected vector and clears c_{j+1}.
// Every block reads the control block before the lead block updates it, or decides only what the
// lead block also decides from h (whether the correction is due): nothing a block branches on
// depends on the order in which the blocks run.
__device__ inline void finish_step(const c128* __restrict__ Vb, long long n, double yy, const c128* __restrict__ hb,
                                   c128* __restrict__ cb, double at,
                                   double* __restrict__ alpha_b, double* __restrict__ beta_b, int j, double tol,
                                   c128* __restrict__ y, double* __restrict__ fpart_b, unsigned* __restrict__ counter_b,
                                   BasisCtl& C, double* sm, c128* gs, int* is_last) {
  const int t = threadIdx.x;
  const double bprev = C.beta_prev;
  double gg = 0.0;
  if (t < j) {
    const c128 v = hb[t];
    gg = v.re * v.re + v.im * v.im;
    gs[t] = v;
  }
  const c128 gam = hb[j];
  const double cprev = j > 0 ? cb[j - 1].re : 0.0;
  const int target = C.target;
  gg = block_sum(gg, sm);   // ends with a barrier: every read above precedes the writes below
  const c128 glast = cmake(gam.re - at, gam.im);
  gg += glast.re * glast.re + glast.im * glast.im;
  const bool lead = blockIdx.x == 0;
  if (lead && t == 0) alpha_b[j] = gam.re - bprev * cprev;
  if (j + 1 >= target) return;                 // the trajectory's last vector: no beta_j
  const double b2 = yy - gg;
  bool fix = !(b2 > kCancel * yy);
  if (!fix) fix = gg > kFixThreshold * kFixThreshold * b2;
  if (!fix) {
    if (!lead) return;
    const double bj = std::sqrt(b2);
    if (t == 0) beta_b[j] = bj;
    if (bj <= tol) {
      if (t == 0) {
        C.active = 0;
        C.size = j + 1;
      }
      return;
    }
    const double inv = 1.0 / bj;
    if (t < j) cb[t] = cscale(gs[t], inv);
    if (t == 0) {
      cb[j] = cscale(glast, inv);
      C.scale = inv;
      C.beta_prev = bj;
    }
    return;
  }
  // explicit correction: y -= V_<=j g
  if (t == 0) gs[j] = glast;
  __syncthreads();
  const int k = j + 1;
  double s = 0.0;
  for (long long i = static_cast<long long>(blockIdx.x) * kThreads + t; i < n;
       i += static_cast<long long>(gridDim.x) * kThreads) {
    c128 yv = y[i];
    for (int c = 0; c < k; ++c) yv = cfms(Vb[static_cast<long long>(c) * n + i], gs[c], yv);
    y[i] = yv;
    s = fma(yv.re, yv.re, fma(yv.im, yv.im, s));
  }
  s = block_sum(s, sm);
  if (t == 0) {
    fpart_b[blockIdx.x] = s;
    __threadfence();
    *is_last = atomicAdd(counter_b, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!*is_last) return;
  __threadfence();
  const volatile double* fp = fpart_b;
  double tot = 0.0;
  for (int i = t; i < static_cast<int>(gridDim.x); i += kThreads) tot += fp[i];
  tot = block_sum(tot, sm);
  if (t < k) cb[t] = cmake(0.0, 0.0);
  if (t == 0) {
    *counter_b = 0;
    C.nfix += 1;
    const double bj = std::sqrt(tot);
    beta_b[j] = bj;
    if (bj <= tol) {
      C.active = 0;
      C.size = j + 1;
    } else {
      C.scale = 1.0 / bj;
      C.beta_prev = bj;
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_lcgs_finish(const c128* __restrict__ V, long long vstride, long long n,
                                                          const double* __restrict__ ysum,
                                                          const c128* __restrict__ h, int hstride,
                                                          c128* __restrict__ cpend, int cstride,
                                                          const double* __restrict__ atilde, double* __restrict__ alpha,
                                                          double* __restrict__ beta, int astride, int j, double tol,
                                                          c128* __restrict__ U, double* __restrict__ fpart,
                                                          unsigned* __restrict__ counter, BasisCtl* ctl) {
  const int b = blockIdx.y;
  if (!basis_live(ctl, b, j, 0)) return;
  __shared__ double sm[kThreads / kWave];
  __shared__ c128 gs[kMaxBasis];
  __shared__ int is_last;
  finish_step(V + b * vstride, n, ysum[b], h + b * hstride, cpend + b * cstride, atilde[b],
              alpha + b * astride, beta + b * astride, j, tol, U + b * n, fpart + static_cast<long long>(b) * gridDim.x,
              counter + b, ctl[b], sm, gs, &is_last);
}

// ---------------------------------------------------------------- paired steps
//
// Steps j and j+1 (j even) share one read of V. With x_j pending (v_j = s x_j - V_<j c_B) and,
// after an earlier pair, v_{j-1} pending as well (v_{j-1} = a z' - V_<j-1 c_A for the previous
// pair's z'), write x^_j = s x_j and x^_{j-1} = a z' (or the materialised v_{j-1}):
//   M1     P = H x^_j and the partials of alpha~_j = <x^_j, P>;
//   prep   z = P - beta_{j-1} x^_{j-1} - alpha~_j x^_j: the three-term step on unprojected vectors,
//          so that z = y_j + (a vector of span V_<j with known coefficients);
//   M2     Q = H z and the partials of <z, Q>;
//   pass   one read of the materialised columns: v_{j-1} and v_j materialised, V^H r (r = z +
//          alpha~_j x^_j = u_j less beta_{j-1} V c_A) and V^H Q, and in-block sums of the columns
//          materialised here and of <v_j, u_j>, |y_j|^2, <v_j, Q>, <z, v_j>, |v_j|^2, |Q|^2, |z|^2;
//   finish step j as a single step (alpha_j, beta_j, g_j = [D; gamma_j - alpha~_j]); then
//          v_{j+1} = z / beta_j - V_<=j c_A' with c_A' = (g_j - alpha~_j c_B) / beta_j (D here
//          without the beta_{j-1} c_A part), and step j+1 by linearity: H x^_{j+1} = Q / beta_j for
//          x^_{j+1} = z / beta_j, V^H u_{j+1} = V^H Q / beta_j - beta_j e_j, gamma_{j+1} =
//          <x^_{j+1}, u_{j+1}> - c_A'^H V^H u_{j+1}, alpha~_{j+1} = <z, Q> / beta_j^2 and |y'|^2 of
//          y' = u_{j+1} - alpha~_{j+1} x^_{j+1} from the in-block sums; then alpha_{j+1}, beta_{j+1}
//          and c_B' as in a single step;
//   close  x_{j+2} = y' = Q / beta_j - beta_j v_j - alpha~_{j+1} / beta_j z, pending with c_B';
//          v_{j+1} stays pending (z, c_A') until the next pass materialises it.
// The explicit correction (|c_A'| or |c_B'| above kPairFixThreshold, cancellation in beta_{j+1}^2, and
// the last pair of a call, which leaves v_{j+1} materialised and one vector pending) materialises
// v_{j+1}, subtracts V_<=j+1 g_{j+1} from x_{j+2} and takes beta_{j+1} from its norm. Cancellation
// in beta_j^2 (a near-breakdown at step j, which a single step corrects before the next matvec)
// stops the call's basis at j + 1 with the corrected vector pending (kBasisShort): the caller
// continues it with another call.
// In the same launches, a trajectory whose window holds only step j (S0) takes a single step on P;
// one whose window holds only step j+1 (S1) takes a single step on M2 applied to its own x_{j+1}.
// Pending vectors: x_p in w[p % 2] (as in single steps); P in w[1]; z alternates between two
// buffers; Q has its own.
// Lanczos index of the first pair. Early steps read few basis columns, so the pair's own traffic
// (prep, close and the extra staged vectors: about eight vectors per trajectory) outweighs the column
// read it saves. Measured on 32 production trajectories over 0.5 us: first pair at 2, 4, 8, 12, 16,
// 20 gives 37.6, 37.9, 38.0, 38.1, 37.6, 36.4 trajectory-us/s.
constexpr int kPairFrom = 12;
// |c_A'|, |c_B'| above which a pair closes with the explicit correction. A pair's pending
// coefficients carry the earlier ones times alpha~ and beta (z = y_j - beta_{j-1} V c_A - alpha~_j
// V c_B), about |H| / beta larger than a single step's for the same lag, so the bound is that
// much looser than kFixThreshold for the same size of lagged component.
constexpr double kPairFixThreshold = 1e-4;
constexpr int kColE = kMaxBasis;             // partial / sum columns of V^H Q
constexpr int kColS = 2 * kMaxBasis;         // in-block scalar sums: (|y|^2, |v|^2), (|Q|^2, |z|^2), <z, v>
constexpr int kPairCols = 2 * kMaxBasis + 8; // columns of the pass partials and sums
constexpr int kCloseGamma = 1, kCloseFix = 2, kCloseTarget = 4, kCloseRetreat = 8, kCloseMask = 15;
static_assert(kBasisShort > kCloseMask, "flag bits overlap");
enum { kModeIdle = 0, kModePair = 1, kModeS0 = 2, kModeS1 = 3 };

__device__ inline int pair_mode(const BasisCtl& C, int j) {
  const bool l0 = basis_step_live(C, j), l1 = basis_step_live(C, j + 1);
  return l0 ? (l1 ? kModePair : kModeS0) : (l1 ? kModeS1 : kModeIdle);
}

// Input of M2: z / s for a pair (M2 scales its input by s, so it forms H z), x_{j+1} for S1.
// kFuse (a single trajectory): every block forms alpha~_j from M1's per-unit partials dots in the
// order of k_alpha, and block 0 stores it, in place of a k_alpha launch. The batched instantiation
// carries none of that code.
template <bool kFuse>
__global__ void __launch_bounds__(kThreads) k_pair_prep(const c128* __restrict__ V, long long vstride, long long n,
                                                        int j, const c128* __restrict__ X0, const c128* __restrict__ X1,
                                                        const c128* __restrict__ ZA, c128* __restrict__ Z,
                                                        double* __restrict__ atilde, const c128* __restrict__ dots,
                                                        int units, const BasisCtl* ctl) {
  const int b = blockIdx.y;
  const BasisCtl& C = ctl[b];
  const int mode = pair_mode(C, j);
  if (mode != kModePair && mode != kModeS1) return;
  double at = 0.0;
  if constexpr (kFuse) {
    if (mode == kModePair) {
      __shared__ c128 sm[kThreads / kWave];
      at = alpha_sum(dots, units, b, sm);
      if (blockIdx.x == 0 && threadIdx.x == 0) atilde[b] = at;
    }
  }
  const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const long long o = b * n + i;
  if (mode == kModeS1) {
    Z[o] = X1[o];
    return;
  }
  if constexpr (!kFuse) at = atilde[b];
  const double s = C.scale, bp = C.beta_prev;
  const c128 xh = cscale(X0[o], s);
  const c128 xp = C.npend == 2 ? cscale(ZA[o], C.zscale) : V[b * vstride + static_cast<long long>(j - 1) * n + i];
  c128 z = X1[o];
  z.re = fma(-bp, xp.re, z.re);
  z.im = fma(-bp, xp.im, z.im);
  z.re = fma(-at, xh.re, z.re);
  z.im = fma(-at, xh.im, z.im);
  Z[o] = cscale(z, 1.0 / s);
}

struct PairPass {
  c128* Vb;                 // this trajectory's basis
  long long n, x0;
  int p, cnt, blk, nblk;
  const c128* xb;           // pending x_p (scale s)
  const c128* za;           // pending z of v_{p-1} (scale a), two pending vectors only
  const c128* r;            // pair: z / s; single: H x^_p
  const c128* q;            // pair: Q
  const c128* cA;           // pending coefficients of v_{p-1} (p - 1 entries) and of v_p (p entries)
  const c128* cB;
  c128* y;                  // single: y = u - alpha~ v_p (the next pending vector)
  c128* part;               // [kPairCols][nblk] partial sums
  double s, a, at, bp;
};

// One block of the paired pass (thread t owns element x0 + t outside the column stream).
template <bool kStream, bool kFull, bool kPair, bool kNp2>
__device__ inline void pair_pass_body(const PairPass& P, c128* s_r, c128* s_q, c128* s_x, c128* s_a, c128* s_z,
                                      c128 (*s_corr)[kPassElems], c128* s_cA, c128* s_cB, double* s_red) {
  const int t = threadIdx.x, lane = t % kWave, wq = t / kWave;
  const long long n = P.n, x0 = P.x0;
  const int p = P.p, cnt = P.cnt;
  const int mc = kNp2 ? p - 1 : p;           // materialised columns, streamed
  const bool mine = t < cnt;
  const int et = mine ? t : cnt - 1;
  const int ncol = mc > wq ? (mc - wq + kPassWaves - 1) / kPassWaves : 0;
  c128 va[kPassPer], vb[kPassPer];
  if (ncol > 0) pass_load<kStream, kFull>(P.Vb + static_cast<long long>(wq) * n + x0, lane, cnt, va);
  {
    // staged: x^_p, a z' (two pending vectors), z (pairs), r and Q; elements past cnt weigh nothing
    const c128 xh = cscale(P.xb[x0 + et], P.s);
    s_x[t] = xh;
    c128 xa = cmake(0.0, 0.0);
    if constexpr (kNp2) {
      xa = cscale(P.za[x0 + et], P.a);
      s_a[t] = xa;
    }
    c128 r1;
    if constexpr (kPair) {
      const c128 z = cscale(P.r[x0 + et], P.s);
      s_z[t] = z;
      r1 = z;
      r1.re = fma(P.at, xh.re, r1.re);
      r1.im = fma(P.at, xh.im, r1.im);
      s_q[t] = mine ? P.q[x0 + et] : cmake(0.0, 0.0);
    } else {
      const c128 xp = kNp2 ? xa : P.Vb[static_cast<long long>(p - 1) * n + x0 + et];
      r1 = P.r[x0 + et];
      r1.re = fma(-P.bp, xp.re, r1.re);
      r1.im = fma(-P.bp, xp.im, r1.im);
    }
    s_r[t] = mine ? r1 : cmake(0.0, 0.0);
  }
  if (t < p) s_cB[t] = P.cB[t];
  if (kNp2 && t < p - 1) s_cA[t] = P.cA[t];
  __syncthreads();
  c128 corrB[kPassPer], corrA[kPassPer];
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) {
    corrB[it] = cmake(0.0, 0.0);
    corrA[it] = cmake(0.0, 0.0);
  }
  auto load = [&](int k, c128(&v)[kPassPer]) {
    pass_load<kStream, kFull>(P.Vb + static_cast<long long>(wq + kPassWaves * k) * n + x0, lane, cnt, v);
  };
  auto use = [&](int k, const c128(&v)[kPassPer]) {
    const int c = wq + kPassWaves * k;
    const c128 cb = s_cB[c];
    const c128 cav = kNp2 ? s_cA[c] : cmake(0.0, 0.0);
    c128 accD = cmake(0.0, 0.0), accE = cmake(0.0, 0.0);
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) {
      accD = cfmaconj(v[it], s_r[it * kWave + lane], accD);
      if constexpr (kPair) accE = cfmaconj(v[it], s_q[it * kWave + lane], accE);
      corrB[it] = cfma(v[it], cb, corrB[it]);
      if constexpr (kNp2) corrA[it] = cfma(v[it], cav, corrA[it]);
    }
    accD = wave_sum(accD);
    if constexpr (kPair) accE = wave_sum(accE);
    if (lane == 0) {
      P.part[static_cast<long long>(c) * P.nblk + P.blk] = accD;
      if constexpr (kPair) P.part[static_cast<long long>(kColE + c) * P.nblk + P.blk] = accE;
    }
  };
  int k = 0;
  for (; k + 2 < ncol; k += 2) {
    load(k + 1, vb);
    use(k, va);
    load(k + 2, va);
    use(k + 1, vb);
  }
  if (k + 1 < ncol) {
    load(k + 1, vb);
    use(k, va);
    use(k + 1, vb);
  } else if (k < ncol) {
    use(k, va);
  }
  // corrections summed over the waves in a fixed order
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) s_corr[wq][it * kWave + lane] = corrB[it];
  __syncthreads();
  c128 sB = cadd(cadd(cadd(s_corr[0][t], s_corr[1][t]), s_corr[2][t]), s_corr[3][t]);
  c128 sA = cmake(0.0, 0.0);
  if constexpr (kNp2) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) s_corr[wq][it * kWave + lane] = corrA[it];
    __syncthreads();
    sA = cadd(cadd(cadd(s_corr[0][t], s_corr[1][t]), s_corr[2][t]), s_corr[3][t]);
  }
  // per element: v_{p-1}, v_p, u_p, y_p and the in-block sums
  constexpr int K = 14;
  double red[K];
#pragma unroll
  for (int i = 0; i < K; ++i) red[i] = 0.0;
  if (mine) {
    const long long e = x0 + t;
    const c128 r1 = s_r[t];
    c128 u = r1;
    c128 vA = cmake(0.0, 0.0);
    if constexpr (kNp2) {
      vA = csub(s_a[t], sA);
      P.Vb[static_cast<long long>(p - 1) * n + e] = vA;
      sB = cfma(vA, s_cB[p - 1], sB);
      u.re = fma(P.bp, sA.re, u.re);
      u.im = fma(P.bp, sA.im, u.im);
    }
    const c128 vB = csub(s_x[t], sB);
    P.Vb[static_cast<long long>(p) * n + e] = vB;
    const c128 y = cmake(fma(-P.at, vB.re, u.re), fma(-P.at, vB.im, u.im));
    const c128 g = cfmaconj(vB, u, cmake(0.0, 0.0));
    red[0] = g.re;
    red[1] = g.im;
    red[2] = fma(y.re, y.re, y.im * y.im);
    if constexpr (kNp2) {
      const c128 d = cfmaconj(vA, u, cmake(0.0, 0.0));
      red[3] = d.re;
      red[4] = d.im;
    }
    if constexpr (kPair) {
      const c128 q = s_q[t];
      const c128 z = s_z[t];
      const c128 eb = cfmaconj(vB, q, cmake(0.0, 0.0));
      const c128 zv = cfmaconj(z, vB, cmake(0.0, 0.0));
      red[5] = eb.re;
      red[6] = eb.im;
      red[7] = zv.re;
      red[8] = zv.im;
      red[9] = fma(vB.re, vB.re, vB.im * vB.im);
      red[10] = fma(q.re, q.re, q.im * q.im);
      red[11] = fma(z.re, z.re, z.im * z.im);
      if constexpr (kNp2) {
        const c128 dq = cfmaconj(vA, q, cmake(0.0, 0.0));
        red[12] = dq.re;
        red[13] = dq.im;
      }
    } else {
      P.y[e] = y;
    }
  }
  // block sums: each wave, then the waves in order
#pragma unroll
  for (int i = 0; i < K; ++i) red[i] = wave_sum(red[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < K; ++i) s_red[wq * K + i] = red[i];
  }
  __syncthreads();
  if (t == 0) {
    double tot[K];
#pragma unroll
    for (int i = 0; i < K; ++i) tot[i] = s_red[i];
    for (int w = 1; w < kPassWaves; ++w) {
#pragma unroll
      for (int i = 0; i < K; ++i) tot[i] += s_red[w * K + i];
    }
    auto put = [&](int col, double re, double im) { P.part[static_cast<long long>(col) * P.nblk + P.blk] = cmake(re, im); };
    put(p, tot[0], tot[1]);
    if constexpr (kNp2) put(p - 1, tot[3], tot[4]);
    if constexpr (kPair) {
      put(kColE + p, tot[5], tot[6]);
      if constexpr (kNp2) put(kColE + p - 1, tot[12], tot[13]);
      put(kColS, tot[2], tot[9]);
      put(kColS + 1, tot[10], tot[11]);
      put(kColS + 2, tot[7], tot[8]);
    } else {
      put(kColS, tot[2], 0.0);
    }
  }
}

template <bool kStream, bool kFull, bool kFuse>
__global__ void __launch_bounds__(kThreads) k_pair_pass(c128* __restrict__ V, long long vstride, long long n, int j,
                                                        c128* __restrict__ X0, c128* __restrict__ X1,
                                                        const c128* __restrict__ ZA, const c128* __restrict__ Z,
                                                        const c128* __restrict__ Q, const c128* __restrict__ cpendA,
                                                        const c128* __restrict__ cpend, int cstride,
                                                        double* __restrict__ atilde,
                                                        double* __restrict__ atilde2, const c128* __restrict__ dots,
                                                        const c128* __restrict__ dots2, int units,
                                                        c128* __restrict__ partial,
                                                        long long pstride, const BasisCtl* ctl, int blk0, int nblk) {
  static_assert(kPassElems == kThreads, "one element per thread outside the column stream");
  const int b = blockIdx.y;
  const BasisCtl& C = ctl[b];
  const int mode = pair_mode(C, j);
  if (mode == kModeIdle) return;
  // kFuse (a single trajectory; k_alpha does not run): the alpha~ this pass or the slot's finish
  // needs comes from the matvec partials dots, dots2, summed in the order of k_alpha: a pair's
  // alpha~_j was stored by k_pair_prep and its alpha~_{j+1} (finish) is stored by block 0 here; a
  // single step S0 or S1 forms its own in every block, and block 0 stores it. The batched
  // instantiations carry none of that code.
  double at_single = 0.0;   // S0 / S1: this block's alpha~ (never re-read from memory here)
  if constexpr (kFuse) {
    __shared__ c128 s_al[kThreads / kWave];
    const bool lead = blk0 + blockIdx.x == 0;
    if (mode == kModePair) {
      if (lead) {
        const double a2 = alpha_sum(dots2, units, b, s_al);
        if (threadIdx.x == 0) atilde2[b] = a2;
      }
    } else if (mode == kModeS0) {
      at_single = alpha_sum(dots, units, b, s_al);
      if (lead && threadIdx.x == 0) atilde[b] = at_single;
    } else {
      at_single = alpha_sum(dots2, units, b, s_al);
      if (lead && threadIdx.x == 0) atilde2[b] = at_single;
    }
  }
  __shared__ c128 s_r[kPassElems];
  __shared__ c128 s_q[kPassElems];
  __shared__ c128 s_x[kPassElems];
  __shared__ c128 s_a[kPassElems];
  __shared__ c128 s_z[kPassElems];
  __shared__ c128 s_corr[kPassWaves][kPassElems];
  __shared__ c128 s_cA[kMaxBasis];
  __shared__ c128 s_cB[kMaxBasis];
  __shared__ double s_red[kPassWaves * 14];
  PairPass P;
  P.Vb = V + b * vstride;
  P.n = n;
  P.blk = blk0 + blockIdx.x;
  P.nblk = nblk;
  P.x0 = static_cast<long long>(P.blk) * kPassElems;
  P.cnt = kFull ? kPassElems : static_cast<int>(n - P.x0 < kPassElems ? n - P.x0 : kPassElems);
  P.s = C.scale;
  P.a = C.zscale;
  P.bp = C.beta_prev;
  P.cA = cpendA + b * cstride;
  P.cB = cpend + b * cstride;
  P.za = ZA + b * n;
  P.part = partial + b * pstride;
  P.q = Q + b * n;
  P.y = nullptr;
  const bool np2 = C.npend == 2;
  if (mode == kModePair) {
    P.p = j;
    P.xb = X0 + b * n;
    P.r = Z + b * n;
    P.at = atilde[b];
    if (np2)
      pair_pass_body<kStream, kFull, true, true>(P, s_r, s_q, s_x, s_a, s_z, s_corr, s_cA, s_cB, s_red);
    else
      pair_pass_body<kStream, kFull, true, false>(P, s_r, s_q, s_x, s_a, s_z, s_corr, s_cA, s_cB, s_red);
  } else if (mode == kModeS0) {
    P.p = j;
    P.xb = X0 + b * n;
    P.r = X1 + b * n;      // P = H x^_j
    P.y = X1 + b * n;      // y_j, the pending x_{j+1}, over P (each element read before it is written)
    P.at = kFuse ? at_single : atilde[b];
    if (np2)
      pair_pass_body<kStream, kFull, false, true>(P, s_r, s_q, s_x, s_a, s_z, s_corr, s_cA, s_cB, s_red);
    else
      pair_pass_body<kStream, kFull, false, false>(P, s_r, s_q, s_x, s_a, s_z, s_corr, s_cA, s_cB, s_red);
  } else {
    P.p = j + 1;           // a window that starts at j + 1 follows a call that left one vector pending
    P.xb = X1 + b * n;
    P.r = Q + b * n;       // M2 on x_{j+1}
    P.y = X0 + b * n;
    P.at = kFuse ? at_single : atilde2[b];
    pair_pass_body<kStream, kFull, false, false>(P, s_r, s_q, s_x, s_a, s_z, s_corr, s_cA, s_cB, s_red);
  }
}

// h[b][col] = sum_blk partial[b][col][blk] over the slot's columns: 0..p, the in-block scalars, and
// (pairs) kColE..kColE + p. Grid (2j + 6, batch). A single step whose pass materialised v_{p-1}
// from a pending z gets D[c] = V_c^H u = (the streamed V_c^H r) + beta_{p-1} cA[c] (c < p - 1); a
// pair's k_pair_finish forms that sum itself.
__global__ void __launch_bounds__(kThreads) k_pair_reduce(const c128* __restrict__ partial, int nblk, long long pstride,
                                                          c128* __restrict__ h, int hstride, int j,
                                                          const c128* __restrict__ cpendA, int cstride,
                                                          const BasisCtl* ctl) {
  const int b = blockIdx.y;
  const BasisCtl& C = ctl[b];
  const int mode = pair_mode(C, j);
  if (mode == kModeIdle) return;
  const int p = mode == kModeS1 ? j + 1 : j;
  const int x = blockIdx.x;
  int col;
  if (x <= p)
    col = x;
  else if (x <= p + 3)
    col = kColS + (x - p - 1);
  else if (mode == kModePair && x - (p + 4) <= p)
    col = kColE + (x - p - 4);
  else
    return;
  __shared__ c128 sm[kThreads / kWave];
  const c128* pp = partial + b * pstride + static_cast<long long>(col) * nblk;
  c128 s{0.0, 0.0};
  for (int i = threadIdx.x; i < nblk; i += kThreads) s = cadd(s, pp[i]);
  s = block_sum(s, sm);
  if (threadIdx.x == 0) {
    if (mode == kModeS0 && C.npend == 2 && col < p - 1) {
      const c128 a = cpendA[b * cstride + col];
      s.re = fma(C.beta_prev, a.re, s.re);
      s.im = fma(C.beta_prev, a.im, s.im);
    }
    h[b * hstride + col] = s;
  }
}

// Scalars of the slot (see above). Single steps (S0, S1) as k_lcgs_finish, with the correction of a
// pending v_{p-1} added to D; a pair's scalars on block 0, its vector work left to k_pair_close.
// psc[b][0..4]: coefficients of x_{j+2} on Q, v_j and the stored z / s, the scale of the stored z
// for v_{j+1}, and s.
__global__ void __launch_bounds__(kThreads) k_pair_finish(const c128* __restrict__ V, long long vstride, long long n,
                                                          const c128* __restrict__ h, int hstride,
                                                          c128* __restrict__ cpendA, c128* __restrict__ cpend,
                                                          c128* __restrict__ gfix, int cstride,
                                                          const double* __restrict__ atilde,
                                                          const double* __restrict__ atilde2, double* __restrict__ alpha,
                                                          double* __restrict__ beta, int astride, int j, double tol,
                                                          c128* __restrict__ X0, c128* __restrict__ X1,
                                                          double* __restrict__ fpart, unsigned* __restrict__ counter,
                                                          BasisCtl* ctl, double* __restrict__ psc, int force_short) {
  const int b = blockIdx.y;
  BasisCtl& C = ctl[b];
  const int mode = pair_mode(C, j);
  if (mode == kModeIdle) return;
  __shared__ double sm[kThreads / kWave];
  __shared__ c128 smc[kThreads / kWave];
  __shared__ c128 gs[kMaxBasis + 2];
  __shared__ c128 ca1[kMaxBasis + 2];
  __shared__ int is_last;
  const c128* hb = h + b * hstride;
  if (mode != kModePair) {
    const int p = mode == kModeS1 ? j + 1 : j;
    finish_step(V + b * vstride, n, hb[kColS].re, hb, cpend + b * cstride,
                mode == kModeS1 ? atilde2[b] : atilde[b], alpha + b * astride, beta + b * astride, p, tol,
                (mode == kModeS1 ? X0 : X1) + b * n, fpart + static_cast<long long>(b) * gridDim.x, counter + b, C, sm,
                gs, &is_last);
    return;
  }
  if (blockIdx.x != 0) return;
  const int t = threadIdx.x;
  const bool np2 = C.npend == 2;
  const double bp = C.beta_prev, s = C.scale, at = atilde[b], zq = atilde2[b];
  c128* ca = cpendA + b * cstride;
  c128* cb = cpend + b * cstride;
  c128* gf = gfix + b * cstride;
  double* ps = psc + b * 8;
  const double yy = hb[kColS].re, vv = hb[kColS].im, qq = hb[kColS + 1].re, zz = hb[kColS + 1].im;
  const c128 zv = hb[kColS + 2];
  // step j: D (with the pending v_{j-1} correction), alpha_j, beta_j
  double gg = 0.0;
  c128 graw = cmake(0.0, 0.0);     // g_j without the beta_{j-1} c_A part, less alpha~_j c_B
  if (t < j) {
    const c128 hv = hb[t];
    c128 v = hv;
    if (np2 && t < j - 1) {
      const c128 a = ca[t];
      v.re = fma(bp, a.re, v.re);
      v.im = fma(bp, a.im, v.im);
    }
    gg = v.re * v.re + v.im * v.im;
    const c128 c = cb[t];
    graw = cmake(fma(-at, c.re, hv.re), fma(-at, c.im, hv.im));
  }
  const c128 gam = hb[j];
  const double cprev = cb[j - 1].re;
  const c128 ej = hb[kColE + j];
  c128 ev = t <= j ? hb[kColE + t] : cmake(0.0, 0.0);
  gg = block_sum(gg, sm);          // ends with a barrier: the reads above precede the writes below
  const c128 glast = cmake(gam.re - at, gam.im);
  gg += glast.re * glast.re + glast.im * glast.im;
  const double aj = gam.re - bp * cprev;
  if (t == 0) alpha[b * astride + j] = aj;
  const double b2 = yy - gg;
  if (!(b2 > kCancel * yy) || force_short) {
    // near-breakdown at step j (or the test hook): k_pair_close corrects z and stops the basis at j + 1
    if (t < j) gf[t] = graw;
    if (t == 0) {
      gf[j] = glast;
      ps[4] = s;
      C.flags = (C.flags & ~kCloseMask) | kCloseRetreat;
    }
    return;
  }
  const double bj = std::sqrt(b2);
  if (t == 0) beta[b * astride + j] = bj;
  if (bj <= tol) {
    if (t == 0) {
      C.active = 0;
      C.size = j + 1;
    }
    return;
  }
  const double ib = 1.0 / bj;
  if (t < j) ca1[t] = cscale(graw, ib);
  if (t == 0) ca1[j] = cscale(glast, ib);
  __syncthreads();
  // step j+1 by linearity
  c128 d1 = cmake(0.0, 0.0);
  if (t <= j) {
    d1 = cscale(ev, ib);
    if (t == j) d1.re = fma(-bj, vv, d1.re);
  }
  const double at1 = zq * ib * ib;
  const c128 cv = t <= j ? ca1[t] : cmake(0.0, 0.0);
  const c128 dot = block_sum(cfmaconj(cv, d1, cmake(0.0, 0.0)), smc);
  const c128 gam1 = cmake(at1 - zv.re - dot.re, -zv.im - dot.im);
  if (t == 0) alpha[b * astride + j + 1] = gam1.re - bj * ca1[j].re;
  const double zsc = s * ib;
  if (t == 0) {
    ps[0] = ib;
    ps[1] = bj;
    ps[2] = at1 * s * ib;
    ps[3] = zsc;
  }
  if (t <= j) ca[t] = cv;
  if (j + 2 >= C.target) {
    // v_{j+1} is the basis's last vector: k_pair_close materialises it
    if (t == 0) C.flags = (C.flags & ~kCloseMask) | kCloseTarget;
    return;
  }
  c128 g1 = cmake(0.0, 0.0);
  if (t <= j) g1 = cmake(fma(-at1, cv.re, d1.re), fma(-at1, cv.im, d1.im));
  if (t == j + 1) g1 = cmake(gam1.re - at1, gam1.im);
  const double gg1 = block_sum(g1.re * g1.re + g1.im * g1.im, sm);
  const double caa = block_sum(cv.re * cv.re + cv.im * cv.im, sm);
  const double yy1 = qq * ib * ib + bj * bj * vv + at1 * at1 * zz * ib * ib - 2.0 * ej.re - 2.0 * at1 * zq * ib * ib +
                     2.0 * at1 * zv.re;
  const double b21 = yy1 - gg1;
  const double thr2 = kPairFixThreshold * kPairFixThreshold;
  bool fix = j + 2 >= C.stop || !(b21 > kCancel * yy1) || gg1 > thr2 * b21 || caa > thr2;
  const double b1 = fix ? 0.0 : std::sqrt(b21);
  if (!fix && b1 <= tol) fix = true;
  if (fix) {
    if (t <= j + 1) gf[t] = g1;
    if (t == 0) C.flags = (C.flags & ~kCloseMask) | kCloseFix;
    return;
  }
  const double i1 = 1.0 / b1;
  if (t <= j + 1) cb[t] = cscale(g1, i1);
  if (t == 0) {
    beta[b * astride + j + 1] = b1;
    C.scale = i1;
    C.beta_prev = b1;
    C.zscale = zsc;
    C.npend = 2;
    C.flags = (C.flags & ~kCloseMask) | kCloseGamma;
  }
}

// Vector work of a pair after k_pair_finish; single steps of the slot leave one vector pending.
// Grid (blocks of kPassElems elements, batch), in one launch, or in two (kFull, then the last
// partial block). Three-term close: x_{j+2} elementwise. Explicit closes stream V_<=j as the pass
// does (waves over columns, the next column's loads in flight) for the corrections V_<=j c_A'
// (v_{j+1}) and V_<=j g (x_{j+2} or, for a retreat, z), and leave their blocks' |x|^2 for k_pair_norm.
template <bool kStream, bool kFull>
__global__ void __launch_bounds__(kThreads) k_pair_close(c128* __restrict__ V, long long vstride, long long n, int j,
                                                         const c128* __restrict__ Z, const c128* __restrict__ Q,
                                                         c128* __restrict__ X0, c128* __restrict__ X1,
                                                         const c128* __restrict__ cpendA, c128* __restrict__ cpend,
                                                         const c128* __restrict__ gfix, int cstride,
                                                         const double* __restrict__ psc, double* __restrict__ fpart,
                                                         BasisCtl* ctl, int blk0, int nblk) {
  const int b = blockIdx.y;
  BasisCtl& C = ctl[b];
  const int mode = pair_mode(C, j);
  const int blk = blk0 + blockIdx.x;
  if (mode == kModeS0 || mode == kModeS1) {
    if (blk == 0 && threadIdx.x == 0) {
      C.npend = 1;
      C.flags &= ~kCloseMask;
    }
    return;
  }
  if (mode != kModePair) return;
  const int cm = C.flags & kCloseMask;
  if (cm == 0) return;
  const int t = threadIdx.x, lane = t % kWave, wq = t / kWave;
  const long long x0 = static_cast<long long>(blk) * kPassElems;
  const int cnt = kFull ? kPassElems : static_cast<int>(n - x0 < kPassElems ? n - x0 : kPassElems);
  const bool mine = t < cnt;
  const long long e = x0 + (mine ? t : cnt - 1);
  const double* ps = psc + b * 8;
  const double kq = ps[0], kv = ps[1], kz = ps[2], zsc = ps[3], sz = ps[4];
  c128* Vb = V + b * vstride;
  const c128 zv = Z[b * n + e];
  if (cm == kCloseGamma) {
    if (mine) {
      const c128 vj = Vb[static_cast<long long>(j) * n + e];
      c128 x = cscale(Q[b * n + e], kq);
      x = cmake(fma(-kv, vj.re, x.re), fma(-kv, vj.im, x.im));
      x = cmake(fma(-kz, zv.re, x.re), fma(-kz, zv.im, x.im));
      X0[b * n + e] = x;
    }
    return;
  }
  __shared__ c128 sa[kMaxBasis + 2];
  __shared__ c128 sg[kMaxBasis + 2];
  __shared__ c128 s_corr[kPassWaves][kPassElems];
  __shared__ double sm[kThreads / kWave];
  const bool needA = cm != kCloseRetreat, needG = cm != kCloseTarget;
  // (a retreat has no c_A', a target close no g: their products are computed and not used)
  if (t <= j) sa[t] = needA ? cpendA[b * cstride + t] : cmake(0.0, 0.0);
  if (t <= j + 1) sg[t] = needG ? gfix[b * cstride + t] : cmake(0.0, 0.0);
  __syncthreads();
  const int ncol = j + 1 > wq ? (j + 1 - wq + kPassWaves - 1) / kPassWaves : 0;
  c128 corrA[kPassPer], corrG[kPassPer], va[kPassPer], vb[kPassPer];
#pragma unroll
  for (int it = 0; it < kPassPer; ++it) {
    corrA[it] = cmake(0.0, 0.0);
    corrG[it] = cmake(0.0, 0.0);
  }
  auto load = [&](int k, c128(&v)[kPassPer]) {
    pass_load<kStream, kFull>(Vb + static_cast<long long>(wq + kPassWaves * k) * n + x0, lane, cnt, v);
  };
  auto use = [&](int k, const c128(&v)[kPassPer]) {
    const int c = wq + kPassWaves * k;
    const c128 a = sa[c], g = sg[c];
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) {
      corrA[it] = cfma(v[it], a, corrA[it]);
      corrG[it] = cfma(v[it], g, corrG[it]);
    }
  };
  // the scheduling barriers keep each column's loads issued before the previous column's use (with
  // few registers in use the compiler would otherwise share va's and vb's registers and wait for
  // every load before the next column's)
  if (ncol > 0) load(0, va);
  int k = 0;
  for (; k + 2 < ncol; k += 2) {
    load(k + 1, vb);
    __builtin_amdgcn_sched_barrier(0);
    use(k, va);
    load(k + 2, va);
    __builtin_amdgcn_sched_barrier(0);
    use(k + 1, vb);
  }
  if (k + 1 < ncol) {
    load(k + 1, vb);
    __builtin_amdgcn_sched_barrier(0);
    use(k, va);
    use(k + 1, vb);
  } else if (k < ncol) {
    use(k, va);
  }
  // corrections summed over the waves in a fixed order
  c128 sA = cmake(0.0, 0.0), sG = cmake(0.0, 0.0);
  if (needA) {
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) s_corr[wq][it * kWave + lane] = corrA[it];
    __syncthreads();
    sA = cadd(cadd(cadd(s_corr[0][t], s_corr[1][t]), s_corr[2][t]), s_corr[3][t]);
    __syncthreads();
  }
  if (needG) {
#pragma unroll
    for (int it = 0; it < kPassPer; ++it) s_corr[wq][it * kWave + lane] = corrG[it];
    __syncthreads();
    sG = cadd(cadd(cadd(s_corr[0][t], s_corr[1][t]), s_corr[2][t]), s_corr[3][t]);
  }
  double ssum = 0.0;
  if (mine) {
    if (cm == kCloseRetreat) {
      const c128 y = csub(cscale(zv, sz), sG);
      X1[b * n + e] = y;
      ssum = fma(y.re, y.re, y.im * y.im);
    } else {
      const c128 v = csub(cscale(zv, zsc), sA);
      Vb[static_cast<long long>(j + 1) * n + e] = v;
      if (cm == kCloseFix) {
        const c128 vj = Vb[static_cast<long long>(j) * n + e];
        c128 x = cscale(Q[b * n + e], kq);
        x = cmake(fma(-kv, vj.re, x.re), fma(-kv, vj.im, x.im));
        x = cmake(fma(-kz, zv.re, x.re), fma(-kz, zv.im, x.im));
        x = csub(x, sG);
        x = cfms(v, sg[j + 1], x);
        X0[b * n + e] = x;
        ssum = fma(x.re, x.re, x.im * x.im);
      }
    }
  }
  if (cm == kCloseTarget) return;
  ssum = block_sum(ssum, sm);
  if (t == 0) fpart[static_cast<long long>(b) * nblk + blk] = ssum;
}

// The norm of an explicit close (k_pair_close's block sums of |x|^2, or |y|^2 for a retreat, summed in
// a fixed order) and the control-block update that goes with it. A kernel of its own: a last-block
// reduction inside k_pair_close would put a device-scope fence (a write-back of the whole L2 on this
// chip) in each of its blocks, which made the close of a prefix flush take 2.5 ms at 32 trajectories.
// Grid (batch), one block per trajectory.
__global__ void __launch_bounds__(kThreads) k_pair_norm(const double* __restrict__ fpart, int nblk, int j,
                                                        c128* __restrict__ cpend, int cstride,
                                                        double* __restrict__ beta, int astride, double tol,
                                                        BasisCtl* ctl) {
  const int b = blockIdx.x;
  BasisCtl& C = ctl[b];
  if (pair_mode(C, j) != kModePair) return;
  const int cm = C.flags & kCloseMask;
  if (cm != kCloseFix && cm != kCloseRetreat) return;
  __shared__ double sm[kThreads / kWave];
  const int t = threadIdx.x;
  const double* fp = fpart + static_cast<long long>(b) * nblk;
  double tot = 0.0;
  for (int i = t; i < nblk; i += kThreads) tot += fp[i];
  tot = block_sum(tot, sm);   // ends with a barrier: every thread has read C before it is written
  const int kk = cm == kCloseRetreat ? j + 1 : j + 2;
  if (t < kk) cpend[b * cstride + t] = cmake(0.0, 0.0);
  if (t == 0) {
    C.nfix += 1;
    const double bn = std::sqrt(tot);
    beta[b * astride + kk - 1] = bn;
    C.npend = 1;
    C.flags &= ~kCloseMask;
    if (bn <= tol) {
      C.active = 0;
      C.size = kk;
    } else {
      C.scale = 1.0 / bn;
      C.beta_prev = bn;
      if (cm == kCloseRetreat) {
        C.stop = j + 1;
        C.flags |= kBasisShort;
      }
    }
  }
}

// info of the trajectories that took part: vectors built so far (size is the breakdown size or the
// eventual target), no-breakdown flag, explicit corrections, and 1 when a paired step stopped the
// basis early at a near-breakdown (the vectors so far are valid; a later call continues it).
__global__ void k_basis_info(const BasisCtl* ctl, int* info, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch && ctl[b].part) {
    info[b * 4 + 0] = min(ctl[b].size, ctl[b].stop);
    info[b * 4 + 1] = ctl[b].active;
    info[b * 4 + 2] = ctl[b].nfix;
    info[b * 4 + 3] = (ctl[b].flags & kBasisShort) ? 1 : 0;
  }
}

// ---------------------------------------------------------------- combine / gram

// out_b[x] = scale_b * sum_{j<k_b} V_b[j][x] c_b[j]   (k_b from the array or the scalar k)
__global__ void __launch_bounds__(kThreads) k_combine(const c128* __restrict__ V, long long vstride, int k,
                                                      const int* __restrict__ kb, long long n,
                                                      const c128* __restrict__ c, int cstride,
                                                      const double* __restrict__ scale, c128* __restrict__ out,
                                                      long long ostride) {
  const int b = blockIdx.y;
  const int kk = kb ? min(kb[b], k) : k;   // with per-trajectory counts, k is their upper limit
  if (kk <= 0) return;
  const long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const c128* Vb = V + b * vstride;
  const c128* cb = c + b * cstride;
  c128 s{0.0, 0.0};
  for (int j = 0; j < kk; ++j) s = cfma(Vb[static_cast<long long>(j) * n + i], cb[j], s);
  const double sc = scale ? scale[b] : 1.0;
  out[b * ostride + i] = cscale(s, sc);
}

constexpr int kGramTile = 64;
constexpr int kGramChunk = 16;

// Partial S[a][c] = sum_n conj(A_a[n]) A_c[n] for the 4x4 blocks on and above the diagonal of the k x k
// matrix, k <= 64, A_a[n] = (V_a[nn + n] - V_a[2 nn + n]) / sqrt2 (k_gram_sym completes the Hermitian
// matrix). Thread t < T (T + 1) / 2, T = ceil(k / 4), computes the t-th such block (ia, ib), ia <= ib,
// in row order; every thread stages. Entries below the diagonal blocks are not written. Grid (nblk, batch).
__global__ void __launch_bounds__(256) k_gram(const c128* __restrict__ V, long long vstride, const int* __restrict__ kb,
                                              int k1, long long nn, long long n_per_block, c128* __restrict__ partial) {
  const int bt = blockIdx.y;
  const int k = kb ? min(kb[bt], k1) : k1;   // per-trajectory counts are clamped to the basis
  if (k <= 0) return;
  __shared__ c128 ra[kGramChunk][kGramTile + 1];
  const int tid = threadIdx.x;
  const int T4 = (k + 3) / 4;
  const bool act = tid < T4 * (T4 + 1) / 2;
  int ia = 0, ib = 0;
  if (act) {
    int q = tid;
    while (q >= T4 - ia) {
      q -= T4 - ia;
      ++ia;
    }
    ib = ia + q;
  }
  const long long N = 4 * nn;
  const c128* Vb = V + bt * vstride;
  const long long n0 = static_cast<long long>(blockIdx.x) * n_per_block;
  const long long n1 = min(nn, n0 + n_per_block);
  const double r2 = 1.0 / std::sqrt(2.0);
  c128 acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = cmake(0.0, 0.0);
  for (long long c0 = n0; c0 < n1; c0 += kGramChunk) {
    for (int s = tid; s < kGramTile * kGramChunk; s += 256) {
      const int row = s / kGramChunk, col = s % kGramChunk;
      const long long nidx = c0 + col;
      c128 va{0.0, 0.0};
      if (nidx < n1 && row < k) {
        const c128* v = Vb + static_cast<long long>(row) * N;
        const c128 p = v[nn + nidx], q = v[2 * nn + nidx];
        va = cmake((p.re - q.re) * r2, (p.im - q.im) * r2);
      }
      ra[col][row] = va;
    }
    __syncthreads();
    for (int col = 0; col < kGramChunk && act; ++col) {
      c128 xa[4], xb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xa[u] = ra[col][ia * 4 + u];
        xb[u] = ra[col][ib * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = cfmaconj(xa[u], xb[v], acc[u][v]);
    }
    __syncthreads();
  }
  if (!act) return;
  c128* out = partial + (static_cast<long long>(bt) * gridDim.x + blockIdx.x) * kGramTile * kGramTile;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) out[(ia * 4 + u) * kGramTile + ib * 4 + v] = acc[u][v];
}

// P_b = the Hermitian matrix whose upper triangle is the sum over blocks of the partial tiles (k x k
// row-major at P + b*pstride), in two steps so that every read is coalesced: k_gram_sum adds the tiles
// (each entry a <= c over the blocks in ascending order) into P, k_gram_sym sets P[c][a] = conj(P[a][c])
// and takes the real part of the diagonal.
__global__ void __launch_bounds__(256) k_gram_sum(const c128* __restrict__ partial, int nblk,
                                                  const int* __restrict__ kb, int k1, c128* __restrict__ P,
                                                  long long pstride) {
  const int bt = blockIdx.y;
  const int k = kb ? min(kb[bt], k1) : k1;   // per-trajectory counts are clamped to the basis
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= k * k) return;
  const int a = e / k, c = e % k;
  if (a > c) return;
  const c128* pb = partial + static_cast<long long>(bt) * nblk * kGramTile * kGramTile + a * kGramTile + c;
  c128 x{0.0, 0.0};
  // unrolled so that the loads run ahead of the (ordered) additions
#pragma unroll 8
  for (int i = 0; i < nblk; ++i) x = cadd(x, pb[static_cast<long long>(i) * kGramTile * kGramTile]);
  P[bt * pstride + e] = x;
}

__global__ void __launch_bounds__(256) k_gram_sym(const int* __restrict__ kb, int k1, c128* __restrict__ P,
                                                  long long pstride) {
  const int bt = blockIdx.y;
  const int k = kb ? min(kb[bt], k1) : k1;   // per-trajectory counts are clamped to the basis
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= k * k) return;
  const int a = e / k, c = e % k;
  if (a > c) return;
  c128* pb = P + bt * pstride;
  const c128 x = pb[a * k + c];
  if (a == c) {
    pb[a * k + a] = cmake(x.re, 0.0);
  } else {
    pb[c * k + a] = cmake(x.re, -x.im);
  }
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
  hipLaunchKernelGGL(k_sumsq, dim3(nblk, batch), dim3(kThreads), 0, s, static_cast<const c128*>(d_x), n, ld, part,
                     nullptr);
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
  hipLaunchKernelGGL(k_singlet_sumsq, dim3(nblk, batch), dim3(kThreads), 0, s, static_cast<const c128*>(d_x), nn, ld,
                     part);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_finish_sum, dim3(batch), dim3(kThreads), 0, s, part, nblk, d_out, 0);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

namespace {
struct BasisWs {
  c128* dots;        // [batch][units] matvec partials of alpha~_j
  int units;
  c128* w[2];        // [batch][n] each: the pending x_p in w[p % 2]
  c128* partial;     // [batch][cols][npass] pass column partials (column j: gamma)
  long long pstride;
  int npass;
  double* ypart;     // [batch][npass] |y|^2 partials
  c128* h;           // [batch][cols] reduced column sums
  int hstride;
  c128* cpend;       // [batch][m] pending correction c_{j+1}
  double* fpart;     // [batch][kFixBlocks]
  double* atilde;    // [batch] alpha~_j
  double* ysum;      // [batch] |y|^2
  unsigned* counter; // [batch]
  double* sq;        // [batch][nchunk]
  BasisCtl* ctl;     // [batch]
  // paired steps only (null otherwise)
  c128* z[2];        // [batch][n] each: z / s of the current and the previous pair
  c128* q;           // [batch][n] H z
  c128* dots2;       // [batch][units] partials of <z, H z>
  double* atilde2;   // [batch]
  c128* cpendA;      // [batch][m] pending correction of the pending z
  c128* gfix;        // [batch][m + 2] coefficients of an explicit correction
  double* psc;       // [batch][8] scalars of the pair close
  size_t bytes;
};

BasisWs basis_layout(char* base, long long n, int m, int batch, int units, bool paired) {
  const int nchunk = ceil_div(n, kChunk);
  const int npass = ceil_div(n, static_cast<long long>(kPassElems));
  const int cols = paired ? kPairCols : m;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~size_t(255);
    return o;
  };
  const size_t B = static_cast<size_t>(batch);
  const size_t o_w0 = take(B * n * 16), o_w1 = take(B * n * 16);
  const size_t o_p = take(B * cols * npass * 16), o_yp = take(B * npass * 8), o_h = take(B * cols * 16);
  const size_t o_c = take(B * m * 16), o_fp = take(B * kFixBlocks * 8), o_cnt = take(B * sizeof(unsigned));
  const size_t o_at = take(B * 8), o_ys = take(B * 8);
  const size_t o_sq = take(B * nchunk * 8), o_ctl = take(B * sizeof(BasisCtl));
  const size_t o_dots = take(B * static_cast<size_t>(units) * 16);
  BasisWs w;
  if (paired) {
    const size_t o_z0 = take(B * n * 16), o_z1 = take(B * n * 16), o_q = take(B * n * 16);
    const size_t o_d2 = take(B * static_cast<size_t>(units) * 16), o_a2 = take(B * 8);
    const size_t o_ca = take(B * m * 16), o_gf = take(B * (m + 2) * 16), o_ps = take(B * 8 * 8);
    w.z[0] = reinterpret_cast<c128*>(base + o_z0);
    w.z[1] = reinterpret_cast<c128*>(base + o_z1);
    w.q = reinterpret_cast<c128*>(base + o_q);
    w.dots2 = reinterpret_cast<c128*>(base + o_d2);
    w.atilde2 = reinterpret_cast<double*>(base + o_a2);
    w.cpendA = reinterpret_cast<c128*>(base + o_ca);
    w.gfix = reinterpret_cast<c128*>(base + o_gf);
    w.psc = reinterpret_cast<double*>(base + o_ps);
  } else {
    w.z[0] = w.z[1] = w.q = w.dots2 = w.cpendA = w.gfix = nullptr;
    w.atilde2 = w.psc = nullptr;
  }
  w.w[0] = reinterpret_cast<c128*>(base + o_w0);
  w.w[1] = reinterpret_cast<c128*>(base + o_w1);
  w.partial = reinterpret_cast<c128*>(base + o_p);
  w.pstride = static_cast<long long>(cols) * npass;
  w.npass = npass;
  w.ypart = reinterpret_cast<double*>(base + o_yp);
  w.h = reinterpret_cast<c128*>(base + o_h);
  w.hstride = cols;
  w.cpend = reinterpret_cast<c128*>(base + o_c);
  w.fpart = reinterpret_cast<double*>(base + o_fp);
  w.atilde = reinterpret_cast<double*>(base + o_at);
  w.ysum = reinterpret_cast<double*>(base + o_ys);
  w.counter = reinterpret_cast<unsigned*>(base + o_cnt);
  w.sq = reinterpret_cast<double*>(base + o_sq);
  w.ctl = reinterpret_cast<BasisCtl*>(base + o_ctl);
  w.dots = reinterpret_cast<c128*>(base + o_dots);
  w.units = units;
  w.bytes = off;
  return w;
}

template <bool kStream>
void launch_pass_kind(const BasisWs& w, c128* V, long long vstride, long long n, int j, const c128* in, long long ild,
                      int m, bool fuse_alpha, c128* out, int batch, hipStream_t s) {
  const int nfull = static_cast<int>(n / kPassElems);
  const c128* adots = fuse_alpha ? w.dots : nullptr;
  if (nfull > 0)
    hipLaunchKernelGGL((k_lcgs_pass<kStream, true>), dim3(nfull, batch), dim3(kThreads), 0, s, V, vstride, n, j, in,
                       ild, w.cpend, m, w.atilde, adots, w.units, out, w.partial, w.pstride, w.ypart, w.ctl, 0,
                       w.npass);
  if (nfull < w.npass)
    hipLaunchKernelGGL((k_lcgs_pass<kStream, false>), dim3(w.npass - nfull, batch), dim3(kThreads), 0, s, V, vstride,
                       n, j, in, ild, w.cpend, m, w.atilde, adots, w.units, out, w.partial, w.pstride, w.ypart, w.ctl,
                       nfull, w.npass);
}

void launch_pass(const BasisWs& w, c128* V, long long vstride, long long n, int j, const c128* in, long long ild, int m,
                 bool fuse_alpha, c128* out, int batch, hipStream_t s) {
  if (batch > 1)
    launch_pass_kind<true>(w, V, vstride, n, j, in, ild, m, fuse_alpha, out, batch, s);
  else
    launch_pass_kind<false>(w, V, vstride, n, j, in, ild, m, fuse_alpha, out, batch, s);
}

// Shared implementation of the single and batched basis builders.
// Steps j_begin <= j < max_stop of the bases of this call (per trajectory: its window from
// k_basis_start); without stops, every basis is built whole in one call (j_begin 0, max_stop the
// target).
// paired_layout: the workspace holds the buffers of paired steps (the prefix builder's layout, so that
// calls with and without paired steps can continue each other's bases); paired: steps from
// kPairFrom on run in pairs; force_short (a test hook): every pair stops its trajectory's call as a
// near-breakdown at its first step would.
int basis_batch(const rpm_operator* op, int batch, const c128* states, long long sld, const int* d_targets,
                int max_target, const int* d_stops, int j_begin, int max_stop, const c128* pairs, c128* V,
                long long vstride, int m, double* alpha, double* beta, int* info, double* norm, double tol,
                void* ws_ptr, size_t ws_bytes, hipStream_t s, bool paired_layout = false, bool paired = false,
                bool force_short = false) {
  const long long n = op->n;
  if (max_target > m || max_target < 1) return fail(RPM_E_INVALID, "target %d outside 1..%d", max_target, m);
  if (max_stop > max_target || max_stop < 1 || j_begin < 0 || j_begin >= max_stop)
    return fail(RPM_E_INVALID, "basis steps %d..%d outside 0..%d", j_begin, max_stop, max_target);
  if (m > kMaxBasis) return fail(RPM_E_UNSUPPORTED, "Krylov dimension above %d is not supported", kMaxBasis);
  BasisWs w = basis_layout(static_cast<char*>(ws_ptr), n, m, batch, apply_units(op), paired_layout);
  if (ws_bytes < w.bytes) return fail(RPM_E_INVALID, "Lanczos workspace %zu < %zu", ws_bytes, w.bytes);
  if (paired && !paired_layout) return fail(RPM_E_INVALID, "paired Lanczos steps need the prefix workspace");
  const int nchunk = ceil_div(n, kChunk);
  const int full = static_cast<int>(max_target < n ? max_target : n);
  const int count = static_cast<int>(max_stop < n ? max_stop : n);
  hipLaunchKernelGGL(k_sumsq, dim3(nchunk, batch), dim3(kThreads), 0, s, states, n, sld, w.sq, d_targets);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_basis_start, dim3(batch), dim3(kThreads), 0, s, w.sq, nchunk, norm, w.ctl, w.counter, d_targets,
                     d_stops, full, n);
  RPM_LAUNCH_CHECK();
  const int single_end = paired ? std::min(count, std::max(j_begin, kPairFrom)) : count;
  for (int j = j_begin; j < single_end; ++j) {
    const c128* in = j == 0 ? states : w.w[j % 2];
    c128* out = w.w[(j + 1) % 2];
    // matvec on x_j = s in: out = H x_j and the per-unit partials of alpha~_j = <x_j, H x_j>; the
    // pass subtracts beta_{j-1} V[j-1] and writes V[j]
    const long long ild = j == 0 ? sld : n;
    ApplyCtl A;
    A.ctl = w.ctl;
    A.j = j;
    A.pairs = pairs;
    A.dots = w.dots;
    A.dots_ld = w.units;
    int rc = operator_apply_ex(op, in, out, batch, ild, A, s);
    if (rc) return rc;
    // A single trajectory forms alpha~_j inside the pass (one launch fewer per step). For a batch
    // the per-block recomputation costs more than the launch, so k_alpha runs; both sum in the
    // same order, so the bases stay bitwise equal.
    const bool fuse_alpha = batch == 1;
    if (!fuse_alpha) {
      hipLaunchKernelGGL(k_alpha, dim3(batch), dim3(kThreads), 0, s, w.dots, w.units, w.atilde, j, w.ctl);
      RPM_LAUNCH_CHECK();
    }
    // one pass over V[0..j): v_j materialised, V^H u, gamma, y = u - alpha~ v_j, |y|^2
    launch_pass(w, V, vstride, n, j, in, ild, m, fuse_alpha, out, batch, s);
    RPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_cols, dim3(j + 2, batch), dim3(kThreads), 0, s, w.partial, w.npass, w.pstride, w.h,
                       w.hstride, nullptr, m, 0, j, 0, w.ctl, w.ypart, w.ysum);
    RPM_LAUNCH_CHECK();
    // alpha_j, beta_j, c_{j+1}; the explicit correction when it is due
    // the last step of a basis forms alpha only (one block); a prefix's last step also forms beta
    // and the pending correction, so that a later call can resume it
    hipLaunchKernelGGL(k_lcgs_finish, dim3(j == full - 1 ? 1 : kFixBlocks, batch), dim3(kThreads), 0, s, V, vstride, n,
                       w.ysum, w.h, w.hstride, w.cpend, m, w.atilde, alpha, beta, m, j, tol, out, w.fpart, w.counter,
                       w.ctl);
    RPM_LAUNCH_CHECK();
  }
  // paired slots (j, j + 1), j even; each trajectory takes a pair, a single step or nothing in a slot
  // from its own window, so its arithmetic does not depend on the other trajectories of the batch
  // (a slot that starts at an odd j_begin - 1 has nothing live at its first index: every window
  // of the call begins at j_begin or later)
  for (int j = single_end & ~1; single_end < count && j < count; j += 2) {
    c128* zc = w.z[(j / 2) % 2];
    const c128* zp = w.z[(j / 2 + 1) % 2];
    ApplyCtl A;
    A.ctl = w.ctl;
    A.j = j;
    A.pairs = pairs;
    A.dots = w.dots;
    A.dots_ld = w.units;
    int rc = operator_apply_ex(op, w.w[0], w.w[1], batch, n, A, s);   // P = H x^_j into w[1]
    if (rc) return rc;
    // A single trajectory forms both alpha~ sums inside the prep and the pass (two launches fewer
    // per slot); a batch runs k_alpha. Both sum in the same order: the bases stay bitwise equal.
    const bool fuse_alpha = batch == 1;
    const c128* fdots = fuse_alpha ? w.dots : nullptr;
    const c128* fdots2 = fuse_alpha ? w.dots2 : nullptr;
    if (!fuse_alpha) {
      hipLaunchKernelGGL(k_alpha, dim3(batch), dim3(kThreads), 0, s, w.dots, w.units, w.atilde, j, w.ctl);
      RPM_LAUNCH_CHECK();
    }
    if (fuse_alpha)
      hipLaunchKernelGGL(k_pair_prep<true>, dim3(ceil_div(n, kThreads), batch), dim3(kThreads), 0, s, V, vstride, n,
                         j, w.w[0], w.w[1], zp, zc, w.atilde, w.dots, w.units, w.ctl);
    else
      hipLaunchKernelGGL(k_pair_prep<false>, dim3(ceil_div(n, kThreads), batch), dim3(kThreads), 0, s, V, vstride, n,
                         j, w.w[0], w.w[1], zp, zc, w.atilde, nullptr, w.units, w.ctl);
    RPM_LAUNCH_CHECK();
    A.j = j + 1;
    A.dots = w.dots2;
    rc = operator_apply_ex(op, zc, w.q, batch, n, A, s);               // Q = H z (S1: H x^_{j+1})
    if (rc) return rc;
    if (!fuse_alpha) {
      hipLaunchKernelGGL(k_alpha, dim3(batch), dim3(kThreads), 0, s, w.dots2, w.units, w.atilde2, j + 1, w.ctl);
      RPM_LAUNCH_CHECK();
    }
    const int nfull = static_cast<int>(n / kPassElems);
    auto pass = [&](auto kernel, int b0, int nb) {
      hipLaunchKernelGGL(kernel, dim3(nb, batch), dim3(kThreads), 0, s, V, vstride, n, j, w.w[0], w.w[1], zp, zc,
                         w.q, w.cpendA, w.cpend, m, w.atilde, w.atilde2, fdots, fdots2, w.units, w.partial, w.pstride,
                         w.ctl, b0, w.npass);
    };
    if (!fuse_alpha) {
      if (nfull > 0) pass(k_pair_pass<true, true, false>, 0, nfull);
      if (nfull < w.npass) pass(k_pair_pass<true, false, false>, nfull, w.npass - nfull);
    } else {
      if (nfull > 0) pass(k_pair_pass<false, true, true>, 0, nfull);
      if (nfull < w.npass) pass(k_pair_pass<false, false, true>, nfull, w.npass - nfull);
    }
    RPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pair_reduce, dim3(2 * j + 6, batch), dim3(kThreads), 0, s, w.partial, w.npass, w.pstride, w.h,
                       w.hstride, j, w.cpendA, m, w.ctl);
    RPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pair_finish, dim3(kFixBlocks, batch), dim3(kThreads), 0, s, V, vstride, n, w.h, w.hstride,
                       w.cpendA, w.cpend, w.gfix, m, w.atilde, w.atilde2, alpha, beta, m, j, tol, w.w[0], w.w[1],
                       w.fpart, w.counter, w.ctl, w.psc, force_short ? 1 : 0);
    RPM_LAUNCH_CHECK();
    // (the explicit close's block sums go to ypart, which single steps alone use)
    if (batch > 1) {
      if (nfull > 0)
        hipLaunchKernelGGL((k_pair_close<true, true>), dim3(nfull, batch), dim3(kThreads), 0, s, V, vstride, n, j, zc,
                           w.q, w.w[0], w.w[1], w.cpendA, w.cpend, w.gfix, m, w.psc, w.ypart, w.ctl, 0, w.npass);
      if (nfull < w.npass)
        hipLaunchKernelGGL((k_pair_close<true, false>), dim3(w.npass - nfull, batch), dim3(kThreads), 0, s, V, vstride,
                           n, j, zc, w.q, w.w[0], w.w[1], w.cpendA, w.cpend, w.gfix, m, w.psc, w.ypart, w.ctl, nfull,
                           w.npass);
    } else {
      if (nfull > 0)
        hipLaunchKernelGGL((k_pair_close<false, true>), dim3(nfull, batch), dim3(kThreads), 0, s, V, vstride, n, j, zc,
                           w.q, w.w[0], w.w[1], w.cpendA, w.cpend, w.gfix, m, w.psc, w.ypart, w.ctl, 0, w.npass);
      if (nfull < w.npass)
        hipLaunchKernelGGL((k_pair_close<false, false>), dim3(w.npass - nfull, batch), dim3(kThreads), 0, s, V,
                           vstride, n, j, zc, w.q, w.w[0], w.w[1], w.cpendA, w.cpend, w.gfix, m, w.psc, w.ypart, w.ctl,
                           nfull, w.npass);
    }
    RPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pair_norm, dim3(batch), dim3(kThreads), 0, s, w.ypart, w.npass, j, w.cpend, m, beta, m, tol,
                       w.ctl);
    RPM_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_basis_info, dim3(ceil_div(batch, 64)), dim3(64), 0, s, w.ctl, info, batch);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}
}  // namespace

extern "C" size_t rpm_lanczos_workspace_bytes(const rpm_operator* op, int m) {
  if (!op || m < 1) return 0;
  return basis_layout(nullptr, op->n, m, 1, apply_units(op), false).bytes;
}

extern "C" size_t rpm_lanczos_batch_workspace_bytes(const rpm_operator* op, int m, int batch) {
  if (!op || m < 1 || batch < 1) return 0;
  return basis_layout(nullptr, op->n, m, batch, apply_units(op), false).bytes;
}

extern "C" size_t rpm_lanczos_prefix_workspace_bytes(const rpm_operator* op, int m, int batch) {
  if (!op || m < 1 || batch < 1) return 0;
  return basis_layout(nullptr, op->n, m, batch, apply_units(op), true).bytes;
}

extern "C" int rpm_lanczos_basis(const rpm_operator* op, const void* d_state, int target, void* d_basis,
                                 double* d_alpha, double* d_beta, int* d_info, double* d_norm, double tol, void* d_ws,
                                 size_t ws_bytes, void* stream) {
  if (!op || !d_state || !d_basis || !d_alpha || !d_beta || !d_info || !d_norm || !d_ws || target < 1)
    return fail(RPM_E_INVALID, "rpm_lanczos_basis: bad arguments");
  const long long n = op->n;
  const int count = static_cast<int>(target < n ? target : n);
  hipStream_t s = static_cast<hipStream_t>(stream);
  BasisWs w = basis_layout(static_cast<char*>(d_ws), n, count, 1, apply_units(op), false);
  if (ws_bytes < w.bytes) return fail(RPM_E_INVALID, "rpm_lanczos_basis: workspace %zu < %zu", ws_bytes, w.bytes);
  return basis_batch(op, 1, static_cast<const c128*>(d_state), n, nullptr, count, nullptr, 0, count, nullptr,
                     static_cast<c128*>(d_basis), static_cast<long long>(count) * n, count, d_alpha, d_beta, d_info,
                     d_norm, tol, d_ws, ws_bytes, s);
}

// ---------------------------------------------------------------- allocation-free basis

// Scratch of rpm_lanczos_basis_noalloc: the single-trajectory basis workspace (two work vectors, the
// pass's column partials, control blocks and counters; basis_layout) lives in this array instead of
// caller or library allocations.  16 MiB holds it for the production system (N = 221,184, m = 64:
// about 8 MB) and for any N up to about 450,000 at m = 64.  Each device keeps its own copy for the
// life of the process.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the current device's copy of `scratch`, looked up once per device (with that
// device current) by hipGetSymbolAddress.  In host code the array's own name is not a device address.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(RPM_E_HIP, "rpm_lanczos_basis_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(RPM_E_UNSUPPORTED, "rpm_lanczos_basis_noalloc: device %d above %d", dev, kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(RPM_E_HIP, "rpm_lanczos_basis_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RPM_OK;
}
}  // namespace

extern "C" int rpm_lanczos_basis_noalloc(const rpm_operator* op, const void* d_state, int target, void* d_basis,
                                         double* d_alpha, double* d_beta, int* d_info, double* d_norm, double tol,
                                         void* stream) {
  if (!op || !d_state || !d_basis || !d_alpha || !d_beta || !d_info || !d_norm || target < 1)
    return fail(RPM_E_INVALID, "rpm_lanczos_basis_noalloc: bad arguments");
  const long long n = op->n;
  const int count = static_cast<int>(target < n ? target : n);
  if (count > kMaxBasis) return fail(RPM_E_UNSUPPORTED, "Krylov dimension above %d is not supported", kMaxBasis);
  const size_t need = basis_layout(nullptr, n, count, 1, apply_units(op), false).bytes;
  if (need > kNoAllocScratchBytes)
    return fail(RPM_E_INVALID,
                "rpm_lanczos_basis_noalloc: this basis (N = %lld, %d vectors) needs %zu bytes of scratch, above the "
                "%zu-byte limit; use rpm_lanczos_basis with a workspace",
                n, count, need, kNoAllocScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  // the counters and control blocks that must start at zero are reset by k_basis_start on every call
  return basis_batch(op, 1, static_cast<const c128*>(d_state), n, nullptr, count, nullptr, 0, count, nullptr,
                     static_cast<c128*>(d_basis), static_cast<long long>(count) * n, count, d_alpha, d_beta, d_info,
                     d_norm, tol, ws, kNoAllocScratchBytes, static_cast<hipStream_t>(stream));
}

extern "C" int rpm_lanczos_basis_batch(const rpm_operator* op, int batch, const void* d_states, long long ld,
                                       const int* d_targets, int max_target, const void* d_pairs, void* d_basis,
                                       int m, double* d_alpha, double* d_beta, int* d_info, double* d_norm, double tol,
                                       void* d_ws, size_t ws_bytes, void* stream) {
  if (!op || !d_states || !d_targets || !d_basis || !d_alpha || !d_beta || !d_info || !d_norm || !d_ws || batch < 1 ||
      batch > 65535 || m < 1)
    return fail(RPM_E_INVALID, "rpm_lanczos_basis_batch: bad arguments");
  // the first matvec reads the states and writes the packed work vectors with one stride
  if (ld != op->n) return fail(RPM_E_INVALID, "rpm_lanczos_basis_batch: states must be packed (ld == N)");
  return basis_batch(op, batch, static_cast<const c128*>(d_states), ld, d_targets, max_target, nullptr, 0, max_target,
                     static_cast<const c128*>(d_pairs), static_cast<c128*>(d_basis),
                     static_cast<long long>(m) * op->n, m, d_alpha, d_beta, d_info, d_norm, tol, d_ws, ws_bytes,
                     static_cast<hipStream_t>(stream));
}

extern "C" int rpm_lanczos_basis_batch_prefix(const rpm_operator* op, int batch, const void* d_states, long long ld,
                                              const int* d_targets, int max_target, const int* d_stops, int j_begin,
                                              int max_stop, const void* d_pairs, void* d_basis, int m,
                                              double* d_alpha, double* d_beta, int* d_info, double* d_norm,
                                              double tol, int flags, void* d_ws, size_t ws_bytes, void* stream) {
  if (!op || !d_states || !d_targets || !d_stops || !d_basis || !d_alpha || !d_beta || !d_info || !d_norm || !d_ws ||
      batch < 1 || batch > 65535 || m < 1 || (flags & ~(RPM_LANCZOS_PAIRED | RPM_LANCZOS_FORCE_SHORT)))
    return fail(RPM_E_INVALID, "rpm_lanczos_basis_batch_prefix: bad arguments");
  if (ld != op->n) return fail(RPM_E_INVALID, "rpm_lanczos_basis_batch_prefix: states must be packed (ld == N)");
  return basis_batch(op, batch, static_cast<const c128*>(d_states), ld, d_targets, max_target, d_stops, j_begin,
                     max_stop, static_cast<const c128*>(d_pairs), static_cast<c128*>(d_basis),
                     static_cast<long long>(m) * op->n, m, d_alpha, d_beta, d_info, d_norm, tol, d_ws, ws_bytes,
                     static_cast<hipStream_t>(stream), true, (flags & RPM_LANCZOS_PAIRED) != 0,
                     (flags & RPM_LANCZOS_FORCE_SHORT) != 0);
}

extern "C" int rpm_krylov_combine(const void* d_basis, int k, long long n, const void* d_c, const double* d_scale,
                                  void* d_out, void* stream) {
  if (!d_basis || !d_c || !d_out || k < 1 || n < 1) return fail(RPM_E_INVALID, "rpm_krylov_combine: bad arguments");
  hipLaunchKernelGGL(k_combine, dim3(ceil_div(n, kThreads), 1), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     static_cast<const c128*>(d_basis), 0LL, k, nullptr, n, static_cast<const c128*>(d_c), 0, d_scale,
                     static_cast<c128*>(d_out), 0LL);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_krylov_combine_batch(const void* d_basis, int m, const int* d_k, int batch, long long n,
                                        const void* d_c, int c_stride, const double* d_scale, void* d_out,
                                        long long out_ld, void* stream) {
  if (!d_basis || !d_k || !d_c || !d_out || batch < 1 || batch > 65535 || n < 1 || m < 1 || c_stride < 1 ||
      out_ld < n)
    return fail(RPM_E_INVALID, "rpm_krylov_combine_batch: bad arguments");
  // the device counts d_k cannot be checked here: the kernel clamps them to the basis and coefficient rows
  hipLaunchKernelGGL(k_combine, dim3(ceil_div(n, kThreads), batch), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const c128*>(d_basis),
                     static_cast<long long>(m) * n, std::min(m, c_stride), d_k, n, static_cast<const c128*>(d_c),
                     c_stride, d_scale, static_cast<c128*>(d_out), out_ld);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

namespace {
long long gram_per_block(long long nn) {
  // about 256 blocks, each an integer number of kGramChunk-wide chunks
  long long per = (nn + 255) / 256;
  per = ((per + kGramChunk - 1) / kGramChunk) * kGramChunk;
  if (per < kGramChunk) per = kGramChunk;
  return per;
}
}  // namespace

extern "C" size_t rpm_singlet_gram_workspace_bytes(int k, long long nn) {
  if (k < 1 || nn < 1) return 0;
  const long long per = gram_per_block(nn);
  const long long nblk = (nn + per - 1) / per;
  return static_cast<size_t>(nblk) * kGramTile * kGramTile * 16 + 512;
}

extern "C" size_t rpm_singlet_gram_batch_workspace_bytes(long long nn, int batch) {
  if (nn < 1 || batch < 1) return 0;
  return rpm_singlet_gram_workspace_bytes(1, nn) * static_cast<size_t>(batch) + 512;
}

namespace {
int gram_batch(const c128* V, long long vstride, const int* d_k, int k1, int batch, long long nn, c128* P,
               long long pstride, void* ws, size_t ws_bytes, hipStream_t s) {
  const long long per = gram_per_block(nn);
  const int nblk = static_cast<int>((nn + per - 1) / per);
  if (ws_bytes < static_cast<size_t>(nblk) * kGramTile * kGramTile * 16 * batch)
    return fail(RPM_E_INVALID, "singlet gram workspace too small");
  c128* partial = static_cast<c128*>(ws);
  hipLaunchKernelGGL(k_gram, dim3(nblk, batch), dim3(256), 0, s, V, vstride, d_k, k1, nn, per, partial);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gram_sum, dim3(kGramTile * kGramTile / 256, batch), dim3(256), 0, s, partial, nblk, d_k, k1, P,
                     pstride);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gram_sym, dim3(kGramTile * kGramTile / 256, batch), dim3(256), 0, s, d_k, k1, P, pstride);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}
}  // namespace

extern "C" int rpm_singlet_gram(const void* d_basis, int k, long long nn, void* d_P, void* d_ws, size_t ws_bytes,
                                void* stream) {
  if (!d_basis || !d_P || !d_ws || k < 1 || nn < 1) return fail(RPM_E_INVALID, "rpm_singlet_gram: bad arguments");
  if (k > kGramTile) return fail(RPM_E_UNSUPPORTED, "singlet projection above dimension %d", kGramTile);
  if (ws_bytes < rpm_singlet_gram_workspace_bytes(k, nn)) return fail(RPM_E_INVALID, "rpm_singlet_gram: workspace");
  return gram_batch(static_cast<const c128*>(d_basis), 0, nullptr, k, 1, nn, static_cast<c128*>(d_P), 0, d_ws, ws_bytes,
                    static_cast<hipStream_t>(stream));
}

extern "C" int rpm_singlet_gram_batch(const void* d_basis, int m, const int* d_k, int batch, long long nn, void* d_P,
                                      void* d_ws, size_t ws_bytes, void* stream) {
  if (!d_basis || !d_k || !d_P || !d_ws || batch < 1 || batch > 65535 || nn < 1 || m < 1 || m > kGramTile)
    return fail(RPM_E_INVALID, "rpm_singlet_gram_batch: bad arguments");
  return gram_batch(static_cast<const c128*>(d_basis), static_cast<long long>(m) * 4 * nn, d_k, m, batch, nn,
                    static_cast<c128*>(d_P), static_cast<long long>(m) * m, d_ws, ws_bytes,
                    static_cast<hipStream_t>(stream));
}
