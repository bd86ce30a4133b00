// Small-matrix side of the Krylov propagator, on the device.
//
// The reference evaluates the projected exponential c(t) = exp(-i T t) e_1 of
// the Lanczos tridiagonal T by a full eigendecomposition
// (R/rpm_statevector_engine.py:569-609).  Here c(t) is evaluated directly by
// the Chebyshev expansion of the exponential (Tal-Ezer & Kosloff, J. Chem.
// Phys. 81, 3967 (1984)):
//   T = c0 + h X with spec(X) in [-1, 1] (Gershgorin bounds),
//   exp(-i T t) v = exp(-i c0 t) sum_k a_k T_k(X) v,  a_0 = J_0(ht),
//   a_k = 2 (-i)^k J_k(ht),
// with the Bessel values from Miller's backward recurrence normalised by
// J_0 + 2 sum J_2k = 1.  T is real, so T_k(X) e_1 is a real vector and the
// recurrence v_{k+1} = 2 X v_k - v_{k-1} costs a few lane operations per term:
// lane j owns component j and the tridiagonal neighbours come from DPP wave
// shifts.  The series stops where the Bessel weights are below 1e-18
// (k > 1.5 ht + 40); for ht beyond kSplitZ the time is split into equal
// sub-steps applied in sequence.  One wavefront evaluates one
// (dimension, time).
//
// Kernels: embedded error estimates of a candidate ladder
// (R/rpm_statevector_engine.py:690-715) and the per-output-time singlet values
// and point errors of an accepted step (861-902).
#include <cmath>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kMaxKrylov = 64;
constexpr int kMaxCand = 16;
constexpr double kSplitZ = 1300.0;                               // largest ht per sub-step
constexpr int kMaxTerms = static_cast<int>(1.5 * kSplitZ) + 41;  // Bessel terms per sub-step

struct KrylovTask {
  int b;       // trajectory
  int dim;     // Krylov dimension of the evaluation (<= that trajectory's basis size)
  int low;     // embedded lower dimension (0: none)
  int out;     // output slot
  double t;    // time offset
  int kind;    // 0 candidate, 1 output value
  int pad;
};
static_assert(sizeof(KrylovTask) == 32, "task layout is part of the C ABI");

struct CandArgs {
  int ncand;
  int dims[kMaxCand];
};

// lane i <- lane i-1 (0 for lane 0)
__device__ inline double shift_up(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
// lane i <- lane i+1 (0 for lane 63)
__device__ inline double shift_down(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

__device__ inline double wave_min(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// Bessel J_0..J_K(z), z > 0, into bess[0..K] by Miller's backward recurrence (one lane).
__device__ void bessel_series(double z, int K, double* bess) {
  const int N = K + 60 + static_cast<int>(sqrt(40.0 * (K + 1)));
  const double two_over_z = 2.0 / z;
  double jp1 = 0.0, j = 1e-300, norm = 0.0;
  for (int k = N; k >= 1; --k) {
    const double jm1 = k * two_over_z * j - jp1;
    jp1 = j;
    j = jm1;                                  // j = J_{k-1} (unnormalised)
    if (k - 1 <= K) bess[k - 1] = j;
    if (((k - 1) & 1) == 0 && k - 1 > 0) norm += 2.0 * j;
    if (fabs(j) > 1e250) {                    // rescale everything kept so far
      j *= 1e-250;
      jp1 *= 1e-250;
      norm *= 1e-250;
      for (int i = k - 1; i <= K; ++i) bess[i] *= 1e-250;
    }
  }
  norm += j;                                  // + J_0
  const double inv = 1.0 / norm;
  for (int i = 0; i <= K; ++i) bess[i] *= inv;
}

// One real Chebyshev recurrence: returns sum_k w_k T_k(X) v0 split into the
// (-i)^k classes: re = sum_{k even} +-w_k T_k v0, im = sum_{k odd} -+w_k T_k v0.
__device__ inline void cheb_real(double v0, double dx, double ox, double oxp, const double* bess, int K,
                                 double& re, double& im) {
  double vprev = v0;
  double vcur = dx * vprev + oxp * shift_up(vprev) + ox * shift_down(vprev);
  re = bess[0] * vprev;
  im = -2.0 * bess[1] * vcur;
  for (int k = 2; k <= K; ++k) {
    const double xv = dx * vcur + oxp * shift_up(vcur) + ox * shift_down(vcur);
    const double vnext = 2.0 * xv - vprev;
    vprev = vcur;
    vcur = vnext;
    const double w = 2.0 * bess[k];
    switch (k & 3) {                                          // (-i)^k
      case 0: re += w * vcur; break;
      case 1: im -= w * vcur; break;
      case 2: re -= w * vcur; break;
      default: im += w * vcur; break;
    }
  }
}

// c = exp(-i T_n t) e_1 for the leading n x n block of (alpha, beta), t >= 0;
// lane j returns component j (0 for j >= n).  One full wavefront; bess: LDS
// scratch of kMaxTerms + 1 doubles private to the wave.
__device__ c128 cheb_expm_e1(int n, const double* __restrict__ alpha, const double* __restrict__ beta, double t,
                             double* bess) {
  const int lane = threadIdx.x % kWave;
  const double a = lane < n ? alpha[lane] : 0.0;
  const double b = lane < n - 1 ? beta[lane] : 0.0;          // coupling (lane, lane+1)
  const double bp = shift_up(b);                              // coupling (lane-1, lane)
  const double lo = wave_min(lane < n ? a - fabs(bp) - fabs(b) : 1e300);
  const double hi = wave_max(lane < n ? a + fabs(bp) + fabs(b) : -1e300);
  const double c0 = 0.5 * (lo + hi);
  double h = 0.5 * (hi - lo);
  // Every wave of the block reaches the one barrier below, also in the trivial case.
  const bool trivial = !(h > 0.0) || t == 0.0;
  h = trivial ? 1.0 : h * (1.0 + 1e-12);
  const double ih = 1.0 / h;
  const double dx = lane < n ? (a - c0) * ih : 0.0;
  const double ox = b * ih, oxp = bp * ih;
  const double ztot = h * t;
  const int nsub = ztot > kSplitZ ? static_cast<int>(ceil(ztot / kSplitZ)) : 1;
  const double tsub = t / nsub, z = h * tsub;
  const int K = static_cast<int>(1.5 * z) + 40;
  if (!trivial && lane == 0) bessel_series(z, K, bess);
  __syncthreads();
  if (trivial) {
    double sn, cs;
    sincos(c0 * t, &sn, &cs);
    return lane == 0 ? cmake(cs, -sn) : cmake(0.0, 0.0);
  }
  double re, im;
  cheb_real(lane == 0 ? 1.0 : 0.0, dx, ox, oxp, bess, K, re, im);
  for (int s = 1; s < nsub; ++s) {
    // apply the same series to the complex vector (re + i im): X is real
    double rr, ri, ir, ii;
    cheb_real(re, dx, ox, oxp, bess, K, rr, ri);
    cheb_real(im, dx, ox, oxp, bess, K, ir, ii);
    re = rr - ii;
    im = ri + ir;
  }
  if (lane >= n) return cmake(0.0, 0.0);
  double sn, cs;
  sincos(c0 * t, &sn, &cs);                                   // exp(-i c0 t) = cs - i sn
  return cmake(cs * re + sn * im, cs * im - sn * re);
}

// One block (two wavefronts: the candidate dimension and its embedded low dimension) per candidate.
__global__ void __launch_bounds__(128) k_candidates(const double* __restrict__ alpha, const double* __restrict__ beta,
                                                    const int* __restrict__ info, const double* __restrict__ norm,
                                                    CandArgs args, int offset, double safety, double t, int m_stride,
                                                    double* __restrict__ err, c128* __restrict__ coeffs) {
  __shared__ double bess[2][kMaxTerms + 1];
  __shared__ c128 cl[kMaxKrylov];
  const int cand = blockIdx.x;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int size = info[0];
  const int dim = min(args.dims[cand], size);
  const int low = max(1, dim - offset);
  const c128 c = cheb_expm_e1(wave == 0 ? dim : low, alpha, beta, t, bess[wave]);
  if (wave == 1) cl[lane] = c;
  __syncthreads();
  if (wave != 0) return;
  for (int i = lane; i < m_stride; i += kWave)
    coeffs[static_cast<long long>(cand) * m_stride + i] = (i < dim) ? c : cmake(0.0, 0.0);
  double s2 = 0.0;
  if (low != dim && lane < dim) {
    c128 df = c;
    if (lane < low) df = csub(df, cl[lane]);
    s2 = df.re * df.re + df.im * df.im;
  }
  s2 = wave_sum(s2);
  if (lane == 0) err[cand] = (low == dim) ? 0.0 : safety * norm[0] * sqrt(s2);
}

// One block (two wavefronts: sel and low) per output time.
__global__ void __launch_bounds__(128) k_outputs(const double* __restrict__ alpha, const double* __restrict__ beta,
                                                 const double* __restrict__ norm, const c128* __restrict__ P, int sel,
                                                 int low, const double* __restrict__ offsets, double safety,
                                                 double* __restrict__ values, double* __restrict__ point_err,
                                                 c128* __restrict__ coeffs_out) {
  __shared__ double bess[2][kMaxTerms + 1];
  __shared__ c128 ch[kMaxKrylov], cl[kMaxKrylov];
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const double t = offsets[blockIdx.x];
  const c128 c = cheb_expm_e1(wave == 0 ? sel : (low > 0 ? low : 1), alpha, beta, t, bess[wave]);
  if (wave == 0) ch[lane] = c;
  else cl[lane] = c;
  __syncthreads();
  if (wave != 0) return;
  double q = 0.0, e2 = 0.0;
  if (lane < sel) {
    if (coeffs_out) coeffs_out[static_cast<long long>(blockIdx.x) * sel + lane] = c;
    // (P c)_i = sum_j P[i][j] c_j = sum_j conj(P[j][i]) c_j   (P Hermitian)
    c128 pc{0.0, 0.0};
    for (int j = 0; j < sel; ++j) pc = cfmaconj(P[j * sel + lane], ch[j], pc);
    q = c.re * pc.re + c.im * pc.im;
    if (low > 0) {
      c128 df = c;
      if (lane < low) df = csub(df, cl[lane]);
      e2 = df.re * df.re + df.im * df.im;
    }
  }
  q = wave_sum(q);
  e2 = wave_sum(e2);
  if (lane == 0) {
    const double nr = norm[0];
    values[blockIdx.x] = nr * nr * q;
    point_err[blockIdx.x] = low > 0 ? safety * nr * sqrt(e2) : 0.0;
  }
}

// Batched form of k_candidates / k_outputs: one block per task, tasks of
// many trajectories in one launch.  Trajectory b's tridiagonal is
// (alpha + b*ab_stride, beta + b*ab_stride), its norm norm[b] and its
// projected singlet matrix P + b*pstride (dim x dim, row-major).
//  kind 0 (candidate): v0[out] = safety*norm*||c_dim - pad(c_low)|| (0 if low >= dim),
//                      coeffs[out*m_stride + i] = c_dim padded with zeros;
//  kind 1 (output):    v0[out] = norm^2 Re(c^H P c), v1[out] = point error (0 if low == 0),
//                      coeffs row out written if coeffs is not NULL.
__global__ void __launch_bounds__(128) k_tasks(const KrylovTask* __restrict__ tasks, const double* __restrict__ alpha,
                                               const double* __restrict__ beta, int ab_stride,
                                               const double* __restrict__ norm, const c128* __restrict__ P,
                                               long long pstride, double safety, double* __restrict__ v0,
                                               double* __restrict__ v1, c128* __restrict__ coeffs, int m_stride) {
  __shared__ double bess[2][kMaxTerms + 1];
  __shared__ c128 ch[kMaxKrylov], cl[kMaxKrylov];
  KrylovTask tk = tasks[blockIdx.x];
  tk.dim = min(max(tk.dim, 1), ab_stride);   // stay inside the trajectory's tridiagonal whatever the task says
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const double* al = alpha + static_cast<long long>(tk.b) * ab_stride;
  const double* be = beta + static_cast<long long>(tk.b) * ab_stride;
  const bool has_low = tk.low > 0 && tk.low < tk.dim;
  const c128 c = cheb_expm_e1(wave == 0 ? tk.dim : (has_low ? tk.low : 1), al, be, tk.t, bess[wave]);
  if (wave == 0) ch[lane] = c;
  else cl[lane] = c;
  __syncthreads();
  if (wave != 0) return;
  const int dim = tk.dim;
  double q = 0.0, e2 = 0.0;
  if (lane < dim) {
    if (has_low) {
      c128 df = c;
      if (lane < tk.low) df = csub(df, cl[lane]);
      e2 = df.re * df.re + df.im * df.im;
    }
    if (tk.kind == 1 && P) {
      const c128* Pb = P + static_cast<long long>(tk.b) * pstride;
      c128 pc{0.0, 0.0};
      for (int j = 0; j < dim; ++j) pc = cfmaconj(Pb[j * dim + lane], ch[j], pc);
      q = c.re * pc.re + c.im * pc.im;
    }
  }
  if (coeffs)
    for (int i = lane; i < m_stride; i += kWave)
      coeffs[static_cast<long long>(tk.out) * m_stride + i] = i < dim ? ch[i] : cmake(0.0, 0.0);
  q = wave_sum(q);
  e2 = wave_sum(e2);
  if (lane == 0) {
    const double nr = norm[tk.b];
    const double pe = has_low ? safety * nr * sqrt(e2) : 0.0;
    if (tk.kind == 0) {
      v0[tk.out] = pe;
    } else {
      v0[tk.out] = nr * nr * q;
      if (v1) v1[tk.out] = pe;
    }
  }
}

// Singlet-yield accumulation.  Trajectory b's outputs of one call are rows row0 .. row0+count-1 of
// values / point_err and weight rows w0 .. w0+count-1 of the (fine, coarse) table; one block per
// trajectory, thread t taking outputs t, t+256, ... and a fixed-shape block sum, so the order of the
// additions depends only on the trajectory's own outputs (batched and single runs agree bitwise).
constexpr int kYieldThreads = 256;

__global__ void __launch_bounds__(kYieldThreads) k_yield_accumulate(const double* __restrict__ values,
                                                                  const double* __restrict__ point_err,
                                                                  long long nvalues,
                                                                  const long long* __restrict__ spans,
                                                                  const double* __restrict__ weights,
                                                                  long long nweights, const double* __restrict__ acc,
                                                                  double* __restrict__ phi) {
  __shared__ double sm[kYieldThreads / kWave];
  const int b = blockIdx.x;
  const long long row0 = spans[3 * b], w0 = spans[3 * b + 2];
  long long count = spans[3 * b + 1];
  // rows and weights outside the buffers contribute nothing (the host checks its own spans)
  if (row0 < 0 || w0 < 0) count = 0;
  count = min(count, min(nvalues - row0, nweights - w0));
  if (count <= 0) return;                       // uniform over the block
  const double a = acc ? acc[b] : 0.0;
  double f = 0.0, c = 0.0, e = 0.0;
  for (long long i = threadIdx.x; i < count; i += kYieldThreads) {
    const double p = values[row0 + i];
    const double wf = weights[2 * (w0 + i)], wc = weights[2 * (w0 + i) + 1];
    const double err = a + (point_err ? point_err[row0 + i] : 0.0);
    const double bound = fmin(1.0, 2.0 * err + err * err);  // state_error_to_population_bound
    f += wf * p;
    c += wc * p;
    e += wf * bound;
  }
  f = block_sum(f, sm);
  c = block_sum(c, sm);
  e = block_sum(e, sm);
  if (threadIdx.x == 0) {
    phi[3 * b] += f;
    phi[3 * b + 1] += c;
    phi[3 * b + 2] += e;
  }
}

}  // namespace rpm

using namespace rpm;

extern "C" size_t rpm_krylov_scratch_bytes(int ncand, int m_stride) {
  (void)ncand;
  (void)m_stride;
  return 256;
}

extern "C" int rpm_krylov_candidates(const double* d_alpha, const double* d_beta, const int* d_info,
                                     const double* d_norm, int ncand, const int* dims, int offset, double safety,
                                     double t, int m_stride, double* d_err, void* d_coeffs, void* d_scratch,
                                     void* stream) {
  (void)d_scratch;
  if (!d_alpha || !d_beta || !d_info || !d_norm || !dims || !d_err || !d_coeffs)
    return fail(RPM_E_INVALID, "rpm_krylov_candidates: NULL argument");
  if (ncand < 1 || ncand > kMaxCand) return fail(RPM_E_INVALID, "rpm_krylov_candidates: 1..%d candidates", kMaxCand);
  if (m_stride < 1 || m_stride > kMaxKrylov)
    return fail(RPM_E_UNSUPPORTED, "Krylov dimension above %d is not supported on the device", kMaxKrylov);
  if (!(t >= 0.0)) return fail(RPM_E_INVALID, "rpm_krylov_candidates: negative time");
  CandArgs a{};
  a.ncand = ncand;
  for (int i = 0; i < ncand; ++i) {
    if (dims[i] < 1 || dims[i] > m_stride) return fail(RPM_E_INVALID, "candidate dimension %d out of range", dims[i]);
    a.dims[i] = dims[i];
  }
  hipLaunchKernelGGL(k_candidates, dim3(ncand), dim3(128), 0, static_cast<hipStream_t>(stream), d_alpha, d_beta, d_info,
                     d_norm, a, offset, safety, t, m_stride, d_err, static_cast<c128*>(d_coeffs));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_krylov_outputs(const double* d_alpha, const double* d_beta, const double* d_norm, const void* d_P,
                                  int sel, int low, int nt, const double* d_offsets, double safety, double* d_values,
                                  double* d_point_err, void* d_coeffs_out, void* d_scratch, size_t scratch_bytes,
                                  void* stream) {
  (void)d_scratch;
  (void)scratch_bytes;
  if (!d_alpha || !d_beta || !d_norm || !d_P || !d_offsets || !d_values || !d_point_err)
    return fail(RPM_E_INVALID, "rpm_krylov_outputs: NULL argument");
  if (sel < 1 || sel > kMaxKrylov || low < 0 || (low > 0 && low >= sel) || nt < 0)
    return fail(RPM_E_INVALID, "rpm_krylov_outputs: sel=%d low=%d", sel, low);
  if (nt == 0) return RPM_OK;
  hipLaunchKernelGGL(k_outputs, dim3(nt), dim3(128), 0, static_cast<hipStream_t>(stream), d_alpha, d_beta, d_norm,
                     static_cast<const c128*>(d_P), sel, low, d_offsets, safety, d_values, d_point_err,
                     static_cast<c128*>(d_coeffs_out));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_krylov_tasks(const void* d_tasks, int ntasks, const double* d_alpha, const double* d_beta,
                                int ab_stride, const double* d_norm, const void* d_P, long long p_stride,
                                double safety, double* d_v0, double* d_v1, void* d_coeffs, int m_stride,
                                void* stream) {
  if (!d_tasks || !d_alpha || !d_beta || !d_norm || !d_v0 || ntasks < 0 || ab_stride < 1)
    return fail(RPM_E_INVALID, "rpm_krylov_tasks: bad arguments");
  if (m_stride < 1 || m_stride > kMaxKrylov || ab_stride > kMaxKrylov)
    return fail(RPM_E_UNSUPPORTED, "Krylov dimension above %d is not supported on the device", kMaxKrylov);
  if (d_P && p_stride < static_cast<long long>(ab_stride) * ab_stride)
    return fail(RPM_E_INVALID, "rpm_krylov_tasks: P stride %lld below %d^2", p_stride, ab_stride);
  if (ntasks == 0) return RPM_OK;
  hipLaunchKernelGGL(k_tasks, dim3(ntasks), dim3(128), 0, static_cast<hipStream_t>(stream),
                     static_cast<const KrylovTask*>(d_tasks), d_alpha, d_beta, ab_stride, d_norm,
                     static_cast<const c128*>(d_P), p_stride, safety, d_v0, d_v1, static_cast<c128*>(d_coeffs),
                     m_stride);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_yield_accumulate(const double* d_values, const double* d_point_err, long long nvalues,
                                    const long long* d_spans, int batch, const double* d_weights, long long nweights,
                                    const double* d_acc, double* d_phi, void* stream) {
  if (!d_values || !d_spans || !d_weights || !d_phi || batch < 1 || batch > 65535 || nvalues < 0 || nweights < 0)
    return fail(RPM_E_INVALID, "rpm_yield_accumulate: bad arguments");
  hipLaunchKernelGGL(k_yield_accumulate, dim3(batch), dim3(kYieldThreads), 0, static_cast<hipStream_t>(stream),
                     d_values, d_point_err, nvalues, d_spans, d_weights, nweights, d_acc, d_phi);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}
