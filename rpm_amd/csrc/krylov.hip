// Small-matrix side of the Krylov propagator, on the device:
//   * symmetric tridiagonal eigendecomposition T = Z diag(w) Z^T by the
//     implicit QL algorithm with Wilkinson-type shifts (textbook method, e.g.
//     Wilkinson & Reinsch, Handbook for Automatic Computation II, tql2), one
//     wavefront per matrix: lane i owns d[i], e[i] and row i of Z;
//   * c(t) = Z (exp(-i w t) * Z[0,:]) -- the projected exponential of
//     R/rpm_statevector_engine.py:569-609;
//   * embedded error estimates of the candidate ladder
//     (R/rpm_statevector_engine.py:690-715) and the per-output-time singlet
//     values and point errors of an accepted step (861-902).
#include <cmath>

#include "rpm_common.hpp"

namespace rpm {

constexpr int kMaxKrylov = 64;
constexpr int kMaxCand = 16;

struct Slot {          // one decomposition in scratch
  double w[kMaxKrylov];                    // eigenvalues
  double zt[kMaxKrylov * kMaxKrylov];      // Zt[k*64 + i] = Z[i][k]
  c128 c[kMaxKrylov];                      // c(t) at the requested time
};

struct CandArgs {
  int ncand;
  int dims[kMaxCand];
};

__device__ inline double lane_get(double v, int i) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), i);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), i);
  return __hiloint2double(hi, lo);
}

// Eigendecomposition of the leading n x n block (n <= 64) of the tridiagonal
// (alpha, beta).  Must be called by one full wavefront.  zrow: LDS row storage
// for this wave, zrow[i * 65 + k] = Z[i][k].  Returns w in lane i's `wout`.
__device__ bool ql_eig(int n, const double* __restrict__ alpha, const double* __restrict__ beta,
                       double* zrow, double& wout) {
  const int lane = threadIdx.x % kWave;
  double d = lane < n ? alpha[lane] : 0.0;
  double e = lane < n - 1 ? beta[lane] : 0.0;
  for (int k = 0; k < n; ++k) zrow[lane * 65 + k] = (lane == k) ? 1.0 : 0.0;
  bool ok = true;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = fabs(lane_get(d, m)) + fabs(lane_get(d, m + 1));
        if (fabs(lane_get(e, m)) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 80) {
        ok = false;
        break;
      }
      const double dl = lane_get(d, l), el = lane_get(e, l);
      double g = (lane_get(d, l + 1) - dl) / (2.0 * el);
      double r = sqrt(g * g + 1.0);
      g = lane_get(d, m) - dl + el / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        const double ei = lane_get(e, i);
        const double f = s * ei, b = c * ei;
        r = sqrt(f * f + g * g);
        if (lane == i + 1) e = r;
        if (r == 0.0) {
          const double dn = lane_get(d, i + 1) - p;
          if (lane == i + 1) d = dn;
          if (lane == m) e = 0.0;
          deflated = true;
          break;
        }
        const double inv = 1.0 / r;
        s = f * inv;
        c = g * inv;
        g = lane_get(d, i + 1) - p;
        r = (lane_get(d, i) - g) * s + 2.0 * c * b;
        p = s * r;
        if (lane == i + 1) d = g + p;
        g = c * r - b;
        if (lane < n) {
          const double zi = zrow[lane * 65 + i], zj = zrow[lane * 65 + i + 1];
          zrow[lane * 65 + i + 1] = s * zi + c * zj;
          zrow[lane * 65 + i] = c * zi - s * zj;
        }
      }
      if (deflated) continue;
      if (lane == l) {
        d -= p;
        e = g;
      }
      if (lane == m) e = 0.0;
    }
    if (!ok) break;
  }
  wout = d;
  return ok;
}

// One block (one wave) per (candidate, high/low).  Writes the decomposition
// and c(t) into slot[cand*2 + which].
__global__ void __launch_bounds__(64) k_tridiag_expm(const double* __restrict__ alpha, const double* __restrict__ beta,
                                                     const int* __restrict__ info, CandArgs args, int offset, double t,
                                                     Slot* __restrict__ slots, int* __restrict__ status) {
  __shared__ double zrow[kMaxKrylov * 65];
  __shared__ c128 u[kMaxKrylov];
  const int cand = blockIdx.x, which = blockIdx.y;
  const int lane = threadIdx.x;
  const int size = info ? info[0] : args.dims[cand];
  const int dim = min(args.dims[cand], size);
  const int low = max(1, dim - offset);
  if (which == 1 && low == dim) return;
  const int n = which == 0 ? dim : low;
  if (n < 1) return;
  double w = 0.0;
  const bool ok = ql_eig(n, alpha, beta, zrow, w);
  if (!ok && lane == 0) atomicOr(status, 1);
  __syncthreads();   // rows of Z written by other lanes are read below
  Slot* S = slots + cand * 2 + which;
  if (lane < n) {
    S->w[lane] = w;
    for (int k = 0; k < n; ++k) S->zt[k * kMaxKrylov + lane] = zrow[lane * 65 + k];
    double sn, cs;
    sincos(w * t, &sn, &cs);
    const double z0 = zrow[0 * 65 + lane];   // Z[0][lane]
    u[lane] = cmake(z0 * cs, -z0 * sn);
  }
  __syncthreads();
  if (lane < n) {
    c128 acc{0.0, 0.0};
    for (int k = 0; k < n; ++k) {
      const double z = zrow[lane * 65 + k];
      acc.re = fma(z, u[k].re, acc.re);
      acc.im = fma(z, u[k].im, acc.im);
    }
    S->c[lane] = acc;
  }
}

__global__ void __launch_bounds__(64) k_cand_err(const int* __restrict__ info, const double* __restrict__ norm,
                                                 CandArgs args, int offset, double safety, const Slot* __restrict__ slots,
                                                 int m_stride, double* __restrict__ err, c128* __restrict__ coeffs) {
  const int cand = blockIdx.x, lane = threadIdx.x;
  const int size = info[0];
  const int dim = min(args.dims[cand], size);
  const int low = max(1, dim - offset);
  const Slot* H = slots + cand * 2;
  const Slot* L = slots + cand * 2 + 1;
  double s2 = 0.0;
  for (int i = lane; i < m_stride; i += kWave) {
    const c128 ch = i < dim ? H->c[i] : cmake(0.0, 0.0);
    coeffs[static_cast<long long>(cand) * m_stride + i] = ch;
    if (low != dim && i < dim) {
      c128 df = ch;
      if (i < low) df = csub(df, L->c[i]);
      s2 = fma(df.re, df.re, s2);
      s2 = fma(df.im, df.im, s2);
    }
  }
  s2 = wave_sum(s2);
  if (lane == 0) err[cand] = (low == dim) ? 0.0 : safety * norm[0] * sqrt(s2);
}

// One block (one wave) per output time.
__global__ void __launch_bounds__(64) k_outputs(const double* __restrict__ norm, const c128* __restrict__ P, int sel,
                                                int low, const double* __restrict__ offsets, double safety,
                                                const Slot* __restrict__ hs, const Slot* __restrict__ ls,
                                                double* __restrict__ values, double* __restrict__ point_err,
                                                c128* __restrict__ coeffs_out) {
  __shared__ c128 uh[kMaxKrylov], ul[kMaxKrylov], ch[kMaxKrylov];
  const int lane = threadIdx.x;
  const double t = offsets[blockIdx.x];
  if (lane < sel) {
    double sn, cs;
    sincos(hs->w[lane] * t, &sn, &cs);
    const double z0 = hs->zt[lane * kMaxKrylov + 0];
    uh[lane] = cmake(z0 * cs, -z0 * sn);
  }
  if (low > 0 && lane < low) {
    double sn, cs;
    sincos(ls->w[lane] * t, &sn, &cs);
    const double z0 = ls->zt[lane * kMaxKrylov + 0];
    ul[lane] = cmake(z0 * cs, -z0 * sn);
  }
  __syncthreads();
  c128 c{0.0, 0.0}, cl{0.0, 0.0};
  if (lane < sel) {
    for (int k = 0; k < sel; ++k) {
      const double z = hs->zt[k * kMaxKrylov + lane];
      c.re = fma(z, uh[k].re, c.re);
      c.im = fma(z, uh[k].im, c.im);
    }
    ch[lane] = c;
    if (coeffs_out) coeffs_out[static_cast<long long>(blockIdx.x) * sel + lane] = c;
  }
  if (low > 0 && lane < low) {
    for (int k = 0; k < low; ++k) {
      const double z = ls->zt[k * kMaxKrylov + lane];
      cl.re = fma(z, ul[k].re, cl.re);
      cl.im = fma(z, ul[k].im, cl.im);
    }
  }
  __syncthreads();
  // (P c)_i = sum_j P[i][j] c_j = sum_j conj(P[j][i]) c_j   (P Hermitian)
  double q = 0.0, e2 = 0.0;
  if (lane < sel) {
    c128 pc{0.0, 0.0};
    for (int j = 0; j < sel; ++j) pc = cfmaconj(P[j * sel + lane], ch[j], pc);
    // Re(conj(c_i) (Pc)_i)
    q = c.re * pc.re + c.im * pc.im;
    if (low > 0) {
      c128 df = c;
      if (lane < low) df = csub(df, cl);
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

}  // namespace rpm

using namespace rpm;

extern "C" size_t rpm_krylov_scratch_bytes(int ncand, int m_stride) {
  (void)m_stride;
  const int n = ncand > 0 ? ncand : 1;
  return static_cast<size_t>(2 * n) * sizeof(Slot) + 256;
}

extern "C" int rpm_krylov_candidates(const double* d_alpha, const double* d_beta, const int* d_info,
                                     const double* d_norm, int ncand, const int* dims, int offset, double safety,
                                     double t, int m_stride, double* d_err, void* d_coeffs, void* d_scratch,
                                     void* stream) {
  if (!d_alpha || !d_beta || !d_info || !d_norm || !dims || !d_err || !d_coeffs || !d_scratch)
    return fail(RPM_E_INVALID, "rpm_krylov_candidates: NULL argument");
  if (ncand < 1 || ncand > kMaxCand) return fail(RPM_E_INVALID, "rpm_krylov_candidates: 1..%d candidates", kMaxCand);
  if (m_stride < 1 || m_stride > kMaxKrylov)
    return fail(RPM_E_UNSUPPORTED, "Krylov dimension above %d is not supported on the device", kMaxKrylov);
  CandArgs a{};
  a.ncand = ncand;
  for (int i = 0; i < ncand; ++i) {
    if (dims[i] < 1 || dims[i] > m_stride) return fail(RPM_E_INVALID, "candidate dimension %d out of range", dims[i]);
    a.dims[i] = dims[i];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  Slot* slots = static_cast<Slot*>(d_scratch);
  int* status = reinterpret_cast<int*>(static_cast<char*>(d_scratch) + 2 * ncand * sizeof(Slot));
  hipLaunchKernelGGL(k_tridiag_expm, dim3(ncand, 2), dim3(64), 0, s, d_alpha, d_beta, d_info, a, offset, t, slots,
                     status);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cand_err, dim3(ncand), dim3(64), 0, s, d_info, d_norm, a, offset, safety, slots, m_stride,
                     d_err, static_cast<c128*>(d_coeffs));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

extern "C" int rpm_krylov_outputs(const double* d_alpha, const double* d_beta, const double* d_norm, const void* d_P,
                                  int sel, int low, int nt, const double* d_offsets, double safety, double* d_values,
                                  double* d_point_err, void* d_coeffs_out, void* d_scratch, size_t scratch_bytes,
                                  void* stream) {
  if (!d_alpha || !d_beta || !d_norm || !d_P || !d_offsets || !d_values || !d_point_err || !d_scratch)
    return fail(RPM_E_INVALID, "rpm_krylov_outputs: NULL argument");
  if (sel < 1 || sel > kMaxKrylov || low < 0 || (low > 0 && low >= sel) || nt < 0)
    return fail(RPM_E_INVALID, "rpm_krylov_outputs: sel=%d low=%d", sel, low);
  if (scratch_bytes < rpm_krylov_scratch_bytes(1, sel)) return fail(RPM_E_INVALID, "rpm_krylov_outputs: scratch");
  if (nt == 0) return RPM_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Slot* slots = static_cast<Slot*>(d_scratch);
  int* status = reinterpret_cast<int*>(static_cast<char*>(d_scratch) + 2 * sizeof(Slot));
  // decompose T_sel (slot 0) and T_low (slot 1): a one-candidate ladder with offset sel-low
  CandArgs a{};
  a.ncand = 1;
  a.dims[0] = sel;
  const int offset = low > 0 ? sel - low : 0;
  hipLaunchKernelGGL(k_tridiag_expm, dim3(1, low > 0 ? 2 : 1), dim3(64), 0, s, d_alpha, d_beta, nullptr, a, offset,
                     0.0, slots, status);
  RPM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_outputs, dim3(nt), dim3(64), 0, s, d_norm, static_cast<const c128*>(d_P), sel, low, d_offsets,
                     safety, slots, slots + 1, d_values, d_point_err, static_cast<c128*>(d_coeffs_out));
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}
