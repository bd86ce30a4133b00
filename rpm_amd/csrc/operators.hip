// Hamiltonian action H·psi on MI355X.
//
// Two operator forms:
//  * pair stencil -- the compiled two-radical form (the production H of
//    R/rpm_statevector_engine.py:132-245).  H is rewritten as
//      diagonal(eA,eB,l,r) + electron flips + per-nucleus off-diagonal couplings
//    so every output amplitude is one fused gather: ~37 complex FMAs per
//    amplitude for the 221,184-dimensional FAD-TrpH pair, no intermediate
//    state passes (the reference makes two 16-axis transposes and two CSR
//    products per matvec).
//  * local terms -- any sum of LocalTerms on any axis order
//    (R/rpm_statevector_engine.py:95-108): one gather per term per amplitude.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rpm_common.hpp"
#include "pair_tile.hpp"

namespace rpm {

// ---------------------------------------------------------------- pair stencil

struct PairTables {
  int NL, NR, KL, KR;
  long long NLR;
  const c128* dA;       // [2][NL]  diagonal from left nuclei
  const c128* dB;       // [2][NR]
  const c128* fA;       // [2][NL]  eA flip from the left nuclei, index eA*NL + l
  const c128* fB;       // [2][NR]
  const c128* pair;     // [16]     the operator's own 4x4 electron matrix
  const int* offL;      // [NL][KL] flat offset of the k-th left neighbour ((l2-l)*NR)
  const int* offR;      // [NR][KR] r2 - r
  const c128* coefL;    // [NL][KL][2][2]  (eA, e')
  const c128* coefR;    // [NR][KR][2][2]  (eB, e')
};

struct PairOp {
  PairTables t;
  TileTables tile;      // valid when use_tile
  bool use_tile;
  int kernel;           // 0 auto, 1 gather, 2 register-tiled, 3 LDS-tiled, 4 LDS-tiled closed first axis
  void* dev;            // single device allocation backing the tables
};

__global__ void __launch_bounds__(256) k_pair_apply(PairTables T, const c128* __restrict__ in,
                                                    c128* __restrict__ out, long long ld, ApplyCtl A) {
  const long long b = blockIdx.y;
  double s = 1.0;
  if (A.ctl) {
    const BasisCtl& c = A.ctl[b];
    if (!basis_step_live(c, A.j)) return;
    s = c.scale;
  }
  const long long t0 = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = t0 < T.NLR;
  const long long t = valid ? t0 : 0;   // idle threads recompute amplitude 0 and store nothing
  PairCoef E;
  load_pair_coef(A.pairs ? A.pairs + b * 16 : T.pair, E);
  const c128* x = in + b * ld;
  c128* y = out + b * ld;
  const int l = static_cast<int>(t / T.NR);
  const int r = static_cast<int>(t - static_cast<long long>(l) * T.NR);
  const long long S = T.NLR;

  c128 p[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) p[e] = x[e * S + t];

  c128 acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int eA = e >> 1, eB = e & 1;
    const c128 d = cadd(cadd(T.dA[eA * T.NL + l], T.dB[eB * T.NR + r]), E.d[e]);
    acc[e] = cfma(d, p[e], cmake(0.0, 0.0));
    acc[e] = cfma(cadd(T.fA[eA * T.NL + l], E.fa[e]), p[e ^ 2], acc[e]);
    acc[e] = cfma(cadd(T.fB[eB * T.NR + r], E.fb[e]), p[e ^ 1], acc[e]);
    acc[e] = cfma(E.bo[e], p[e ^ 3], acc[e]);
  }

  // left nuclei: couple (eA, m_n) with (e', m'), eB and r unchanged
  for (int k = 0; k < T.KL; ++k) {
    const long long o = t + T.offL[l * T.KL + k];
    const c128* C = T.coefL + (static_cast<long long>(l) * T.KL + k) * 4;
    const c128 c00 = C[0], c01 = C[1], c10 = C[2], c11 = C[3];
    const c128 q0 = x[0 * S + o], q1 = x[1 * S + o], q2 = x[2 * S + o], q3 = x[3 * S + o];
    // e = eA*2+eB ; source e' = e'A*2 + eB
    acc[0] = cfma(c00, q0, acc[0]); acc[0] = cfma(c01, q2, acc[0]);
    acc[1] = cfma(c00, q1, acc[1]); acc[1] = cfma(c01, q3, acc[1]);
    acc[2] = cfma(c10, q0, acc[2]); acc[2] = cfma(c11, q2, acc[2]);
    acc[3] = cfma(c10, q1, acc[3]); acc[3] = cfma(c11, q3, acc[3]);
  }
  // right nuclei: couple (eB, m_n) with (e', m'), eA and l unchanged
  for (int k = 0; k < T.KR; ++k) {
    const long long o = t + T.offR[r * T.KR + k];
    const c128* C = T.coefR + (static_cast<long long>(r) * T.KR + k) * 4;
    const c128 c00 = C[0], c01 = C[1], c10 = C[2], c11 = C[3];
    const c128 q0 = x[0 * S + o], q1 = x[1 * S + o], q2 = x[2 * S + o], q3 = x[3 * S + o];
    acc[0] = cfma(c00, q0, acc[0]); acc[0] = cfma(c01, q1, acc[0]);
    acc[1] = cfma(c10, q0, acc[1]); acc[1] = cfma(c11, q1, acc[1]);
    acc[2] = cfma(c00, q2, acc[2]); acc[2] = cfma(c01, q3, acc[2]);
    acc[3] = cfma(c10, q2, acc[3]); acc[3] = cfma(c11, q3, acc[3]);
  }

  const double bprev = A.prev ? A.ctl[b].beta_prev : 0.0;
  c128 dot{0.0, 0.0};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    c128 v = apply_epilogue(A, acc[e], p[e], s, b, ld, e * S + t);
    if (A.dots && valid) dot = cfmaconj(cscale(p[e], s), v, dot);
    if (A.prev) {
      const c128 q = A.prev[b * A.prev_ld + e * S + t];
      v.re = fma(-bprev, q.re, v.re);
      v.im = fma(-bprev, q.im, v.im);
    }
    if (valid) y[e * S + t] = v;
  }
  if (A.copy && valid) {
    c128* z = A.copy + b * A.copy_ld;
#pragma unroll
    for (int e = 0; e < 4; ++e) z[e * S + t] = cscale(p[e], s);
  }
  if (A.dots) {
    __shared__ c128 red[256 / kWave];
    block_dot_store(dot, red, A.dots + b * A.dots_ld + blockIdx.x);
  }
}

// Host-side compilation of the per-nucleus blocks into the gather tables.
static int build_pair(int nL, const int* ldims, int nR, const int* rdims, const double* blocks,
                      const double* pair, PairOp** out) {
  if (nL < 0 || nR < 0 || nL + nR > 40) return fail(RPM_E_INVALID, "bad nucleus counts");
  long long NL = 1, NR = 1;
  int KL = 0, KR = 0;
  for (int i = 0; i < nL; ++i) {
    if (ldims[i] < 1 || ldims[i] > 16) return fail(RPM_E_INVALID, "left dim %d out of range", ldims[i]);
    NL *= ldims[i];
    KL += ldims[i] - 1;
  }
  for (int i = 0; i < nR; ++i) {
    if (rdims[i] < 1 || rdims[i] > 16) return fail(RPM_E_INVALID, "right dim %d out of range", rdims[i]);
    NR *= rdims[i];
    KR += rdims[i] - 1;
  }
  if (NL * NR * 4 > (1LL << 31)) return fail(RPM_E_INVALID, "state too large");
  // block pointers
  std::vector<const c128*> blk(nL + nR);
  {
    const c128* p = reinterpret_cast<const c128*>(blocks);
    for (int i = 0; i < nL + nR; ++i) {
      blk[i] = p;
      const int d = i < nL ? ldims[i] : rdims[i - nL];
      p += 4 * d * d;
    }
  }
  const c128* P = reinterpret_cast<const c128*>(pair);

  // one side's tables: digits by C-order over dims
  auto side = [&](int n, const int* dims, long long N, int K, int first_block, long long unit,
                  std::vector<c128>& dtab, std::vector<c128>& ftab, std::vector<int>& off,
                  std::vector<c128>& coef) {
    dtab.assign(2 * N, c128{0, 0});
    ftab.assign(2 * N, c128{0, 0});   // [e][idx]: e -> 1-e, same nuclear index
    off.assign(N * K > 0 ? N * K : 1, 0);
    coef.assign(N * K > 0 ? N * K * 4 : 4, c128{0, 0});
    std::vector<long long> stride(n);
    long long s = 1;
    for (int i = n - 1; i >= 0; --i) { stride[i] = s; s *= dims[i]; }
    for (long long idx = 0; idx < N; ++idx) {
      int k = 0;
      for (int i = 0; i < n; ++i) {
        const int d = dims[i];
        const int m = static_cast<int>((idx / stride[i]) % d);
        const c128* B = blk[first_block + i];       // (2d x 2d), row = e*d + m
        const int D = 2 * d;
        for (int e = 0; e < 2; ++e) {
          dtab[e * N + idx] = cadd(dtab[e * N + idx], B[(e * d + m) * D + (e * d + m)]);
          const c128 f = B[(e * d + m) * D + ((1 - e) * d + m)];
          ftab[e * N + idx] = cadd(ftab[e * N + idx], f);
        }
        for (int m2 = 0; m2 < d; ++m2) {
          if (m2 == m) continue;
          off[idx * K + k] = static_cast<int>((m2 - m) * stride[i] * unit);
          for (int e = 0; e < 2; ++e)
            for (int e2 = 0; e2 < 2; ++e2)
              coef[(idx * K + k) * 4 + e * 2 + e2] = B[(e * d + m) * D + (e2 * d + m2)];
          ++k;
        }
      }
    }
  };

  std::vector<c128> dA, dB;
  std::vector<c128> fA, fB, coefL, coefR;
  std::vector<int> offL, offR;
  side(nL, ldims, NL, KL, 0, NR, dA, fA, offL, coefL);
  side(nR, rdims, NR, KR, nL, 1, dB, fB, offR, coefR);

  std::vector<c128> pairv(P, P + 16);

  // ---- tiled-kernel tables: the last NC = 2 left axes (spin-1/2) are closed in the tile
  const int NC = 2;
  bool tile_ok = nL >= NC;
  for (int j = 0; j < NC && tile_ok; ++j) tile_ok = ldims[nL - 1 - j] == 2;
  int KO = 0;
  for (int i = 0; i < nL - NC && tile_ok; ++i) KO += ldims[i] - 1;
  tile_ok = tile_ok && pair_tile_supported(static_cast<int>(NR), NC, KO, KR);
  const int LG = 1 << NC;
  const long long ngroups = tile_ok ? NL / LG : 0;
  std::vector<unsigned char> flagR(KR > 0 ? KR : 1, 0), flagC(NC, 0);
  std::vector<c128> coefC(NC * 2 * 4);
  std::vector<int> ofsO(ngroups * KO > 0 ? ngroups * KO : 1, 0);
  std::vector<c128> coefO(ngroups * KO > 0 ? ngroups * KO * 4 : 4);
  std::vector<unsigned> maskO(ngroups > 0 ? ngroups : 1, 0u);
  std::vector<unsigned> nzO(ngroups > 0 ? ngroups : 1, 0u), hselO(ngroups > 0 ? ngroups : 1, 0u);
  unsigned bothO = 0;   // bit k: some group's open slot k reads both input electron halves
  unsigned halfO = 0, nzCt = 0;
  // one nibble per open slot in a 32-bit word (nzO) and per closed (axis, digit) pair (nzC)
  if (KO > 8 || KR > 16 || NC > 4) tile_ok = false;
  auto nz = [](c128 c) { return c.re != 0.0 || c.im != 0.0; };
  auto flags4 = [&](const c128* C) {
    int f = 0;
    if (nz(C[0]) || nz(C[3])) f |= kSameE;
    if (nz(C[1]) || nz(C[2])) f |= kFlipE;
    return static_cast<unsigned char>(f);
  };
  std::vector<int> packR(tile_ok ? NR * KR : 1, 0);
  std::vector<c128> coefT(tile_ok ? KR * 3 * 4 : 4, c128{0, 0});
  if (tile_ok) {
    for (long long r = 0; r < NR; ++r)
      for (int k = 0; k < KR; ++k) flagR[k] |= flags4(&coefR[(r * KR + k) * 4]);
    // right neighbours by (slot k, digit m): slot order follows the right nuclei, (d-1) slots each
    std::vector<long long> rstride(nR);
    {
      long long s = 1;
      for (int i = nR - 1; i >= 0; --i) { rstride[i] = s; s *= rdims[i]; }
    }
    int k0 = 0;
    for (int i = 0; i < nR; ++i) {
      const int d = rdims[i];
      const c128* B = blk[nL + i];
      for (int m = 0; m < d; ++m) {
        int slot = 0;
        for (int m2 = 0; m2 < d; ++m2) {
          if (m2 == m) continue;
          for (int e = 0; e < 2; ++e)
            for (int e2 = 0; e2 < 2; ++e2)
              coefT[((k0 + slot) * 3 + m) * 4 + e * 2 + e2] = B[(e * d + m) * (2 * d) + (e2 * d + m2)];
          ++slot;
        }
      }
      for (long long r = 0; r < NR; ++r) {
        const int m = static_cast<int>((r / rstride[i]) % d);
        int slot = 0;
        for (int m2 = 0; m2 < d; ++m2) {
          if (m2 == m) continue;
          const int off = static_cast<int>((m2 - m) * rstride[i]);
          packR[r * KR + k0 + slot] = (off & 0xffff) | (m << 16);
          ++slot;
        }
      }
      k0 += d - 1;
    }
    for (int j = 0; j < NC; ++j) {
      const c128* B = blk[nL - 1 - j];      // d = 2, row = e*2 + m
      for (int m = 0; m < 2; ++m) {
        for (int e = 0; e < 2; ++e)
          for (int e2 = 0; e2 < 2; ++e2) coefC[(j * 2 + m) * 4 + e * 2 + e2] = B[(e * 2 + m) * 4 + (e2 * 2 + (1 - m))];
        flagC[j] |= flags4(&coefC[(j * 2 + m) * 4]);
      }
    }
    std::vector<long long> lstride(nL);
    {
      long long s = 1;
      for (int i = nL - 1; i >= 0; --i) { lstride[i] = s; s *= ldims[i]; }
    }
    for (long long g = 0; g < ngroups; ++g) {
      const long long l0 = g * LG;
      int k = 0;
      for (int i = 0; i < nL - NC; ++i) {
        const int d = ldims[i];
        const int m = static_cast<int>((l0 / lstride[i]) % d);
        const c128* B = blk[i];
        for (int m2 = 0; m2 < d; ++m2) {
          if (m2 == m) continue;
          ofsO[g * KO + k] = static_cast<int>((m2 - m) * lstride[i]);
          c128* C = &coefO[(g * KO + k) * 4];
          for (int e = 0; e < 2; ++e)
            for (int e2 = 0; e2 < 2; ++e2) C[e * 2 + e2] = B[(e * d + m) * (2 * d) + (e2 * d + m2)];
          maskO[g] |= static_cast<unsigned>(flags4(C)) << (2 * k);
          unsigned bits = 0;
          for (int q = 0; q < 4; ++q) bits |= nz(C[q]) ? (1u << q) : 0u;
          nzO[g] |= bits << (4 * k);
          if ((bits & 10u) != 0 && (bits & 5u) == 0) hselO[g] |= 1u << k;
          if ((bits & 5u) != 0 && (bits & 10u) != 0) bothO |= 1u << k;
          ++k;
        }
      }
    }
  }

  if (tile_ok) {
    for (int k = 0; k < KO; ++k)
      if (!((bothO >> k) & 1u)) halfO |= 1u << k;
    for (int j = 0; j < NC; ++j)
      for (int m = 0; m < 2; ++m) {
        unsigned bits = 0;
        for (int q = 0; q < 4; ++q) bits |= nz(coefC[(j * 2 + m) * 4 + q]) ? (1u << q) : 0u;
        nzCt |= bits << (4 * (j * 2 + m));
      }
  }
  // ---- closed-first-axis tile tables (k_pair_tile5): the first left axis (dimension kD0) is closed,
  // every other left axis is open; group g holds the rows l = g + t * ngroups5
  bool t5_ok = tile_ok && nL >= 2 && ldims[0] == kD0;
  int KO5 = 0;
  for (int i = 1; i < nL && t5_ok; ++i) KO5 += ldims[i] - 1;
  t5_ok = t5_ok && pair_tile5_supported(static_cast<int>(NR), KO5, KR) && KO5 <= 8;
  const long long ngroups5 = t5_ok ? NL / kD0 : 0;
  std::vector<int> ofsO5(ngroups5 * KO5 > 0 ? ngroups5 * KO5 : 1, 0);
  std::vector<c128> coefO5(ngroups5 * KO5 > 0 ? ngroups5 * KO5 * 4 : 4, c128{0, 0});
  std::vector<unsigned> nzO5(ngroups5 > 0 ? ngroups5 : 1, 0u), hselO5(ngroups5 > 0 ? ngroups5 : 1, 0u);
  std::vector<c128> coefF(kD0 * (kD0 - 1) * 4, c128{0, 0});
  unsigned halfO5 = 0, nzF = 0;
  if (t5_ok) {
    std::vector<long long> lstride(nL);
    {
      long long st = 1;
      for (int i = nL - 1; i >= 0; --i) { lstride[i] = st; st *= ldims[i]; }
    }
    unsigned both5 = 0;
    for (long long g = 0; g < ngroups5; ++g) {
      int k = 0;
      for (int i = 1; i < nL; ++i) {
        const int d = ldims[i];
        const int m = static_cast<int>((g / lstride[i]) % d);
        const c128* B = blk[i];
        for (int m2 = 0; m2 < d; ++m2) {
          if (m2 == m) continue;
          ofsO5[g * KO5 + k] = static_cast<int>((m2 - m) * lstride[i]);
          c128* C = &coefO5[(g * KO5 + k) * 4];
          for (int e = 0; e < 2; ++e)
            for (int e2 = 0; e2 < 2; ++e2) C[e * 2 + e2] = B[(e * d + m) * (2 * d) + (e2 * d + m2)];
          unsigned bits = 0;
          for (int q = 0; q < 4; ++q) bits |= nz(C[q]) ? (1u << q) : 0u;
          nzO5[g] |= bits << (4 * k);
          if ((bits & 10u) != 0 && (bits & 5u) == 0) hselO5[g] |= 1u << k;
          if ((bits & 5u) != 0 && (bits & 10u) != 0) both5 |= 1u << k;
          ++k;
        }
      }
    }
    for (int k = 0; k < KO5; ++k)
      if (!((both5 >> k) & 1u)) halfO5 |= 1u << k;
    const c128* B = blk[0];      // (2 kD0) x (2 kD0), row = e * kD0 + t
    for (int t = 0; t < kD0; ++t) {
      int sidx = 0;
      for (int t2 = 0; t2 < kD0; ++t2) {
        if (t2 == t) continue;
        c128* C = &coefF[(t * (kD0 - 1) + sidx) * 4];
        for (int e = 0; e < 2; ++e)
          for (int e2 = 0; e2 < 2; ++e2) C[e * 2 + e2] = B[(e * kD0 + t) * (2 * kD0) + (e2 * kD0 + t2)];
        unsigned bits = 0;
        for (int q = 0; q < 4; ++q) bits |= nz(C[q]) ? (1u << q) : 0u;
        nzF |= bits << (4 * (t * (kD0 - 1) + sidx));
        ++sidx;
      }
    }
  }
  PairOp* op = new PairOp();
  PairTables& T = op->t;
  T.NL = static_cast<int>(NL);
  T.NR = static_cast<int>(NR);
  T.KL = KL;
  T.KR = KR;
  T.NLR = NL * NR;
  // pack device buffer
  size_t bytes = 0;
  auto reserve = [&](size_t n) { size_t o = bytes; bytes += (n + 255) & ~size_t(255); return o; };
  const size_t o_dA = reserve(dA.size() * 16), o_dB = reserve(dB.size() * 16);
  const size_t o_cA = reserve(fA.size() * 16), o_cB = reserve(fB.size() * 16), o_P = reserve(16 * 16);
  const size_t o_offL = reserve(offL.size() * 4), o_offR = reserve(offR.size() * 4);
  const size_t o_coL = reserve(coefL.size() * 16), o_coR = reserve(coefR.size() * 16);
  const size_t o_cC = reserve(coefC.size() * 16);
  const size_t o_oO = reserve(ofsO.size() * 4), o_cO = reserve(coefO.size() * 16), o_fO = reserve(maskO.size() * 4);
  const size_t o_pR = reserve(packR.size() * 4), o_cT = reserve(coefT.size() * 16);
  const size_t o_nzO = reserve(nzO.size() * 4), o_hsO = reserve(hselO.size() * 4);
  const size_t o_oO5 = reserve(ofsO5.size() * 4), o_cO5 = reserve(coefO5.size() * 16);
  const size_t o_nzO5 = reserve(nzO5.size() * 4), o_hsO5 = reserve(hselO5.size() * 4);
  const size_t o_cF = reserve(coefF.size() * 16);
  std::vector<char> host(bytes);
  std::memcpy(host.data() + o_oO5, ofsO5.data(), ofsO5.size() * 4);
  std::memcpy(host.data() + o_cO5, coefO5.data(), coefO5.size() * 16);
  std::memcpy(host.data() + o_nzO5, nzO5.data(), nzO5.size() * 4);
  std::memcpy(host.data() + o_hsO5, hselO5.data(), hselO5.size() * 4);
  std::memcpy(host.data() + o_cF, coefF.data(), coefF.size() * 16);
  std::memcpy(host.data() + o_cC, coefC.data(), coefC.size() * 16);
  std::memcpy(host.data() + o_oO, ofsO.data(), ofsO.size() * 4);
  std::memcpy(host.data() + o_cO, coefO.data(), coefO.size() * 16);
  std::memcpy(host.data() + o_fO, maskO.data(), maskO.size() * 4);
  std::memcpy(host.data() + o_pR, packR.data(), packR.size() * 4);
  std::memcpy(host.data() + o_cT, coefT.data(), coefT.size() * 16);
  std::memcpy(host.data() + o_nzO, nzO.data(), nzO.size() * 4);
  std::memcpy(host.data() + o_hsO, hselO.data(), hselO.size() * 4);
  std::memcpy(host.data() + o_dA, dA.data(), dA.size() * 16);
  std::memcpy(host.data() + o_dB, dB.data(), dB.size() * 16);
  std::memcpy(host.data() + o_cA, fA.data(), fA.size() * 16);
  std::memcpy(host.data() + o_cB, fB.data(), fB.size() * 16);
  std::memcpy(host.data() + o_P, pairv.data(), 16 * 16);
  std::memcpy(host.data() + o_offL, offL.data(), offL.size() * 4);
  std::memcpy(host.data() + o_offR, offR.data(), offR.size() * 4);
  std::memcpy(host.data() + o_coL, coefL.data(), coefL.size() * 16);
  std::memcpy(host.data() + o_coR, coefR.data(), coefR.size() * 16);
  void* dev = nullptr;
  if (hipMalloc(&dev, bytes) != hipSuccess) {
    delete op;
    return fail(RPM_E_NOMEM, "hipMalloc(%zu) for pair tables failed", bytes);
  }
  if (hipMemcpy(dev, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dev);
    delete op;
    return fail(RPM_E_HIP, "copy of pair tables failed");
  }
  char* base = static_cast<char*>(dev);
  T.dA = reinterpret_cast<const c128*>(base + o_dA);
  T.dB = reinterpret_cast<const c128*>(base + o_dB);
  T.fA = reinterpret_cast<const c128*>(base + o_cA);
  T.fB = reinterpret_cast<const c128*>(base + o_cB);
  T.pair = reinterpret_cast<const c128*>(base + o_P);
  T.offL = reinterpret_cast<const int*>(base + o_offL);
  T.offR = reinterpret_cast<const int*>(base + o_offR);
  T.coefL = reinterpret_cast<const c128*>(base + o_coL);
  T.coefR = reinterpret_cast<const c128*>(base + o_coR);
  op->use_tile = tile_ok;
  op->kernel = 0;
  if (tile_ok) {
    TileTables& U = op->tile;
    U.NL = T.NL;
    U.NR = T.NR;
    U.NC = NC;
    U.KO = KO;
    U.KR = KR;
    U.ngroups = static_cast<int>(ngroups);
    U.NLR = T.NLR;
    U.dA = T.dA;
    U.dB = T.dB;
    U.fA = T.fA;
    U.fB = T.fB;
    U.pair = T.pair;
    U.offR = T.offR;
    U.coefR = T.coefR;
    U.maskR = 0;
    for (int k = 0; k < KR; ++k) U.maskR |= static_cast<unsigned>(flagR[k]) << (2 * k);
    U.coefC = reinterpret_cast<const c128*>(base + o_cC);
    U.maskC = 0;
    for (int j = 0; j < NC; ++j) U.maskC |= static_cast<unsigned>(flagC[j]) << (2 * j);
    U.ofsO = reinterpret_cast<const int*>(base + o_oO);
    U.coefO = reinterpret_cast<const c128*>(base + o_cO);
    U.maskO = reinterpret_cast<const unsigned*>(base + o_fO);
    U.packR = reinterpret_cast<const int*>(base + o_pR);
    U.coefT = reinterpret_cast<const c128*>(base + o_cT);
    U.nzO = reinterpret_cast<const unsigned*>(base + o_nzO);
    U.hselO = reinterpret_cast<const unsigned*>(base + o_hsO);
    U.halfO = halfO;
    U.nzC = nzCt;
    U.t5 = t5_ok;
    U.ngroups5 = static_cast<int>(ngroups5);
    U.KO5 = KO5;
    U.ofsO5 = reinterpret_cast<const int*>(base + o_oO5);
    U.coefO5 = reinterpret_cast<const c128*>(base + o_cO5);
    U.nzO5 = reinterpret_cast<const unsigned*>(base + o_nzO5);
    U.hselO5 = reinterpret_cast<const unsigned*>(base + o_hsO5);
    U.halfO5 = halfO5;
    U.coefF = reinterpret_cast<const c128*>(base + o_cF);
    U.nzF = nzF;
  }
  op->dev = dev;
  *out = op;
  return RPM_OK;
}

// ---------------------------------------------------------------- generic local terms

constexpr int kMaxTermAxes = 8;
constexpr int kMaxLocal = 64;

struct LocalTermDev {
  int naxes, D;
  long long stride[kMaxTermAxes];  // state strides of the term axes (as listed)
  int dim[kMaxTermAxes];
  int op_offset;                   // into ops (complex elements)
  int col_offset;                  // into cols (long long elements)
};

struct LocalOp {
  int nterms;
  long long n;
  LocalTermDev* terms;   // device
  c128* ops;             // device
  long long* cols;       // device: flat offset of local column b relative to the base
  void* dev;
};

__global__ void __launch_bounds__(256) k_local_apply(int nterms, long long n,
                                                     const LocalTermDev* __restrict__ terms,
                                                     const c128* __restrict__ ops,
                                                     const long long* __restrict__ cols,
                                                     const c128* __restrict__ in,
                                                     c128* __restrict__ out, long long ld, ApplyCtl A) {
  const long long b = blockIdx.y;
  double s = 1.0;
  if (A.ctl) {
    const BasisCtl& c = A.ctl[b];
    if (!basis_step_live(c, A.j)) return;
    s = c.scale;
  }
  const long long i0 = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool valid = i0 < n;
  const long long i = valid ? i0 : 0;   // idle threads recompute element 0 and store nothing
  const c128* x = in + b * ld;
  c128 acc{0.0, 0.0};
  for (int t = 0; t < nterms; ++t) {
    const LocalTermDev& T = terms[t];
    int a = 0;
    long long base = i;
    for (int k = 0; k < T.naxes; ++k) {
      const int dig = static_cast<int>((i / T.stride[k]) % T.dim[k]);
      a = a * T.dim[k] + dig;
      base -= dig * T.stride[k];
    }
    const c128* row = ops + T.op_offset + static_cast<long long>(a) * T.D;
    const long long* cc = cols + T.col_offset;
    for (int c = 0; c < T.D; ++c) acc = cfma(row[c], x[base + cc[c]], acc);
  }
  c128 v = cscale(acc, s);
  c128 dot{0.0, 0.0};
  if (A.dots && valid) dot = cfmaconj(cscale(x[i], s), v, dot);
  if (A.prev) {
    const double bprev = A.ctl[b].beta_prev;
    const c128 q = A.prev[b * A.prev_ld + i];
    v.re = fma(-bprev, q.re, v.re);
    v.im = fma(-bprev, q.im, v.im);
  }
  if (valid) {
    out[b * ld + i] = v;
    if (A.copy) A.copy[b * A.copy_ld + i] = cscale(x[i], s);
  }
  if (A.dots) {
    __shared__ c128 red[256 / kWave];
    block_dot_store(dot, red, A.dots + b * A.dots_ld + blockIdx.x);
  }
}

static int build_local(int naxes, const int* dims, int nterms, const int* tnaxes, const int* taxes,
                       const double* tops, LocalOp** out) {
  if (naxes < 1 || naxes > 64) return fail(RPM_E_INVALID, "naxes out of range");
  std::vector<long long> stride(naxes);
  long long n = 1;
  for (int a = naxes - 1; a >= 0; --a) {
    if (dims[a] < 1) return fail(RPM_E_INVALID, "Every tensor dimension must be positive.");
    stride[a] = n;
    n *= dims[a];
  }
  if (n > (1LL << 31)) return fail(RPM_E_INVALID, "state too large");
  std::vector<LocalTermDev> terms(nterms);
  std::vector<c128> ops;
  std::vector<long long> cols;
  const int* ax = taxes;
  const c128* op = reinterpret_cast<const c128*>(tops);
  for (int t = 0; t < nterms; ++t) {
    LocalTermDev& T = terms[t];
    T.naxes = tnaxes[t];
    if (T.naxes < 1 || T.naxes > kMaxTermAxes)
      return fail(RPM_E_INVALID, "term %d has %d axes (1..%d supported)", t, T.naxes, kMaxTermAxes);
    int D = 1;
    for (int k = 0; k < T.naxes; ++k) {
      const int a = ax[k];
      if (a < 0 || a >= naxes) return fail(RPM_E_INVALID, "term %d contains an out-of-range axis.", t);
      for (int k2 = 0; k2 < k; ++k2)
        if (ax[k2] == a) return fail(RPM_E_INVALID, "term %d must have unique active axes.", t);
      T.stride[k] = stride[a];
      T.dim[k] = dims[a];
      D *= dims[a];
    }
    if (D > kMaxLocal) return fail(RPM_E_UNSUPPORTED, "term %d local dimension %d > %d", t, D, kMaxLocal);
    T.D = D;
    T.op_offset = static_cast<int>(ops.size());
    T.col_offset = static_cast<int>(cols.size());
    for (int e = 0; e < D * D; ++e) ops.push_back(op[e]);
    for (int c = 0; c < D; ++c) {
      long long o = 0;
      int rem = c;
      for (int k = T.naxes - 1; k >= 0; --k) {
        o += static_cast<long long>(rem % T.dim[k]) * T.stride[k];
        rem /= T.dim[k];
      }
      cols.push_back(o);
    }
    ax += T.naxes;
    op += D * D;
  }
  if (ops.empty()) ops.push_back(c128{0, 0});
  if (cols.empty()) cols.push_back(0);
  const size_t b_terms = (terms.size() * sizeof(LocalTermDev) + 255) & ~size_t(255);
  const size_t b_ops = (ops.size() * 16 + 255) & ~size_t(255);
  const size_t b_cols = cols.size() * 8;
  void* dev = nullptr;
  if (hipMalloc(&dev, b_terms + b_ops + b_cols + 256) != hipSuccess)
    return fail(RPM_E_NOMEM, "hipMalloc for local-term tables failed");
  char* base = static_cast<char*>(dev);
  bool ok = true;
  if (!terms.empty())
    ok = ok && hipMemcpy(base, terms.data(), terms.size() * sizeof(LocalTermDev), hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(base + b_terms, ops.data(), ops.size() * 16, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(base + b_terms + b_ops, cols.data(), b_cols, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    (void)hipFree(dev);
    return fail(RPM_E_HIP, "copy of local-term tables failed");
  }
  LocalOp* L = new LocalOp();
  L->nterms = nterms;
  L->n = n;
  L->terms = reinterpret_cast<LocalTermDev*>(base);
  L->ops = reinterpret_cast<c128*>(base + b_terms);
  L->cols = reinterpret_cast<long long*>(base + b_terms + b_ops);
  L->dev = dev;
  *out = L;
  return RPM_OK;
}

int apply_units(const rpm_operator* op) {
  if (op->kind == 1) {
    const PairOp* P = static_cast<const PairOp*>(op->impl);
    if (P->use_tile && P->kernel == 4) return P->tile.ngroups5;
    if (P->use_tile && P->kernel != 1) return P->tile.ngroups;
    return ceil_div(P->t.NLR, 256);
  }
  return ceil_div(static_cast<const LocalOp*>(op->impl)->n, 256);
}

int operator_apply_ex(const rpm_operator* op, const c128* in, c128* out, int batch, long long ld,
                      const ApplyCtl& A, hipStream_t stream) {
  if (!op) return fail(RPM_E_INVALID, "NULL operator");
  if ((A.prev || A.dots) && (!A.ctl || A.affine)) return fail(RPM_E_INVALID, "three-term options need a basis ctl");
  if (A.dots && A.dots_ld < apply_units(op)) return fail(RPM_E_INVALID, "dots stride below the unit count");
  if (batch < 1 || batch > 65535) return fail(RPM_E_INVALID, "batch %d out of range", batch);
  if (ld < op->n) return fail(RPM_E_INVALID, "ld %lld smaller than the dimension %lld", ld, op->n);
  if (!in || !out) return fail(RPM_E_INVALID, "NULL vector");
  if (op->kind == 1) {
    const PairOp* P = static_cast<const PairOp*>(op->impl);
    if (P->use_tile && P->kernel != 1) {
      if (static_cast<long long>(P->tile.ngroups5 > P->tile.ngroups ? P->tile.ngroups5 : P->tile.ngroups) * batch >
          0x7fffffffLL)
        return fail(RPM_E_INVALID, "batch %d too large for one launch", batch);
      if (P->kernel == 2) return launch_pair_tile(P->tile, in, out, batch, ld, A, stream);
      if (P->kernel == 4) return launch_pair_tile5(P->tile, in, out, batch, ld, A, stream);
      return launch_pair_tile3(P->tile, in, out, batch, ld, A, stream);
    }
    dim3 grid(ceil_div(P->t.NLR, 256), batch);
    hipLaunchKernelGGL(k_pair_apply, grid, dim3(256), 0, stream, P->t, in, out, ld, A);
  } else {
    if (A.pairs) return fail(RPM_E_UNSUPPORTED, "per-vector electron matrices need a pair operator");
    if (A.affine) return fail(RPM_E_UNSUPPORTED, "the affine epilogue needs a pair operator");
    const LocalOp* L = static_cast<const LocalOp*>(op->impl);
    dim3 grid(ceil_div(L->n, 256), batch);
    hipLaunchKernelGGL(k_local_apply, grid, dim3(256), 0, stream, L->nterms, L->n, L->terms, L->ops,
                       L->cols, in, out, ld, A);
  }
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

}  // namespace rpm

using namespace rpm;

extern "C" int rpm_operator_create_pair(int n_left, const int* left_dims, int n_right,
                                        const int* right_dims, const double* nuclear_blocks,
                                        const double* electron_pair, rpm_operator** out) {
  if (!out || !electron_pair || (!nuclear_blocks && n_left + n_right > 0))
    return fail(RPM_E_INVALID, "NULL argument");
  PairOp* P = nullptr;
  int rc = build_pair(n_left, left_dims, n_right, right_dims, nuclear_blocks, electron_pair, &P);
  if (rc) return rc;
  rpm_operator* h = new rpm_operator();
  h->kind = 1;
  h->n = 4LL * P->t.NLR;
  h->impl = P;
  *out = h;
  return RPM_OK;
}

extern "C" int rpm_operator_create_local(int naxes, const int* dims, int nterms, const int* term_naxes,
                                         const int* term_axes, const double* term_ops,
                                         rpm_operator** out) {
  if (!out || !dims || (nterms > 0 && (!term_naxes || !term_axes || !term_ops)) || nterms < 0)
    return fail(RPM_E_INVALID, "NULL argument");
  LocalOp* L = nullptr;
  int rc = build_local(naxes, dims, nterms, term_naxes, term_axes, term_ops, &L);
  if (rc) return rc;
  rpm_operator* h = new rpm_operator();
  h->kind = 2;
  h->n = L->n;
  h->impl = L;
  *out = h;
  return RPM_OK;
}

extern "C" int rpm_operator_destroy(rpm_operator* op) {
  if (!op) return RPM_OK;
  if (op->kind == 1) {
    PairOp* P = static_cast<PairOp*>(op->impl);
    if (P->dev) (void)hipFree(P->dev);
    delete P;
  } else if (op->kind == 2) {
    LocalOp* L = static_cast<LocalOp*>(op->impl);
    if (L->dev) (void)hipFree(L->dev);
    delete L;
  }
  delete op;
  return RPM_OK;
}

extern "C" long long rpm_operator_dimension(const rpm_operator* op) { return op ? op->n : -1; }

extern "C" int rpm_operator_apply(const rpm_operator* op, const void* d_in, void* d_out, int batch,
                                  long long ld, void* stream) {
  if (d_in == d_out) return fail(RPM_E_INVALID, "d_in and d_out must not overlap");
  return operator_apply_ex(op, static_cast<const c128*>(d_in), static_cast<c128*>(d_out), batch, ld, ApplyCtl{},
                           static_cast<hipStream_t>(stream));
}

extern "C" int rpm_operator_apply_pairs(const rpm_operator* op, const void* d_in, void* d_out, int batch,
                                        long long ld, const void* d_pairs, void* stream) {
  if (d_in == d_out) return fail(RPM_E_INVALID, "d_in and d_out must not overlap");
  if (!d_pairs) return fail(RPM_E_INVALID, "rpm_operator_apply_pairs: NULL electron matrices");
  ApplyCtl A;
  A.pairs = static_cast<const c128*>(d_pairs);
  return operator_apply_ex(op, static_cast<const c128*>(d_in), static_cast<c128*>(d_out), batch, ld, A,
                           static_cast<hipStream_t>(stream));
}

extern "C" int rpm_operator_apply_affine(const rpm_operator* op, const void* d_in, void* d_out, int batch,
                                         long long ld, const void* d_pairs, double s, double a_in, const void* d_x0,
                                         double c0_re, double c0_im, const void* d_x1, double c1, void* stream) {
  if (d_in == d_out || (d_x0 && d_x0 == d_in) || (d_x1 && d_x1 == d_in))
    return fail(RPM_E_INVALID, "rpm_operator_apply_affine: the input may not alias the output or x0/x1");
  if (d_x0 && d_x0 == d_out) return fail(RPM_E_INVALID, "rpm_operator_apply_affine: x0 may not alias the output");
  if (op && op->kind != 1) return fail(RPM_E_UNSUPPORTED, "the affine epilogue needs a pair operator");
  ApplyCtl A;
  A.pairs = static_cast<const c128*>(d_pairs);
  A.affine = 1;
  A.aff_s = s;
  A.aff_in = a_in;
  A.x0 = static_cast<const c128*>(d_x0);
  A.aff_c0 = cmake(c0_re, c0_im);
  A.x1 = static_cast<const c128*>(d_x1);
  A.aff_c1 = c1;
  return operator_apply_ex(op, static_cast<const c128*>(d_in), static_cast<c128*>(d_out), batch, ld, A,
                           static_cast<hipStream_t>(stream));
}

extern "C" int rpm_operator_set_kernel(rpm_operator* op, int kernel) {
  if (!op || kernel < 0 || kernel > 4) return fail(RPM_E_INVALID, "rpm_operator_set_kernel: bad arguments");
  if (op->kind != 1) return kernel >= 2 ? fail(RPM_E_UNSUPPORTED, "tiled kernels need a pair operator") : RPM_OK;
  PairOp* P = static_cast<PairOp*>(op->impl);
  if (kernel >= 2 && !P->use_tile) return fail(RPM_E_UNSUPPORTED, "the operator shape has no tiled kernel");
  if (kernel == 4 && !P->tile.t5)
    return fail(RPM_E_UNSUPPORTED, "the operator shape has no closed-first-axis tiled kernel");
  P->kernel = kernel;
  return RPM_OK;
}

extern "C" int rpm_operator_kernel(const rpm_operator* op) {
  if (!op) return -1;
  if (op->kind != 1) return 9;
  const PairOp* P = static_cast<const PairOp*>(op->impl);
  if (!P->use_tile || P->kernel == 1) return 1;
  if (P->kernel == 4) return 4;
  return P->kernel == 2 ? 2 : 3;
}
