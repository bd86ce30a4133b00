// Tiled pair stencils for the full-size two-radical layout (the production
// FAD-TrpH operator and random-tensor systems of the same shape).
//
// Layout: psi[(e * NL + l) * NR + r], e = eA*2+eB.  One workgroup owns a tile
// of all 4 electron states x LG = 2^NC consecutive left indices (the NC
// lowest left axes, all spin-1/2, are "closed" inside the tile) x all NR right
// indices:
//   * diagonal and electron flips: registers;
//   * closed left axes and right-radical neighbours r': the tile, staged in LDS
//     (k_pair_tile3) or held in registers (k_pair_tile);
//   * the other left axes (open, KO neighbour offsets): read from global
//     memory (L2), the same offsets and coefficients for the whole tile.
// Structurally zero couplings are skipped with scalar branches: k_pair_tile
// masks whole halves (same-electron or electron-flip) of each 2x2 electron
// block; k_pair_tile3 skips single zero entries of the left blocks (a planar
// radical's hyperfine tensors, A_xz = A_yz = 0, leave one or two of the four)
// and loads only the input electron half an open neighbour's nonzero entries
// read.  Workgroups are mapped XCD-aware: all tiles of a vector run on one XCD,
// so the re-read neighbours stay in that XCD's L2.
#include <cstring>
#include <vector>

#include "rpm_common.hpp"
#include "pair_tile.hpp"

namespace rpm {

// Operator tables are never written while a kernel runs: reading them through the constant address
// space lets the compiler use scalar loads after the kernel's own stores.
template <typename T>
using kconst = const __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ kconst<T>* kc(const T* p) {
  return (kconst<T>*)(p);
}

__device__ __forceinline__ c128 kload(kconst<c128>* p) { return c128{p->re, p->im}; }

// out[eA,eB] += sum_e' C[eA][e'] q[e',eB], entry eA*2+e' taken when bit (eA*2+e') of nz is set
// (nz wave-uniform).  h0 = (q[0], q[1]), the input half eA' = 0; h1 = (q[2], q[3]), eA' = 1.
template <typename CP>
__device__ __forceinline__ void mac_left_nz(CP C, unsigned nz, const c128* h0, const c128* h1, c128 acc[4]) {
  if (nz & 1u) {
    const c128 a = kload(C);
    acc[0] = cfma(a, h0[0], acc[0]);
    acc[1] = cfma(a, h0[1], acc[1]);
  }
  if (nz & 2u) {
    const c128 a = kload(C + 1);
    acc[0] = cfma(a, h1[0], acc[0]);
    acc[1] = cfma(a, h1[1], acc[1]);
  }
  if (nz & 4u) {
    const c128 a = kload(C + 2);
    acc[2] = cfma(a, h0[0], acc[2]);
    acc[3] = cfma(a, h0[1], acc[3]);
  }
  if (nz & 8u) {
    const c128 a = kload(C + 3);
    acc[2] = cfma(a, h1[0], acc[2]);
    acc[3] = cfma(a, h1[1], acc[3]);
  }
}

// XCD-aware mapping of the linear workgroup id onto (vector, group).
__device__ inline void tile_map(int G, int batch, int& b, int& grp) {
  const int g = blockIdx.x;
  if (batch % kXcds == 0) {
    const int x = g % kXcds, k = g / kXcds;
    b = (k / G) * kXcds + x;
    grp = k % G;
  } else {
    b = g / G;
    grp = g % G;
  }
}

// Per-vector control: false when vector b is idle in this launch; s = input scale.
__device__ inline bool tile_begin(const ApplyCtl& A, int b, double& s) {
  s = 1.0;
  if (A.ctl) {
    const BasisCtl& c = A.ctl[b];
    if (!basis_step_live(c, A.j)) return false;
    s = c.scale;
  }
  return true;
}

// Raw buffer access to one vector's amplitudes: a wave-uniform scalar offset (the electron state's
// block) plus a 32-bit per-lane byte offset, so each load's address costs one 32-bit add.
typedef int i32x4 __attribute__((ext_vector_type(4)));
constexpr int kBufferWord3 = 0x00020000;   // raw 32-bit data, gfx9 family buffer resource

__device__ __forceinline__ __amdgpu_buffer_rsrc_t vector_rsrc(const c128* base, long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<c128*>(base), 0, static_cast<int>(bytes), kBufferWord3);
}

__device__ __forceinline__ c128 bld16(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff) {
  const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>(voff), static_cast<int>(soff), 0);
  return __builtin_bit_cast(c128, v);
}

__device__ inline void mac2x2_left(const c128* __restrict__ C, int flags, const c128 q[4], c128 acc[4]) {
  // out[eA,eB] += sum_e' C[eA][e'] q[e',eB]   (q index e'*2+eB)
  if (flags & kSameE) {
    acc[0] = cfma(C[0], q[0], acc[0]);
    acc[1] = cfma(C[0], q[1], acc[1]);
    acc[2] = cfma(C[3], q[2], acc[2]);
    acc[3] = cfma(C[3], q[3], acc[3]);
  }
  if (flags & kFlipE) {
    acc[0] = cfma(C[1], q[2], acc[0]);
    acc[1] = cfma(C[1], q[3], acc[1]);
    acc[2] = cfma(C[2], q[0], acc[2]);
    acc[3] = cfma(C[2], q[1], acc[3]);
  }
}

__device__ inline void mac2x2_right(const c128 C[4], int flags, const c128 q[4], c128 acc[4]) {
  // out[eA,eB] += sum_e' C[eB][e'] q[eA,e']   (q index eA*2+e')
  if (flags & kSameE) {
    acc[0] = cfma(C[0], q[0], acc[0]);
    acc[1] = cfma(C[3], q[1], acc[1]);
    acc[2] = cfma(C[0], q[2], acc[2]);
    acc[3] = cfma(C[3], q[3], acc[3]);
  }
  if (flags & kFlipE) {
    acc[0] = cfma(C[1], q[1], acc[0]);
    acc[1] = cfma(C[2], q[0], acc[1]);
    acc[2] = cfma(C[1], q[3], acc[2]);
    acc[3] = cfma(C[2], q[2], acc[3]);
  }
}

// Register-tiled variant: thread r holds the tile column psi[e][l][r] (4*LG
// amplitudes) and all 4*LG outputs.
template <int NR, int NC, int KO, int KR>
__global__ void __launch_bounds__(NR) k_pair_tile(TileTables T, const c128* __restrict__ in, c128* __restrict__ out,
                                                  long long ld, int batch, ApplyCtl A) {
  constexpr int LG = 1 << NC;
  __shared__ c128 tile[4 * LG * NR];
  int b, grp;
  tile_map(T.ngroups, batch, b, grp);
  double s;
  if (!tile_begin(A, b, s)) return;
  PairCoef E;
  load_pair_coef(A.pairs ? A.pairs + b * 16 : T.pair, E);
  const int r = threadIdx.x;
  const long long S = T.NLR;
  const c128* __restrict__ x = in + b * ld;
  const int l0 = grp * LG;

  c128 psi[LG][4];
#pragma unroll
  for (int l = 0; l < LG; ++l)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      psi[l][e] = x[e * S + static_cast<long long>(l0 + l) * NR + r];
      tile[(e * LG + l) * NR + r] = psi[l][e];
    }

  c128 acc[LG][4];
  {
    const c128 dB0 = T.dB[r], dB1 = T.dB[NR + r], fB0 = T.fB[r], fB1 = T.fB[NR + r];
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const c128 dA0 = T.dA[l0 + l], dA1 = T.dA[T.NL + l0 + l];
      const c128 fA0 = T.fA[l0 + l], fA1 = T.fA[T.NL + l0 + l];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const c128 d = cadd(cadd((e >> 1) ? dA1 : dA0, (e & 1) ? dB1 : dB0), E.d[e]);
        c128 a = cfma(d, psi[l][e], cmake(0.0, 0.0));
        a = cfma(cadd((e >> 1) ? fA1 : fA0, E.fa[e]), psi[l][e ^ 2], a);
        a = cfma(cadd((e & 1) ? fB1 : fB0, E.fb[e]), psi[l][e ^ 1], a);
        a = cfma(E.bo[e], psi[l][e ^ 3], a);
        acc[l][e] = a;
      }
    }
  }

#pragma unroll
  for (int j = 0; j < NC; ++j) {
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const int m = (l >> j) & 1;
      mac2x2_left(T.coefC + (j * 2 + m) * 4, (T.maskC >> (2 * j)) & 3, psi[l ^ (1 << j)], acc[l]);
    }
  }

  {
    const int* __restrict__ ofs = T.ofsO + grp * KO;
    const c128* __restrict__ cof = T.coefO + static_cast<long long>(grp) * KO * 4;
    const unsigned fm = T.maskO[grp];
#pragma unroll 1
    for (int k = 0; k < KO; ++k) {
      const long long o = static_cast<long long>(l0 + ofs[k]) * NR + r;
      const int f = (fm >> (2 * k)) & 3;
#pragma unroll
      for (int l = 0; l < LG; ++l) {
        c128 q[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = x[e * S + o + static_cast<long long>(l) * NR];
        mac2x2_left(cof + k * 4, f, q, acc[l]);
      }
    }
  }

  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < KR; ++k) {
    const int r2 = r + T.offR[r * KR + k];
    c128 C[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) C[i] = T.coefR[(r * KR + k) * 4 + i];
    const int f = (T.maskR >> (2 * k)) & 3;
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      c128 q[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) q[e] = tile[(e * LG + l) * NR + r2];
      mac2x2_right(C, f, q, acc[l]);
    }
  }

  c128* __restrict__ y = out + b * ld;
  const double bprev = A.prev ? A.ctl[b].beta_prev : 0.0;
  c128 dot{0.0, 0.0};
#pragma unroll
  for (int l = 0; l < LG; ++l)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long long idx = e * S + static_cast<long long>(l0 + l) * NR + r;
      c128 v = apply_epilogue(A, acc[l][e], psi[l][e], s, b, ld, idx);
      if (A.dots) dot = cfmaconj(cscale(psi[l][e], s), v, dot);
      if (A.prev) {
        const c128 p = A.prev[b * A.prev_ld + idx];
        v.re = fma(-bprev, p.re, v.re);
        v.im = fma(-bprev, p.im, v.im);
      }
      __builtin_nontemporal_store(v.re, &y[idx].re);
      __builtin_nontemporal_store(v.im, &y[idx].im);
    }
  if (A.dots) {
    __shared__ c128 red[NR / kWave];
    block_dot_store(dot, red, A.dots + b * A.dots_ld + grp);
  }
  if (A.copy) {
    c128* z = A.copy + b * A.copy_ld;
#pragma unroll
    for (int l = 0; l < LG; ++l)
#pragma unroll
      for (int e = 0; e < 4; ++e) z[e * S + static_cast<long long>(l0 + l) * NR + r] = cscale(psi[l][e], s);
  }
}

// LDS-tiled variant (default): one (l, r) per thread, 4 outputs per thread,
// closed left axes and right neighbours from LDS, right-radical coefficients
// from a (slot, digit) table staged in LDS.  HALF marks the open slots that
// read one input electron half in every group (bit k: slot k; the half is the
// group's hselO bit).
template <int NR, int NC, int KO, int KR, unsigned HALF>
__global__ void __launch_bounds__(NR << NC) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_pair_tile3(TileTables T, const c128* __restrict__ in, c128* __restrict__ out, long long ld, int batch, ApplyCtl A) {
  constexpr int LG = 1 << NC;
  constexpr int NT = NR * LG;
  static_assert(NR % kWave == 0, "a wave must hold a single left index");
  static_assert(KR % 4 == 0, "right neighbour descriptors are fetched four at a time");
  __shared__ c128 tile[4 * LG * NR];
  __shared__ c128 coefT[KR * 3 * 4];
  __shared__ c128 tabB[4 * NR];   // dB[0], dB[1], fB[0], fB[1] over r
  int b, grp;
  tile_map(T.ngroups, batch, b, grp);
  double s;
  if (!tile_begin(A, b, s)) return;
  const int tid = threadIdx.x;
  // Vector-memory instructions and FP64 VALU work, not HBM bandwidth, bound this kernel, so every
  // per-thread load that is not amplitude data is taken off the vector-memory path: l is
  // wave-uniform (NR is a multiple of the wave size), so the left tables are scalar loads; the right
  // tables are staged once per block in LDS; the right neighbour descriptors come in 16-byte pieces.
  const int l = __builtin_amdgcn_readfirstlane(tid / NR);
  const int r = tid - l * NR;
  const long long S = T.NLR;
  const c128* __restrict__ x = in + b * ld;
  const int l0 = grp * LG;
  const long long own = static_cast<long long>(l0 + l) * NR + r;
  // Amplitude loads go through a buffer resource over the vector (4 S amplitudes, under 2^31 bytes):
  // electron state e at scalar offset e * Sb, the lane's amplitude at vector offset ob.
  const long long Sb = S * static_cast<long long>(sizeof(c128));
  const __amdgpu_buffer_rsrc_t xr = vector_rsrc(x, 4 * Sb);
  const unsigned ob = static_cast<unsigned>(own * static_cast<long long>(sizeof(c128)));

  for (int i = tid; i < KR * 3 * 4; i += NT) coefT[i] = T.coefT[i];
  for (int i = tid; i < 4 * NR; i += NT) {
    const int q = i / NR;
    tabB[i] = (q < 2 ? T.dB : T.fB)[(q & 1) * NR + (i - q * NR)];
  }
  c128 psi[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    psi[e] = bld16(xr, ob, static_cast<unsigned>(e * Sb));
    tile[(e * LG + l) * NR + r] = psi[e];
  }

  // open left axes (global / L2): offsets and coefficients uniform over the tile
  c128 acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = cmake(0.0, 0.0);
  {
    const int* __restrict__ ofs = T.ofsO + grp * KO;
    kconst<c128>* cof = kc(T.coefO + static_cast<long long>(grp) * KO * 4);
    const unsigned nzo = T.nzO[grp];
    const unsigned hs = T.hselO[grp];
#pragma unroll
    for (int k = 0; k < KO; ++k) {
      const unsigned o = ob + static_cast<unsigned>(ofs[k] * NR * static_cast<int>(sizeof(c128)));
      const unsigned nzk = (nzo >> (4 * k)) & 0xfu;
      c128 q[4];
      if ((HALF >> k) & 1u) {
        // only the input half this group's nonzero entries read
        const unsigned h = 2 * ((hs >> k) & 1u);
        q[0] = bld16(xr, o, static_cast<unsigned>(h * Sb));
        q[1] = bld16(xr, o, static_cast<unsigned>((h + 1) * Sb));
        mac_left_nz(cof + k * 4, nzk, &q[0], &q[0], acc);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = bld16(xr, o, static_cast<unsigned>(e * Sb));
        mac_left_nz(cof + k * 4, nzk, &q[0], &q[2], acc);
      }
    }
  }

  __syncthreads();
  {
    PairCoef E;
    load_pair_coef(A.pairs ? A.pairs + b * 16 : T.pair, E);
    const c128 dA0 = T.dA[l0 + l], dA1 = T.dA[T.NL + l0 + l];
    const c128 fA0 = T.fA[l0 + l], fA1 = T.fA[T.NL + l0 + l];
    const c128 dB0 = tabB[r], dB1 = tabB[NR + r], fB0 = tabB[2 * NR + r], fB1 = tabB[3 * NR + r];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const c128 d = cadd(cadd((e >> 1) ? dA1 : dA0, (e & 1) ? dB1 : dB0), E.d[e]);
      c128 a = cfma(d, psi[e], acc[e]);
      a = cfma(cadd((e >> 1) ? fA1 : fA0, E.fa[e]), psi[e ^ 2], a);
      a = cfma(cadd((e & 1) ? fB1 : fB0, E.fb[e]), psi[e ^ 1], a);
      a = cfma(E.bo[e], psi[e ^ 3], a);
      acc[e] = a;
    }
  }
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const int l2 = l ^ (1 << j);
    const int m = (l >> j) & 1;
    const unsigned nzj = (T.nzC >> (4 * (j * 2 + m))) & 0xfu;
    c128 q[4];
    if ((nzj & 5u) == 0 || (nzj & 10u) == 0) {
      // the nonzero entries read one input electron half (l, hence m, is wave-uniform)
      const int h = (nzj & 10u) ? 2 : 0;
      q[0] = tile[(h * LG + l2) * NR + r];
      q[1] = tile[((h + 1) * LG + l2) * NR + r];
      mac_left_nz(kc(T.coefC) + (j * 2 + m) * 4, nzj, &q[0], &q[0], acc);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) q[e] = tile[(e * LG + l2) * NR + r];
      mac_left_nz(kc(T.coefC) + (j * 2 + m) * 4, nzj, &q[0], &q[2], acc);
    }
  }
  int pkr[KR];
#pragma unroll
  for (int k4 = 0; k4 < KR / 4; ++k4) {
    const int4 v = reinterpret_cast<const int4*>(T.packR + r * KR)[k4];
    pkr[4 * k4] = v.x;
    pkr[4 * k4 + 1] = v.y;
    pkr[4 * k4 + 2] = v.z;
    pkr[4 * k4 + 3] = v.w;
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int pk = pkr[k];
    const int off = static_cast<short>(pk & 0xffff);
    const int m = pk >> 16;
    c128 C[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) C[i] = coefT[(k * 3 + m) * 4 + i];
#pragma unroll
    for (int e = 0; e < 4; ++e) q[e] = tile[(e * LG + l) * NR + r + off];
    mac2x2_right(C, (T.maskR >> (2 * k)) & 3, q, acc);
  }

  char* __restrict__ yb = reinterpret_cast<char*>(out + b * ld);
  const double bprev = A.prev ? A.ctl[b].beta_prev : 0.0;
  const bool plain = !A.ctl && !A.affine;   // s == 1: the output is H in itself
  c128 dot{0.0, 0.0};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    c128 v = plain ? acc[e] : apply_epilogue(A, acc[e], psi[e], s, b, ld, e * S + own);
    if (A.dots) dot = cfmaconj(cscale(psi[e], s), v, dot);
    if (A.prev) {
      const c128 p = A.prev[b * A.prev_ld + e * S + own];
      v.re = fma(-bprev, p.re, v.re);
      v.im = fma(-bprev, p.im, v.im);
    }
    c128* yo = reinterpret_cast<c128*>(yb + e * Sb + ob);
    __builtin_nontemporal_store(v.re, &yo->re);
    __builtin_nontemporal_store(v.im, &yo->im);
  }
  if (A.copy) {
    c128* z = A.copy + b * A.copy_ld;
#pragma unroll
    for (int e = 0; e < 4; ++e) z[e * S + own] = cscale(psi[e], s);
  }
  if (A.dots) {
    __shared__ c128 red[NT / kWave];
    block_dot_store(dot, red, A.dots + b * A.dots_ld + grp);
  }
}


// Closed-first-axis variant: one workgroup holds the kD0 = 3 rows l = g + t * ngroups5 of the first
// left axis (the 14N5 of the production layout) over all NR right indices, 36 KB of amplitudes, so
// that axis's two neighbour slabs are read from LDS; every other left axis is open (L2), its
// offsets and coefficients uniform over the tile. Against k_pair_tile3 it loads one amplitude
// fewer per output (two full 14N5 slabs off, the H8c and Hbeta half slots on) and its 576-thread,
// 50 KB workgroups fit three to a CU. One (t, r) per thread, 4 outputs per thread.
template <int NR, int KO, int KR, unsigned HALF>
__global__ void __launch_bounds__(kD0 * NR) __attribute__((amdgpu_waves_per_eu(7, 8)))
k_pair_tile5(TileTables T, const c128* __restrict__ in, c128* __restrict__ out, long long ld, int batch, ApplyCtl A) {
  constexpr int NT = kD0 * NR;
  static_assert(NR % kWave == 0, "a wave must hold a single left index");
  static_assert(KR % 4 == 0, "right neighbour descriptors are fetched four at a time");
  __shared__ c128 tile[4 * kD0 * NR];
  __shared__ c128 coefT[KR * 3 * 4];
  __shared__ c128 tabB[4 * NR];   // dB[0], dB[1], fB[0], fB[1] over r
  int b, grp;
  tile_map(T.ngroups5, batch, b, grp);
  double s;
  if (!tile_begin(A, b, s)) return;
  const int tid = threadIdx.x;
  const int t = __builtin_amdgcn_readfirstlane(tid / NR);     // first-axis digit, wave-uniform
  const int r = tid - t * NR;
  const int l = grp + t * T.ngroups5;
  const long long S = T.NLR;
  const c128* __restrict__ x = in + b * ld;
  const long long own = static_cast<long long>(l) * NR + r;
  const long long Sb = S * static_cast<long long>(sizeof(c128));
  const __amdgpu_buffer_rsrc_t xr = vector_rsrc(x, 4 * Sb);
  const unsigned ob = static_cast<unsigned>(own * static_cast<long long>(sizeof(c128)));

  for (int i = tid; i < KR * 3 * 4; i += NT) coefT[i] = T.coefT[i];
  for (int i = tid; i < 4 * NR; i += NT) {
    const int q = i / NR;
    tabB[i] = (q < 2 ? T.dB : T.fB)[(q & 1) * NR + (i - q * NR)];
  }
  {
    c128 psi[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      psi[e] = bld16(xr, ob, static_cast<unsigned>(e * Sb));
      tile[(e * kD0 + t) * NR + r] = psi[e];
    }
  }

  // open left axes (global / L2): offsets and coefficients uniform over the tile
  c128 acc[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = cmake(0.0, 0.0);
  {
    const int* __restrict__ ofs = T.ofsO5 + grp * KO;
    kconst<c128>* cof = kc(T.coefO5 + static_cast<long long>(grp) * KO * 4);
    const unsigned nzo = T.nzO5[grp];
    const unsigned hs = T.hselO5[grp];
#pragma unroll
    for (int k = 0; k < KO; ++k) {
      const unsigned o = ob + static_cast<unsigned>(ofs[k] * NR * static_cast<int>(sizeof(c128)));
      const unsigned nzk = (nzo >> (4 * k)) & 0xfu;
      c128 q[4];
      if ((HALF >> k) & 1u) {
        const unsigned h = 2 * ((hs >> k) & 1u);
        q[0] = bld16(xr, o, static_cast<unsigned>(h * Sb));
        q[1] = bld16(xr, o, static_cast<unsigned>((h + 1) * Sb));
        mac_left_nz(cof + k * 4, nzk, &q[0], &q[0], acc);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) q[e] = bld16(xr, o, static_cast<unsigned>(e * Sb));
        mac_left_nz(cof + k * 4, nzk, &q[0], &q[2], acc);
      }
    }
  }

  __syncthreads();
  {
    c128 psi[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) psi[e] = tile[(e * kD0 + t) * NR + r];
    PairCoef E;
    load_pair_coef(A.pairs ? A.pairs + b * 16 : T.pair, E);
    const c128 dA0 = T.dA[l], dA1 = T.dA[T.NL + l];
    const c128 fA0 = T.fA[l], fA1 = T.fA[T.NL + l];
    const c128 dB0 = tabB[r], dB1 = tabB[NR + r], fB0 = tabB[2 * NR + r], fB1 = tabB[3 * NR + r];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const c128 d = cadd(cadd((e >> 1) ? dA1 : dA0, (e & 1) ? dB1 : dB0), E.d[e]);
      c128 a = cfma(d, psi[e], acc[e]);
      a = cfma(cadd((e >> 1) ? fA1 : fA0, E.fa[e]), psi[e ^ 2], a);
      a = cfma(cadd((e & 1) ? fB1 : fB0, E.fb[e]), psi[e ^ 1], a);
      a = cfma(E.bo[e], psi[e ^ 3], a);
      acc[e] = a;
    }
  }
  // the closed first axis: the other two rows of the tile (t uniform over a wave)
#pragma unroll
  for (int sidx = 0; sidx < kD0 - 1; ++sidx) {
    const int t2 = sidx < t ? sidx : sidx + 1;
    const unsigned nzf = (T.nzF >> (4 * (t * (kD0 - 1) + sidx))) & 0xfu;
    kconst<c128>* C = kc(T.coefF) + (t * (kD0 - 1) + sidx) * 4;
    c128 q[4];
    if ((nzf & 5u) == 0 || (nzf & 10u) == 0) {
      const int h = (nzf & 10u) ? 2 : 0;
      q[0] = tile[(h * kD0 + t2) * NR + r];
      q[1] = tile[((h + 1) * kD0 + t2) * NR + r];
      mac_left_nz(C, nzf, &q[0], &q[0], acc);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) q[e] = tile[(e * kD0 + t2) * NR + r];
      mac_left_nz(C, nzf, &q[0], &q[2], acc);
    }
  }
  int pkr[KR];
#pragma unroll
  for (int k4 = 0; k4 < KR / 4; ++k4) {
    const int4 v = reinterpret_cast<const int4*>(T.packR + r * KR)[k4];
    pkr[4 * k4] = v.x;
    pkr[4 * k4 + 1] = v.y;
    pkr[4 * k4 + 2] = v.z;
    pkr[4 * k4 + 3] = v.w;
  }
#pragma unroll
  for (int k = 0; k < KR; ++k) {
    const int pk = pkr[k];
    const int off = static_cast<short>(pk & 0xffff);
    const int m = pk >> 16;
    c128 C[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) C[i] = coefT[(k * 3 + m) * 4 + i];
#pragma unroll
    for (int e = 0; e < 4; ++e) q[e] = tile[(e * kD0 + t) * NR + r + off];
    mac2x2_right(C, (T.maskR >> (2 * k)) & 3, q, acc);
  }

  char* __restrict__ yb = reinterpret_cast<char*>(out + b * ld);
  const double bprev = A.prev ? A.ctl[b].beta_prev : 0.0;
  const bool plain = !A.ctl && !A.affine;   // s == 1: the output is H in itself
  c128 dot{0.0, 0.0};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const c128 pe = tile[(e * kD0 + t) * NR + r];
    c128 v = plain ? acc[e] : apply_epilogue(A, acc[e], pe, s, b, ld, e * S + own);
    if (A.dots) dot = cfmaconj(cscale(pe, s), v, dot);
    if (A.prev) {
      const c128 p = A.prev[b * A.prev_ld + e * S + own];
      v.re = fma(-bprev, p.re, v.re);
      v.im = fma(-bprev, p.im, v.im);
    }
    c128* yo = reinterpret_cast<c128*>(yb + e * Sb + ob);
    __builtin_nontemporal_store(v.re, &yo->re);
    __builtin_nontemporal_store(v.im, &yo->im);
  }
  if (A.copy) {
    c128* z = A.copy + b * A.copy_ld;
#pragma unroll
    for (int e = 0; e < 4; ++e) z[e * S + own] = cscale(tile[(e * kD0 + t) * NR + r], s);
  }
  if (A.dots) {
    __shared__ c128 red[NT / kWave];
    block_dot_store(dot, red, A.dots + b * A.dots_ld + grp);
  }
}

bool pair_tile_supported(int NR, int NC, int KO, int KR) { return NR == 192 && NC == 2 && KO == 7 && KR == 8; }

int launch_pair_tile(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                     hipStream_t stream) {
  if (!pair_tile_supported(T.NR, T.NC, T.KO, T.KR))
    return fail(RPM_E_UNSUPPORTED, "no tiled pair kernel for NR=%d NC=%d KO=%d KR=%d", T.NR, T.NC, T.KO, T.KR);
  hipLaunchKernelGGL((k_pair_tile<192, 2, 7, 8>), dim3(T.ngroups * batch), dim3(192), 0, stream, T, in, out, ld,
                     batch, A);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}



int launch_pair_tile3(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                      hipStream_t stream) {
  if (!pair_tile_supported(T.NR, T.NC, T.KO, T.KR))
    return fail(RPM_E_UNSUPPORTED, "no tiled pair kernel for NR=%d NC=%d KO=%d KR=%d", T.NR, T.NC, T.KO, T.KR);
  // the amplitude reads address a vector through a buffer resource with a 32-bit size
  if (4 * T.NLR * static_cast<long long>(sizeof(c128)) > 0x7fffffffLL)
    return fail(RPM_E_UNSUPPORTED, "vector of %lld amplitudes too large for the tiled kernel", 4 * T.NLR);
  constexpr unsigned kHalf = 0x60u;   // open slots 5 and 6 (the two open spin-1/2 methyl protons) read one half
  if ((T.halfO & kHalf) == kHalf)
    hipLaunchKernelGGL((k_pair_tile3<192, 2, 7, 8, kHalf>), dim3(T.ngroups * batch), dim3(192 * 4), 0, stream, T, in,
                       out, ld, batch, A);
  else
    hipLaunchKernelGGL((k_pair_tile3<192, 2, 7, 8, 0u>), dim3(T.ngroups * batch), dim3(192 * 4), 0, stream, T, in,
                       out, ld, batch, A);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

bool pair_tile5_supported(int NR, int KO5, int KR) { return NR == 192 && KO5 == 7 && KR == 8; }

int launch_pair_tile5(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                      hipStream_t stream) {
  if (!T.t5 || !pair_tile5_supported(T.NR, T.KO5, T.KR))
    return fail(RPM_E_UNSUPPORTED, "no closed-first-axis tiled kernel for NR=%d KO=%d KR=%d", T.NR, T.KO5, T.KR);
  if (4 * T.NLR * static_cast<long long>(sizeof(c128)) > 0x7fffffffLL)
    return fail(RPM_E_UNSUPPORTED, "vector of %lld amplitudes too large for the tiled kernel", 4 * T.NLR);
  // open slots 3-6 (the four open spin-1/2 protons H8a, H8b, H8c, Hbeta of the production layout) read one half
  constexpr unsigned kHalf5 = 0x78u;
  if ((T.halfO5 & kHalf5) == kHalf5)
    hipLaunchKernelGGL((k_pair_tile5<192, 7, 8, kHalf5>), dim3(T.ngroups5 * batch), dim3(192 * kD0), 0, stream, T, in,
                       out, ld, batch, A);
  else
    hipLaunchKernelGGL((k_pair_tile5<192, 7, 8, 0u>), dim3(T.ngroups5 * batch), dim3(192 * kD0), 0, stream, T, in,
                       out, ld, batch, A);
  RPM_LAUNCH_CHECK();
  return RPM_OK;
}

}  // namespace rpm
