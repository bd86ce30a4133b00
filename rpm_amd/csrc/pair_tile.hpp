// Tables of the tiled pair stencil (see pair_tile.hip).
#pragma once

#include "rpm_common.hpp"

namespace rpm {

constexpr int kXcds = 8;         // MI355X: 8 accelerator complex dies, workgroups dispatched round-robin
constexpr int kSameE = 1;        // coefficient block has a nonzero same-electron half
constexpr int kFlipE = 2;        // ... a nonzero electron-flip half

struct TileTables {
  int NL, NR, NC, KO, KR, ngroups;
  long long NLR;
  const c128* dA;                  // [2][NL] diagonal from the left nuclei, index eA*NL + l
  const c128* dB;                  // [2][NR]
  const c128* fA;                  // [2][NL] eA flip from the left nuclei (same nuclear index)
  const c128* fB;                  // [2][NR]
  const c128* pair;                // [16]    the operator's own 4x4 electron matrix
  const int* offR;                 // [NR][KR]
  const c128* coefR;               // [NR][KR][4]
  unsigned maskR;                  // 2 bits per right slot k: bit 2k same-electron, 2k+1 electron-flip
  const c128* coefC;               // [NC][2][4]   closed left axes (spin-1/2), j = 0 is the last left axis
  unsigned maskC;                  // 2 bits per closed axis j
  const int* ofsO;                 // [ngroups][KO]  left-index offset of each open-axis neighbour
  const c128* coefO;               // [ngroups][KO][4]
  const unsigned* maskO;           // [ngroups]  2 bits per open neighbour k
  const int* packR;                // [NR][KR]  (r2 - r) & 0xffff | m << 16 : neighbour offset and digit
  const c128* coefT;               // [KR][3][4] right coefficients by (slot, digit m)
  // per-entry form of the left couplings: bit eA*2+e' of a 2x2 block's nibble is set when that entry
  // is nonzero; a half slot reads only the input electron half (eA') that its nonzero entries use
  const unsigned* nzO;             // [ngroups] one nibble per open neighbour k
  const unsigned* hselO;           // [ngroups] bit k: the input half a half slot k reads
  unsigned halfO;                  // bit k: open slot k is a half slot in every group
  unsigned nzC;                    // one nibble per closed (axis j, digit m) at index j * 2 + m
  // Closed-first-axis tile (k_pair_tile5): the first left axis (dimension kD0 = 3, the 14N5 of the
  // production layout) is closed in the tile and every other left axis is open. Group g holds the
  // rows l = g + t * ngroups5 (t < kD0) over all NR right indices.
  bool t5;                         // the tables below are valid
  int ngroups5, KO5;
  const int* ofsO5;                // [ngroups5][KO5]  left-index offset of each open neighbour
  const c128* coefO5;              // [ngroups5][KO5][4]
  const unsigned* nzO5;            // [ngroups5] one nibble per open slot
  const unsigned* hselO5;          // [ngroups5] bit k: the input half a half slot k reads
  unsigned halfO5;                 // bit k: open slot k is a half slot in every group
  const c128* coefF;               // [kD0][kD0 - 1][4]  first-axis block of digit t to its s-th other digit
  unsigned nzF;                    // one nibble per (t, s) at index t * (kD0 - 1) + s
};

constexpr int kD0 = 3;             // dimension of the closed first left axis of k_pair_tile5

int launch_pair_tile(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                     hipStream_t stream);
int launch_pair_tile3(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                      hipStream_t stream);
int launch_pair_tile5(const TileTables& T, const c128* in, c128* out, int batch, long long ld, const ApplyCtl& A,
                      hipStream_t stream);
bool pair_tile_supported(int NR, int NC, int KO, int KR);
bool pair_tile5_supported(int NR, int KO5, int KR);

}  // namespace rpm
