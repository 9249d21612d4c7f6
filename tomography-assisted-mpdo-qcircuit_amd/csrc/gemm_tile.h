// Output-tile enumeration shared by the GEMM kernels (gemm.hip, gemm_bf16.hip).
#pragma once
#include "mpdo_internal.h"

namespace mpdo {

// Origin (m0, n0) of output tile e (blockIdx.x): tiles in row-major order (n fastest); a
// Hermitian output (BM == BN) enumerates the tiles with n-tile <= m-tile only; with a triangular
// B (tri = 1) the work grows with the n-tile, so m runs fastest: consecutive workgroups (dealt
// round-robin to the XCDs) share an n-tile and every XCD gets every n-tile.
template <int BM, int BN>
__device__ inline void gemm_tile_origin(const GemmArgs& g, int e, int& m0, int& n0) {
  if (g.herm) {
    int mt = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (mt * (mt + 1) / 2 > e) --mt;
    while ((mt + 1) * (mt + 2) / 2 <= e) ++mt;
    m0 = mt * BM;
    n0 = (e - mt * (mt + 1) / 2) * BN;
  } else if (g.tri == 1) {
    const int ntm = (g.M + BM - 1) / BM;
    m0 = (e % ntm) * BM;
    n0 = (e / ntm) * BN;
  } else {
    const int ntn = (g.N + BN - 1) / BN;
    m0 = (e / ntn) * BM;
    n0 = (e % ntn) * BN;
  }
}

// Reduction length that can be nonzero for the tile at (m0, n0): a triangular operand leaves
// only part of it.
template <int BM, int BN>
__device__ inline int gemm_tile_kend(const GemmArgs& g, int m0, int n0) {
  return g.tri == 1 ? min(g.K, n0 + BN) : g.tri == 2 ? min(g.K, m0 + BM) : g.K;
}

}  // namespace mpdo
