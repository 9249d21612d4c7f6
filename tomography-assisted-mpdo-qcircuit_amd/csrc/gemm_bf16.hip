// Complex64 GEMM on the bf16 matrix cores of gfx950 (v_mfma_f32_16x16x32_bf16) with split
// operands: every fp32 component x of A and B is written as a sum of NS bf16 parts
//   x = p0 + p1 (+ p2) + r,   p_t = bf16(x - p0 - ... - p_{t-1}),   |r| <= 2^-8NS |x|
// (each part exact in fp32, each residual exact), and a real product a b is taken as the sum of
// the part products p_s(a) p_t(b) with s + t < NS:
//   NS = 2 ("bf16x3"): a0 b0 + a0 b1 + a1 b0, error ~2^-16 |a||b|;
//   NS = 3 ("bf16x6"): a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0, error ~2^-24 |a||b|, the
//   level of an fp32 product.
// bf16 x bf16 products are exact in the fp32 accumulators, so the result carries fp32
// accumulation error on top. Complex products use four real products (Ar Br, Ai Bi, Ar Bi + Ai Br
// in three accumulators; Re = P_rr - P_ii at the end), which avoids the cancellation of the
// three-product form in the imaginary part.
//
// The split happens once per element as the workgroup stages a BK = 32 slice of A and B into LDS:
// plane c * NS + t holds part t of the real (c = 0) or imaginary (c = 1) component, k fastest,
// so that each lane's MFMA operand (8 consecutive k of one row or column) is one 16-byte read.
// 256 threads, 4 waves in a 2 x 2 grid over the BM x BN output tile; the next slice is loaded
// into registers while the current one feeds the matrix cores.
#include "gemm_tile.h"
#include "mpdo_internal.h"

namespace mpdo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int W> struct BfVec;
template <> struct BfVec<4> { typedef bf16x4 T; };
template <> struct BfVec<8> { typedef bf16x8 T; };

template <int NS, int BM, int BN>
__global__ __launch_bounds__(256, 2) void gemm_bf16_kernel(GemmArgs g) {
  constexpr int BK = 32, LDK = BK + 8;         // bf16 per LDS row (16 bytes of padding)
  constexpr int WM = BM / 2, WN = BN / 2;      // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;    // 16 x 16 MFMA tiles per wave
  constexpr int PA = BM * BK / 256, PB = BN * BK / 256;   // elements per thread per slice
  constexpr int WA = PA >= 8 ? 8 : 4, WB = PB >= 8 ? 8 : 4;   // bf16 per LDS write
  static_assert(PA % WA == 0 && PB % WB == 0 && 256 % BM == 0 && 256 % BN == 0, "tile shape");
  __shared__ __attribute__((aligned(16))) __bf16 As[2 * NS][BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2 * NS][BN][LDK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int b = blockIdx.z;
  int m0, n0;
  gemm_tile_origin<BM, BN>(g, blockIdx.x, m0, n0);
  const int ks = blockIdx.y;   // reduction slice (0 without a split)
  const float2* A = (const float2*)g.A + (long long)b * g.bsA + ks * g.ksA;
  const float2* B = (const float2*)g.B + (long long)b * g.bsB + ks * g.ksB;
  const bool aK = (g.csA == 1);   // A contiguous along k
  const bool bK = (g.rsB == 1);   // B contiguous along k
  const float sA = g.conjA ? -1.f : 1.f, sB = g.conjB ? -1.f : 1.f;
  // each thread stages PA consecutive k of one row of A (PB of one column of B): along k in
  // memory when the operand is contiguous along k, else one row per lane (coalesced along m)
  const int ra_row = aK ? tid * PA / BK : tid % BM, ra_k = aK ? tid * PA % BK : tid / BM * PA;
  const int rb_col = bK ? tid * PB / BK : tid % BN, rb_k = bK ? tid * PB % BK : tid / BN * PB;

  floatx4 prr[TM][TN], pii[TM][TN], pim[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) { prr[i][j][q] = 0.f; pii[i][j][q] = 0.f; pim[i][j][q] = 0.f; }

  float2 ra[PA], rb[PB];
  auto load = [&](int k0) {
    const int gm = m0 + ra_row, gn = n0 + rb_col;
#pragma unroll
    for (int i = 0; i < PA; ++i) {
      const int gk = k0 + ra_k + i;
      ra[i] = (gm < g.M && gk < g.K) ? A[(long long)gm * g.rsA + (long long)gk * g.csA] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const int gk = k0 + rb_k + i;
      rb[i] = (gn < g.N && gk < g.K) ? B[(long long)gk * g.rsB + (long long)gn * g.csB] : make_float2(0.f, 0.f);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int c0 = 0; c0 < PA; c0 += WA) {
      typename BfVec<WA>::T re[NS], im[NS];
#pragma unroll
      for (int u = 0; u < WA; ++u) {
        float xr = ra[c0 + u].x, xi = sA * ra[c0 + u].y;
#pragma unroll
        for (int t = 0; t < NS; ++t) {
          const __bf16 pr = (__bf16)xr, pi = (__bf16)xi;
          re[t][u] = pr;
          im[t][u] = pi;
          xr -= (float)pr;
          xi -= (float)pi;
        }
      }
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        *(typename BfVec<WA>::T*)&As[t][ra_row][ra_k + c0] = re[t];
        *(typename BfVec<WA>::T*)&As[NS + t][ra_row][ra_k + c0] = im[t];
      }
    }
#pragma unroll
    for (int c0 = 0; c0 < PB; c0 += WB) {
      typename BfVec<WB>::T re[NS], im[NS];
#pragma unroll
      for (int u = 0; u < WB; ++u) {
        float xr = rb[c0 + u].x, xi = sB * rb[c0 + u].y;
#pragma unroll
        for (int t = 0; t < NS; ++t) {
          const __bf16 pr = (__bf16)xr, pi = (__bf16)xi;
          re[t][u] = pr;
          im[t][u] = pi;
          xr -= (float)pr;
          xi -= (float)pi;
        }
      }
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        *(typename BfVec<WB>::T*)&Bs[t][rb_col][rb_k + c0] = re[t];
        *(typename BfVec<WB>::T*)&Bs[NS + t][rb_col][rb_k + c0] = im[t];
      }
    }
  };

  // MFMA operands: lane (lr, lk) holds row (column) lr of a 16-row (16-column) block at
  // k = 8 lk .. 8 lk + 7, the same k for A and B
  const int lr = lane & 15, lk = lane >> 4;
  auto compute = [&]() {
    bf16x8 fa[TM][2 * NS], fb[TN][2 * NS];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int p = 0; p < 2 * NS; ++p) fa[i][p] = *(const bf16x8*)&As[p][wm + i * 16 + lr][lk * 8];
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int p = 0; p < 2 * NS; ++p) fb[j][p] = *(const bf16x8*)&Bs[p][wn + j * 16 + lr][lk * 8];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int s = NS - 1; s >= 0; --s)   // smaller part products first
#pragma unroll
          for (int t = NS - 1 - s; t >= 0; --t) {
            prr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s], fb[j][t], prr[i][j], 0, 0, 0);
            pii[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][NS + s], fb[j][NS + t], pii[i][j], 0, 0, 0);
            pim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s], fb[j][NS + t], pim[i][j], 0, 0, 0);
            pim[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][NS + s], fb[j][t], pim[i][j], 0, 0, 0);
          }
  };

  const int kend = gemm_tile_kend<BM, BN>(g, m0, n0);
  load(0);
  for (int k0 = 0; k0 < kend; k0 += BK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);
    compute();
  }

  // output of a 16 x 16 f32 MFMA: D[i][j] in lane j + 16 (i / 4), item i % 4
  float2* C = (float2*)g.C + (long long)b * g.bsC + ks * g.ksC;
  const bool beta0 = (g.beta_re == 0.0 && g.beta_im == 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + i * 16 + lk * 4 + q, col = n0 + wn + j * 16 + lr;
        if (row < g.M && col < g.N) {
          const double xr = (double)prr[i][j][q] - (double)pii[i][j][q], xi = pim[i][j][q];
          double yr = g.alpha_re * xr - g.alpha_im * xi;
          double yi = g.alpha_re * xi + g.alpha_im * xr;
          const long long off = (long long)row * g.rsC + (long long)col * g.csC;
          if (!beta0) {
            const float2 o = C[off];
            yr += g.beta_re * o.x - g.beta_im * o.y;
            yi += g.beta_re * o.y + g.beta_im * o.x;
          }
          C[off] = make_float2((float)yr, (float)yi);
        }
      }
}

template <int NS, int BM, int BN>
static hipError_t launch_bf16_tile(const GemmArgs& g0, hipStream_t s) {
  GemmArgs g = g0;
  g.herm = (g0.herm && BM == BN) ? 1 : 0;
  const int split = g0.ksplit > 1 ? g0.ksplit : 1;
  if (split == 1) g.ksA = g.ksB = g.ksC = 0;
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int tiles = g.herm ? tm * (tm + 1) / 2 : tm * tn;
  hipLaunchKernelGGL((gemm_bf16_kernel<NS, BM, BN>), dim3(tiles, split, g.batch), dim3(256), 0, s, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !g.herm || split > 1) return e;
  return herm_mirror_b(g.C, false, g.N, g.bsC, g.batch, s);
}

// Complex64 GEMM through split-bf16 products (nsplit = 2: three part products per real
// product, nsplit = 3: six), fp32 accumulation, complex64 output. A Hermitian output needs C
// row-major and contiguous (rsC = N, csC = 1).
hipError_t gemm_bf16(int nsplit, const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return hipSuccess;
  if (g.herm && (g.rsC != g.N || g.csC != 1)) return hipErrorInvalidValue;
  const bool few_rows = g.M <= 32 && g.N >= 512 && !g.herm && !g.tri;
  if (nsplit == 2) {
    if (few_rows) return launch_bf16_tile<2, 32, 128>(g, s);
    if (g.herm) return launch_bf16_tile<2, 64, 64>(g, s);
    return launch_bf16_tile<2, 128, 64>(g, s);
  }
  if (nsplit == 3) {
    if (few_rows) return launch_bf16_tile<3, 32, 128>(g, s);
    return launch_bf16_tile<3, 64, 64>(g, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mpdo
