// Complex64 GEMM on the bf16 matrix cores of gfx950 (v_mfma_f32_16x16x32_bf16) with split
// operands: every fp32 component x of A and B is written as a sum of NS bf16 parts
//   x = p0 + p1 (+ p2) + r,   p_t = bf16(x - p0 - ... - p_{t-1})  (round to nearest),
// each residual exact in fp32, |r| <= 2^-8NS |x| (r = 0 for NS = 3: three parts hold the 24-bit
// significand), and a real product a b is taken as the sum of the part products p_s(a) p_t(b)
// with s + t < NS:
//   NS = 2 ("bf16x3"): a0 b0 + a0 b1 + a1 b0, error ~2^-16 |a||b|;
//   NS = 3 ("bf16x6"): a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0, the dropped products
//   ~3 * 2^-24 |a||b|: the level of an fp32 product.
// bf16 x bf16 products are exact in the fp32 accumulators, so the result carries fp32
// accumulation error on top. Complex products use four real products (Ar Br, Ai Bi, Ar Bi + Ai Br
// in three accumulators; Re = P_rr - P_ii at the end), which avoids the cancellation of the
// three-product form in the imaginary part.
//
// The split happens once per element as the workgroup stages a BK = 32 slice of A and B: each
// thread loads its share of the next slice and splits it in registers while the current slice
// feeds the matrix cores, then writes it to LDS, where plane c * NS + t holds part t of the real
// (c = 0) or imaginary (c = 1) components, k fastest, so that each lane's MFMA operand (8
// consecutive k of one row or column) is one 16-byte read. 256 threads, 4 waves in a 2 x 2 grid
// over the BM x BN output tile.
#include "gemm_tile.h"
#include "mpdo_internal.h"

namespace mpdo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// The NS parts of two fp32 values, two bf16 per dword (value 0 in the low half): part t is the
// bf16 rounding (to nearest) of what parts 0 .. t-1 leave. Every residual is exact in fp32, and
// three parts hold the whole 24-bit significand.
template <int NS>
__device__ inline void split_pair(float x0, float x1, unsigned* out, int stride) {
#pragma unroll
  for (int t = 0; t < NS; ++t) {
    const unsigned u = __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){x0, x1}, bf16x2));
    out[t * stride] = u;
    if (t + 1 < NS) {
      x0 -= __uint_as_float(u << 16);
      x1 -= __uint_as_float(u & 0xffff0000u);
    }
  }
}

// LDS rows of 32 bf16 (four 16-byte chunks) with chunk c of row r stored at chunk c ^ s(r % 16),
// s(q) = ((q >> 1) + (q >> 3)) & 3: the 16 rows x 16 bytes that 16 consecutive lanes read for one
// MFMA operand (rows 16 i .. 16 i + 15) then fall on distinct banks with 64 banks of 4 bytes, and
// so do any 8 consecutive lanes with 32 banks. Returns the dword offset within the row.
__device__ inline int lds_chunk(int r, int c) {
  const int q = r & 15;
  return (c ^ (((q >> 1) + (q >> 3)) & 3)) * 4;
}

// workgroups per CU the LDS allows (160 KiB per CU), at most 2 for the register budget
template <int NS, int BM, int BN>
constexpr int bf16_wg_per_cu() {
  return 2 * (2 * NS * (BM + BN) * 64) <= 160 * 1024 ? 2 : 1;
}

template <int NS, int BM, int BN>
__global__ __launch_bounds__(256, (bf16_wg_per_cu<NS, BM, BN>())) void gemm_bf16_kernel(GemmArgs g) {
  constexpr int BK = 32, LDW = BK / 2;         // dwords per LDS row: 32 bf16
  constexpr int WM = BM / 2, WN = BN / 2;      // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;    // 16 x 16 MFMA tiles per wave
  constexpr int PA = BM * BK / 256, PB = BN * BK / 256;   // elements per thread per slice
  static_assert(PA % 4 == 0 && PB % 4 == 0 && 256 % BM == 0 && 256 % BN == 0, "tile shape");
  // plane c * NS + t: part t of the real (c = 0) or imaginary (c = 1) components, k fastest
  __shared__ __attribute__((aligned(16))) unsigned As[2 * NS][BM][LDW];
  __shared__ __attribute__((aligned(16))) unsigned Bs[2 * NS][BN][LDW];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int b = blockIdx.z;
  int m0, n0;
  gemm_tile_origin<BM, BN>(g, blockIdx.x, m0, n0);
  const int ks = blockIdx.y;   // reduction slice (0 without a split)
  // operands through buffer descriptors of this batch entry and slice: an element outside the
  // logical M x K (K x N) range is given an offset past the descriptor's extent, which the
  // hardware answers with zero (no branch); byte offsets fit 31 bits (gemm_bf16_fits)
  const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
      (void*)((const float2*)g.A + (long long)b * g.bsA + ks * g.ksA), (short)0, 0x7fffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
      (void*)((const float2*)g.B + (long long)b * g.bsB + ks * g.ksB), (short)0, 0x7fffffff, 0x00020000);
  const bool aK = (g.csA == 1);   // A contiguous along k
  const bool bK = (g.rsB == 1);   // B contiguous along k
  const float sA = g.conjA ? -1.f : 1.f, sB = g.conjB ? -1.f : 1.f;
  // each thread stages PA consecutive k of one row of A (PB of one column of B): along k in
  // memory when the operand is contiguous along k, else one row per lane (coalesced along m)
  const int ra_row = aK ? tid * PA / BK : tid % BM, ra_k = aK ? tid * PA % BK : tid / BM * PA;
  const int rb_col = bK ? tid * PB / BK : tid % BN, rb_k = bK ? tid * PB % BK : tid / BN * PB;
  const bool a_in = m0 + ra_row < g.M, b_in = n0 + rb_col < g.N;
  const unsigned a_off = a_in ? (unsigned)((m0 + ra_row) * g.rsA + ra_k * g.csA) * 8u : 0u;
  const unsigned b_off = b_in ? (unsigned)((n0 + rb_col) * g.csB + rb_k * g.rsB) * 8u : 0u;
  const unsigned a_step = (unsigned)g.csA * 8u, b_step = (unsigned)g.rsB * 8u;   // bytes per k
  constexpr unsigned OUT = 0x80000000u;   // past the descriptors' extent

  floatx4 prr[TM][TN], pii[TM][TN], pim[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) { prr[i][j][q] = 0.f; pii[i][j][q] = 0.f; pim[i][j][q] = 0.f; }

  // operand slices in registers: raw complex64, then split into the planes' dwords
  float2 ra[PA], rb[PB];
  unsigned sa[2 * NS][PA / 2], sb[2 * NS][PB / 2];
  auto load = [&](int k0) {
    const unsigned ao = a_off + (unsigned)k0 * a_step, bo = b_off + (unsigned)k0 * b_step;
#pragma unroll
    for (int i = 0; i < PA; ++i) {
      const bool ok = a_in && k0 + ra_k + i < g.K;
      ra[i] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rA, ok ? ao + i * a_step : OUT, 0, 0));
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const bool ok = b_in && k0 + rb_k + i < g.K;
      rb[i] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rB, ok ? bo + i * b_step : OUT, 0, 0));
    }
  };
  auto split = [&]() {
#pragma unroll
    for (int i = 0; i < PA; i += 2) {
      split_pair<NS>(ra[i].x, ra[i + 1].x, &sa[0][i / 2], PA / 2);
      split_pair<NS>(sA * ra[i].y, sA * ra[i + 1].y, &sa[NS][i / 2], PA / 2);
    }
#pragma unroll
    for (int i = 0; i < PB; i += 2) {
      split_pair<NS>(rb[i].x, rb[i + 1].x, &sb[0][i / 2], PB / 2);
      split_pair<NS>(sB * rb[i].y, sB * rb[i + 1].y, &sb[NS][i / 2], PB / 2);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < 2 * NS; ++p) {
      if constexpr (PA % 8 == 0) {
#pragma unroll
        for (int c = 0; c < PA / 2; c += 4)
          *(uint4*)&As[p][ra_row][lds_chunk(ra_row, ra_k / 8 + c / 4)] =
              make_uint4(sa[p][c], sa[p][c + 1], sa[p][c + 2], sa[p][c + 3]);
      } else {
        *(uint2*)&As[p][ra_row][lds_chunk(ra_row, ra_k / 8) + ra_k % 8 / 2] = make_uint2(sa[p][0], sa[p][1]);
      }
      if constexpr (PB % 8 == 0) {
#pragma unroll
        for (int c = 0; c < PB / 2; c += 4)
          *(uint4*)&Bs[p][rb_col][lds_chunk(rb_col, rb_k / 8 + c / 4)] =
              make_uint4(sb[p][c], sb[p][c + 1], sb[p][c + 2], sb[p][c + 3]);
      } else {
        *(uint2*)&Bs[p][rb_col][lds_chunk(rb_col, rb_k / 8) + rb_k % 8 / 2] = make_uint2(sb[p][0], sb[p][1]);
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
      for (int p = 0; p < 2 * NS; ++p)
        fa[i][p] = __builtin_bit_cast(bf16x8, *(const uint4*)&As[p][wm + i * 16 + lr][lds_chunk(lr, lk)]);
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int p = 0; p < 2 * NS; ++p)
        fb[j][p] = __builtin_bit_cast(bf16x8, *(const uint4*)&Bs[p][wn + j * 16 + lr][lds_chunk(lr, lk)]);
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

  // slice k0 + BK is loaded and split while slice k0 feeds the matrix cores; the last round
  // loads and splits a slice past kend that is never stored (zeros past K)
  const int kend = gemm_tile_kend<BM, BN>(g, m0, n0);
  load(0);
  split();
  for (int k0 = 0; k0 < kend; k0 += BK) {
    __syncthreads();
    store();
    __syncthreads();
    load(k0 + BK);
    compute();
    split();
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

// Whether the split-bf16 kernel can address both operands: nonnegative strides and every element
// of one batch entry (and reduction slice) within 2^31 - 16 bytes of its first.
bool gemm_bf16_fits(const GemmArgs& g) {
  if (g.rsA < 0 || g.csA < 0 || g.rsB < 0 || g.csB < 0) return false;
  const long long lim = (0x7fffffffLL - 16) / 8;
  const long long kx = g.K + 31;   // the loads run up to a slice past the last one
  return (long long)(g.M - 1) * g.rsA + kx * g.csA <= lim && kx * g.rsB + (long long)(g.N - 1) * g.csB <= lim;
}

// Complex64 GEMM through split-bf16 products (nsplit = 2: three part products per real
// product, nsplit = 3: six), fp32 accumulation, complex64 output. A Hermitian output needs C
// row-major and contiguous (rsC = N, csC = 1); the operands must pass gemm_bf16_fits.
// Tile shape of the split-bf16 GEMMs: 0 = by the rules below, 1 = 64 x 64, 2 = 128 x 64,
// 3 = 64 x 128, 4 = 32 x 128, 5 = 128 x 32 (a Hermitian output always takes 64 x 64). Measured at
// the benchmark's site-contraction shapes (2048 x 8192 x 512, 256 x 16384 x 256, 8192 x 512 x 512,
// 1024 x 512 x 1024, 32 batches), three-part split: 64 x 128 ahead of 64 x 64 by 7-16 %,
// 128 x 64 and 32 x 128 behind it; at most 32 rows (16 x 4096 x 256): 32 x 128; at most 32
// columns: 128 x 32.
int g_bf16_tile = 0;

template <int NS>
static hipError_t launch_bf16(const GemmArgs& g, hipStream_t s) {
  int v = g_bf16_tile;
  if (v == 0) v = (g.M <= 32 && g.N >= 512 && !g.tri) ? 4 : (g.N <= 32 && g.M >= 512 && !g.tri) ? 5 : 3;
  if (g.herm) v = 1;
  switch (v) {
    case 2: return launch_bf16_tile<NS, 128, 64>(g, s);
    case 3: return launch_bf16_tile<NS, 64, 128>(g, s);
    case 4: return launch_bf16_tile<NS, 32, 128>(g, s);
    case 5: return launch_bf16_tile<NS, 128, 32>(g, s);
    default: return launch_bf16_tile<NS, 64, 64>(g, s);
  }
}

hipError_t gemm_bf16(int nsplit, const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return hipSuccess;
  if (g.herm && (g.rsC != g.N || g.csC != 1)) return hipErrorInvalidValue;
  if (!gemm_bf16_fits(g)) return hipErrorInvalidValue;
  if (nsplit == 2) return launch_bf16<2>(g, s);
  if (nsplit == 3) return launch_bf16<3>(g, s);
  return hipErrorInvalidValue;
}

}  // namespace mpdo

extern "C" int mpdo_debug_gemm_bf16_tiles(int v) {
  const int old = mpdo::g_bf16_tile;
  if (v >= 0 && v <= 5) mpdo::g_bf16_tile = v;
  return old;
}
