// Memory-bound kernels of the MPDO update path (gfx950): Kraus absorption into the inner index,
// the two-qubit channel on the QR-reduced core, tensor permutations, the truncation rule of
// decompositions.svd, and the conversion of fp64 eigenvectors into complex64 factors.
#include "site_ops.h"
#include "prof.h"

namespace mpdo {

static int grid1(long long tot, int cap = 8192) {
  long long b = (tot + 255) / 256;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}


// ---------------------------------------------------------------------------------------
// 1q gate / Kraus absorption (Circuit.py:138-178):
//   T'[l, s', a*K + k, r] = sum_s E_k[s', s] T[l, s, a, r]
// One thread per (l, a, r) input column pair; writes 2K outputs, coalesced along r.
// ---------------------------------------------------------------------------------------
__global__ void apply_1q_kernel(const Site1q* sites, const float2* kraus) {
  const Site1q S = sites[blockIdx.y];
  const long long tot = (long long)S.l * S.a * S.r;
  const float2* E = kraus + S.koff;          // [k][2][2]
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int r = (int)(e % S.r);
    long long t = e / S.r;
    int a = (int)(t % S.a);
    int l = (int)(t / S.a);
    const float2* in = S.in + ((long long)l * 2 * S.a + a) * S.r + r;
    float2 x0 = in[0], x1 = in[(long long)S.a * S.r];
    const long long aK = (long long)S.a * S.K;
    float2* out = S.out + ((long long)l * 2 * aK + (long long)a * S.K) * S.r + r;
    for (int k = 0; k < S.K; ++k) {
      float2 e00 = E[k * 4 + 0], e01 = E[k * 4 + 1], e10 = E[k * 4 + 2], e11 = E[k * 4 + 3];
      float2 y0, y1;
      y0.x = e00.x * x0.x - e00.y * x0.y + e01.x * x1.x - e01.y * x1.y;
      y0.y = e00.x * x0.y + e00.y * x0.x + e01.x * x1.y + e01.y * x1.x;
      y1.x = e10.x * x0.x - e10.y * x0.y + e11.x * x1.x - e11.y * x1.y;
      y1.y = e10.x * x0.y + e10.y * x0.x + e11.x * x1.y + e11.y * x1.x;
      out[(long long)k * S.r] = y0;
      out[(aK + k) * (long long)S.r] = y1;
    }
  }
}

hipError_t apply_1q(int nsites, const Site1q* d_sites, const float2* d_kraus, long long max_elems,
                    hipStream_t s) {
  if (nsites <= 0) return hipSuccess;
  int blocks = (int)((max_elems + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(apply_1q_kernel, dim3(blocks, nsites), dim3(256), 0, s, d_sites, d_kraus);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Two-qubit channel on the reduced core (Circuit.py:87-106):
//   C[(p*2 + s1'), ((s2'*K + k)*P2 + q)] = sum_{s1,s2} G[s1', s2', s1, s2, k] Y[s1][s2][p][q]
// ---------------------------------------------------------------------------------------
__global__ void channel2q_kernel(const float2* Y, const float2* G, int K, int P1, int P2, float2* C, long long bsY,
                                 long long bsC) {
  Y += blockIdx.y * bsY;
  C += blockIdx.y * bsC;
  const long long tot = (long long)P1 * P2;
  const long long ldc = 2LL * K * P2;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int p = (int)(e / P2), q = (int)(e % P2);
    float2 y[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = Y[(long long)i * tot + e];
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int o = 0; o < 4; ++o) {   // o = s1'*2 + s2'
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int i = 0; i < 4; ++i) {   // i = s1*2 + s2
          float2 g = G[((o * 4 + i) * K) + k];
          acc.x += g.x * y[i].x - g.y * y[i].y;
          acc.y += g.x * y[i].y + g.y * y[i].x;
        }
        int s1 = o >> 1, s2 = o & 1;
        C[((long long)p * 2 + s1) * ldc + ((long long)s2 * K + k) * P2 + q] = acc;
      }
    }
  }
}

hipError_t channel2q(const float2* Y, const float2* G, int K, int P1, int P2, float2* C, hipStream_t s, int batch,
                     long long bsY, long long bsC) {
  long long tot = (long long)P1 * P2;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(channel2q_kernel, dim3(blocks, batch), dim3(256), 0, s, Y, G, K, P1, P2, C, bsY, bsC);
  return hipGetLastError();
}

// out[b][r][r'] = sum_a M[b][(a R + r)][(a R + r')]  (sum of the diagonal R x R blocks; complex64 in,
// complex128 out)
__global__ void block_diag_sum_kernel(const float2* M, int A, int R, double2* out, long long bsM, long long bsO) {
  M += blockIdx.y * bsM;
  out += blockIdx.y * bsO;
  const long long ld = (long long)A * R;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)R * R;
       e += (long long)gridDim.x * blockDim.x) {
    int r = (int)(e / R), q = (int)(e % R);
    double2 acc = make_double2(0, 0);
    for (int a = 0; a < A; ++a) {
      float2 v = M[((long long)a * R + r) * ld + (long long)a * R + q];
      acc.x += v.x;
      acc.y += v.y;
    }
    out[e] = acc;
  }
}

hipError_t block_diag_sum(const float2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch,
                          hipStream_t s) {
  hipLaunchKernelGGL(block_diag_sum_kernel, dim3(grid1((long long)R * R), batch), dim3(256), 0, s, M, A, R, out, bsM,
                     bsO);
  return hipGetLastError();
}

// the same with a complex128 input
__global__ void block_diag_sum64_kernel(const double2* M, int A, int R, double2* out, long long bsM, long long bsO) {
  M += blockIdx.y * bsM;
  out += blockIdx.y * bsO;
  const long long ld = (long long)A * R;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)R * R;
       e += (long long)gridDim.x * blockDim.x) {
    int r = (int)(e / R), q = (int)(e % R);
    double2 acc = make_double2(0, 0);
    for (int a = 0; a < A; ++a) {
      double2 v = M[((long long)a * R + r) * ld + (long long)a * R + q];
      acc.x += v.x;
      acc.y += v.y;
    }
    out[e] = acc;
  }
}

hipError_t block_diag_sum64(const double2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch,
                            hipStream_t s) {
  hipLaunchKernelGGL(block_diag_sum64_kernel, dim3(grid1((long long)R * R), batch), dim3(256), 0, s, M, A, R, out,
                     bsM, bsO);
  return hipGetLastError();
}

// out[b][i][j] = V[b][i][j] for j < k[b] (a per-matrix count held on the device), else 0;
// complex128 in, complex64 out
__global__ void mask_cols_kernel(const double2* V, long long ldv, long long bsV, int rows, int cols, const int* k,
                                 float2* out, long long ldo, long long bsO) {
  const int b = blockIdx.y;
  V += b * bsV;
  out += b * bsO;
  const int kb = k[b];
  const long long tot = (long long)rows * cols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int i = (int)(e / cols), j = (int)(e % cols);
    double2 v = V[(long long)i * ldv + j];
    out[(long long)i * ldo + j] = j < kb ? make_float2((float)v.x, (float)v.y) : make_float2(0.f, 0.f);
  }
}

hipError_t mask_cols_b(const double2* V, long long ldv, long long bsV, int rows, int cols, const int* k, float2* out,
                       long long ldo, long long bsO, int batch, hipStream_t s) {
  hipLaunchKernelGGL(mask_cols_kernel, dim3(grid1((long long)rows * cols), batch), dim3(256), 0, s, V, ldv, bsV, rows,
                     cols, k, out, ldo, bsO);
  return hipGetLastError();
}

// complex64 -> complex128 copy of a batch of contiguous blocks
__global__ void widen_kernel(const float2* in, double2* out, long long n, long long bsI, long long bsO) {
  in += blockIdx.y * bsI;
  out += blockIdx.y * bsO;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    float2 v = in[e];
    out[e] = make_double2(v.x, v.y);
  }
}

hipError_t widen(const float2* in, double2* out, long long n, long long bsI, long long bsO, int batch, hipStream_t s) {
  hipLaunchKernelGGL(widen_kernel, dim3(grid1(n), batch), dim3(256), 0, s, in, out, n, bsI, bsO);
  return hipGetLastError();
}

// Gram of the two-site core C = G(Y) without forming C's Kraus columns. With
// C[(p, o1), (o2, k, q)] = sum_i G[(o1 o2), i, k] Y[i][p][q] (i = input pair (s1 s2)),
// C C^H [(p,o1),(p',o1')] = sum_{i,i'} E[o1, i, o1', i'] Pg[(i, p), (i', p')] with
// E = sum_{o2,k} G[(o1 o2), i, k] conj(G[(o1' o2), i', k]) and Pg = Ycat Ycat^H (Ycat the four
// P1 x P2 blocks of Y stacked, 4 P1 x 4 P1, complex128). Gt: 16 K complex64, [(o1 o2)][i][k].
__global__ void kraus_gram_kernel(const double2* Pg, const float2* Gt, int K, int P1, double2* out, long long bsP,
                                  long long bsO) {
  __shared__ double2 E[64];   // [o1][i][o1'][i']
  Pg += blockIdx.y * bsP;
  out += blockIdx.y * bsO;
  if (threadIdx.x < 64) {
    const int o1 = threadIdx.x >> 5, i = (threadIdx.x >> 3) & 3, o1p = (threadIdx.x >> 2) & 1, ip = threadIdx.x & 3;
    double2 acc = make_double2(0.0, 0.0);
    for (int o2 = 0; o2 < 2; ++o2)
      for (int k = 0; k < K; ++k) {
        const float2 g = Gt[((o1 * 2 + o2) * 4 + i) * K + k];
        const float2 h = Gt[((o1p * 2 + o2) * 4 + ip) * K + k];
        acc.x += (double)g.x * h.x + (double)g.y * h.y;
        acc.y += (double)g.y * h.x - (double)g.x * h.y;
      }
    E[threadIdx.x] = acc;
  }
  __syncthreads();
  const long long n = 2LL * P1, ldp = 4LL * P1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n * n; e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / n), c = (int)(e % n);
    const int p = r >> 1, o1 = r & 1, pp = c >> 1, o1p = c & 1;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int ip = 0; ip < 4; ++ip) {
        const double2 w = E[((o1 * 4 + i) * 2 + o1p) * 4 + ip];
        const double2 v = Pg[((long long)i * P1 + p) * ldp + (long long)ip * P1 + pp];
        acc.x += w.x * v.x - w.y * v.y;
        acc.y += w.x * v.y + w.y * v.x;
      }
    out[e] = acc;
  }
}

hipError_t kraus_gram(const double2* Pg, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s) {
  const long long n = 2LL * P1;
  hipLaunchKernelGGL(kraus_gram_kernel, dim3(grid1(n * n), batch), dim3(256), 0, s, Pg, Gt, K, P1, out,
                     16LL * P1 * P1, n * n);
  return hipGetLastError();
}

// The dual contraction: for C as in kraus_gram and N (2 P1 x 2 P1, rows (p, o1)),
// sum_{o2,k} C_{o2 k}^H N C_{o2 k} = Ycat^H Nt Ycat with
// Nt[(i, p), (i', p')] = sum_{o1,o1'} conj(E[o1, i, o1', i']) N[(p, o1), (p', o1')]  (4 P1 x 4 P1).
// One thread per (p, p'): its four N entries serve all sixteen (i, i') blocks.
__global__ void kraus_dual_kernel(const double2* N, const float2* Gt, int K, int P1, double2* out, long long bsN,
                                  long long bsO) {
  __shared__ double2 E[64];   // [o1][i][o1'][i'], as in kraus_gram_kernel
  N += blockIdx.y * bsN;
  out += blockIdx.y * bsO;
  if (threadIdx.x < 64) {
    const int o1 = threadIdx.x >> 5, i = (threadIdx.x >> 3) & 3, o1p = (threadIdx.x >> 2) & 1, ip = threadIdx.x & 3;
    double2 acc = make_double2(0.0, 0.0);
    for (int o2 = 0; o2 < 2; ++o2)
      for (int k = 0; k < K; ++k) {
        const float2 g = Gt[((o1 * 2 + o2) * 4 + i) * K + k];
        const float2 h = Gt[((o1p * 2 + o2) * 4 + ip) * K + k];
        acc.x += (double)g.x * h.x + (double)g.y * h.y;
        acc.y += (double)g.y * h.x - (double)g.x * h.y;
      }
    E[threadIdx.x] = acc;
  }
  __syncthreads();
  const long long ldn = 2LL * P1, ldo = 4LL * P1, pp = (long long)P1 * P1;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < pp; e += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(e / P1), q = (int)(e % P1);
    double2 nv[2][2];
#pragma unroll
    for (int o1 = 0; o1 < 2; ++o1)
#pragma unroll
      for (int o1p = 0; o1p < 2; ++o1p) nv[o1][o1p] = N[(2LL * p + o1) * ldn + 2LL * q + o1p];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int ip = 0; ip < 4; ++ip) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int o1 = 0; o1 < 2; ++o1)
#pragma unroll
          for (int o1p = 0; o1p < 2; ++o1p) {
            const double2 w = E[((o1 * 4 + i) * 2 + o1p) * 4 + ip];   // conj(w) * n
            const double2 v = nv[o1][o1p];
            acc.x += w.x * v.x + w.y * v.y;
            acc.y += w.x * v.y - w.y * v.x;
          }
        out[((long long)i * P1 + p) * ldo + (long long)ip * P1 + q] = acc;
      }
  }
}

hipError_t kraus_dual(const double2* N, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s) {
  const long long pp = (long long)P1 * P1;
  hipLaunchKernelGGL(kraus_dual_kernel, dim3(grid1(pp), batch), dim3(256), 0, s, N, Gt, K, P1, out, 4 * pp, 16 * pp);
  return hipGetLastError();
}

// out[(p, s), (p', s')] = delta_{s s'} R[p, p'] (2 P x 2 Q, complex128) from R (P x Q, complex64)
__global__ void kron_i2_kernel(const float2* R, int P, int Q, double2* out, long long bsR, long long bsO) {
  R += blockIdx.y * bsR;
  out += blockIdx.y * bsO;
  const long long tot = 4LL * P * Q;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const long long row = e / (2LL * Q), col = e % (2LL * Q);
    const float2 v = R[(row >> 1) * Q + (col >> 1)];
    out[e] = ((row & 1) == (col & 1)) ? make_double2(v.x, v.y) : make_double2(0.0, 0.0);
  }
}

hipError_t kron_i2(const float2* R, int P, int Q, double2* out, int batch, hipStream_t s) {
  const long long tot = 4LL * P * Q;
  hipLaunchKernelGGL(kron_i2_kernel, dim3(grid1(tot), batch), dim3(256), 0, s, R, P, Q, out, (long long)P * Q, tot);
  return hipGetLastError();
}

// Kappa Gram of a new max site held in Kraus-factored form. The site is
// T[k, s2, K2, a2, r] = sum_{oi} g[s2][oi][K2] W[k, oi, a2, r] with oi = (o1, i), i the pair's
// input physical index, g[s2][(o1,i)][K2] = Gt[((o1 2 + s2) 4 + i) K + K2]; given
// Hg = Wp Wp^H (Wp = W as rows (oi, a2) over columns (k, r); 8 A x 8 A, complex128), its Gram over
// the inner index (a2, K2) is
// out[(a2 K2), (a2' K2')] = sum_{s2, oi, oi'} conj(g[s2][oi][K2]) g[s2][oi'][K2'] conj(Hg[(oi a2), (oi' a2')]).
__global__ void kappa_core_gram_kernel(const double2* Hg, const float2* Gt, int K, int A, double2* out, long long bsH,
                                       long long bsO) {
  extern __shared__ double2 gs[];   // [s2][oi][K2]
  Hg += blockIdx.y * bsH;
  out += blockIdx.y * bsO;
  for (int t = threadIdx.x; t < 16 * K; t += blockDim.x) {
    const int s2 = t / (8 * K), oi = (t / K) % 8, k2 = t % K;
    const int o1 = oi >> 2, i = oi & 3;
    const float2 v = Gt[((o1 * 2 + s2) * 4 + i) * K + k2];
    gs[t] = make_double2(v.x, v.y);
  }
  __syncthreads();
  const long long n = (long long)A * K, ldh = 8LL * A;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n * n; e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / n), c = (int)(e % n);
    const int a2 = r / K, k2 = r % K, a2p = c / K, k2p = c % K;
    double2 acc = make_double2(0.0, 0.0);
    for (int s2 = 0; s2 < 2; ++s2)
      for (int oi = 0; oi < 8; ++oi) {
        const double2 g1 = gs[(s2 * 8 + oi) * K + k2];
        double2 inner = make_double2(0.0, 0.0);   // sum_oi' g[s2][oi'][K2'] conj(Hg[(oi a2), (oi' a2')])
        for (int oj = 0; oj < 8; ++oj) {
          const double2 g2 = gs[(s2 * 8 + oj) * K + k2p];
          const double2 h = Hg[((long long)oi * A + a2) * ldh + (long long)oj * A + a2p];
          inner.x += g2.x * h.x + g2.y * h.y;
          inner.y += g2.y * h.x - g2.x * h.y;
        }
        acc.x += g1.x * inner.x + g1.y * inner.y;   // conj(g1) * inner
        acc.y += g1.x * inner.y - g1.y * inner.x;
      }
    out[e] = acc;
  }
}

// The same Gram one K x K block (a2, a2') per workgroup, through the intermediate
// U[s2][oi][K2'] = sum_oi' g[s2][oi'][K2'] conj(Hg[(oi a2), (oi' a2')]) held in LDS:
// out[(a2 K2), (a2' K2')] = sum_{s2, oi} conj(g[s2][oi][K2]) U[s2][oi][K2'] (a fifth of the
// arithmetic of the direct sum, every operand read once). K <= 16.
__global__ __launch_bounds__(256) void kappa_core_gram_blk_kernel(const double2* Hg, const float2* Gt, int K, int A,
                                                                  double2* out, long long bsH, long long bsO) {
  __shared__ double2 g[2][8][16];
  __shared__ double2 H[8][9];
  __shared__ double2 U[2][8][16];
  const int a2 = blockIdx.x / A, a2p = blockIdx.x % A, b = blockIdx.y, tid = threadIdx.x;
  Hg += (long long)b * bsH;
  out += (long long)b * bsO;
  const long long ldh = 8LL * A;
  if (tid < 16 * K) {
    const int s2 = tid / (8 * K), oi = (tid / K) % 8, k2 = tid % K;
    const int o1 = oi >> 2, i = oi & 3;
    const float2 v = Gt[((o1 * 2 + s2) * 4 + i) * K + k2];
    g[s2][oi][k2] = make_double2(v.x, v.y);
  }
  if (tid < 64) {
    const int oi = tid >> 3, oj = tid & 7;
    H[oi][oj] = Hg[((long long)oi * A + a2) * ldh + (long long)oj * A + a2p];
  }
  __syncthreads();
  if (tid < 16 * K) {
    const int s2 = tid / (8 * K), oi = (tid / K) % 8, k2p = tid % K;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int oj = 0; oj < 8; ++oj) {   // g[s2][oj][k2p] conj(H[oi][oj])
      const double2 x = g[s2][oj][k2p], h = H[oi][oj];
      acc.x += x.x * h.x + x.y * h.y;
      acc.y += x.y * h.x - x.x * h.y;
    }
    U[s2][oi][k2p] = acc;
  }
  __syncthreads();
  if (tid < K * K) {
    const int k2 = tid / K, k2p = tid % K;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
      for (int oi = 0; oi < 8; ++oi) {   // conj(g[s2][oi][k2]) U[s2][oi][k2p]
        const double2 x = g[s2][oi][k2], u = U[s2][oi][k2p];
        acc.x += x.x * u.x + x.y * u.y;
        acc.y += x.x * u.y - x.y * u.x;
      }
    const long long n = (long long)A * K;
    out[((long long)a2 * K + k2) * n + (long long)a2p * K + k2p] = acc;
  }
}

hipError_t kappa_core_gram(const double2* Hg, const float2* Gt, int K, int A, double2* out, int batch, hipStream_t s) {
  const long long n = (long long)A * K;
  if (K <= 16) {
    hipLaunchKernelGGL(kappa_core_gram_blk_kernel, dim3(A * A, batch), dim3(256), 0, s, Hg, Gt, K, A, out, 64LL * A * A,
                       n * n);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(kappa_core_gram_kernel, dim3(grid1(n * n), batch), dim3(256), sizeof(double2) * 16 * K, s, Hg,
                     Gt, K, A, out, 64LL * A * A, n * n);
  return hipGetLastError();
}

// The kept inner directions mapped through the Kraus factorisation:
// Vt[(oi a2), (s2 j)] = sum_K2 g[s2][oi][K2] Vk[(a2 K2), j]  (Vk: A K x kk, Vt: 8 A x 2 kk, complex64)
__global__ void kappa_core_vec_kernel(const float2* Vk, const float2* Gt, int K, int A, int kk, float2* Vt, long long bsV,
                                      long long bsT) {
  Vk += blockIdx.y * bsV;
  Vt += blockIdx.y * bsT;
  const long long tot = 8LL * A * 2 * kk;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    const int row = (int)(e / (2 * kk)), col = (int)(e % (2 * kk));
    const int oi = row / A, a2 = row % A, s2 = col / kk, j = col % kk;
    const int o1 = oi >> 2, i = oi & 3;
    double ax = 0.0, ay = 0.0;
    for (int k2 = 0; k2 < K; ++k2) {
      const float2 g = Gt[((o1 * 2 + s2) * 4 + i) * K + k2];
      const float2 v = Vk[((long long)a2 * K + k2) * kk + j];
      ax += (double)g.x * v.x - (double)g.y * v.y;
      ay += (double)g.x * v.y + (double)g.y * v.x;
    }
    Vt[e] = make_float2((float)ax, (float)ay);
  }
}

hipError_t kappa_core_vec(const float2* Vk, const float2* Gt, int K, int A, int kk, float2* Vt, int batch,
                          hipStream_t s) {
  const long long tot = 8LL * A * 2 * kk;
  hipLaunchKernelGGL(kappa_core_vec_kernel, dim3(grid1(tot), batch), dim3(256), 0, s, Vk, Gt, K, A, kk, Vt,
                     (long long)A * K * kk, tot);
  return hipGetLastError();
}

// Complex64 copies of a factor stored by columns (L[v n + i] = L_iv): mode 0 out[u][j] = conj(L_ju)
// (the factor's conjugate transpose, row-major), mode 1 out[i][v] = L_iv (row-major)
__global__ void narrow_factor_kernel(const double2* L, int n, int mode, float2* out) {
  const long long nn = (long long)n * n;
  L += blockIdx.y * nn;
  out += blockIdx.y * nn;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nn; e += (long long)gridDim.x * blockDim.x) {
    if (mode == 0) {
      const double2 v = L[e];
      out[e] = make_float2((float)v.x, -(float)v.y);
    } else {
      const long long i = e / n, v = e % n;
      const double2 x = L[v * n + i];
      out[e] = make_float2((float)x.x, (float)x.y);
    }
  }
}

hipError_t narrow_factor(const double2* L, int n, int mode, float2* out, int batch, hipStream_t s) {
  hipLaunchKernelGGL(narrow_factor_kernel, dim3(grid1((long long)n * n), batch), dim3(256), 0, s, L, n, mode, out);
  return hipGetLastError();
}

// out[b][e] = sum_s in[b S + s][e], e < n (partial products of a split reduction)
__global__ void sum_parts_kernel(const double2* in, int S, long long n, double2* out) {
  const int b = blockIdx.y;
  const double2* src = in + (long long)b * S * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int t = 0; t < S; ++t) {
      const double2 v = src[(long long)t * n + e];
      acc.x += v.x;
      acc.y += v.y;
    }
    out[(long long)b * n + e] = acc;
  }
}

hipError_t sum_parts(const double2* in, int S, long long n, double2* out, int batch, hipStream_t s) {
  hipLaunchKernelGGL(sum_parts_kernel, dim3(grid1(n), batch), dim3(256), 0, s, in, S, n, out);
  return hipGetLastError();
}

// batched identity
__global__ void eye_b_kernel(float2* out, int rows, int cols, long long bs) {
  out += blockIdx.y * bs;
  long long tot = (long long)rows * cols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int i = (int)(e / cols), j = (int)(e % cols);
    out[e] = make_float2(i == j ? 1.f : 0.f, 0.f);
  }
}

hipError_t eye_b(float2* out, int rows, int cols, long long bs, int batch, hipStream_t s) {
  hipLaunchKernelGGL(eye_b_kernel, dim3(grid1((long long)rows * cols), batch), dim3(256), 0, s, out, rows, cols, bs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Permutation of a row-major tensor of rank <= 6 (complex64):
//   out[idx permuted by perm] = in[idx]; out dims are dims[perm[0]], dims[perm[1]], ...
// ---------------------------------------------------------------------------------------
__global__ void permute_kernel(const float2* in, float2* out, PermDesc d, long long tot, long long bstride) {
  in += blockIdx.y * bstride;
  out += blockIdx.y * bstride;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    // decompose the OUTPUT index, gather from the input (coalesced writes)
    long long t = e, src = 0;
    for (int i = d.rank - 1; i >= 0; --i) {
      int c = (int)(t % d.odims[i]);
      t /= d.odims[i];
      src += (long long)c * d.istride[i];
    }
    out[e] = in[src];
  }
}

hipError_t permute(const float2* in, float2* out, int rank, const int* dims, const int* perm, hipStream_t s,
                   int batch, long long bstride) {
  PermDesc d;
  d.rank = rank;
  long long st[8];
  st[rank - 1] = 1;
  for (int i = rank - 2; i >= 0; --i) st[i] = st[i + 1] * dims[i + 1];
  long long tot = 1;
  for (int i = 0; i < rank; ++i) {
    d.odims[i] = dims[perm[i]];
    d.istride[i] = st[perm[i]];
    tot *= dims[i];
  }
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  prof::Scope ps(prof::PERMUTE, s, 0.0, 16.0 * (double)tot * batch);
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    int nb = batch - b0 < 65535 ? batch - b0 : 65535;
    hipLaunchKernelGGL(permute_kernel, dim3(blocks, nb), dim3(256), 0, s, in + b0 * bstride, out + b0 * bstride, d,
                       tot, bstride);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// decompositions.svd keep-count (decompositions.py:116-133) on s = sqrt(max(lambda, 0)),
// lambda the descending Gram eigenvalues; entries past `nev` are zero singular values.
// Writes k and the discarded weight sum_{j >= k} s_j^2 (the reference computes and drops it).
// ---------------------------------------------------------------------------------------
__global__ void keep_count_kernel(const double* lam, long long ldw, int nev, int len_s, int msv,
                                  double err, int relative, int* k_out, double* disc_out, int batch) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const double* L = lam + (long long)b * ldw;
  int n = len_s;
  int maxk = (msv <= 0 || msv > n) ? n : msv;   // msv <= 0: no cap (None)
  int kerr = maxk;
  if (err >= 0 && n > 0) {
    double tot = 0;
    for (int j = 0; j < nev && j < n; ++j) tot += fmax(L[j], 0.0);
    double te_last = sqrt(tot);
    double s0 = sqrt(fmax(L[0], 0.0));
    double bound = relative ? err * s0 : err;
    double cum = 0;
    for (int i = 0; i < n - 1; ++i) {
      cum += (i < nev) ? fmax(L[i], 0.0) : 0.0;
      if (te_last - sqrt(cum) <= bound) { kerr = i + 1; break; }
    }
  }
  int k = kerr < maxk ? kerr : maxk;
  double d = 0;
  for (int j = k; j < nev && j < n; ++j) d += fmax(L[j], 0.0);
  k_out[b] = k;
  disc_out[b] = d;
}

hipError_t keep_count(const double* lam, long long ldw, int nev, int len_s, int msv, double err, int relative,
                      int* k_out, double* disc_out, int batch, hipStream_t s) {
  hipLaunchKernelGGL(keep_count_kernel, dim3((batch + 63) / 64), dim3(64), 0, s, lam, ldw, nev, len_s, msv, err,
                     relative, k_out, disc_out, batch);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Factor from an eigendecomposition (V columns = eigenvectors, lambda descending):
//   trans = 0:  out[i*ldo + j] = V[i][j] * f(lambda_j)           i < rows, j < k
//   trans = 1:  out[j*ldo + i] = conj(V[i][j]) * f(lambda_j)
// with f(lambda) = lambda^power when lambda > floor_rel * lambda_0 (else 0), and zero for
// j >= nev (padded zero singular values).
// ---------------------------------------------------------------------------------------
template <typename TO>
__global__ void factor_kernel(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL,
                              int rows, int nev, int k, double power, double floor_rel, int trans, TO* out,
                              long long ldo, long long bsO) {
  const long long tot = (long long)rows * k;
  V += blockIdx.y * bsV;
  lam += blockIdx.y * bsL;
  out += blockIdx.y * bsO;
  const double l0 = lam[0] > 0 ? lam[0] : 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int i, j;
    if (trans) { j = (int)(e / rows); i = (int)(e % rows); }
    else { i = (int)(e / k); j = (int)(e % k); }
    TO o;
    o.x = 0; o.y = 0;
    if (j < nev) {
      double lj = lam[j];
      if (lj > floor_rel * l0 && lj > 0) {
        double f = (power == 0.0) ? 1.0 : pow(lj, power);
        double2 v = V[(long long)i * ldv + j];
        o.x = v.x * f;
        o.y = (trans ? -v.y : v.y) * f;
      }
    }
    if (trans) out[(long long)j * ldo + i] = o;
    else out[(long long)i * ldo + j] = o;
  }
}

hipError_t factor(const double2* V, long long ldv, const double* lam, int rows, int nev, int k, double power,
                  double floor_rel, int trans, float2* out, long long ldo, hipStream_t s) {
  return factor_b(V, ldv, 0, lam, 0, rows, nev, k, power, floor_rel, trans, out, ldo, 0, 1, s);
}

hipError_t factor64(const double2* V, long long ldv, const double* lam, int rows, int nev, int k, double power,
                    double floor_rel, int trans, double2* out, long long ldo, hipStream_t s) {
  return factor64_b(V, ldv, 0, lam, 0, rows, nev, k, power, floor_rel, trans, out, ldo, 0, 1, s);
}

hipError_t factor_b(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL, int rows,
                    int nev, int k, double power, double floor_rel, int trans, float2* out, long long ldo,
                    long long bsO, int batch, hipStream_t s) {
  long long tot = (long long)rows * k;
  if (tot <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(factor_kernel<float2>, dim3(grid1(tot), batch), dim3(256), 0, s, V, ldv, bsV, lam, bsL, rows,
                     nev, k, power, floor_rel, trans, out, ldo, bsO);
  return hipGetLastError();
}

hipError_t factor64_b(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL, int rows,
                      int nev, int k, double power, double floor_rel, int trans, double2* out, long long ldo,
                      long long bsO, int batch, hipStream_t s) {
  long long tot = (long long)rows * k;
  if (tot <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(factor_kernel<double2>, dim3(grid1(tot), batch), dim3(256), 0, s, V, ldv, bsV, lam, bsL, rows,
                     nev, k, power, floor_rel, trans, out, ldo, bsO);
  return hipGetLastError();
}

// identity into a (rows x cols) complex64 row-major matrix
__global__ void eye_kernel(float2* out, int rows, int cols) {
  long long tot = (long long)rows * cols;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int i = (int)(e / cols), j = (int)(e % cols);
    out[e] = make_float2(i == j ? 1.f : 0.f, 0.f);
  }
}

hipError_t eye(float2* out, int rows, int cols, hipStream_t s) {
  long long tot = (long long)rows * cols;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(eye_kernel, dim3(blocks), dim3(256), 0, s, out, rows, cols);
  return hipGetLastError();
}

// P (2x2) applied to the physical index of X[l][s][rest]: out[l][s'][rest] = sum_s P[s'][s] X[l][s][rest]
__global__ void phys_op_kernel(const float2* X, float2* out, int L, long long rest, float2 p00, float2 p01,
                               float2 p10, float2 p11) {
  long long tot = (long long)L * rest;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    long long l = e / rest, c = e % rest;
    float2 x0 = X[(l * 2) * rest + c], x1 = X[(l * 2 + 1) * rest + c];
    float2 y0, y1;
    y0.x = p00.x * x0.x - p00.y * x0.y + p01.x * x1.x - p01.y * x1.y;
    y0.y = p00.x * x0.y + p00.y * x0.x + p01.x * x1.y + p01.y * x1.x;
    y1.x = p10.x * x0.x - p10.y * x0.y + p11.x * x1.x - p11.y * x1.y;
    y1.y = p10.x * x0.y + p10.y * x0.x + p11.x * x1.y + p11.y * x1.x;
    out[(l * 2) * rest + c] = y0;
    out[(l * 2 + 1) * rest + c] = y1;
  }
}

hipError_t phys_op(const float2* X, float2* out, int L, long long rest, const float2* P, hipStream_t s) {
  long long tot = (long long)L * rest;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(phys_op_kernel, dim3(blocks), dim3(256), 0, s, X, out, L, rest, P[0], P[1], P[2], P[3]);
  return hipGetLastError();
}

// sum_{i,j} A[i][j] B[j][i] (complex, fp64 accumulation), one workgroup
__global__ __launch_bounds__(1024) void trace_prod_kernel(const float2* A, const float2* B, int n, double2* out) {
  __shared__ double2 red[1024];
  double2 acc = make_double2(0, 0);
  const long long tot = (long long)n * n;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    int i = (int)(e / n), j = (int)(e % n);
    float2 x = A[e], y = B[(long long)j * n + i];
    acc.x += (double)x.x * y.x - (double)x.y * y.y;
    acc.y += (double)x.x * y.y + (double)x.y * y.x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[threadIdx.x].x += red[threadIdx.x + st].x;
      red[threadIdx.x].y += red[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

hipError_t trace_prod(const float2* A, const float2* B, int n, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(trace_prod_kernel, dim3(1), dim3(1024), 0, s, A, B, n, out);
  return hipGetLastError();
}

}  // namespace mpdo

namespace mpdo {

// ---------------------------------------------------------------- Pauli strings, many at once
// E[t][0] = 1 for t < count (the left boundary environment of every string)
__global__ void env_one_kernel(float2* E, long long bs, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) E[(long long)t * bs] = make_float2(1.f, 0.f);
}

hipError_t env_one(float2* E, long long bs, int count, hipStream_t s) {
  hipLaunchKernelGGL(env_one_kernel, dim3((count + 255) / 256), dim3(256), 0, s, E, bs, count);
  return hipGetLastError();
}

// String t's operator at this site applied to the physical index of X[t][l][s][rest]:
// out[t][l][s'][rest] = sum_s P[s'][s] X[t][l][s][rest], P from codes[t * n + site] (0 X, 1 Y,
// 2 Z, 3 |0><0|, 4 |1><1|; -1 the identity: a copy)
__global__ void phys_op_codes_kernel(const float2* X, float2* out, int L, long long rest, long long bs, int count,
                                     const int* codes, int n, int site) {
  const long long per = (long long)L * rest, tot = per * count;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int t = (int)(e / per);
    const long long w = e % per, l = w / rest, c = w % rest;
    const int code = codes[(long long)t * n + site];
    const float2 x0 = X[t * bs + (l * 2) * rest + c], x1 = X[t * bs + (l * 2 + 1) * rest + c];
    float2 y0 = x0, y1 = x1;
    switch (code) {
      case 0: y0 = x1; y1 = x0; break;                                                  // X
      case 1: y0 = make_float2(x1.y, -x1.x); y1 = make_float2(-x0.y, x0.x); break;     // Y: (-i x1, i x0)
      case 2: y1 = make_float2(-x1.x, -x1.y); break;                                    // Z
      case 3: y1 = make_float2(0.f, 0.f); break;                                        // |0><0|
      case 4: y0 = make_float2(0.f, 0.f); break;                                        // |1><1|
      default: break;
    }
    out[t * bs + (l * 2) * rest + c] = y0;
    out[t * bs + (l * 2 + 1) * rest + c] = y1;
  }
}

hipError_t phys_op_codes(const float2* X, float2* out, int L, long long rest, long long bs, int count, const int* codes,
                         int n, int site, hipStream_t s) {
  const long long tot = (long long)L * rest * count;
  int blocks = (int)std::min<long long>((tot + 255) / 256, 16384);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(phys_op_codes_kernel, dim3(blocks), dim3(256), 0, s, X, out, L, rest, bs, count, codes, n, site);
  return hipGetLastError();
}

// res[t] = Re E[t][0]
__global__ void take_re_kernel(const float2* E, long long bs, int count, double* res) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) res[t] = (double)E[(long long)t * bs].x;
}

hipError_t take_re(const float2* E, long long bs, int count, double* res, hipStream_t s) {
  hipLaunchKernelGGL(take_re_kernel, dim3((count + 255) / 256), dim3(256), 0, s, E, bs, count, res);
  return hipGetLastError();
}

}  // namespace mpdo
