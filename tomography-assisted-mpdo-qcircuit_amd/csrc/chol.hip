// Pivoted Cholesky of batched positive semidefinite Gram matrices (complex128) and the inverse of
// its pivot-ordered triangular factor.
//
// pchol_b: G = L L^H with L stored by columns (L[k n + i] = L_ik in the original row order), the
// pivot of column k the row with the largest remaining diagonal, stopped once that diagonal is at
// or below tol * max_i G_ii (the rank). Blocked: a panel kernel produces PC_NB pivot columns
// (left-looking within the panel, one workgroup per matrix), then one fp64 MFMA GEMM subtracts the
// panel's outer products from the trailing matrix, so the sequential part reads only the panel.
//
// pchol_inverse_b: with Lt[t][u] = L[piv[t]][u] (lower triangular in pivot order, t, u < rank),
// the inverse Lt^{-1} by 64 x 64 diagonal-block inversions in LDS and block merges
// inv([A 0; C B]) = [A^-1 0; -B^-1 C A^-1  B^-1] on the matrix cores, scattered back to the
// original index order as the factor of a Cholesky QR: Q = M X with X[piv[t]][u] = conj(Lt^-1[u][t])
// (M^H M = L L^H) or Qt = X M with X[u][piv[t]] = Lt^-1[u][t] (M M^H = L L^H).
#include "mpdo_internal.h"
#include "prof.h"

#include <algorithm>
#include <vector>

namespace mpdo {
namespace {

inline size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr int PC_NB = 64;      // columns per panel; the trailing update is one GEMM per panel
constexpr int PC_THREADS = 512;
constexpr int PC_ROWS = 4;     // rows per thread of the panel kernel: n <= PC_ROWS * PC_THREADS

// Per-matrix state of the blocked pivoted Cholesky, kept in global memory across panels.
struct PcState {
  double* d;             // [batch][n] remaining diagonal A_ii - sum_t |L_it|^2
  unsigned char* used;   // [batch][n] 1 once row i has been a pivot
  int* piv;              // [batch][n] pivot row of column k
  int* rank;             // [batch] columns produced so far (final: the numerical rank)
  int* stop;             // [batch] 1 once the largest remaining diagonal is at or below thr
  double* thr;           // [batch] tol * max_i A_ii
};

__global__ __launch_bounds__(PC_THREADS) void pchol_init_kernel(const double2* A, int n, long long lda, long long bsA,
                                                                 double tol_rel, double2* L, PcState st) {
  __shared__ double red[PC_THREADS];
  const int b = blockIdx.x, tid = threadIdx.x;
  A += (long long)b * bsA;
  double m = 0;
  for (int i = tid; i < n; i += PC_THREADS) {
    const double v = A[(long long)i * lda + i].x;
    st.d[(long long)b * n + i] = v;
    st.used[(long long)b * n + i] = 0;
    st.piv[(long long)b * n + i] = 0;
    m = fmax(m, v);
  }
  double2* Lb = L + (long long)b * n * n;
  for (long long e = tid; e < (long long)n * n; e += PC_THREADS) Lb[e] = make_double2(0.0, 0.0);
  red[tid] = m;
  __syncthreads();
  for (int s2 = PC_THREADS / 2; s2 > 0; s2 >>= 1) {
    if (tid < s2) red[tid] = fmax(red[tid], red[tid + s2]);
    __syncthreads();
  }
  if (tid == 0) {
    st.thr[b] = tol_rel * red[0];
    st.rank[b] = 0;
    st.stop[b] = 0;
  }
}

// One panel of up to PC_NB pivot columns of matrix b (left-looking within the panel). A holds
// the Gram matrix minus the outer products of every earlier panel (the trailing updates), so a
// column needs only the panel's own earlier columns: L_ik = (A_ip - sum_{t in panel, t<k}
// L_it conj(L_pt)) / sqrt(d_p). Rows that were pivots get 0. Thread tid owns rows tid + m
// PC_THREADS (m < PC_ROWS) and keeps their remaining diagonal and pivot flags in registers for
// the whole panel. Two barriers per column: the argmax candidates of the waves, then the pivot
// row's panel entries.
__global__ __launch_bounds__(PC_THREADS) void pchol_panel_kernel(const double2* A, int n, long long lda, long long bsA,
                                                                  int j0, double2* L, PcState st) {
  __shared__ double wv[PC_THREADS / 64];
  __shared__ int wi[PC_THREADS / 64];
  __shared__ double2 Lp[PC_NB];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (st.stop[b]) return;
  A += (long long)b * bsA;
  double2* Lb = L + (long long)b * n * n;
  double* d = st.d + (long long)b * n;
  unsigned char* used = st.used + (long long)b * n;
  const double thr = st.thr[b];
  double dr[PC_ROWS];
  bool ur[PC_ROWS];
#pragma unroll
  for (int m = 0; m < PC_ROWS; ++m) {
    const int i = tid + m * PC_THREADS;
    dr[m] = i < n ? d[i] : 0.0;
    ur[m] = i < n ? used[i] != 0 : true;
  }
  int k = j0;
  const int kend = min(j0 + PC_NB, n);
  for (; k < kend; ++k) {
    // argmax of the remaining diagonal (ties: lowest index) over the rows this thread owns
    double best = -1.0;
    int bi = n;
#pragma unroll
    for (int m = 0; m < PC_ROWS; ++m)
      if (!ur[m] && dr[m] > best) { best = dr[m]; bi = tid + m * PC_THREADS; }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
    __syncthreads();
    // every thread reduces the per-wave candidates itself: no broadcast barrier needed
    double bv = wv[0];
    int bx = wi[0];
    for (int w = 1; w < PC_THREADS / 64; ++w)
      if (wv[w] > bv || (wv[w] == bv && wi[w] < bx)) { bv = wv[w]; bx = wi[w]; }
    const int p = (bv > thr && bx < n) ? bx : -1;
    if (p < 0) break;
    if (tid == 0) st.piv[(long long)b * n + k] = p;
    for (int t = tid; t < k - j0; t += PC_THREADS) Lp[t] = Lb[(long long)(j0 + t) * n + p];
    __syncthreads();
    const double piv = sqrt(bv);
    const double2* Ap = A + (long long)p * lda;   // row p: A_ip = conj(A_pi)
#pragma unroll
    for (int m = 0; m < PC_ROWS; ++m) {
      const int i = tid + m * PC_THREADS;
      if (i >= n) continue;
      double2 v;
      if (i == p) {
        v = make_double2(piv, 0.0);
        ur[m] = true;
      } else if (ur[m]) {
        v = make_double2(0.0, 0.0);
      } else {
        const double2 api = Ap[i];
        double2 acc = make_double2(api.x, -api.y);
        for (int t = 0; t < k - j0; ++t) {   // acc -= L_it conj(L_pt)
          const double2 x = Lb[(long long)(j0 + t) * n + i], y = Lp[t];
          acc.x -= x.x * y.x + x.y * y.y;
          acc.y -= x.y * y.x - x.x * y.y;
        }
        v = make_double2(acc.x / piv, acc.y / piv);
        dr[m] -= v.x * v.x + v.y * v.y;
      }
      Lb[(long long)k * n + i] = v;
    }
  }
#pragma unroll
  for (int m = 0; m < PC_ROWS; ++m) {
    const int i = tid + m * PC_THREADS;
    if (i < n) {
      d[i] = dr[m];
      used[i] = ur[m] ? 1 : 0;
    }
  }
  if (tid == 0) {
    st.rank[b] = k;
    if (k < kend || k == n) st.stop[b] = 1;
  }
}

// Lt[b] (R x R) = the pivot-ordered factor, the identity past the matrix's own rank
__global__ void gather_lt_kernel(const double2* L, const int* piv, const int* rank, int n, int R, double2* Lt) {
  const int b = blockIdx.y;
  const int rb = rank[b];
  const double2* Lb = L + (long long)b * n * n;
  const int* pb = piv + (long long)b * n;
  double2* T = Lt + (long long)b * R * R;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)R * R;
       e += (long long)gridDim.x * blockDim.x) {
    const int t = (int)(e / R), u = (int)(e % R);
    double2 v = make_double2(t == u ? 1.0 : 0.0, 0.0);
    if (t < rb && u < rb) v = u <= t ? Lb[(long long)u * n + pb[t]] : make_double2(0.0, 0.0);
    T[e] = v;
  }
}

// In-place inverse of the 64 x 64 lower-triangular diagonal blocks of Lt (one wave per block):
// lane j solves Lt_blk x = e_j by forward substitution.
__global__ __launch_bounds__(64) void inv64_kernel(double2* Lt, int R) {
  __shared__ double2 A[64][65];
  __shared__ double2 X[64][65];
  const int blk = blockIdx.x, b = blockIdx.y, j = threadIdx.x;
  double2* T = Lt + (long long)b * R * R + (long long)blk * 64 * R + blk * 64;
  for (int r = 0; r < 64; ++r) A[r][j] = T[(long long)r * R + j];
  __syncthreads();
  for (int i = 0; i < 64; ++i) {
    double2 v = make_double2(0.0, 0.0);
    if (i == j) {
      const double2 a = A[i][i];
      const double m = a.x * a.x + a.y * a.y;
      v = make_double2(a.x / m, -a.y / m);
    } else if (i > j) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = j; k < i; ++k) {   // acc += A_ik X_kj
        const double2 a = A[i][k], x = X[k][j];
        acc.x += a.x * x.x - a.y * x.y;
        acc.y += a.x * x.y + a.y * x.x;
      }
      const double2 a = A[i][i];
      const double m = a.x * a.x + a.y * a.y;
      // v = -acc / a_ii
      v = make_double2(-(acc.x * a.x + acc.y * a.y) / m, -(acc.y * a.x - acc.x * a.y) / m);
    }
    X[i][j] = v;
  }
  for (int r = 0; r < 64; ++r) T[(long long)r * R + j] = X[r][j];
}

// Pivot-order gathers of complex64 data: cols = 1: out[i][t] = M[i][piv[t]] (rows x n), cols = 0:
// out[t][j] = M[piv[t]][j] (n x cols); zero for t >= rank
__global__ void gather_piv_kernel(const float2* M, const int* piv, const int* rank, int n, long long other, int cols,
                                  float2* out) {
  const int b = blockIdx.y;
  const int rb = rank[b];
  const int* pb = piv + (long long)b * n;
  const long long per = (long long)n * other;
  const float2* Mb = M + (long long)b * per;
  float2* Ob = out + (long long)b * per;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (long long)gridDim.x * blockDim.x) {
    float2 v = make_float2(0.f, 0.f);
    if (cols) {
      const long long i = e / n;
      const int t = (int)(e % n);
      if (t < rb) v = Mb[i * n + pb[t]];
    } else {
      const int t = (int)(e / other);
      const long long j = e % other;
      if (t < rb) v = Mb[(long long)pb[t] * other + j];
    }
    Ob[e] = v;
  }
}

__global__ void zero_c128_kernel(double2* X, long long per) {
  double2* Xb = X + (long long)blockIdx.y * per;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (long long)gridDim.x * blockDim.x)
    Xb[e] = make_double2(0.0, 0.0);
}

// X (n x n complex128, zero-filled before) from Linv = Lt^-1 (R x R): mode 0
// X[piv[t]][u] = conj(Linv[u][t]), mode 1 X[u][piv[t]] = Linv[u][t], for t, u < rank; modes 2
// and 3 the same in pivot order (X[t][u] = conj(Linv[u][t]), upper triangular; X[u][t] =
// Linv[u][t], lower triangular), for products with the pivot-gathered input
__global__ void scatter_x_kernel(const double2* Linv, const int* piv, const int* rank, int n, int R, int mode,
                                 double2* X) {
  const int b = blockIdx.y;
  const int rb = rank[b];
  const int* pb = piv + (long long)b * n;
  const double2* Li = Linv + (long long)b * R * R;
  double2* Xb = X + (long long)b * n * n;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)rb * rb;
       e += (long long)gridDim.x * blockDim.x) {
    const int t = (int)(e / rb), u = (int)(e % rb);
    const double2 v = Li[(long long)u * R + t];
    if (mode == 0) Xb[(long long)pb[t] * n + u] = make_double2(v.x, -v.y);
    else if (mode == 1) Xb[(long long)u * n + pb[t]] = v;
    else if (mode == 2) Xb[(long long)t * n + u] = make_double2(v.x, -v.y);
    else Xb[(long long)u * n + t] = v;
  }
}

// w[b][rank_b + j] for j < n - rank_b: the diagonal left in the Schur complement of matrix b (the
// rows never pivoted), clamped at 0 and sorted descending: the Gram mass below the pivot threshold
// stays in the eigenvalue list. One 1024-thread block per matrix (n <= 1024), bitonic sort in LDS.
__global__ __launch_bounds__(1024) void pchol_tail_kernel(const double* d, const unsigned char* used, const int* rank,
                                                          int n, double* w, long long ldw) {
  __shared__ double t[1024];
  const int b = blockIdx.x, tid = threadIdx.x;
  const double* db = d + (long long)b * n;
  const unsigned char* ub = used + (long long)b * n;
  t[tid] = (tid < n && !ub[tid]) ? fmax(db[tid], 0.0) : -1.0;
  __syncthreads();
  for (int k = 2; k <= 1024; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = tid ^ j;
      if (p > tid) {
        const bool desc = (tid & k) == 0;
        const double x = t[tid], y = t[p];
        if (desc ? (x < y) : (x > y)) { t[tid] = y; t[p] = x; }
      }
      __syncthreads();
    }
  const int rb = rank[b];
  if (tid < n - rb) w[(long long)b * ldw + rb + tid] = t[tid];
}

// cond2[b] = (L_00 / L_{r-1,r-1})^2 in pivot order (first over last pivot of matrix b), 1 when r <= 1
__global__ void pchol_cond_kernel(const double2* L, const int* piv, const int* rank, int n, int batch, double* cond2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int r = rank[b];
  double c = 1.0;
  if (r > 1) {
    const double2* Lb = L + (long long)b * n * n;
    const int* pb = piv + (long long)b * n;
    const double d0 = Lb[pb[0]].x, dl = Lb[(long long)(r - 1) * n + pb[r - 1]].x;
    c = dl > 0 ? (d0 / dl) * (d0 / dl) : 1e300;
  }
  cond2[b] = c;
}

int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 2048)); }

}  // namespace

size_t pchol_workspace_bytes(int n, int batch) {
  return a256((size_t)n * sizeof(double) * batch) + a256((size_t)n * batch) + 3 * a256(sizeof(double) * batch);
}

hipError_t pchol_b(int n, int batch, void* A, long long lda, long long bsA, double tol_rel, double2* L, int* piv,
                   int* rank, void* ws, hipStream_t st, bool early_exit) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  if (n > PC_ROWS * PC_THREADS) return hipErrorInvalidValue;
  hipError_t err;
  char* p = (char*)ws;
  PcState pc;
  pc.d = (double*)p; p += a256((size_t)n * sizeof(double) * batch);
  pc.used = (unsigned char*)p; p += a256((size_t)n * batch);
  pc.stop = (int*)p; p += a256(sizeof(double) * batch);
  pc.thr = (double*)p; p += a256(sizeof(double) * batch);
  pc.piv = piv;
  pc.rank = rank;
  prof::Scope ps(prof::EIG_SMALL, st, 0.0, 16.0 * n * n * batch);
  hipLaunchKernelGGL(pchol_init_kernel, dim3(batch), dim3(PC_THREADS), 0, st, (const double2*)A, n, lda, bsA, tol_rel,
                     L, pc);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  std::vector<int> hs(batch);
  for (int j0 = 0; j0 < n; j0 += PC_NB) {
    hipLaunchKernelGGL(pchol_panel_kernel, dim3(batch), dim3(PC_THREADS), 0, st, (const double2*)A, n, lda, bsA, j0, L,
                       pc);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    const int nb = std::min(PC_NB, n - j0);
    if (j0 + nb >= n) break;
    // trailing update A <- A - L[:, j0:j0+nb] L[:, j0:j0+nb]^H (zero columns of stopped matrices add nothing)
    GemmArgs g{};
    g.M = n; g.N = n; g.K = nb; g.batch = batch;
    g.A = L + (long long)j0 * n; g.rsA = 1; g.csA = n; g.bsA = (long long)n * n;
    g.B = L + (long long)j0 * n; g.rsB = n; g.csB = 1; g.bsB = (long long)n * n; g.conjB = 1;
    g.C = A; g.rsC = lda; g.csC = 1; g.bsC = bsA;
    g.alpha_re = -1.0; g.beta_re = 1.0;
    g.herm = 1;   // the updated matrix stays Hermitian: lower tiles, then the mirror
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
    if (early_exit && (j0 + nb) % 128 == 0) {   // every 128 columns: stop once every matrix has stopped
      if ((err = hipMemcpyAsync(hs.data(), pc.stop, sizeof(int) * batch, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return err;
      if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
      bool all = true;
      for (int b = 0; b < batch; ++b) all = all && hs[b];
      if (all) break;
    }
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------------------
// Unpivoted blocked Cholesky for well-conditioned Hermitian matrices (the Gram matrices of nearly
// orthonormal blocks, and positive-definiteness tests): per 64-column panel one workgroup per
// matrix factors the diagonal block in LDS and inverts its triangle, one GEMM solves the rows
// below (A_ip Lpp^{-H}, triangular operand), one Hermitian GEMM updates the trailing matrix. No
// host synchronisation. A matrix whose pivot falls to tol_rel x max_i A_ii stops there (rank).
// Output in pchol_b's layout with the identity pivot order, for pchol_inverse_b.
// ---------------------------------------------------------------------------------------
namespace {

struct ChState {
  int* rank;
  int* stop;
  double* thr;
  double2* Linv;   // [batch][64][64] inverse of the current diagonal block (row-major)
};

__global__ __launch_bounds__(256) void chol_init_kernel(const double2* A, int n, long long lda, long long bsA,
                                                        double tol_rel, double2* L, int* piv, ChState st) {
  __shared__ double red[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  A += (long long)b * bsA;
  double m = 0;
  for (int i = tid; i < n; i += 256) {
    m = fmax(m, A[(long long)i * lda + i].x);
    piv[(long long)b * n + i] = i;
  }
  double2* Lb = L + (long long)b * n * n;
  for (long long e = tid; e < (long long)n * n; e += 256) Lb[e] = make_double2(0.0, 0.0);
  red[tid] = m;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (tid < s2) red[tid] = fmax(red[tid], red[tid + s2]);
    __syncthreads();
  }
  if (tid == 0) {
    st.thr[b] = tol_rel * red[0];
    st.rank[b] = 0;
    st.stop[b] = 0;
  }
}

// Diagonal block j0 .. j0 + nb - 1 of matrix b (already updated by the earlier panels): Cholesky,
// written to L (column k at L[k n + i]), and its triangle's inverse to st.Linv. 256 threads as a
// 16 x 16 grid; thread (ty, tx) keeps the elements (ty + 16 a, tx + 16 c), a, c < 4, in registers.
// Factor: at column j the owners of column j publish it (unscaled) to LDS, and every thread
// subtracts S_ij conj(S_kj) / S_jj from its elements with j < k <= i; the columns are scaled by
// 1 / sqrt(S_jj) at the end. Inverse X = L^{-1}, row by row: the owners of row i divide it by L_ii
// and publish it, every thread subtracts L_ri X_it from its rows r > i. Both publish into
// alternating LDS buffers: one barrier per column (row).
__global__ __launch_bounds__(256) void chol_diag_kernel(const double2* A, int n, long long lda, long long bsA, int j0,
                                                        double2* L, ChState st) {
  __shared__ double2 vec[2][64];
  __shared__ double2 Ls[64][65];
  const int b = blockIdx.x, tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  if (st.stop[b]) return;
  const int nb = min(64, n - j0);
  A += (long long)b * bsA + (long long)j0 * lda + j0;
  double2 v[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = ty + 16 * a, k = tx + 16 * c;
      v[a][c] = (i < nb && k <= i) ? A[(long long)i * lda + k] : make_double2(0.0, 0.0);
    }
  const double thr = st.thr[b];
  double dk[4] = {0.0, 0.0, 0.0, 0.0};   // S_kk of the thread's columns k = tx + 16 c
  int fail = -1;
  for (int j = 0; j < nb; ++j) {
    double2* col = vec[j & 1];
    // (register arrays indexed by unrolled constants only: the owner tests run over every c)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (tx + 16 * c == j) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int i = ty + 16 * a;
          if (i >= j && i < nb) col[i] = v[a][c];
        }
      }
    __syncthreads();
    const double d = col[j].x;
    if (!(d > thr)) { fail = j; break; }   // uniform: every thread read the same value
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (tx + 16 * c == j) dk[c] = d;
    const double inv = 1.0 / d;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = ty + 16 * a;
      if (i <= j || i >= nb) continue;
      const double2 x = col[i];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = tx + 16 * c;
        if (k <= j || k > i) continue;
        const double2 y = col[k];
        v[a][c].x -= (x.x * y.x + x.y * y.y) * inv;
        v[a][c].y -= (x.y * y.x - x.x * y.y) * inv;
      }
    }
  }
  double2* Lb = L + (long long)b * n * n;
  double2* Li = st.Linv + (long long)b * 64 * 64;
  // the factor's columns k < kend (all nb, or those before a failed pivot): L_ik = S_ik / sqrt(S_kk)
  // below the diagonal, sqrt(S_kk) on it
  const int kend = fail >= 0 ? fail : nb;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = tx + 16 * c;
    const double rk = k < kend ? sqrt(dk[c]) : 1.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = ty + 16 * a;
      double2 l = make_double2(0.0, 0.0);
      if (i < nb && k < kend && k <= i) {
        l = (k == i) ? make_double2(rk, 0.0) : make_double2(v[a][c].x / rk, v[a][c].y / rk);
        Lb[(long long)(j0 + k) * n + j0 + i] = l;
      }
      Ls[i][k] = l;
      // the inverse starts from the identity (the right-hand sides)
      v[a][c] = make_double2(i == k ? 1.0 : 0.0, 0.0);
    }
  }
  if (fail >= 0) {
    for (int e = tid; e < 64 * 64; e += 256) Li[e] = make_double2(0.0, 0.0);
    if (tid == 0) {
      st.rank[b] = j0 + fail;
      st.stop[b] = 1;
    }
    return;
  }
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    double2* row = vec[i & 1];
#pragma unroll
    for (int a = 0; a < 4; ++a)
      if (ty + 16 * a == i) {
        const double dii = Ls[i][i].x;   // real and positive
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int t = tx + 16 * c;
          if (t <= i) {
            v[a][c] = make_double2(v[a][c].x / dii, v[a][c].y / dii);
            row[t] = v[a][c];
          }
        }
      }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int r = ty + 16 * a;
      if (r <= i || r >= nb) continue;
      const double2 l = Ls[r][i];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int t = tx + 16 * c;
        if (t > i) continue;
        const double2 x = row[t];
        v[a][c].x -= l.x * x.x - l.y * x.y;
        v[a][c].y -= l.x * x.y + l.y * x.x;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) Li[(ty + 16 * a) * 64 + tx + 16 * c] = v[a][c];
  if (tid == 0) st.rank[b] = j0 + nb;
}

}  // namespace

size_t chol_workspace_bytes(int batch) {
  return 2 * a256(sizeof(int) * batch) + a256(sizeof(double) * batch) + a256((size_t)64 * 64 * sizeof(double2) * batch);
}

hipError_t chol_b(int n, int batch, void* A, long long lda, long long bsA, double tol_rel, double2* L, int* piv,
                  int* rank, void* ws, hipStream_t st) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipError_t err;
  char* p = (char*)ws;
  ChState cs;
  cs.rank = rank;
  cs.stop = (int*)p; p += a256(sizeof(int) * batch);
  p += a256(sizeof(int) * batch);
  cs.thr = (double*)p; p += a256(sizeof(double) * batch);
  cs.Linv = (double2*)p;
  prof::Scope ps(prof::EIG_SMALL, st, 0.0, 16.0 * n * n * batch);
  hipLaunchKernelGGL(chol_init_kernel, dim3(batch), dim3(256), 0, st, (const double2*)A, n, lda, bsA, tol_rel, L, piv,
                     cs);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  double2* Ab = (double2*)A;
  for (int j0 = 0; j0 < n; j0 += 64) {
    hipLaunchKernelGGL(chol_diag_kernel, dim3(batch), dim3(256), 0, st, (const double2*)A, n, lda, bsA, j0, L, cs);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    const int nb = std::min(64, n - j0), rest = n - j0 - nb;
    if (rest <= 0) break;
    // L[j0+nb:, j0:j0+nb] = A[j0+nb:, j0:j0+nb] Linv^H (Linv^H upper triangular)
    GemmArgs g{};
    g.M = rest; g.N = nb; g.K = nb; g.batch = batch;
    g.A = Ab + (long long)(j0 + nb) * lda + j0; g.rsA = lda; g.csA = 1; g.bsA = bsA;
    g.B = cs.Linv; g.rsB = 1; g.csB = 64; g.bsB = 64 * 64; g.conjB = 1;
    g.C = L + (long long)j0 * n + j0 + nb; g.rsC = 1; g.csC = n; g.bsC = (long long)n * n;
    g.alpha_re = 1.0;
    g.tri = 1;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
    // A[j0+nb:, j0+nb:] -= L[j0+nb:, j0:j0+nb] L[j0+nb:, j0:j0+nb]^H
    GemmArgs h{};
    h.M = rest; h.N = rest; h.K = nb; h.batch = batch;
    h.A = L + (long long)j0 * n + j0 + nb; h.rsA = 1; h.csA = n; h.bsA = (long long)n * n;
    h.B = L + (long long)j0 * n + j0 + nb; h.rsB = n; h.csB = 1; h.bsB = (long long)n * n; h.conjB = 1;
    h.C = Ab + (long long)(j0 + nb) * lda + j0 + nb; h.rsC = lda; h.csC = 1; h.bsC = bsA;
    h.alpha_re = -1.0; h.beta_re = 1.0;
    h.herm = 1;
    if ((err = gemm(C128_F64ACC, h, st)) != hipSuccess) return err;
  }
  return hipSuccess;
}

hipError_t pchol_tail_eigs(int n, int batch, const void* ws, const int* rank, double* w, long long ldw, hipStream_t st) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  if (n > 1024) return hipErrorInvalidValue;
  const char* p = (const char*)ws;
  const double* d = (const double*)p;
  const unsigned char* used = (const unsigned char*)(p + a256((size_t)n * sizeof(double) * batch));
  hipLaunchKernelGGL(pchol_tail_kernel, dim3(batch), dim3(1024), 0, st, d, used, rank, n, w, ldw);
  return hipGetLastError();
}

hipError_t pchol_cond2(int n, int batch, const double2* L, const int* piv, const int* rank, double* cond2,
                       hipStream_t st) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(pchol_cond_kernel, dim3((batch + 255) / 256), dim3(256), 0, st, L, piv, rank, n, batch, cond2);
  return hipGetLastError();
}

int pinv_order(int n) {
  int R = 64;
  while (R < n) R *= 2;
  return R;
}

size_t pchol_inverse_workspace_bytes(int n, int batch) {
  const size_t R = (size_t)pinv_order(n);
  return a256(R * R * sizeof(double2) * batch) + a256(R * R / 4 * sizeof(double2) * batch);
}

hipError_t pchol_inverse_b(int n, int batch, const double2* L, const int* piv, const int* rank, int mode, double2* X,
                           void* ws, hipStream_t st) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipError_t err;
  const int R = pinv_order(n);
  double2* Lt = (double2*)ws;
  double2* T = (double2*)((char*)ws + a256((size_t)R * R * sizeof(double2) * batch));
  prof::Scope ps(prof::EIG_SMALL, st, 2.7 * R * R * R * batch, 16.0 * R * R * batch);
  hipLaunchKernelGGL(gather_lt_kernel, dim3(grid_for((long long)R * R), batch), dim3(256), 0, st, L, piv, rank, n, R,
                     Lt);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  hipLaunchKernelGGL(inv64_kernel, dim3(R / 64, batch), dim3(64), 0, st, Lt, R);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  for (int s = 64; s < R; s *= 2) {
    for (int m = 0; m < R / (2 * s); ++m) {
      double2* Ainv = Lt + (long long)(2 * m) * s * (R + 1);
      double2* Binv = Lt + (long long)(2 * m + 1) * s * (R + 1);
      double2* C = Lt + (long long)(2 * m + 1) * s * R + (long long)(2 * m) * s;
      // T = C Ainv
      GemmArgs g{};
      g.M = s; g.N = s; g.K = s; g.batch = batch;
      g.A = C; g.rsA = R; g.csA = 1; g.bsA = (long long)R * R;
      g.B = Ainv; g.rsB = R; g.csB = 1; g.bsB = (long long)R * R;
      g.C = T; g.rsC = s; g.csC = 1; g.bsC = (long long)R * R / 4;
      g.alpha_re = 1.0;
      if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
      // C <- -Binv T
      GemmArgs h{};
      h.M = s; h.N = s; h.K = s; h.batch = batch;
      h.A = Binv; h.rsA = R; h.csA = 1; h.bsA = (long long)R * R;
      h.B = T; h.rsB = s; h.csB = 1; h.bsB = (long long)R * R / 4;
      h.C = C; h.rsC = R; h.csC = 1; h.bsC = (long long)R * R;
      h.alpha_re = -1.0;
      if ((err = gemm(C128_F64ACC, h, st)) != hipSuccess) return err;
    }
  }
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(zero_c128_kernel, dim3(grid_for(nn), batch), dim3(256), 0, st, X, nn);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  hipLaunchKernelGGL(scatter_x_kernel, dim3(grid_for(nn), batch), dim3(256), 0, st, (const double2*)Lt, piv, rank, n, R,
                     mode, X);
  return hipGetLastError();
}

hipError_t gather_piv(const float2* M, const int* piv, const int* rank, int n, long long other, int cols, float2* out,
                      int batch, hipStream_t st) {
  if (n <= 0 || other <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_piv_kernel, dim3(grid_for((long long)n * other), batch), dim3(256), 0, st, M, piv, rank, n,
                     other, cols, out);
  return hipGetLastError();
}

}  // namespace mpdo
