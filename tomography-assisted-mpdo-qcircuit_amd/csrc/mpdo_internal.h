// Internal declarations shared by the gfx950 kernel files of libmpdo.
// Every complex number is stored interleaved (re, im): float2 for complex64, double2 for
// complex128. Site tensors use the fixed layout T[l][s=2][a][r] (r fastest).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpdo.h"

namespace mpdo {

// General strided batched complex GEMM:  C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b]
// with op(A)[i,k] = A[b*bsA + i*rsA + k*csA] (conjugated when conjA), similarly for B[k,j]
// and C[i,j] = C[b*bsC + i*rsC + j*csC]. Strides are in complex elements.
struct GemmArgs {
  int M, N, K, batch;
  const void* A; long long rsA, csA, bsA; int conjA;
  const void* B; long long rsB, csB, bsB; int conjB;
  void* C; long long rsC, csC, bsC;
  double alpha_re, alpha_im, beta_re, beta_im;
  // herm: C is Hermitian (M == N, the caller's product is Hermitian in exact arithmetic): only
  // the tiles on and below the diagonal are computed, then the strict upper triangle is set to
  // the conjugate transpose of the lower one (C exactly Hermitian)
  int herm;
  // operand storage overriding the precision mode's input type (fp64-accumulating modes only):
  // 0 = as the mode says, 1 = complex64, 2 = complex128
  int inA, inB;
  // split reduction: ksplit > 1 runs ksplit slices of K each (the caller's K is one slice),
  // slice t reading A + t ksA, B + t ksB and writing its partial product to C + t ksC (a
  // Hermitian output is then left unmirrored: the caller sums the slices and mirrors)
  int ksplit;
  long long ksA, ksB, ksC;
  // triangular operand: 1 = B upper triangular (B[k][j] = 0 for k > j), 2 = A lower triangular
  // (A[i][k] = 0 for k > i); each output tile then reduces over the k range that can be nonzero
  int tri;
};

enum Prec { C64_F32ACC = 0,   // complex64 in, fp32 MFMA accumulate, complex64 out
            C64_F64ACC = 1,   // complex64 in, fp64 MFMA accumulate, complex128 out
            C128_F64ACC = 2,  // complex128 in, fp64 MFMA accumulate, complex128 out
            C128_IN_C64_OUT = 3,    // complex128 in, fp64 accumulate, complex64 out
            C64_F64ACC_C64OUT = 4,    // complex64 in, fp64 accumulate, complex64 out
            C64_F32ACC_3M = 5,        // complex64 in, fp32 MFMA accumulate of 3 real products, complex64 out
            C64_BF16X3 = 6,           // complex64 in, bf16 MFMA on 2-part split operands (3 part products), fp32 accumulate, complex64 out
            C64_BF16X6 = 7 };         // the same with 3-part split operands (6 part products)

hipError_t gemm(Prec p, const GemmArgs& g, hipStream_t s);
// Complex64 GEMM on the bf16 matrix cores with each fp32 component split into nsplit (2 or 3) bf16
// parts; fp32 accumulation, complex64 output (gemm_bf16.hip)
hipError_t gemm_bf16(int nsplit, const GemmArgs& g, hipStream_t s);
// whether gemm_bf16 can address the operands (nonnegative strides, one batch entry's elements
// within 2 GiB of its first)
bool gemm_bf16_fits(const GemmArgs& g);
// C[b][i][j] = conj(C[b][j][i]) for j > i (n x n row-major, batch stride bs)
hipError_t herm_mirror_b(void* C, bool c128, int n, long long bs, int batch, hipStream_t s);

// Convenience: C = alpha*A*B (+beta C) with row-major operands of the given leading dims.
GemmArgs gemm_rowmajor(int M, int N, int K,
                       const void* A, long long lda, bool transA, bool conjA,
                       const void* B, long long ldb, bool transB, bool conjB,
                       void* C, long long ldc);

// Batched Hermitian eigensolver (complex128): for each of `batch` matrices
// A[b] (n x n, row-major, leading dim lda, batch stride bsA) computes eigenvalues in
// descending order w[b*ldw + j] and eigenvectors V[b*bsV + i*ldv + j] (column j = vector j).
// A is destroyed. `ws` must hold eigh_workspace_bytes(n, batch) bytes.
size_t eigh_workspace_bytes(int n, int batch);
hipError_t eigh(int n, int batch, void* A, long long lda, long long bsA,
                double* w, long long ldw, void* V, long long ldv, long long bsV,
                void* ws, hipStream_t s, int* sweeps_out);

// The same for positive semidefinite Gram matrices, preconditioned by pivoted Cholesky (pivots
// down to tol_rel of the largest diagonal) and solved at that rank r: the first r eigenpairs come
// from the Jacobi method on L^H L; the remaining n - r eigenvalues are the diagonal left in the
// Schur complement (sorted, so the eigenvalue sum stays the trace) with zero vectors.
// Synchronises `s` once (the rank). `ws` must hold eigh_gram_workspace_bytes(n, batch) bytes.
size_t eigh_gram_workspace_bytes(int n, int batch);
hipError_t eigh_gram(int n, int batch, void* A, long long lda, long long bsA,
                     double* w, long long ldw, void* V, long long ldv, long long bsV,
                     void* ws, hipStream_t s, int* sweeps_out, double tol_rel = 1e-14);

// Whether the absolute-error cut of decompositions.svd (te = sqrt(cumsum lambda), keep idx + 1
// values for the first idx with te[-1] - te[idx] <= eps; decompositions.py:121-131) keeps every
// one of the n directions of each Gram matrix A[b]: *all = true when, for every b, A[b] - D*_b I is
// positive definite, D*_b = tr A[b] - (sqrt(tr A[b]) - eps)^2 being the largest eigenvalue the
// rule may drop (decided by a Cholesky factorisation of the shifted matrix completing all n columns,
// with a small margin towards "no"). A is not modified; `ws`: eigh_gram_workspace_bytes(n, batch).
// 64 < n <= 1024 (else *all = false). Synchronises once.
hipError_t gram_keeps_all(int n, int batch, const void* A, long long lda, long long bsA, double eps, void* ws,
                          hipStream_t st, bool* all);

// The same cut decided from the bottom of the spectrum (tailcut.hip): subspace iteration with the
// inverse of a pivoted Cholesky factor of G[b] (n x n complex128, contiguous batch entries; not
// modified) on m columns for q steps, Rayleigh-Ritz, the rule on the Ritz values and a
// positive-definiteness certificate for the values it rests on. On success (res->ok) kept[b] and
// disc[b] (device) hold the rule's kept count and dropped eigenvalue sum, and where some chain
// drops a direction (res->any_drop) P[b] (complex64, leading dim ldP, batch stride bsP) = I - V_d
// V_d^H, V_d the dropped eigenvectors. Otherwise nothing is written and the caller runs the full
// eigensolver (res->reason says which step declined). 64 < n <= 1024, 16 <= m <= n / 2.
// Synchronises the stream four times. `ws`: gram_tail_workspace_bytes(n, m, batch).
struct TailResult {
  bool ok = false, any_drop = false;
  int max_kept = 0, reason = 0;
};
size_t gram_tail_workspace_bytes(int n, int m, int batch);
hipError_t gram_tail_cut(int n, int batch, const double2* G, double eps, int m, int q, void* ws, hipStream_t st,
                         int* kept, double* disc, float2* P, long long ldP, long long bsP, TailResult* res);

// The k largest eigenpairs of Gram matrices G[b] (n x n complex128, contiguous; not modified) by
// subspace iteration on m columns (q steps), Rayleigh-Ritz and a positive-definiteness certificate
// that the k leading Ritz values are the k largest eigenvalues (tailcut.hip). On success (res->ok)
// w[b] (leading dim ldw) holds the m Ritz values, descending, then the trace's remainder, then
// zeros (n values), and V[b] (leading dim ldv, batch stride bsV) the m Ritz vectors in its first m
// columns; otherwise nothing is written (res->reason says which step declined). 64 < n <= 1024,
// k + 16 <= m <= n - 64. Synchronises the stream three times. `ws`: gram_top_workspace_bytes.
struct TopResult {
  bool ok = false;
  int reason = 0;
};
size_t gram_top_workspace_bytes(int n, int m, int batch);
hipError_t gram_top_eig(int n, int batch, const double2* G, int k, int m, int q, void* ws, hipStream_t st, double* w,
                        long long ldw, double2* V, long long ldv, long long bsV, TopResult* res);

// Pivoted Cholesky of positive semidefinite Gram matrices A[b] (n x n, leading dim lda, batch
// stride bsA; destroyed): L[b] (n x n, stored by columns: L[b][k n + i] = L_ik) with A = L L^H,
// piv[b][k] the row pivoted at column k, rank[b] the columns produced before the largest remaining
// diagonal fell to tol_rel * max_i A_ii (later columns zero). `ws`: pchol_workspace_bytes.
size_t pchol_workspace_bytes(int n, int batch);
// early_exit: every 128 columns read back whether every matrix has stopped (saves the rest of the
// work for low-rank inputs at the price of a synchronisation).
hipError_t pchol_b(int n, int batch, void* A, long long lda, long long bsA, double tol_rel, double2* L, int* piv,
                   int* rank, void* ws, hipStream_t s, bool early_exit = true);
// Unpivoted blocked Cholesky of well-conditioned Hermitian matrices, in pchol_b's output layout
// with the identity pivot order (piv[b][k] = k): rank[b] = n when every pivot stays above
// tol_rel * max_i A_ii, else the column where it did not (later columns meaningless). A is
// destroyed. No host synchronisation. `ws`: chol_workspace_bytes(batch).
size_t chol_workspace_bytes(int batch);
hipError_t chol_b(int n, int batch, void* A, long long lda, long long bsA, double tol_rel, double2* L, int* piv,
                  int* rank, void* ws, hipStream_t s);
// After pchol_b (same ws, rank): w[b][rank[b] + j], j < n - rank[b], receives the diagonal left in
// matrix b's Schur complement, sorted descending (the Gram mass below the pivot threshold); n <= 1024.
hipError_t pchol_tail_eigs(int n, int batch, const void* ws, const int* rank, double* w, long long ldw, hipStream_t s);
// cond2[b] (device) = (first pivot / last pivot) of matrix b's factor: the squared ratio of the
// extreme diagonal entries of the pivot-ordered triangle, an estimate of the condition number of
// the Gram matrix restricted to its numerical rank.
hipError_t pchol_cond2(int n, int batch, const double2* L, const int* piv, const int* rank, double* cond2,
                       hipStream_t s);
// The Cholesky-QR factor of a pivoted factor, as complex128 X[b] (n x n): with Lt[t][u] =
// L[piv[t]][u] (t, u < rank), mode 0 gives X[piv[t]][u] = conj(Lt^-1[u][t]) (Q = M X has
// orthonormal columns when M^H M = L L^H and M = Q L^H), mode 1 X[u][piv[t]] = Lt^-1[u][t]
// (Qt = X M when M M^H = L L^H, M = L Qt); zero elsewhere. Modes 2 and 3 give the same factors
// in pivot order, X[t][u] = conj(Lt^-1[u][t]) (upper triangular) and X[u][t] = Lt^-1[u][t] (lower
// triangular), for Q = gather_piv(M, cols) X and Qt = X gather_piv(M, rows).
// `ws`: pchol_inverse_workspace_bytes.
size_t pchol_inverse_workspace_bytes(int n, int batch);
// Pivot-order gather of complex64 data for the triangular products of modes 2/3: cols = 1
// out[i][t] = M[i][piv[t]] (M: other x n), cols = 0 out[t][j] = M[piv[t]][j] (M: n x other); zero
// for t >= rank. Contiguous batch entries of n * other elements.
hipError_t gather_piv(const float2* M, const int* piv, const int* rank, int n, long long other, int cols, float2* out,
                      int batch, hipStream_t st);
hipError_t pchol_inverse_b(int n, int batch, const double2* L, const int* piv, const int* rank, int mode, double2* X,
                           void* ws, hipStream_t s);

// Pauli-string expectations of many strings at once through a caller-provided scratch region
// (api.hip; the arithmetic of mpdo_pauli_expect): expect_check validates the arguments and gives
// the scratch bytes one string needs; expect_chunk_bytes the bytes of a chunk of `chunk` strings;
// expect_run evaluates the strings `chunk` at a time in the scratch at `base` (synchronises once
// per chunk). Failures set the thread's error text.
constexpr int EXPECT_MAX_CHUNK = 4096;
mpdo_status expect_check(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out, size_t* per);
size_t expect_chunk_bytes(int n, const mpdo_site* sites, int chunk);
mpdo_status expect_run(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out, char* base,
                       int chunk, hipStream_t s);
}  // namespace mpdo
