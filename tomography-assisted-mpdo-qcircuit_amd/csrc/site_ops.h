// Declarations of the memory-bound kernels (site_ops.hip).
#pragma once
#include "mpdo_internal.h"

namespace mpdo {

struct Site1q {
  const float2* in;
  float2* out;
  int l, a, r, K;
  long long koff;   // offset of this site's [K][2][2] Kraus block in the Kraus buffer
};

struct PermDesc {
  int rank;
  int odims[8];
  long long istride[8];
};

hipError_t apply_1q(int nsites, const Site1q* d_sites, const float2* d_kraus, long long max_elems,
                    hipStream_t s);
hipError_t channel2q(const float2* Y, const float2* G, int K, int P1, int P2, float2* C, hipStream_t s, int batch = 1,
                     long long bsY = 0, long long bsC = 0);
hipError_t block_diag_sum(const float2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch,
                          hipStream_t s);
hipError_t block_diag_sum64(const double2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch,
                            hipStream_t s);
hipError_t mask_cols_b(const double2* V, long long ldv, long long bsV, int rows, int cols, const int* k, float2* out,
                       long long ldo, long long bsO, int batch, hipStream_t s);
hipError_t widen(const float2* in, double2* out, long long n, long long bsI, long long bsO, int batch, hipStream_t s);
// C C^H of the two-site core C = G(Y) (channel2q layout) from Pg = Ycat Ycat^H (4 P1 x 4 P1 per
// batch entry) and the channel tensor: out (2 P1 x 2 P1, complex128)
hipError_t kraus_gram(const double2* Pg, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s);
// The dual: sum over the channel's output/Kraus columns of C^H N C = Ycat^H Nt Ycat, Nt (4 P1 x 4 P1,
// complex128) from N (2 P1 x 2 P1, rows (p, o1)) and the channel tensor
hipError_t kraus_dual(const double2* N, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s);
// out = R (x) I_2 with rows (p, s), columns (p', s') (complex64 in, complex128 out)
hipError_t kron_i2(const float2* R, int P, int Q, double2* out, int batch, hipStream_t s);
// Kappa cut of a Kraus-factored max site T[k,s2,K2,a2,r] = sum_oi g[s2][oi][K2] W[k,oi,a2,r]:
// the inner Gram (A K x A K) from Hg = Wp Wp^H (8 A x 8 A, rows (oi, a2)), and the kept
// directions Vk (A K x kk) mapped to Vt[(oi a2), (s2 j)] (8 A x 2 kk)
hipError_t kappa_core_gram(const double2* Hg, const float2* Gt, int K, int A, double2* out, int batch, hipStream_t s);
hipError_t kappa_core_vec(const float2* Vk, const float2* Gt, int K, int A, int kk, float2* Vt, int batch,
                          hipStream_t s);
// complex64 copy of a column-stored n x n factor: mode 0 its conjugate transpose, mode 1 itself (row-major)
hipError_t narrow_factor(const double2* L, int n, int mode, float2* out, int batch, hipStream_t s);
// out[b] = sum over t < S of in[b S + t] (n complex128 each)
hipError_t sum_parts(const double2* in, int S, long long n, double2* out, int batch, hipStream_t s);
hipError_t eye_b(float2* out, int rows, int cols, long long bs, int batch, hipStream_t s);
hipError_t factor_b(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL, int rows,
                    int nev, int k, double power, double floor_rel, int trans, float2* out, long long ldo,
                    long long bsO, int batch, hipStream_t s);
hipError_t factor64_b(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL, int rows,
                      int nev, int k, double power, double floor_rel, int trans, double2* out, long long ldo,
                      long long bsO, int batch, hipStream_t s);
hipError_t permute(const float2* in, float2* out, int rank, const int* dims, const int* perm, hipStream_t s,
                   int batch = 1, long long bstride = 0);
hipError_t keep_count(const double* lam, long long ldw, int nev, int len_s, int msv, double err, int relative,
                      int* k_out, double* disc_out, int batch, hipStream_t s);
hipError_t factor(const double2* V, long long ldv, const double* lam, int rows, int nev, int k, double power,
                  double floor_rel, int trans, float2* out, long long ldo, hipStream_t s);
hipError_t factor64(const double2* V, long long ldv, const double* lam, int rows, int nev, int k, double power,
                    double floor_rel, int trans, double2* out, long long ldo, hipStream_t s);
hipError_t eye(float2* out, int rows, int cols, hipStream_t s);
hipError_t phys_op(const float2* X, float2* out, int L, long long rest, const float2* P, hipStream_t s);

hipError_t trace_prod(const float2* A, const float2* B, int n, double2* out, hipStream_t s);

// Pauli-string expectations of many strings at once (batch entry t = string t, stride bs):
// E[t][0] = 1; string t's single-site operator (codes[t n + site]: -1 identity, 0 X, 1 Y, 2 Z,
// 3 |0><0|, 4 |1><1|) on the physical index of X[t][L][2][rest]; res[t] = Re E[t][0]
hipError_t env_one(float2* E, long long bs, int count, hipStream_t s);
hipError_t phys_op_codes(const float2* X, float2* out, int L, long long rest, long long bs, int count, const int* codes,
                         int n, int site, hipStream_t s);
hipError_t take_re(const float2* E, long long bs, int count, double* res, hipStream_t s);

}  // namespace mpdo
