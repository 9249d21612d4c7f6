// The gate-time cut of the fused layer decided from the bottom of the spectrum, without a full
// eigen-decomposition of its Gram matrix.
//
// The absolute-error rule of decompositions.svd (decompositions.py:121-131, Circuit.py:120-124:
// te = sqrt(cumsum s^2) over the descending singular values, keep idx + 1 for the first idx with
// te[-1] - te[idx] <= eps) drops the largest set of smallest eigenvalues of the Gram matrix G whose
// sum stays within D* = tr - (sqrt(tr) - eps)^2 (about 2 eps sqrt(tr)). At the circuit's steady
// state it drops 0 - 7 of 1024 directions, so only the bottom of the spectrum matters:
//
//  1. G + tau I = P Lt Lt^H P^T (pivoted Cholesky, tau = 1e-12 tr), X = Lt^{-H} (pivot order);
//  2. subspace iteration with (G + tau I)^{-1} = X X^H on m = 128 columns, starting from the last
//     m columns of X (the directions the rank-revealing pivot order leaves to the end),
//     orthonormalised by Cholesky QR each step, q steps;
//  3. Rayleigh-Ritz with G itself on that subspace: Ritz values theta (each an upper bound of the
//     eigenvalue of the same rank from the bottom) and vectors, with residual norms;
//  4. the rule on the Ritz values: d dropped; then a certificate that the bottom k >= d + 1 Ritz
//     values are the bottom k eigenvalues to within the residual norm: for sigma in the gap after
//     theta_k, G + tr V_k V_k^H - sigma I positive definite (a rank-k positive update raises
//     eigenvalue j at most to eigenvalue j + k of G, so lambda_{k+1}(G) > sigma), checked by a
//     Cholesky factorisation completing all n columns; the decision must clear D* by the residual and
//     rounding margins on both sides.
// A chain that fails any step (rank loss, no gap, a decision too close to D*, the certificate)
// makes the caller run the full eigensolver for the batch. On success the projector onto the kept
// directions, P = I - V_d V_d^H, is written in complex64 (any orthonormal basis of the kept space
// gives the layer the same result; P is one, padded to n columns).
#include "mpdo_internal.h"
#include "prof.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mpdo {

namespace {

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
inline unsigned blocks_for(long long n) {
  long long g = (n + 255) / 256;
  return (unsigned)std::max<long long>(1, std::min<long long>(g, 2048));
}

__global__ void tail_trace_kernel(const double2* G, int n, double* tr) {
  __shared__ double red[256];
  const double2* Gb = G + (long long)blockIdx.x * n * n;
  double t = 0;
  for (int i = threadIdx.x; i < n; i += 256) t += Gb[(long long)i * n + i].x;
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) tr[blockIdx.x] = red[0];
}

// Gs[b] = sg G[b] + (ftr tr[b] - sig[b]) I  (sig may be null)
__global__ void tail_shift_kernel(const double2* G, int n, double sg, const double* tr, double ftr, const double* sig,
                                  double2* Gs) {
  const int b = blockIdx.y;
  const long long nn = (long long)n * n;
  const double sh = ftr * tr[b] - (sig ? sig[b] : 0.0);
  const double2* Gb = G + b * nn;
  double2* Sb = Gs + b * nn;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < nn; e += (long long)gridDim.x * 256) {
    double2 v = Gb[e];
    v.x *= sg;
    v.y *= sg;
    const long long i = e / n;
    if (e - i * n == i) v.x += sh;
    Sb[e] = v;
  }
}

// Z[b] (n x m) = columns c0 .. c0 + m - 1 of X[b] (n x n)
__global__ void tail_cols_kernel(const double2* X, int n, int c0, int m, double2* Z) {
  const int b = blockIdx.y;
  const long long tot = (long long)n * m;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
    const long long i = e / m, j = e - i * m;
    Z[b * tot + e] = X[(long long)b * n * n + i * n + c0 + j];
  }
}

// Qo[b][piv[b][t]][j] = Q[b][t][j]: pivot order back to the original index order
__global__ void tail_unpivot_kernel(const double2* Q, const int* piv, int n, int m, double2* Qo) {
  const int b = blockIdx.y;
  const long long tot = (long long)n * m;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
    const long long t = e / m, j = e - t * m;
    Qo[b * tot + (long long)piv[(long long)b * n + t] * m + j] = Q[b * tot + e];
  }
}

// rn[b][j] = || GV[b][:, j] - w[b][j] V[b][:, j] ||^2
__global__ void tail_resid_kernel(const double2* GV, const double2* V, const double* w, int n, int m, double* rn) {
  __shared__ double red[256];
  const int j = blockIdx.x, b = blockIdx.y;
  const long long off = (long long)b * n * m;
  const double wj = w[(long long)b * m + j];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double2 g = GV[off + (long long)i * m + j], v = V[off + (long long)i * m + j];
    const double dx = g.x - wj * v.x, dy = g.y - wj * v.y;
    s += dx * dx + dy * dy;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) rn[(long long)b * m + j] = red[0];
}

// Per chain: the rule on the Ritz values (ascending theta_i = w[m - 1 - i]), the certificate's k
// and sigma, and whether every margin holds (status 0) or the full eigensolver must decide (1).
__global__ void tail_decide_kernel(const double* wH, const double* rn, const double* tr, double eps, int n, int m,
                                   int* kept, double* disc, int* dcnt, int* kcnt, double* sig, int* status,
                                   int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const double* w = wH + (long long)b * m;
  const double* r = rn + (long long)b * m;
  status[b] = 1;
  dcnt[b] = 0;
  kcnt[b] = 0;
  sig[b] = 0.0;
  const double tot = tr[b];
  const double rt = sqrt(fmax(tot, 0.0));
  if (!(rt > eps)) return;
  const double dstar = tot - (rt - eps) * (rt - eps);
  // rounding of the trace against the eigenvalue sum, and of the full rule's top-down sums
  const double slack = 1e-13 * tot;
  double tail = 0.0;
  int d = 0;
  while (d < m - 2 && tail + fmax(w[m - 1 - d], 0.0) <= dstar) {
    tail += fmax(w[m - 1 - d], 0.0);
    ++d;
  }
  if (d >= m - 2) return;
  // the first k >= d + 1 whose gap to the next Ritz value clears the residual margin
  double rf2 = 0.0;
  for (int i = 0; i < d + 1; ++i) rf2 += r[m - 1 - i];
  int k = d + 1;
  for (; k < m - 1 && k <= d + 32; ++k) {
    const double gap = w[m - 1 - k] - w[m - k];
    if (gap > 4.0 * sqrt(rf2) + 2.0 * slack) break;
    rf2 += r[m - 1 - k];
  }
  if (k >= m - 1 || k > d + 32) return;
  const double rf = sqrt(rf2);
  const double tail1 = tail + fmax(w[m - 1 - d], 0.0);
  if (tail1 - (d + 1) * rf <= dstar + slack) return;   // the (d+1)-th value must not be droppable
  if (dstar - tail <= slack) return;                   // the first d must be, with room to spare
  kept[b] = n - d;
  disc[b] = tail;
  dcnt[b] = d;
  kcnt[b] = k;
  sig[b] = 0.5 * (w[m - k] + w[m - 1 - k]);
  status[b] = 0;
}

// Vk[b] (n x kmax) = sqrt(scale[b]) x the first cnt[b] Ritz vectors of Vr[b] (leading dim ldv, batch
// stride bsV, m columns in descending order): the ascending ones (columns m - 1, m - 2, ...) when
// rev, else columns 0, 1, ...; zero beyond. cnt null: kmax for every chain.
__global__ void tail_pick_kernel(const double2* Vr, int n, int m, long long ldv, long long bsV, int rev, const int* cnt,
                                 int kmax, const double* scale, double2* Vk) {
  const int b = blockIdx.y;
  const long long tot = (long long)n * kmax;
  const int c = cnt ? cnt[b] : kmax;
  const double f = scale ? sqrt(fmax(scale[b], 0.0)) : 1.0;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
    const long long i = e / kmax, j = e - i * kmax;
    double2 v = make_double2(0.0, 0.0);
    if (j < c) {
      v = Vr[(long long)b * bsV + i * ldv + (rev ? m - 1 - j : j)];
      v.x *= f;
      v.y *= f;
    }
    Vk[b * tot + e] = v;
  }
}

// P[b] = I + Tmp[b] in complex64 (Tmp = -V_d V_d^H)
__global__ void tail_proj_kernel(const double2* Tmp, int n, long long ldP, long long bsP, float2* P) {
  const int b = blockIdx.y;
  const long long nn = (long long)n * n;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < nn; e += (long long)gridDim.x * 256) {
    const long long i = e / n, j = e - i * n;
    double2 t = Tmp[b * nn + e];
    if (i == j) t.x += 1.0;
    P[b * bsP + i * ldP + j] = make_float2((float)t.x, (float)t.y);
  }
}

struct TailWS {
  double2 *Gs, *L, *X, *Z, *Q, *T, *Y, *Vr, *GV, *C, *Lc, *Xc, *H, *U;
  double *tr, *sig, *wH, *rn, *disc;
  int *piv, *rank, *pivc, *rankc, *dcnt, *kcnt, *status, *kept;
  void *pws, *pinv, *ews, *cws;
};

size_t carve_tail(int n, int m, int batch, char* base, TailWS* t) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += al256(bytes);
    return p;
  };
  const size_t B = batch, nn = (size_t)n * n * B * sizeof(double2), nm = (size_t)n * m * B * sizeof(double2),
               mm = (size_t)m * m * B * sizeof(double2);
  TailWS w{};
  w.Gs = (double2*)take(nn);
  w.L = (double2*)take(nn);
  w.X = (double2*)take(nn);
  w.Z = (double2*)take(nm);
  w.Q = (double2*)take(nm);
  w.T = (double2*)take(nm);
  w.Y = (double2*)take(nm);
  w.Vr = (double2*)take(nm);
  w.GV = (double2*)take(nm);
  w.C = (double2*)take(mm);
  w.Lc = (double2*)take(mm);
  w.Xc = (double2*)take(mm);
  w.H = (double2*)take(mm);
  w.U = (double2*)take(mm);
  w.tr = (double*)take(B * sizeof(double));
  w.sig = (double*)take(B * sizeof(double));
  w.disc = (double*)take(B * sizeof(double));
  w.wH = (double*)take((size_t)m * B * sizeof(double));
  w.rn = (double*)take((size_t)m * B * sizeof(double));
  w.piv = (int*)take((size_t)n * B * sizeof(int));
  w.rank = (int*)take(B * sizeof(int));
  w.pivc = (int*)take((size_t)m * B * sizeof(int));
  w.rankc = (int*)take(B * sizeof(int));
  w.dcnt = (int*)take(B * sizeof(int));
  w.kcnt = (int*)take(B * sizeof(int));
  w.status = (int*)take(B * sizeof(int));
  w.kept = (int*)take(B * sizeof(int));
  w.pws = take(pchol_workspace_bytes(n, batch));
  w.pinv = take(pchol_inverse_workspace_bytes(n, batch));
  w.ews = take(eigh_workspace_bytes(m, batch));
  w.cws = take(chol_workspace_bytes(batch));
  if (t) *t = w;
  return off;
}

GemmArgs mat(int M, int N, int K, int batch, const void* A, long long rsA, long long csA, long long bsA,
             const void* B, long long rsB, long long csB, long long bsB, void* C, long long rsC, long long bsC) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.A = A; g.rsA = rsA; g.csA = csA; g.bsA = bsA;
  g.B = B; g.rsB = rsB; g.csB = csB; g.bsB = bsB;
  g.C = C; g.rsC = rsC; g.csC = 1; g.bsC = bsC;
  g.alpha_re = 1.0;
  return g;
}

// Q = Z R^{-1} with Z^H Z = R^H R (one Cholesky-QR pass over n x m complex128 blocks)
template <typename WS>
hipError_t cholqr(int n, int m, int batch, const double2* Z, double2* Q, const WS& w, hipStream_t st) {
  hipError_t err;
  const long long nm = (long long)n * m, mm = (long long)m * m;
  GemmArgs g = mat(m, m, n, batch, Z, 1, m, nm, Z, m, 1, nm, w.C, m, mm);
  g.conjA = 1;
  g.herm = 1;
  if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  if ((err = chol_b(m, batch, w.C, m, mm, 1e-14, w.Lc, w.pivc, w.rankc, w.cws, st)) != hipSuccess) return err;
  if ((err = pchol_inverse_b(m, batch, w.Lc, w.pivc, w.rankc, 0, w.Xc, w.pinv, st)) != hipSuccess) return err;
  return gemm(C128_F64ACC, mat(n, m, m, batch, Z, m, 1, nm, w.Xc, m, 1, mm, Q, m, nm), st);
}

template <typename T>
hipError_t fetch(std::vector<T>& h, const T* d, int count, hipStream_t st) {
  h.resize(count);
  hipError_t err = hipMemcpyAsync(h.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

}  // namespace

size_t gram_tail_workspace_bytes(int n, int m, int batch) { return carve_tail(n, m, batch, nullptr, nullptr) + 256; }

hipError_t gram_tail_cut(int n, int batch, const double2* G, double eps, int m, int q, void* ws, hipStream_t st,
                         int* kept, double* disc, float2* P, long long ldP, long long bsP, TailResult* res) {
  *res = TailResult{};
  if (n <= 64 || n > 1024 || batch <= 0 || m < 16 || 2 * m > n || q < 1) return hipSuccess;
  hipError_t err;
  TailWS w;
  carve_tail(n, m, batch, (char*)ws, &w);
  const long long nn = (long long)n * n, nm = (long long)n * m, mm = (long long)m * m;
  const dim3 gnn(blocks_for(nn), batch), gnm(blocks_for(nm), batch);
  // (1) G + tau I = P Lt Lt^H P^T and X = Lt^{-H} (pivot order, upper triangular)
  hipLaunchKernelGGL(tail_trace_kernel, dim3(batch), dim3(256), 0, st, G, n, w.tr);
  hipLaunchKernelGGL(tail_shift_kernel, gnn, dim3(256), 0, st, G, n, 1.0, (const double*)w.tr, 1e-12,
                     (const double*)nullptr, w.Gs);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if ((err = pchol_b(n, batch, w.Gs, n, nn, 1e-14, w.L, w.piv, w.rank, w.pws, st, false)) != hipSuccess) return err;
  std::vector<int> hr;
  if ((err = fetch(hr, (const int*)w.rank, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != n) { res->reason = 1; return hipSuccess; }
  if ((err = pchol_inverse_b(n, batch, w.L, w.piv, w.rank, 2, w.X, w.pinv, st)) != hipSuccess) return err;
  // (2) subspace iteration with X X^H = (G + tau I)^{-1} (pivot order) from X's last m columns
  hipLaunchKernelGGL(tail_cols_kernel, gnm, dim3(256), 0, st, (const double2*)w.X, n, n - m, m, w.Z);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  for (int it = 0; it < q; ++it) {
    if ((err = cholqr(n, m, batch, w.Z, w.Q, w, st)) != hipSuccess) return err;
    GemmArgs g = mat(n, m, n, batch, w.X, 1, n, nn, w.Q, m, 1, nm, w.T, m, nm);   // T = X^H Q
    g.conjA = 1;
    g.tri = 2;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
    if ((err = gemm(C128_F64ACC, mat(n, m, n, batch, w.X, n, 1, nn, w.T, m, 1, nm, w.Z, m, nm), st)) !=
        hipSuccess)
      return err;   // Z = X T
  }
  if ((err = cholqr(n, m, batch, w.Z, w.Q, w, st)) != hipSuccess) return err;
  if ((err = cholqr(n, m, batch, w.Q, w.Z, w, st)) != hipSuccess) return err;   // orthonormal basis in Z
  if ((err = fetch(hr, (const int*)w.rankc, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != m) { res->reason = 2; return hipSuccess; }
  // (3) Rayleigh-Ritz with G: Qo (original order), Y = G Qo, H = Qo^H Y, H = U diag(wH) U^H
  hipLaunchKernelGGL(tail_unpivot_kernel, gnm, dim3(256), 0, st, (const double2*)w.Z, (const int*)w.piv, n, m, w.Q);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if ((err = gemm(C128_F64ACC, mat(n, m, n, batch, G, n, 1, nn, w.Q, m, 1, nm, w.Y, m, nm), st)) != hipSuccess)
    return err;
  {
    GemmArgs g = mat(m, m, n, batch, w.Q, 1, m, nm, w.Y, m, 1, nm, w.H, m, mm);
    g.conjA = 1;
    g.herm = 1;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  int sweeps = 0;
  if ((err = eigh(m, batch, w.H, m, mm, w.wH, m, w.U, m, mm, w.ews, st, &sweeps)) != hipSuccess) return err;
  if (sweeps > 60) { res->reason = 3; return hipSuccess; }
  if ((err = gemm(C128_F64ACC, mat(n, m, m, batch, w.Q, m, 1, nm, w.U, m, 1, mm, w.Vr, m, nm), st)) != hipSuccess)
    return err;   // Ritz vectors
  if ((err = gemm(C128_F64ACC, mat(n, m, m, batch, w.Y, m, 1, nm, w.U, m, 1, mm, w.GV, m, nm), st)) != hipSuccess)
    return err;   // G times the Ritz vectors
  hipLaunchKernelGGL(tail_resid_kernel, dim3(m, batch), dim3(256), 0, st, (const double2*)w.GV, (const double2*)w.Vr,
                     (const double*)w.wH, n, m, w.rn);
  // (4) the rule on the Ritz values and the certificate's parameters
  hipLaunchKernelGGL(tail_decide_kernel, dim3((batch + 63) / 64), dim3(64), 0, st, (const double*)w.wH,
                     (const double*)w.rn, (const double*)w.tr, eps, n, m, w.kept, w.disc, w.dcnt, w.kcnt, w.sig,
                     w.status, batch);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  std::vector<int> hs, hd, hk;
  if ((err = fetch(hs, (const int*)w.status, batch, st)) != hipSuccess) return err;
  if ((err = fetch(hd, (const int*)w.dcnt, batch, st)) != hipSuccess) return err;
  if ((err = fetch(hk, (const int*)w.kcnt, batch, st)) != hipSuccess) return err;
  int kmax = 0, dmax = 0, dmin = n;
  for (int b = 0; b < batch; ++b) {
    if (hs[b] != 0) { res->reason = 4; return hipSuccess; }
    kmax = std::max(kmax, hk[b]);
    dmax = std::max(dmax, hd[b]);
    dmin = std::min(dmin, hd[b]);
  }
  // certificate: G + tr V_k V_k^H - sigma I positive definite in every chain
  hipLaunchKernelGGL(tail_shift_kernel, gnn, dim3(256), 0, st, G, n, 1.0, (const double*)w.tr, 0.0,
                     (const double*)w.sig, w.Gs);
  hipLaunchKernelGGL(tail_pick_kernel, dim3(blocks_for((long long)n * kmax), batch), dim3(256), 0, st,
                     (const double2*)w.Vr, n, m, (long long)m, nm, 1, (const int*)w.kcnt, kmax, (const double*)w.tr, w.Z);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  {
    GemmArgs g = mat(n, n, kmax, batch, w.Z, kmax, 1, (long long)n * kmax, w.Z, 1, kmax, (long long)n * kmax, w.Gs, n,
                     nn);
    g.conjB = 1;
    g.herm = 1;
    g.beta_re = 1.0;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  if ((err = chol_b(n, batch, w.Gs, n, nn, 1e-14, w.L, w.piv, w.rank, w.cws, st)) != hipSuccess) return err;
  if ((err = fetch(hr, (const int*)w.rank, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != n) { res->reason = 5; return hipSuccess; }
  // accepted: kept counts, dropped weights, and P = I - V_d V_d^H where a chain drops something
  if ((err = hipMemcpyAsync(kept, w.kept, sizeof(int) * batch, hipMemcpyDeviceToDevice, st)) != hipSuccess) return err;
  if ((err = hipMemcpyAsync(disc, w.disc, sizeof(double) * batch, hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return err;
  if (dmax > 0) {
    hipLaunchKernelGGL(tail_pick_kernel, dim3(blocks_for((long long)n * dmax), batch), dim3(256), 0, st,
                       (const double2*)w.Vr, n, m, (long long)m, nm, 1, (const int*)w.dcnt, dmax, (const double*)nullptr,
                       w.Z);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    GemmArgs g = mat(n, n, dmax, batch, w.Z, dmax, 1, (long long)n * dmax, w.Z, 1, dmax, (long long)n * dmax, w.X, n,
                     nn);
    g.conjB = 1;
    g.herm = 1;
    g.alpha_re = -1.0;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
    hipLaunchKernelGGL(tail_proj_kernel, gnn, dim3(256), 0, st, (const double2*)w.X, n, ldP, bsP, P);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  res->ok = true;
  res->any_drop = dmax > 0;
  res->max_kept = n - dmin;
  return hipSuccess;
}

// =======================================================================================
// The top k eigenpairs of Gram matrices (the chi step at a pair bond keeps k = chi of about 4 chi
// directions): subspace iteration with G on m = 2k columns from a pseudo-random start (Z = G Omega),
// Cholesky QR each step, Rayleigh-Ritz by the Cholesky-preconditioned Jacobi solver on the m x m
// projection, then a certificate that the k leading Ritz values are the k largest eigenvalues to
// within the residual norm: for sigma in the gap below theta_{k-1}, sigma I - G + tr V_k V_k^H
// positive definite (a rank-k positive update lowers eigenvalue j of -G at most to eigenvalue j + k,
// so lambda_{k+1}(G) < sigma), checked by a Cholesky factorisation completing all n columns.
// =======================================================================================
namespace {

__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Om[b] (n x m): complex Gaussian entries from a counter hash (Box-Muller), the same for every call
__global__ void top_start_kernel(int n, int m, double2* Om) {
  const int b = blockIdx.y;
  const long long tot = (long long)n * m;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
    const unsigned long long h1 = mix64(2 * (unsigned long long)e + 1), h2 = mix64(2 * (unsigned long long)e + 2);
    const double u1 = ((h1 >> 11) + 1.0) * (1.0 / 9007199254740993.0), u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    Om[b * tot + e] = make_double2(r * cos(6.283185307179586 * u2), r * sin(6.283185307179586 * u2));
  }
}

// dst[b][i][j] = src[b][i][j] for i < rows, j < cols with each side's leading dim and batch stride
__global__ void top_copy_kernel(const double2* src, long long lds, long long bss, int rows, int cols, double2* dst,
                                long long ldd, long long bsd) {
  const int b = blockIdx.y;
  const long long tot = (long long)rows * cols;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < tot; e += (long long)gridDim.x * 256) {
    const long long i = e / cols, j = e - i * cols;
    dst[b * bsd + i * ldd + j] = src[b * bss + i * lds + j];
  }
}

// Per chain: the k leading Ritz values (descending wH) must stand clear of the rest by the residual
// margin; sig[b] = -sigma (the shift the certificate adds to -G), status 0 when they do.
__global__ void top_decide_kernel(const double* wH, const double* rn, const double* tr, int m, int k, double* sig,
                                  int* status, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const double* w = wH + (long long)b * m;
  const double* r = rn + (long long)b * m;
  status[b] = 1;
  sig[b] = 0.0;
  double rf2 = 0.0;
  for (int i = 0; i < k; ++i) rf2 += r[i];
  const double slack = 1e-13 * tr[b];
  const double gap = w[k - 1] - w[k];
  if (!(gap > 4.0 * sqrt(rf2) + 2.0 * slack)) return;
  sig[b] = -0.5 * (w[k - 1] + w[k]);
  status[b] = 0;
}

// The eigenvalue list the keep rule and the factors read: w[b][j] = theta_j for j < m, the
// remaining trace at j = m (so that every tail sum past the m-th value is the trace's remainder),
// zero beyond
__global__ void top_eigs_kernel(const double* wH, const double* tr, int m, int n, double* w, long long ldw) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  double t = 0;
  for (int j = threadIdx.x; j < m; j += 256) t += wH[(long long)b * m + j];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double rem = fmax(tr[b] - red[0], 0.0);
  for (int j = threadIdx.x; j < n; j += 256)
    w[(long long)b * ldw + j] = j < m ? wH[(long long)b * m + j] : (j == m ? rem : 0.0);
}

struct TopWS {
  double2 *Gs, *L, *Z, *Q, *Y, *Vr, *GV, *C, *Lc, *Xc, *H, *U;
  double *tr, *sig, *wH, *rn;
  int *piv, *rank, *pivc, *rankc, *status;
  void *pws, *pinv, *ews, *cws;
};

size_t carve_top(int n, int m, int batch, char* base, TopWS* t) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> char* {
    char* p = base ? base + off : nullptr;
    off += al256(bytes);
    return p;
  };
  const size_t B = batch, nn = (size_t)n * n * B * sizeof(double2), nm = (size_t)n * m * B * sizeof(double2),
               mm = (size_t)m * m * B * sizeof(double2);
  TopWS w{};
  w.Gs = (double2*)take(nn);
  w.L = (double2*)take(nn);
  w.Z = (double2*)take(nm);
  w.Q = (double2*)take(nm);
  w.Y = (double2*)take(nm);
  w.Vr = (double2*)take(nm);
  w.GV = (double2*)take(nm);
  w.C = (double2*)take(mm);
  w.Lc = (double2*)take(mm);
  w.Xc = (double2*)take(mm);
  w.H = (double2*)take(mm);
  w.U = (double2*)take(mm);
  w.tr = (double*)take(B * sizeof(double));
  w.sig = (double*)take(B * sizeof(double));
  w.wH = (double*)take((size_t)m * B * sizeof(double));
  w.rn = (double*)take((size_t)m * B * sizeof(double));
  w.piv = (int*)take((size_t)n * B * sizeof(int));
  w.rank = (int*)take(B * sizeof(int));
  w.pivc = (int*)take((size_t)m * B * sizeof(int));
  w.rankc = (int*)take(B * sizeof(int));
  w.status = (int*)take(B * sizeof(int));
  w.pws = take(pchol_workspace_bytes(n, batch));
  w.pinv = take(pchol_inverse_workspace_bytes(m, batch));
  w.ews = take(eigh_gram_workspace_bytes(m, batch));
  w.cws = take(chol_workspace_bytes(batch));
  if (t) *t = w;
  return off;
}

}  // namespace

size_t gram_top_workspace_bytes(int n, int m, int batch) { return carve_top(n, m, batch, nullptr, nullptr) + 256; }

hipError_t gram_top_eig(int n, int batch, const double2* G, int k, int m, int q, void* ws, hipStream_t st, double* w,
                        long long ldw, double2* V, long long ldv, long long bsV, TopResult* res) {
  *res = TopResult{};
  if (n <= 64 || n > 1024 || batch <= 0 || k < 1 || m < k + 16 || m + 64 > n || q < 1) return hipSuccess;
  hipError_t err;
  TopWS t;
  carve_top(n, m, batch, (char*)ws, &t);
  const long long nn = (long long)n * n, nm = (long long)n * m, mm = (long long)m * m;
  const dim3 gnn(blocks_for(nn), batch), gnm(blocks_for(nm), batch);
  hipLaunchKernelGGL(tail_trace_kernel, dim3(batch), dim3(256), 0, st, G, n, t.tr);
  // Z = G Omega, then q steps of Q = cholqr(Z), Z = G Q
  hipLaunchKernelGGL(top_start_kernel, gnm, dim3(256), 0, st, n, m, t.Q);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if ((err = gemm(C128_F64ACC, mat(n, m, n, batch, G, n, 1, nn, t.Q, m, 1, nm, t.Z, m, nm), st)) != hipSuccess)
    return err;
  for (int it = 0; it < q; ++it) {
    if ((err = cholqr(n, m, batch, t.Z, t.Q, t, st)) != hipSuccess) return err;
    if ((err = gemm(C128_F64ACC, mat(n, m, n, batch, G, n, 1, nn, t.Q, m, 1, nm, t.Z, m, nm), st)) != hipSuccess)
      return err;
  }
  if ((err = cholqr(n, m, batch, t.Z, t.Q, t, st)) != hipSuccess) return err;
  if ((err = cholqr(n, m, batch, t.Q, t.Z, t, st)) != hipSuccess) return err;   // orthonormal basis in Z
  std::vector<int> hr;
  if ((err = fetch(hr, (const int*)t.rankc, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != m) { res->reason = 2; return hipSuccess; }
  // Rayleigh-Ritz: Y = G Z, H = Z^H Y = U diag(wH) U^H (descending), Ritz vectors Vr = Z U
  if ((err = gemm(C128_F64ACC, mat(n, m, n, batch, G, n, 1, nn, t.Z, m, 1, nm, t.Y, m, nm), st)) != hipSuccess)
    return err;
  {
    GemmArgs g = mat(m, m, n, batch, t.Z, 1, m, nm, t.Y, m, 1, nm, t.H, m, mm);
    g.conjA = 1;
    g.herm = 1;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  int sweeps = 0;
  if ((err = eigh_gram(m, batch, t.H, m, mm, t.wH, m, t.U, m, mm, t.ews, st, &sweeps, 1e-14)) != hipSuccess)
    return err;
  if (sweeps > 60) { res->reason = 3; return hipSuccess; }
  if ((err = gemm(C128_F64ACC, mat(n, m, m, batch, t.Z, m, 1, nm, t.U, m, 1, mm, t.Vr, m, nm), st)) != hipSuccess)
    return err;
  if ((err = gemm(C128_F64ACC, mat(n, m, m, batch, t.Y, m, 1, nm, t.U, m, 1, mm, t.GV, m, nm), st)) != hipSuccess)
    return err;
  hipLaunchKernelGGL(tail_resid_kernel, dim3(m, batch), dim3(256), 0, st, (const double2*)t.GV, (const double2*)t.Vr,
                     (const double*)t.wH, n, m, t.rn);
  hipLaunchKernelGGL(top_decide_kernel, dim3((batch + 63) / 64), dim3(64), 0, st, (const double*)t.wH,
                     (const double*)t.rn, (const double*)t.tr, m, k, t.sig, t.status, batch);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if ((err = fetch(hr, (const int*)t.status, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != 0) { res->reason = 4; return hipSuccess; }
  // certificate: sigma I - G + tr V_k V_k^H positive definite in every chain
  hipLaunchKernelGGL(tail_shift_kernel, gnn, dim3(256), 0, st, G, n, -1.0, (const double*)t.tr, 0.0,
                     (const double*)t.sig, t.Gs);
  hipLaunchKernelGGL(tail_pick_kernel, dim3(blocks_for((long long)n * k), batch), dim3(256), 0, st,
                     (const double2*)t.Vr, n, m, (long long)m, nm, 0, (const int*)nullptr, k, (const double*)t.tr, t.Q);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  {
    GemmArgs g = mat(n, n, k, batch, t.Q, k, 1, (long long)n * k, t.Q, 1, k, (long long)n * k, t.Gs, n, nn);
    g.conjB = 1;
    g.herm = 1;
    g.beta_re = 1.0;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  if ((err = chol_b(n, batch, t.Gs, n, nn, 1e-14, t.L, t.piv, t.rank, t.cws, st)) != hipSuccess) return err;
  if ((err = fetch(hr, (const int*)t.rank, batch, st)) != hipSuccess) return err;
  for (int b = 0; b < batch; ++b)
    if (hr[b] != n) { res->reason = 5; return hipSuccess; }
  // accepted: the eigenvalue list and the m Ritz vectors in V's leading columns
  hipLaunchKernelGGL(top_eigs_kernel, dim3(batch), dim3(256), 0, st, (const double*)t.wH, (const double*)t.tr, m, n, w,
                     ldw);
  hipLaunchKernelGGL(top_copy_kernel, gnm, dim3(256), 0, st, (const double2*)t.Vr, (long long)m, nm, n, m, V, ldv,
                     bsV);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  res->ok = true;
  return hipSuccess;
}

}  // namespace mpdo
