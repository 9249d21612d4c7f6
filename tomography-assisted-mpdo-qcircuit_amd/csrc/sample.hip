// Conditional marginals for shot sampling (MPDOSimulator/Circuit.py:297-467).
//
// The reference recontracts the whole density-matrix network once per (group, qubit) with the
// projectors of the group's history. Here the chain is swept once from the right to build the
// traced tail environments R_i, and the sampling sweep carries one left environment E per group:
//   E'_s[r, r'] = sum_{l, l', a} T[l, s, a, r] E[l, l'] conj(T[l', s, a, r'])   (projected site)
//   E'[r, r']   = E'_0 + E'_1                                                   (traced site)
//   P_s         = sum_{r, r'} E'_s[r, r'] R_{i+1}[r, r']
// P_s are the unnormalised diagonal entries the reference reads off its contraction
// (Circuit.py:321-326). Environments and products are complex128 (fp64 MFMA accumulation).
#include "mpdo.h"
#include "mpdo_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace mpdo {
mpdo_status layer_fail(mpdo_status s, const char* msg);
mpdo_status layer_ws(size_t bytes, hipStream_t s, char** base);
}  // namespace mpdo

using namespace mpdo;

namespace {

#define SCK(x)                                          \
  do {                                                  \
    if ((x) != hipSuccess) return layer_fail(MPDO_EHIP, #x); \
  } while (0)

inline size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

bool bad(const mpdo_site& t) { return !t.data || t.l <= 0 || t.a <= 0 || t.r <= 0; }

// T (complex64, [l][2][a][r]) -> W (complex128, same order) and Wp (complex128, [2][l][a][r])
__global__ void widen_split_kernel(const float2* T, int l, int a, int r, double2* W, double2* Wp) {
  const long long ar = (long long)a * r, total = (long long)l * 2 * ar;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long li = e / (2 * ar), rest = e % (2 * ar), s = rest / ar, k = rest % ar;
    const float2 v = T[e];
    const double2 w = make_double2(v.x, v.y);
    W[e] = w;
    Wp[(s * l + li) * ar + k] = w;
  }
}

// out[g] = prev[sel[g]] for l x l complex128 matrices
__global__ void gather_kernel(const double2* prev, const int* sel, long long m, double2* out) {
  const int g = blockIdx.y;
  const double2* src = prev + (long long)sel[g] * m;
  double2* dst = out + (long long)g * m;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (long long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// P[b] = Re sum_{ij} E[b][i][j] R[i][j], one block per matrix
__global__ __launch_bounds__(256) void pair_dot_kernel(const double2* E, const double2* R, long long m, double* P) {
  __shared__ double red[256];
  const double2* e = E + (long long)blockIdx.x * m;
  double acc = 0;
  for (long long k = threadIdx.x; k < m; k += blockDim.x) acc += e[k].x * R[k].x - e[k].y * R[k].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) P[blockIdx.x] = red[0];
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

}  // namespace

extern "C" {

long long mpdo_env_right_elems(int n, const mpdo_site* sites) {
  if (n <= 0 || !sites) return 0;
  long long t = 1;
  for (int i = 0; i < n; ++i) t += (long long)sites[i].l * sites[i].l;
  return t;
}

mpdo_status mpdo_env_right(int n, const mpdo_site* sites, void* out, void* stream) {
  if (n <= 0) return MPDO_OK;
  if (!sites || !out) return layer_fail(MPDO_EINVAL, "mpdo_env_right: null argument");
  size_t maxT = 0;
  for (int i = 0; i < n; ++i) {
    if (bad(sites[i])) return layer_fail(MPDO_EINVAL, "mpdo_env_right: bad site descriptor");
    if (i + 1 < n && sites[i].r != sites[i + 1].l) return layer_fail(MPDO_EINVAL, "mpdo_env_right: bond mismatch");
    maxT = std::max(maxT, (size_t)sites[i].l * 2 * sites[i].a * sites[i].r);
  }
  if (sites[0].l != 1 || sites[n - 1].r != 1)
    return layer_fail(MPDO_EINVAL, "mpdo_env_right: open boundary bonds must be 1");
  hipStream_t s = (hipStream_t)stream;
  char* base;
  const size_t bT = a256(maxT * sizeof(double2));
  mpdo_status st = layer_ws(3 * bT, s, &base);
  if (st != MPDO_OK) return st;
  double2* W = (double2*)base;
  double2* Wp = (double2*)(base + bT);
  double2* Y = (double2*)(base + 2 * bT);
  std::vector<long long> off(n + 1);
  off[0] = 0;
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (long long)sites[i].l * sites[i].l;
  double2* R = (double2*)out;
  const double2 one = make_double2(1.0, 0.0);
  SCK(hipMemcpyAsync(R + off[n], &one, sizeof one, hipMemcpyHostToDevice, s));
  for (int i = n - 1; i >= 0; --i) {
    const int l = sites[i].l, a = sites[i].a, r = sites[i].r;
    const long long tot = (long long)l * 2 * a * r;
    hipLaunchKernelGGL(widen_split_kernel, dim3(grid_for(tot)), dim3(256), 0, s, (const float2*)sites[i].data, l, a, r,
                       W, Wp);
    SCK(hipGetLastError());
    // Y[(l s a), r'] = sum_r W[(l s a), r] R_{i+1}[r, r']
    GemmArgs g = gemm_rowmajor(l * 2 * a, r, r, W, r, false, false, R + off[i + 1], r, false, false, Y, r);
    g.alpha_re = 1;
    SCK(gemm(C128_F64ACC, g, s));
    // R_i[l, l''] = sum_{(s a r')} Y[l, (s a r')] conj(W[l'', (s a r')])
    GemmArgs h = gemm_rowmajor(l, l, 2 * a * r, Y, 2 * a * r, false, false, W, 2 * a * r, true, true, R + off[i], l);
    h.alpha_re = 1;
    SCK(gemm(C128_F64ACC, h, s));
  }
  return MPDO_OK;
}

mpdo_status mpdo_sample_step(const mpdo_site* site, int groups, const void* prev, int prev_count, const int* sel,
                             int sampled, const void* R_next, void* out, double* probs, void* stream) {
  if (groups <= 0) return MPDO_OK;
  if (!site || bad(*site) || !prev || !out) return layer_fail(MPDO_EINVAL, "mpdo_sample_step: null argument");
  if (sampled && (!R_next || !probs)) return layer_fail(MPDO_EINVAL, "mpdo_sample_step: sampled site needs R_next and probs");
  if (prev_count <= 0) return layer_fail(MPDO_EINVAL, "mpdo_sample_step: prev_count must be positive");
  if (sel)
    for (int g = 0; g < groups; ++g)
      if (sel[g] < 0 || sel[g] >= prev_count) return layer_fail(MPDO_EINVAL, "mpdo_sample_step: selector out of range");
  if (!sel && groups > prev_count) return layer_fail(MPDO_EINVAL, "mpdo_sample_step: more groups than environments");
  hipStream_t s = (hipStream_t)stream;
  const int l = site->l, a = site->a, r = site->r;
  const long long ar = (long long)a * r, tot = (long long)l * 2 * ar, mm = (long long)l * l, rr = (long long)r * r;
  const size_t bT = a256(tot * sizeof(double2));
  // groups per pass: the per-group product X (l x 2ar complex128) is kept under 1 GiB per pass
  const long long per = tot * (long long)sizeof(double2);
  const int chunk = (int)std::max<long long>(
      1, std::min<long long>(std::min(groups, 65535), (1LL << 30) / std::max<long long>(per, 1)));
  const size_t bE = a256((size_t)chunk * mm * sizeof(double2));
  const size_t bX = a256((size_t)chunk * per);
  const size_t bS = a256((size_t)chunk * sizeof(int));
  const size_t bP = a256((size_t)chunk * 2 * sizeof(double));
  char* base;
  mpdo_status st = layer_ws(2 * bT + bE + bX + bS + bP, s, &base);
  if (st != MPDO_OK) return st;
  double2* W = (double2*)base;
  double2* Wp = (double2*)(base + bT);
  double2* Ein = (double2*)(base + 2 * bT);
  double2* X = (double2*)(base + 2 * bT + bE);
  int* dsel = (int*)(base + 2 * bT + bE + bX);
  double* dP = (double*)(base + 2 * bT + bE + bX + bS);
  hipLaunchKernelGGL(widen_split_kernel, dim3(grid_for(tot)), dim3(256), 0, s, (const float2*)site->data, l, a, r, W, Wp);
  SCK(hipGetLastError());
  const double2* P = (const double2*)prev;
  double2* O = (double2*)out;
  for (int g0 = 0; g0 < groups; g0 += chunk) {
    const int G = std::min(chunk, groups - g0);
    const double2* E = P + (long long)g0 * mm;
    if (sel) {
      SCK(hipMemcpyAsync(dsel, sel + g0, sizeof(int) * G, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(gather_kernel, dim3(grid_for(mm), G), dim3(256), 0, s, P, (const int*)dsel, mm, Ein);
      SCK(hipGetLastError());
      E = Ein;
    }
    if (!sampled) {
      // X[g][l', (s a r)] = sum_l E_g[l, l'] W[l, (s a r)]; out[g][r, r'] = sum_{(l' s a)} X[g] conj(W)
      GemmArgs g = gemm_rowmajor(l, 2 * ar, l, E, l, true, false, W, 2 * ar, false, false, X, 2 * ar);
      g.batch = G; g.bsA = mm; g.bsB = 0; g.bsC = tot; g.alpha_re = 1;
      SCK(gemm(C128_F64ACC, g, s));
      GemmArgs h = gemm_rowmajor(r, r, l * 2 * a, X, r, true, false, W, r, false, true, O + (long long)g0 * rr, r);
      h.batch = G; h.bsA = tot; h.bsB = 0; h.bsC = rr; h.alpha_re = 1;
      SCK(gemm(C128_F64ACC, h, s));
      continue;
    }
    for (int sv = 0; sv < 2; ++sv) {
      const double2* Ws = Wp + sv * l * ar;   // [l][(a r)]
      double2* Xs = X + sv * l * ar;           // X[g][s][l'][(a r)]
      GemmArgs g = gemm_rowmajor(l, (int)ar, l, E, l, true, false, Ws, ar, false, false, Xs, ar);
      g.batch = G; g.bsA = mm; g.bsB = 0; g.bsC = tot; g.alpha_re = 1;
      SCK(gemm(C128_F64ACC, g, s));
      // out[g][s][r, r'] = sum_{(l' a)} Xs[g][(l' a), r] conj(Ws[(l' a), r'])
      GemmArgs h = gemm_rowmajor(r, r, l * a, Xs, r, true, false, Ws, r, false, true, O + ((long long)g0 * 2 + sv) * rr, r);
      h.batch = G; h.bsA = tot; h.bsB = 0; h.bsC = 2 * rr; h.alpha_re = 1;
      SCK(gemm(C128_F64ACC, h, s));
    }
    hipLaunchKernelGGL(pair_dot_kernel, dim3(2 * G), dim3(256), 0, s, (const double2*)(O + (long long)g0 * 2 * rr),
                       (const double2*)R_next, rr, dP);
    SCK(hipGetLastError());
    SCK(hipMemcpyAsync(probs + (long long)g0 * 2, dP, sizeof(double) * 2 * G, hipMemcpyDeviceToHost, s));
    SCK(hipStreamSynchronize(s));   // dP and dsel are reused by the next pass
  }
  if (!sampled && sel) SCK(hipStreamSynchronize(s));   // dsel is host-staged per pass
  return MPDO_OK;
}

}  // extern "C"
