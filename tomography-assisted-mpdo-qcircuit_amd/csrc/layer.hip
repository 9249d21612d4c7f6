// One brickwork circuit-layer for a batch of MPDO chains with identical shapes, without ever
// forming the (up to 4*chi)-bond tensor that the reference's gate-time split gives each pair.
//
// Reference schedule being reproduced (Circuit.py:74-178, :469-489; TNNOptimizer.py:72-197):
//   1q channels on every site -> 2q channel + SVD split on each pair (i, i+1) -> truncate():
//   QR sweep left to right, sqrt(S)-split SVD sweep right to left with the chi rule, kappa SVD.
// Equivalent computation (exact in exact arithmetic; shapes identical to the per-gate path):
//   The gate-time split of the reference (Circuit.py:110-124: SVD of the two-site tensor theta,
//   absolute tail 2.7e-8, no chi cap) is reproduced in compressed form: T_min = Q1t R1t (QR),
//   T_max = L2 Q2 (LQ), theta = (Q1t (x) I) C_theta (I (x) Q2) with C_theta = G(R1t (x)_m L2); the
//   Gram of C_theta gives theta's singular values, the kept count m' and, when a direction is
//   dropped, the projector P onto the kept left singular vectors. The split tensors themselves
//   (bond up to 4 chi) are never formed: the QR sweep needs from the pair only
//   X1r = (R (x) I) Q1t = Q1 Ra and R_max, a factor of the right-bond Gram of R theta', which is
//   sum_a Q2_a^H (Cu^H Cu summed over the channel/physical index) Q2_a with the core
//   Cu = (Ra (x) I) P C_theta. The chi step at the pair bond is the truncated SVD of
//   Theta = R theta' R_max^{-1} P_right = (Q1 (x) I) Cu (I (x) Q2 P_right), read off the Gram
//   Cu (I (x) H) Cu^H of the core (H the Gram of Q2 P_right, so that factor is never factored):
//   U sqrt(S) gives the new min site, S^{-1/2} U^H Cu (I (x) Q2 P_right) the new max site. The
//   kappa cut of every site runs as soon as the site is final, so the large max-site tensor
//   exists for one pair at a time. Cu's 2 K2 p2 columns are the channel's image of the 4 blocks
//   of Y = R1t L2, so both Grams of Cu come from 4 p1t x 4 p1t Grams of Y's blocks and the 8 x 8
//   Gram of the channel; Cu is formed only when the new max site is not kappa-cut in factored
//   form.
// Every stage is batched over the chains (same shapes; rank rules deterministic because no
// error threshold is used).
#include "mpdo.h"
#include "mpdo_internal.h"
#include "site_ops.h"

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <string>
#include <vector>

namespace mpdo {
// shared with api.hip
mpdo_status layer_fail(mpdo_status s, const char* msg);
mpdo_status layer_ws(size_t bytes, hipStream_t s, char** base);
extern Prec g_tensor_prec;
}  // namespace mpdo

using namespace mpdo;

namespace {

#define LCK(x)                                                                              \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      char b_[512];                                                                         \
      snprintf(b_, sizeof b_, "%s:%d %s -> %s %s", __FILE__, __LINE__, #x, hipGetErrorString(e_),       \
               t_oob.c_str());                                                                      \
      return layer_fail(MPDO_EHIP, b_);                                                     \
    }                                                                                       \
  } while (0)
#define LST(x)                     \
  do {                             \
    mpdo_status s_ = (x);          \
    if (s_ != MPDO_OK) return s_;  \
  } while (0)

constexpr double FLOOR = 1e-13;         // relative eigenvalue floor of the QR-type factorisations
constexpr double SPLIT_FLOOR = 1e-12;   // the same floor as the per-gate sqrt(S) splits (s_j < 1e-6 s_0)
// absolute error bound of the gate-time split (Circuit.py:22,120-124): with te = sqrt(cumsum s^2)
// the split keeps idx+1 singular values for the first idx with te[-1] - te[idx] <= GATE_EPS
// (decompositions.py:121-131)
constexpr double GATE_EPS = 2.718281828459045e-8;

// ------------------------------------------------------------------ host-side bounds checks
// Every device buffer the layer touches (each carved workspace buffer, each input and output
// site) is registered as a span; every launch below checks the extent of each operand it
// reads or writes against the span its pointer falls in and refuses to launch on overflow.
struct Span {
  uintptr_t lo, hi;
};
thread_local std::vector<Span>* t_spans = nullptr;
thread_local std::string t_oob;
// dry run: every check is made, nothing is launched or copied (host-only validation of a plan)
thread_local bool t_dry = false;
thread_local size_t t_ws_bytes = 0;   // workspace size of the last planned layer
thread_local std::vector<int> t_kept;   // per pair of the last layer: max kept count of its gate-time split
// spectra of chain 0 recorded on request: per gate-time cut and per chi step, its order n followed
// by its n eigenvalues (descending)
thread_local bool t_rec_gate = false;
thread_local std::vector<double> t_gate_eigs;
thread_local std::vector<double> t_chi_eigs;
mpdo_status record_eigs(std::vector<double>& dst, const double* w, int n, hipStream_t s) {
  const size_t o = dst.size();
  dst.resize(o + 1 + n);
  dst[o] = n;
  LCK(hipMemcpyAsync(dst.data() + o + 1, w, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  LCK(hipStreamSynchronize(s));
  return MPDO_OK;
}
thread_local std::vector<int> t_gemms;  // dry run: (source line, M, N, K, batch, precision) of every GEMM
thread_local int t_gate_skips = 0;     // gate-time cuts decided by the shifted Cholesky (last layer)
thread_local int t_gate_tails = 0;     // gate-time cuts decided from the bottom of the spectrum (last layer)
bool g_gate_skip = true;               // mpdo_debug_layer_gate_skip
bool g_gate_tail = true;               // mpdo_debug_layer_gate_tail
// the bottom-of-the-spectrum cut: subspace columns and iteration steps (mpdo_debug_layer_tail_shape);
// below order PARTIAL_MIN the full eigensolver costs less than the partial solvers' launches and
// readbacks. Measured on two steady-state layers of the 32-qubit chi = 256,
// kappa = 8 benchmark (18 cuts decided this way): 128 columns with 3, 4 or 5 steps and 96 with 4
// accept every cut (8.39-8.46 circuit-layers/s, within the run-to-run spread; the fifth step keeps
// the benchmark's observables closest to the fp64 per-gate path); 64 columns decline 2-4 at 4-5
// steps, 32 most.
constexpr int PARTIAL_MIN = 512;
int g_tail_m = 128, g_tail_q = 5;
// the chi step at a pair bond by its top k = chi eigenpairs: k + g_top_extra subspace columns
// (rounded up to a multiple of 64; 2k when g_top_extra = 0) and g_top_q iteration steps
// (mpdo_debug_layer_top_shape). Measured on two steady-state layers of the 32-qubit chi = 256,
// kappa = 8 benchmark (k = 256 of 1024 and 128 of 512): k + 128 columns and 5 steps certify all
// 29 chi steps at 8.39 circuit-layers/s; 2k columns and 4 steps 7.61 (2 declined: the wider
// subspace reaches the part of the spectrum below the Rayleigh-Ritz solver's floor); k + 128 with
// 4 steps 7.27 (5 declined); k + 192 with 4 or 5 steps 8.01 / 7.86; k + 32 with 6 or 7 steps
// about 5 (most declined).
int g_top_extra = 128;
int g_top_q = 5;
thread_local int t_chi_tops = 0;       // pair-bond chi steps decided that way (last layer)
bool g_chi_top = true;                 // mpdo_debug_layer_chi_top
int top_m(int n, int k) {
  const int m = g_top_extra > 0 ? (k + g_top_extra + 63) / 64 * 64 : 2 * k;
  return (k >= 16 && m + 64 <= n && n >= PARTIAL_MIN && n <= 1024) ? m : 0;
}

void add_span(const void* p, size_t bytes) {
  if (t_spans && p && bytes) t_spans->push_back(Span{(uintptr_t)p, (uintptr_t)p + bytes});
}
// operand of `elems` elements of `esz` bytes starting at p lies inside one registered span
bool fits(const void* p, long long elems, size_t esz, const char* what, int line) {
  if (!t_spans || elems <= 0) return true;
  const uintptr_t lo = (uintptr_t)p, hi = lo + (uintptr_t)elems * esz;
  for (const Span& sp : *t_spans)
    if (lo >= sp.lo && lo < sp.hi) {
      if (hi <= sp.hi) return true;
      char b[256];
      snprintf(b, sizeof b, "layer.hip:%d %s overruns its buffer by %llu bytes", line, what,
               (unsigned long long)(hi - sp.hi));
      t_oob = b;
      return false;
    }
  char b[256];
  snprintf(b, sizeof b, "layer.hip:%d %s points outside every registered buffer", line, what);
  t_oob = b;
  return false;
}
// elements spanned by a strided batch of rows x cols matrices (non-negative strides)
long long ext(int batch, long long bs, long long rows, long long rs, long long cols, long long cs) {
  if (batch <= 0 || rows <= 0 || cols <= 0) return 0;
  return (batch - 1) * bs + (rows - 1) * rs + (cols - 1) * cs + 1;
}
#define FITS(p, n, T, what) \
  if (!fits((p), (n), sizeof(T), (what), __LINE__)) return hipErrorInvalidValue

size_t in_bytes(Prec p) { return (p == C128_F64ACC || p == C128_IN_C64_OUT) ? 16 : 8; }
size_t out_bytes(Prec p) { return (p == C64_F64ACC || p == C128_F64ACC) ? 16 : 8; }

size_t op_bytes(Prec p, int flag) { return flag ? (flag == 2 ? 16 : 8) : in_bytes(p); }

hipError_t k_gemm(Prec p, const GemmArgs& g, hipStream_t s, int line) {
  const long long sp = g.ksplit > 1 ? g.ksplit - 1 : 0;   // extra reduction slices
  if (!fits(g.A, ext(g.batch, g.bsA, g.M, g.rsA, g.K, g.csA) + sp * g.ksA, op_bytes(p, g.inA), "GEMM operand A",
            line) ||
      !fits(g.B, ext(g.batch, g.bsB, g.K, g.rsB, g.N, g.csB) + sp * g.ksB, op_bytes(p, g.inB), "GEMM operand B",
            line) ||
      !fits(g.C, ext(g.batch, g.bsC, g.M, g.rsC, g.N, g.csC) + sp * g.ksC, out_bytes(p), "GEMM result C", line))
    return hipErrorInvalidValue;
  if (sp && (g.ksA < 0 || g.ksB < 0 || g.ksC < 0)) return hipErrorInvalidValue;
  if (g.herm && g.M != g.N) {
    char b[128];
    snprintf(b, sizeof b, "layer.hip:%d Hermitian GEMM output is not square", line);
    t_oob = b;
    return hipErrorInvalidValue;
  }
  if (t_dry) {
    t_gemms.insert(t_gemms.end(), {line, g.M, g.N, g.K * (g.ksplit > 1 ? g.ksplit : 1), g.batch, (int)p});
    return hipSuccess;
  }
  return gemm(p, g, s);
}

// Partial-product buffer of the split-reduction Hermitian GEMMs of the running layer
// (SPLITK_PARTS complex128 per chain: 8 slices of 256 x 256, or 2 of 512 x 512)
constexpr long long SPLITK_PARTS = 8LL * 256 * 256;
thread_local double2* t_split_parts = nullptr;
thread_local long long t_split_cap = 0;   // complex128 elements

// A Hermitian product G (n x n complex128, contiguous per batch entry): with few output tiles
// and a long reduction, the reduction is split into S slices whose partial products run as one
// launch (enough workgroups to fill the chip), then summed in a fixed order and mirrored.
hipError_t herm_gemm_(int line, Prec p, GemmArgs g, hipStream_t s) {
  const int n = g.M;
  const long long nn = (long long)n * n;
  const long long T = (n + 63) / 64, tiles = T * (T + 1) / 2;
  int S = 1;
  if (g.herm && t_split_parts && out_bytes(p) == 16 && g.rsC == n && g.csC == 1 && g.bsC == nn)
    while (S < 8 && g.K % (2 * S) == 0 && g.K / (2 * S) >= 2048 && tiles * g.batch * 2 * S <= 4096 &&
           2LL * S * nn * g.batch <= t_split_cap)
      S *= 2;
  if (S == 1) return k_gemm(p, g, s, line);
  double2* G = (double2*)g.C;
  GemmArgs h = g;
  h.K = g.K / S;
  h.ksplit = S;
  h.ksA = (long long)h.K * g.csA;
  h.ksB = (long long)h.K * g.rsB;
  h.C = t_split_parts;
  h.bsC = S * nn;
  h.ksC = nn;
  hipError_t e = k_gemm(p, h, s, line);
  if (e != hipSuccess) return e;
  if (!fits(t_split_parts, g.batch * S * nn, sizeof(double2), "split-reduction partials", line) ||
      !fits(G, g.batch * nn, sizeof(double2), "split-reduction output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  if ((e = sum_parts(t_split_parts, S, nn, G, g.batch, s)) != hipSuccess) return e;
  return herm_mirror_b(G, true, n, nn, g.batch, s);
}
#define K_GEMM(p, g, s) k_gemm((p), (g), (s), __LINE__)

hipError_t k_factor(const double2* V, long long ldv, long long bsV, const double* lam, long long bsL, int rows,
                    int nev, int k, double power, double floor_rel, int trans, void* out, long long ldo,
                    long long bsO, int batch, hipStream_t s, bool f64, int line) {
  const int kv = std::min(k, nev);
  if (kv > 0 && (!fits(V, ext(batch, bsV, rows, ldv, kv, 1), sizeof(double2), "factor V", line) ||
                 !fits(lam, ext(batch, bsL, 1, 0, kv, 1), sizeof(double), "factor lambda", line)))
    return hipErrorInvalidValue;
  const long long eo = trans ? ext(batch, bsO, k, ldo, rows, 1) : ext(batch, bsO, rows, ldo, k, 1);
  if (!fits(out, eo, f64 ? sizeof(double2) : sizeof(float2), "factor output", line)) return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return f64 ? factor64_b(V, ldv, bsV, lam, bsL, rows, nev, k, power, floor_rel, trans, (double2*)out, ldo, bsO,
                          batch, s)
             : factor_b(V, ldv, bsV, lam, bsL, rows, nev, k, power, floor_rel, trans, (float2*)out, ldo, bsO, batch,
                        s);
}
#define K_FACTOR(V, ldv, bsV, lam, bsL, rows, nev, k, pw, fl, tr, out, ldo, bsO, batch, s) \
  k_factor(V, ldv, bsV, lam, bsL, rows, nev, k, pw, fl, tr, out, ldo, bsO, batch, s, false, __LINE__)
#define K_FACTOR64(V, ldv, bsV, lam, bsL, rows, nev, k, pw, fl, tr, out, ldo, bsO, batch, s) \
  k_factor(V, ldv, bsV, lam, bsL, rows, nev, k, pw, fl, tr, out, ldo, bsO, batch, s, true, __LINE__)

hipError_t k_permute(const float2* in, float2* out, int rank, const int* dims, const int* perm, hipStream_t s,
                     int batch, long long bstride, int line) {
  long long tot = 1;
  for (int i = 0; i < rank; ++i) tot *= dims[i];
  const long long e = (batch - 1) * bstride + tot;
  if (!fits(in, e, sizeof(float2), "permute input", line) || !fits(out, e, sizeof(float2), "permute output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return permute(in, out, rank, dims, perm, s, batch, bstride);
}
#define K_PERMUTE(in, out, rank, dims, perm, s, batch, bs) k_permute(in, out, rank, dims, perm, s, batch, bs, __LINE__)

hipError_t k_keep(const double* lam, long long ldw, int nev, int len_s, int msv, double err, int relative, int* k_out,
                  double* disc_out, int batch, hipStream_t s, int line) {
  if (!fits(lam, ext(batch, ldw, 1, 0, std::min(nev, len_s), 1), sizeof(double), "keep_count lambda", line) ||
      !fits(k_out, batch, sizeof(int), "keep_count k", line) ||
      !fits(disc_out, batch, sizeof(double), "keep_count discarded", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return keep_count(lam, ldw, nev, len_s, msv, err, relative, k_out, disc_out, batch, s);
}
#define K_KEEP(lam, ldw, nev, len_s, msv, err, rel, k_out, disc, batch, s) \
  k_keep(lam, ldw, nev, len_s, msv, err, rel, k_out, disc, batch, s, __LINE__)

hipError_t k_eye(float2* out, int rows, int cols, long long bs, int batch, hipStream_t s, int line) {
  FITS(out, (batch - 1) * bs + (long long)rows * cols, float2, "identity output");
  (void)line;
  if (t_dry) return hipSuccess;
  return eye_b(out, rows, cols, bs, batch, s);
}
#define K_EYE(out, rows, cols, bs, batch, s) k_eye(out, rows, cols, bs, batch, s, __LINE__)

hipError_t k_channel2q(const float2* Y, const float2* Gt, int K, int P1, int P2, float2* C, hipStream_t s, int batch,
                       long long bsY, long long bsC, int line) {
  if (!fits(Y, (batch - 1) * bsY + 4LL * P1 * P2, sizeof(float2), "channel input", line) ||
      !fits(Gt, 16LL * K, sizeof(float2), "channel tensor", line) ||
      !fits(C, (batch - 1) * bsC + 4LL * P1 * K * P2, sizeof(float2), "channel output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return channel2q(Y, Gt, K, P1, P2, C, s, batch, bsY, bsC);
}

hipError_t k_bds(const float2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch, hipStream_t s,
                 int line) {
  const long long ar = (long long)A * R;
  if (!fits(M, (batch - 1) * bsM + ar * ar, sizeof(float2), "block-diagonal input", line) ||
      !fits(out, (batch - 1) * bsO + (long long)R * R, sizeof(double2), "block-diagonal output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return block_diag_sum(M, A, R, out, bsM, bsO, batch, s);
}

hipError_t k_bds64(const double2* M, int A, int R, double2* out, long long bsM, long long bsO, int batch,
                   hipStream_t s, int line) {
  const long long ar = (long long)A * R;
  if (!fits(M, (batch - 1) * bsM + ar * ar, sizeof(double2), "block-diagonal input", line) ||
      !fits(out, (batch - 1) * bsO + (long long)R * R, sizeof(double2), "block-diagonal output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return block_diag_sum64(M, A, R, out, bsM, bsO, batch, s);
}

// out[b][e] = in[b][e] (conjugated when conj) in complex64, e < n
__global__ void narrow_c128_kernel(const double2* in, float2* out, long long n, int conj) {
  const int b = blockIdx.y;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const double2 v = in[b * n + e];
    out[b * n + e] = make_float2((float)v.x, conj ? (float)-v.y : (float)v.y);
  }
}

hipError_t k_narrow(const double2* in, float2* out, long long n, int conj, int batch, hipStream_t s, int line) {
  if (!fits(in, (long long)batch * n, sizeof(double2), "narrow input", line) ||
      !fits(out, (long long)batch * n, sizeof(float2), "narrow output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  const unsigned g = (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 2048));
  hipLaunchKernelGGL(narrow_c128_kernel, dim3(g, batch), dim3(256), 0, s, in, out, n, conj);
  return hipGetLastError();
}

hipError_t k_widen(const float2* in, double2* out, long long n, long long bsI, long long bsO, int batch, hipStream_t s,
                   int line) {
  if (!fits(in, (batch - 1) * bsI + n, sizeof(float2), "widen input", line) ||
      !fits(out, (batch - 1) * bsO + n, sizeof(double2), "widen output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return widen(in, out, n, bsI, bsO, batch, s);
}

hipError_t k_mask(const double2* V, long long ldv, long long bsV, int rows, int cols, const int* k, float2* out,
                  long long ldo, long long bsO, int batch, hipStream_t s, int line) {
  if (!fits(V, ext(batch, bsV, rows, ldv, cols, 1), sizeof(double2), "mask input", line) ||
      !fits(k, batch, sizeof(int), "mask counts", line) ||
      !fits(out, ext(batch, bsO, rows, ldo, cols, 1), sizeof(float2), "mask output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return mask_cols_b(V, ldv, bsV, rows, cols, k, out, ldo, bsO, batch, s);
}

hipError_t k_kraus_gram(const double2* Pg, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s,
                        int line) {
  const long long n = 2LL * P1;
  if (!fits(Pg, batch * 4 * n * n, sizeof(double2), "core Gram blocks", line) ||
      !fits(Gt, 16LL * K, sizeof(float2), "core Gram channel", line) ||
      !fits(out, batch * n * n, sizeof(double2), "core Gram output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return kraus_gram(Pg, Gt, K, P1, out, batch, s);
}

hipError_t k_kraus_dual(const double2* N, const float2* Gt, int K, int P1, double2* out, int batch, hipStream_t s,
                        int line) {
  const long long n = 2LL * P1;
  if (!fits(N, batch * n * n, sizeof(double2), "dual core input", line) ||
      !fits(Gt, 16LL * K, sizeof(float2), "dual core channel", line) ||
      !fits(out, batch * 4 * n * n, sizeof(double2), "dual core output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return kraus_dual(N, Gt, K, P1, out, batch, s);
}

hipError_t k_kron_i2(const float2* R, int P, int Q, double2* out, int batch, hipStream_t s, int line) {
  if (!fits(R, (long long)batch * P * Q, sizeof(float2), "kron input", line) ||
      !fits(out, (long long)batch * 4 * P * Q, sizeof(double2), "kron output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return kron_i2(R, P, Q, out, batch, s);
}

hipError_t k_kappa_gram(const double2* Hg, const float2* Gt, int K, int A, double2* out, int batch, hipStream_t s,
                        int line) {
  const long long n = (long long)A * K;
  if (!fits(Hg, batch * 64LL * A * A, sizeof(double2), "kappa core Gram input", line) ||
      !fits(Gt, 16LL * K, sizeof(float2), "kappa core channel", line) ||
      !fits(out, batch * n * n, sizeof(double2), "kappa core Gram output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return kappa_core_gram(Hg, Gt, K, A, out, batch, s);
}

hipError_t k_kappa_vec(const float2* Vk, const float2* Gt, int K, int A, int kk, float2* Vt, int batch, hipStream_t s,
                       int line) {
  if (!fits(Vk, (long long)batch * A * K * kk, sizeof(float2), "kappa core vectors", line) ||
      !fits(Gt, 16LL * K, sizeof(float2), "kappa core channel", line) ||
      !fits(Vt, (long long)batch * 16 * A * kk, sizeof(float2), "kappa core mapped vectors", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return kappa_core_vec(Vk, Gt, K, A, kk, Vt, batch, s);
}

hipError_t k_narrow_factor(const double2* L, int n, int mode, float2* out, int batch, hipStream_t s, int line) {
  if (!fits(L, (long long)batch * n * n, sizeof(double2), "factor input", line) ||
      !fits(out, (long long)batch * n * n, sizeof(float2), "factor output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return narrow_factor(L, n, mode, out, batch, s);
}

hipError_t k_copy(void* dst, const void* src, size_t bytes, hipStream_t s, int line) {
  if (!fits(dst, (long long)bytes, 1, "copy destination", line) || !fits(src, (long long)bytes, 1, "copy source", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
}
#define K_COPY(dst, src, bytes, s) k_copy(dst, src, bytes, s, __LINE__)

struct Carver {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* get(size_t n) {
    size_t b = (n * sizeof(T) + 255) & ~size_t(255);
    T* p = base ? (T*)(base + off) : nullptr;
    if (p) add_span(p, n * sizeof(T));
    off += b;
    return p;
  }
};

// ------------------------------------------------------------------ batched helpers
hipError_t mmb_(int line, Prec p, int M, int N, int K, const void* A, long long lda, long long bsA, const void* B,
               long long ldb, long long bsB, void* C, long long ldc, long long bsC, int batch, hipStream_t s,
               bool conjTransA = false, bool herm = false) {
  GemmArgs g{};
  g.herm = herm;
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.A = A; g.rsA = conjTransA ? 1 : lda; g.csA = conjTransA ? lda : 1; g.bsA = bsA; g.conjA = conjTransA;
  g.B = B; g.rsB = ldb; g.csB = 1; g.bsB = bsB;
  g.C = C; g.rsC = ldc; g.csC = 1; g.bsC = bsC;
  g.alpha_re = 1;
  return k_gemm(p, g, s, line);
}
#define mmb(...) mmb_(__LINE__, __VA_ARGS__)
// the same with the storage of each operand given (1 = complex64, 2 = complex128), row-major
hipError_t mmx_(int line, Prec p, int M, int N, int K, const void* A, long long lda, long long bsA, int inA,
                const void* B, long long ldb, long long bsB, int inB, void* C, long long ldc, long long bsC, int batch,
                hipStream_t s, bool conjTransA = false, bool herm = false) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.A = A; g.rsA = conjTransA ? 1 : lda; g.csA = conjTransA ? lda : 1; g.bsA = bsA; g.conjA = conjTransA;
  g.B = B; g.rsB = ldb; g.csB = 1; g.bsB = bsB;
  g.C = C; g.rsC = ldc; g.csC = 1; g.bsC = bsC;
  g.alpha_re = 1;
  g.inA = inA; g.inB = inB; g.herm = herm;
  return k_gemm(p, g, s, line);
}
#define mmx(...) mmx_(__LINE__, __VA_ARGS__)
// G[b] (n x n, complex128) = X[b]^H X[b], X (rows x n) row-major
hipError_t gram_cols_b_(int line, const float2* X, int rows, int n, long long bsX, double2* G, int batch, hipStream_t s) {
  GemmArgs g{};
  g.M = n; g.N = n; g.K = rows; g.batch = batch;
  g.A = X; g.rsA = 1; g.csA = n; g.bsA = bsX; g.conjA = 1;
  g.B = X; g.rsB = n; g.csB = 1; g.bsB = bsX;
  g.C = G; g.rsC = n; g.csC = 1; g.bsC = (long long)n * n;
  g.alpha_re = 1;
  g.herm = 1;
  return herm_gemm_(line, C64_F64ACC, g, s);
}
// G[b] (n x n, complex128) = X[b] X[b]^H, X (n x cols) row-major
hipError_t gram_rows_b_(int line, const float2* X, int n, int cols, long long bsX, double2* G, int batch, hipStream_t s) {
  GemmArgs g{};
  g.M = n; g.N = n; g.K = cols; g.batch = batch;
  g.A = X; g.rsA = cols; g.csA = 1; g.bsA = bsX;
  g.B = X; g.rsB = 1; g.csB = cols; g.bsB = bsX; g.conjB = 1;
  g.C = G; g.rsC = n; g.csC = 1; g.bsC = (long long)n * n;
  g.alpha_re = 1;
  g.herm = 1;
  return herm_gemm_(line, C64_F64ACC, g, s);
}

#define gram_cols_b(...) gram_cols_b_(__LINE__, __VA_ARGS__)
#define gram_rows_b(...) gram_rows_b_(__LINE__, __VA_ARGS__)

// Gram of a tall, narrow batch (n <= SPLIT_N columns, contiguous X of rows x n per matrix): the
// rows are split into S slices whose partial Grams are one batched GEMM (batch B S, enough
// workgroups to fill the chip), then summed in a fixed order.
constexpr int SPLIT_N = 64;
constexpr int SPLIT_S = 64;
hipError_t gram_cols_split_b_(int line, const float2* X, int rows, int n, double2* parts, double2* G, int batch,
                              hipStream_t s) {
  int S = SPLIT_S;
  while (S > 1 && (rows % S != 0 || rows / S < 256)) S /= 2;
  if (S == 1) return gram_cols_b_(line, X, rows, n, (long long)rows * n, G, batch, s);
  const long long nn = (long long)n * n;
  GemmArgs g{};
  g.M = n; g.N = n; g.K = rows / S; g.batch = batch * S;
  g.A = X; g.rsA = 1; g.csA = n; g.bsA = (long long)(rows / S) * n; g.conjA = 1;
  g.B = X; g.rsB = n; g.csB = 1; g.bsB = (long long)(rows / S) * n;
  g.C = parts; g.rsC = n; g.csC = 1; g.bsC = nn;
  g.alpha_re = 1;
  g.herm = 1;
  hipError_t e = k_gemm(C64_F64ACC, g, s, line);
  if (e != hipSuccess) return e;
  if (!fits(parts, (long long)batch * S * nn, sizeof(double2), "split Gram partials", line) ||
      !fits(G, (long long)batch * nn, sizeof(double2), "split Gram output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return sum_parts(parts, S, nn, G, batch, s);
}
#define gram_cols_split_b(...) gram_cols_split_b_(__LINE__, __VA_ARGS__)

// The Gram matrices whose consumers read only the leading eigenvectors and the eigenvalue sums
// (the gate-time cut, the chi steps and the kappa cuts) use the pivoted-Cholesky-preconditioned
// solver (eigh_gram); MPDO_LAYER_GRAM_EIG=0 selects the plain Jacobi solver for them.
bool use_gram_eig() {
  static const bool on = [] {
    const char* v = getenv("MPDO_LAYER_GRAM_EIG");
    return !(v && v[0] == '0');
  }();
  return on;
}

struct EigB {
  double2* V;
  double* w;
  void* ws;
};
// workspace of the eigensolvers at order n: the full one, the bottom-of-the-spectrum cut and the
// top-k solver for k up to kmax
size_t eig_ws_bytes(int n, int batch, int kmax) {
  size_t b = eigh_gram_workspace_bytes(n, batch);
  if (n >= PARTIAL_MIN && 2 * g_tail_m <= n) b = std::max(b, gram_tail_workspace_bytes(n, g_tail_m, batch));
  const int k = std::min(kmax, (n - 64) / 2);
  if (top_m(n, k)) b = std::max(b, gram_top_workspace_bytes(n, top_m(n, k), batch));
  return b;
}
void carve_eig(Carver& c, int n, int batch, int kmax, EigB& e) {
  e.V = c.get<double2>((size_t)n * n * batch);
  e.w = c.get<double>((size_t)n * batch);
  e.ws = c.get<char>(eig_ws_bytes(n, batch, kmax));
}
mpdo_status eig_b_(int line, int n, int batch, double2* G, EigB& e, hipStream_t s, bool gram = false) {
  int sweeps = 0;
  const long long nn = (long long)n * n;
  if (!fits(G, batch * nn, sizeof(double2), "eigensolver input", line) ||
      !fits(e.V, batch * nn, sizeof(double2), "eigenvectors", line) ||
      !fits(e.w, (long long)batch * n, sizeof(double), "eigenvalues", line) ||
      !fits(e.ws, (long long)eigh_gram_workspace_bytes(n, batch), 1, "eigensolver workspace", line)) {
    return layer_fail(MPDO_EINVAL, t_oob.c_str());
  }
  if (t_dry) return MPDO_OK;
  if (gram && use_gram_eig())   // pivots down to 1e-12 of the largest diagonal; the rest stays as tail mass
    LCK(eigh_gram(n, batch, G, n, (long long)n * n, e.w, n, e.V, n, (long long)n * n, e.ws, s, &sweeps, 1e-12));
  else
    LCK(eigh(n, batch, G, n, (long long)n * n, e.w, n, e.V, n, (long long)n * n, e.ws, s, &sweeps));
  if (sweeps > 60) {
    char b[160];
    snprintf(b, sizeof b, "Jacobi eigensolver did not converge (layer.hip:%d, order %d, batch %d)", line, n, batch);
    return layer_fail(MPDO_ENOCONV, b);
  }
  return MPDO_OK;
}
#define eig_b(...) eig_b_(__LINE__, __VA_ARGS__)

// Cholesky QR with column pivoting, twice (the second pass restores orthonormality lost to
// the conditioning of the first): the Gram matrix of the input, its pivoted Cholesky factor and
// the inverse of that factor's pivot-ordered triangle (pchol_b, pchol_inverse_b), then one GEMM.
// Directions whose pivot falls below FLOOR of the largest diagonal get zero columns of Q and zero
// rows of R, as the zero-filled null directions of the reference's QR.
struct QRWS {
  double2 *G, *L1, *L2, *F;
  int *piv1, *piv2, *rank1, *rank2;
  double* cond2;
  void *pws, *iws;
  float2* A1;
};
void carve_qr(Carver& c, int rows, int cols, int batch, QRWS& w) {
  const size_t n = (size_t)cols;
  w.G = c.get<double2>(n * n * batch);
  w.L1 = c.get<double2>(n * n * batch);
  w.L2 = c.get<double2>(n * n * batch);
  w.piv1 = c.get<int>(n * batch);
  w.piv2 = c.get<int>(n * batch);
  w.rank1 = c.get<int>((size_t)batch);
  w.rank2 = c.get<int>((size_t)batch);
  w.cond2 = c.get<double>((size_t)batch);
  w.pws = c.get<char>(pchol_workspace_bytes(cols, batch));
  w.iws = c.get<char>(pchol_inverse_workspace_bytes(cols, batch));
  w.F = c.get<double2>(n * n * batch);
  w.A1 = c.get<float2>((size_t)rows * cols * batch);
}
hipError_t k_pchol(int n, int batch, double2* G, double tol, double2* L, int* piv, int* rank, QRWS& w, hipStream_t s,
                   int line) {
  const long long nn = (long long)n * n;
  if (!fits(G, batch * nn, sizeof(double2), "pchol input", line) ||
      !fits(L, batch * nn, sizeof(double2), "pchol factor", line) ||
      !fits(piv, (long long)batch * n, sizeof(int), "pchol pivots", line) ||
      !fits(rank, batch, sizeof(int), "pchol ranks", line) ||
      !fits(w.pws, (long long)pchol_workspace_bytes(n, batch), 1, "pchol workspace", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return pchol_b(n, batch, G, n, nn, tol, L, piv, rank, w.pws, s);
}
hipError_t k_pinv(int n, int batch, const double2* L, const int* piv, const int* rank, int mode, double2* X, QRWS& w,
                  hipStream_t s, int line) {
  const long long nn = (long long)n * n;
  if (!fits(L, batch * nn, sizeof(double2), "pivoted inverse factor", line) ||
      !fits(piv, (long long)batch * n, sizeof(int), "pivoted inverse pivots", line) ||
      !fits(rank, batch, sizeof(int), "pivoted inverse ranks", line) ||
      !fits(X, batch * nn, sizeof(double2), "pivoted inverse output", line) ||
      !fits(w.iws, (long long)pchol_inverse_workspace_bytes(n, batch), 1, "pivoted inverse workspace", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return pchol_inverse_b(n, batch, L, piv, rank, mode, X, w.iws, s);
}
#define K_PCHOL(n, batch, G, tol, L, piv, rank, w, s) k_pchol(n, batch, G, tol, L, piv, rank, w, s, __LINE__)

// One pass of Cholesky QR leaves the columns orthonormal to about 1e-16 times the squared
// condition number of the input; below QR1_COND2 that is under the complex64 rounding of the
// stored factor and the second pass is skipped. Reads the estimates back (one sync).
constexpr double QR1_COND2 = 1e8;
mpdo_status one_pass_enough(int n, int batch, QRWS& w, bool& ok, hipStream_t s, int line) {
  ok = false;
  if (!fits(w.cond2, batch, sizeof(double), "QR condition estimates", line)) return layer_fail(MPDO_EINVAL, t_oob.c_str());
  if (t_dry) return MPDO_OK;
  LCK(pchol_cond2(n, batch, w.L1, w.piv1, w.rank1, w.cond2, s));
  std::vector<double> c(batch);
  LCK(hipMemcpyAsync(c.data(), w.cond2, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
  LCK(hipStreamSynchronize(s));
  ok = true;
  for (double v : c) ok = ok && v < QR1_COND2;
  return MPDO_OK;
}
#define K_PINV(n, batch, L, piv, rank, mode, X, w, s) k_pinv(n, batch, L, piv, rank, mode, X, w, s, __LINE__)

hipError_t k_gather(const float2* M, const int* piv, const int* rank, int n, long long other, int cols, float2* out,
                    int batch, hipStream_t s, int line) {
  if (!fits(M, batch * n * other, sizeof(float2), "pivot gather input", line) ||
      !fits(piv, (long long)batch * n, sizeof(int), "pivot gather pivots", line) ||
      !fits(rank, batch, sizeof(int), "pivot gather ranks", line) ||
      !fits(out, batch * n * other, sizeof(float2), "pivot gather output", line))
    return hipErrorInvalidValue;
  if (t_dry) return hipSuccess;
  return gather_piv(M, piv, rank, n, other, cols, out, batch, s);
}

// Y (rows x cols, complex64) = G (rows x cols, complex64 data gathered to pivot order) times the
// upper-triangular complex128 factor U (cols x cols)
hipError_t tri_right_(int line, int rows, int cols, const float2* Gm, const double2* U, float2* Y, int batch,
                      hipStream_t s) {
  GemmArgs g{};
  g.M = rows; g.N = cols; g.K = cols; g.batch = batch;
  g.A = Gm; g.rsA = cols; g.csA = 1; g.bsA = (long long)rows * cols; g.inA = 1;
  g.B = U; g.rsB = cols; g.csB = 1; g.bsB = (long long)cols * cols; g.inB = 2;
  g.C = Y; g.rsC = cols; g.csC = 1; g.bsC = (long long)rows * cols;
  g.alpha_re = 1;
  g.tri = 1;
  return k_gemm(C128_IN_C64_OUT, g, s, line);
}
// Yt (rows x cols, complex64) = the lower-triangular complex128 factor Lw (rows x rows) times G
// (rows x cols, complex64 data gathered to pivot order)
hipError_t tri_left_(int line, int rows, int cols, const double2* Lw, const float2* Gm, float2* Yt, int batch,
                     hipStream_t s) {
  GemmArgs g{};
  g.M = rows; g.N = cols; g.K = rows; g.batch = batch;
  g.A = Lw; g.rsA = rows; g.csA = 1; g.bsA = (long long)rows * rows; g.inA = 2;
  g.B = Gm; g.rsB = cols; g.csB = 1; g.bsB = (long long)rows * cols; g.inB = 1;
  g.C = Yt; g.rsC = cols; g.csC = 1; g.bsC = (long long)rows * cols;
  g.alpha_re = 1;
  g.tri = 2;
  return k_gemm(C128_IN_C64_OUT, g, s, line);
}

// Q (rows x cols) with orthonormal columns, Rf (cols x cols) complex64 with M = Q Rf; M is tall.
// M = Q1 L1^H (Q1 = M X1), Q1 = Q L2^H (Q = Q1 X2), so Rf = L2^H L1^H. With M's columns gathered
// to pivot order the inverse factor X is upper triangular (half the product's work); the
// products with it (entries growing with the condition number) take it in complex128.
mpdo_status qr_b(const float2* M, int rows, int cols, float2* Q, float2* Rf, int batch, QRWS& w, hipStream_t s) {
  const long long cc = (long long)cols * cols;
  LCK(gram_cols_b(M, rows, cols, (long long)rows * cols, w.G, batch, s));
  LCK(K_PCHOL(cols, batch, w.G, FLOOR, w.L1, w.piv1, w.rank1, w, s));
  LCK(K_PINV(cols, batch, w.L1, w.piv1, w.rank1, 2, w.F, w, s));
  LCK(k_gather(M, w.piv1, w.rank1, cols, rows, 1, w.A1, batch, s, __LINE__));
  bool one = false;
  LST(one_pass_enough(cols, batch, w, one, s, __LINE__));
  if (one) {   // Q = M X1, Rf = L1^H
    LCK(tri_right_(__LINE__, rows, cols, w.A1, w.F, Q, batch, s));
    LCK(k_narrow_factor(w.L1, cols, 0, Rf, batch, s, __LINE__));
    return MPDO_OK;
  }
  LCK(tri_right_(__LINE__, rows, cols, w.A1, w.F, Q, batch, s));   // Q1, held in Q
  LCK(gram_cols_b(Q, rows, cols, (long long)rows * cols, w.G, batch, s));
  LCK(K_PCHOL(cols, batch, w.G, FLOOR, w.L2, w.piv2, w.rank2, w, s));
  LCK(K_PINV(cols, batch, w.L2, w.piv2, w.rank2, 2, w.F, w, s));
  LCK(k_gather(Q, w.piv2, w.rank2, cols, rows, 1, w.A1, batch, s, __LINE__));
  LCK(tri_right_(__LINE__, rows, cols, w.A1, w.F, Q, batch, s));
  // Rf[u][j] = sum_t conj(L2[t][u]) conj(L1[j][t]); L stored by columns: L[x][y] at L[y cols + x]
  GemmArgs g{};
  g.M = cols; g.N = cols; g.K = cols; g.batch = batch;
  g.A = w.L2; g.rsA = cols; g.csA = 1; g.bsA = cc; g.conjA = 1;
  g.B = w.L1; g.rsB = cols; g.csB = 1; g.bsB = cc; g.conjB = 1;
  g.C = Rf; g.rsC = cols; g.csC = 1; g.bsC = cc;
  g.alpha_re = 1;
  LCK(K_GEMM(C128_IN_C64_OUT, g, s));
  return MPDO_OK;
}
// Qt (rows x cols) with orthonormal rows, Lf (rows x rows) complex64 with X = Lf Qt; X is wide.
// X = L1 Qt1 (Qt1 = X1 X), Qt1 = L2 Qt (Qt = X2 Qt1), so Lf = L1 L2. With X's rows gathered to
// pivot order the inverse factor is lower triangular.
mpdo_status lq_b(const float2* X, int rows, int cols, float2* Qt, float2* Lf, int batch, QRWS& w, hipStream_t s) {
  const long long rr = (long long)rows * rows;
  LCK(gram_rows_b(X, rows, cols, (long long)rows * cols, w.G, batch, s));
  LCK(K_PCHOL(rows, batch, w.G, FLOOR, w.L1, w.piv1, w.rank1, w, s));
  LCK(K_PINV(rows, batch, w.L1, w.piv1, w.rank1, 3, w.F, w, s));
  LCK(k_gather(X, w.piv1, w.rank1, rows, cols, 0, w.A1, batch, s, __LINE__));
  bool one = false;
  LST(one_pass_enough(rows, batch, w, one, s, __LINE__));
  if (one) {   // Qt = X1 X, Lf = L1
    LCK(tri_left_(__LINE__, rows, cols, w.F, w.A1, Qt, batch, s));
    LCK(k_narrow_factor(w.L1, rows, 1, Lf, batch, s, __LINE__));
    return MPDO_OK;
  }
  LCK(tri_left_(__LINE__, rows, cols, w.F, w.A1, Qt, batch, s));   // Qt1, held in Qt
  LCK(gram_rows_b(Qt, rows, cols, (long long)rows * cols, w.G, batch, s));
  LCK(K_PCHOL(rows, batch, w.G, FLOOR, w.L2, w.piv2, w.rank2, w, s));
  LCK(K_PINV(rows, batch, w.L2, w.piv2, w.rank2, 3, w.F, w, s));
  LCK(k_gather(Qt, w.piv2, w.rank2, rows, cols, 0, w.A1, batch, s, __LINE__));
  LCK(tri_left_(__LINE__, rows, cols, w.F, w.A1, Qt, batch, s));
  // Lf[i][v] = sum_u L1[i][u] L2[u][v]; L stored by columns
  GemmArgs g{};
  g.M = rows; g.N = rows; g.K = rows; g.batch = batch;
  g.A = w.L1; g.rsA = 1; g.csA = rows; g.bsA = rr;
  g.B = w.L2; g.rsB = 1; g.csB = rows; g.bsB = rr;
  g.C = Lf; g.rsC = rows; g.csC = 1; g.bsC = rr;
  g.alpha_re = 1;
  LCK(K_GEMM(C128_IN_C64_OUT, g, s));
  return MPDO_OK;
}

struct Dims {
  int l, a, r;
};

}  // namespace

namespace mpdo {

// --------------------------------------------------------------------------------------------
// Plan: output dims and applicability (host only).
// --------------------------------------------------------------------------------------------
struct LayerPlan {
  int n, K1, K2, parity, chi;  // chi <= 0: none
  int kappa;                   // <= 0: none
  std::vector<int> role;       // 0 single, 1 pair min, 2 pair max
  std::vector<Dims> in, a1;    // input dims, dims after the 1q channels
  std::vector<int> kq;         // QR bond right of site i (singles and max sites)
  std::vector<Dims> chi_out, out;
  std::vector<int> step_len;   // len(s) of each chi step (bond i|i+1), index i
  std::string why;
  bool ok = false;
};

static int capk(int chi, long long len) { return (int)(chi > 0 ? std::min<long long>(chi, len) : len); }

LayerPlan plan_layer(int n, const mpdo_site* in, int K1, int parity, int K2, int chi, int kappa) {
  LayerPlan P;
  P.n = n; P.K1 = K1; P.K2 = K2; P.parity = parity; P.chi = chi; P.kappa = kappa;
  if (n < 2 || K1 < 1 || K2 < 1 || (parity != 0 && parity != 1)) { P.why = "bad arguments"; return P; }
  P.role.assign(n, 0);
  for (int i = parity; i + 1 < n; i += 2) { P.role[i] = 1; P.role[i + 1] = 2; }
  for (int i = 0; i < n; ++i) {
    P.in.push_back(Dims{in[i].l, in[i].a, in[i].r});
    if (i + 1 < n && in[i].r != in[i + 1].l) { P.why = "bond mismatch"; return P; }
  }
  if (in[0].l != 1 || in[n - 1].r != 1) { P.why = "open boundary bonds must be 1"; return P; }
  for (int i = 0; i < n; ++i) P.a1.push_back(Dims{in[i].l, in[i].a * K1, in[i].r});
  auto afin = [&](int i) { return (long long)P.a1[i].a * (P.role[i] == 2 ? K2 : 1); };
  // QR sweep bonds
  P.kq.assign(n, 0);
  long long kprev = 1;
  for (int i = 0; i < n; ++i) {
    if (P.role[i] == 1) continue;   // internal to the pair
    if (i == n - 1) break;
    if (P.role[i] == 0) {
      long long rows = kprev * 2 * P.a1[i].a;
      P.kq[i] = (int)std::min<long long>(rows, P.in[i].r);
    } else {
      // pair max: QR bond = min(D' * 2 a K2, r) = r when 2 a K2 >= r (D' >= 1)
      if (2 * afin(i) < P.in[i].r) { P.why = "pair max site too thin for the fused path"; return P; }
      P.kq[i] = P.in[i].r;
    }
    kprev = P.kq[i];
  }
  auto kq_left = [&](int i) -> long long { return i == 0 ? 1 : P.kq[i - 1]; };   // site i-1 single or max
  P.chi_out.assign(n, Dims{0, 0, 0});
  P.step_len.assign(n, 0);
  long long kright = 1;
  // chi sweep, right to left
  for (int j = n - 1; j >= 0;) {
    if (P.role[j] == 0) {
      P.chi_out[j].r = (int)kright;
      P.chi_out[j].a = (int)afin(j);
      if (j == 0) { P.chi_out[0].l = 1; break; }
      long long cols = 2 * afin(j) * kright;
      long long len;
      if (P.role[j - 1] == 2) {
        long long lower = 2 * afin(j - 1);
        if (chi > 0 ? lower < std::min<long long>(chi, cols) : lower < cols) { P.why = "bond rank depends on the pair split"; return P; }
        len = cols;
      } else {
        len = std::min<long long>(kq_left(j - 1) * 2 * afin(j - 1), cols);
      }
      int k = capk(chi, len);
      P.step_len[j - 1] = (int)len;
      P.chi_out[j].l = k;
      kright = k;
      j -= 1;
    } else {
      // pair (i = j-1, j)
      int i = j - 1;
      if (P.role[i] != 1) { P.why = "internal: pair layout"; return P; }
      long long len_p = std::min<long long>(kq_left(i) * 2 * afin(i), 2 * afin(j) * kright);
      int kp = capk(chi, len_p);
      P.step_len[i] = (int)len_p;
      P.chi_out[j] = Dims{kp, (int)afin(j), (int)kright};
      P.chi_out[i].a = (int)afin(i);
      P.chi_out[i].r = kp;
      if (i == 0) { P.chi_out[0].l = 1; break; }
      long long cols = 2 * afin(i) * kp;
      long long len;
      if (P.role[i - 1] == 2) {
        long long lower = 2 * afin(i - 1);
        if (chi > 0 ? lower < std::min<long long>(chi, cols) : lower < cols) { P.why = "bond rank depends on the pair split"; return P; }
        len = cols;
      } else {
        len = std::min<long long>(kq_left(i - 1) * 2 * afin(i - 1), cols);
      }
      int k = capk(chi, len);
      P.step_len[i - 1] = (int)len;
      P.chi_out[i].l = k;
      kright = k;
      j = i - 1;
    }
  }
  P.out = P.chi_out;
  for (int i = 0; i < n; ++i) {
    const Dims& d = P.chi_out[i];
    if (kappa > 0 && d.a > kappa) {
      long long len_s = std::min<long long>(2LL * d.l * d.r, d.a);
      P.out[i].a = (int)std::min<long long>(kappa, len_s);
    } else if (kappa <= 0) {
      P.out[i].a = d.a;   // no kappa cap and no error bound: svdKappa is a no-op
    }
  }
  P.ok = true;
  return P;
}

// --------------------------------------------------------------------------------------------
struct PairStore {
  int i;                        // min site
  int l, rows1t, p1t;           // T_min as [(l a1) x (s1 m)] = Q1t R1t, p1t = min(l a1, 2m)
  int kprev, rows1, cols1, p1;  // X1r = (R (x) I_a1) Q1t: [(kprev a1) x p1t] = Q1 Ra, p1 = min(rows1, p1t)
  int m, rows2, cols2, p2;      // T_max as [(m s2) x (a2 r)] = L2 Q2
  bool q1tid, q1id, q2id;
  float2 *Q1, *Ra, *Y, *Um, *Q2, *L2;   // Y: the theta core before the channel, [2][2][p1t][p2];
                                        // Um: kept left singular vectors of the theta core [nT x mq]
  float2 *Qz, *Rz;              // (Ra (x) I) Um = Qz Rz when the gate-time cut is active
  int mq, pz, nc;               // kept directions (max over chains), rows of Rz, rows of the core
  bool qzid;
  double2* WR;    // right inverse of R_max (r x r, complex128), kept for the chi sweep
  double* lamR;
  bool has_R, has_P;            // has_P: the gate-time cut drops a direction in some chain
};

mpdo_status run_layer(int batch, const LayerPlan& P, const mpdo_site* in, const float* kraus1, const float* gates2,
                      const mpdo_site* out, double* disc_host, hipStream_t s, bool dry) {
  struct DryScope {
    explicit DryScope(bool d) { t_dry = d; }
    ~DryScope() { t_dry = false; }
  } dry_scope(dry);
  if (dry) t_gemms.clear();
  const int n = P.n, B = batch, K1 = P.K1, K2 = P.K2;
  // ---------------- layout ----------------
  std::vector<PairStore> pairs;
  for (int i = 0; i < n; ++i)
    if (P.role[i] == 1) {
      PairStore ps{};
      ps.i = i;
      ps.m = P.in[i].r;
      ps.rows2 = 2 * ps.m;
      ps.cols2 = P.a1[i + 1].a * P.in[i + 1].r;
      ps.q2id = ps.cols2 <= ps.rows2;
      ps.p2 = std::min(ps.rows2, ps.cols2);
      ps.l = P.in[i].l;
      ps.rows1t = ps.l * P.a1[i].a;
      ps.q1tid = ps.rows1t <= 2 * ps.m;
      ps.p1t = std::min(ps.rows1t, 2 * ps.m);
      ps.kprev = (i == 0) ? 1 : P.kq[i - 1];
      ps.rows1 = ps.kprev * P.a1[i].a;
      ps.cols1 = ps.p1t;
      ps.q1id = ps.rows1 <= ps.cols1;
      ps.p1 = std::min(ps.rows1, ps.cols1);
      ps.has_R = (i + 1 < n - 1);
      pairs.push_back(ps);
    }
  // sizes (elements per chain) of the shared scratch buffers: each the maximum over its uses
  auto kapplies = [&](int i) { return P.out[i].a != P.chi_out[i].a; };
  size_t mxPg = 1, mxZm = 1, mxV = 1, mxHk = 1, mxVt = 1, mxRs = 1, mxNn = 1, mxCp = 1;
  size_t mxQR_rows = 1, mxQR_cols = 1, mxR = 1, mxY = 1, mxOm = 1, mxZ = 1, mxQ2P = 1, mxC = 1;
  size_t mxTmin = 1, mxWm = 1, mxTm = 1, mxFL = 1, mxPn = 1, mxRT = 1, mxX1 = 1, mxLast = 1, mxTcur = 1;
  size_t mxTchi = 1, mxMp = 1, mxKout = 1, mxX1t = 1, mxQ1t = 1, mxR1t = 1;
  int mxEigQR = 1, mxEigC = 1, mxEigChi = 1, mxEigK = 1, mxEigR = 1;
  for (int i = 0; i < n; ++i) {
    const Dims& c = P.chi_out[i];
    const size_t szc = (size_t)c.l * 2 * c.a * c.r;
    mxEigChi = std::max(mxEigChi, c.l);
    mxPn = std::max(mxPn, (size_t)std::max(c.l, P.in[i].r) * std::max(c.l, c.r));
    if (P.role[i] != 2) mxTchi = std::max(mxTchi, szc);
    if (kapplies(i)) {
      mxEigK = std::max(mxEigK, c.a);
      mxMp = std::max(mxMp, szc);
      mxKout = std::max(mxKout, (size_t)2 * c.l * c.r * P.out[i].a);
    }
  }
  long long kprev = 1;
  for (int i = 0; i < n; ++i) {
    if (P.role[i] == 0) {
      const size_t m = (size_t)(kprev * 2 * P.a1[i].a * P.in[i].r);
      if (i == n - 1) {
        mxLast = std::max(mxLast, m);
      } else {
        long long rows = kprev * 2 * P.a1[i].a, cols = P.in[i].r;
        mxRT = std::max(mxRT, m);
        if (rows > cols) { mxQR_rows = std::max(mxQR_rows, (size_t)rows); mxQR_cols = std::max(mxQR_cols, (size_t)cols); }
        mxEigQR = std::max(mxEigQR, (int)cols);
        mxTcur = std::max(mxTcur, (size_t)(kprev * 2 * P.chi_out[i].a) * P.chi_out[i].r);   // A_i Pn
        kprev = P.kq[i];
      }
    } else if (P.role[i] == 2 && i < n - 1) {
      kprev = P.kq[i];
    }
    mxR = std::max(mxR, (size_t)(kprev * std::max(P.in[i].l, P.in[i].r)));
  }
  for (auto& ps : pairs) {
    const int kp = P.chi_out[ps.i + 1].l, kr = P.chi_out[ps.i + 1].r, a2 = P.a1[ps.i + 1].a;
    mxX1t = std::max(mxX1t, (size_t)ps.rows1t * 2 * ps.m);
    mxQ1t = std::max(mxQ1t, (size_t)ps.rows1t * ps.p1t);
    mxR1t = std::max(mxR1t, (size_t)ps.p1t * 2 * ps.m);
    mxX1 = std::max(mxX1, (size_t)ps.rows1 * ps.p1t);
    if (!ps.q1tid) { mxQR_rows = std::max(mxQR_rows, (size_t)ps.rows1t); mxQR_cols = std::max(mxQR_cols, (size_t)2 * ps.m); }
    if (!ps.q1id) { mxQR_rows = std::max(mxQR_rows, (size_t)ps.rows1); mxQR_cols = std::max(mxQR_cols, (size_t)ps.cols1); }
    if (!ps.q2id) { mxQR_rows = std::max(mxQR_rows, (size_t)ps.cols2); mxQR_cols = std::max(mxQR_cols, (size_t)ps.rows2); }
    mxEigQR = std::max(mxEigQR, std::max({2 * ps.m, ps.cols1, ps.rows2}));
    mxEigC = std::max(mxEigC, 2 * ps.p1t);
    mxPg = std::max(mxPg, (size_t)16 * ps.p1t * ps.p1t);
    // QR of (Ra (x) I) Um, [2 p1 x mq] with mq <= 2 p1t, factored only when 2 p1 > mq
    mxQR_rows = std::max(mxQR_rows, (size_t)2 * ps.p1);
    mxQR_cols = std::max(mxQR_cols, (size_t)std::min(2 * ps.p1, 2 * ps.p1t));
    mxY = std::max(mxY, (size_t)4 * ps.p1t * ps.p2);
    mxRs = std::max(mxRs, (size_t)2 * ps.p1 * 2 * ps.p1t);
    mxNn = std::max(mxNn, (size_t)4 * ps.p1t * ps.p1t);
    mxCp = std::max(mxCp, (size_t)2 * ps.p1 * 2 * ps.p1t);
    mxOm = std::max(mxOm, (size_t)ps.p2 * ps.p2);
    mxZ = std::max(mxZ, (size_t)ps.p2 * ps.cols2);
    mxEigR = std::max(mxEigR, P.in[ps.i + 1].r);
    const int c3 = a2 * kr;
    mxQ2P = std::max(mxQ2P, (size_t)ps.p2 * c3);

    mxEigC = std::max(mxEigC, 2 * ps.p1);
    mxFL = std::max(mxFL, (size_t)2 * ps.p1 * kp);
    mxTmin = std::max(mxTmin, (size_t)ps.rows1 * 2 * kp);
    mxTcur = std::max(mxTcur, (size_t)ps.kprev * 2 * P.a1[ps.i].a * kp);
    if (!(kapplies(ps.i + 1) && 2 * K2 > 8)) {   // the max site formed from the explicit core
      mxC = std::max(mxC, (size_t)2 * ps.p1t * 2 * K2 * ps.p2);
      mxCp = std::max(mxCp, mxC);
      mxWm = std::max(mxWm, (size_t)kp * 2 * K2 * ps.p2);
      mxTm = std::max(mxTm, (size_t)kp * 2 * K2 * c3);
    } else {                                     // Kraus-factored kappa cut of the max site
      mxWm = std::max(mxWm, (size_t)kp * 2 * ps.p1t);
      mxTm = std::max(mxTm, (size_t)8 * kp * c3);
      mxZm = std::max(mxZm, (size_t)kp * 2 * ps.p1t);
      mxV = std::max(mxV, (size_t)kp * 8 * ps.p2);
      mxHk = std::max(mxHk, (size_t)64 * a2 * a2);
      mxVt = std::max(mxVt, (size_t)16 * a2 * P.out[ps.i + 1].a);
    }
    mxPn = std::max(mxPn, (size_t)P.in[ps.i + 1].r * std::max(kr, P.in[ps.i + 1].r));
  }
  const int mxEig = std::max({mxEigQR, mxEigC, mxEigChi, mxEigK, mxEigR});
  int mxKp = 0;   // the largest kept count of a pair-bond chi step
  for (auto& ps : pairs) mxKp = std::max(mxKp, P.chi_out[ps.i + 1].l);
  size_t mxBond = 1;
  for (int i = 0; i < n; ++i) {
    mxBond = std::max(mxBond, (size_t)std::max(P.in[i].l, P.in[i].r));
    mxBond = std::max(mxBond, (size_t)std::max(P.chi_out[i].l, P.chi_out[i].r));
  }
  const size_t mxBond2 = mxBond * mxBond;
  mxFL = std::max(mxFL, mxBond2);
  mxPn = std::max(mxPn, mxBond2);

  // device buffers
  std::vector<float2*> T1(n, nullptr), Astore(n, nullptr);
  Site1q* dtab = nullptr;
  float2 *Mlast = nullptr, *FLq, *Rcur, *Rnext, *RT, *X1, *X1t, *Q1t, *R1t, *Cs, *Cp, *Cu, *Zb, *Q2P, *FL, *FR, *Tmin, *Wm,
         *Tm, *Pn, *Pn2, *PW, *Pp, *Tchi, *Mp, *Vk, *Kout, *dG, *Tcur2, *dK;
  double2 *G, *GR, *Om64, *Z64, *H64, *Gparts, *Pg, *Hk, *Tw, *Rs64, *Zr, *Nn, *Sparts;
  double2* LR;          // pivoted Cholesky factor of the right-bond Gram (R_max = L^H)
  int *pivR, *rankR;
  char *pwsR, *iwsR;
  float2 *Zm, *Vf, *Vt;
  double* ddisc;
  int* dk;
  EigB eE{};
  QRWS qw{};
  const int nsteps = n;
  auto layout = [&](Carver& c) {
    for (int i = 0; i < n; ++i) T1[i] = c.get<float2>((size_t)P.a1[i].l * 2 * P.a1[i].a * P.a1[i].r * B);
    long long kp = 1;
    for (int i = 0; i < n - 1; ++i) {
      if (P.role[i] == 0) {
        long long rows = kp * 2 * P.a1[i].a;
        Astore[i] = c.get<float2>((size_t)(rows * P.kq[i]) * B);
        kp = P.kq[i];
      } else if (P.role[i] == 2) {
        kp = P.kq[i];
      }
    }
    for (auto& ps : pairs) {
      ps.Q1 = ps.q1id ? nullptr : c.get<float2>((size_t)ps.rows1 * ps.p1 * B);
      ps.Ra = c.get<float2>((size_t)ps.p1 * ps.p1t * B);
      ps.Y = c.get<float2>((size_t)4 * ps.p1t * ps.p2 * B);
      ps.Um = c.get<float2>((size_t)4 * ps.p1t * ps.p1t * B);
      ps.Qz = c.get<float2>((size_t)4 * ps.p1 * ps.p1 * B);
      ps.Rz = c.get<float2>((size_t)2 * ps.p1 * 2 * ps.p1t * B);
      ps.Q2 = ps.q2id ? nullptr : c.get<float2>((size_t)ps.p2 * ps.cols2 * B);
      ps.L2 = ps.q2id ? nullptr : c.get<float2>((size_t)ps.rows2 * ps.p2 * B);
      const int r = P.in[ps.i + 1].r;
      ps.WR = ps.has_R ? c.get<double2>((size_t)r * r * B) : nullptr;
      ps.lamR = ps.has_R ? c.get<double>((size_t)r * B) : nullptr;
    }
    Mlast = c.get<float2>(mxLast * B);
    Rcur = c.get<float2>(std::max(mxR, mxBond2) * B);
    Rnext = c.get<float2>(std::max(mxR, mxBond2) * B);
    RT = c.get<float2>(mxRT * B);
    X1 = c.get<float2>(mxX1 * B);
    X1t = c.get<float2>(mxX1t * B);
    Q1t = c.get<float2>(mxQ1t * B);
    R1t = c.get<float2>(mxR1t * B);
    Cs = c.get<float2>(mxC * B);
    Cp = c.get<float2>(mxCp * B);
    Cu = c.get<float2>(mxC * B);
    Tw = c.get<double2>(mxY * B);
    Rs64 = c.get<double2>(mxRs * B);
    Zr = c.get<double2>(mxRs * B);
    Nn = c.get<double2>(mxNn * B);
    Om64 = c.get<double2>(mxOm * B);
    Z64 = c.get<double2>(mxZ * B);
    Zb = c.get<float2>(std::max(mxZ, mxBond2) * B);
    dtab = c.get<Site1q>((size_t)n * B);
    Q2P = c.get<float2>(mxQ2P * B);
    H64 = c.get<double2>(mxOm * B);
    FL = c.get<float2>(mxFL * B);
    FLq = c.get<float2>(mxFL * B);
    FR = c.get<float2>(mxFL * B);
    Tmin = c.get<float2>(mxTmin * B);
    Tcur2 = c.get<float2>(mxTcur * B);
    Wm = c.get<float2>(mxWm * B);
    Tm = c.get<float2>(mxTm * B);
    Pn = c.get<float2>(mxPn * B);
    Pn2 = c.get<float2>(mxPn * B);
    PW = c.get<float2>(mxBond2 * B);
    Pp = c.get<float2>(mxPn * B);
    Tchi = c.get<float2>(mxTchi * B);
    Mp = c.get<float2>(mxMp * B);
    Vk = c.get<float2>((size_t)mxEigK * mxEigK * B);
    Kout = c.get<float2>(mxKout * B);
    dG = c.get<float2>((size_t)pairs.size() * 16 * K2 + 16);
    dK = c.get<float2>((size_t)n * B * 4 * K1);
    G = c.get<double2>((size_t)mxEig * mxEig * B);
    Gparts = c.get<double2>((size_t)SPLIT_S * SPLIT_N * SPLIT_N * B);
    Sparts = c.get<double2>((size_t)SPLITK_PARTS * B);
    Pg = c.get<double2>(mxPg * B);
    Hk = c.get<double2>(mxHk * B);
    Zm = c.get<float2>(mxZm * B);
    Vf = c.get<float2>(mxV * B);
    Vt = c.get<float2>(mxVt * B);
    GR = c.get<double2>((size_t)mxEig * mxEig * B);
    LR = c.get<double2>((size_t)mxEigR * mxEigR * B);
    pivR = c.get<int>((size_t)mxEigR * B);
    rankR = c.get<int>((size_t)B);
    pwsR = c.get<char>(pchol_workspace_bytes(mxEigR, B));
    iwsR = c.get<char>(pchol_inverse_workspace_bytes(mxEigR, B));
    carve_eig(c, mxEig, B, mxKp, eE);
    carve_qr(c, (int)mxQR_rows, (int)mxQR_cols, B, qw);
    dk = c.get<int>((size_t)B * 4);
    ddisc = c.get<double>((size_t)B * (nsteps + 2 * n));
  };
  Carver c0;
  layout(c0);
  t_ws_bytes = c0.off;
  char* base;
  if (dry) base = (char*)(uintptr_t(1) << 44);   // never dereferenced: nothing is launched
  else LST(layer_ws(c0.off, s, &base));
  std::vector<Span> spans;
  struct SpanScope {
    explicit SpanScope(std::vector<Span>* p) { t_spans = p; t_oob.clear(); }
    ~SpanScope() { t_spans = nullptr; }
  } span_scope(&spans);
  for (int i = 0; i < n; ++i) {
    add_span(in[i].data, sizeof(float2) * (size_t)P.in[i].l * 2 * P.in[i].a * P.in[i].r * B);
    add_span(out[i].data, sizeof(float2) * (size_t)P.out[i].l * 2 * P.out[i].a * P.out[i].r * B);
  }
  Carver c1;
  c1.base = base;
  layout(c1);
  struct SplitScope {
    SplitScope(double2* p, long long cap) { t_split_parts = p; t_split_cap = cap; }
    ~SplitScope() { t_split_parts = nullptr; t_split_cap = 0; }
  } split_scope(Sparts, (long long)SPLITK_PARTS * B);
#define HFITS(p, cnt, T, what) \
  if (!fits((p), (long long)(cnt), sizeof(T), (what), __LINE__)) return layer_fail(MPDO_EINVAL, t_oob.c_str())
  HFITS(ddisc, (size_t)B * (nsteps + 2 * n), double, "discarded weights");
  if (!dry) LCK(hipMemsetAsync(ddisc, 0, sizeof(double) * B * (nsteps + 2 * n), s));

  // ---------------- phase 0: single-qubit channels ----------------
  {
    std::vector<Site1q> tab((size_t)n * B);
    for (int b = 0; b < B; ++b)
      for (int i = 0; i < n; ++i) {
        const Dims& d = P.in[i];
        const long long sz_in = (long long)d.l * 2 * d.a * d.r, sz_out = sz_in * K1;
        tab[(size_t)b * n + i] = Site1q{(const float2*)in[i].data + b * sz_in, T1[i] + b * sz_out, d.l, d.a, d.r, K1,
                                        ((long long)b * n + i) * 4 * K1};
      }
    for (const Site1q& e : tab) {
      HFITS(e.in, (long long)e.l * 2 * e.a * e.r, float2, "single-qubit channel input");
      HFITS(e.out, (long long)e.l * 2 * e.a * e.K * e.r, float2, "single-qubit channel output");
      HFITS(dK + e.koff, 4LL * e.K, float2, "single-qubit Kraus operators");
    }
    HFITS(dtab, tab.size(), Site1q, "site table");
    HFITS(dK, (size_t)n * B * 4 * K1, float2, "Kraus buffer");
    if (!dry) {
      LCK(hipMemcpyAsync(dtab, tab.data(), sizeof(Site1q) * tab.size(), hipMemcpyHostToDevice, s));
      LCK(hipMemcpyAsync(dK, kraus1, sizeof(float2) * (size_t)n * B * 4 * K1, hipMemcpyHostToDevice, s));
    }
    long long maxe = 0;
    for (int i = 0; i < n; ++i) maxe = std::max(maxe, (long long)P.in[i].l * P.in[i].a * P.in[i].r);
    if (!dry) {
      LCK(apply_1q(n * B, dtab, dK, maxe, s));
      LCK(hipStreamSynchronize(s));   // tab is pageable host memory
    }
  }
  HFITS(dG, pairs.size() * 16 * K2, float2, "two-qubit channel buffer");
  if (!dry) LCK(hipMemcpyAsync(dG, gates2, sizeof(float2) * pairs.size() * 16 * K2, hipMemcpyHostToDevice, s));

  // ---------------- phase 1: LQ of every pair's max site ----------------
  for (auto& ps : pairs) {
    if (ps.q2id) continue;
    // X2 = T1[max] as [(m s2) x (a2 r)] with rows2 < cols2: X2 = L2 Q2
    LST(lq_b(T1[ps.i + 1], ps.rows2, ps.cols2, ps.Q2, ps.L2, B, qw, s));
  }

  // The pair core in a row basis of Q1's (p s1) space: (Ra (x) I_s) P G(Y) = Qc Rc. Without an
  // active gate-time cut P = I, Qc = I and Rc = (Ra (x) I_s) G(Y) has nc = 2 p1 rows; with it
  // P = Um Um^H and Rc = Rz (Um^H G(Y)) has nc = pz <= mq rows (Qc = Qz). Rc goes to Cu.
  auto make_core = [&](PairStore& ps, int gidx) -> mpdo_status {
    const int p1t = ps.p1t, nT = 2 * p1t;
    const long long cols = 2LL * K2 * ps.p2, szCt = (long long)nT * cols;
    LCK(k_channel2q(ps.Y, dG + (size_t)gidx * 16 * K2, K2, p1t, ps.p2, Cs, s, B, 4LL * p1t * ps.p2, szCt,
                    __LINE__));
    if (!ps.has_P) {
      // Cs viewed as [p1t x (s1 cols)]
      LCK(mmb(g_tensor_prec, ps.p1, (int)(2 * cols), p1t, ps.Ra, p1t, (long long)ps.p1 * p1t, Cs, 2 * cols, szCt,
              Cu, 2 * cols, 2LL * ps.p1 * cols, B, s));
      ps.nc = 2 * ps.p1;
      return MPDO_OK;
    }
    const int mq = ps.mq;
    LCK(mmb(g_tensor_prec, mq, (int)cols, nT, ps.Um, mq, (long long)nT * mq, Cs, cols, szCt, Cp, cols,
            (long long)mq * cols, B, s, true));
    LCK(mmb(g_tensor_prec, ps.pz, (int)cols, mq, ps.Rz, mq, (long long)ps.pz * mq, Cp, cols, (long long)mq * cols,
            Cu, cols, (long long)ps.pz * cols, B, s));
    ps.nc = ps.pz;
    return MPDO_OK;
  };
  // Rs: the core's rows as a map from the channel output's rows (p, o1) (nT = 2 p1t), so that
  // Rc = Rs G(Y): Rz Um^H with the gate-time cut, Ra (x) I_2 without it; complex128 in Rs64
  // (nc x nT). Sets nc.
  auto core_rows = [&](PairStore& ps) -> mpdo_status {
    const int nT = 2 * ps.p1t;
    if (ps.has_P) {
      GemmArgs g{};
      g.M = ps.pz; g.N = nT; g.K = ps.mq; g.batch = B;
      g.A = ps.Rz; g.rsA = ps.mq; g.csA = 1; g.bsA = (long long)ps.pz * ps.mq;
      g.B = ps.Um; g.rsB = 1; g.csB = ps.mq; g.bsB = (long long)nT * ps.mq; g.conjB = 1;
      g.C = Rs64; g.rsC = nT; g.csC = 1; g.bsC = (long long)ps.pz * nT;
      g.alpha_re = 1;
      LCK(K_GEMM(C64_F64ACC, g, s));
      ps.nc = ps.pz;
    } else {
      LCK(k_kron_i2(ps.Ra, ps.p1, ps.p1t, Rs64, B, s, __LINE__));
      ps.nc = 2 * ps.p1;
    }
    return MPDO_OK;
  };

  t_kept.clear();
  t_gate_skips = 0;
  t_gate_tails = 0;
  t_chi_tops = 0;
  t_gate_eigs.clear();
  t_chi_eigs.clear();
  // ---------------- phase 2: QR sweep left to right ----------------
  // Rcur: [B][kp][l] with kp the QR bond so far (1 at the left end)
  int kcur = 1;
  {
    std::vector<float2> one(B, make_float2(1.f, 0.f));
    HFITS(Rcur, B, float2, "left factor");
    if (!dry) {
      LCK(hipMemcpyAsync(Rcur, one.data(), sizeof(float2) * B, hipMemcpyHostToDevice, s));
      LCK(hipStreamSynchronize(s));
    }
  }
  size_t pi = 0;
  for (int i = 0; i < n; ++i) {
    if (P.role[i] == 0) {
      const Dims& d = P.a1[i];
      const long long colsT = 2LL * d.a * d.r, rows = (long long)kcur * 2 * d.a;
      float2* M = (i == n - 1) ? Mlast : RT;
      LCK(mmb(g_tensor_prec, kcur, (int)colsT, d.l, Rcur, d.l, (long long)kcur * d.l, T1[i], colsT,
              (long long)d.l * colsT, M, colsT, kcur * colsT, B, s));
      if (i == n - 1) break;
      const int cols = d.r, kq = P.kq[i];
      if (rows <= cols) {
        LCK(K_EYE(Astore[i], (int)rows, (int)rows, rows * rows, B, s));
        // R = M (rows x cols)
        LCK(K_COPY(Rnext, M, sizeof(float2) * rows * cols * B, s));
      } else {
        LST(qr_b(M, (int)rows, cols, Astore[i], Rnext, B, qw, s));
      }
      std::swap(Rcur, Rnext);
      kcur = kq;
      continue;
    }
    if (P.role[i] == 2) continue;
    // pair (i, i+1)
    PairStore& ps = pairs[pi++];
    const Dims& dmin = P.a1[i];
    const int m = ps.m, a1 = dmin.a, a2 = P.a1[i + 1].a, r = P.in[i + 1].r;
    const long long colsT = 2LL * a1 * m;
    // (1) the reference's gate-time split, in compressed form. T_min as [(l a1) x (s1 m)] = Q1t R1t
    //     and T_max = L2 Q2 give theta = (Q1t (x) I) C_theta (I (x) Q2) with C_theta = G(R1t L2):
    //     its Gram gives the singular values of theta and the kept count m' (absolute tail).
    const int l = ps.l, p1t = ps.p1t, m2 = 2 * m;
    {
      int dims[4] = {l, 2, a1, m}, perm[4] = {0, 2, 1, 3};
      LCK(K_PERMUTE(T1[i], X1t, 4, dims, perm, s, B, (long long)l * colsT));
    }
    const float2* R1tp;
    long long bsR1t;
    if (ps.q1tid) {   // Q1t = identity
      LCK(K_EYE(Q1t, ps.rows1t, ps.rows1t, (long long)ps.rows1t * ps.rows1t, B, s));
      R1tp = X1t;
      bsR1t = (long long)ps.rows1t * m2;
    } else {
      LST(qr_b(X1t, ps.rows1t, m2, Q1t, R1t, B, qw, s));
      R1tp = R1t;
      bsR1t = (long long)m2 * m2;
    }
    const float2* L2p = ps.q2id ? T1[i + 1] : ps.L2;
    const long long bsL2 = ps.q2id ? (long long)ps.rows2 * ps.cols2 : (long long)ps.rows2 * ps.p2;
    // Y[s1][s2] = R1t[:, s1 m:(s1+1) m] L2[(m s2), :]
    for (int s1 = 0; s1 < 2; ++s1)
      for (int s2 = 0; s2 < 2; ++s2) {
        GemmArgs g{};
        g.M = p1t; g.N = ps.p2; g.K = m; g.batch = B;
        g.A = R1tp + (long long)s1 * m; g.rsA = m2; g.csA = 1; g.bsA = bsR1t;
        g.B = L2p + (long long)s2 * ps.p2; g.rsB = 2LL * ps.p2; g.csB = 1; g.bsB = bsL2;
        g.C = ps.Y + (long long)(s1 * 2 + s2) * p1t * ps.p2; g.rsC = ps.p2; g.csC = 1;
        g.bsC = 4LL * p1t * ps.p2;
        g.alpha_re = 1;
        LCK(K_GEMM(g_tensor_prec, g, s));
      }
    {
      const int nT = 2 * p1t;
      const int len_gate = (int)std::min(2LL * l * a1, 2LL * a2 * K2 * r);
      // Gram of the core through its Kraus structure: the 4 p1t x 4 p1t Gram of Y's blocks and the
      // 8 x 8 Gram of the channel (16x fewer flops than C C^H over the 2 K2 p2 columns)
      LCK(gram_rows_b(ps.Y, 4 * p1t, ps.p2, 4LL * p1t * ps.p2, Pg, B, s));
      LCK(k_kraus_gram(Pg, dG + (pi - 1) * 16 * K2, K2, p1t, G, B, s, __LINE__));
      // the cut keeps every direction in every chain (then no eigenvector is needed: P = I, nothing
      // is discarded) exactly when G - D* I is positive definite, D* the largest eigenvalue the
      // absolute-error rule may drop: one Cholesky factorisation of the shifted matrix instead of the eigensolver
      bool keeps_all = false;
      if (!dry && !t_rec_gate && use_gram_eig() && len_gate >= nT && g_gate_skip)
        LCK(gram_keeps_all(nT, B, G, nT, (long long)nT * nT, GATE_EPS, eE.ws, s, &keeps_all));
      // otherwise the few directions the rule drops come from the bottom of the spectrum (subspace
      // iteration, Rayleigh-Ritz and a certificate; P = I - V_d V_d^H over all nT columns), or,
      // where that declines, from the full eigensolver
      TailResult tail;
      if (!keeps_all && !dry && !t_rec_gate && use_gram_eig() && len_gate >= nT && g_gate_tail && nT >= PARTIAL_MIN &&
          2 * g_tail_m <= nT) {
        if (!fits(eE.ws, (long long)gram_tail_workspace_bytes(nT, g_tail_m, B), 1, "gate-time cut workspace", __LINE__))
          return layer_fail(MPDO_EINVAL, t_oob.c_str());
        if (!fits(ps.Um, (long long)nT * nT * B, sizeof(float2), "gate-time projector", __LINE__))
          return layer_fail(MPDO_EINVAL, t_oob.c_str());
        LCK(gram_tail_cut(nT, B, G, GATE_EPS, g_tail_m, g_tail_q, eE.ws, s, dk, ddisc + (size_t)B * (nsteps + n + i), ps.Um,
                          nT, (long long)nT * nT, &tail));
      }
      if (keeps_all) {
        ps.has_P = false;
        ps.mq = nT;
        t_kept.push_back(nT);
        ++t_gate_skips;
      } else if (tail.ok) {
        ps.has_P = tail.any_drop;
        ps.mq = nT;
        t_kept.push_back(tail.any_drop ? tail.max_kept : nT);
        ++t_gate_tails;
      } else {
        LST(eig_b(nT, B, G, eE, s, true));
        if (t_rec_gate && !dry) LST(record_eigs(t_gate_eigs, eE.w, nT, s));
        LCK(K_KEEP(eE.w, nT, nT, len_gate, -1, GATE_EPS, 0, dk, ddisc + (size_t)B * (nsteps + n + i), B, s));
        // the projector onto the kept directions is applied only when some chain drops one; the
        // core then lives in the mq-dimensional span of the kept directions (mq = max over chains)
        ps.has_P = dry;
        ps.mq = nT;
        if (!dry) {
          std::vector<int> kept(B);
          LCK(hipMemcpyAsync(kept.data(), dk, sizeof(int) * B, hipMemcpyDeviceToHost, s));
          LCK(hipStreamSynchronize(s));
          ps.mq = 1;
          for (int b = 0; b < B; ++b) {
            ps.has_P = ps.has_P || kept[b] < nT;
            ps.mq = std::max(ps.mq, std::min(kept[b], nT));
          }
        }
        if (!dry) t_kept.push_back(ps.has_P ? ps.mq : nT);
        if (ps.has_P)
          LCK(k_mask(eE.V, nT, (long long)nT * nT, nT, ps.mq, dk, ps.Um, ps.mq, (long long)nT * ps.mq, B, s,
                     __LINE__));
      }
    }
    // (2) QR sweep through the pair: X1r = (R (x) I_a1) Q1t = Q1 Ra, [(kcur a1) x p1t]
    LCK(mmb(g_tensor_prec, kcur, a1 * p1t, l, Rcur, l, (long long)kcur * l, Q1t, a1 * p1t,
            (long long)ps.rows1t * p1t, X1, a1 * p1t, (long long)kcur * a1 * p1t, B, s));
    if (ps.q1id) {
      LCK(K_COPY(ps.Ra, X1, sizeof(float2) * (size_t)ps.rows1 * p1t * B, s));
    } else {
      LST(qr_b(X1, ps.rows1, p1t, ps.Q1, ps.Ra, B, qw, s));
    }
    if (ps.has_P) {
      // Za = (Ra (x) I_s) Um: Um viewed as [p1t x (s1 mq)] -> [p1 x (s1 mq)] = [(p s1) x mq] = Qz Rz
      const int mq = ps.mq, nz = 2 * ps.p1;
      LCK(mmb(g_tensor_prec, ps.p1, 2 * mq, p1t, ps.Ra, p1t, (long long)ps.p1 * p1t, ps.Um, 2 * mq,
              (long long)2 * p1t * mq, Cp, 2 * mq, (long long)nz * mq, B, s));
      ps.qzid = nz <= mq;
      ps.pz = std::min(nz, mq);
      if (ps.qzid) {
        LCK(K_COPY(ps.Rz, Cp, sizeof(float2) * (size_t)nz * mq * B, s));
      } else {
        LST(qr_b(Cp, nz, mq, ps.Qz, ps.Rz, B, qw, s));
      }
    }
    if (!ps.has_R) continue;
    // (3) the core of R theta is (Q1 (x) I) Qc Rc with Rc = Rs G(Y) = [nc x (s2 K2 p2)]; its
    //     right-bond Gram, all in complex128 so that it is Hermitian to fp64 rounding, is
    //     GR = sum_{q a2} conj(Q2[q a2, r]) (Om Q2)[q a2, r'] with Om = Rc_v^H Rc_v
    //     (Rc_v = [(nc s2 K2) x p2]) = sum_{s2 K2} C_{s2 K2}^H (Rs^H Rs) C_{s2 K2} = Ycat^H Nt Ycat,
    //     Nt the Kraus dual of Rs^H Rs: the core itself is never formed
    LST(core_rows(ps));
    {
      const int nT = 2 * p1t, n4 = 4 * p1t;
      const long long bsY = (long long)n4 * ps.p2;
      GemmArgs g{};
      g.M = nT; g.N = nT; g.K = ps.nc; g.batch = B;
      g.A = Rs64; g.rsA = 1; g.csA = nT; g.bsA = (long long)ps.nc * nT; g.conjA = 1;
      g.B = Rs64; g.rsB = nT; g.csB = 1; g.bsB = (long long)ps.nc * nT;
      g.C = Nn; g.rsC = nT; g.csC = 1; g.bsC = (long long)nT * nT;
      g.alpha_re = 1;
      g.herm = 1;
      LCK(K_GEMM(C128_F64ACC, g, s));
      LCK(k_kraus_dual(Nn, dG + (pi - 1) * 16 * K2, K2, p1t, Pg, B, s, __LINE__));
      LCK(mmx(C128_F64ACC, n4, ps.p2, n4, Pg, n4, (long long)n4 * n4, 2, ps.Y, ps.p2, bsY, 1, Tw, ps.p2, bsY, B, s));
      LCK(mmx(C128_F64ACC, ps.p2, ps.p2, n4, ps.Y, ps.p2, bsY, 1, Tw, ps.p2, bsY, 2, Om64, ps.p2,
              (long long)ps.p2 * ps.p2, B, s, true, true));
    }
    if (ps.q2id) {
      LCK(k_bds64(Om64, a2, r, GR, (long long)ps.p2 * ps.p2, (long long)r * r, B, s, __LINE__));
    } else {
      const long long qz = (long long)ps.p2 * ps.cols2;
      LCK(mmx(C128_F64ACC, ps.p2, ps.cols2, ps.p2, Om64, ps.p2, (long long)ps.p2 * ps.p2, 2, ps.Q2, ps.cols2, qz, 1,
              Z64, ps.cols2, qz, B, s));
      GemmArgs g{};
      const int rowsv = ps.p2 * a2;
      g.M = r; g.N = r; g.K = rowsv; g.batch = B;
      g.A = ps.Q2; g.rsA = 1; g.csA = r; g.bsA = qz; g.conjA = 1; g.inA = 1;
      g.B = Z64; g.rsB = r; g.csB = 1; g.bsB = qz;
      g.C = GR; g.rsC = r; g.csC = 1; g.bsC = (long long)r * r;
      g.alpha_re = 1;
      g.herm = 1;
      LCK(herm_gemm_(__LINE__, C128_F64ACC, g, s));
    }
    // R_max with R_max^H R_max = GR: the pivoted Cholesky factor GR = L L^H (pivots down to FLOOR
    // of the largest diagonal; rows of R_max past the rank zero), R_max = L^H; its right inverse
    // P Lt^{-H} (Lt the pivot-ordered triangle), kept for the chi sweep, maps the kept directions
    // back to the bond. Any factor of GR gives the layer the same result (the QR sweep's bond
    // basis is free); this one needs no eigensolver.
    if (!fits(GR, (long long)r * r * B, sizeof(double2), "right-bond Gram", __LINE__) ||
        !fits(LR, (long long)r * r * B, sizeof(double2), "right-bond factor", __LINE__) ||
        !fits(ps.WR, (long long)r * r * B, sizeof(double2), "right-bond inverse", __LINE__) ||
        !fits(pivR, (long long)r * B, sizeof(int), "right-bond pivots", __LINE__) ||
        !fits(pwsR, (long long)pchol_workspace_bytes(r, B), 1, "right-bond Cholesky workspace", __LINE__) ||
        !fits(iwsR, (long long)pchol_inverse_workspace_bytes(r, B), 1, "right-bond inverse workspace", __LINE__))
      return layer_fail(MPDO_EINVAL, t_oob.c_str());
    if (!dry) {
      LCK(pchol_b(r, B, GR, r, (long long)r * r, FLOOR, LR, pivR, rankR, pwsR, s, false));
      LCK(pchol_inverse_b(r, B, LR, pivR, rankR, 0, ps.WR, iwsR, s));
    }
    // R_max[k][i] = conj(L_ik): L is stored by columns (L[k r + i] = L_ik), so elementwise
    LCK(k_narrow(LR, Rcur, (long long)r * r, 1, B, s, __LINE__));
    kcur = r;
  }

  // ---------------- phase 3: chi sweep right to left, kappa as sites become final ----------
  // kappa cut of site i given as Mq = [l][2][r][a] (inner index last): out[i] = (Mq V_k) permuted
  auto kappa_on = [&](int i, const float2* Mq) -> mpdo_status {
    const Dims& c = P.chi_out[i];
    const long long sz = (long long)c.l * 2 * c.a * c.r;
    const int rows = 2 * c.l * c.r, a = c.a, k = P.out[i].a;
    if (a <= SPLIT_N) LCK(gram_cols_split_b(Mq, rows, a, Gparts, G, B, s));
    else LCK(gram_cols_b(Mq, rows, a, sz, G, B, s));
    LST(eig_b(a, B, G, eE, s, true));
    LCK(K_KEEP(eE.w, a, a, std::min(rows, a), k, -1.0, 1, dk, ddisc + (size_t)B * (nsteps + i), B, s));
    LCK(K_FACTOR(eE.V, a, (long long)a * a, eE.w, a, a, a, k, 0.0, 0.0, 0, Vk, k, (long long)a * k, B, s));
    LCK(mmb(g_tensor_prec, rows, k, a, Mq, a, sz, Vk, k, (long long)a * k, Kout, k, (long long)rows * k, B, s));
    int d2[4] = {c.l, 2, c.r, k}, p2[4] = {0, 1, 3, 2};
    LCK(K_PERMUTE(Kout, (float2*)out[i].data, 4, d2, p2, s, B, (long long)rows * k));
    return MPDO_OK;
  };
  // Kappa cut of a pair's new max site without forming it. The site is
  //   T = S^{-1/2} U^H Rc (I (x) Q2P),  Rc = Rs C,  C[(p, o1), (s2, K2, q)] = sum_i G[(o1 s2), i, K2] Y[i][p][q]
  // with Rs the map from C's rows to the core's (Rz Um^H, or Ra (x) I without the gate-time cut).
  // With Zm = S^{-1/2} U^H Rs (kp x 2 p1t), V[k, (o1 i), q] = sum_p Zm[k, (p o1)] Y[i][p][q] and
  // W = V Q2P [(k, o1 i) x (a2 kr)], T[k, s2, K2, a2, kr] = sum_{o1 i} G[(o1 s2), i, K2] W[k, (o1 i), a2, kr]:
  // the inner index (a2, K2) of T is the image of (o1 i, a2) (8 a2 directions instead of
  // 2 K2 a2), so T's Gram over the inner index comes from the 8 a2 x 8 a2 Gram of W and the
  // channel (kappa_core_gram), and the kept directions act on W directly (kappa_core_vec).
  // FR (kp x nC) is S^{-1/2} U^H of the chi step.
  auto kappa_factored = [&](PairStore& ps, int j, int gidx, int kp, int kr, int a2, const float2* Qright) -> mpdo_status {
    const int p1t = ps.p1t, nT = 2 * p1t, nC = ps.nc, k = P.out[j].a, a = a2 * K2;
    const long long bsFR = (long long)kp * nC, bsZm = (long long)kp * nT;
    const float2* Gt = dG + (size_t)gidx * 16 * K2;
    // Zm = FR Rs
    if (ps.has_P) {
      // FR Rz (kp x mq), then (FR Rz) Um^H
      LCK(mmb(g_tensor_prec, kp, ps.mq, ps.pz, FR, nC, bsFR, ps.Rz, ps.mq, (long long)ps.pz * ps.mq, Wm, ps.mq,
              (long long)kp * ps.mq, B, s));
      GemmArgs g{};
      g.M = kp; g.N = nT; g.K = ps.mq; g.batch = B;
      g.A = Wm; g.rsA = ps.mq; g.csA = 1; g.bsA = (long long)kp * ps.mq;
      g.B = ps.Um; g.rsB = 1; g.csB = ps.mq; g.bsB = (long long)nT * ps.mq; g.conjB = 1;
      g.C = Zm; g.rsC = nT; g.csC = 1; g.bsC = bsZm;
      g.alpha_re = 1;
      LCK(K_GEMM(g_tensor_prec, g, s));
    } else {
      // Zm[k, (p', o1)] = sum_p FR[k, (p, o1)] Ra[p, p']
      for (int o1 = 0; o1 < 2; ++o1) {
        GemmArgs g{};
        g.M = kp; g.N = p1t; g.K = ps.p1; g.batch = B;
        g.A = FR + o1; g.rsA = nC; g.csA = 2; g.bsA = bsFR;
        g.B = ps.Ra; g.rsB = p1t; g.csB = 1; g.bsB = (long long)ps.p1 * p1t;
        g.C = Zm + o1; g.rsC = nT; g.csC = 2; g.bsC = bsZm;
        g.alpha_re = 1;
        LCK(K_GEMM(g_tensor_prec, g, s));
      }
    }
    // V[k, (o1 i), q] = sum_p Zm[k, (p o1)] Y[i][p][q]
    const long long bsV = (long long)kp * 8 * ps.p2;
    for (int o1 = 0; o1 < 2; ++o1)
      for (int i = 0; i < 4; ++i) {
        GemmArgs g{};
        g.M = kp; g.N = ps.p2; g.K = p1t; g.batch = B;
        g.A = Zm + o1; g.rsA = nT; g.csA = 2; g.bsA = bsZm;
        g.B = ps.Y + (long long)i * p1t * ps.p2; g.rsB = ps.p2; g.csB = 1; g.bsB = 4LL * p1t * ps.p2;
        g.C = Vf + (long long)(o1 * 4 + i) * ps.p2; g.rsC = 8LL * ps.p2; g.csC = 1; g.bsC = bsV;
        g.alpha_re = 1;
        LCK(K_GEMM(g_tensor_prec, g, s));
      }
    // W = V Q2P: [(k, o1 i) x (a2 kr)] in Tm, then rows (o1 i, a2) over columns (k, kr) in Mp
    const long long c3 = (long long)a2 * kr, bsW = 8LL * kp * c3;
    LCK(mmb(g_tensor_prec, 8 * kp, (int)c3, ps.p2, Vf, ps.p2, bsV, Qright, c3, (long long)ps.p2 * c3, Tm, c3, bsW, B,
            s));
    {
      int dims[4] = {kp, 8, a2, kr}, perm[4] = {1, 2, 0, 3};
      LCK(K_PERMUTE(Tm, Mp, 4, dims, perm, s, B, bsW));
    }
    LCK(gram_rows_b(Mp, 8 * a2, kp * kr, bsW, Hk, B, s));
    LCK(k_kappa_gram(Hk, Gt, K2, a2, G, B, s, __LINE__));
    LST(eig_b(a, B, G, eE, s, true));
    LCK(K_KEEP(eE.w, a, a, (int)std::min(2LL * kp * kr, (long long)a), k, -1.0, 1, dk, ddisc + (size_t)B * (nsteps + j), B,
               s));
    LCK(K_FACTOR(eE.V, a, (long long)a * a, eE.w, a, a, a, k, 0.0, 0.0, 0, Vk, k, (long long)a * k, B, s));
    LCK(k_kappa_vec(Vk, Gt, K2, a2, k, Vt, B, s, __LINE__));
    // out2[(s2 j), (k' kr)] = sum_{(o1 i a2)} Vt[(o1 i a2), (s2 j)] W[(o1 i a2), (k' kr)]
    {
      GemmArgs g{};
      g.M = 2 * k; g.N = kp * kr; g.K = 8 * a2; g.batch = B;
      g.A = Vt; g.rsA = 1; g.csA = 2 * k; g.bsA = 16LL * a2 * k;
      g.B = Mp; g.rsB = (long long)kp * kr; g.csB = 1; g.bsB = bsW;
      g.C = Kout; g.rsC = (long long)kp * kr; g.csC = 1; g.bsC = 2LL * k * kp * kr;
      g.alpha_re = 1;
      LCK(K_GEMM(g_tensor_prec, g, s));
    }
    int dims[4] = {2, k, kp, kr}, perm[4] = {2, 0, 1, 3};
    LCK(K_PERMUTE(Kout, (float2*)out[j].data, 4, dims, perm, s, B, 2LL * k * kp * kr));
    return MPDO_OK;
  };
  // where the chi step writes site i: straight into the output when no kappa cut follows
  auto chi_target = [&](int i) -> float2* { return kapplies(i) ? Tchi : (float2*)out[i].data; };
  // kappa cut of site i held as T = [l][2][a][r] in chi_target(i)
  auto kappa_site = [&](int i, const float2* T) -> mpdo_status {
    const Dims& c = P.chi_out[i];
    if (!kapplies(i)) {
      if (T != out[i].data) LCK(K_COPY(out[i].data, T, sizeof(float2) * (size_t)c.l * 2 * c.a * c.r * B, s));
      return MPDO_OK;
    }
    int dims[4] = {c.l, 2, c.a, c.r}, perm[4] = {0, 1, 3, 2};
    LCK(K_PERMUTE(T, Mp, 4, dims, perm, s, B, (long long)c.l * 2 * c.a * c.r));
    return kappa_on(i, Mp);
  };
  // canonical chi step on B (D x colsB): out T_right = S^{-1/2} U^H B (k x colsB), Pn = U sqrt(S) (D x k)
  auto chi_step = [&](const float2* Bt, int D, long long colsB, int k, int len, float2* Tright, float2* PnOut,
                      int disc_slot) -> mpdo_status {
    LCK(gram_rows_b(Bt, D, (int)colsB, D * colsB, G, B, s));
    LST(eig_b(D, B, G, eE, s, true));
    LCK(K_KEEP(eE.w, D, D, len, k, -1.0, 1, dk, ddisc + (size_t)B * disc_slot, B, s));
    LCK(K_FACTOR(eE.V, D, (long long)D * D, eE.w, D, D, D, k, -0.25, SPLIT_FLOOR, 1, FR, D, (long long)k * D, B, s));
    LCK(mmb(g_tensor_prec, k, (int)colsB, D, FR, D, (long long)k * D, Bt, colsB, D * colsB, Tright, colsB,
            k * colsB, B, s));
    LCK(K_FACTOR(eE.V, D, (long long)D * D, eE.w, D, D, D, k, 0.25, SPLIT_FLOOR, 0, PnOut, k, (long long)D * k, B, s));
    return MPDO_OK;
  };

  // Bcur: the current explicit tensor of the site right of the next step
  const float2* Bcur = nullptr;
  int pidx = (int)pairs.size() - 1;
  int kP = 1;   // columns of Pn
  bool havePn = false;
  for (int j = n - 1; j >= 0;) {
    if (P.role[j] == 0) {
      const Dims& c = P.chi_out[j];
      // current tensor of site j: [l_qr][2][a][kright]
      const int lq = (j == 0) ? 1 : ((P.role[j - 1] == 2 || P.role[j - 1] == 0) ? P.kq[j - 1] : 1);
      if (j == n - 1) {
        Bcur = Mlast;
      } else {
        // T_j^{cur} = A_j Pn : [(lq 2 a) x kq] [kq x kP]
        const long long rowsA = (long long)lq * 2 * c.a;
        LCK(mmb(g_tensor_prec, (int)rowsA, kP, P.kq[j], Astore[j], P.kq[j], rowsA * P.kq[j], Pn, kP,
                (long long)P.kq[j] * kP, Tcur2, kP, rowsA * kP, B, s));
        Bcur = Tcur2;
      }
      if (j == 0) {
        LST(kappa_site(0, Bcur));
        break;
      }
      const long long colsB = 2LL * c.a * c.r;
      LST(chi_step(Bcur, lq, colsB, c.l, P.step_len[j - 1], chi_target(j), Pn2, j - 1));
      LST(kappa_site(j, chi_target(j)));
      std::swap(Pn, Pn2);
      kP = c.l;
      havePn = true;
      j -= 1;
      continue;
    }
    // pair (i, j = i+1)
    const int i = j - 1;
    PairStore& ps = pairs[pidx--];
    const Dims& cmax = P.chi_out[j];
    const Dims& cmin = P.chi_out[i];
    const int kp = cmax.l, kr = cmax.r, a1 = P.a1[i].a, a2 = P.a1[j].a, r = P.in[j].r;
    const int c3 = a2 * kr;
    // right projector in the original bond: P = W_R L_R^{-1/2} Pn (r x kr), or the identity
    const float2* Qright;   // Q2 P : [p2][a2][kr] = [p2 x c3]
    if (ps.has_R) {
      LCK(k_narrow(ps.WR, PW, (long long)r * r, 0, B, s, __LINE__));   // R_max's right inverse
      LCK(mmb(g_tensor_prec, r, kr, r, PW, r, (long long)r * r, Pn, kP, (long long)r * kP, Pp, kr, (long long)r * kr, B,
              s));
      const float2* Q2p = ps.Q2;
      long long bsQ2 = (long long)ps.p2 * ps.cols2;
      if (ps.q2id) {   // Q2 = identity (p2 = cols2): one shared copy
        LCK(K_EYE(Zb, ps.p2, ps.cols2, 0, 1, s));
        Q2p = Zb;
        bsQ2 = 0;
      }
      LCK(mmb(g_tensor_prec, ps.p2 * a2, kr, r, Q2p, r, bsQ2, Pp, kr, (long long)r * kr, Q2P, kr,
              (long long)ps.p2 * a2 * kr, B, s));
      Qright = Q2P;
    } else {
      // last pair: r = kr = 1, Q2 P = Q2
      if (ps.q2id) {
        LCK(K_EYE(Q2P, ps.p2, c3, (long long)ps.p2 * c3, B, s));
      } else {
        LCK(K_COPY(Q2P, ps.Q2, sizeof(float2) * (size_t)ps.p2 * c3 * B, s));
      }
      Qright = Q2P;
    }
    // Theta = (Q1 (x) I) Qc Rc (I (x) Q2P): its core's Gram is Rc (I (x) H) Rc^H with
    // H = Q2P Q2P^H [p2 x p2], = Rs kraus_gram(Ycat H Ycat^H) Rs^H, all in complex128 (neither
    // Q2P nor the core is factored or formed)
    LCK(gram_rows_b(Qright, ps.p2, c3, (long long)ps.p2 * c3, H64, B, s));
    LST(core_rows(ps));
    const int nC = ps.nc;
    {
      const int p1t = ps.p1t, nT = 2 * p1t, n4 = 4 * p1t;
      const long long bsY = (long long)n4 * ps.p2, bsRs = (long long)nC * nT;
      LCK(mmx(C128_F64ACC, n4, ps.p2, ps.p2, ps.Y, ps.p2, bsY, 1, H64, ps.p2, (long long)ps.p2 * ps.p2, 2, Tw, ps.p2,
              bsY, B, s));
      GemmArgs g{};
      g.M = n4; g.N = n4; g.K = ps.p2; g.batch = B;
      g.A = Tw; g.rsA = ps.p2; g.csA = 1; g.bsA = bsY;
      g.B = ps.Y; g.rsB = 1; g.csB = ps.p2; g.bsB = bsY; g.conjB = 1; g.inB = 1;
      g.C = Pg; g.rsC = n4; g.csC = 1; g.bsC = (long long)n4 * n4;
      g.alpha_re = 1;
      g.herm = 1;
      LCK(K_GEMM(C128_F64ACC, g, s));
      LCK(k_kraus_gram(Pg, dG + (size_t)(pidx + 1) * 16 * K2, K2, p1t, Nn, B, s, __LINE__));
      LCK(mmb(C128_F64ACC, nC, nT, nT, Rs64, nT, bsRs, Nn, nT, (long long)nT * nT, Zr, nT, bsRs, B, s));
      GemmArgs h{};
      h.M = nC; h.N = nC; h.K = nT; h.batch = B;
      h.A = Zr; h.rsA = nT; h.csA = 1; h.bsA = bsRs;
      h.B = Rs64; h.rsB = 1; h.csB = nT; h.bsB = bsRs; h.conjB = 1;
      h.C = G; h.rsC = nC; h.csC = 1; h.bsC = (long long)nC * nC;
      h.alpha_re = 1;
      h.herm = 1;
      LCK(K_GEMM(C128_F64ACC, h, s));
    }
    // the step keeps kp of about 4 kp directions: its top kp eigenpairs by subspace iteration and
    // a certificate, or, where that declines, the full eigensolver
    TopResult top;
    const int mt = top_m(nC, kp);
    if (mt && !dry && !t_rec_gate && use_gram_eig() && g_chi_top && kp <= P.step_len[i]) {
      if (!fits(eE.ws, (long long)gram_top_workspace_bytes(nC, mt, B), 1, "chi step workspace", __LINE__))
        return layer_fail(MPDO_EINVAL, t_oob.c_str());
      LCK(gram_top_eig(nC, B, G, kp, mt, g_top_q, eE.ws, s, eE.w, nC, eE.V, nC, (long long)nC * nC, &top));
      if (top.ok) ++t_chi_tops;
    }
    if (!top.ok) LST(eig_b(nC, B, G, eE, s, true));
    if (t_rec_gate && !dry) LST(record_eigs(t_chi_eigs, eE.w, nC, s));
    LCK(K_KEEP(eE.w, nC, nC, P.step_len[i], kp, -1.0, 1, dk, ddisc + (size_t)B * i, B, s));
    // new min site: (Q1 (x) I) U sqrt(S) -> (kprev, a1, s1, kp) -> (kprev, s1, a1, kp)
    LCK(K_FACTOR(eE.V, nC, (long long)nC * nC, eE.w, nC, nC, nC, kp, 0.25, SPLIT_FLOOR, 0, FL, kp, (long long)nC * kp, B, s));
    const float2* Zs = FL;
    long long bsZ = (long long)nC * kp;
    if (ps.has_P && !ps.qzid) {   // back to Q1's (p s1) basis: Qz [2 p1 x pz] U sqrt(S)
      LCK(mmb(g_tensor_prec, 2 * ps.p1, kp, ps.pz, ps.Qz, ps.pz, (long long)2 * ps.p1 * ps.pz, FL, kp, bsZ, FLq, kp,
              (long long)2 * ps.p1 * kp, B, s));
      Zs = FLq;
      bsZ = (long long)2 * ps.p1 * kp;
    }
    if (!ps.q1id) {
      LCK(mmb(g_tensor_prec, ps.rows1, 2 * kp, ps.p1, ps.Q1, ps.p1, (long long)ps.rows1 * ps.p1, Zs, 2 * kp,
              bsZ, Tmin, 2 * kp, (long long)ps.rows1 * 2 * kp, B, s));
      Zs = Tmin;
      bsZ = (long long)ps.rows1 * 2 * kp;
    }
    {
      int dims[4] = {ps.kprev, a1, 2, kp}, perm[4] = {0, 2, 1, 3};
      LCK(K_PERMUTE(Zs, Tcur2, 4, dims, perm, s, B, bsZ));
    }
    // new max site: S^{-1/2} U^H Rc (I (x) Q2P) -> (kp, s2, K2, a2, kr) -> (kp, s2, a2, K2, kr)
    LCK(K_FACTOR(eE.V, nC, (long long)nC * nC, eE.w, nC, nC, nC, kp, -0.25, SPLIT_FLOOR, 1, FR, nC, (long long)kp * nC, B, s));
    if (kapplies(j) && 2 * K2 > 8) {
      LST(kappa_factored(ps, j, pidx + 1, kp, kr, a2, Qright));
    } else {
      LST(make_core(ps, pidx + 1));
      const long long colsR = 2LL * K2 * ps.p2, szC = (long long)nC * colsR;
      LCK(mmb(g_tensor_prec, kp, (int)colsR, nC, FR, nC, (long long)kp * nC, Cu, colsR, szC, Wm, colsR,
              kp * colsR, B, s));
      LCK(mmb(g_tensor_prec, kp * 2 * K2, c3, ps.p2, Wm, ps.p2, kp * colsR, Qright, c3, (long long)ps.p2 * c3, Tm,
              c3, (long long)kp * 2 * K2 * c3, B, s));
      const float2* Ts = Tm;
      const long long bsT = (long long)kp * 2 * K2 * c3;
      if (kapplies(j)) {   // (kp, s2, K2, a2, kr) -> (kp, s2, kr, a2, K2): inner index last for the kappa cut
        int dims[5] = {kp, 2, K2, a2, kr}, perm[5] = {0, 1, 4, 3, 2};
        LCK(K_PERMUTE(Ts, Mp, 5, dims, perm, s, B, bsT));
        LST(kappa_on(j, Mp));
      } else {             // (kp, s2, K2, a2, kr) -> (kp, s2, a2, K2, kr)
        int dims[5] = {kp, 2, K2, a2, kr}, perm[5] = {0, 1, 3, 2, 4};
        LCK(K_PERMUTE(Ts, (float2*)out[j].data, 5, dims, perm, s, B, bsT));
      }
    }
    // step at bond (i-1 | i) on the new min-site tensor
    if (i == 0) {
      LST(kappa_site(0, Tcur2));
      break;
    }
    const long long colsB = 2LL * a1 * kp;
    LST(chi_step(Tcur2, ps.kprev, colsB, cmin.l, P.step_len[i - 1], chi_target(i), Pn2, i - 1));
    LST(kappa_site(i, chi_target(i)));
    std::swap(Pn, Pn2);
    kP = cmin.l;
    havePn = true;
    j = i - 1;
  }
  (void)havePn;
  if (dry) return MPDO_OK;
  // discarded weights
  if (disc_host) {
    std::vector<double> h((size_t)B * (nsteps + 2 * n));
    HFITS(ddisc, h.size(), double, "discarded weights");
    LCK(hipMemcpyAsync(h.data(), ddisc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    LCK(hipStreamSynchronize(s));
    // device layout: slot-major [slot][B]; host layout [B][slot]: chi steps by bond, kappa cuts
    // by site, gate-time cuts by pair min site
    for (int b = 0; b < B; ++b)
      for (int t = 0; t < nsteps + 2 * n; ++t) disc_host[(size_t)b * (nsteps + 2 * n) + t] = h[(size_t)t * B + b];
  } else {
    LCK(hipStreamSynchronize(s));
  }
  return MPDO_OK;
}

}  // namespace mpdo

extern "C" {

int mpdo_layer_dims(int n, const mpdo_site* in, int K1, int parity, int K2, int chi, int kappa, mpdo_site* out) {
  if (!in || !out || n <= 0) return 0;
  LayerPlan P = plan_layer(n, in, K1, parity, K2, chi, kappa);
  if (!P.ok) {
    layer_fail(MPDO_EINVAL, P.why.c_str());
    return 0;
  }
  for (int i = 0; i < n; ++i) {
    out[i].l = P.out[i].l;
    out[i].a = P.out[i].a;
    out[i].r = P.out[i].r;
  }
  return 1;
}

mpdo_status mpdo_layer_batched(int batch, int n, const mpdo_site* in, const float* kraus1, int K1, int parity,
                               const float* gates2, int K2, int chi, int kappa, const mpdo_site* out,
                               double* discarded, void* stream) {
  if (batch <= 0 || n <= 0 || !in || !out || !kraus1 || !gates2) return layer_fail(MPDO_EINVAL, "null argument");
  for (int i = 0; i < n; ++i)
    if (!in[i].data || !out[i].data) return layer_fail(MPDO_EINVAL, "null site buffer");
  LayerPlan P = plan_layer(n, in, K1, parity, K2, chi, kappa);
  if (!P.ok) return layer_fail(MPDO_EINVAL, P.why.c_str());
  for (int i = 0; i < n; ++i)
    if (out[i].l != P.out[i].l || out[i].a != P.out[i].a || out[i].r != P.out[i].r)
      return layer_fail(MPDO_EINVAL, "output descriptors must carry mpdo_layer_dims dims");
  return run_layer(batch, P, in, kraus1, gates2, out, discarded, (hipStream_t)stream, false);
}

int mpdo_debug_layer_gemm_log(int* out, int cap) {
  const int n = (int)t_gemms.size() / 6;
  for (int i = 0; i < 6 * n && i < cap; ++i) out[i] = t_gemms[i];
  return n;
}

int mpdo_debug_layer_gate_eigs(int record, double* out, int cap) {
  t_rec_gate = record != 0;
  const int n = (int)t_gate_eigs.size();
  for (int i = 0; i < n && i < cap; ++i) out[i] = t_gate_eigs[i];
  return n;
}

int mpdo_debug_layer_chi_eigs(double* out, int cap) {
  const int n = (int)t_chi_eigs.size();
  for (int i = 0; i < n && i < cap; ++i) out[i] = t_chi_eigs[i];
  return n;
}

int mpdo_debug_layer_gate_skip(int on) {
  const int old = g_gate_skip ? 1 : 0;
  if (on >= 0) g_gate_skip = on != 0;
  return old;
}

int mpdo_debug_layer_gate_skips(void) { return t_gate_skips; }

int mpdo_debug_layer_gate_tail(int on) {
  const int old = g_gate_tail ? 1 : 0;
  if (on >= 0) g_gate_tail = on != 0;
  return old;
}

int mpdo_debug_layer_gate_tails(void) { return t_gate_tails; }

int mpdo_debug_layer_chi_top(int on) {
  const int old = g_chi_top ? 1 : 0;
  if (on >= 0) g_chi_top = on != 0;
  return old;
}

int mpdo_debug_layer_chi_tops(void) { return t_chi_tops; }

int mpdo_debug_layer_tail_shape(int m, int q) {
  const int old = g_tail_m;
  if (m >= 16 && m <= 256) g_tail_m = m;
  if (q >= 1) g_tail_q = q;
  return old;
}

int mpdo_debug_layer_top_shape(int extra, int q) {
  const int old = g_top_extra;
  if (extra >= 0) g_top_extra = extra;
  if (q >= 1) g_top_q = q;
  return old;
}

int mpdo_debug_layer_kept(int* out, int cap) {
  const int n = (int)t_kept.size();
  for (int i = 0; i < n && i < cap; ++i) out[i] = t_kept[i];
  return n;
}

size_t mpdo_layer_workspace_bytes(int batch, int n, const mpdo_site* in, int K1, int parity, int K2, int chi,
                                  int kappa) {
  if (batch <= 0 || n <= 0 || !in) {
    layer_fail(MPDO_EINVAL, "null argument");
    return 0;
  }
  LayerPlan P = plan_layer(n, in, K1, parity, K2, chi, kappa);
  if (!P.ok) {
    layer_fail(MPDO_EINVAL, P.why.c_str());
    return 0;
  }
  // disjoint stand-in addresses for the chains (the plan needs their extents, not their data)
  std::vector<mpdo_site> fin(n), fout(n);
  uintptr_t addr = uintptr_t(1) << 40;
  for (int i = 0; i < n; ++i) {
    fin[i] = mpdo_site{(void*)addr, P.in[i].l, P.in[i].a, P.in[i].r};
    addr += ((sizeof(float2) * (size_t)batch * P.in[i].l * 2 * P.in[i].a * P.in[i].r) | 4095) + 1;
    fout[i] = mpdo_site{(void*)addr, P.out[i].l, P.out[i].a, P.out[i].r};
    addr += ((sizeof(float2) * (size_t)batch * P.out[i].l * 2 * P.out[i].a * P.out[i].r) | 4095) + 1;
  }
  t_ws_bytes = 0;
  if (run_layer(batch, P, fin.data(), nullptr, nullptr, fout.data(), nullptr, nullptr, true) != MPDO_OK) return 0;
  return t_ws_bytes;
}

mpdo_status mpdo_debug_layer_dry_run(int batch, int n, const mpdo_site* in, int K1, int parity, int K2, int chi,
                                     int kappa, const mpdo_site* out, unsigned long long* workspace_bytes) {
  if (batch <= 0 || n <= 0 || !in || !out) return layer_fail(MPDO_EINVAL, "null argument");
  LayerPlan P = plan_layer(n, in, K1, parity, K2, chi, kappa);
  if (!P.ok) return layer_fail(MPDO_EINVAL, P.why.c_str());
  for (int i = 0; i < n; ++i)
    if (out[i].l != P.out[i].l || out[i].a != P.out[i].a || out[i].r != P.out[i].r)
      return layer_fail(MPDO_EINVAL, "output descriptors must carry mpdo_layer_dims dims");
  t_ws_bytes = 0;
  mpdo_status st = run_layer(batch, P, in, nullptr, nullptr, out, nullptr, nullptr, true);
  if (workspace_bytes) *workspace_bytes = t_ws_bytes;
  return st;
}

}  // extern "C"
