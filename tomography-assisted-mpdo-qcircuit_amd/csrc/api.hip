// C-ABI entry points of libmpdo_gfx950 (declared in include/mpdo.h): host-side drivers that
// sequence the gfx950 kernels of one reference operation each.
//
// Every SVD of the reference path is computed from an fp64 Gram matrix (complex64 data
// accumulated in fp64 on the matrix cores) and the fp64 Jacobi eigensolver; factors are then
// rebuilt in complex64 and applied with the fp32 MFMA GEMM. The decompositions.svd keep rule
// runs on the device; its count is read back where later launch shapes depend on it.
#include "mpdo.h"
#include "mpdo_debug.h"
#include "mpdo_internal.h"
#include "site_ops.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace mpdo;

namespace mpdo {
// Accumulation precision of the complex64 tensor contractions (mpdo_set_math_mode).
Prec g_tensor_prec = C64_F64ACC_C64OUT;
mpdo_status layer_fail(mpdo_status s, const char* msg);
mpdo_status layer_ws(size_t bytes, hipStream_t s, char** base);
}  // namespace mpdo

namespace {

thread_local std::string g_err;

mpdo_status fail(mpdo_status s, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define CK(x)                                                                                   \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) return fail(MPDO_EHIP, "%s:%d %s -> %s", __FILE__, __LINE__, #x,     \
                                      hipGetErrorString(e_));                                   \
  } while (0)
#define ST(x)                                  \
  do {                                         \
    mpdo_status s_ = (x);                      \
    if (s_ != MPDO_OK) return s_;              \
  } while (0)

// ---------------------------------------------------------------- workspace (per device and stream)
// Each call carves its scratch from one device buffer. A caller may bind its own buffer to a stream
// (mpdo_bind_workspace; the Python host binds one from torch's caching allocator): calls on that
// stream then use it and never allocate, and a call that needs more returns MPDO_ENOMEM before
// launching anything, with the size it needs in mpdo_workspace_needed(). Without a binding the
// library keeps a grow-only buffer of its own for the (device, stream) pair. Calls on different
// streams use different buffers, so they may run concurrently; calls on one stream are ordered.
struct DevWS {
  void* p = nullptr;
  size_t cap = 0;
  bool bound = false;   // the caller's buffer (never freed here)
};
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, DevWS> g_ws;
thread_local size_t t_ws_needed = 0;

mpdo_status ws_reserve(size_t bytes, hipStream_t s, char** base) {
  int dev = 0;
  CK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  DevWS& w = g_ws[{dev, s}];
  if (w.bound) {
    if (w.cap < bytes) {
      t_ws_needed = bytes;
      return fail(MPDO_ENOMEM, "the workspace bound to this stream holds %zu bytes, the call needs %zu", w.cap, bytes);
    }
    *base = (char*)w.p;
    return MPDO_OK;
  }
  if (w.cap < bytes) {
    if (w.p) {
      CK(hipStreamSynchronize(s));
      CK(hipFree(w.p));
      w.p = nullptr;
      w.cap = 0;
    }
    size_t nb = bytes + bytes / 8 + (1 << 20);
    if (hipMalloc(&w.p, nb) != hipSuccess) {
      (void)hipGetLastError();
      w.p = nullptr;
      t_ws_needed = bytes;
      return fail(MPDO_ENOMEM, "device workspace of %zu bytes", nb);
    }
    w.cap = nb;
  }
  *base = (char*)w.p;
  return MPDO_OK;
}

// Two-pass carver: a dry pass measures, the real pass hands out 256-byte aligned slices.
struct Carver {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* get(size_t n) {
    size_t b = (n * sizeof(T) + 255) & ~size_t(255);
    T* p = base ? (T*)(base + off) : nullptr;
    off += b;
    return p;
  }
};

inline hipStream_t S(void* s) { return (hipStream_t)s; }

// ---------------------------------------------------------------- linear-algebra helpers
// G (n x n, complex128) = X^H X with X (rows x n) row-major, leading dim ldx
hipError_t gram_cols(const float2* X, int rows, int n, long long ldx, double2* G, hipStream_t s) {
  GemmArgs g{};
  g.M = n; g.N = n; g.K = rows; g.batch = 1;
  g.A = X; g.rsA = 1; g.csA = ldx; g.conjA = 1;
  g.B = X; g.rsB = ldx; g.csB = 1; g.conjB = 0;
  g.C = G; g.rsC = n; g.csC = 1;
  g.alpha_re = 1;
  return gemm(C64_F64ACC, g, s);
}
// G (n x n, complex128) = X X^H with X (n x cols) row-major, leading dim ldx
hipError_t gram_rows(const float2* X, int n, int cols, long long ldx, double2* G, hipStream_t s) {
  GemmArgs g{};
  g.M = n; g.N = n; g.K = cols; g.batch = 1;
  g.A = X; g.rsA = ldx; g.csA = 1; g.conjA = 0;
  g.B = X; g.rsB = 1; g.csB = ldx; g.conjB = 1;
  g.C = G; g.rsC = n; g.csC = 1;
  g.alpha_re = 1;
  return gemm(C64_F64ACC, g, s);
}

// C (M x N) = A (M x K) * B (K x N), all row-major complex64 with the given leading dims
hipError_t mm(int M, int N, int K, const float2* A, long long lda, const float2* B, long long ldb, float2* C,
              long long ldc, hipStream_t s) {
  GemmArgs g = gemm_rowmajor(M, N, K, A, lda, false, false, B, ldb, false, false, C, ldc);
  return gemm(g_tensor_prec, g, s);
}
hipError_t mm64(Prec p, int M, int N, int K, const void* A, long long lda, bool tA, bool cA, const void* B,
                long long ldb, void* C, long long ldc, hipStream_t s) {
  GemmArgs g = gemm_rowmajor(M, N, K, A, lda, tA, cA, B, ldb, false, false, C, ldc);
  return gemm(p, g, s);
}

struct Eig {
  double2* V;
  double* w;
  void* ws;
};
void carve_eig(Carver& c, int n, Eig& e) {
  e.V = c.get<double2>((size_t)n * n);
  e.w = c.get<double>((size_t)n);
  e.ws = c.get<char>(eigh_workspace_bytes(n, 1));
}
// test hook (mpdo_debug_fail_eig): the next g_fail_eig eigensolver calls of the per-gate entry
// points run to completion and then report non-convergence, as a sweep cap would
std::atomic<int> g_fail_eig{0};

mpdo_status run_eig(int n, double2* G, Eig& e, hipStream_t s) {
  int sweeps = 0;
  CK(eigh(n, 1, G, n, 0, e.w, n, e.V, n, 0, e.ws, s, &sweeps));
  if (sweeps > 60) return fail(MPDO_ENOCONV, "Jacobi eigensolver did not converge (n=%d)", n);
  int pending = g_fail_eig.load();
  while (pending > 0 && !g_fail_eig.compare_exchange_weak(pending, pending - 1)) {
  }
  if (pending > 0) return fail(MPDO_ENOCONV, "Jacobi eigensolver did not converge (n=%d; forced by the test hook)", n);
  return MPDO_OK;
}

mpdo_status read_k(const int* dk, const double* dd, int* k, double* disc, hipStream_t s) {
  CK(hipMemcpyAsync(k, dk, sizeof(int), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(disc, dd, sizeof(double), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

constexpr double QR_FLOOR = 1e-13;   // relative eigenvalue floor of the QR-type factorisations
// Relative Gram-eigenvalue floor of the sqrt(S) splits: directions with s_j < 1e-6 s_0 are below
// what complex64 data resolves (a float32 SVD returns them as noise of that size); their rows
// S^{-1/2} U^H B would be rounding noise amplified by lambda^{-1/4}, so they are set to zero
// (the kept count is unchanged: those singular values are zero-filled).
constexpr double SPLIT_FLOOR = 1e-12;

// ---------------------------------------------------------------- truncating SVD step
// theta = A (rowsA x D) * B (D x colsB). Writes T_left = U sqrt(S) (rowsA x k) to outL and
// T_right = sqrt(S) V^H (k x colsB) to outR. `canonical`: A has orthonormal columns.
struct ChiWS {
  double2 *GB, *GA, *Phi, *T1, *P, *Q;
  float2 *FL, *FR;
  Eig eA, eZ;
  int* dk;
  double* dd;
};
void carve_chi(Carver& c, int D, int kmax, ChiWS& w) {
  size_t DD = (size_t)D * D;
  w.GB = c.get<double2>(DD);
  w.GA = c.get<double2>(DD);
  w.Phi = c.get<double2>(DD);
  w.T1 = c.get<double2>(DD);
  w.P = c.get<double2>(DD);
  w.Q = c.get<double2>((size_t)D * std::max(kmax, D));
  w.FL = c.get<float2>((size_t)D * kmax);
  w.FR = c.get<float2>((size_t)D * kmax);
  carve_eig(c, D, w.eA);
  carve_eig(c, D, w.eZ);
  w.dk = c.get<int>(4);
  w.dd = c.get<double>(4);
}

mpdo_status chi_step(const float2* A, int rowsA, int D, const float2* B, int colsB, bool canonical, int msv,
                     double err, int len_ref, float2* outL, float2* outR, int* k_out, double* disc, ChiWS& w,
                     hipStream_t s) {
  CK(gram_rows(B, D, colsB, colsB, w.GB, s));
  double2* Z = w.GB;
  if (!canonical) {
    CK(gram_cols(A, rowsA, D, D, w.GA, s));
    ST(run_eig(D, w.GA, w.eA, s));
    // Phi = W Lambda^{1/2};  Z = Phi^H GB Phi
    CK(factor64(w.eA.V, D, w.eA.w, D, D, D, 0.5, QR_FLOOR, 0, w.Phi, D, s));
    CK(mm64(C128_F64ACC, D, D, D, w.GB, D, false, false, w.Phi, D, w.T1, D, s));
    CK(mm64(C128_F64ACC, D, D, D, w.Phi, D, true, true, w.T1, D, w.GA, D, s));
    Z = w.GA;
  }
  ST(run_eig(D, Z, w.eZ, s));
  CK(keep_count(w.eZ.w, D, D, len_ref, msv, err, 1, w.dk, w.dd, 1, s));
  int k = 0;
  ST(read_k(w.dk, w.dd, &k, disc, s));
  if (canonical) {
    CK(factor(w.eZ.V, D, w.eZ.w, D, D, k, 0.25, SPLIT_FLOOR, 0, w.FL, k, s));
    CK(factor(w.eZ.V, D, w.eZ.w, D, D, k, -0.25, SPLIT_FLOOR, 1, w.FR, D, s));
  } else {
    // FL = (W Lambda^{-1/2}) (U' S^{1/2});  FR = (S^{-1/2} U'^H) (Lambda^{1/2} W^H)
    CK(factor64(w.eA.V, D, w.eA.w, D, D, D, -0.5, QR_FLOOR, 0, w.P, D, s));
    CK(factor64(w.eZ.V, D, w.eZ.w, D, D, k, 0.25, SPLIT_FLOOR, 0, w.Q, k, s));
    CK(mm64(C128_IN_C64_OUT, D, k, D, w.P, D, false, false, w.Q, k, w.FL, k, s));
    CK(factor64(w.eZ.V, D, w.eZ.w, D, D, k, -0.25, SPLIT_FLOOR, 1, w.Q, D, s));
    CK(factor64(w.eA.V, D, w.eA.w, D, D, D, 0.5, QR_FLOOR, 1, w.P, D, s));
    CK(mm64(C128_IN_C64_OUT, k, D, D, w.Q, D, false, false, w.P, D, w.FR, D, s));
  }
  CK(mm(rowsA, k, D, A, D, w.FL, k, outL, k, s));
  CK(mm(k, colsB, D, w.FR, D, B, colsB, outR, colsB, s));
  *k_out = k;
  return MPDO_OK;
}

bool bad_site(const mpdo_site& t) { return !t.data || t.l <= 0 || t.a <= 0 || t.r <= 0; }

}  // namespace

namespace mpdo {
mpdo_status layer_fail(mpdo_status s, const char* msg) { return fail(s, "%s", msg); }
mpdo_status layer_ws(size_t bytes, hipStream_t s, char** base) { return ws_reserve(bytes, s, base); }
}  // namespace mpdo

namespace {
// ---------------------------------------------------------------- Pauli strings
// Every string advances through the chain together: per site, X[t] = E[t]^T T (one strided
// batched GEMM over the strings, T shared), string t's operator on X[t]'s physical index, then
// E[t] <- X[t]^T conj(T) (the next environment, complex64 stored, fp64 accumulated): the
// arithmetic of dmOperations.py:174-201 for each string, with the launches shared.
struct ExpectLayout {
  float2 *E0, *E1, *X, *Xp;
  int* codes;
  double* res;
  long long bsE, bsX;
};

void expect_dims(int n, const mpdo_site* sites, size_t& maxX, int& maxr) {
  maxX = 1;
  maxr = 1;
  for (int i = 0; i < n; ++i) {
    maxX = std::max(maxX, (size_t)sites[i].l * 2 * sites[i].a * sites[i].r);
    maxr = std::max(maxr, std::max(sites[i].r, sites[i].l));
  }
}

ExpectLayout expect_layout(Carver& c, int n, const mpdo_site* sites, int chunk) {
  size_t maxX;
  int maxr;
  expect_dims(n, sites, maxX, maxr);
  ExpectLayout L;
  L.bsE = (long long)maxr * maxr;
  L.bsX = (long long)maxX;
  L.E0 = c.get<float2>((size_t)L.bsE * chunk);
  L.E1 = c.get<float2>((size_t)L.bsE * chunk);
  L.X = c.get<float2>((size_t)L.bsX * chunk);
  L.Xp = c.get<float2>((size_t)L.bsX * chunk);
  L.codes = c.get<int>((size_t)n * chunk);
  L.res = c.get<double>((size_t)chunk);
  return L;
}

}  // namespace

namespace mpdo {
size_t expect_chunk_bytes(int n, const mpdo_site* sites, int chunk) {
  Carver c;
  expect_layout(c, n, sites, chunk);
  return c.off;
}

// argument checks; *per = scratch bytes per string (the layout of a chunk of one)
mpdo_status expect_check(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out, size_t* per) {
  if (n <= 0 || nstrings <= 0) return MPDO_OK;
  if (!sites || !codes || !out) return fail(MPDO_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) {
    if (bad_site(sites[i])) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (i + 1 < n && sites[i].r != sites[i + 1].l) return fail(MPDO_EINVAL, "bond %d mismatch", i);
  }
  if (sites[0].l != 1 || sites[n - 1].r != 1) return fail(MPDO_EINVAL, "open boundary bonds must be 1");
  for (long long e = 0; e < (long long)n * nstrings; ++e)
    if (codes[e] < -1 || codes[e] > 4) return fail(MPDO_EINVAL, "operator code %d", codes[e]);
  *per = expect_chunk_bytes(n, sites, 1);
  return MPDO_OK;
}

// the strings in chunks of `chunk` through a scratch region laid out by expect_layout at `base`
mpdo_status expect_run(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out, char* base,
                       int chunk, hipStream_t s) {
  Carver c;
  c.base = base;
  const ExpectLayout L = expect_layout(c, n, sites, chunk);
  for (int t0 = 0; t0 < nstrings; t0 += chunk) {
    const int cnt = std::min(chunk, nstrings - t0);
    CK(hipMemcpyAsync(L.codes, codes + (size_t)t0 * n, sizeof(int) * (size_t)cnt * n, hipMemcpyHostToDevice, s));
    CK(env_one(L.E0, L.bsE, cnt, s));
    float2* E = L.E0;
    float2* En = L.E1;
    for (int i = 0; i < n; ++i) {
      const int l = sites[i].l, a = sites[i].a, r = sites[i].r;
      const long long cols = 2LL * a * r;
      // X[t][l', (s a r)] = sum_l E[t][l, l'] T[l, (s a r)]
      GemmArgs g{};
      g.M = l; g.N = (int)cols; g.K = l; g.batch = cnt;
      g.A = E; g.rsA = 1; g.csA = l; g.bsA = L.bsE;
      g.B = sites[i].data; g.rsB = cols; g.csB = 1; g.bsB = 0;
      g.C = L.X; g.rsC = cols; g.csC = 1; g.bsC = L.bsX;
      g.alpha_re = 1;
      CK(gemm(C64_F64ACC_C64OUT, g, s));
      CK(phys_op_codes(L.X, L.Xp, l, (long long)a * r, L.bsX, cnt, L.codes, n, i, s));
      // En[t][r, r'] = sum_{l' s a} Xp[t][l' s a, r] conj(T[l' s a, r'])
      GemmArgs h{};
      h.M = r; h.N = r; h.K = l * 2 * a; h.batch = cnt;
      h.A = L.Xp; h.rsA = 1; h.csA = r; h.bsA = L.bsX;
      h.B = sites[i].data; h.rsB = r; h.csB = 1; h.bsB = 0; h.conjB = 1;
      h.C = En; h.rsC = r; h.csC = 1; h.bsC = L.bsE;
      h.alpha_re = 1;
      CK(gemm(C64_F64ACC_C64OUT, h, s));
      std::swap(E, En);
    }
    CK(take_re(E, L.bsE, cnt, L.res, s));
    CK(hipMemcpyAsync(out + t0, L.res, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
  }
  return MPDO_OK;
}
}  // namespace mpdo

// ==========================================================================================
extern "C" {

int mpdo_version(void) { return 1; }

int mpdo_set_math_mode(int mode) {
  const int old = g_tensor_prec == C64_F32ACC      ? 1
                  : g_tensor_prec == C64_F32ACC_3M ? 2
                  : g_tensor_prec == C64_BF16X3    ? 3
                  : g_tensor_prec == C64_BF16X6    ? 4
                                                   : 0;
  if (mode == 0) g_tensor_prec = C64_F64ACC_C64OUT;
  else if (mode == 1) g_tensor_prec = C64_F32ACC;
  else if (mode == 2) g_tensor_prec = C64_F32ACC_3M;
  else if (mode == 3) g_tensor_prec = C64_BF16X3;
  else if (mode == 4) g_tensor_prec = C64_BF16X6;
  return old;
}
const char* mpdo_last_error(void) { return g_err.c_str(); }

void mpdo_release_workspace(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (it->first.first == dev && !it->second.bound) {
      if (it->second.p) {
        (void)hipDeviceSynchronize();
        (void)hipFree(it->second.p);
      }
      it = g_ws.erase(it);
    } else {
      ++it;
    }
  }
}

mpdo_status mpdo_bind_workspace(void* stream, void* ptr, size_t bytes) {
  int dev = 0;
  CK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, (hipStream_t)stream);
  auto it = g_ws.find(key);
  if (it != g_ws.end() && !it->second.bound && it->second.p) {
    CK(hipStreamSynchronize((hipStream_t)stream));
    CK(hipFree(it->second.p));
  }
  if (!ptr) {
    if (it != g_ws.end()) g_ws.erase(it);
    return MPDO_OK;
  }
  if (((uintptr_t)ptr & 255) != 0) return fail(MPDO_EINVAL, "the workspace must be 256-byte aligned");
  DevWS& w = g_ws[key];
  w.p = ptr;
  w.cap = bytes;
  w.bound = true;
  return MPDO_OK;
}

size_t mpdo_workspace_needed(void) { return t_ws_needed; }

// ------------------------------------------------------------------------------ 1q gates
mpdo_status mpdo_apply_1q(int nsites, const mpdo_site* in, const mpdo_site* out, const float* kraus,
                          const int* K, void* stream) {
  if (nsites <= 0) return MPDO_OK;
  if (!in || !out || !kraus || !K) return fail(MPDO_EINVAL, "null argument");
  std::vector<Site1q> tab(nsites);
  long long koff = 0, maxe = 0;
  for (int i = 0; i < nsites; ++i) {
    if (bad_site(in[i]) || !out[i].data) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (K[i] < 1 || K[i] > 4096) return fail(MPDO_EINVAL, "site %d: K=%d", i, K[i]);
    if (out[i].l != in[i].l || out[i].r != in[i].r || out[i].a != in[i].a * K[i])
      return fail(MPDO_EINVAL, "site %d: output dims (%d,%d,%d) != (%d,%d,%d)", i, out[i].l, out[i].a, out[i].r,
                  in[i].l, in[i].a * K[i], in[i].r);
    if (out[i].data == in[i].data) return fail(MPDO_EINVAL, "site %d: output aliases input", i);
    tab[i] = Site1q{(const float2*)in[i].data, (float2*)out[i].data, in[i].l, in[i].a, in[i].r, K[i], koff};
    koff += 4LL * K[i];
    maxe = std::max(maxe, (long long)in[i].l * in[i].a * in[i].r);
  }
  hipStream_t s = S(stream);
  Carver c;
  c.get<Site1q>(nsites);
  c.get<float2>(koff);
  char* base;
  ST(ws_reserve(c.off, s, &base));
  Carver r;
  r.base = base;
  Site1q* dtab = r.get<Site1q>(nsites);
  float2* dk = r.get<float2>(koff);
  CK(hipMemcpyAsync(dtab, tab.data(), sizeof(Site1q) * nsites, hipMemcpyHostToDevice, s));
  CK(hipMemcpyAsync(dk, kraus, sizeof(float2) * koff, hipMemcpyHostToDevice, s));
  CK(apply_1q(nsites, dtab, dk, maxe, s));
  // the host arrays above are pageable: make sure the copies completed before they go away
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ 2q gates
int mpdo_apply_2q_capacity(const mpdo_site* tmin, const mpdo_site* tmax, int K) {
  if (!tmin || !tmax || K < 1) return 0;
  long long rows1 = (long long)tmin->l * tmin->a, cols1 = 2LL * tmin->r;
  long long rows2 = 2LL * tmax->l, cols2 = (long long)tmax->a * tmax->r;
  long long p1 = std::min(rows1, cols1), p2 = std::min(rows2, cols2);
  long long cap = std::min(2 * p1, 2LL * K * p2);
  long long lref = std::min(2LL * tmin->l * tmin->a, 2LL * tmax->a * K * tmax->r);
  return (int)std::min(cap, lref);
}

mpdo_status mpdo_apply_2q(const mpdo_site* tmin, const mpdo_site* tmax, const float* gate, int K, double eps_abs,
                          mpdo_site* out_min, mpdo_site* out_max, double* discarded, void* stream) {
  if (!tmin || !tmax || !gate || !out_min || !out_max) return fail(MPDO_EINVAL, "null argument");
  if (bad_site(*tmin) || bad_site(*tmax)) return fail(MPDO_EINVAL, "bad input site");
  if (tmin->r != tmax->l) return fail(MPDO_EINVAL, "bond mismatch %d != %d", tmin->r, tmax->l);
  if (K < 1 || K > 4096) return fail(MPDO_EINVAL, "K=%d", K);
  if (!out_min->data || !out_max->data) return fail(MPDO_EINVAL, "null output buffer");
  hipStream_t s = S(stream);
  const int l = tmin->l, a1 = tmin->a, m = tmin->r, a2 = tmax->a, r = tmax->r;
  const int rows1 = l * a1, cols1 = 2 * m, rows2 = 2 * m, cols2 = a2 * r;
  const bool q1id = rows1 <= cols1, q2id = cols2 <= rows2;
  const int p1 = std::min(rows1, cols1), p2 = std::min(rows2, cols2);
  const int nC = 2 * p1;
  const long long colsC = 2LL * K * p2;
  const int cap = mpdo_apply_2q_capacity(tmin, tmax, K);
  const int len_ref = (int)std::min(2LL * l * a1, 2LL * a2 * K * r);

  float2 *X1, *R1 = nullptr, *F1 = nullptr, *Q1 = nullptr, *L2 = nullptr, *F2 = nullptr, *Q2 = nullptr, *Y, *C, *dG;
  float2 *FL, *FR, *Z, *Wm, *Tm;
  double2 *G1 = nullptr, *G2 = nullptr, *GC;
  Eig e1{}, e2{}, eC{};
  int* dk;
  double* dd;
  auto layout = [&](Carver& c) {
    X1 = c.get<float2>((size_t)rows1 * cols1);
    if (!q1id) {
      G1 = c.get<double2>((size_t)cols1 * cols1);
      carve_eig(c, cols1, e1);
      R1 = c.get<float2>((size_t)p1 * cols1);
      F1 = c.get<float2>((size_t)cols1 * p1);
      Q1 = c.get<float2>((size_t)rows1 * p1);
    }
    if (!q2id) {
      G2 = c.get<double2>((size_t)rows2 * rows2);
      carve_eig(c, rows2, e2);
      L2 = c.get<float2>((size_t)rows2 * p2);
      F2 = c.get<float2>((size_t)p2 * rows2);
      Q2 = c.get<float2>((size_t)p2 * cols2);
    }
    Y = c.get<float2>((size_t)4 * p1 * p2);
    C = c.get<float2>((size_t)nC * colsC);
    dG = c.get<float2>((size_t)16 * K);
    GC = c.get<double2>((size_t)nC * nC);
    carve_eig(c, nC, eC);
    FL = c.get<float2>((size_t)nC * cap);
    FR = c.get<float2>((size_t)nC * cap);
    Z = c.get<float2>((size_t)rows1 * 2 * cap);
    Wm = c.get<float2>((size_t)cap * colsC);
    Tm = c.get<float2>((size_t)cap * 2 * K * cols2);
    dk = c.get<int>(4);
    dd = c.get<double>(4);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);

  CK(hipMemcpyAsync(dG, gate, sizeof(float2) * 16 * K, hipMemcpyHostToDevice, s));
  // X1 = T_min as [(l a1) x (s1 m)]
  {
    int dims[4] = {l, 2, a1, m}, perm[4] = {0, 2, 1, 3};
    CK(permute((const float2*)tmin->data, X1, 4, dims, perm, s));
  }
  const float2* R1p = X1;     // (p1 x cols1)
  if (!q1id) {
    CK(gram_cols(X1, rows1, cols1, cols1, G1, s));
    ST(run_eig(cols1, G1, e1, s));
    CK(factor(e1.V, cols1, e1.w, cols1, cols1, cols1, 0.5, QR_FLOOR, 1, R1, cols1, s));
    CK(factor(e1.V, cols1, e1.w, cols1, cols1, cols1, -0.5, QR_FLOOR, 0, F1, cols1, s));
    CK(mm(rows1, p1, cols1, X1, cols1, F1, p1, Q1, p1, s));
    R1p = R1;
  }
  const float2* X2 = (const float2*)tmax->data;   // [(m s2) x (a2 r)]
  const float2* L2p = X2;    // (rows2 x p2)
  if (!q2id) {
    CK(gram_rows(X2, rows2, cols2, cols2, G2, s));
    ST(run_eig(rows2, G2, e2, s));
    CK(factor(e2.V, rows2, e2.w, rows2, rows2, rows2, 0.5, QR_FLOOR, 0, L2, p2, s));
    CK(factor(e2.V, rows2, e2.w, rows2, rows2, rows2, -0.5, QR_FLOOR, 1, F2, rows2, s));
    CK(mm(p2, cols2, rows2, F2, rows2, X2, cols2, Q2, cols2, s));
    L2p = L2;
  }
  // Y[s1][s2] = R1[:, s1 m : (s1+1) m] * L2[(m s2), :]
  for (int s1 = 0; s1 < 2; ++s1)
    for (int s2 = 0; s2 < 2; ++s2) {
      GemmArgs g{};
      g.M = p1; g.N = p2; g.K = m; g.batch = 1;
      g.A = R1p + (long long)s1 * m; g.rsA = cols1; g.csA = 1;
      g.B = L2p + (long long)s2 * p2; g.rsB = 2LL * p2; g.csB = 1;
      g.C = Y + (long long)(s1 * 2 + s2) * p1 * p2; g.rsC = p2; g.csC = 1;
      g.alpha_re = 1;
      CK(gemm(g_tensor_prec, g, s));
    }
  CK(channel2q(Y, dG, K, p1, p2, C, s));
  CK(gram_rows(C, nC, (int)colsC, colsC, GC, s));
  ST(run_eig(nC, GC, eC, s));
  CK(keep_count(eC.w, nC, nC, len_ref, -1, eps_abs, 0, dk, dd, 1, s));
  int k = 0;
  double disc = 0;
  ST(read_k(dk, dd, &k, &disc, s));
  if (k > cap) return fail(MPDO_EINVAL, "internal: rank %d above capacity %d", k, cap);
  // T_min' = (Q1 (x) I) U sqrt(S)  -> (l, a1, s1', k) -> (l, s1', a1, k)
  CK(factor(eC.V, nC, eC.w, nC, nC, k, 0.25, SPLIT_FLOOR, 0, FL, k, s));
  const float2* Zs = FL;
  if (!q1id) {
    CK(mm(rows1, 2 * k, p1, Q1, p1, FL, 2 * k, Z, 2 * k, s));
    Zs = Z;
  }
  {
    int dims[4] = {l, a1, 2, k}, perm[4] = {0, 2, 1, 3};
    CK(permute(Zs, (float2*)out_min->data, 4, dims, perm, s));
  }
  // T_max' = S^{-1/2} U^H C (I (x) Q2) -> (k, s2', kk, a2, r) -> (k, s2', a2, kk, r)
  CK(factor(eC.V, nC, eC.w, nC, nC, k, -0.25, SPLIT_FLOOR, 1, FR, nC, s));
  CK(mm(k, (int)colsC, nC, FR, nC, C, colsC, Wm, colsC, s));
  const float2* Ts = Wm;
  if (!q2id) {
    CK(mm(k * 2 * K, cols2, p2, Wm, p2, Q2, cols2, Tm, cols2, s));
    Ts = Tm;
  }
  {
    int dims[5] = {k, 2, K, a2, r}, perm[5] = {0, 1, 3, 2, 4};
    CK(permute(Ts, (float2*)out_max->data, 5, dims, perm, s));
  }
  out_min->l = l; out_min->a = a1; out_min->r = k;
  out_max->l = k; out_max->a = a2 * K; out_max->r = r;
  if (discarded) *discarded = disc;
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ QR sweep
void mpdo_qr_left2right_dims(int n, const mpdo_site* in, mpdo_site* out) {
  int l = n > 0 ? in[0].l : 1;
  for (int i = 0; i < n; ++i) {
    out[i].l = l;
    out[i].a = in[i].a;
    if (i == n - 1) {
      out[i].r = in[i].r;
    } else {
      long long rows = (long long)l * 2 * in[i].a;
      int k = (int)std::min(rows, (long long)in[i].r);
      out[i].r = k;
      l = k;
    }
  }
}

mpdo_status mpdo_qr_left2right(int n, const mpdo_site* in, mpdo_site* out, void* stream) {
  if (n <= 0) return MPDO_OK;
  if (!in || !out) return fail(MPDO_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) {
    if (bad_site(in[i]) || !out[i].data) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (i + 1 < n && in[i].r != in[i + 1].l) return fail(MPDO_EINVAL, "bond %d mismatch", i);
  }
  std::vector<mpdo_site> d(n);
  mpdo_qr_left2right_dims(n, in, d.data());
  hipStream_t s = S(stream);
  size_t maxM = 0;
  int maxc = 1;
  for (int i = 0; i < n; ++i) {
    maxM = std::max(maxM, (size_t)d[i].l * 2 * d[i].a * in[i].r);
    maxc = std::max(maxc, in[i].r);
  }
  float2 *buf0, *buf1, *A1, *F, *R;
  double2 *G, *P1, *P2;
  Eig e{}, e2{};
  auto layout = [&](Carver& c) {
    buf0 = c.get<float2>(maxM);
    buf1 = c.get<float2>(maxM);
    A1 = c.get<float2>(maxM);
    F = c.get<float2>((size_t)maxc * maxc);
    R = c.get<float2>((size_t)maxc * maxc);
    G = c.get<double2>((size_t)maxc * maxc);
    P1 = c.get<double2>((size_t)maxc * maxc);
    P2 = c.get<double2>((size_t)maxc * maxc);
    carve_eig(c, maxc, e);
    carve_eig(c, maxc, e2);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);
  const float2* M = (const float2*)in[0].data;
  float2* nxt = buf0;
  for (int i = 0; i < n - 1; ++i) {
    const int rows = d[i].l * 2 * d[i].a, cols = in[i].r, k = d[i].r;
    const float2* Rm;
    long long ldr;
    if (rows <= cols) {
      CK(eye((float2*)out[i].data, rows, rows, s));
      Rm = M;
      ldr = cols;
    } else {
      // pass 1: M^H M = W L W^H, A1 = M W L^{-1/2};  pass 2 re-orthonormalises A1 the same way
      // (A1^H A1 = W2 L2 W2^H, Q = A1 W2 L2^{-1/2}), R = (L2^{1/2} W2^H)(L^{1/2} W^H).
      CK(gram_cols(M, rows, cols, cols, G, s));
      ST(run_eig(cols, G, e, s));
      CK(factor(e.V, cols, e.w, cols, cols, cols, -0.5, QR_FLOOR, 0, F, cols, s));
      CK(mm(rows, cols, cols, M, cols, F, cols, A1, cols, s));
      CK(gram_cols(A1, rows, cols, cols, G, s));
      ST(run_eig(cols, G, e2, s));
      CK(factor(e2.V, cols, e2.w, cols, cols, cols, -0.5, QR_FLOOR, 0, F, cols, s));
      CK(mm(rows, cols, cols, A1, cols, F, cols, (float2*)out[i].data, cols, s));
      CK(factor64(e2.V, cols, e2.w, cols, cols, cols, 0.5, QR_FLOOR, 1, P2, cols, s));
      CK(factor64(e.V, cols, e.w, cols, cols, cols, 0.5, QR_FLOOR, 1, P1, cols, s));
      CK(mm64(C128_IN_C64_OUT, cols, cols, cols, P2, cols, false, false, P1, cols, R, cols, s));
      Rm = R;
      ldr = cols;
    }
    const int colsN = 2 * in[i + 1].a * in[i + 1].r;
    CK(mm(k, colsN, cols, Rm, ldr, (const float2*)in[i + 1].data, colsN, nxt, colsN, s));
    M = nxt;
    nxt = (nxt == buf0) ? buf1 : buf0;
  }
  CK(hipMemcpyAsync(out[n - 1].data, M, sizeof(float2) * (size_t)d[n - 1].l * 2 * d[n - 1].a * d[n - 1].r,
                    hipMemcpyDeviceToDevice, s));
  for (int i = 0; i < n; ++i) { out[i].l = d[i].l; out[i].a = d[i].a; out[i].r = d[i].r; }
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ SVD sweeps
void mpdo_svd_sweep_dims(int n, const mpdo_site* in, int direction, int msv, mpdo_site* out) {
  for (int i = 0; i < n; ++i) out[i] = in[i];
  if (direction == 0) {
    for (int idx = n - 1; idx >= 1; --idx) {
      long long lr = std::min(2LL * out[idx - 1].l * out[idx - 1].a, 2LL * out[idx].a * out[idx].r);
      int k = (int)(msv > 0 ? std::min((long long)msv, lr) : lr);
      out[idx].l = k;
      out[idx - 1].r = k;
    }
  } else {
    for (int idx = 0; idx + 1 < n; ++idx) {
      long long lr = std::min(2LL * out[idx].l * out[idx].a, 2LL * out[idx + 1].a * out[idx + 1].r);
      int k = (int)(msv > 0 ? std::min((long long)msv, lr) : lr);
      out[idx].r = k;
      out[idx + 1].l = k;
    }
  }
}

mpdo_status mpdo_svd_sweep(int n, const mpdo_site* in, mpdo_site* out, int direction, int msv, double err,
                           int left_canonical, double* discarded, void* stream) {
  if (n <= 0) return MPDO_OK;
  if (!in || !out) return fail(MPDO_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) {
    if (bad_site(in[i]) || !out[i].data) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (i + 1 < n && in[i].r != in[i + 1].l) return fail(MPDO_EINVAL, "bond %d mismatch", i);
  }
  hipStream_t s = S(stream);
  std::vector<mpdo_site> cap(n);
  mpdo_svd_sweep_dims(n, in, direction, msv, cap.data());
  // workspace bounds: the running tensor and the bond dimension D of each step
  size_t maxT = 0;
  int maxD = 1, maxk = 1;
  for (int i = 0; i < n; ++i) {
    maxT = std::max(maxT, (size_t)cap[i].l * 2 * cap[i].a * cap[i].r);
    maxT = std::max(maxT, (size_t)in[i].l * 2 * in[i].a * cap[i].r);
    maxT = std::max(maxT, (size_t)cap[i].l * 2 * in[i].a * in[i].r);
    maxD = std::max(maxD, in[i].r);
    maxk = std::max(maxk, std::max(cap[i].l, cap[i].r));
  }
  float2 *cur0, *cur1;
  ChiWS w{};
  auto layout = [&](Carver& c) {
    cur0 = c.get<float2>(maxT);
    cur1 = c.get<float2>(maxT);
    carve_chi(c, maxD, std::max(maxk, maxD), w);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);
  std::vector<mpdo_site> cur(in, in + n);
  float2* nxt = cur0;
  if (n == 1) {
    CK(hipMemcpyAsync(out[0].data, in[0].data, sizeof(float2) * (size_t)in[0].l * 2 * in[0].a * in[0].r,
                      hipMemcpyDeviceToDevice, s));
    out[0].l = in[0].l; out[0].a = in[0].a; out[0].r = in[0].r;
    return MPDO_OK;
  }
  if (direction == 0) {
    for (int idx = n - 1; idx >= 1; --idx) {
      mpdo_site& A = cur[idx - 1];
      mpdo_site& B = cur[idx];
      const int rowsA = A.l * 2 * A.a, D = A.r, colsB = 2 * B.a * B.r;
      const int len_ref = (int)std::min((long long)rowsA, (long long)colsB);
      int k = 0;
      double disc = 0;
      ST(chi_step((const float2*)A.data, rowsA, D, (const float2*)B.data, colsB, left_canonical != 0, msv, err,
                  len_ref, nxt, (float2*)out[idx].data, &k, &disc, w, s));
      out[idx].l = k; out[idx].a = B.a; out[idx].r = B.r;
      if (discarded) discarded[n - 1 - idx] = disc;
      A.data = nxt;
      A.r = k;
      nxt = (nxt == cur0) ? cur1 : cur0;
    }
    CK(hipMemcpyAsync(out[0].data, cur[0].data, sizeof(float2) * (size_t)cur[0].l * 2 * cur[0].a * cur[0].r,
                      hipMemcpyDeviceToDevice, s));
    out[0].l = cur[0].l; out[0].a = cur[0].a; out[0].r = cur[0].r;
  } else {
    for (int idx = 0; idx + 1 < n; ++idx) {
      mpdo_site& A = cur[idx];
      mpdo_site& B = cur[idx + 1];
      const int rowsA = A.l * 2 * A.a, D = A.r, colsB = 2 * B.a * B.r;
      const int len_ref = (int)std::min((long long)rowsA, (long long)colsB);
      int k = 0;
      double disc = 0;
      ST(chi_step((const float2*)A.data, rowsA, D, (const float2*)B.data, colsB, false, msv, err, len_ref,
                  (float2*)out[idx].data, nxt, &k, &disc, w, s));
      out[idx].l = A.l; out[idx].a = A.a; out[idx].r = k;
      if (discarded) discarded[idx] = disc;
      B.data = nxt;
      B.l = k;
      nxt = (nxt == cur0) ? cur1 : cur0;
    }
    CK(hipMemcpyAsync(out[n - 1].data, cur[n - 1].data,
                      sizeof(float2) * (size_t)cur[n - 1].l * 2 * cur[n - 1].a * cur[n - 1].r,
                      hipMemcpyDeviceToDevice, s));
    out[n - 1].l = cur[n - 1].l; out[n - 1].a = cur[n - 1].a; out[n - 1].r = cur[n - 1].r;
  }
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ kappa sweep
mpdo_status mpdo_svd_kappa(int nsites, const mpdo_site* in, mpdo_site* out, int msv, double err,
                           double* discarded, void* stream) {
  if (nsites <= 0) return MPDO_OK;
  if (!in || !out) return fail(MPDO_EINVAL, "null argument");
  size_t maxT = 0;
  int maxa = 1;
  for (int i = 0; i < nsites; ++i) {
    if (bad_site(in[i]) || !out[i].data) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    maxT = std::max(maxT, (size_t)in[i].l * 2 * in[i].a * in[i].r);
    maxa = std::max(maxa, in[i].a);
  }
  hipStream_t s = S(stream);
  float2 *Mp, *Vk;
  double2* G;
  Eig e{};
  int* dk;
  double* dd;
  auto layout = [&](Carver& c) {
    Mp = c.get<float2>(maxT);
    Vk = c.get<float2>((size_t)maxa * maxa);
    G = c.get<double2>((size_t)maxa * maxa);
    carve_eig(c, maxa, e);
    dk = c.get<int>(4);
    dd = c.get<double>(4);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);
  for (int i = 0; i < nsites; ++i) {
    const int l = in[i].l, a = in[i].a, r = in[i].r;
    const int rows = l * 2 * r;
    const int len_s = std::min(rows, a);
    int dims[4] = {l, 2, a, r}, perm[4] = {0, 1, 3, 2};
    CK(permute((const float2*)in[i].data, Mp, 4, dims, perm, s));
    CK(gram_cols(Mp, rows, a, a, G, s));
    ST(run_eig(a, G, e, s));
    CK(keep_count(e.w, a, a, len_s, msv, err, 1, dk, dd, 1, s));
    int k = 0;
    double disc = 0;
    ST(read_k(dk, dd, &k, &disc, s));
    CK(factor(e.V, a, e.w, a, a, k, 0.0, 0.0, 0, Vk, k, s));
    // out[l, s, c, r] = sum_a Vk[a, c] T[l, s, a, r]
    GemmArgs g{};
    g.M = k; g.N = r; g.K = a; g.batch = 2 * l;
    g.A = Vk; g.rsA = 1; g.csA = k; g.bsA = 0;
    g.B = in[i].data; g.rsB = r; g.csB = 1; g.bsB = (long long)a * r;
    g.C = out[i].data; g.rsC = r; g.csC = 1; g.bsC = (long long)k * r;
    g.alpha_re = 1;
    CK(gemm(g_tensor_prec, g, s));
    out[i].l = l; out[i].a = k; out[i].r = r;
    if (discarded) discarded[i] = disc;
  }
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ observables
mpdo_status mpdo_pauli_expect(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out,
                              void* stream) {
  size_t per = 0;
  ST(expect_check(n, sites, nstrings, codes, out, &per));
  if (n <= 0 || nstrings <= 0) return MPDO_OK;
  // strings per chunk: every string at once up to 512 MiB of environments and intermediates
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nstrings, (size_t)EXPECT_MAX_CHUNK,
                                                                  (size_t)(512u << 20) / per}));
  char* base;
  ST(ws_reserve(expect_chunk_bytes(n, sites, chunk), S(stream), &base));
  return expect_run(n, sites, nstrings, codes, out, base, chunk, S(stream));
}

mpdo_status mpdo_reduced_dm(int n, const mpdo_site* sites, const int* keep, void* out, void* stream) {
  if (n <= 0) return MPDO_OK;
  if (!sites || !keep || !out) return fail(MPDO_EINVAL, "null argument");
  int nk = 0;
  for (int i = 0; i < n; ++i) {
    if (bad_site(sites[i])) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (i + 1 < n && sites[i].r != sites[i + 1].l) return fail(MPDO_EINVAL, "bond %d mismatch", i);
    nk += keep[i] ? 1 : 0;
  }
  if (nk > 12) return fail(MPDO_EINVAL, "reduced density matrix of %d qubits is too large", nk);
  if (sites[0].l != 1 || sites[n - 1].r != 1) return fail(MPDO_EINVAL, "open boundary bonds must be 1");
  hipStream_t s = S(stream);
  // bound the environment sizes: env[KB][l][l'], X[KB][l'][2 a r]
  size_t maxEnv = 1, maxX = 1, maxT = 1;
  {
    long long kb = 1;
    for (int i = 0; i < n; ++i) {
      const long long l = sites[i].l, a = sites[i].a, r = sites[i].r;
      maxEnv = std::max(maxEnv, (size_t)(kb * l * l));
      maxX = std::max(maxX, (size_t)(kb * l * 2 * a * r));
      maxT = std::max(maxT, (size_t)(l * 2 * a * r));
      if (keep[i]) kb *= 4;
      maxEnv = std::max(maxEnv, (size_t)(kb * r * r));
    }
  }
  float2 *env0, *env1, *X, *Xp, *Tp;
  auto layout = [&](Carver& c) {
    env0 = c.get<float2>(maxEnv);
    env1 = c.get<float2>(maxEnv);
    X = c.get<float2>(maxX);
    Xp = c.get<float2>(maxX);
    Tp = c.get<float2>(maxT);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);
  float2 one = make_float2(1.f, 0.f);
  CK(hipMemcpyAsync(env0, &one, sizeof(float2), hipMemcpyHostToDevice, s));
  float2* env = env0;
  float2* nenv = env1;
  long long kket = 1;   // ket (= bra) index range so far
  for (int i = 0; i < n; ++i) {
    const int l = sites[i].l, a = sites[i].a, r = sites[i].r;
    const long long KB = kket * kket;
    const int cols = 2 * a * r;
    // Tp = T permuted (l, s, a, r) -> (s, l, a, r)
    {
      int dims[4] = {l, 2, a, r}, perm[4] = {1, 0, 2, 3};
      CK(permute((const float2*)sites[i].data, Tp, 4, dims, perm, s));
    }
    // X_b[l', (s a r)] = sum_l env_b[l, l'] T[l, (s a r)]
    for (long long b0 = 0; b0 < KB; b0 += 65535) {
      GemmArgs g{};
      g.M = l; g.N = cols; g.K = l; g.batch = (int)std::min(65535LL, KB - b0);
      g.A = env + b0 * l * l; g.rsA = 1; g.csA = l; g.bsA = (long long)l * l;
      g.B = sites[i].data; g.rsB = cols; g.csB = 1; g.bsB = 0;
      g.C = X + b0 * l * cols; g.rsC = cols; g.csC = 1; g.bsC = (long long)l * cols;
      g.alpha_re = 1;
      CK(gemm(C64_F64ACC_C64OUT, g, s));
    }
    // Xp = X permuted (KB, l', s, a, r) -> (KB, s, l', a, r)
    {
      int dims[4] = {l, 2, a, r}, perm[4] = {1, 0, 2, 3};
      CK(permute(X, Xp, 4, dims, perm, s, (int)KB, (long long)l * cols));
    }
    const long long la = (long long)l * a;
    if (!keep[i]) {
      // nenv_b[r, r'] = sum_{s, l', a} Xp_b[(s l' a), r] conj(Tp[(s l' a), r'])
      for (long long b0 = 0; b0 < KB; b0 += 65535) {
        GemmArgs g{};
        g.M = r; g.N = r; g.K = (int)(2 * la); g.batch = (int)std::min(65535LL, KB - b0);
        g.A = Xp + b0 * l * cols; g.rsA = 1; g.csA = r; g.bsA = (long long)l * cols;
        g.B = Tp; g.rsB = r; g.csB = 1; g.bsB = 0; g.conjB = 1;
        g.C = nenv + b0 * r * r; g.rsC = r; g.csC = 1; g.bsC = (long long)r * r;
        g.alpha_re = 1;
        CK(gemm(C64_F64ACC_C64OUT, g, s));
      }
    } else {
      // tmp[b][s][s'][r][r'] then regroup (ket, bra, s, s') -> (ket, s, bra, s')
      float2* tmp = X;   // X is free again (Xp holds the data)
      for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2)
          for (long long b0 = 0; b0 < KB; b0 += 65535) {
            GemmArgs g{};
            g.M = r; g.N = r; g.K = (int)la; g.batch = (int)std::min(65535LL, KB - b0);
            g.A = Xp + b0 * l * cols + (long long)s1 * la * r; g.rsA = 1; g.csA = r; g.bsA = (long long)l * cols;
            g.B = Tp + (long long)s2 * la * r; g.rsB = r; g.csB = 1; g.bsB = 0; g.conjB = 1;
            g.C = tmp + b0 * 4 * r * r + (long long)(s1 * 2 + s2) * r * r; g.rsC = r; g.csC = 1;
            g.bsC = 4LL * r * r;
            g.alpha_re = 1;
            CK(gemm(C64_F64ACC_C64OUT, g, s));
          }
      int dims[6] = {(int)kket, (int)kket, 2, 2, r, r}, perm[6] = {0, 2, 1, 3, 4, 5};
      CK(permute(tmp, nenv, 6, dims, perm, s));
      kket *= 2;
    }
    std::swap(env, nenv);
  }
  CK(hipMemcpyAsync(out, env, sizeof(float2) * (size_t)(kket * kket), hipMemcpyDeviceToDevice, s));
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

mpdo_status mpdo_trace_rho_rho(int n, const mpdo_site* a, const mpdo_site* b, double* out, void* stream) {
  return mpdo_trace_rho_rho_masked(n, a, b, nullptr, out, stream);
}

mpdo_status mpdo_contract_state(int n, const mpdo_site* sites, void* out, void* stream) {
  if (n <= 0) return MPDO_OK;
  if (!sites || !out) return fail(MPDO_EINVAL, "null argument");
  if (n > 30) return fail(MPDO_EINVAL, "state vector of %d qubits is too large", n);
  size_t maxv = 2;
  for (int i = 0; i < n; ++i) {
    if (bad_site(sites[i])) return fail(MPDO_EINVAL, "site %d: bad descriptor", i);
    if (sites[i].a != 1) return fail(MPDO_EINVAL, "site %d carries an inner index (a=%d): not a pure state", i, sites[i].a);
    if (i + 1 < n && sites[i].r != sites[i + 1].l) return fail(MPDO_EINVAL, "bond %d mismatch", i);
    maxv = std::max(maxv, ((size_t)2 << i) * sites[i].r);
  }
  if (sites[0].l != 1 || sites[n - 1].r != 1) return fail(MPDO_EINVAL, "open boundary bonds must be 1");
  hipStream_t s = S(stream);
  float2 *v0, *v1;
  auto layout = [&](Carver& c) {
    v0 = c.get<float2>(maxv);
    v1 = c.get<float2>(maxv);
  };
  Carver c0;
  layout(c0);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  Carver c1;
  c1.base = base;
  layout(c1);
  // psi_0 = T_0 as (2 x r0); psi_{i} = psi_{i-1} (2^i x r_{i-1}) * T_i (r_{i-1} x 2 r_i)
  const float2* psi = (const float2*)sites[0].data;
  float2* nxt = v0;
  long long rows = 2;
  for (int i = 1; i < n; ++i) {
    const int l = sites[i].l, r = sites[i].r;
    float2* dst = (i == n - 1) ? (float2*)out : nxt;
    CK(mm((int)rows, 2 * r, l, psi, l, (const float2*)sites[i].data, 2 * r, dst, 2 * r, s));
    psi = dst;
    rows *= 2;
    nxt = (nxt == v0) ? v1 : v0;
  }
  if (n == 1) CK(hipMemcpyAsync(out, sites[0].data, 2 * sizeof(float2), hipMemcpyDeviceToDevice, s));
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ eigensolver
mpdo_status mpdo_eigh(int n, int batch, void* A, double* w, void* V, void* stream) {
  if (n <= 0 || batch <= 0) return MPDO_OK;
  if (!A || !w || !V) return fail(MPDO_EINVAL, "null argument");
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(eigh_workspace_bytes(n, batch));
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  int sw = 0;
  CK(eigh(n, batch, A, n, (long long)n * n, w, n, V, n, (long long)n * n, base, s, &sw));
  CK(hipStreamSynchronize(s));
  if (sw > 60) return fail(MPDO_ENOCONV, "Jacobi eigensolver did not converge (n=%d, batch %d)", n, batch);
  return MPDO_OK;
}

// ------------------------------------------------------------------------------ test hooks
void mpdo_debug_fail_eig(int count) { g_fail_eig.store(count > 0 ? count : 0); }

mpdo_status mpdo_debug_gemm_ex(int prec, int herm, int inA, int inB, int tri, int M, int N, int K, int batch,
                               const void* A,
                               long long rsA, long long csA, long long bsA, int conjA, const void* B, long long rsB,
                               long long csB, long long bsB, int conjB, void* C, long long rsC, long long csC,
                               long long bsC, double alpha_re, double alpha_im, double beta_re, double beta_im,
                               void* stream) {
  if (prec < 0 || prec > 7) return fail(MPDO_EINVAL, "prec %d", prec);
  if (herm && M != N) return fail(MPDO_EINVAL, "Hermitian output needs M == N (%d, %d)", M, N);
  if (inA < 0 || inA > 2 || inB < 0 || inB > 2) return fail(MPDO_EINVAL, "operand storage %d %d", inA, inB);
  if (tri < 0 || tri > 2 || (tri && herm)) return fail(MPDO_EINVAL, "tri %d", tri);
  GemmArgs g{};
  g.tri = tri;
  g.herm = herm != 0;
  g.inA = inA; g.inB = inB;
  g.M = M; g.N = N; g.K = K; g.batch = batch;
  g.A = A; g.rsA = rsA; g.csA = csA; g.bsA = bsA; g.conjA = conjA;
  g.B = B; g.rsB = rsB; g.csB = csB; g.bsB = bsB; g.conjB = conjB;
  g.C = C; g.rsC = rsC; g.csC = csC; g.bsC = bsC;
  g.alpha_re = alpha_re; g.alpha_im = alpha_im; g.beta_re = beta_re; g.beta_im = beta_im;
  CK(gemm((Prec)prec, g, S(stream)));
  CK(hipStreamSynchronize(S(stream)));
  return MPDO_OK;
}

mpdo_status mpdo_debug_gram_split(int n, int K, int batch, int slices, const void* X, void* parts, void* G,
                                  void* stream) {
  if (n <= 0 || K <= 0 || batch <= 0 || slices < 1 || slices > 64 || K % slices)
    return fail(MPDO_EINVAL, "n=%d K=%d batch=%d slices=%d", n, K, batch, slices);
  hipStream_t st = S(stream);
  const long long nn = (long long)n * n;
  GemmArgs g{};
  g.M = n; g.N = n; g.K = K / slices; g.batch = batch;
  g.A = X; g.rsA = K; g.csA = 1; g.bsA = (long long)n * K;
  g.B = X; g.rsB = 1; g.csB = K; g.bsB = (long long)n * K; g.conjB = 1;
  g.C = parts; g.rsC = n; g.csC = 1; g.bsC = slices * nn;
  g.alpha_re = 1;
  g.herm = 1;
  g.ksplit = slices; g.ksA = K / slices; g.ksB = K / slices; g.ksC = nn;
  CK(gemm(C64_F64ACC, g, st));
  CK(sum_parts((const double2*)parts, slices, nn, (double2*)G, batch, st));
  CK(herm_mirror_b(G, true, n, nn, batch, st));
  CK(hipStreamSynchronize(st));
  return MPDO_OK;
}

mpdo_status mpdo_debug_gemm(int prec, int M, int N, int K, int batch, const void* A, long long rsA, long long csA,
                            long long bsA, int conjA, const void* B, long long rsB, long long csB, long long bsB,
                            int conjB, void* C, long long rsC, long long csC, long long bsC, double alpha_re,
                            double alpha_im, double beta_re, double beta_im, void* stream) {
  const int herm = prec >= 10;
  return mpdo_debug_gemm_ex(herm ? prec - 10 : prec, herm, 0, 0, 0, M, N, K, batch, A, rsA, csA, bsA, conjA, B, rsB, csB,
                            bsB, conjB, C, rsC, csC, bsC, alpha_re, alpha_im, beta_re, beta_im, stream);
}

mpdo_status mpdo_debug_eigh(int n, int batch, void* A, double* w, void* V, int* sweeps, void* stream) {
  if (n <= 0 || batch <= 0) return fail(MPDO_EINVAL, "n=%d batch=%d", n, batch);
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(eigh_workspace_bytes(n, batch));
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  int sw = 0;
  CK(eigh(n, batch, A, n, (long long)n * n, w, n, V, n, (long long)n * n, base, s, &sw));
  CK(hipStreamSynchronize(s));
  if (sweeps) *sweeps = sw;
  return MPDO_OK;
}

mpdo_status mpdo_debug_gram_tail_cut(int n, int batch, const void* G, double eps, int m, int q, int* kept,
                                     double* disc, void* P, int* info, void* stream) {
  if (n <= 0 || batch <= 0 || !G || !kept || !disc || !P || !info)
    return fail(MPDO_EINVAL, "n=%d batch=%d or a null pointer", n, batch);
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(gram_tail_workspace_bytes(n, m, batch));
  const size_t o_kept = c0.off;
  c0.get<int>(batch);
  const size_t o_disc = c0.off;
  c0.get<double>(batch);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  int* dkept = (int*)(base + o_kept);
  double* ddisc = (double*)(base + o_disc);
  TailResult r;
  CK(gram_tail_cut(n, batch, (const double2*)G, eps, m, q, base, s, dkept, ddisc, (float2*)P, n, (long long)n * n,
                   &r));
  if (r.ok) {
    CK(hipMemcpyAsync(kept, dkept, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    CK(hipMemcpyAsync(disc, ddisc, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
  }
  CK(hipStreamSynchronize(s));
  info[0] = r.ok;
  info[1] = r.any_drop;
  info[2] = r.max_kept;
  info[3] = r.reason;
  return MPDO_OK;
}

mpdo_status mpdo_debug_gram_top_eig(int n, int batch, const void* G, int k, int m, int q, double* w, void* V, int* info,
                                    void* stream) {
  if (n <= 0 || batch <= 0 || !G || !w || !V || !info) return fail(MPDO_EINVAL, "n=%d batch=%d or a null pointer", n, batch);
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(gram_top_workspace_bytes(n, m, batch));
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  TopResult r;
  CK(gram_top_eig(n, batch, (const double2*)G, k, m, q, base, s, w, n, (double2*)V, n, (long long)n * n, &r));
  CK(hipStreamSynchronize(s));
  info[0] = r.ok;
  info[1] = r.reason;
  return MPDO_OK;
}

mpdo_status mpdo_debug_chol(int n, int batch, void* A, double tol_rel, void* L, int* rank, void* stream) {
  if (n <= 0 || batch <= 0 || !A || !L || !rank) return fail(MPDO_EINVAL, "n=%d batch=%d or a null pointer", n, batch);
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(chol_workspace_bytes(batch));
  const size_t o_piv = c0.off;
  c0.get<int>((size_t)n * batch);
  const size_t o_rank = c0.off;
  c0.get<int>((size_t)batch);
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  int* dr = (int*)(base + o_rank);
  CK(chol_b(n, batch, A, n, (long long)n * n, tol_rel, (double2*)L, (int*)(base + o_piv), dr, base, s));
  CK(hipMemcpyAsync(rank, dr, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  return MPDO_OK;
}

mpdo_status mpdo_debug_eigh_gram(int n, int batch, void* A, double* w, void* V, int* sweeps, void* stream) {
  if (n <= 0 || batch <= 0) return fail(MPDO_EINVAL, "n=%d batch=%d", n, batch);
  hipStream_t s = S(stream);
  Carver c0;
  c0.get<char>(eigh_gram_workspace_bytes(n, batch));
  char* base;
  ST(ws_reserve(c0.off, s, &base));
  int sw = 0;
  CK(eigh_gram(n, batch, A, n, (long long)n * n, w, n, V, n, (long long)n * n, base, s, &sw));
  CK(hipStreamSynchronize(s));
  if (sweeps) *sweeps = sw;
  return MPDO_OK;
}

}  // extern "C"
