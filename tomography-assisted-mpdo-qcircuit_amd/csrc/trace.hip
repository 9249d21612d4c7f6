// Tr(rho0 rho1) of two MPDOs by a left-to-right transfer sweep (dmOperations.py:51-66, the network
// trace_rho_rho hands to contractors.auto), for any number of qubits.
//
// The environment F[l0][l1][m0][m1] carries the four open bonds: ket and bra of rho0 (A, conj A)
// and ket and bra of rho1 (B, conj B), complex128. At a site both states keep, the physical index
// of A meets conj(B) and that of conj(A) meets B (rho0[s, s'] rho1[s', s]):
//   F'[r0, r1, n0, n1] = sum F[l0, l1, m0, m1] A[l0, s, a, r0] conj(B[m1, s, b, n1])
//                                            conj(A[m0, s', a, n0]) B[l1, s', b, r1]
// At a site both states trace (reduce_dmNodes), each state's physical index meets its own bra:
//   F'[r0, r1, n0, n1] = sum F[l0, l1, m0, m1] A[l0, s, a, r0] conj(A[m0, s, a, n0])
//                                            B[l1, t, b, r1] conj(B[m1, t, b, n1])
// Each update is a chain of strided complex128 GEMMs (fp64 MFMA) with two small permutes; the
// per-l1 intermediates are processed in slices that fit a fixed budget.
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

#define TCK(x)                                                \
  do {                                                        \
    if ((x) != hipSuccess) return layer_fail(MPDO_EHIP, #x);  \
  } while (0)

inline size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr long long SLICE_BUDGET = 1LL << 28;    // complex128 elements of the per-slice intermediates

struct P6 {
  long long dim[6];
  long long ost[6];   // output stride of each input axis
};

// out[sum_k i_k ost_k] = in[e], e the row-major index of (i_0 .. i_5) over dim
__global__ void permute64_kernel(const double2* in, double2* out, long long total, P6 p) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    long long rem = e, off = 0;
    for (int k = 5; k >= 0; --k) {
      const long long i = rem % p.dim[k];
      rem /= p.dim[k];
      off += i * p.ost[k];
    }
    out[off] = in[e];
  }
}

__global__ void widen_kernel(const float2* in, double2* out, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
    out[e] = make_double2(in[e].x, in[e].y);
}

int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192)); }

// permute a row-major tensor of the given dims (rank <= 6) into the axis order `order`
// (output axis j is input axis order[j])
hipError_t permute64(const double2* in, double2* out, std::vector<long long> dims, std::vector<int> order,
                     hipStream_t s) {
  const int r = (int)dims.size();
  P6 p;
  long long total = 1;
  for (long long d : dims) total *= d;
  std::vector<long long> ostride(r);
  long long st = 1;
  for (int j = r - 1; j >= 0; --j) {
    ostride[order[j]] = st;
    st *= dims[order[j]];
  }
  for (int k = 0; k < 6; ++k) {
    const int src = k - (6 - r);
    p.dim[k] = src >= 0 ? dims[src] : 1;
    p.ost[k] = src >= 0 ? ostride[src] : 0;
  }
  hipLaunchKernelGGL(permute64_kernel, dim3(grid_for(total)), dim3(256), 0, s, in, out, total, p);
  return hipGetLastError();
}

GemmArgs ga(int M, int N, int K, const void* A, long long rsA, long long csA, int conjA, const void* B, long long rsB,
            long long csB, int conjB, void* C, long long rsC, long long csC) {
  GemmArgs g{};
  g.M = M; g.N = N; g.K = K; g.batch = 1;
  g.A = A; g.rsA = rsA; g.csA = csA; g.conjA = conjA;
  g.B = B; g.rsB = rsB; g.csB = csB; g.conjB = conjB;
  g.C = C; g.rsC = rsC; g.csC = csC;
  g.alpha_re = 1;
  return g;
}

bool bad(const mpdo_site& t) { return !t.data || t.l <= 0 || t.a <= 0 || t.r <= 0; }

}  // namespace

extern "C" mpdo_status mpdo_trace_rho_rho_masked(int n, const mpdo_site* sa, const mpdo_site* sb, const int* keep,
                                                 double* out, void* stream) {
  if (n <= 0) return layer_fail(MPDO_EINVAL, "trace_rho_rho: no qubits");
  if (!sa || !sb || !out) return layer_fail(MPDO_EINVAL, "trace_rho_rho: null argument");
  long long maxF = 1, maxT = 1, maxY = 1, maxQ = 1, maxX = 1;
  for (int i = 0; i < n; ++i) {
    if (bad(sa[i]) || bad(sb[i])) return layer_fail(MPDO_EINVAL, "trace_rho_rho: bad site descriptor");
    if (i + 1 < n && (sa[i].r != sa[i + 1].l || sb[i].r != sb[i + 1].l))
      return layer_fail(MPDO_EINVAL, "trace_rho_rho: bond mismatch");
    const long long L0 = sa[i].l, L1 = sb[i].l, R0 = sa[i].r, R1 = sb[i].r, a = sa[i].a, b = sb[i].a;
    maxF = std::max({maxF, L0 * L1 * L0 * L1, R0 * R1 * R0 * R1});
    maxT = std::max({maxT, L0 * 2 * a * R0, L1 * 2 * b * R1});
    maxY = std::max(maxY, L1 * L1 * R0 * R0);
    maxQ = std::max(maxQ, L1 * R1 * L1 * R1);
    // one l1 slice: X1 (L0 L1 2a R0) + X2 (L0 a R0 b R1) + X3, X3p (2 R0 b R1 R0 each)
    maxX = std::max({maxX, L0 * L1 * 2 * a * R0 + L0 * a * R0 * b * R1 + 4 * R0 * b * R1 * R0,
                     2 * L0 * L1 * 2 * a * R0});
  }
  if (sa[0].l != 1 || sb[0].l != 1 || sa[n - 1].r != 1 || sb[n - 1].r != 1)
    return layer_fail(MPDO_EINVAL, "trace_rho_rho: open boundary bonds must be 1");
  if (maxF > (1LL << 30) || maxX > SLICE_BUDGET)
    return layer_fail(MPDO_EINVAL, "trace_rho_rho: bonds too large for the transfer sweep (chi0 * chi1 above ~ 2^15)");
  hipStream_t s = (hipStream_t)stream;
  const size_t bF = a256(maxF * sizeof(double2)), bT = a256(maxT * sizeof(double2)),
               bY = a256(maxY * sizeof(double2)), bQ = a256(maxQ * sizeof(double2)),
               bX = a256(SLICE_BUDGET * sizeof(double2));
  char* base;
  mpdo_status st = layer_ws(3 * bF + 3 * bT + bY + 2 * bQ + bX, s, &base);
  if (st != MPDO_OK) return st;
  double2* F = (double2*)base;
  double2* F2 = (double2*)(base + bF);
  double2* F3 = (double2*)(base + 2 * bF);
  double2* WA = (double2*)(base + 3 * bF);
  double2* WAp = (double2*)(base + 3 * bF + bT);
  double2* WB = (double2*)(base + 3 * bF + 2 * bT);
  double2* Y = (double2*)(base + 3 * bF + 3 * bT);
  double2* Q = (double2*)(base + 3 * bF + 3 * bT + bY);
  double2* Q2 = (double2*)(base + 3 * bF + 3 * bT + bY + bQ);
  double2* X = (double2*)(base + 3 * bF + 3 * bT + bY + 2 * bQ);
  const double2 one = make_double2(1.0, 0.0);
  TCK(hipMemcpyAsync(F, &one, sizeof one, hipMemcpyHostToDevice, s));
  for (int i = 0; i < n; ++i) {
    const int L0 = sa[i].l, L1 = sb[i].l, R0 = sa[i].r, R1 = sb[i].r, a = sa[i].a, b = sb[i].a;
    const long long nA = (long long)L0 * 2 * a * R0, nB = (long long)L1 * 2 * b * R1;
    hipLaunchKernelGGL(widen_kernel, dim3(grid_for(nA)), dim3(256), 0, s, (const float2*)sa[i].data, WA, nA);
    TCK(hipGetLastError());
    hipLaunchKernelGGL(widen_kernel, dim3(grid_for(nB)), dim3(256), 0, s, (const float2*)sb[i].data, WB, nB);
    TCK(hipGetLastError());
    const long long colsF = (long long)L1 * L0 * L1;     // (l1 m0 m1)
    if (keep && !keep[i]) {
      // rho0 and rho1 traced at this site
      const long long perL1 = (long long)L0 * L1 * 2 * a * R0;   // X1 + X1p per l1
      const int chunk = (int)std::max<long long>(
          1, std::min<long long>({(long long)L1, SLICE_BUDGET / (2 * perL1), 65535LL / L1}));
      double2* X1 = X;
      double2* X1p = X + (long long)chunk * perL1;
      for (int c0 = 0; c0 < L1; c0 += chunk) {
        const int cn = std::min(chunk, L1 - c0);
        // X1[(l1 m0 m1), (s a r0)] = sum_l0 F[l0, (l1 m0 m1)] A[l0, (s a r0)]
        GemmArgs g1 = ga(cn * L0 * L1, 2 * a * R0, L0, F + (long long)c0 * L0 * L1, 1, colsF, 0, WA, 2 * a * R0, 1, 0,
                         X1, 2 * a * R0, 1);
        TCK(gemm(C128_F64ACC, g1, s));
        // X1p[l1][m1][m0][(s a r0)]
        TCK(permute64(X1, X1p, {cn, L0, L1, 2LL * a * R0}, {0, 2, 1, 3}, s));
        // Y[l1][m1][r0][n0] = sum_{(m0 s a)} X1p[(m0 s a), r0] conj(A[(m0 s a), n0])
        GemmArgs g3 = ga(R0, R0, L0 * 2 * a, X1p, 1, R0, 0, WA, R0, 1, 1, Y + (long long)c0 * L1 * R0 * R0, R0, 1);
        g3.batch = cn * L1; g3.bsA = (long long)L0 * 2 * a * R0; g3.bsB = 0; g3.bsC = (long long)R0 * R0;
        TCK(gemm(C128_F64ACC, g3, s));
      }
      // Q[(l1 r1), (m1 n1)] = sum_{(t b)} Bp[(l1 r1), (t b)] conj(Bp[(m1 n1), (t b)]), Bp = B permuted to [l1][r1][t b]
      TCK(permute64(WB, WAp, {L1, 2LL * b, R1}, {0, 2, 1}, s));
      GemmArgs g4 = ga(L1 * R1, L1 * R1, 2 * b, WAp, 2 * b, 1, 0, WAp, 1, 2 * b, 1, Q, (long long)L1 * R1, 1);
      TCK(gemm(C128_F64ACC, g4, s));
      // Q2[l1][m1][r1][n1]
      TCK(permute64(Q, Q2, {L1, R1, L1, R1}, {0, 2, 1, 3}, s));
      // F3[(r0 n0), (r1 n1)] = sum_{(l1 m1)} Y[(l1 m1), (r0 n0)] Q2[(l1 m1), (r1 n1)]
      GemmArgs g5 = ga(R0 * R0, R1 * R1, L1 * L1, Y, 1, (long long)R0 * R0, 0, Q2, (long long)R1 * R1, 1, 0, F3,
                       (long long)R1 * R1, 1);
      TCK(gemm(C128_F64ACC, g5, s));
      // F[r0][r1][n0][n1] from F3[r0][n0][r1][n1]
      TCK(permute64(F3, F, {R0, R0, R1, R1}, {0, 2, 1, 3}, s));
      continue;
    }
    // rho0 and rho1 kept at this site: A permuted to [s][l0][a][r0] for the conj(A) step
    TCK(permute64(WA, WAp, {L0, 2, (long long)a * R0}, {1, 0, 2}, s));
    const long long x1 = (long long)L0 * L1 * 2 * a * R0, x2 = (long long)L0 * a * R0 * b * R1,
                    x3 = 2LL * R0 * b * R1 * R0;
    const long long perL1 = x1 + x2 + 2 * x3;
    const int chunk = (int)std::max<long long>(1, std::min<long long>({(long long)L1, SLICE_BUDGET / perL1, 65535LL / L0}));
    double2* X1 = X;
    double2* X2 = X1 + (long long)chunk * x1;
    double2* X3 = X2 + (long long)chunk * x2;
    double2* X3p = X3 + (long long)chunk * x3;
    const long long MF = (long long)R0 * R1 * R0;        // (r0 n1 n0)
    for (int c0 = 0; c0 < L1; c0 += chunk) {
      const int cn = std::min(chunk, L1 - c0);
      // X1[(l1 m0 m1), (s a r0)] = sum_l0 F[l0, (l1 m0 m1)] A[l0, (s a r0)]
      GemmArgs g1 = ga(cn * L0 * L1, 2 * a * R0, L0, F + (long long)c0 * L0 * L1, 1, colsF, 0, WA, 2 * a * R0, 1, 0, X1,
                       2 * a * R0, 1);
      TCK(gemm(C128_F64ACC, g1, s));
      // X2[(l1 m0)][(a r0), (b n1)] = sum_{(m1 s)} X1[(l1 m0)][(m1 s), (a r0)] conj(B[(m1 s), (b n1)])
      GemmArgs g2 = ga(a * R0, b * R1, L1 * 2, X1, 1, (long long)a * R0, 0, WB, (long long)b * R1, 1, 1, X2,
                       (long long)b * R1, 1);
      g2.batch = cn * L0; g2.bsA = (long long)L1 * 2 * a * R0; g2.bsB = 0; g2.bsC = (long long)a * R0 * b * R1;
      TCK(gemm(C128_F64ACC, g2, s));
      // X3[l1][s'][(r0 b n1), n0] = sum_{(m0 a)} X2[l1][(m0 a), (r0 b n1)] conj(Ap[s'][(m0 a), n0])
      for (int sp = 0; sp < 2; ++sp) {
        GemmArgs g3 = ga(R0 * b * R1, R0, L0 * a, X2, 1, (long long)R0 * b * R1, 0, WAp + (long long)sp * L0 * a * R0,
                         R0, 1, 1, X3 + (long long)sp * R0 * b * R1 * R0, R0, 1);
        g3.batch = cn; g3.bsA = x2; g3.bsB = 0; g3.bsC = x3;
        TCK(gemm(C128_F64ACC, g3, s));
      }
      // X3p[l1][s'][b][r0][n1][n0]
      TCK(permute64(X3, X3p, {cn, 2, R0, b, R1, R0}, {0, 1, 3, 2, 4, 5}, s));
      // F2[(r0 n1 n0), r1] (+)= sum_{(l1 s' b)} X3p[(l1 s' b), (r0 n1 n0)] B[(l1 s' b), r1]
      GemmArgs g4 = ga((int)MF, R1, cn * 2 * b, X3p, 1, MF, 0, WB + (long long)c0 * 2 * b * R1, R1, 1, 0, F2, R1, 1);
      if (c0 > 0) g4.beta_re = 1;
      TCK(gemm(C128_F64ACC, g4, s));
    }
    // F[r0][r1][n0][n1] from F2[r0][n1][n0][r1]
    TCK(permute64(F2, F, {R0, R1, R0, R1}, {0, 3, 2, 1}, s));
  }
  double2 h = make_double2(0, 0);
  TCK(hipMemcpyAsync(&h, F, sizeof(double2), hipMemcpyDeviceToHost, s));
  TCK(hipStreamSynchronize(s));
  *out = h.x;
  return MPDO_OK;
}
