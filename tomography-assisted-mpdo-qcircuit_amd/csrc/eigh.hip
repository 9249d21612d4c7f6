// Batched Hermitian eigensolver in fp64 for the Gram matrices behind every SVD of the path.
//
// n <= 64: one workgroup per matrix holds A and V in LDS (2 x 64x65 complex128 = 130 KiB of
//   the 160 KiB) and runs parallel cyclic Jacobi (round-robin pairing, n/2 disjoint complex
//   rotations per round) until a sweep performs no rotation.
// n > 64: two-sided block Jacobi on the zero-padded N x N matrix (N a multiple of 64) with
//   32-wide blocks: each round pairs the blocks round-robin, one workgroup per block pair
//   diagonalises its 64x64 principal submatrix in LDS (one inner sweep) and emits the 64x64
//   unitary W; A <- J^H A J and V <- V J are then applied tile by tile.
// Rotations skip |a_pq| <= 1e-15 sqrt(a_pp a_qq) (relative criterion; for the positive
// semidefinite Gram matrices this yields eigenvalues with small relative error).
#include "mpdo_internal.h"
#include "prof.h"
#include "site_ops.h"

#include <vector>

namespace mpdo {

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int NS = 64;       // LDS matrix order
constexpr int LD = NS + 1;   // padded LDS leading dimension
constexpr double REL_TOL = 1e-15;
// Off-diagonal entries below ABS_TOL * max_i |a_ii| of the whole matrix are left alone: the
// rotations themselves leave fp64 rounding of order n * eps * max|a_ii| behind, so a smaller
// floor would keep a rank-deficient null space churning forever. Eigenvalues are resolved to
// 1e-14 of the largest, i.e. singular values to 1e-7 of the largest - the rounding level of
// the complex64 data the Gram matrices come from. The blocked solver (n > 64) applies each
// 64x64 block rotation as 64-term fp64 sums, once per round, so its residual grows with the
// number of block columns: its floor is ABS_TOL * N / 64 (1.6e-13 at N = 1024, singular
// values still resolved to 4e-7 of the largest).
constexpr double ABS_TOL = 1e-14;

struct Rot { double cs, sn, wr, wi; int on; };

__device__ inline void pair_of(int N, int r, int i, int& p, int& q) {
  if (i == 0) { p = N - 1; q = r; }
  else { p = (r + i) % (N - 1); q = (r - i + (N - 1)) % (N - 1); }
}

// One parallel Jacobi round on the LDS matrix A (order N, even) accumulating into V: rotations
// for the round-robin pairs (one per thread, skipped below the thresholds), then one pass over
// A: the round's pairs partition the indices, so A <- J^H A J is a 2x2-block update
// B <- J_i^H B J_j of every block (pair i, pair j), each owned by one thread; V <- V J likewise
// by (row, pair). Two barriers per round; `cnt` (LDS) counts the rotations applied.
__device__ inline double2 rot_x(const Rot& R, double2 x, double2 y, bool left, bool second) {
  // right (columns, w conjugated): x' = cs x - sn conj(w) y,  y' = sn x + cs conj(w) y
  // left (rows):                   x' = cs x - sn w y,        y' = sn x + cs w y
  const double wi = left ? R.wi : -R.wi;
  const double2 wy = make_double2(R.wr * y.x - wi * y.y, R.wr * y.y + wi * y.x);
  return second ? make_double2(R.sn * x.x + R.cs * wy.x, R.sn * x.y + R.cs * wy.y)
                : make_double2(R.cs * x.x - R.sn * wy.x, R.cs * x.y - R.sn * wy.y);
}

__device__ void jacobi_round_blk(double2 (*A)[LD], double2 (*V)[LD], Rot* rot, int N, int r, double floor_abs,
                                 int* cnt) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = N / 2;
  if (tid < np) {
    int p, q;
    pair_of(N, r, tid, p, q);
    if (p > q) { int t = p; p = q; q = t; }
    double a = A[p][p].x, b = A[q][q].x;
    double2 c = A[p][q];
    double ac = sqrt(c.x * c.x + c.y * c.y);
    Rot R; R.on = 0; R.cs = 1; R.sn = 0; R.wr = 1; R.wi = 0;
    if (ac > floor_abs && ac > REL_TOL * sqrt(fabs(a * b))) {
      double zeta = (b - a) / (2.0 * ac);
      double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      R.cs = 1.0 / sqrt(1.0 + t * t);
      R.sn = t * R.cs;
      R.wr = c.x / ac; R.wi = c.y / ac;
      R.on = 1;
      atomicAdd(cnt, 1);
    }
    rot[tid] = R;
  }
  __syncthreads();
  for (int task = tid; task < np * np; task += nt) {
    const int bi = task / np, bj = task % np;
    const Rot Ri = rot[bi], Rj = rot[bj];
    if (!Ri.on && !Rj.on) continue;
    int pi, qi, pj, qj;
    pair_of(N, r, bi, pi, qi);
    if (pi > qi) { int t = pi; pi = qi; qi = t; }
    pair_of(N, r, bj, pj, qj);
    if (pj > qj) { int t = pj; pj = qj; qj = t; }
    double2 b00 = A[pi][pj], b01 = A[pi][qj], b10 = A[qi][pj], b11 = A[qi][qj];
    if (Rj.on) {   // B J_j
      double2 x0 = rot_x(Rj, b00, b01, false, false), y0 = rot_x(Rj, b00, b01, false, true);
      double2 x1 = rot_x(Rj, b10, b11, false, false), y1 = rot_x(Rj, b10, b11, false, true);
      b00 = x0; b01 = y0; b10 = x1; b11 = y1;
    }
    if (Ri.on) {   // J_i^H B
      double2 x0 = rot_x(Ri, b00, b10, true, false), y0 = rot_x(Ri, b00, b10, true, true);
      double2 x1 = rot_x(Ri, b01, b11, true, false), y1 = rot_x(Ri, b01, b11, true, true);
      b00 = x0; b10 = y0; b01 = x1; b11 = y1;
    }
    if (bi == bj) {   // the rotated pair: exactly diagonal, real diagonal
      b01 = make_double2(0, 0);
      b10 = make_double2(0, 0);
      b00.y = 0;
      b11.y = 0;
    }
    A[pi][pj] = b00; A[pi][qj] = b01; A[qi][pj] = b10; A[qi][qj] = b11;
  }
  for (int task = tid; task < np * N; task += nt) {
    const int i = task / N, k = task % N;
    const Rot R = rot[i];
    if (!R.on) continue;
    int p, q;
    pair_of(N, r, i, p, q);
    if (p > q) { int t = p; p = q; q = t; }
    const double2 x = V[k][p], y = V[k][q];
    V[k][p] = rot_x(R, x, y, false, false);
    V[k][q] = rot_x(R, x, y, false, true);
  }
  __syncthreads();
}

// The same round without V: A <- J^H A J, and the round's rotations go to the log rl[0 .. N/2)
// as (cs, sn, wr, wi) (the identity (1, 0, 1, 0) where a pair is not rotated), for replay_round.
__device__ void jacobi_round_log(double2 (*A)[LD], Rot* rot, double4* rl, int N, int r, double floor_abs, int* cnt) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = N / 2;
  if (tid < np) {
    int p, q;
    pair_of(N, r, tid, p, q);
    if (p > q) { int t = p; p = q; q = t; }
    double a = A[p][p].x, b = A[q][q].x;
    double2 c = A[p][q];
    double ac = sqrt(c.x * c.x + c.y * c.y);
    Rot R; R.on = 0; R.cs = 1; R.sn = 0; R.wr = 1; R.wi = 0;
    if (ac > floor_abs && ac > REL_TOL * sqrt(fabs(a * b))) {
      double zeta = (b - a) / (2.0 * ac);
      double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      R.cs = 1.0 / sqrt(1.0 + t * t);
      R.sn = t * R.cs;
      R.wr = c.x / ac; R.wi = c.y / ac;
      R.on = 1;
      atomicAdd(cnt, 1);
    }
    rot[tid] = R;
    rl[tid] = make_double4(R.cs, R.sn, R.wr, R.wi);
  }
  __syncthreads();
  for (int task = tid; task < np * np; task += nt) {
    const int bi = task / np, bj = task % np;
    const Rot Ri = rot[bi], Rj = rot[bj];
    if (!Ri.on && !Rj.on) continue;
    int pi, qi, pj, qj;
    pair_of(N, r, bi, pi, qi);
    if (pi > qi) { int t = pi; pi = qi; qi = t; }
    pair_of(N, r, bj, pj, qj);
    if (pj > qj) { int t = pj; pj = qj; qj = t; }
    double2 b00 = A[pi][pj], b01 = A[pi][qj], b10 = A[qi][pj], b11 = A[qi][qj];
    if (Rj.on) {   // B J_j
      double2 x0 = rot_x(Rj, b00, b01, false, false), y0 = rot_x(Rj, b00, b01, false, true);
      double2 x1 = rot_x(Rj, b10, b11, false, false), y1 = rot_x(Rj, b10, b11, false, true);
      b00 = x0; b01 = y0; b10 = x1; b11 = y1;
    }
    if (Ri.on) {   // J_i^H B
      double2 x0 = rot_x(Ri, b00, b10, true, false), y0 = rot_x(Ri, b00, b10, true, true);
      double2 x1 = rot_x(Ri, b01, b11, true, false), y1 = rot_x(Ri, b01, b11, true, true);
      b00 = x0; b10 = y0; b01 = x1; b11 = y1;
    }
    if (bi == bj) {   // the rotated pair: exactly diagonal, real diagonal
      b01 = make_double2(0, 0);
      b10 = make_double2(0, 0);
      b00.y = 0;
      b11.y = 0;
    }
    A[pi][pj] = b00; A[pi][qj] = b01; A[qi][pj] = b10; A[qi][qj] = b11;
  }
  __syncthreads();
}

// V <- V J with round r's logged rotations (the V update of jacobi_round_blk, same arithmetic in
// the same order per element, so the accumulated V is the same to the bit)
__device__ void replay_round(double2 (*V)[LD], const double4* rl, int N, int r) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = N / 2;
  for (int task = tid; task < np * N; task += nt) {
    const int i = task / N, k = task % N;
    const double4 g = rl[i];
    if (g.x == 1.0 && g.y == 0.0 && g.z == 1.0 && g.w == 0.0) continue;   // not rotated
    Rot R; R.cs = g.x; R.sn = g.y; R.wr = g.z; R.wi = g.w; R.on = 1;
    int p, q;
    pair_of(N, r, i, p, q);
    if (p > q) { int t = p; p = q; q = t; }
    const double2 x = V[k][p], y = V[k][q];
    V[k][p] = rot_x(R, x, y, false, false);
    V[k][q] = rot_x(R, x, y, false, true);
  }
  __syncthreads();
}

__device__ double max_diag(double2 (*A)[LD], int N, double* red) {
  const int tid = threadIdx.x;
  double m = 0;
  for (int i = tid; i < N; i += blockDim.x) m = fmax(m, fabs(A[i][i].x));
  red[tid] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------------------
// n <= 64: whole matrix in LDS
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void eigh_small_kernel(int n, double2* Ag, long long lda, long long bsA,
                                                         double* w, long long ldw, double2* Vg,
                                                         long long ldv, long long bsV, int max_sweeps,
                                                         int* sweeps_out) {
  __shared__ double2 A[NS][LD];
  __shared__ double2 V[NS][LD];
  __shared__ Rot rot[NS / 2];
  __shared__ double red[256];
  __shared__ int cnt;
  __shared__ int order[NS];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int N = (n + 1) & ~1;
  double2* Ab = Ag + (long long)b * bsA;
  for (int e = tid; e < N * N; e += blockDim.x) {
    int i = e / N, j = e % N;
    A[i][j] = (i < n && j < n) ? Ab[(long long)i * lda + j] : make_double2(0, 0);
    V[i][j] = make_double2(i == j ? 1.0 : 0.0, 0);
  }
  __syncthreads();
  double fl = max_diag(A, N, red) * ABS_TOL;
  int sw = 0;
  for (; sw < max_sweeps; ++sw) {
    if (tid == 0) cnt = 0;
    __syncthreads();
    for (int r = 0; r < N - 1; ++r) jacobi_round_blk(A, V, rot, N, r, fl, &cnt);
    if (cnt == 0) break;
    __syncthreads();
  }
  // sort descending (stable by index) over the n real positions: the zero padding never mixes
  // with them, and ranking it could displace slightly negative null-space eigenvalues
  for (int i = tid; i < n; i += blockDim.x) {
    double li = A[i][i].x;
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      double lj = A[j][j].x;
      rk += (lj > li) || (lj == li && j < i);
    }
    order[rk] = i;
  }
  __syncthreads();
  for (int j = tid; j < n; j += blockDim.x) w[(long long)b * ldw + j] = A[order[j]][order[j]].x;
  double2* Vb = Vg + (long long)b * bsV;
  for (int e = tid; e < n * n; e += blockDim.x) {
    int i = e / n, j = e % n;
    Vb[(long long)i * ldv + j] = V[i][order[j]];
  }
  if (tid == 0 && sweeps_out) atomicMax(sweeps_out, sw + 1);
}

// ---------------------------------------------------------------------------------------
// n > 64: block Jacobi on an N x N padded copy (N % 64 == 0), block width 32
// ---------------------------------------------------------------------------------------
// threads per bj_solve workgroup (mpdo_debug_bj_threads)
int g_bj_threads = 1024;
// rounds with at most this many subproblems solve them with W in LDS (bj_solve_w): one per CU
int g_bj_fused_max = 256;

struct BJ {
  int n, N, batch;
  double2* A;   // [batch][N][N]
  double2* V;   // [batch][N][N]
  double2* W;   // [batch][N/64][64][64]
  int* flag;    // [batch]
  int* ident;   // [batch][N/64]: the round's rotation of block pair i is the identity
  double* floor_abs;   // [batch]: ABS_TOL * max diagonal of the whole matrix
  double4* rlog;       // [batch][N/64][63][32]: the rotations of each subproblem's sweep
};

__global__ __launch_bounds__(256) void bj_floor(BJ s) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const double2* Ab = s.A + (long long)b * s.N * s.N;
  double m = 0;
  for (int i = threadIdx.x; i < s.N; i += blockDim.x) m = fmax(m, fabs(Ab[(long long)i * s.N + i].x));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) s.floor_abs[b] = red[0] * ABS_TOL * (s.N / NS);
}

__global__ void bj_init(BJ s, const double2* Ag, long long lda, long long bsA) {
  long long tot = (long long)s.N * s.N;
  int b = blockIdx.y;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot; e += (long long)gridDim.x * blockDim.x) {
    int i = (int)(e / s.N), j = (int)(e % s.N);
    double2 v = make_double2(0, 0);
    if (i < s.n && j < s.n) v = Ag[(long long)b * bsA + (long long)i * lda + j];
    s.A[(long long)b * tot + e] = v;
    s.V[(long long)b * tot + e] = make_double2(i == j ? 1.0 : 0.0, 0);
  }
}

__device__ inline void block_pair(int P, int r, int i, int& I, int& J) {
  pair_of(P, r, i, I, J);
  if (I > J) { int t = I; I = J; J = t; }
}

// one workgroup per (block pair, matrix): one sweep of the Jacobi method on the 64x64 principal
// submatrix (any multiple of 64 threads up to 1024: the round's rotations, 2x2-block updates and
// V columns are strided over the workgroup). The sweep's rotations are logged to global memory
// and then replayed on the identity in the same LDS array, which is free once A's rounds are
// done: one 64x64 complex128 array (67 KB of LDS with the rotations) lets two workgroups share a
// CU, so a round's 512 subproblems at order 1024 run in one wave instead of two.
__global__ __launch_bounds__(1024) void bj_solve(BJ s, int r) {
  __shared__ double2 A[NS][LD];   // the subproblem, then W
  __shared__ Rot rot[NS / 2];
  __shared__ int cnt, need;
  const int P = s.N / 32, pi = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  int I, J;
  block_pair(P, r, pi, I, J);
  const double2* Ab = s.A + (long long)b * s.N * s.N;
  // A keeps its lower block triangle only: the (I, J) block (I < J) is the conjugate transpose
  // of the stored (J, I) block; every 32 x 32 block is read along its stored rows (coalesced)
  for (int q = tid; q < NS * NS; q += blockDim.x) {
    const int sb = q >> 10, e = q & 1023, sr = sb >> 1, sc = sb & 1;
    const int R = sr ? J : I, C = sc ? J : I, u = e >> 5, v = e & 31;
    if (R >= C) {
      A[sr * 32 + u][sc * 32 + v] = Ab[(long long)(R * 32 + u) * s.N + C * 32 + v];
    } else {
      const double2 t = Ab[(long long)(C * 32 + u) * s.N + R * 32 + v];
      A[sr * 32 + v][sc * 32 + u] = make_double2(t.x, -t.y);
    }
  }
  if (tid == 0) cnt = 0;
  __syncthreads();
  const double fl = s.floor_abs[b];
  // no off-diagonal entry above the rotation thresholds: the sweep would rotate nothing, so the
  // round's rotation of this block pair is the identity (W is not written; its users skip it)
  if (tid == 0) need = 0;
  __syncthreads();
  for (int e = tid; e < NS * NS; e += blockDim.x) {
    const int i = e / NS, j = e % NS;
    if (j > i) {
      const double2 c = A[i][j];
      const double ac = sqrt(c.x * c.x + c.y * c.y);
      if (ac > fl && ac > REL_TOL * sqrt(fabs(A[i][i].x * A[j][j].x))) need = 1;
    }
  }
  __syncthreads();
  if (need == 0) {
    if (tid == 0) s.ident[b * (P / 2) + pi] = 1;
    return;
  }
  double4* rl = s.rlog + ((long long)b * (P / 2) + pi) * (NS - 1) * (NS / 2);
  for (int rr = 0; rr < NS - 1; ++rr) jacobi_round_log(A, rot, rl + rr * (NS / 2), NS, rr, fl, &cnt);
  // W = J_0 J_1 ... J_62 accumulated in A's storage (every round above ended with a barrier)
  for (int e = tid; e < NS * NS; e += blockDim.x) A[e / NS][e % NS] = make_double2(e / NS == e % NS ? 1.0 : 0.0, 0);
  __threadfence_block();   // the log entries written by other threads of the workgroup
  __syncthreads();
  for (int rr = 0; rr < NS - 1; ++rr) replay_round(A, rl + rr * (NS / 2), NS, rr);
  double2* Wb = s.W + ((long long)b * (P / 2) + pi) * NS * NS;
  for (int e = tid; e < NS * NS; e += blockDim.x) Wb[e] = A[e / NS][e % NS];
  if (tid == 0) {
    if (cnt > 0) atomicOr(&s.flag[b], 1);
    s.ident[b * (P / 2) + pi] = cnt == 0;
  }
}

// The same subproblem sweep with W accumulated in LDS alongside A, round by round (two 64x64
// arrays, 133 KB of LDS: one workgroup per CU), for launches with no more subproblems than CUs:
// one pass of 63 rounds instead of the logged sweep and its replay.
__global__ __launch_bounds__(1024) void bj_solve_w(BJ s, int r) {
  __shared__ double2 A[NS][LD];
  __shared__ double2 W[NS][LD];
  __shared__ Rot rot[NS / 2];
  __shared__ int cnt, need;
  const int P = s.N / 32, pi = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  int I, J;
  block_pair(P, r, pi, I, J);
  const double2* Ab = s.A + (long long)b * s.N * s.N;
  for (int q = tid; q < NS * NS; q += blockDim.x) {
    const int sb = q >> 10, e = q & 1023, sr = sb >> 1, sc = sb & 1;
    const int R = sr ? J : I, C = sc ? J : I, u = e >> 5, v = e & 31;
    if (R >= C) {
      A[sr * 32 + u][sc * 32 + v] = Ab[(long long)(R * 32 + u) * s.N + C * 32 + v];
    } else {
      const double2 t = Ab[(long long)(C * 32 + u) * s.N + R * 32 + v];
      A[sr * 32 + v][sc * 32 + u] = make_double2(t.x, -t.y);
    }
  }
  if (tid == 0) {
    cnt = 0;
    need = 0;
  }
  __syncthreads();
  const double fl = s.floor_abs[b];
  for (int e = tid; e < NS * NS; e += blockDim.x) {
    const int i = e / NS, j = e % NS;
    if (j > i) {
      const double2 c = A[i][j];
      const double ac = sqrt(c.x * c.x + c.y * c.y);
      if (ac > fl && ac > REL_TOL * sqrt(fabs(A[i][i].x * A[j][j].x))) need = 1;
    }
  }
  __syncthreads();
  if (need == 0) {
    if (tid == 0) s.ident[b * (P / 2) + pi] = 1;
    return;
  }
  for (int e = tid; e < NS * NS; e += blockDim.x) W[e / NS][e % NS] = make_double2(e / NS == e % NS ? 1.0 : 0.0, 0);
  __syncthreads();
  for (int rr = 0; rr < NS - 1; ++rr) jacobi_round_blk(A, W, rot, NS, rr, fl, &cnt);
  double2* Wb = s.W + ((long long)b * (P / 2) + pi) * NS * NS;
  for (int e = tid; e < NS * NS; e += blockDim.x) Wb[e] = W[e / NS][e % NS];
  if (tid == 0) {
    if (cnt > 0) atomicOr(&s.flag[b], 1);
    s.ident[b * (P / 2) + pi] = cnt == 0;
  }
}

// Two-sided update of A by the round's block rotations, one 64x64 block pair per workgroup:
// A[R, C] <- W_R^H A[R, C] W_C for the block pairs pr >= pc, R = I_pr u J_pr, C = I_pc u J_pc
// (two 32-wide blocks each). A is Hermitian and only its lower block triangle is kept (32 x 32
// blocks (a, b) with a >= b), so every stored block is read and written once per round.
__device__ inline void mfma_tile64(const double2 (*X)[LD], const double2* Wg, bool w_left, doublex4 (&re)[2][2],
                                   doublex4 (&im)[2][2], int r0, int c0, int li, int lk) {
  // w_left: out = W^H X (P[i][k] = conj(W[k][i])); else out = X W (Q[k][j] = W[k][j])
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      re[a][c] = doublex4{0, 0, 0, 0};
      im[a][c] = doublex4{0, 0, 0, 0};
    }
  for (int kk = 0; kk < NS; kk += 4) {
    const int k = kk + lk;
    double pr[2], pi_[2], pn[2], qr[2], qi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = r0 + 16 * a + li;
      double2 v;
      if (w_left) { v = Wg[k * NS + row]; v.y = -v.y; }
      else v = X[row][k];
      pr[a] = v.x; pi_[a] = v.y; pn[a] = -v.y;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = c0 + 16 * c + li;
      double2 v = w_left ? X[k][col] : Wg[k * NS + col];
      qr[c] = v.x; qi[c] = v.y;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        re[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[a], qr[c], re[a][c], 0, 0, 0);
        re[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pn[a], qi[c], re[a][c], 0, 0, 0);
        im[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[a], qi[c], im[a][c], 0, 0, 0);
        im[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi_[a], qr[c], im[a][c], 0, 0, 0);
      }
  }
}

__device__ __forceinline__ void sym_body(BJ s, int r, int e, int b, double2 (*X)[LD]) {
  const int P = s.N / 32, h = P / 2, tid = threadIdx.x;
  // lower-triangle block pair (pr >= pc) from the linear index
  int pr = 0;
  while (e >= pr + 1) { e -= pr + 1; ++pr; }
  const int pc = e;
  (void)h;
  int Ir, Jr, Ic, Jc;
  block_pair(P, r, pr, Ir, Jr);
  block_pair(P, r, pc, Ic, Jc);
  const bool idR = s.ident[b * (P / 2) + pr] != 0, idC = s.ident[b * (P / 2) + pc] != 0;
  if (idR && idC) return;   // W_R^H A W_C = A (uniform over the workgroup)
  double2* A = s.A + (long long)b * s.N * s.N;
  const double2* Wr = s.W + ((long long)b * (P / 2) + pr) * NS * NS;
  const double2* Wc = s.W + ((long long)b * (P / 2) + pc) * NS * NS;
  // the four 32 x 32 sub-blocks (row block Rb, column block Cb) of the tile, each held in its
  // stored orientation: directly when Rb >= Cb, else as the conjugate transpose of (Cb, Rb)
  for (int q = tid; q < NS * NS; q += blockDim.x) {
    const int sb = q >> 10, e = q & 1023, sr = sb >> 1, sc = sb & 1;
    const int Rb = sr ? Jr : Ir, Cb = sc ? Jc : Ic, u = e >> 5, v = e & 31;
    if (Rb >= Cb) {
      X[sr * 32 + u][sc * 32 + v] = A[(long long)(Rb * 32 + u) * s.N + Cb * 32 + v];
    } else {
      const double2 t = A[(long long)(Cb * 32 + u) * s.N + Rb * 32 + v];
      X[sr * 32 + v][sc * 32 + u] = make_double2(t.x, -t.y);
    }
  }
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64, li = lane % 16, lk = lane / 16;
  const int r0 = (wave / 2) * 32, c0 = (wave % 2) * 32;
  doublex4 re[2][2], im[2][2];
  if (!idR) {
    mfma_tile64(X, Wr, true, re, im, r0, c0, li, lk);     // T = W_R^H X
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          X[r0 + 16 * a + 4 * q + lk][c0 + 16 * c + li] = make_double2(re[a][c][q], im[a][c][q]);
    __syncthreads();
  }
  if (!idC) {
    mfma_tile64(X, Wc, false, re, im, r0, c0, li, lk);    // T W_C
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double2 v = X[r0 + 16 * a + 4 * q + lk][c0 + 16 * c + li];
          re[a][c][q] = v.x;
          im[a][c][q] = v.y;
        }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        X[r0 + 16 * a + 4 * q + lk][c0 + 16 * c + li] = make_double2(re[a][c][q], im[a][c][q]);
  __syncthreads();
  // back in the stored orientations; on the diagonal tile (pr == pc) the upper sub-block's
  // storage is the lower sub-block of the same tile, written directly
  for (int q = tid; q < NS * NS; q += blockDim.x) {
    const int sb = q >> 10, e = q & 1023, sr = sb >> 1, sc = sb & 1;
    const int Rb = sr ? Jr : Ir, Cb = sc ? Jc : Ic, u = e >> 5, v = e & 31;
    if (Rb >= Cb) {
      A[(long long)(Rb * 32 + u) * s.N + Cb * 32 + v] = X[sr * 32 + u][sc * 32 + v];
    } else if (pr != pc) {
      const double2 t = X[sr * 32 + v][sc * 32 + u];
      A[(long long)(Cb * 32 + u) * s.N + Rb * 32 + v] = make_double2(t.x, -t.y);
    }
  }
}

// A's two-sided update by the round's rotations: workgroup x updates block pair x (sym_body);
// blockIdx.y = matrix. V's update (independent of it) is bj_rotate_v_direct.
__global__ __launch_bounds__(256) void bj_update(BJ s, int r) {
  __shared__ double2 X[NS][LD];
  sym_body(s, r, blockIdx.x, blockIdx.y, X);
}

// V <- V W on the columns of the round's block pairs without LDS: each wave reads its rows of the
// 64 x 64 tile and the pair's rotation straight into the matrix-core operands (no LDS, so three
// workgroups per CU), and every wave has read the whole tile before any wave writes it back.
__global__ __launch_bounds__(256, 3) void bj_rotate_v_direct(BJ s, int r, int ntile) {
  const int P = s.N / 32, t = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int tile = t % ntile, pi = t / ntile, t0 = tile * NS;
  if (s.ident[b * (P / 2) + pi]) return;   // V W = V (uniform over the workgroup)
  int I, J;
  block_pair(P, r, pi, I, J);
  double2* M = s.V + (long long)b * s.N * s.N;
  const double2* Wb = s.W + ((long long)b * (P / 2) + pi) * NS * NS;
  const int wave = tid / 64, lane = tid % 64, li = lane % 16, lk = lane / 16;
  const int r0 = (wave / 2) * 32, c0 = (wave % 2) * 32;
  doublex4 re[2][2], im[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      re[a][c] = doublex4{0, 0, 0, 0};
      im[a][c] = doublex4{0, 0, 0, 0};
    }
#pragma unroll 4
  for (int kk = 0; kk < NS; kk += 4) {
    const int k = kk + lk;
    const int gk = k < 32 ? I * 32 + k : J * 32 + k - 32;
    double pr[2], pim[2], pn[2], qr[2], qi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double2 v = M[(long long)(t0 + r0 + 16 * a + li) * s.N + gk];
      pr[a] = v.x; pim[a] = v.y; pn[a] = -v.y;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double2 v = Wb[k * NS + c0 + 16 * c + li];
      qr[c] = v.x; qi[c] = v.y;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        re[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[a], qr[c], re[a][c], 0, 0, 0);
        re[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pn[a], qi[c], re[a][c], 0, 0, 0);
        im[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr[a], qi[c], im[a][c], 0, 0, 0);
        im[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(pim[a], qr[c], im[a][c], 0, 0, 0);
      }
  }
  __syncthreads();   // the tile is rewritten in place: every read above precedes every write below
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = r0 + 16 * a + 4 * q + lk, j = c0 + 16 * c + li;
        const int gj = j < 32 ? I * 32 + j : J * 32 + j - 32;
        M[(long long)(t0 + i) * s.N + gj] = make_double2(re[a][c][q], im[a][c][q]);
      }
}

// eigenvalues = diag(A), sorted descending (stable); the matching column order of V goes to
// order_ws for bj_gather_v
__global__ __launch_bounds__(256) void bj_finish(BJ s, double* w, long long ldw, double2* Vg, long long ldv,
                                                 long long bsV, int* order_ws) {
  const int b = blockIdx.x, tid = threadIdx.x;
  const double2* Ab = s.A + (long long)b * s.N * s.N;
  const double2* Vb = s.V + (long long)b * s.N * s.N;
  int* order = order_ws + (long long)b * s.N;
  for (int i = tid; i < s.n; i += blockDim.x) {
    double li = Ab[(long long)i * s.N + i].x;
    int rk = 0;
    for (int j = 0; j < s.n; ++j) {
      double lj = Ab[(long long)j * s.N + j].x;
      rk += (lj > li) || (lj == li && j < i);
    }
    order[rk] = i;
  }
  __syncthreads();
  for (int j = tid; j < s.n; j += blockDim.x) {
    int o = order[j];
    w[(long long)b * ldw + j] = Ab[(long long)o * s.N + o].x;
  }
  (void)Vb; (void)Vg; (void)ldv; (void)bsV;
}

// eigenvectors = V's columns in the order bj_finish found (grid over the elements of each matrix)
__global__ void bj_gather_v(BJ s, double2* Vg, long long ldv, long long bsV, const int* order_ws) {
  const int b = blockIdx.y;
  const double2* Vb = s.V + (long long)b * s.N * s.N;
  const int* order = order_ws + (long long)b * s.N;
  double2* Vo = Vg + (long long)b * bsV;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)s.n * s.n;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / s.n), j = (int)(e % s.n);
    Vo[(long long)i * ldv + j] = Vb[(long long)i * s.N + order[j]];
  }
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

size_t eigh_workspace_bytes(int n, int batch) {
  if (n <= NS) return 256;
  size_t N = (size_t)((n + 63) / 64) * 64;
  size_t mats = align256(N * N * sizeof(double2) * batch);
  size_t w = align256((N / 64) * NS * NS * sizeof(double2) * batch);
  size_t rlog = align256((N / 64) * (NS - 1) * (NS / 2) * sizeof(double4) * batch);
  return 2 * mats + w + rlog + align256(sizeof(int) * batch) + align256(sizeof(int) * (N / 64) * batch) +
         align256(sizeof(double) * batch) + align256(sizeof(int) * N * batch) + 256;
}

hipError_t eigh(int n, int batch, void* A, long long lda, long long bsA, double* w, long long ldw, void* V,
                long long ldv, long long bsV, void* ws, hipStream_t st, int* sweeps_out) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  const int max_sweeps = 60;
  hipError_t err;
  if (n <= NS) {
    int* sw_dev = (int*)ws;
    if ((err = hipMemsetAsync(sw_dev, 0, sizeof(int), st)) != hipSuccess) return err;
    {
      prof::Scope ps(prof::EIG_SMALL, st, 0.0, 2.0 * 16.0 * n * n * batch);
      hipLaunchKernelGGL(eigh_small_kernel, dim3(batch), dim3(256), 0, st, n, (double2*)A, lda, bsA, w, ldw,
                         (double2*)V, ldv, bsV, max_sweeps, sw_dev);
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if (sweeps_out) {
      if ((err = hipMemcpyAsync(sweeps_out, sw_dev, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return err;
      if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
    }
    return hipSuccess;
  }
  BJ s;
  s.n = n;
  s.N = ((n + 63) / 64) * 64;
  s.batch = batch;
  char* p = (char*)ws;
  size_t mats = align256((size_t)s.N * s.N * sizeof(double2) * batch);
  s.A = (double2*)p; p += mats;
  s.V = (double2*)p; p += mats;
  s.W = (double2*)p; p += align256((size_t)(s.N / 64) * NS * NS * sizeof(double2) * batch);
  s.rlog = (double4*)p; p += align256((size_t)(s.N / 64) * (NS - 1) * (NS / 2) * sizeof(double4) * batch);
  s.flag = (int*)p; p += align256(sizeof(int) * batch);
  s.ident = (int*)p; p += align256(sizeof(int) * (s.N / 64) * batch);
  s.floor_abs = (double*)p; p += align256(sizeof(double) * batch);
  int* order = (int*)p;
  const int P = s.N / 32;
  hipLaunchKernelGGL(bj_init, dim3(256, batch), dim3(256), 0, st, s, (const double2*)A, lda, bsA);
  hipLaunchKernelGGL(bj_floor, dim3(batch), dim3(256), 0, st, s);
  std::vector<int> hflag(batch);
  int sw = 0;
  for (; sw < max_sweeps; ++sw) {
    if ((err = hipMemsetAsync(s.flag, 0, sizeof(int) * batch, st)) != hipSuccess) return err;
    // algorithmic work: one inner sweep of a 64x64 subproblem applies 63 x 32 rotations, each
    // updating 192 element pairs at 18 flops; a rotation tile is a 64x64x64 complex product
    const double solve_flops = (double)(P / 2) * batch * 63.0 * 32.0 * 192.0 * 18.0;
    const double solve_bytes = (double)(P / 2) * batch * 2.0 * NS * NS * 16.0;
    const double tile_flops = 8.0 * NS * NS * NS, tile_bytes = 3.0 * NS * NS * 16.0;
    const int npairs_sym = (P / 2) * (P / 2 + 1) / 2;
    const double tiles_sym = (double)npairs_sym * batch, tiles_r = (double)(s.N / NS) * (P / 2) * batch;
    const int ntile = s.N / NS;
    for (int r = 0; r < P - 1; ++r) {
      {
        prof::Scope ps(prof::EIG_SOLVE, st, solve_flops, solve_bytes);
        if ((long long)(P / 2) * batch <= g_bj_fused_max)
          hipLaunchKernelGGL(bj_solve_w, dim3(P / 2, batch), dim3(1024), 0, st, s, r);
        else
          hipLaunchKernelGGL(bj_solve, dim3(P / 2, batch), dim3(g_bj_threads), 0, st, s, r);
      }
      {   // A <- J^H A J block pair by block pair (Hermitian: lower triangle + mirror)
        prof::Scope ps(prof::EIG_UPDATE, st, tiles_sym * 2 * tile_flops, tiles_sym * tile_bytes);
        hipLaunchKernelGGL(bj_update, dim3(npairs_sym, batch), dim3(256), 0, st, s, r);
      }
      {   // V <- V J on the columns of the round's block pairs
        prof::Scope ps(prof::EIG_ROTATE, st, tiles_r * tile_flops, tiles_r * tile_bytes);
        hipLaunchKernelGGL(bj_rotate_v_direct, dim3(ntile * (P / 2), batch), dim3(256), 0, st, s, r, ntile);
      }
      if (prof::enabled()) {
        // the work the three launches executed: only subproblems with a rotation run their sweep,
        // and only the tiles of rotated block pairs are multiplied (one readback per round, timing
        // runs only)
        std::vector<int> id((size_t)(P / 2) * batch);
        if ((err = hipMemcpyAsync(id.data(), s.ident, sizeof(int) * id.size(), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
          return err;
        if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
        double nsolve = 0, nsym = 0, nsym_tiles = 0, nv = 0;
        for (int b = 0; b < batch; ++b) {
          const int* idb = id.data() + (size_t)b * (P / 2);
          for (int i = 0; i < P / 2; ++i) nsolve += idb[i] == 0;
          for (int pr = 0; pr < P / 2; ++pr)
            for (int pc = 0; pc <= pr; ++pc) {
              const int k = (idb[pr] == 0) + (idb[pc] == 0);
              if (k) { nsym += 1; nsym_tiles += k; }
            }
        }
        nv = nsolve * ntile;
        const double per_solve = solve_flops / ((double)(P / 2) * batch);
        prof::add_executed(prof::EIG_SOLVE, nsolve * per_solve, nsolve * 2.0 * NS * NS * 16.0);
        prof::add_executed(prof::EIG_UPDATE, nsym_tiles * tile_flops, (2.0 * nsym + nsym_tiles) * NS * NS * 16.0);
        prof::add_executed(prof::EIG_ROTATE, nv * tile_flops, nv * tile_bytes);
      }
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if ((err = hipMemcpyAsync(hflag.data(), s.flag, sizeof(int) * batch, hipMemcpyDeviceToHost, st)) != hipSuccess)
      return err;
    if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
    int any = 0;
    for (int b = 0; b < batch; ++b) any |= hflag[b];
    if (!any) break;
  }
  hipLaunchKernelGGL(bj_finish, dim3(batch), dim3(256), 0, st, s, w, ldw, (double2*)V, ldv, bsV, order);
  {
    const long long nn = (long long)n * n;
    const int blocks = (int)std::min<long long>((nn + 255) / 256, 1024);
    hipLaunchKernelGGL(bj_gather_v, dim3(blocks, batch), dim3(256), 0, st, s, (double2*)V, ldv, bsV, order);
  }
  if (sweeps_out) *sweeps_out = sw + 1;
  prof::eig_log(n, batch, sw + 1);
  return hipGetLastError();
}

}  // namespace mpdo

extern "C" int mpdo_debug_bj_fused_max(int v) {
  const int old = mpdo::g_bj_fused_max;
  if (v >= 0) mpdo::g_bj_fused_max = v;
  return old;
}

extern "C" int mpdo_debug_bj_threads(int v) {
  const int old = mpdo::g_bj_threads;
  if (v == 256 || v == 512 || v == 1024) mpdo::g_bj_threads = v;
  return old;
}

// =======================================================================================
// Eigen-decomposition of positive semidefinite Gram matrices, preconditioned by pivoted
// Cholesky: G = P L L^H P^T stopped at the numerical rank r (largest remaining diagonal at
// or below 1e-14 of the largest diagonal), the Jacobi method on the r x r matrix M = L^H L (same nonzero
// eigenvalues, strongly diagonally dominant for graded spectra, so it converges in a few sweeps),
// eigenvectors L U Lambda^{-1/2}. Directions beyond the rank (and eigenvalues below the floor)
// come back as zero vectors with zero eigenvalues.
// =======================================================================================
namespace mpdo {
namespace {

constexpr int PC_MAX = 1024;
// w[b][j] <- w_M[b][j] for j < r, 0 beyond
__global__ void pad_eigs_kernel(const double* wm, int r, double* w, long long ldw, int n) {
  const int b = blockIdx.x;
  for (int j = threadIdx.x; j < n; j += blockDim.x) w[(long long)b * ldw + j] = j < r ? wm[(long long)b * r + j] : 0.0;
}

// shift[b] = D*_b + margin: D*_b = tr G - (sqrt(tr G) - eps)^2 is the largest eigenvalue the
// absolute-error cut may drop (te[-1] - sqrt(te[-1]^2 - lambda) <= eps), margin = 1e-9 D* +
// 1e-13 tr G keeps rounding on the side of running the eigensolver
__global__ void gate_shift_kernel(const double2* G, long long lda, long long bsA, int n, double eps, double* shift) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const double2* Gb = G + (long long)b * bsA;
  double t = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t += Gb[(long long)i * lda + i].x;
  red[threadIdx.x] = t;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tr = red[0];
    const double rt = sqrt(fmax(tr, 0.0));
    const double dstar = rt > eps ? tr - (rt - eps) * (rt - eps) : tr;
    shift[b] = dstar * (1.0 + 1e-9) + 1e-13 * tr;
  }
}

// Gs[b] = G[b] - shift[b] I (n x n, contiguous)
__global__ void shift_gram_kernel(const double2* G, long long lda, long long bsA, int n, const double* shift,
                                  double2* Gs) {
  const int b = blockIdx.y;
  const double2* Gb = G + (long long)b * bsA;
  double2* Sb = Gs + (long long)b * n * n;
  const double sh = shift[b];
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)n * n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e / n, j = e - i * n;
    double2 v = Gb[i * lda + j];
    if (i == j) v.x -= sh;
    Sb[e] = v;
  }
}

}  // namespace

hipError_t gram_keeps_all(int n, int batch, const void* A, long long lda, long long bsA, double eps, void* ws,
                          hipStream_t st, bool* all) {
  *all = false;
  if (n <= NS || n > PC_MAX || batch <= 0) return hipSuccess;
  hipError_t err;
  char* p = (char*)ws;
  const size_t mat = align256((size_t)n * n * sizeof(double2) * batch);
  p += mat;                                     // (eigh_gram's L)
  double2* Gs = (double2*)p; p += mat;          // (M)
  double2* L = (double2*)p; p += mat;           // (U)
  void* cws = p; p += mat;                      // (Y): the Cholesky's state, n > 64 so it fits
  double* shift = (double*)p; p += align256((size_t)n * sizeof(double) * batch);
  int* piv = (int*)p; p += align256((size_t)n * sizeof(int) * batch);
  int* rank = (int*)p; p += align256(sizeof(int) * batch);
  hipLaunchKernelGGL(gate_shift_kernel, dim3(batch), dim3(256), 0, st, (const double2*)A, lda, bsA, n, eps, shift);
  {
    const long long nn = (long long)n * n;
    const int blocks = (int)std::min<long long>((nn + 255) / 256, 2048);
    hipLaunchKernelGGL(shift_gram_kernel, dim3(blocks, batch), dim3(256), 0, st, (const double2*)A, lda, bsA, n,
                       (const double*)shift, Gs);
  }
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // positive definite <=> the (unpivoted) Cholesky completes all n columns with pivots above 1e-14
  // of the largest diagonal
  if ((err = chol_b(n, batch, Gs, n, (long long)n * n, 1e-14, L, piv, rank, cws, st)) != hipSuccess) return err;
  std::vector<int> hr(batch);
  if ((err = hipMemcpyAsync(hr.data(), rank, sizeof(int) * batch, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return err;
  if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
  bool ok = true;
  for (int b = 0; b < batch; ++b) ok = ok && hr[b] == n;
  *all = ok;
  return hipSuccess;
}

size_t eigh_gram_workspace_bytes(int n, int batch) {
  const size_t mat = align256((size_t)n * n * sizeof(double2) * batch);
  const size_t vec = align256((size_t)n * sizeof(double) * batch);
  return 4 * mat + vec + align256((size_t)n * sizeof(int) * batch) + align256(sizeof(int) * batch) +
         pchol_workspace_bytes(n, batch) + eigh_workspace_bytes(n, batch) + 256;
}

hipError_t eigh_gram(int n, int batch, void* A, long long lda, long long bsA, double* w, long long ldw, void* V,
                     long long ldv, long long bsV, void* ws, hipStream_t st, int* sweeps_out, double tol_rel) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  if (n <= NS || n > PC_MAX) return eigh(n, batch, A, lda, bsA, w, ldw, V, ldv, bsV, ws, st, sweeps_out);
  hipError_t err;
  char* p = (char*)ws;
  const size_t mat = align256((size_t)n * n * sizeof(double2) * batch);
  double2* L = (double2*)p; p += mat;
  double2* M = (double2*)p; p += mat;
  double2* U = (double2*)p; p += mat;
  double2* Y = (double2*)p; p += mat;
  double* wm = (double*)p; p += align256((size_t)n * sizeof(double) * batch);
  int* piv = (int*)p; p += align256((size_t)n * sizeof(int) * batch);
  int* rank = (int*)p; p += align256(sizeof(int) * batch);
  void* pws = p; p += pchol_workspace_bytes(n, batch);
  void* ws_inner = p;
  // eigenvectors are zeroed for eigenvalues of L^H L below 1e-14 of the largest (the Jacobi floor)
  const double tol = 1e-14;
  if ((err = pchol_b(n, batch, A, lda, bsA, tol_rel, L, piv, rank, pws, st)) != hipSuccess) return err;
  std::vector<int> hr(batch);
  if ((err = hipMemcpyAsync(hr.data(), rank, sizeof(int) * batch, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return err;
  if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
  int r = 0;
  for (int b = 0; b < batch; ++b) r = hr[b] > r ? hr[b] : r;
  if (r == 0) r = 1;   // all-zero input: one zero column, zero eigenvalue
  // M = L[:, :r]^H L[:, :r]  (r x r); columns of L past a matrix's own rank are zero
  {
    GemmArgs g{};
    g.M = r; g.N = r; g.K = n; g.batch = batch;
    g.A = L; g.rsA = n; g.csA = 1; g.bsA = (long long)n * n; g.conjA = 1;
    g.B = L; g.rsB = 1; g.csB = n; g.bsB = (long long)n * n;
    g.C = M; g.rsC = r; g.csC = 1; g.bsC = (long long)r * r;
    g.alpha_re = 1;
    g.herm = 1;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  int sw = 0;
  if ((err = eigh(r, batch, M, r, (long long)r * r, wm, r, U, r, (long long)r * r, ws_inner, st, &sw)) != hipSuccess)
    return err;
  if (sweeps_out) *sweeps_out = sw;
  // Y = L[:, :r] U  (n x r), then V = Y Lambda^{-1/2} (zero past r and below the floor)
  {
    GemmArgs g{};
    g.M = n; g.N = r; g.K = r; g.batch = batch;
    g.A = L; g.rsA = 1; g.csA = n; g.bsA = (long long)n * n;
    g.B = U; g.rsB = r; g.csB = 1; g.bsB = (long long)r * r;
    g.C = Y; g.rsC = r; g.csC = 1; g.bsC = (long long)n * r;
    g.alpha_re = 1;
    if ((err = gemm(C128_F64ACC, g, st)) != hipSuccess) return err;
  }
  if ((err = factor64_b(Y, r, (long long)n * r, wm, r, n, r, n, -0.5, tol, 0, (double2*)V, ldv, bsV, batch, st)) !=
      hipSuccess)
    return err;
  hipLaunchKernelGGL(pad_eigs_kernel, dim3(batch), dim3(256), 0, st, wm, r, w, ldw, n);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  return pchol_tail_eigs(n, batch, pws, rank, w, ldw, st);
}

}  // namespace mpdo
