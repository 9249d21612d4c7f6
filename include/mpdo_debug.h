/*
 * Kernel-level test hooks of libmpdo_gfx950. They have no counterpart in the reference: they
 * expose the building blocks behind include/mpdo.h (the strided complex GEMM on the matrix
 * cores and the batched fp64 Jacobi eigensolver) so that tests can check them in isolation.
 */
#ifndef MPDO_DEBUG_H
#define MPDO_DEBUG_H

#include "mpdo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b]; strides in complex elements.
 * prec: 0 = complex64 in / fp32 accumulate / complex64 out, 1 = complex64 in / fp64 / complex128 out,
 *       2 = complex128 in / fp64 / complex128 out, 3 = complex128 in / fp64 / complex64 out,
 *       4 = complex64 in / fp64 / complex64 out, 5 = complex64 in / fp32 accumulation of three real
 *       products (3M) / complex64 out, 6 and 7 = complex64 in / bf16 matrix cores on operands split
 *       into 2 (6) or 3 (7) bf16 parts, fp32 accumulation / complex64 out;
 * prec + 10: the same with a Hermitian output (M == N; the tiles on and below the diagonal are
 *       computed and the strict upper triangle set to the conjugate transpose of the lower). */
mpdo_status mpdo_debug_gemm(int prec, int M, int N, int K, int batch,
                            const void* A, long long rsA, long long csA, long long bsA, int conjA,
                            const void* B, long long rsB, long long csB, long long bsB, int conjB,
                            void* C, long long rsC, long long csC, long long bsC,
                            double alpha_re, double alpha_im, double beta_re, double beta_im, void* stream);

/* G[b] = X[b] X[b]^H (X: n x K complex64 row-major, G: n x n complex128) through the
 * split-reduction path: `slices` partial products of K / slices each in one launch (only the
 * tiles on and below the diagonal), written to parts (slices n^2 complex128 per batch entry),
 * summed in order and mirrored. */
mpdo_status mpdo_debug_gram_split(int n, int K, int batch, int slices, const void* X, void* parts, void* G,
                                  void* stream);

/* mpdo_debug_gemm with each setting explicit: herm = 1 for a Hermitian output (as prec + 10
 * above); inA / inB the storage of each operand (0 = as prec says, 1 = complex64, 2 = complex128;
 * a complex128 operand needs an fp64-accumulating prec); tri = 1 when op(B) is upper triangular,
 * 2 when op(A) is lower triangular (each output tile reduces only over its nonzero k range). */
mpdo_status mpdo_debug_gemm_ex(int prec, int herm, int inA, int inB, int tri, int M, int N, int K, int batch,
                               const void* A, long long rsA, long long csA, long long bsA, int conjA,
                               const void* B, long long rsB, long long csB, long long bsB, int conjB,
                               void* C, long long rsC, long long csC, long long bsC,
                               double alpha_re, double alpha_im, double beta_re, double beta_im, void* stream);

/* Eigen-decomposition of `batch` Hermitian complex128 n x n matrices (row-major, contiguous,
 * destroyed): w[b][n] descending, V[b][n][n] with eigenvectors in columns. *sweeps = sweeps used. */
mpdo_status mpdo_debug_eigh(int n, int batch, void* A, double* w, void* V, int* sweeps, void* stream);

/* The same for positive semidefinite (Gram) matrices through the pivoted-Cholesky-preconditioned
 * solver the fused layer uses: eigenpairs below the solver's floor (about 1e-14 x n/64 of the
 * largest eigenvalue) or past the numerical rank come back as zero eigenvalues with zero vectors.
 * *sweeps = Jacobi sweeps spent on the rank-sized problem. */
mpdo_status mpdo_debug_eigh_gram(int n, int batch, void* A, double* w, void* V, int* sweeps, void* stream);

/* The unpivoted blocked Cholesky the partial eigensolvers' certificates and Cholesky QR use, on
 * `batch` Hermitian complex128 n x n matrices (row-major, contiguous, the lower triangle read;
 * destroyed): L (n x n complex128, stored by columns: L[k n + i] = L_ik) with A = L L^H, stopping at
 * the first pivot at or below tol_rel x max_i A_ii (later columns zero); rank[b] = the columns
 * produced. */
mpdo_status mpdo_debug_chol(int n, int batch, void* A, double tol_rel, void* L, int* rank, void* stream);

/* Host-only rehearsal of mpdo_layer_batched: plans the layer and walks every launch with fake
 * device addresses, checking each operand's extent against the buffer it falls in, without
 * touching a GPU. `in[i].data`/`out[i].data` may be any distinct non-overlapping addresses.
 * *workspace_bytes receives the device workspace the layer needs. */
mpdo_status mpdo_debug_layer_dry_run(int batch, int n, const mpdo_site* in, int K1, int parity, int K2, int chi,
                                     int kappa, const mpdo_site* out, unsigned long long* workspace_bytes);

/* Tile shape of the fp64-accumulating GEMMs for tuning sweeps: 0 = 64x64x16 (default),
 * 1 = 64x64x32, 2 = 128x64x16, 3 = 64x128x16, 4 = 128x128x16, 5 = 64x64x16 with two LDS
 * stages; returns the previous choice
 * (an out-of-range v only reads it). */
int mpdo_debug_gemm_tiles(int v);
/* Tile shape of the split-bf16 GEMMs (precisions 6 and 7) for tuning sweeps: 0 = the built-in
 * rules (default), 1 = 64x64, 2 = 128x64, 3 = 64x128, 4 = 32x128, 5 = 128x32; a Hermitian output
 * always takes 64x64. Returns the previous choice (an out-of-range v only reads it). */
int mpdo_debug_gemm_bf16_tiles(int v);

/* Threads per workgroup of the block-Jacobi subproblem kernel (256, 512 or 1024); returns the
 * previous choice (any other v only reads it). */
int mpdo_debug_bj_threads(int v);
/* Block-Jacobi rounds with at most v subproblems (block pairs x matrices) solve them with the
 * rotation product accumulated in LDS next to the subproblem (one workgroup per CU, one pass);
 * larger rounds log the rotations and replay them (two workgroups per CU). Default 256; returns
 * the previous value (v < 0 only reads it). */
int mpdo_debug_bj_fused_max(int v);

/* The GEMMs the last dry run of the fused layer would launch, in order: out [6 n] = source line,
 * M, N, K, batch, precision mode; returns n (fills at most cap ints). */
int mpdo_debug_layer_gemm_log(int* out, int cap);

/* Per pair of the last mpdo_layer_batched call on this thread: the largest number of singular
 * values its gate-time split kept in any chain (2 p1t when none was dropped). Writes up to `cap`
 * values, returns the number of pairs. */
int mpdo_debug_layer_kept(int* out, int cap);

/* Recording of the spectra of chain 0 in the fused layers that follow on this thread (record = 1
 * on, 0 off: each gate-time cut and chi step then costs one extra readback, and every gate-time cut
 * runs its eigensolver). gate_eigs copies the last layer's gate-time cut spectra into out, per pair
 * its order n followed by its n eigenvalues (descending), and returns the number of doubles;
 * chi_eigs does the same for the chi steps at the pair bonds. */
int mpdo_debug_layer_gate_eigs(int record, double* out, int cap);
int mpdo_debug_layer_chi_eigs(double* out, int cap);

/* The fused layer's gate-time cut decides "every direction kept in every chain" by one Cholesky
 * factorisation of G - D* I (D* the largest eigenvalue the absolute-error rule may drop) and then needs
 * no eigensolver; on = 0 always runs the eigensolver, on < 0 leaves the switch. Returns the
 * previous setting. mpdo_debug_layer_gate_skips: how many pairs of the last fused layer on this
 * thread were decided that way. */
int mpdo_debug_layer_gate_skip(int on);
/* Where the shifted Cholesky does not settle the cut, the fused layer first decides it from the
 * bottom of the spectrum (subspace iteration with the inverse Cholesky factor, Rayleigh-Ritz and a
 * positive-definiteness certificate; the full eigensolver only where that declines); on = 0 always
 * runs the eigensolver, on < 0 leaves the switch. Returns the previous setting.
 * mpdo_debug_layer_gate_tails: how many pairs of the last fused layer on this thread were decided
 * that way. */
int mpdo_debug_layer_gate_tail(int on);
int mpdo_debug_layer_gate_tails(void);
/* Subspace of the gate-time cut from the bottom of the spectrum: m columns (16..256, default 128;
 * other values leave it) and q >= 1 inverse-iteration steps (default 5). Returns the previous m. */
int mpdo_debug_layer_tail_shape(int m, int q);
/* The chi step at a pair bond keeps chi of about 4 chi directions: the fused layer takes its top chi
 * eigenpairs by subspace iteration on 2 chi columns, Rayleigh-Ritz and a positive-definiteness
 * certificate (the full eigensolver only where that declines); on = 0 always runs the eigensolver,
 * on < 0 leaves the switch. Returns the previous setting. mpdo_debug_layer_chi_tops: how many
 * pair-bond chi steps of the last fused layer on this thread were decided that way. */
int mpdo_debug_layer_chi_top(int on);
int mpdo_debug_layer_chi_tops(void);
/* Subspace of the chi step's top-eigenpair solver: extra > 0 takes k + extra columns (rounded up
 * to a multiple of 64; default 128), extra = 0 takes 2k, extra < 0 leaves it; q >= 1 sets the
 * subspace-iteration steps (default 5). Returns the previous extra. */
int mpdo_debug_layer_top_shape(int extra, int q);
/* That top-k solver on `batch` Hermitian positive semidefinite complex128 Gram matrices G (device,
 * n x n each, contiguous; not modified): m subspace columns, q steps. info (host, 2 ints): accepted,
 * the step that declined. When accepted, w (device, n doubles per matrix) holds the m Ritz values
 * (descending), the trace's remainder, then zeros, and V (device, complex128 n x n per matrix) the
 * m Ritz vectors in its first m columns. */
mpdo_status mpdo_debug_gram_top_eig(int n, int batch, const void* G, int k, int m, int q, double* w, void* V, int* info,
                                    void* stream);
/* That bottom-of-the-spectrum decision on `batch` Hermitian positive semidefinite complex128 Gram
 * matrices G (device, n x n each, contiguous): m subspace columns, q iteration steps, absolute
 * error eps. info (host, 4 ints): accepted, some chain drops a direction, the largest kept count,
 * the step that declined (0 when accepted). When accepted, kept and disc (host, batch each) hold the
 * rule's kept counts and dropped eigenvalue sums, and where a direction is dropped P (device,
 * complex64 n x n each) holds I - V_d V_d^H. */
mpdo_status mpdo_debug_gram_tail_cut(int n, int batch, const void* G, double eps, int m, int q, int* kept,
                                     double* disc, void* P, int* info, void* stream);
int mpdo_debug_layer_gate_skips(void);

/* The next `count` eigensolver calls of the per-gate entry points (mpdo_apply_2q, the sweeps,
 * the kappa cut) run to completion and then return MPDO_ENOCONV, as if the sweep cap had been
 * hit: the host-side route of the reference's perturb-and-retry (Circuit.py:118-136) without any
 * kernel that fails to finish. 0 turns it off. */
void mpdo_debug_fail_eig(int count);

/* Launch timing by kernel class (0 complex GEMM, 1 block-Jacobi subproblem solve (bj_solve),
 * 2 block-Jacobi eigenvector update (bj_rotate_v_direct), 3 single-workgroup Jacobi and Cholesky
 * panels, 4 permutation, 5 block-Jacobi two-sided matrix update (bj_update)): HIP events recorded on each
 * launch's own stream. enable(1) resets the totals and starts recording, enable(0) stops;
 * read() waits for the recorded launches and returns the totals of one class: launches, summed
 * duration in ms, algorithmic flops (the arithmetic the launches are asked to do) and operand
 * bytes. */
mpdo_status mpdo_debug_prof_enable(int on);
mpdo_status mpdo_debug_prof_read(int cls, long long* launches, double* ms, double* flops, double* bytes);
/* Per-shape totals of one timing class while timing is on: shapes [5 n] (for GEMMs: M, N, K,
 * batch, precision mode), stats [3 n] (launches, ms, algorithmic flops); returns the number of
 * distinct shapes (fills at most cap). */
int mpdo_debug_prof_shapes(int cls, int* shapes, double* stats, int cap);
/* Work the launches of class cls executed while timing was on, as counted from the device's
 * skip flags (the block-Jacobi kernels skip identity rotations; mpdo_debug_prof_read gives the
 * work they were asked to do). Zero for classes without such flags. */
mpdo_status mpdo_debug_prof_read_executed(int cls, double* flops, double* bytes);
/* The blocked eigensolver calls (order > 64) made while timing was on: (order, batch, sweeps)
 * triples, up to cap ints written; returns the number of calls. */
int mpdo_debug_prof_eig_log(int* out, int cap);

#ifdef __cplusplus
}
#endif

#endif /* MPDO_DEBUG_H */
