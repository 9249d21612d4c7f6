/*
 * libmpdo_gfx950 — C ABI of the MI355X (gfx950) implementation of the MPDO noisy-circuit
 * update path of WeiguoMa/Tomography-assisted-MPDO-QCircuit (package MPDOSimulator).
 *
 * Data model: one site tensor per qubit, complex64 (interleaved re, im floats), row-major
 * T[l][2][a][r] in device memory: l = bond to the left neighbour (bond_{i-1}_{i}), 2 = physics_i,
 * a = inner/Kraus index (I_i), r = bond to the right neighbour. A missing edge is a dimension 1.
 * Output tensors are written into caller-allocated device buffers whose capacity the matching
 * *_dims / *_capacity query gives; the entry points return the actual dimensions in the
 * output descriptors. Every entry point is stream-ordered on `stream` (a hipStream_t, NULL =
 * default stream) and returns after the results are complete when it has to read a
 * data-dependent rank back (documented per call). Scratch comes from a workspace per (device,
 * stream): the caller's (mpdo_bind_workspace) or the library's own grow-only buffer; calls on
 * different streams may run concurrently from several host threads.
 */
#ifndef MPDO_H
#define MPDO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MPDO_OK = 0,
  MPDO_EINVAL = 1,   /* bad shape/argument (maps to ValueError/TypeError)                    */
  MPDO_ENOCONV = 2,  /* eigensolver sweep cap hit (maps to torch.linalg.LinAlgError)          */
  MPDO_ENOMEM = 3,   /* device allocation of the workspace failed                            */
  MPDO_EHIP = 4      /* HIP runtime error                                                    */
} mpdo_status;

typedef struct {
  void* data; /* device pointer, complex64 T[l][2][a][r] */
  int l;
  int a;
  int r;
} mpdo_site;

int mpdo_version(void);
/* Thread-local text of the last failure. */
const char* mpdo_last_error(void);
/* Frees the library's own workspace buffers of the current device (not bound ones). */
void mpdo_release_workspace(void);
/*
 * Caller-owned workspace (no counterpart in the reference, whose tensors torch allocates). Binds
 * the device buffer ptr (bytes long, 256-byte aligned) to `stream` on the current device: every
 * later call on that stream carves its scratch from it and allocates nothing; ptr NULL unbinds.
 * A call that needs more than the bound buffer returns MPDO_ENOMEM before launching anything and
 * mpdo_workspace_needed() (thread-local) then gives the bytes it needs. Without a binding the
 * library keeps a grow-only buffer of its own per (device, stream). Calls on different streams
 * use different buffers and may overlap in time; calls on one stream must not.
 */
mpdo_status mpdo_bind_workspace(void* stream, void* ptr, size_t bytes);
size_t mpdo_workspace_needed(void);
/*
 * Accumulation precision of the complex64 tensor contractions whose results are complex64 site
 * data (no counterpart in the reference, which computes in the dtype it is given): 0 = fp64
 * matrix-core accumulation (default; complex64 storage), 1 = fp32 matrix-core accumulation of
 * four real products per complex product (twice the fp64 peak rate), 2 = fp32 accumulation of
 * three real products (3M: Re = P1 - P2, Im = P3 - P1 - P2 with P1 = Ar Br, P2 = Ai Bi,
 * P3 = (Ar + Ai)(Br + Bi); 8/3 of the fp64 rate), 3 = bf16 matrix cores on operands split into
 * two bf16 parts (x = p0 + p1 + r, |r| <= 2^-16 |x|; three part products per real product, fp32
 * accumulation: about 2^-16 relative per product), 4 = the same with three parts (six part
 * products: fp32-level products). Gram matrices, eigensolves and the QR/LQ factors are always
 * fp64. Returns the previous mode.
 */
int mpdo_set_math_mode(int mode);
/*
 * Tile autotuner of the fp64-accumulating complex GEMMs (no counterpart in the reference, which
 * calls the vendor's GEMMs). A key is a GEMM's (M, N, K, batch, precision mode, Hermitian output,
 * triangular operand, A stored complex128, B stored complex128, reduction slices, A contiguous
 * along k, B contiguous along k); the table maps keys to tile shapes (0 = 64x64x16 one LDS stage,
 * 1 = 64x64x32, 2 = 128x64x16, 3 = 64x128x16, 4 = 128x128x16, 5 = 64x64x16 two LDS stages,
 * 6 = 32x128x16). Keys absent from the table use the built-in rules.
 * record(1) starts recording the first launch of every new key (clearing the previous record),
 * record(0) stops; returns the previous state. tune_run replays every recorded GEMM with each tile
 * shape that fits it, `reps` timed launches on `stream` after one untimed, keeps the fastest in the
 * table and returns the number of keys tuned (-1 on a HIP error). The replays overwrite the
 * recorded GEMMs' outputs, whose buffers must still be allocated: run it right after the recorded
 * work and discard that work's results. get writes (12 key ints, tile) per entry into out (at most
 * cap entries) and returns the number of entries; set replaces the table with `count` such entries
 * and returns how many it accepted.
 */
int mpdo_gemm_tune_record(int on);
int mpdo_gemm_tune_run(int reps, void* stream);
int mpdo_gemm_tune_get(int* out, int cap);
int mpdo_gemm_tune_set(const int* in, int count);

/*
 * Single-qubit gate / Kraus absorption, batched over sites.
 * Replaces Circuit.py:138-178 (TensorCircuit._apply_single_qubit_gate) for each target qubit:
 * T'[l, s', a*K + k, r] = sum_s E_k[s', s] T[l, s, a, r] (old inner index major, the
 * tn.flatten_edges order of Circuit.py:174-178). A plain gate is K = 1.
 * kraus: host array, per site K[i] matrices E_k[s'][s] (complex64), concatenated.
 * out[i].data must hold l*2*(a*K[i])*r elements and must not alias in[i].data.
 */
mpdo_status mpdo_apply_1q(int nsites, const mpdo_site* in, const mpdo_site* out,
                          const float* kraus, const int* K, void* stream);

/*
 * Two-qubit gate (+ Kraus channel) on adjacent sites (min, max) and the SVD split.
 * Replaces Circuit.py:74-136 (TensorCircuit._apply_two_qubits_gate) with tn.split_node's
 * U*sqrt(S) | sqrt(S)*V^H convention: theta[(l, s_min, a_min), (s_max, a_max*K + k, r)] =
 * sum G[s_min', s_max', s_min, s_max, k] T_min T_max, split keeping the singular values of
 * decompositions.svd's rule with absolute error `eps_abs` (the reference passes
 * GLOBAL_MINIMUM = 2.718281828459045e-8, Circuit.py:22,123) and no count cap.
 * gate: host (2,2,2,2,K) complex64 ordered (o_min, o_max, i_min, i_max, k).
 * out_min->data / out_max->data: capacity mpdo_apply_2q_capacity(...) for the new bond.
 * On return out_min = (l, a_min, k) and out_max = (k, a_max*K, r); *discarded = sum s_rest^2.
 * Synchronises `stream` once (the kept rank is data dependent).
 */
int mpdo_apply_2q_capacity(const mpdo_site* tmin, const mpdo_site* tmax, int K);
mpdo_status mpdo_apply_2q(const mpdo_site* tmin, const mpdo_site* tmax, const float* gate, int K,
                          double eps_abs, mpdo_site* out_min, mpdo_site* out_max, double* discarded,
                          void* stream);

/*
 * Left-to-right QR sweep. Replaces TNNOptimizer.py:87-108 (qr_left2right) with
 * decompositions.py:149-195 (qr): for i = 0..n-2, T_i as [(l s a) x r] = Q R with
 * min(rows, cols) columns, T_i <- Q, T_{i+1} <- R T_{i+1}. mpdo_qr_left2right_dims fills the
 * output dims (deterministic); out[i].data must hold them.
 */
void mpdo_qr_left2right_dims(int n, const mpdo_site* in, mpdo_site* out);
mpdo_status mpdo_qr_left2right(int n, const mpdo_site* in, mpdo_site* out, void* stream);

/*
 * Truncating SVD sweep with the sqrt(S) split.
 * direction 0 replaces TNNOptimizer.py:111-134 (svd_right2left), direction 1 replaces
 * TNNOptimizer.py:137-161 (svd_left2right); keep rule decompositions.py:116-133 with
 * relative=True; max_singular_values <= 0 means None, max_truncation_err < 0 means None.
 * left_canonical != 0 asserts that sites 0..n-2 are left-orthonormal (true right after
 * mpdo_qr_left2right, the bondTruncate pairing of TNNOptimizer.py:72-84), which lets each step
 * skip the factorisation of the left tensor. mpdo_svd_sweep_dims gives capacity dims.
 * On return out[i] holds actual dims; discarded[j] (n-1 entries, bond order of the sweep)
 * receives sum s_rest^2 of each step. Synchronises once per step.
 */
void mpdo_svd_sweep_dims(int n, const mpdo_site* in, int direction, int max_singular_values, mpdo_site* out);
mpdo_status mpdo_svd_sweep(int n, const mpdo_site* in, mpdo_site* out, int direction,
                           int max_singular_values, double max_truncation_err, int left_canonical,
                           double* discarded, void* stream);

/*
 * Inner-index (kappa) truncation of independent sites. Replaces the loop body of
 * TNNOptimizer.py:164-197 (svdKappa_left2right): SVD of T as [(l s r) x a], keep
 * min(kappa, k_err) (relative=True), T <- U*S. out[i].data capacity: l*2*min(kappa, a)*r
 * (or a when kappa <= 0). discarded[i] = sum s_rest^2. Synchronises once.
 */
mpdo_status mpdo_svd_kappa(int nsites, const mpdo_site* in, mpdo_site* out, int max_singular_values,
                           double max_truncation_err, double* discarded, void* stream);

/*
 * Unnormalised Re Tr(O rho~) for products of single-site operators, rho~ = sum_a T_a T_a^H.
 * Replaces dmOperations.py:174-201 (pauli_expect) and :40-48 (trace_rho).
 * codes: nstrings x n, per site -1 = identity (traced), 0 = X, 1 = Y, 2 = Z (PAULI_DICT,
 * dmOperations.py:11-15), 3 = |0><0|, 4 = |1><1| (projectors for bitstring probabilities).
 * out: host, nstrings doubles. Synchronises once.
 */
mpdo_status mpdo_pauli_expect(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out,
                              void* stream);

/*
 * mpdo_pauli_expect without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and synchronisation as
 * mpdo_pauli_expect; the strings' environments live in a fixed 32 MiB device array of the library
 * (one per GPU, held until the process ends) instead of the grow-only workspace. Returns
 * MPDO_ENOMEM, before launching anything, when one string needs more than that at these site
 * shapes (two l*2*a*r complex64 intermediates of the largest site plus two bond^2 environments:
 * n = 32, chi = 256, a = 8 fits). Every call on one GPU shares that array: calls that overlap in
 * time on one GPU (from several host threads or streams) race and must not be made.
 */
mpdo_status mpdo_pauli_expect_noalloc(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out,
                                      void* stream);

/*
 * Reduced density matrix of the qubits with keep[i] != 0 (others traced), rows = ket bits of
 * the kept qubits in increasing index order, lowest index most significant, unnormalised.
 * Replaces Circuit.py:226-262 (_create_dmNodes/_contract_dm) + dmOperations.py:18-37.
 * out: device complex64 (2^k x 2^k), k = number of kept qubits (k <= 12).
 */
mpdo_status mpdo_reduced_dm(int n, const mpdo_site* sites, const int* keep, void* out, void* stream);

/*
 * Re Tr(rho0 rho1) of two MPDOs over the same qubits, by a transfer sweep over the four open
 * bonds (any n; bonds up to chi0 * chi1 ~ 2^15). Replaces dmOperations.py:51-66 (trace_rho_rho).
 * mpdo_trace_rho_rho_masked: qubits with keep[i] == 0 are traced within each state first (both
 * states reduced on the same qubits, reduce_dmNodes, dmOperations.py:18-37); keep NULL = all kept.
 * out: host double. Synchronises once.
 */
mpdo_status mpdo_trace_rho_rho(int n, const mpdo_site* a, const mpdo_site* b, double* out, void* stream);
mpdo_status mpdo_trace_rho_rho_masked(int n, const mpdo_site* a, const mpdo_site* b, const int* keep, double* out,
                                      void* stream);

/*
 * One brickwork circuit-layer for `batch` chains with identical shapes: the single-qubit
 * channels of every site, the two-qubit channel of every pair (i, i+1), i = parity, parity+2, ...,
 * then truncate() (QR sweep, chi sweep, kappa cut). Replaces the sequence
 * Circuit.py:138-178 (per qubit) + Circuit.py:74-136 (per pair) + Circuit.py:476-481
 * (TNNOptimizer.py:72-197) of one layer of the reference's evolve loop, with the same
 * truncation decisions and output shapes as calling mpdo_apply_1q, mpdo_apply_2q,
 * mpdo_qr_left2right, mpdo_svd_sweep(left_canonical=1) and mpdo_svd_kappa in that order (no
 * error threshold: only count caps). The two-site tensors of the pairs are never materialised.
 * in[i].data / out[i].data: complex64 [batch][l][2][a][r]; out dims from mpdo_layer_dims, which
 * returns 0 (and sets the error text) when the layer does not qualify (then use the per-gate
 * entry points). kraus1: host [batch][n][K1][2][2]; gates2: host [npairs][2][2][2][2][K2] in
 * (o_min, o_max, i_min, i_max, k) order. discarded: host [batch][3n]: n chi-step slots by bond
 * index, n kappa slots by site, n gate-time split slots by pair min site (the weight the
 * reference's split of each pair drops, Circuit.py:120-124), or NULL. Synchronises once per pair
 * (the kept count of its gate-time split) and at the end.
 */
int mpdo_layer_dims(int n, const mpdo_site* in, int K1, int parity, int K2, int chi, int kappa, mpdo_site* out);
/* Workspace bytes mpdo_layer_batched needs for these chains (dims of `in` only; data ignored);
 * 0 when the layer does not qualify (error text set). */
size_t mpdo_layer_workspace_bytes(int batch, int n, const mpdo_site* in, int K1, int parity, int K2, int chi,
                                  int kappa);
mpdo_status mpdo_layer_batched(int batch, int n, const mpdo_site* in, const float* kraus1, int K1, int parity,
                               const float* gates2, int K2, int chi, int kappa, const mpdo_site* out,
                               double* discarded, void* stream);

/*
 * Shot sampling (replaces the contractions of Circuit.py:297-332, _conditional_prb, driven by
 * _conditional_batch_sample / _conditional_sample, Circuit.py:334-391).
 *
 * mpdo_env_right: traced tail environments R_i[l, l'] = sum over sites i..n-1 of the network
 * rho~ with every physical index traced (R_n = 1), complex128, packed back to back: R_i starts at
 * element sum_{j<i} l_j^2 and is l_i x l_i (ket bond, bra bond); mpdo_env_right_elems gives the
 * total count. Open boundary bonds must be 1.
 *
 * mpdo_sample_step: one site of the left-to-right sweep for `groups` histories. Group g starts
 * from environment prev[sel[g]] (sel: host array of `groups` indices < prev_count; NULL = prev[g]),
 * each environment l x l complex128 with l = site->l (the first site starts from the 1 x 1
 * identity). sampled != 0: out[g][s] = E projected on physical value s (r x r each, s = 0, 1) and
 * probs[g][s] (host) = Re sum E_s .* R_next, the unnormalised P(history, s) the reference reads
 * off the diagonal of its contraction; the call synchronises. sampled == 0 (a qubit that is
 * traced, Circuit.py:299-300,313): out[g] = E_0 + E_1 and probs may be NULL.
 */
long long mpdo_env_right_elems(int n, const mpdo_site* sites);
mpdo_status mpdo_env_right(int n, const mpdo_site* sites, void* out, void* stream);
mpdo_status mpdo_sample_step(const mpdo_site* site, int groups, const void* prev, int prev_count, const int* sel,
                             int sampled, const void* R_next, void* out, double* probs, void* stream);

/*
 * Batched Hermitian eigendecomposition in fp64 (the Jacobi solver every SVD and QR of the path is
 * built on; decompositions.py:51-195 computes those factorisations with torch.linalg.svd/qr).
 * A: device complex128 [batch][n][n] row-major, Hermitian, destroyed; w: device double
 * [batch][n], eigenvalues in descending order; V: device complex128 [batch][n][n], column j the
 * eigenvector of w[j]. MPDO_ENOCONV when the sweep cap is hit. Synchronises once. The
 * differentiable evolution builds its QR, SVD and kappa cuts on it.
 */
mpdo_status mpdo_eigh(int n, int batch, void* A, double* w, void* V, void* stream);

/*
 * State vector of a pure MPDO (every inner dimension 1), amplitudes in ket-bit order with the
 * lowest qubit index most significant. Replaces Circuit.py:264-272 (cal_vector).
 * out: device complex64, 2^n entries (n <= 30).
 */
mpdo_status mpdo_contract_state(int n, const mpdo_site* sites, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPDO_H */
