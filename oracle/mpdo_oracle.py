"""CPU ORACLE — test infrastructure, never the product.

A numpy restatement of the reference's MPDO noisy-circuit update path
(WeiguoMa/Tomography-assisted-MPDO-QCircuit, package ``MPDOSimulator``), written to be read next to
the reference files it cites. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and only as the checker.

State layout (fixed, replaces tensornetwork's named edges): every site tensor is ``T[l, s, a, r]``
with ``l`` = bond to site i-1 (``bond_{i-1}_{i}``), ``s`` = 2 (``physics_i``), ``a`` = inner/Kraus
index (``I_i``), ``r`` = bond to site i+1. An absent edge is a dimension-1 axis, plus explicit flags
(``has_I``, ``bond``) for the two places where the reference's behaviour depends on the presence of
an edge rather than on its dimension (``Circuit.py:83,140`` and ``TNNOptimizer.py:51-61,176-181``).

Parity pinning (how each part is checked, ``tests/test_oracle.py``):
  * gate matrices: matrix exponentials of their generators (the reference's RYY is the XY
    rotation exp(-i t/2 X (x) Y), YGates.py:138-143), fixed gates from projector identities;
  * Kraus builders: Bloch-vector / matrix-element actions of depolarizing, amplitude damping
    and dephasing, CPTP identities, the noisy-gate composition order; RealNoise: ``ideal_cz.mat``
    = CZ with one operator and the recorded sum E^H E diagonal of ``czDefault.mat`` (K=16);
  * decompositions.svd keep-count rule: hand-computed spectra (ties, relative mode,
    min(chi, k), the last index); the rank-aware split against the explicit SVD;
  * gate application, sweeps without truncation, observables: the independent dense complex128
    density-matrix simulator (``oracle/dense.py``);
  * chi cut: the Schmidt values of a dense pure state at the first cut of the sweep;
    kappa cut: the eigen-decomposition of a single site's density matrix;
  * sampling: conditional probabilities against marginals of the dense density matrix (every
    history, reduced qubits, rotated bases), the draws against the exact distribution
    (chi-square), the reference's rules for p1 within GLOBAL_MINIMUM of 0 or 1 (no draw).
Functions whose numbers no such check reaches are marked "parity unpinned" in their docstring.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Circuit.py:22-23
GLOBAL_MINIMUM = 2.718281828459045 * 1e-8
MAX_ATTEMPTS = 4

# ----------------------------------------------------------------------------------------------
# Gate matrices (QuantumGates/*.py). 1q gates: G[out, in]; 2q gates: G[o0, o1, i0, i1] with the
# first listed qubit as the most significant bit (row-major reshape of the 4x4 matrix).
# ----------------------------------------------------------------------------------------------
_S2 = 1.0 / math.sqrt(2.0)


def gate_matrix(name: str, params: Sequence[float] = ()) -> np.ndarray:
    """Return the tensor a reference gate's ``.tensor`` property builds (complex128)."""
    n = name.upper()
    p = [float(x) for x in params]
    c = lambda t: math.cos(t / 2.0)
    s = lambda t: math.sin(t / 2.0)
    one = {
        'I': [[1, 0], [0, 1]],                                       # SingleGates.py:17-47
        'H': [[_S2, _S2], [_S2, -_S2]],                              # SingleGates.py:50-80
        'X': [[0, 1], [1, 0]],                                       # XGates.py:15-45
        'Y': [[0, -1j], [1j, 0]],                                    # YGates.py:15-45
        'Z': [[1, 0], [0, -1]],                                      # ZGates.py:15-45
        'S': [[1, 0], [0, 1j]],                                      # PhaseGates.py:16-48
        'SDG': [[1, 0], [0, -1j]],                                   # PhaseGates.py:49-81
        'T': [[1, 0], [0, (1 + 1j) * _S2]],                          # PhaseGates.py:82-114
        'MEASUREX': [[_S2, _S2], [_S2, -_S2]],                       # SingleGates.py:240-271
        'MEASUREY': [[_S2, -1j * _S2], [_S2, 1j * _S2]],             # SingleGates.py:274-305
        'MEASUREZ': [[1, 0], [0, 1]],                                # SingleGates.py:308-339
        'RESET0': [[1, 0], [0, 0]],                                  # SingleGates.py:342-373
        'RESET1': [[0, 0], [0, 1]],                                  # SingleGates.py:376-407
    }
    if n in one:
        return np.array(one[n], dtype=np.complex128)
    if n == 'RX':                                                    # XGates.py:48-85
        t = p[0]
        return np.array([[c(t), -1j * s(t)], [-1j * s(t), c(t)]], dtype=np.complex128)
    if n == 'RY':                                                    # YGates.py:48-85
        t = p[0]
        return np.array([[c(t), -s(t)], [s(t), c(t)]], dtype=np.complex128)
    if n == 'RZ':                                                    # ZGates.py:48-84
        t = p[0]
        return np.array([[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]], dtype=np.complex128)
    if n in ('P', 'U1'):                                             # PhaseGates.py:115-153, SingleGates.py:83-119
        return np.array([[1, 0], [0, np.exp(1j * p[0])]], dtype=np.complex128)
    if n == 'U2':                                                    # SingleGates.py:122-159 (uses phi, lam)
        phi, lam = p
        return np.array([[1, -np.exp(1j * lam)], [np.exp(1j * phi), np.exp(1j * (lam + phi))]],
                        dtype=np.complex128) * _S2
    if n == 'U3':                                                    # SingleGates.py:162-201
        t, phi, lam = p
        return np.array([[c(t), -np.exp(1j * lam) * s(t)],
                         [np.exp(1j * phi) * s(t), np.exp(1j * (lam + phi)) * c(t)]], dtype=np.complex128)
    two = None
    if n in ('CNOT', 'CX'):                                          # DoubleGates.py:50-82, XGates.py:88-120
        two = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]
    elif n == 'CY':                                                  # YGates.py:88-121
        two = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]
    elif n == 'CZ':                                                  # ZGates.py:87-119
        two = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]
    elif n == 'II':                                                  # DoubleGates.py:15-47
        two = np.eye(4)
    elif n == 'SWAP':                                                # DoubleGates.py:120-152
        two = [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]
    elif n == 'ISWAP':                                               # DoubleGates.py:85-117
        two = [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]
    elif n == 'CP':                                                  # PhaseGates.py:154-190
        two = np.diag([1, 1, 1, np.exp(1j * p[0])])
    elif n == 'RXX':                                                 # XGates.py:123-160
        t = p[0]
        a, b = c(t), -1j * s(t)
        two = [[a, 0, 0, b], [0, a, b, 0], [0, b, a, 0], [b, 0, 0, a]]
    elif n == 'RYY':                                                 # YGates.py:124-161 (:138-143)
        # the reference's matrix is exp(-i t/2 X (x) Y), not exp(-i t/2 Y (x) Y)
        t = p[0]
        a, b = c(t), s(t)
        two = [[a, 0, 0, -b], [0, a, b, 0], [0, -b, a, 0], [b, 0, 0, a]]
    elif n == 'RZZ':                                                 # ZGates.py:122-159
        t = p[0]
        m, pl = np.exp(-0.5j * t), np.exp(0.5j * t)
        two = np.diag([m, pl, pl, m])
    elif n == 'PSWAP':                                               # DoubleGates.py:155-190
        t = p[0]
        a, b = math.cos(t), math.sin(t)
        two = [[1, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, 1]]
    elif n == 'XXPLUSYY':                                            # DoubleGates.py:229-262
        t, beta = p
        a, b = math.cos(t / 2), math.sin(t / 2)
        two = [[1, 0, 0, 0],
               [0, a, -1j * b * np.exp(1j * beta), 0],
               [0, -1j * b * np.exp(-1j * beta), a, 0],
               [0, 0, 0, 1]]
    if two is None:
        raise ValueError(f'unknown gate {name}')
    return np.asarray(two, dtype=np.complex128).reshape(2, 2, 2, 2)


# ----------------------------------------------------------------------------------------------
# Noise (NoiseChannel.py, ChipInfo.py)
# ----------------------------------------------------------------------------------------------
PAULIS = [np.eye(2, dtype=np.complex128),
          np.array([[0, 1], [1, 0]], dtype=np.complex128),
          np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
          np.array([[1, 0], [0, -1]], dtype=np.complex128)]

CHIPS = {  # ChipInfo.py:44-87 (gate_time, bath_rate, decay_rate, dephasing_rate, T1, T2, dpc)
    'best': dict(gate_time=30.0, bath_rate=0.0, decay_rate=0.0, dephasing_rate=0.0,
                 T1=2e11, T2=2e10, dpc_error_rate=11e-4),
    'medium': dict(gate_time=1.0, bath_rate=0.01, decay_rate=0.98, dephasing_rate=0.02,
                   T1=2e11, T2=2e10, dpc_error_rate=5e-2),
    'worst': dict(gate_time=30.0, bath_rate=0.0, decay_rate=0.0, dephasing_rate=0.0,
                  T1=2e2, T2=2e1, dpc_error_rate=11e-2),
}


def depolarizing_tensor(p: float, qn: int = 1) -> np.ndarray:
    """NoiseChannel.py:51-91: Kraus tensor of shape (2,)*2qn + (4**qn,), Pauli strings with the
    first qubit's Pauli as the major index, weight sqrt(1-(4^q-1)p/4^q) on I..I, sqrt(p/4^q) else."""
    if not 0 <= p <= 1:
        raise ValueError('Probability p must be in the range [0, 1].')
    d = 4 ** qn
    w = [math.sqrt(1 - (d - 1) * p / d)] + [math.sqrt(p / d)] * (d - 1)
    mats = list(PAULIS)
    for _ in range(qn - 1):
        mats = [np.kron(e, b) for e in mats for b in PAULIS]
    st = np.stack(mats) * np.asarray(w)[:, None, None]             # (K, 2^q, 2^q)
    return np.transpose(st, (1, 2, 0)).reshape([2] * (2 * qn) + [d])


def decay_tensor(param: float) -> np.ndarray:
    """NoiseChannel.py:144-155 with _param = 1-exp(-bath*rate*t): K1=diag(1,sqrt(1-g)), K2=sqrt(g)|0><1|."""
    k = np.array([[[1, 0], [0, math.sqrt(1 - param)]], [[0, math.sqrt(param)], [0, 0]]], dtype=np.complex128)
    return np.transpose(k, (1, 2, 0))


def dephasing_tensor(param: float) -> np.ndarray:
    """NoiseChannel.py:157-170: sqrt(1-p) I, sqrt(p)|0><0|, sqrt(p)|1><1|."""
    a, b = math.sqrt(param), math.sqrt(1 - param)
    k = np.array([[[b, 0], [0, b]], [[a, 0], [0, 0]], [[0, 0], [0, a]]], dtype=np.complex128)
    return np.transpose(k, (1, 2, 0))


def amp_phase_damping_tensor(time: float, T1: float, T2: float) -> np.ndarray:
    """NoiseChannel.py:93-142 (built by the reference but never applied on the path)."""
    T2p = 2 * T1 * T2 / (2 * T1 - T2)
    p1 = 1 - math.exp(-time / T1)
    p2 = 1 - math.exp(-time / T2p)
    k = np.array([[[1, 0], [0, math.sqrt(1 - (p1 + p2))]],
                  [[0, 0], [0, math.sqrt(p2)]],
                  [[0, math.sqrt(p1)], [0, 0]]], dtype=np.complex128)
    return np.transpose(k, (1, 2, 0))


@dataclass
class NoiseSpec:
    """Noise parameters of one TensorCircuit. ``single`` selects the 1q noisy-gate channel:
    'decay_dephasing' is the reference's (Circuit.py:144-155); 'depolarizing' is an extension
    (1q depolarizing Kraus of NoiseChannel.py:51-91 applied after each noisy 1q gate)."""
    chip: str = 'worst'
    decay_param: Optional[float] = None
    dephasing_param: Optional[float] = None
    dpc_error_rate: Optional[float] = None
    single: str = 'decay_dephasing'
    p1: float = 0.0

    def params(self):
        c = CHIPS[self.chip]
        dec = self.decay_param
        if dec is None:
            dec = 1 - math.exp(-c['bath_rate'] * c['decay_rate'] * c['gate_time'])
        dep = self.dephasing_param
        if dep is None:
            dep = 1 - math.exp(-c['bath_rate'] * c['dephasing_rate'] * c['gate_time'])
        dpc = self.dpc_error_rate if self.dpc_error_rate is not None else c['dpc_error_rate']
        return dec, dep, dpc

    def single_kraus(self, G: np.ndarray) -> np.ndarray:
        """Circuit.py:144-155: einsum('nlm, ljk, ji -> nimk', decay, dephasing, G).reshape(2,2,-1)."""
        dec, dep, _ = self.params()
        if self.single == 'depolarizing':
            D = depolarizing_tensor(self.p1, 1)
            return np.einsum('nlm,li->nim', D, G)
        return np.einsum('nlm,ljk,ji->nimk', decay_tensor(dec), dephasing_tensor(dep), G).reshape(2, 2, -1)

    def double_kraus(self, G: np.ndarray) -> np.ndarray:
        """Circuit.py:91-99: einsum('ijklp, klmn -> ijmnp', dpCTensor2, G)."""
        _, _, dpc = self.params()
        return np.einsum('ijklp,klmn->ijmnp', depolarizing_tensor(dpc, 2), G)


# ----------------------------------------------------------------------------------------------
# RealNoise (RealNoise.py:22-179, Tools.py:466-509)
# ----------------------------------------------------------------------------------------------
def _name2matrix(name: str) -> np.ndarray:
    """Tools.py:488-509; note 'Y' is -1j * Y = [[0,-1],[1,0]] (Tools.py:504)."""
    ops = {'I': np.eye(2, dtype=np.complex128), 'X': PAULIS[1], 'Y': -1j * PAULIS[2], 'Z': PAULIS[3]}
    m = np.array([[1.0 + 0j]])
    for ch in name:
        m = np.kron(m, ops[ch])
    return m


def chi_to_kraus(chi: np.ndarray) -> np.ndarray:
    """RealNoise.py:35-161 + :432-438: chi (4^N x 4^N) -> Kraus tensor (2,2,2,2,K).

    Eigen-decompose chi (numpy eig), sort eigenvalues descending (numpy complex sort), count
    |lambda| > 1e-12, and build E_i = sum_j sqrt(lambda_i) U_ji P_j for the first ``count``
    sorted eigenpairs, zeroing coefficients below 1e-12 (RealNoise.py:68-104)."""
    N = int(round(math.log(chi.shape[0], 4)))
    basis = [''.join(g) for g in itertools.product('IXYZ', repeat=N)]
    s, u = np.linalg.eig(chi)
    idx = s.argsort()[::-1]
    s, u = s[idx], u[:, idx]
    count = int(np.sum(np.abs(s) > 1e-12))
    mats = []
    for i in range(count):
        e = u.T[i] * np.sqrt(s[i])
        e = np.where(np.abs(e) < 1e-12, 0, e)
        m = sum(e[j] * _name2matrix(basis[j]) for j in range(len(basis)) if e[j] != 0)
        mats.append(np.asarray(m).reshape([2] * (2 * N)))
    return np.moveaxis(np.stack(mats), 0, -1)


def load_chi(path: str) -> np.ndarray:
    """RealNoise.py:22-32 ('.mat' key 'exp', '.npz' key 'chi')."""
    if path.endswith('.mat'):
        from scipy.io import loadmat
        return loadmat(path)['exp']
    if path.endswith('.npz'):
        return np.load(path)['chi']
    raise TypeError('Current file-type is not supported.')


# ----------------------------------------------------------------------------------------------
# decompositions.py (the repo's replacement of tensornetwork's pytorch backend file)
# ----------------------------------------------------------------------------------------------
def keep_count(s: np.ndarray, max_singular_values: Optional[int], max_truncation_error: Optional[float],
               relative: bool) -> int:
    """decompositions.py:116-133: number of singular values kept.

    trunc_errs = sqrt(cumsum(s^2)); with an error bound, keep idx+1 for the first idx (excluding
    the last) where trunc_errs[-1] - trunc_errs[idx] <= err (err * s[0] when relative); result
    is min(max_singular_values or len(s), that count), at most len(s)."""
    n = s.shape[0]
    msv = n if max_singular_values is None else max_singular_values
    k_err = msv
    if max_truncation_error is not None and n > 0:
        te = np.sqrt(np.cumsum(s.astype(np.float64) ** 2))
        bound = max_truncation_error * s[0] if relative else max_truncation_error
        for i in range(n - 1):
            if te[-1] - te[i] <= bound:
                k_err = i + 1
                break
    # the reference slices s[:k] (decompositions.py:135-139), which caps k at len(s)
    return min(msv, k_err, n)


def randomized_svd(M: np.ndarray, n_components: int, n_over: int = 7, rng_draw=None):
    """decompositions.py:23-48, the branch decompositions.py:111-114 takes when numel >= 10000 and a
    cap is given.

    ``rng_draw(m, k)`` supplies the Gaussian test matrix; the reference draws it from a fresh
    default-seeded CPU ``torch.Generator`` (parity unpinned for the draw itself, torch behaviour;
    the algorithm is checked by closed forms in tests/test_oracle.py)."""
    m, n = M.shape
    k = n_components + n_over
    Q = rng_draw(m, k) if rng_draw is not None else (
        np.random.default_rng(0).standard_normal((m, k)) + 1j * np.random.default_rng(1).standard_normal((m, k))
    ) / math.sqrt(2)
    it = 3 if m >= n else 2
    for _ in range(it):
        Q = M @ (M.conj().T @ Q)
    Q, _ = np.linalg.qr(Q)
    B = Q.conj().T @ M
    u, s, vh = np.linalg.svd(B, full_matrices=False)
    return Q @ u, s, vh


def svd(M: np.ndarray, max_singular_values=None, max_truncation_error=None, relative=False,
        randomized: bool = False, rng_draw=None):
    """decompositions.py:51-146 on an already-flattened matrix: returns (u, s, vh, s_rest).

    ``randomized`` reproduces the reference's branch for numel >= 10000 with a cap; the oracle's
    default is the exact SVD (the reference's algorithm with exact SVD)."""
    if randomized and M.size >= 10000 and max_singular_values is not None:
        u, s, vh = randomized_svd(M, max_singular_values, rng_draw=rng_draw)
    else:
        u, s, vh = np.linalg.svd(M, full_matrices=False)
    k = keep_count(s, max_singular_values, max_truncation_error, relative)
    return u[:, :k], s[:k], vh[:k, :], s[k:]


def qr(M: np.ndarray):
    """decompositions.py:149-195: reduced QR, no sign fixing (non_negative_diagonal=False)."""
    return np.linalg.qr(M, mode='reduced')


def _keep_and_pad(u, s, vh, m, n, max_singular_values, max_truncation_error, relative):
    """Apply the keep rule to singular values padded with zeros to min(m, n) (the length the
    reference's full SVD of the m x n matrix reports) and pad u / vh with zero vectors for the
    zero singular values that are kept (they enter the outputs multiplied by sqrt(0) or 0)."""
    L = min(m, n)
    r = s.shape[0]
    sp = np.concatenate([s, np.zeros(max(L - r, 0))]) if L > r else s[:L]
    k = keep_count(sp, max_singular_values, max_truncation_error, relative)
    if k > r:
        u = np.concatenate([u, np.zeros((u.shape[0], k - r), dtype=u.dtype)], axis=1)
        vh = np.concatenate([vh, np.zeros((k - r, vh.shape[1]), dtype=vh.dtype)], axis=0)
    return u[:, :k], sp[:k], vh[:k, :], sp[k:]


def svd_product(A: np.ndarray, B: np.ndarray, max_singular_values=None, max_truncation_error=None,
                relative=False):
    """decompositions.svd of theta = A @ B (A: m x D, B: D x n) without forming theta: reduced QR
    of A and of B^H, exact SVD of the small core, factors multiplied back. Identical singular
    values and singular subspaces to the SVD of theta; pinned against the explicit product in
    the oracle tests."""
    m, n = A.shape[0], B.shape[1]
    QA, RA = np.linalg.qr(A, mode='reduced')
    QB, RB = np.linalg.qr(B.conj().T, mode='reduced')
    uc, sc, vhc = np.linalg.svd(RA @ RB.conj().T, full_matrices=False)
    return _keep_and_pad(QA @ uc, sc, vhc @ QB.conj().T, m, n, max_singular_values, max_truncation_error,
                         relative)


# ----------------------------------------------------------------------------------------------
# MPDO state and the engine (Circuit.py, TNNOptimizer.py)
# ----------------------------------------------------------------------------------------------
@dataclass
class MPDO:
    sites: List[np.ndarray]                      # T[l, s, a, r]
    has_I: List[bool]
    bond: List[bool]                             # bond[i]: edge bond_{i}_{i+1} exists
    discarded: List[Tuple[str, int, float]] = field(default_factory=list)  # (kind, index, sum s_rest^2)

    @property
    def n(self):
        return len(self.sites)

    def copy(self):
        return MPDO([t.copy() for t in self.sites], list(self.has_I), list(self.bond), list(self.discarded))


def product_state(n: int, vec=(1.0, 0.0), dtype=np.complex128) -> MPDO:
    """Tools.py:110-198: create_ket*Series — one (2,) node per qubit, no bonds, no inner index."""
    v = np.asarray(vec, dtype=dtype)
    return MPDO([v.reshape(1, 2, 1, 1).copy() for _ in range(n)], [False] * n, [False] * (n - 1))


def check_connectivity(st: MPDO) -> bool:
    """TNNOptimizer.py:51-61 (tn.check_connected on a chain: every bond present)."""
    return st.n <= 1 or all(st.bond)


def apply_1q(T: np.ndarray, E: np.ndarray) -> np.ndarray:
    """Circuit.py:138-178 for one qubit: contract ``inner`` with ``physics``; a Kraus tensor
    E[out, in, k] gives T'[l, s', (a, k), r] with the old inner index major (flatten_edges)."""
    if E.ndim == 2:
        return np.einsum('ij,ljar->liar', E, T)
    l, _, a, r = T.shape
    K = E.shape[2]
    return np.einsum('ijk,ljar->liakr', E, T).reshape(l, 2, a * K, r)


def compose_compressed(kraus_sets: Sequence[np.ndarray], rel_tol: float = 1e-14) -> np.ndarray:
    """One Kraus tensor (2, 2, K') for the channels ``kraus_sets`` applied in order (each
    (2,2[,K])), with K' the Choi rank of the composition: the stacked vec(E_k) rows (K x 4) are
    replaced by the rows of S V^H of their SVD. The new set is related to the sequential one
    (Circuit.py:174-178 flattens each new Kraus axis into I_i) by an isometry on the inner
    index, which no gauge-invariant quantity, kept count or discarded weight of the path sees.
    Test infrastructure for fixtures whose uncompressed inner index would not fit in memory;
    pinned by the channel action on random density matrices."""
    ops = [np.eye(2, dtype=np.complex128)]
    for E in kraus_sets:
        E = np.asarray(E, dtype=np.complex128)
        if E.ndim == 2:
            E = E[..., None]
        ops = [E[:, :, k] @ F for F in ops for k in range(E.shape[2])]
    rows = np.stack([F.reshape(4) for F in ops])                 # K x 4
    _, s, vh = np.linalg.svd(rows, full_matrices=False)
    keep = s > rel_tol * max(s[0], 1e-300)
    new = (s[keep][:, None] * vh[keep]).reshape(-1, 2, 2)
    return np.transpose(new, (1, 2, 0)).copy()


def _split_sqrt(theta: np.ndarray, **kw):
    """tn.split_node: U*sqrt(S) and sqrt(S)*V^H (tensornetwork 0.4.6 convention)."""
    u, s, vh, s_rest = svd(theta, **kw)
    sq = np.sqrt(np.maximum(s, 0))
    return u * sq[None, :], sq[:, None] * vh, s, s_rest


def _split_sqrt_product(A: np.ndarray, B: np.ndarray, explicit: bool = False, **kw):
    """_split_sqrt of theta = A @ B (rank-aware unless ``explicit``)."""
    if explicit:
        return _split_sqrt(A @ B, **kw)
    u, s, vh, s_rest = svd_product(A, B, **kw)
    sq = np.sqrt(np.maximum(s, 0))
    return u * sq[None, :], sq[:, None] * vh, s, s_rest


def apply_2q(Tmin: np.ndarray, Tmax: np.ndarray, G: np.ndarray, kraus_on_max: bool = True,
             svd_fail_hook=None, rng=None, explicit_theta: bool = False):
    """Circuit.py:74-136 for adjacent qubits (min, max): G[o_min, o_max, i_min, i_max(, k)].

    theta[(l, s1, a1), (s2, a2, k, r)], Kraus index flattened old-major into I_max (:109-114);
    split without chi cap, absolute tail GLOBAL_MINIMUM (:120-124), left U*sqrt(S), right
    sqrt(S)*V^H; on LinAlgError perturb by eps64*max(|theta|,1)*GLOBAL_MINIMUM*1e-2*(N+iN) and
    retry up to MAX_ATTEMPTS (:118-136). Returns (Tmin', Tmax', s_rest)."""
    l, _, a1, m = Tmin.shape
    m2, _, a2, r = Tmax.shape
    assert m == m2
    if G.ndim == 4:
        G = G[..., None]
    K = G.shape[4]
    # theta[(l x a), (y b k r)] = sum_{u v m} G[x y u v k] Tmin[l u a m] Tmax[m v b r]
    #   = Left[(l x a), (u m)] @ Right[(u m), (y b k r)] with Left = I_x (x) Tmin-part: write
    #   theta = sum_u Tmin[:, u] (l a m) . (G_{x y, u v k} Tmax[m v b r]) as one product of rank
    #   <= 2*2*m: Left[(l x a), (x' u m)] = delta_{x x'} Tmin[l u a m].
    Left = np.einsum('xz,luam->lxazum', np.eye(2), Tmin).reshape(l * 2 * a1, 2 * 2 * m)
    Right = np.einsum('xyuvk,mvbr->xumybkr', G, Tmax).reshape(2 * 2 * m, 2 * a2 * K * r)
    explicit = explicit_theta or (l * 2 * a1) * (2 * a2 * K * r) <= 4096
    rng = rng or np.random.default_rng(12345)
    for attempt in range(MAX_ATTEMPTS):
        try:
            if svd_fail_hook is not None:
                svd_fail_hook(attempt)
            us, svh, s, s_rest = _split_sqrt_product(Left, Right, explicit, max_truncation_error=GLOBAL_MINIMUM,
                                                     relative=False)
            break
        except np.linalg.LinAlgError:
            norm = float(np.linalg.norm(Left @ Right)) if explicit else float(np.linalg.norm(Left) * np.linalg.norm(Right))
            mag = np.finfo(np.float64).eps * max(norm, 1.0) * (GLOBAL_MINIMUM * 1e-2)
            Right = Right + mag * (rng.standard_normal(Right.shape) + 1j * rng.standard_normal(Right.shape))
    else:
        raise np.linalg.LinAlgError(f'SVD converges failed. Multiple times ({MAX_ATTEMPTS}) of perturbation are tried.')
    k = s.shape[0]
    new_min = us.reshape(l, 2, a1, k)
    new_max = svh.reshape(k, 2, a2 * K, r)
    return new_min, new_max, s_rest


def qr_left2right(st: MPDO) -> None:
    """TNNOptimizer.py:87-108: for i=0..n-2, QR of T_i as [(l s a) x r]; T_i <- Q, T_{i+1} <- R T_{i+1}."""
    if not check_connectivity(st):
        raise ValueError('qr_left2right needs every bond_{i}_{i+1} edge')
    for i in range(st.n - 1):
        T = st.sites[i]
        l, _, a, r = T.shape
        q, R = qr(T.reshape(l * 2 * a, r))
        k = q.shape[1]
        st.sites[i] = q.reshape(l, 2, a, k)
        nx = st.sites[i + 1]
        st.sites[i + 1] = np.tensordot(R, nx, axes=([1], [0]))


def _sweep_split(A, B, msv, err, svd_kw):
    """One sweep step: the split of theta = A @ B with relative=True. The randomized branch
    (svd_kw randomized=True) forms theta as the reference does; otherwise the exact SVD is
    computed rank-aware (explicitly for small theta)."""
    svd_kw = dict(svd_kw)
    explicit = svd_kw.pop('explicit_theta', False) or A.shape[0] * B.shape[1] <= 4096
    if svd_kw.get('randomized'):
        return _split_sqrt(A @ B, max_singular_values=msv, max_truncation_error=err, relative=True, **svd_kw)
    return _split_sqrt_product(A, B, explicit, max_singular_values=msv, max_truncation_error=err, relative=True)


def svd_right2left(st: MPDO, max_singular_values=None, max_truncation_err=None, **svd_kw) -> None:
    """TNNOptimizer.py:111-134: for idx=n-1..1, SVD of theta = T_{idx-1} T_idx with rows
    (l s a) of idx-1, cols (s a r) of idx, keep min(chi, k_err) (relative=True), split U*sqrt(S) /
    sqrt(S)*V^H. The discarded weight sum(s_rest^2) is recorded (the reference drops it)."""
    for idx in range(st.n - 1, 0, -1):
        A, B = st.sites[idx - 1], st.sites[idx]
        l, _, a1, m = A.shape
        _, _, a2, r = B.shape
        us, svh, s, s_rest = _sweep_split(A.reshape(l * 2 * a1, m), B.reshape(m, 2 * a2 * r),
                                          max_singular_values, max_truncation_err, svd_kw)
        k = s.shape[0]
        st.sites[idx - 1] = us.reshape(l, 2, a1, k)
        st.sites[idx] = svh.reshape(k, 2, a2, r)
        st.discarded.append(('chi', idx, float(np.sum(np.abs(s_rest) ** 2))))


def svd_left2right(st: MPDO, max_singular_values, max_truncation_err=None, **svd_kw) -> None:
    """TNNOptimizer.py:137-161: the same split swept idx=0..n-2 (exported, unused by evolve)."""
    for idx in range(st.n - 1):
        A, B = st.sites[idx], st.sites[idx + 1]
        l, _, a1, m = A.shape
        _, _, a2, r = B.shape
        us, svh, s, s_rest = _sweep_split(A.reshape(l * 2 * a1, m), B.reshape(m, 2 * a2 * r),
                                          max_singular_values, max_truncation_err, svd_kw)
        k = s.shape[0]
        st.sites[idx] = us.reshape(l, 2, a1, k)
        st.sites[idx + 1] = svh.reshape(k, 2, a2, r)
        st.discarded.append(('chi_l2r', idx, float(np.sum(np.abs(s_rest) ** 2))))


def bond_truncate(st: MPDO, max_singular_values=None, max_truncation_err=None, regularization=False, **svd_kw):
    """TNNOptimizer.py:72-84."""
    if max_singular_values is None and max_truncation_err is None and not regularization:
        return None
    qr_left2right(st)
    svd_right2left(st, max_singular_values, max_truncation_err, **svd_kw)


def kappa_us(M: np.ndarray, msv, err, svd_kw=None, gram_min_size: int = 4_000_000):
    """U and S of decompositions.svd(M) with relative=True (V^H is dropped by the caller,
    TNNOptimizer.py:186-191). Large wide matrices use the eigen-decomposition of M M^H (the
    same left singular vectors and s = sqrt(eigenvalues)); pinned against the SVD path."""
    svd_kw = svd_kw or {}
    m, n = M.shape
    if svd_kw.get('randomized') or M.size < gram_min_size or m > n:
        u, s, vh, s_rest = svd(M, max_singular_values=msv, max_truncation_error=err, relative=True, **svd_kw)
        return u, s, s_rest
    w, V = np.linalg.eigh(M @ M.conj().T)
    w, V = w[::-1], V[:, ::-1]
    s = np.sqrt(np.maximum(w, 0.0))
    u, s, _, s_rest = _keep_and_pad(V, s, np.zeros((m, 1), dtype=M.dtype), m, n, msv, err, True)
    return u, s, s_rest


def svd_kappa_left2right(st: MPDO, max_singular_values=None, max_truncation_err=None, **svd_kw) -> None:
    """TNNOptimizer.py:164-197: for each site with I_i and dim(I_i) > kappa: SVD of [(l s r) x a],
    keep min(kappa, k_err) (relative=True), T <- U*S (V^H dropped). Sites without an inner index
    are skipped (the reference raises KeyError there when kappa is None but an error is given)."""
    if max_singular_values is None and max_truncation_err is None:
        return None
    for i, T in enumerate(st.sites):
        if not st.has_I[i]:
            continue
        l, _, a, r = T.shape
        if max_singular_values is not None and a <= max_singular_values:
            continue
        M = np.transpose(T, (0, 1, 3, 2)).reshape(l * 2 * r, a)
        u, s, s_rest = kappa_us(M, max_singular_values, max_truncation_err, svd_kw)
        k = s.shape[0]
        st.sites[i] = np.transpose((u * s[None, :]).reshape(l, 2, r, k), (0, 1, 3, 2))
        st.discarded.append(('kappa', i, float(np.sum(np.abs(s_rest) ** 2))))


# ----------------------------------------------------------------------------------------------
# Circuit description, builder rules and evolve
# ----------------------------------------------------------------------------------------------
@dataclass
class Op:
    """One layer of a reference circuit (AbstractCircuit.py:138-141 ``_add_module``)."""
    name: str                        # gate name as the reference's gate.name (upper case) or
    qubits: List[int] = field(default_factory=list)   # 'TRUNCATE' / 'BARRIER'
    params: List[float] = field(default_factory=list)
    ideal: Optional[bool] = None     # the gate's ``_ideal`` flag (None = noisy when noise is on)
    matrix: Optional[np.ndarray] = None   # ArbSingle/ArbDouble/CZEXP/CPEXP tensors
    variational: bool = False


VARIATIONAL = {'RX', 'RY', 'RZ', 'RXX', 'RYY', 'RZZ', 'U1', 'U2', 'U3', 'P', 'CP', 'PSWAP', 'XXPLUSYY',
               'ARBDOUBLEGATE'}
TWO_QUBIT = {'CNOT', 'CX', 'CY', 'CZ', 'II', 'SWAP', 'ISWAP', 'CP', 'RXX', 'RYY', 'RZZ', 'PSWAP',
             'XXPLUSYY', 'ARBDOUBLEGATE', 'CZEXP', 'CPEXP'}


class CircuitBuilder:
    """Restates the builder rules of AbstractCircuit.py:156-532 that change the layer list:
    under realNoise, cx/cnot -> ry(-pi/2, t, ideal) cz ry(pi/2, t, ideal) (:291-301,:332-342),
    cz -> CZEXP with the tomographed tensor (:313-330), cp -> CPEXP (:404-422), rxx/ryy/rzz
    decomposed (:231-275)."""

    def __init__(self, qn: int, noise_type: str = 'no', exp_tensors: Optional[Dict[str, np.ndarray]] = None):
        self.qn = qn
        self.real = noise_type.lower() == 'realnoise'
        self.exp = exp_tensors or {}
        self.ops: List[Op] = []

    def add(self, name, qubits, params=(), ideal=None, matrix=None):
        name = name.upper()
        q = [qubits] if isinstance(qubits, int) else list(qubits)
        p = list(params)
        if self.real and not ideal and name in ('CNOT', 'CX'):
            self.add('RY', [q[1]], [-math.pi / 2], True)
            self.add('CZ', q)
            self.add('RY', [q[1]], [math.pi / 2], True)
            return
        if self.real and not ideal and name == 'CZ':
            t = self.exp.get(f'{q[0]}{q[1]}', self.exp.get(f'{q[1]}{q[0]}'))
            self.ops.append(Op('CZEXP', q, [], False, t))
            return
        if self.real and not ideal and name == 'CP':
            t = self.exp.get('CP' + f'{q[0]}{q[1]}', self.exp.get('CP' + f'{q[1]}{q[0]}'))
            self.ops.append(Op('CPEXP', q, [], False, t))
            return
        if self.real and not ideal and name == 'RXX':
            self.add('H', q, (), True); self.add('CX', q); self.add('RZ', [q[1]], p, True)
            self.add('CX', q); self.add('H', q, (), True)
            return
        if self.real and not ideal and name == 'RYY':
            self.add('RX', q, [math.pi / 2], True); self.add('CX', q); self.add('RZ', [q[1]], p, True)
            self.add('CX', q); self.add('RX', q, [-math.pi / 2], True)
            return
        if self.real and not ideal and name == 'RZZ':
            self.add('CX', q); self.add('RZ', [q[1]], p, True); self.add('CX', q)
            return
        self.ops.append(Op(name, q, p, ideal, matrix, name in VARIATIONAL))

    def truncate(self):
        self.ops.append(Op('TRUNCATE'))

    def barrier(self):
        self.ops.append(Op('BARRIER'))


def _adjacent(q):
    if abs(q[0] - q[1]) != 1:
        raise NotImplementedError('only nearest-neighbour two-qubit gates are supported')


def evolve(ops: Sequence[Op], st: MPDO, noise_type: str = 'no', ideal: bool = True,
           noise: Optional[NoiseSpec] = None, chi=None, kappa=None, max_truncation_err=None,
           svd_kw: Optional[dict] = None, cache: Optional[dict] = None) -> MPDO:
    """Circuit.py:469-491 ``evolve`` with ``_add_gate`` (:180-224) and the 1q/2q application
    rules (:74-178). The state is updated in place and returned."""
    svd_kw = svd_kw or {}
    nt = noise_type.lower()
    if not ideal and nt not in ('unified', 'realnoise', 'idealnoise'):
        raise ValueError(f'Unsupported noise type: {nt}')
    unified, real, idealn = (not ideal and nt == 'unified'), (not ideal and nt == 'realnoise'), \
        (not ideal and nt == 'idealnoise')
    noise = noise or NoiseSpec()
    cache = {} if cache is None else cache        # noiseTensorDict, keyed by gate.name (:92-96,:145-151)
    n = st.n
    for op in ops:
        if op.name == 'TRUNCATE':
            if check_connectivity(st):
                bond_truncate(st, chi, max_truncation_err, **svd_kw)
                if not ideal:
                    svd_kappa_left2right(st, kappa, max_truncation_err, **svd_kw)
            continue
        if op.name == 'BARRIER':
            continue
        if max(op.qubits) >= n:
            raise ValueError(f'Qubit index out of range, max index is Q{max(op.qubits)}.')
        if op.name == 'MEASUREZ':
            continue                                  # stopping condition (:197-198)
        if op.matrix is not None:
            G = np.asarray(op.matrix, dtype=np.complex128)
        else:
            G = gate_matrix(op.name, op.params)
        if op.name not in TWO_QUBIT and (op.name != 'ARBDOUBLEGATE'):
            gnoise = (idealn or unified) and not op.ideal
            if gnoise:
                if op.variational:
                    E = noise.single_kraus(G)
                else:
                    E = cache.setdefault(op.name, noise.single_kraus(G))
            else:
                E = G
            for q in op.qubits:
                st.sites[q] = apply_1q(st.sites[q], E)
                if gnoise:
                    st.has_I[q] = True
            continue
        q = op.qubits
        if len(q) != 2 or q[0] == q[1]:
            raise ValueError('Invalid operating qubits for a two-qubit gate.')
        _adjacent(q)
        gnoise = (idealn and not op.ideal) or real
        if gnoise and not real:
            if op.variational:
                G = noise.double_kraus(G)
            else:
                G = cache.setdefault(op.name, noise.double_kraus(G))
        if real and G.ndim != 5:
            raise ValueError('realNoise two-qubit gates must carry a Kraus axis (CZEXP/CPEXP)')
        if q[0] > q[1]:                               # reorder G to (min, max)
            G = np.swapaxes(np.swapaxes(G, 0, 1), 2, 3)
        mn = min(q)
        a, b, _ = apply_2q(st.sites[mn], st.sites[mn + 1], G, True)
        st.sites[mn], st.sites[mn + 1] = a, b
        st.bond[mn] = True
        if G.ndim == 5:
            st.has_I[mn + 1] = True
    if not ideal and ops and ops[-1].name != 'TRUNCATE':
        svd_kappa_left2right(st, kappa, max_truncation_err, **svd_kw)   # :487-489
    return st


# ----------------------------------------------------------------------------------------------
# Observables (Circuit.py:226-295, dmOperations.py)
# ----------------------------------------------------------------------------------------------
def density_matrix(st: MPDO, reduced_index: Sequence[int] = ()) -> np.ndarray:
    """Circuit.py:226-262 + dmOperations.py:18-37: rho~ = sum_a T_a conj(T_a), the qubits in
    ``reduced_index`` traced, rows = ket bits of the kept qubits (lowest index most significant),
    unnormalised."""
    red = set(reduced_index)
    keep = [i for i in range(st.n) if i not in red]
    # MPO of rho with bond (l, l'): W[(l l'), s, s', (r r')]
    env = np.ones((1, 1, 1, 1), dtype=np.complex128)    # [ket_bits, bra_bits, l, l']
    kb = 1
    for i, T in enumerate(st.sites):
        W = np.einsum('lsar,mtan->lmstrn', T, T.conj())  # (l, l', s, s', r, r')
        if i in red:
            W = np.einsum('lmssrn->lmrn', W)
            env = np.einsum('abln,lnrm->abrm', env, W)
        else:
            env = np.einsum('abln,lnstrm->asbtrm', env, W)
            kb *= 2
            env = env.reshape(kb, kb, W.shape[4], W.shape[5])
    return env[:, :, 0, 0]


def expect(st: MPDO, observables, oqs) -> List[float]:
    """dmOperations.py:135-171: Re Tr(O rho~) per observable; O acts on qubits ``oq`` in the
    given order (first listed = most significant), other qubits traced."""
    obs = [observables] if isinstance(observables, np.ndarray) else list(observables)
    oqs = [oqs] if isinstance(oqs, int) else list(oqs)
    out = []
    for O, oq in zip(obs, oqs):
        oq = [oq] if isinstance(oq, int) else list(oq)
        k = len(oq)
        O = np.asarray(O, dtype=np.complex128).reshape(2 ** k, 2 ** k)
        rho = density_matrix(st, [i for i in range(st.n) if i not in oq])
        srt = sorted(oq)
        # rho is in sorted-qubit order; permute O into it
        perm = [oq.index(q) for q in srt]
        Ot = O.reshape([2] * (2 * k)).transpose(perm + [p + k for p in perm]).reshape(2 ** k, 2 ** k)
        out.append(float(np.real(np.trace(Ot @ rho))))
    return out


def pauli_expect(st: MPDO, observables, oqs) -> float:
    """dmOperations.py:174-201: Re Tr((P_1 x ... x P_k) rho~), 0=X, 1=Y, 2=Z (:11-15)."""
    obs = [observables] if isinstance(observables, int) else list(observables)
    oqs = [oqs] if isinstance(oqs, int) else list(oqs)
    env = np.ones((1, 1), dtype=np.complex128)
    for i, T in enumerate(st.sites):
        P = np.eye(2, dtype=np.complex128)
        if i in oqs:
            P = PAULIS[1 + obs[oqs.index(i)]]
        env = np.einsum('ln,lsar,ts,ntam->rm', env, T, P, T.conj(), optimize=True)
    return float(np.real(env[0, 0]))


def bitstring_probs(st: MPDO, bits: np.ndarray) -> np.ndarray:
    """Unnormalised P(b) = Re Tr(|b><b| rho~) for each row b of ``bits`` (shape (count, n), 0/1),
    the projector contraction of Circuit.py:297-332 (_conditional_prb with every qubit sampled in
    the computational basis). All bitstrings advance together: env[c, l, l'] picks the physical
    slice bits[c, i] of each site. Pinned against the diagonal of the dense density matrix."""
    bits = np.asarray(bits, dtype=np.int64)
    cnt = bits.shape[0]
    env = np.ones((cnt, 1, 1), dtype=np.complex128)
    for i, T in enumerate(st.sites):
        l, _, a, r = T.shape
        nxt = np.empty((cnt, r, r), dtype=np.complex128)
        for s in (0, 1):
            sel = np.nonzero(bits[:, i] == s)[0]
            if sel.size == 0:
                continue
            Ts = T[:, s].reshape(l, a * r)                       # [l, (a r)]
            X = np.einsum('cln,lk->cnk', env[sel], Ts)           # [c, l', (a r)]
            X = X.reshape(sel.size, l, a, r)
            nxt[sel] = np.einsum('cnar,nam->crm', X, T[:, s].conj())
        env = nxt
    return env[:, 0, 0].real.copy()


def trace_rho(st: MPDO) -> float:
    """dmOperations.py:40-48."""
    env = np.ones((1, 1), dtype=np.complex128)
    for T in st.sites:
        env = np.einsum('ln,lsar,nsam->rm', env, T, T.conj(), optimize=True)
    return float(np.real(env[0, 0]))


def trace_rho_rho(st0: MPDO, st1: Optional[MPDO] = None, reduced: Sequence[int] = ()) -> float:
    """dmOperations.py:51-66: Re Tr(rho0 rho1); the qubits in ``reduced`` are traced within each
    state first (their physical edges are no longer dangling, so the reference leaves them)."""
    st1 = st0 if st1 is None else st1
    red = set(reduced)
    env = np.ones((1, 1, 1, 1), dtype=np.complex128)      # [ket0, bra0, ket1, bra1]
    o = dict(optimize=True)
    for i, (A, B) in enumerate(zip(st0.sites, st1.sites)):
        if i in red:     # each state's physical index meets its own bra
            t = np.einsum('abcd,asxe->bcdsxe', env, A, **o)
            t = np.einsum('bcdsxe,bsxf->cdef', t, A.conj(), **o)
            t = np.einsum('cdef,ctyg->deftyg', t, B, **o)
            env = np.einsum('deftyg,dtyh->efgh', t, B.conj(), **o)
        else:            # rho0[s, s'] rho1[s', s]: ket0 s with bra1 s, bra0 s' with ket1 s'
            t = np.einsum('abcd,asxe->bcdsxe', env, A, **o)
            t = np.einsum('bcdsxe,dsyh->bcxeyh', t, B.conj(), **o)
            t = np.einsum('bcxeyh,btxf->ceyhtf', t, A.conj(), **o)
            env = np.einsum('ceyhtf,ctyg->efgh', t, B, **o)
    return float(np.real(env[0, 0, 0, 0]))


def trace_composited_rho2(states: Sequence[MPDO], reduced: Sequence[int] = ()) -> float:
    """dmOperations.py:90-132: Tr(rho^2) of rho = sum_i rho_i / S (diagonal and cross terms)."""
    S = len(states)
    tot = sum(trace_rho_rho(states[i], states[j], reduced) for i in range(S) for j in range(S))
    return tot / S ** 2


# ----------------------------------------------------------------------------------------------
# Shot sampling (Circuit.py:297-467). The draws use torch's default CPU generator with the
# reference's own calls, so a seeded run gives the reference's bitstrings for the same
# probabilities.
# ----------------------------------------------------------------------------------------------
MEASURE_ROTATIONS = {0: gate_matrix('H'),                                       # SingleGates.py:240-271
                     1: np.array([[_S2, -1j * _S2], [_S2, 1j * _S2]])}          # SingleGates.py:274-305


def sampling_state(st: MPDO, orientation: Sequence[int]) -> MPDO:
    """Circuit.py:413-419: a copy of the state with MeasureX (orientation 0) and MeasureY (1)
    applied as ideal gates. The reference passes the positions in ``orientation`` as the qubit
    indices (``_indices`` enumerates the orientation list), so with qubits reduced the rotation
    lands on qubit ``position``; that behaviour is kept."""
    out = st.copy()
    for pos, o in enumerate(orientation):
        if o in MEASURE_ROTATIONS:
            out.sites[pos] = apply_1q(out.sites[pos], MEASURE_ROTATIONS[o])
    return out


def conditional_probs(st: MPDO, order: Sequence[int], history: Sequence[int]) -> np.ndarray:
    """Circuit.py:297-326: the unnormalised diagonal the reference reads off its contraction for
    qubit ``order[len(history)]``: qubits ``order[:k]`` projected on ``history``, the open qubit
    kept, every other qubit traced (``reduce_dmNodes`` on the reduced and later qubits)."""
    k = len(history)
    proj = {q: int(h) for q, h in zip(order[:k], history)}
    open_q = order[k]
    env = np.ones((1, 1, 1), dtype=np.complex128)            # [s_open, l, l']
    for i, T in enumerate(st.sites):
        if i in proj:
            Ts = T[:, proj[i]]
            env = np.einsum('xln,lar,nam->xrm', env, Ts, Ts.conj(), optimize=True)
        elif i == open_q:
            env = np.einsum('ln,lsar,nsam->srm', env[0], T, T.conj(), optimize=True)
        else:
            env = np.einsum('xln,lsar,nsam->xrm', env, T, T.conj(), optimize=True)
    return np.real(env[:, 0, 0])


def normalise_prb(P: Sequence[float]) -> float:
    """Circuit.py:328-332 in the arithmetic of a complex64 state: real parts in float32 plus
    GLOBAL_MINIMUM, RuntimeError on a negative entry, the normalised last entry."""
    import torch
    probs = torch.tensor([float(P[0]), float(P[1])], dtype=torch.float32) + GLOBAL_MINIMUM
    if (probs < 0).sum() > 0:
        raise RuntimeError(f"State is illegal, and your probability distribution is {probs}.")
    return (probs / probs.sum())[-1].item()


def conditional_batch_sample(st: MPDO, order: Sequence[int], shots: int, bool_result: bool = False,
                             trace: Optional[list] = None) -> list:
    """Circuit.py:355-391: groups of shots sharing a history; per group one Bernoulli count
    (skipped when p1 is within GLOBAL_MINIMUM of 0 or 1), the ones first."""
    import torch
    L = len(order)
    seq = torch.zeros((shots, L), dtype=torch.bool if bool_result else torch.int)
    val = True if bool_result else 1
    groups = [([], 0, shots)]
    for j in range(L):
        new = []
        for hist, start, length in groups:
            p1 = normalise_prb(conditional_probs(st, order, hist))
            if trace is not None:
                trace.append((tuple(hist), p1))
            if p1 < GLOBAL_MINIMUM:
                k1, k0 = 0, length
            elif p1 > 1 - GLOBAL_MINIMUM:
                k1, k0 = length, 0
            else:
                k1 = int(torch.bernoulli(torch.full((length,), p1)).sum())
                k0 = length - k1
            seq[start:start + k1, j] = val
            if j < L - 1:
                if k1 > 0:
                    new.append((hist + [1], start, k1))
                if k0 > 0:
                    new.append((hist + [0], start + k1, k0))
        groups = new
    return seq.tolist()


def conditional_sample(st: MPDO, order: Sequence[int], shots: int) -> list:
    """Circuit.py:334-342: shot by shot, qubit by qubit, one two-outcome multinomial draw."""
    import torch
    out = []
    for _ in range(shots):
        ch = [-1] * len(order)
        for j in range(len(order)):
            p = normalise_prb(conditional_probs(st, order, ch[:j]))
            ch[j] = torch.multinomial(torch.tensor([1 - p, p]), num_samples=1).item()
        out.append(ch)
    return out


def count_item(data) -> dict:
    """Tools.py:554-570: occurrence counts sorted by the bitstring's value."""
    from collections import Counter

    def key(x):
        if isinstance(x[0], tuple):
            return int(''.join(str(int(v)) for v in x[0]))
        if isinstance(x[0], str):
            return x[0]
        return int(x[0])
    counted = dict(Counter(tuple(i) if isinstance(i, list) else i for i in data))
    return dict(sorted(counted.items(), key=key))


def sample(st: MPDO, shots: Optional[int] = None, orientation: Optional[Sequence[int]] = None,
           reduced: Optional[Sequence[int]] = None, sample_string: bool = True, sequential: bool = False,
           bool_result: bool = False, counts: bool = True):
    """Circuit.py:393-445 (TensorCircuit.sample)."""
    shots = shots or 1024
    reduced = list(reduced or [])
    order = [q for q in range(st.n) if q not in reduced]
    orientation = list(orientation or [2] * len(order))
    if len(orientation) != len(order):
        raise ValueError('Length of orientation must match the sample length. Check reduced or unmeasured qubits.')
    rot = sampling_state(st, orientation)
    if sequential:
        bits = conditional_sample(rot, order, shots)
    else:
        bits = conditional_batch_sample(rot, order, shots, bool_result)
    if sample_string and not bool_result:
        bits = [''.join(map(str, b)) for b in bits]
    return (bits, count_item(bits)) if counts else bits
