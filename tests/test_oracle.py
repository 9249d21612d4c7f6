"""Pins of the CPU oracle (runs without a GPU). Every check compares the oracle with something the
mathematics fixes independently of its code: matrix exponentials, Bloch-vector actions of the
channels, CPTP identities, the recorded known answers of the reference's tomographed CZ data,
hand-computed keep counts, an exact dense density-matrix simulator, Schmidt decompositions of
dense state vectors and dense eigen-decompositions of single-site density matrices."""
import math
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import dense as D
from oracle import mpdo_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)
P1 = [I2, X, Y, Z]
P2 = [np.kron(a, b) for a in P1 for b in P1]


def _rand_rho(d, rng):
    A = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    r = A @ A.conj().T
    return r / np.trace(r)


def _channel(E, rho):
    """rho -> sum_k E_k rho E_k^H for a (d, d[, K]) Kraus tensor in (out, in, k) layout."""
    if E.ndim == 2:
        return E @ rho @ E.conj().T
    return sum(E[..., k] @ rho @ E[..., k].conj().T for k in range(E.shape[-1]))


# ---------------------------------------------------------------------------------- gates
@pytest.mark.parametrize('theta', [0.0, 0.37, -1.9, math.pi, 5.5])
def test_rotations_are_exponentials(theta):
    assert np.allclose(O.gate_matrix('RX', [theta]), expm(-0.5j * theta * X))
    assert np.allclose(O.gate_matrix('RY', [theta]), expm(-0.5j * theta * Y))
    assert np.allclose(O.gate_matrix('RZ', [theta]), expm(-0.5j * theta * Z))
    assert np.allclose(O.gate_matrix('RZZ', [theta]).reshape(4, 4), expm(-0.5j * theta * np.kron(Z, Z)))
    # PhaseGates.py:154-190 / DoubleGates.py:155-190,229-262
    assert np.allclose(O.gate_matrix('CP', [theta]).reshape(4, 4), expm(1j * theta * np.diag([0, 0, 0, 1])))
    xy = (np.kron(X, X) + np.kron(Y, Y)) / 2
    assert np.allclose(O.gate_matrix('XXPLUSYY', [theta, 0.0]).reshape(4, 4), expm(-0.5j * theta * xy))
    swap_gen = np.array([[0, 0, 0, 0], [0, 0, -1, 0], [0, 1, 0, 0], [0, 0, 0, 0]], dtype=complex)
    assert np.allclose(O.gate_matrix('PSWAP', [theta]).reshape(4, 4) @ np.diag([1, 1, 1, 1]),
                       np.eye(4) + (np.cos(theta) - 1) * np.diag([0, 1, 1, 0]) + np.sin(theta) * np.array(
                           [[0, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 0]]))
    assert np.allclose(O.gate_matrix('U3', [theta, 0.4, -0.7]),
                       np.diag([1, np.exp(0.4j)]) @ O.gate_matrix('RY', [theta]) @ np.diag([1, np.exp(-0.7j)]))
    assert np.allclose(O.gate_matrix('RXX', [theta]).reshape(4, 4), expm(-0.5j * theta * np.kron(X, X)))
    # YGates.py:138-143 builds [[c,0,0,-s],[0,c,s,0],[0,-s,c,0],[s,0,0,c]]: the XY rotation
    assert np.allclose(O.gate_matrix('RYY', [theta]).reshape(4, 4), expm(-0.5j * theta * np.kron(X, Y)))


def test_fixed_gates():
    P0, P1_ = np.diag([1, 0]).astype(complex), np.diag([0, 1]).astype(complex)
    assert np.allclose(O.gate_matrix('CNOT').reshape(4, 4), np.kron(P0, I2) + np.kron(P1_, X))
    assert np.allclose(O.gate_matrix('CZ').reshape(4, 4), np.kron(P0, I2) + np.kron(P1_, Z))
    assert np.allclose(O.gate_matrix('H'), (X + Z) / math.sqrt(2))
    for nm, M in (('X', X), ('Y', Y), ('Z', Z), ('I', I2)):
        assert np.allclose(O.gate_matrix(nm), M)
    sw = O.gate_matrix('SWAP').reshape(4, 4)
    for a in P1:
        for b in P1:
            assert np.allclose(sw @ np.kron(a, b) @ sw, np.kron(b, a))
    for nm in ('CNOT', 'CZ', 'SWAP', 'ISWAP', 'II'):
        U = O.gate_matrix(nm).reshape(4, 4)
        assert np.allclose(U @ U.conj().T, np.eye(4))


# ---------------------------------------------------------------------------------- noise
@pytest.mark.parametrize('p', [0.0, 0.002, 0.05, 0.3])
def test_depolarizing_shrinks_bloch_vector(p):
    rng = np.random.default_rng(1)
    E = O.depolarizing_tensor(p, 1)
    for _ in range(3):
        rho = _rand_rho(2, rng)
        out = _channel(E, rho)
        assert np.allclose(out, (1 - p) * rho + p * np.trace(rho) * I2 / 2)
    E2 = O.depolarizing_tensor(p, 2).reshape(4, 4, 16)
    rho = _rand_rho(4, rng)
    assert np.allclose(_channel(E2, rho), (1 - p) * rho + p * np.trace(rho) * np.eye(4) / 4)


@pytest.mark.parametrize('g', [0.0, 0.1, 0.7])
def test_decay_and_dephasing_actions(g):
    rng = np.random.default_rng(2)
    rho = _rand_rho(2, rng)
    out = _channel(O.decay_tensor(g), rho)          # amplitude damping
    assert np.isclose(out[1, 1], (1 - g) * rho[1, 1])
    assert np.isclose(out[0, 0], rho[0, 0] + g * rho[1, 1])
    assert np.isclose(out[0, 1], math.sqrt(1 - g) * rho[0, 1])
    out = _channel(O.dephasing_tensor(g), rho)
    assert np.allclose(np.diag(out), np.diag(rho))
    assert np.isclose(out[0, 1], (1 - g) * rho[0, 1])


def test_channels_trace_preserving():
    for E in (O.depolarizing_tensor(0.1, 1), O.decay_tensor(0.2), O.dephasing_tensor(0.3),
              O.amp_phase_damping_tensor(30, 2e2, 2e1)):
        assert np.allclose(np.einsum('ijk,ilk->jl', E.conj(), E), I2)
    E2 = O.depolarizing_tensor(0.1, 2).reshape(4, 4, 16)
    assert np.allclose(np.einsum('ijk,ilk->jl', E2.conj(), E2), np.eye(4))


def test_noisy_gate_kraus_composition():
    """single_kraus: the gate first, then the noise channel (Circuit.py:144-155)."""
    rng = np.random.default_rng(3)
    ns = O.NoiseSpec(chip='best', dpc_error_rate=0.02, single='depolarizing', p1=0.04)
    G = O.gate_matrix('RY', [0.8])
    rho = _rand_rho(2, rng)
    want = (1 - 0.04) * G @ rho @ G.conj().T + 0.04 * I2 / 2
    assert np.allclose(_channel(ns.single_kraus(G), rho), want)
    G2 = O.gate_matrix('CNOT')
    rho4 = _rand_rho(4, rng)
    U = G2.reshape(4, 4)
    want = (1 - 0.02) * U @ rho4 @ U.conj().T + 0.02 * np.eye(4) / 4
    assert np.allclose(_channel(ns.double_kraus(G2).reshape(4, 4, -1), rho4), want)


def test_tomographed_cz_known_answers():
    """The two chi matrices shipped with the reference: ideal_cz is exactly CZ (one Kraus
    operator, up to phase); czDefault has 16 operators whose sum E^H E has the recorded
    diagonal (a trace-non-preserving tomographic estimate)."""
    E = O.chi_to_kraus(O.load_chi(os.path.join(HERE, 'golden', 'ideal_cz.mat')))
    assert E.shape[-1] == 1
    M = E[..., 0].reshape(4, 4)
    cz = np.diag([1, 1, 1, -1]).astype(complex)
    ph = np.vdot(cz.ravel(), M.ravel()) / 4
    assert abs(abs(ph) - 1) < 1e-6 and np.allclose(M, ph * cz, atol=1e-6)
    E = O.chi_to_kraus(O.load_chi(os.path.join(HERE, 'golden', 'czDefault.mat')))
    assert E.shape[-1] == 16
    S = np.einsum('ijk,ilk->jl', E.reshape(4, 4, 16).conj(), E.reshape(4, 4, 16))
    assert np.allclose(np.real(np.diag(S)), [1.018895, 0.995842, 1.001604, 0.983660], atol=5e-6)


# ---------------------------------------------------------------------------------- keep rule
# te = sqrt(cumsum([9, 4, 1, 0.25])) = [3, 3.6056, 3.7417, 3.7749]; te[-1] - te[idx] = 0.7749, 0.1694, 0.0333
@pytest.mark.parametrize('s,msv,err,rel,want', [
    ([3.0, 2.0, 1.0, 0.5], None, None, False, 4),
    ([3.0, 2.0, 1.0, 0.5], 2, None, False, 2),
    ([3.0, 2.0, 1.0, 0.5], 9, None, False, 4),          # capped at len(s)
    ([3.0, 2.0, 1.0, 0.5], None, 0.6, False, 2),        # 0.1694 <= 0.6 at idx 1
    ([3.0, 2.0, 1.0, 0.5], None, 1.2, False, 1),        # 0.7749 <= 1.2 at idx 0
    ([3.0, 2.0, 1.0, 0.5], None, 0.4, True, 1),         # bound 0.4 * 3 = 1.2
    ([3.0, 2.0, 1.0, 0.5], None, 0.1, True, 2),         # bound 0.3
    ([3.0, 2.0, 1.0, 0.5], 1, 0.6, False, 1),           # min(chi, k_err)
    ([3.0, 2.0, 1.0, 0.5], None, 0.1, False, 3),        # 0.0333 <= 0.1 at idx 2
    ([3.0, 2.0, 1.0, 0.5], None, 0.01, False, 4),       # nothing within the bound
    ([1.0, 1.0, 1.0], None, 0.5, False, 2),             # te = [1, 1.4142, 1.7321]: 0.3178 at idx 1
    ([2.0], None, 10.0, False, 1),                      # the last index never qualifies
])
def test_keep_count_hand_cases(s, msv, err, rel, want):
    """decompositions.py:117-140 (rule at :121-131): keep idx+1 for the first idx < len-1 with
    sqrt(cumsum s^2)[-1] - sqrt(cumsum s^2)[idx] <= err (err * s[0] when relative)."""
    assert O.keep_count(np.asarray(s), msv, err, rel) == want


@pytest.mark.parametrize('m,n', [(120, 80), (60, 90)])   # n_iter 'auto': 3 power steps if m >= n, else 2
def test_randomized_svd_closed_forms(m, n):
    """decompositions.py:23-48 restated (the Gaussian draw is the oracle's own, not torch's):
    (1) for M of rank 6 <= n_components + 7 the sketch spans M's range, so the singular values
    and the reconstruction u s vh are exact; (2) on a full-rank spectrum 0.7^j the power steps
    Q <- M (M^H Q) resolve the 10 leading singular values to the accuracy the sketch guarantees
    ((sigma_18 / sigma_10)^(2q+1) for the range, squared for the values). A transposed Q in
    B = Q^H M breaks (1); a missing conjugate in M^H leaves errors near 1e-5 in (2)."""
    rng = np.random.default_rng(3 + m)
    X = (rng.standard_normal((m, 6)) + 1j * rng.standard_normal((m, 6))) @ \
        (rng.standard_normal((6, n)) + 1j * rng.standard_normal((6, n)))
    u, s, vh = O.randomized_svd(X, 4)
    ref = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(s[:6], ref[:6], rtol=1e-10)
    assert np.abs(s[6:]).max() <= 1e-10 * ref[0]
    assert np.abs((u * s) @ vh - X).max() <= 1e-10 * ref[0]
    assert np.abs(u.conj().T @ u - np.eye(u.shape[1])).max() <= 1e-10
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
    sig = 0.7 ** np.arange(r)
    u, s, vh = O.randomized_svd((U * sig) @ V.conj().T, 10)
    assert np.abs(s[:10] - sig[:10]).max() <= 1e-8


@pytest.mark.parametrize('m,n,rank', [(30, 50, 7), (64, 20, 20), (12, 12, 3)])
def test_svd_product_matches_explicit_svd(m, n, rank):
    rng = np.random.default_rng(m * n)
    A = rng.standard_normal((m, rank)) + 1j * rng.standard_normal((m, rank))
    B = rng.standard_normal((rank, n)) + 1j * rng.standard_normal((rank, n))
    # bonus zero columns in the middle dimension: a rank-deficient product
    A2 = np.concatenate([A, np.zeros((m, 3))], axis=1)
    B2 = np.concatenate([B, rng.standard_normal((3, n))], axis=0)
    for msv in (None, 2, rank + 2):
        u, s, vh, rest = O.svd_product(A2, B2, max_singular_values=msv)
        s_ref = np.linalg.svd(A @ B, compute_uv=False)
        k = min(msv or min(m, n), min(m, n))
        assert u.shape[1] == k and vh.shape[0] == k
        sp = np.concatenate([s_ref, np.zeros(min(m, n) - len(s_ref))])
        assert np.allclose(s, sp[:k], atol=1e-9)
        assert np.isclose(np.sum(np.abs(rest) ** 2), np.sum(sp[k:] ** 2))
        # reconstruction of the truncated product
        U, S, VH = np.linalg.svd(A @ B)
        kk = min(k, len(S))
        assert np.allclose((u * s) @ vh, (U[:, :kk] * S[:kk]) @ VH[:kk], atol=1e-9)


def test_kappa_gram_path_matches_svd():
    rng = np.random.default_rng(5)
    M = rng.standard_normal((40, 300)) + 1j * rng.standard_normal((40, 300))
    u1, s1, r1 = O.kappa_us(M, 6, None, gram_min_size=1)
    u2, s2, r2 = O.kappa_us(M, 6, None, gram_min_size=10 ** 9)
    assert np.allclose(s1, s2) and np.isclose(np.sum(r1 ** 2), np.sum(r2 ** 2))
    assert np.allclose(u1 @ u1.conj().T, u2 @ u2.conj().T)


# ---------------------------------------------------------------------------------- evolution
def _random_ops(n, depth, seed, two='CNOT'):
    """Brickwork RY (noisy) RZ (ideal) + CNOT/CZ; without a kappa cut every noisy gate multiplies
    an inner dimension, so the exact-regime circuits stay shallow."""
    rng = np.random.default_rng(seed)
    b = O.CircuitBuilder(n, 'idealNoise')
    for d in range(depth):
        for q in range(n):
            b.add('RY', [q], [float(rng.uniform(0, 6.28))])
            b.add('RZ', [q], [float(rng.uniform(0, 6.28))], ideal=True)
        for a in range(d % 2, n - 1, 2):
            b.add(two if d % 2 == 0 else 'CZ', [a, a + 1])
        b.truncate()
    return b.ops


@pytest.mark.parametrize('noisy', [False, True])
@pytest.mark.parametrize('n', [3, 4])
def test_evolution_matches_dense_simulator(n, noisy):
    """No truncation active (chi, kappa None): the MPDO must equal the dense evolution."""
    ops = _random_ops(n, 2, 11 + n)
    ns = O.NoiseSpec(chip='best', dpc_error_rate=0.05, single='depolarizing', p1=0.03)
    st = O.evolve(ops, O.product_state(n), 'idealNoise', not noisy, ns)
    rho = D.dense_evolve(ops, n, 'idealNoise', not noisy, ns)
    assert np.allclose(O.density_matrix(st), rho, atol=1e-12)
    for q in range(n):
        assert np.isclose(O.pauli_expect(st, [2], [q]), D.pauli_string_expect(rho, [2], [q], n))
    assert np.isclose(O.trace_rho(st), np.real(np.trace(rho)))
    assert np.isclose(O.trace_rho_rho(st), np.real(np.trace(rho @ rho)))
    red = O.density_matrix(st, [0])
    assert np.allclose(red, D.partial_trace(rho, list(range(1, n)), n), atol=1e-12)


def test_chi_cut_is_the_schmidt_truncation_of_a_pure_state():
    """Ideal 2-qubit... n-qubit pure state: after the QR sweep the first step of the chi sweep cuts
    bond (n-2 | n-1) at the Schmidt values of the state there."""
    n = 4
    rng = np.random.default_rng(7)
    ops = _random_ops(n, 4, 21)
    st = O.evolve([o for o in ops if o.name != 'TRUNCATE'], O.product_state(n), 'no', True)
    psi = np.einsum('i,j,k,l->ijkl', *[np.ones(1)] * 4)  # placeholder shape, replaced below
    rho = O.density_matrix(st)
    w, V = np.linalg.eigh(rho)
    psi = V[:, -1] * math.sqrt(max(w[-1], 0))
    S = np.linalg.svd(psi.reshape(2 ** (n - 1), 2), compute_uv=False)
    st2 = st.copy()
    O.qr_left2right(st2)
    A = st2.sites[n - 2]
    B = st2.sites[n - 1]
    l, _, a1, m = A.shape
    u, s, vh, rest = O.svd_product(A.reshape(l * 2 * a1, m), B.reshape(m, -1), max_singular_values=1)
    assert np.isclose(s[0], S[0]) and np.isclose(np.sum(np.abs(rest) ** 2), S[1] ** 2)
    # chi = 1 everywhere: the state after bond_truncate is a product along that cut
    st3 = st.copy()
    O.bond_truncate(st3, 1)
    chi_steps = [d for d in st3.discarded if d[0] == 'chi']
    assert np.isclose(chi_steps[0][2], S[1] ** 2)


@pytest.mark.parametrize('seed', [31, 32, 33])
def test_second_chi_cut_follows_the_sqrt_s_split(seed):
    """3-qubit pure state, chi = 1: the sweep's first cut (bond 1|2) splits psi[(s0 s1), s2] =
    U S V^H and hands U sqrt(S) (kept part) to the left; the second cut (bond 0|1) is the SVD of
    that tensor as [s0 x (s1 k)] (TNNOptimizer.py:111-134 with tn.split_node's U sqrt(S) /
    sqrt(S) V^H). Both cuts' discarded weights from dense linear algebra on psi."""
    n = 3
    ops = _random_ops(n, 3, seed)
    st = O.evolve([o for o in ops if o.name != 'TRUNCATE'], O.product_state(n), 'no', True)
    rho = O.density_matrix(st)
    w, V = np.linalg.eigh(rho)
    psi = V[:, -1] * math.sqrt(max(w[-1], 0))
    U, S, VH = np.linalg.svd(psi.reshape(4, 2), full_matrices=False)
    left = (U[:, :1] * np.sqrt(S[:1])).reshape(2, 2)          # [s0, (s1 k)]
    S2 = np.linalg.svd(left, compute_uv=False)
    st2 = st.copy()
    O.bond_truncate(st2, 1)
    chi = [d[2] for d in st2.discarded if d[0] == 'chi']
    assert np.isclose(chi[0], S[1] ** 2, atol=1e-12)
    assert np.isclose(chi[1], S2[1] ** 2, atol=1e-12)


def test_kappa_cut_of_a_single_site_keeps_its_top_eigenvectors():
    """One qubit with an inner index (no bonds): the kappa cut keeps the top-kappa eigenpairs of
    its density matrix rho = sum_a T_a T_a^H."""
    rng = np.random.default_rng(8)
    T = (rng.standard_normal((1, 2, 5, 1)) + 1j * rng.standard_normal((1, 2, 5, 1)))
    st = O.MPDO([T.copy()], [True], [])
    rho = O.density_matrix(st)
    w, V = np.linalg.eigh(rho)
    O.svd_kappa_left2right(st, 1)
    rho1 = O.density_matrix(st)
    assert np.allclose(rho1, w[-1] * np.outer(V[:, -1], V[:, -1].conj()))
    assert np.isclose(st.discarded[-1][2], w[0])


def test_golden_fixture_is_reproduced():
    """tests/golden/brickwork8_chi32_oracle.npz (written by tests/golden/make_golden.py) is what the oracle
    gives for the 8-qubit configuration now."""
    import bridge
    from mpdo_amd import workloads as W
    g = np.load(os.path.join(HERE, 'golden', 'brickwork8_chi32_oracle.npz'))
    st = bridge.run_oracle(W.brickwork8_chi32())
    assert np.allclose(O.density_matrix(st), g['rho'], atol=1e-12)


def test_bitstring_probs_are_the_dense_diagonal():
    """oracle.bitstring_probs (projector contraction of every qubit) against the diagonal of the
    dense density matrix of a noisy 4-qubit evolution, all 16 bitstrings in a shuffled order."""
    n = 4
    ops = _random_ops(n, 2, 41)
    ns = O.NoiseSpec(chip='best', dpc_error_rate=0.05, single='depolarizing', p1=0.03)
    st = O.evolve(ops, O.product_state(n), 'idealNoise', False, ns)
    rho = D.dense_evolve(ops, n, 'idealNoise', False, ns)
    idx = np.random.default_rng(5).permutation(2 ** n)
    bits = (idx[:, None] >> np.arange(n - 1, -1, -1)[None, :]) & 1      # qubit 0 = most significant
    got = O.bitstring_probs(st, bits)
    assert np.allclose(got, np.real(np.diag(rho))[idx], atol=1e-13)
    assert np.isclose(got.sum(), np.real(np.trace(rho)))


def test_expect_with_unsorted_qubits_matches_dense():
    """oracle.expect with oqs listed out of order (O's first factor acts on the first listed
    qubit): Tr(O rho) from the dense density matrix with O embedded on those qubits."""
    n = 4
    ops = _random_ops(n, 2, 43)
    ns = O.NoiseSpec(chip='best', dpc_error_rate=0.05, single='depolarizing', p1=0.03)
    st = O.evolve(ops, O.product_state(n), 'idealNoise', False, ns)
    rho = D.dense_evolve(ops, n, 'idealNoise', False, ns)
    rng = np.random.default_rng(6)
    A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    B = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    C = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    for oq in ([2, 0], [3, 1, 0], [1, 3]):
        mats = [A, B, C][:len(oq)]
        Ofull = mats[0]
        for m in mats[1:]:
            Ofull = np.kron(Ofull, m)
        # dense embedding: factor j acts on qubit oq[j]; identity elsewhere
        op = np.ones((1, 1), dtype=np.complex128)
        for q in range(n):
            op = np.kron(op, mats[oq.index(q)] if q in oq else np.eye(2))
        want = float(np.real(np.trace(op @ rho)))
        got = O.expect(st, [Ofull], [oq])[0]
        assert np.isclose(got, want, atol=1e-12), oq


def _mirror(st):
    """The same state with the chain read right to left (site i -> n-1-i, l <-> r)."""
    sites = [np.transpose(T, (3, 1, 2, 0)).copy() for T in reversed(st.sites)]
    return O.MPDO(sites, list(reversed(st.has_I)), list(reversed(st.bond)))


@pytest.mark.parametrize('seed', [51, 52])
def test_svd_left2right_pins(seed):
    """svd_left2right (TNNOptimizer.py:137-161): (1) on a pure state whose sites 1..n-1 are
    right-orthonormal, the first cut (bond 0|1) discards the Schmidt tail of psi[s0, rest];
    (2) the whole sweep equals the right-to-left sweep of the mirrored chain, mirrored back
    (same density matrix, same discarded weights in the mirrored order): a transposed split or
    a sweep in the wrong direction fails one of the two."""
    n = 4
    ops = _random_ops(n, 3, seed)
    st = O.evolve([o for o in ops if o.name != 'TRUNCATE'], O.product_state(n), 'no', True)
    # right-orthonormal gauge for sites 1..n-1: mirrored QR sweep
    m = _mirror(st)
    O.qr_left2right(m)
    st = _mirror(m)
    rho = O.density_matrix(st)
    w, V = np.linalg.eigh(rho)
    psi = V[:, -1] * math.sqrt(max(w[-1], 0))
    S = np.linalg.svd(psi.reshape(2, 2 ** (n - 1)), compute_uv=False)
    st1 = st.copy()
    O.svd_left2right(st1, 1)
    d = [x for x in st1.discarded if x[0] == 'chi_l2r']
    assert np.isclose(d[0][2], S[1] ** 2, atol=1e-12)
    for chi in (1, 2):
        a = st.copy()
        O.svd_left2right(a, chi)
        b = _mirror(st)
        O.svd_right2left(b, chi)
        db = [x[2] for x in b.discarded if x[0] == 'chi']
        b = _mirror(b)
        assert np.allclose(O.density_matrix(a), O.density_matrix(b), atol=1e-12)
        da = [x[2] for x in a.discarded if x[0] == 'chi_l2r']
        assert len(da) == n - 1
        assert np.allclose(da, db, atol=1e-12)


def test_compose_compressed_is_the_same_channel():
    """oracle.compose_compressed: the composed, Choi-rank-reduced Kraus set acts on random
    density matrices as the channels applied one after the other, with at most 4 operators;
    a reversed composition order fails (the channels below do not commute)."""
    rng = np.random.default_rng(9)
    ns = O.NoiseSpec(chip='best', single='depolarizing', p1=0.02)
    ud = O.NoiseSpec(chip='best', decay_param=0.03, dephasing_param=0.02)
    sets = [ns.single_kraus(O.gate_matrix('RY', [0.7])), ns.single_kraus(O.gate_matrix('RZ', [1.9])),
            ud.single_kraus(O.gate_matrix('RX', [0.4]))]
    E = O.compose_compressed(sets)
    assert E.shape[:2] == (2, 2) and E.shape[2] <= 4
    for _ in range(3):
        A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        rho = A @ A.conj().T
        want = rho
        for S in sets:
            want = sum(S[:, :, k] @ want @ S[:, :, k].conj().T for k in range(S.shape[2]))
        got = sum(E[:, :, k] @ rho @ E[:, :, k].conj().T for k in range(E.shape[2]))
        assert np.allclose(got, want, atol=1e-13)
        bad = O.compose_compressed(sets[::-1])
        wrong = sum(bad[:, :, k] @ rho @ bad[:, :, k].conj().T for k in range(bad.shape[2]))
        assert not np.allclose(wrong, want, atol=1e-6)


@pytest.mark.parametrize('noise', ['idealNoise', 'unified'])
def test_fused_channel_oracle_matches_sequential(noise):
    """bridge.run_oracle_fused (single-qubit channel runs composed and Choi-reduced) against
    oracle.evolve applying every Kraus set in sequence, with chi and kappa both truncating:
    the same density matrix, the same discarded weights at every cut."""
    import bridge
    from mpdo_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(61))
    n = 5
    ops = W.brickwork(n, 4, rng, two=('CNOT', 'CZ'), noisy_identity_after_2q=(noise == 'unified'))
    if noise == 'unified':
        w = W.Workload('t', n, ops, 'unified', False, 3, 2, chip=dict(decay_param=0.05, dephasing_param=0.03), depth=4)
    else:
        w = W.Workload('t', n, ops, 'idealNoise', False, 3, 2, chip=dict(dpc_error_rate=0.04),
                       single_qubit_noise='depolarizing', p1=0.02, depth=4)
    a = bridge.run_oracle(w)
    b = bridge.run_oracle_fused(w)
    assert np.allclose(O.density_matrix(a), O.density_matrix(b), atol=1e-12)
    da = [d for d in a.discarded if d[0] in ('chi', 'kappa')]
    db = [d for d in b.discarded if d[0] in ('chi', 'kappa')]
    assert [d[:2] for d in da] == [d[:2] for d in db]
    assert np.allclose([d[2] for d in da], [d[2] for d in db], atol=1e-12)
    assert any(d[2] > 1e-6 for d in da)
