"""CPU ORACLE — exact complex128 dense density-matrix simulator (test infrastructure only).

rho -> sum_k K_k rho K_k^dagger on the listed qubits, qubit 0 the most significant bit. It shares
no code with ``mpdo_oracle``'s tensor-network engine: it is the independent check that pins gate
application, Kraus absorption and the observables whenever no truncation is active.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np


def zero_state(n: int) -> np.ndarray:
    rho = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    rho[0, 0] = 1.0
    return rho


def _apply_op_left(rho_t: np.ndarray, K: np.ndarray, qubits: Sequence[int], n: int) -> np.ndarray:
    """K acting on the ket (row) indices of rho viewed as a 2n-index tensor."""
    k = len(qubits)
    Kt = K.reshape([2] * (2 * k))
    out = np.tensordot(Kt, rho_t, axes=(list(range(k, 2 * k)), list(qubits)))
    # tensordot puts K's output axes first; move them back to the qubit positions
    rest = [i for i in range(2 * n) if i not in qubits]
    order = list(qubits) + rest
    inv = np.argsort(order)
    return np.transpose(out, inv)


def apply_channel(rho: np.ndarray, kraus: List[np.ndarray], qubits: Sequence[int], n: int) -> np.ndarray:
    """rho <- sum_k K rho K^dagger, each K a (2^k x 2^k) matrix on ``qubits`` (first = MSB)."""
    rt = rho.reshape([2] * (2 * n))
    bra = [n + q for q in qubits]
    acc = np.zeros_like(rt)
    for K in kraus:
        t = _apply_op_left(rt, K, qubits, n)
        t = _apply_op_left(t, K.conj(), bra, n)
        acc = acc + t
    return acc.reshape(2 ** n, 2 ** n)


def kraus_list_1q(E: np.ndarray) -> List[np.ndarray]:
    """(2,2) gate or (2,2,K) Kraus tensor -> list of 2x2 matrices."""
    if E.ndim == 2:
        return [E]
    return [E[:, :, k] for k in range(E.shape[2])]


def kraus_list_2q(G: np.ndarray) -> List[np.ndarray]:
    """(2,2,2,2) gate or (2,2,2,2,K) Kraus tensor -> list of 4x4 matrices."""
    if G.ndim == 4:
        return [G.reshape(4, 4)]
    return [G[..., k].reshape(4, 4) for k in range(G.shape[4])]


def partial_trace(rho: np.ndarray, keep: Sequence[int], n: int) -> np.ndarray:
    rt = rho.reshape([2] * (2 * n))
    drop = [q for q in range(n) if q not in keep]
    letters = 'abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ'
    ket = [letters[i] for i in range(n)]
    bra = [letters[n + i] for i in range(n)]
    for q in drop:
        bra[q] = ket[q]
    out = ''.join(ket[q] for q in keep) + ''.join(bra[q] for q in keep)
    red = np.einsum(''.join(ket) + ''.join(bra) + '->' + out, rt)
    d = 2 ** len(keep)
    return red.reshape(d, d)


def pauli_string_expect(rho: np.ndarray, paulis: Sequence[int], qubits: Sequence[int], n: int) -> float:
    P = [np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]])]
    op = np.array([[1.0 + 0j]])
    for q in range(n):
        m = np.eye(2)
        if q in qubits:
            m = P[paulis[list(qubits).index(q)]]
        op = np.kron(op, m)
    return float(np.real(np.trace(op @ rho)))


def dense_evolve(ops, n: int, noise_type: str = 'no', ideal: bool = True, noise=None,
                 rho: np.ndarray = None) -> np.ndarray:
    """Exact evolution of the same layer list (no truncation). Which gates are noisy and which
    Kraus set they carry follows the reference rules (Circuit.py:84,141); the Kraus builders are
    the oracle's (pinned by the closed-form tests)."""
    from .mpdo_oracle import NoiseSpec, TWO_QUBIT, gate_matrix
    nt = noise_type.lower()
    unified, real, idealn = (not ideal and nt == 'unified'), (not ideal and nt == 'realnoise'), \
        (not ideal and nt == 'idealnoise')
    noise = noise or NoiseSpec()
    rho = zero_state(n) if rho is None else rho
    cache = {}
    for op in ops:
        if op.name in ('TRUNCATE', 'BARRIER', 'MEASUREZ'):
            continue
        G = np.asarray(op.matrix, dtype=np.complex128) if op.matrix is not None else gate_matrix(op.name, op.params)
        if op.name not in TWO_QUBIT:
            if (idealn or unified) and not op.ideal:
                E = noise.single_kraus(G) if op.variational else cache.setdefault(op.name, noise.single_kraus(G))
            else:
                E = G
            for q in op.qubits:
                rho = apply_channel(rho, kraus_list_1q(E), [q], n)
            continue
        if ((idealn and not op.ideal) or real) and not real:
            G = noise.double_kraus(G) if op.variational else cache.setdefault(op.name, noise.double_kraus(G))
        rho = apply_channel(rho, kraus_list_2q(G), list(op.qubits), n)
    return rho
