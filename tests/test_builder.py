"""Builder API under realNoise and the layer markers (AbstractCircuit.py:156-532): the gate lists
that cx/cnot/rxx/ryy/rzz expand into are pinned by composing them into 4x4 unitaries with the
ideal CZ channel (closed forms exp(-i theta/2 P (x) P)); the product builder's layer list must
equal the oracle builder's; measure/reset/barrier/truncate append what the reference appends.
Host only (no GPU)."""
import math
import os

import numpy as np
import pytest
import torch
from scipy.linalg import expm

from oracle import mpdo_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
IDEAL_CZ = os.path.join(HERE, 'golden', 'ideal_cz.mat')
PX = np.array([[0, 1], [1, 0]], complex)
PY = np.array([[0, -1j], [1j, 0]], complex)
PZ = np.array([[1, 0], [0, -1]], complex)


def _embed(m, q):
    """One-qubit matrix on qubit q of a pair (q = 0 the most significant bit)."""
    return np.kron(m, np.eye(2)) if q == 0 else np.kron(np.eye(2), m)


def _compose(ops, q0=0):
    """Unitary of a layer list on qubits (q0, q0 + 1), applied in list order."""
    U = np.eye(4, dtype=complex)
    for op in ops:
        if op.name in ('TRUNCATE', 'BARRIER'):
            continue
        if op.name not in ('CZEXP', 'CPEXP') and O.gate_matrix(op.name, op.params).shape == (2, 2):
            G = np.eye(4, dtype=complex)   # a one-qubit gate acts on every listed qubit
            for q in op.qubits:
                G = _embed(O.gate_matrix(op.name, op.params), q - q0) @ G
        elif op.name == 'CZEXP':
            assert op.matrix.shape[-1] == 1
            G = op.matrix[..., 0].reshape(4, 4)
            if op.qubits[0] > op.qubits[1]:
                G = G.reshape(2, 2, 2, 2).transpose(1, 0, 3, 2).reshape(4, 4)
        else:
            G = O.gate_matrix(op.name, op.params).reshape(4, 4)
        U = G @ U
    return U


def _equal_up_to_phase(A, B, tol=1e-12):
    k = np.unravel_index(np.argmax(np.abs(B)), B.shape)
    ph = A[k] / B[k]
    return abs(abs(ph) - 1) < tol and np.abs(A - ph * B).max() < tol


def _exp_tensors():
    return {'01': O.chi_to_kraus(O.load_chi(IDEAL_CZ)), '10': O.chi_to_kraus(O.load_chi(IDEAL_CZ))}


@pytest.mark.parametrize('theta', [0.3, -1.1, 2.5])
@pytest.mark.parametrize('name,pauli', [('RXX', PX), ('RYY', PY), ('RZZ', PZ)])
def test_realnoise_rotation_decompositions(name, pauli, theta):
    """AbstractCircuit.py:231-275: under realNoise rxx/ryy/rzz become CX-conjugated RZ (with
    H / RX(pi/2) basis changes); with the ideal CZ channel the list is exp(-i theta/2 P (x) P).
    (The non-realNoise RYY gate is the reference's X (x) Y rotation, YGates.py:138-143; the
    realNoise decomposition is the true Y (x) Y rotation.)"""
    b = O.CircuitBuilder(2, 'realNoise', _exp_tensors())
    b.add(name, [0, 1], [theta])
    assert [op.name for op in b.ops].count('CZEXP') == 2
    assert _equal_up_to_phase(_compose(b.ops), expm(-0.5j * theta * np.kron(pauli, pauli)))


@pytest.mark.parametrize('name', ['CX', 'CNOT'])
@pytest.mark.parametrize('q', [(0, 1), (1, 0)])
def test_realnoise_cnot_decomposition(name, q):
    """cx/cnot -> ry(-pi/2, target) CZEXP ry(pi/2, target) (AbstractCircuit.py:291-301, :332-342)."""
    b = O.CircuitBuilder(2, 'realNoise', _exp_tensors())
    b.add(name, list(q))
    cnot = np.eye(4, dtype=complex)[[0, 1, 3, 2]]
    if q == (1, 0):
        cnot = np.eye(4, dtype=complex)[[0, 3, 2, 1]]
    assert [op.name for op in b.ops] == ['RY', 'CZEXP', 'RY']
    assert [op.ideal for op in b.ops] == [True, False, True]
    assert _equal_up_to_phase(_compose(b.ops), cnot)


def _product_ops(c):
    """(name, qubits, params, ideal) of every layer of a product-path circuit."""
    out = []
    for gate, oqs in zip(c.layers, c._oqs_list):
        if gate.name in ('truncate', 'barrier'):
            out.append((gate.name.upper(), [], [], None))
            continue
        para = gate.para
        params = [] if para is None else [float(p) for p in (para if isinstance(para, list) else [para])]
        out.append((gate.name.upper(), list(oqs), params, gate.ideal))
    return out


def _oracle_ops(ops):
    return [(op.name, list(op.qubits), [float(p) for p in op.params], op.ideal) for op in ops]


def test_product_builder_matches_oracle_builder():
    """The product path's TensorCircuit builds the same realNoise layer list as the oracle's
    builder (names, qubits, parameters, ideal flags), including the tomographed CZ tensors."""
    import mpdo_amd as M
    n = 4
    files = {'CZ': {f'{i}{i + 1}': IDEAL_CZ for i in range(n - 1)}, 'CP': {}}
    c = M.TensorCircuit(n, ideal=False, noiseType='realNoise', chiFileDict=files, chi=8, kappa=4, chip='best',
                        device='cpu')
    b = O.CircuitBuilder(n, 'realNoise', {f'{i}{i + 1}': O.chi_to_kraus(O.load_chi(IDEAL_CZ)) for i in range(n - 1)})
    c.h(0); b.add('H', [0])
    c.rxx(torch.tensor(0.4), 0, 1); b.add('RXX', [0, 1], [0.4])
    c.ryy(torch.tensor(-0.7), 1, 2); b.add('RYY', [1, 2], [-0.7])
    c.rzz(torch.tensor(1.3), 2, 3); b.add('RZZ', [2, 3], [1.3])
    c.cnot(3, 2); b.add('CNOT', [3, 2])
    c.cz(1, 2); b.add('CZ', [1, 2])
    c.barrier(); b.barrier()
    c.truncate(); b.truncate()
    got, want = _product_ops(c), _oracle_ops(b.ops)
    assert [g[:2] for g in got] == [w[:2] for w in want]
    for g, w in zip(got, want):
        assert np.allclose(g[2], w[2], atol=1e-6), (g, w)
        assert bool(g[3]) == bool(w[3]), (g, w)
    # the tomographed tensors carried by the CZEXP layers equal the oracle's conversion
    exp = O.chi_to_kraus(O.load_chi(IDEAL_CZ))
    for gate in c.layers:
        if gate.name == 'CZEXP':
            t = gate.tensor.detach().cpu().numpy()
            assert np.abs(t - exp.astype(t.dtype)).max() < 1e-6


def test_measure_reset_and_markers():
    """measure(oqs, orientations) appends one MeasureX/Y/Z layer per qubit (default Z) with the
    reference's headline, rejects an orientation outside 0..2; reset0/reset1 append Reset0/1;
    barrier/truncate append markers on [None] (AbstractCircuit.py:483-532)."""
    import mpdo_amd as M
    c = M.TensorCircuit(3, device='cpu')
    c.measure([0, 2], [0, 1])
    c.measure([1])
    c.reset0(2)
    c.reset1([0, 1])
    c.barrier()
    c.truncate()
    names = [g.name for g in c.layers]
    assert names == ['MeasureX', 'MeasureY', 'MeasureZ', 'Reset0', 'Reset1', 'barrier', 'truncate']
    assert c._oqs_list == [[0], [2], [1], [2], [0, 1], [None], [None]]
    T = [g.tensor.detach().cpu().numpy().astype(complex) for g in c.layers[:5]]
    # MeasureX/MeasureY rotate the +1 eigenstate of X/Y onto |0> and the -1 one onto |1>
    # (SingleGates.py:240-305); MeasureZ is the identity; Reset0/Reset1 are the projectors
    # |0><0| and |1><1| (SingleGates.py:342-407: a projection, not a reset channel)
    s2 = 1 / math.sqrt(2)
    for M_, plus, minus in ((T[0], [s2, s2], [s2, -s2]), (T[1], [s2, 1j * s2], [s2, -1j * s2])):
        assert np.allclose(M_ @ np.array(plus), [1, 0], atol=1e-6)
        assert np.allclose(np.abs(M_ @ np.array(minus)), [0, 1], atol=1e-6)
        assert np.allclose(M_.conj().T @ M_, np.eye(2), atol=1e-6)
    assert np.allclose(T[2], np.eye(2))
    assert np.allclose(T[3], np.diag([1, 0])) and np.allclose(T[4], np.diag([0, 1]))
    with pytest.raises(ValueError):
        c.measure([0], [3])
