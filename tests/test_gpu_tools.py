"""The product-side observables around the kernels, on states evolved on the GPU, against the dense
density-matrix simulator (oracle/dense.py) with operators built by numpy in the listed qubit order:

* dmOperations.expect (dmOperations.py:135-171): arbitrary Hermitian observables on unsorted qubit
  lists (axis t of the observable acts on qubit oqs[t]), Re Tr(O rho~) within 1e-5 of max(1, |ref|);
* TensorCircuit.cal_vector (Circuit.py:264-272) of a noise-free circuit: |psi><psi| equal to the
  dense density matrix within 1e-5 elementwise, |psi| = 1 within 1e-5;
* Tools.cal_fidelity (Tools.py:528-551) between the GPU density matrix and the dense one (1 within
  1e-5) and between two different circuits' states against numpy's fidelity of the dense pair."""
import numpy as np
import pytest
import torch

from oracle import dense as D

import bridge
from test_gpu_circuits import _random_workload, _run_product

pytestmark = pytest.mark.gpu


def _dense(w):
    return D.dense_evolve(bridge.oracle_ops(w), w.n, w.noise_type, w.ideal, bridge.oracle_noise(w))


def _embed(O, qubits, n):
    """O (2^k x 2^k, axis t on qubit qubits[t]) as a 2^n x 2^n operator in the natural qubit order."""
    k = len(qubits)
    rest = [q for q in range(n) if q not in qubits]
    full = np.kron(O, np.eye(2 ** (n - k))).reshape([2] * (2 * n))
    order = list(qubits) + rest
    perm = [order.index(q) for q in range(n)]
    return full.transpose(perm + [p + n for p in perm]).reshape(2 ** n, 2 ** n)


def _hermitian(rng, d):
    X = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return (X + X.conj().T) / 2


def test_expect_unsorted_qubits_vs_dense(cuda):
    from mpdo_amd import dmOperations
    w = _random_workload(7, 5, 3, 'idealNoise', None, None, two=('CNOT', 'CZ'))
    c, _ = _run_product(w, cuda)
    dm = c.cal_dmNodes()
    rho = _dense(w)
    rng = np.random.default_rng(3)
    oqs = [[3, 1], [4, 0, 2], [2], [0, 4]]
    obs = [_hermitian(rng, 2 ** len(q)) for q in oqs]
    vals = dmOperations.expect(dm, [torch.from_numpy(o) for o in obs], oqs)
    for v, o, q in zip(vals, obs, oqs):
        ref = float(np.trace(_embed(o, q, w.n) @ rho).real)
        assert abs(float(v) - ref) <= 1e-5 * max(1.0, abs(ref)), (q, float(v), ref)
    # a repeated qubit is refused as in the reference's node bookkeeping
    with pytest.raises(ValueError):
        dmOperations.expect(dm, torch.from_numpy(obs[0]), [[1, 1]])


def test_cal_vector_of_a_noise_free_circuit(cuda):
    w = _random_workload(11, 6, 3, 'ideal', None, None, two=('CNOT', 'CZ'))
    c, _ = _run_product(w, cuda)
    psi = c.cal_vector().cpu().numpy().astype(np.complex128).reshape(-1)
    rho = _dense(w)
    assert abs(np.linalg.norm(psi) - 1.0) <= 1e-5
    assert np.abs(np.outer(psi, psi.conj()) - rho).max() <= 1e-5


def test_cal_fidelity_of_gpu_states(cuda):
    from mpdo_amd import Tools
    w1 = _random_workload(13, 4, 2, 'idealNoise', None, None)
    w2 = _random_workload(14, 4, 2, 'idealNoise-decay', None, None)
    c1, _ = _run_product(w1, cuda)
    c2, _ = _run_product(w2, cuda)
    r1 = c1.cal_dm().cpu().to(torch.complex128)
    r2 = c2.cal_dm().cpu().to(torch.complex128)
    d1, d2 = _dense(w1), _dense(w2)
    f_same = float(Tools.cal_fidelity(r1, torch.from_numpy(d1)))
    assert abs(f_same - 1.0) <= 1e-5, f_same
    # numpy's fidelity of the dense pair: (tr sqrt(sqrt(a) b sqrt(a)))^2
    ev, V = np.linalg.eigh(d1)
    sa = (V * np.sqrt(np.clip(ev, 0, None))) @ V.conj().T
    ref = np.sum(np.sqrt(np.clip(np.linalg.eigvalsh(sa @ d2 @ sa), 0, None))) ** 2
    f = float(Tools.cal_fidelity(r1, r2))
    assert abs(f - ref) <= 1e-5, (f, ref)
