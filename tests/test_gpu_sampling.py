"""Shot sampling on the device (mpdo_env_right + mpdo_sample_step through TensorCircuit.sample)
against the oracle's restatement of Circuit.py:297-467 on the same evolved state: the
conditional probabilities within 1e-9 of the history's total (1e-6 when the oracle also applies
the basis rotations in complex128), and, with the same torch seed, the same bitstrings."""
import numpy as np
import pytest
import torch

from oracle import mpdo_oracle as O

pytestmark = pytest.mark.gpu


def _evolved(cuda, n, depth, chi, kappa, seed=5):
    import mpdo_amd as M
    from mpdo_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = W.brickwork(n, depth, rng)
    w = W.Workload('s', n, ops, 'idealNoise', False, chi, kappa, chip=dict(dpc_error_rate=0.03),
                   single_qubit_noise='depolarizing', p1=0.02, depth=depth)
    c = M.TensorCircuit(device=str(cuda), **w.circuit_kwargs())
    W.apply_ops(c, ops)
    c.evolve(M.Tools.create_ket0Series(n, device=str(cuda)))
    return c


def _oracle_state(c):
    st = c.stateNodes
    return O.MPDO([s.tensor.detach().cpu().numpy().astype(np.complex128) for s in st],
                  [s.has_I for s in st], [s.bond_right for s in st[:-1]])


@pytest.mark.parametrize('reduced,orientation', [([], None), ([], 'mixed'), ([2, 5], 'mixed')])
def test_conditional_probabilities_match_oracle(cuda, reduced, orientation):
    from mpdo_amd import sampling
    c = _evolved(cuda, 8, 6, 8, 2)
    n = 8
    order = [q for q in range(n) if q not in reduced]
    ori = [2] * len(order) if orientation is None else [(3 * i + 1) % 3 for i in range(len(order))]
    c.sample(shots=8, orientation=ori, reduced=reduced, _tqdm_disable=True)
    sweep = c._nodes4samples
    # the basis rotations are complex64 contractions on the device, as in the reference: the
    # sweep is checked on the rotated sites it was given, the rotation against the oracle's own
    rotated = O.MPDO([t.detach().cpu().numpy().astype(np.complex128) for t in sweep.sites],
                     [True] * n, [True] * (n - 1))
    ost = O.sampling_state(_oracle_state(c), ori)
    rng = np.random.default_rng(0)
    for k in range(len(order)):
        for _ in range(3):
            hist = [int(b) for b in rng.integers(0, 2, k)]
            got = np.array(sweep.probs_of(hist))
            want = O.conditional_probs(rotated, order, hist)
            assert np.abs(got - want).max() <= 1e-9 * max(want.sum(), 1e-300) + 1e-15, (hist, got, want)
            ref = O.conditional_probs(ost, order, hist)
            assert np.abs(got - ref).max() <= 1e-6 * max(ref.sum(), 1e-300) + 1e-15, (hist, got, ref)
            assert sampling.normalise_prb(*got) == pytest.approx(O.normalise_prb(want), abs=1e-6)


@pytest.mark.parametrize('kw', [dict(), dict(orientation=[0, 1, 2, 0, 1, 2, 0, 1]), dict(reduced=[0, 3]),
                                dict(_require_bool_result=True, _require_counts=False)])
def test_batch_sample_reproduces_oracle_bitstrings(cuda, kw):
    c = _evolved(cuda, 8, 6, 8, 2)
    ost = _oracle_state(c)
    torch.manual_seed(2024)
    got = c.sample(shots=512, _tqdm_disable=True, **kw)
    torch.manual_seed(2024)
    want = O.sample(ost, shots=512, orientation=kw.get('orientation'), reduced=kw.get('reduced'),
                    bool_result=kw.get('_require_bool_result', False), counts=kw.get('_require_counts', True))
    assert got == want


def test_sequential_sample_reproduces_oracle_bitstrings(cuda):
    c = _evolved(cuda, 6, 5, 8, 2)
    ost = _oracle_state(c)
    torch.manual_seed(77)
    got = c.sample(shots=40, orientation=[1, 0, 2, 2, 0, 1], _tqdm_disable=True, _require_sequential_sample=True)
    torch.manual_seed(77)
    want = O.sample(ost, shots=40, orientation=[1, 0, 2, 2, 0, 1], sequential=True)
    assert got == want


def test_random_sample_schemes(cuda):
    import mpdo_amd as M
    c = _evolved(cuda, 6, 4, 8, 2)
    schemes = M.Tools.random_measurementScheme(6, 3)
    out = c.randomSample(schemes, shots_per_scheme=64)
    assert len(out) == 3 and all(len(r) == 64 and all(len(b) == 6 for b in r) for r in out)
    assert all(v in (0, 1) for r in out for b in r for v in b)


def test_sampled_z_means_match_expectations_at_chi_32(cuda):
    """16 qubits at chi 32: the sample means of Z_i agree with <Z_i>/Tr(rho) within 5 sigma."""
    from mpdo_amd import dmOperations as dm
    n, shots = 16, 4096
    c = _evolved(cuda, n, 10, 32, 4, seed=11)
    torch.manual_seed(5)
    bits = c.sample(shots=shots, sample_string=False, _tqdm_disable=True, _require_counts=False)
    b = np.array(bits, dtype=float)
    dmn = c.cal_dmNodes()
    tr = float(dm.trace_rho(dmn))
    for q in range(n):
        z = float(dm.pauli_expect(dmn, [2], [q])) / tr
        mean = float(np.mean(1 - 2 * b[:, q]))
        sigma = np.sqrt(max(1 - z * z, 1e-6) / shots)
        assert abs(mean - z) < 5 * sigma + 1e-6, (q, mean, z)
