"""Pins of the oracle's shot sampler (Circuit.py:297-467) on the CPU: conditional probabilities
against marginals of the dense density matrix, draws against the exact distribution, and the
reference's group and edge rules."""
import itertools

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import mpdo_oracle as O


def _state(n, seed, depth=2):
    rng = np.random.default_rng(seed)
    b = O.CircuitBuilder(n, 'idealNoise')
    for d in range(depth):
        for q in range(n):
            b.add('RY', [q], [float(rng.uniform(0, 6.28))])
            b.add('RZ', [q], [float(rng.uniform(0, 6.28))], ideal=True)
        for a in range(d % 2, n - 1, 2):
            b.add('CNOT', [a, a + 1])
        b.truncate()
    ns = O.NoiseSpec(chip='best', dpc_error_rate=0.05, single='depolarizing', p1=0.03)
    st = O.evolve(b.ops, O.product_state(n), 'idealNoise', False, ns)
    rho = D.dense_evolve(b.ops, n, 'idealNoise', False, ns)
    return st, rho


def _rotated_dense(rho, orientation, n):
    for pos, o in enumerate(orientation):
        if o in (0, 1):
            rho = D.apply_channel(rho, [O.MEASURE_ROTATIONS[o]], [pos], n)
    return rho


def _joint(rho, order, n):
    """Diagonal of the reduced density matrix of ``order`` as an array indexed by bits."""
    red = D.partial_trace(rho, list(order), n)
    return np.real(np.diag(red)).reshape([2] * len(order))


@pytest.mark.parametrize('reduced,orientation', [([], [2, 2, 2, 2]), ([], [0, 1, 2, 0]), ([1], [1, 0, 2]),
                                                 ([0, 2], [0, 1])])
def test_conditional_probs_are_dense_marginals(reduced, orientation):
    n = 4
    st, rho = _state(n, 5)
    order = [q for q in range(n) if q not in reduced]
    rst = O.sampling_state(st, orientation)
    joint = _joint(_rotated_dense(rho, orientation, n), order, n)
    for k in range(len(order)):
        for hist in itertools.product([0, 1], repeat=k):
            P = O.conditional_probs(rst, order, list(hist))
            want = joint[hist].reshape(2, -1).sum(axis=1)
            assert np.allclose(P, want, rtol=1e-10, atol=1e-14), (hist, P, want)


@pytest.mark.parametrize('sequential,shots', [(False, 20000), (True, 2000)])
def test_draws_follow_the_exact_distribution(sequential, shots):
    n = 3
    st, rho = _state(n, 9)
    p = _joint(rho, range(n), n).reshape(-1)
    p = p / p.sum()
    torch.manual_seed(1234)
    bits, counts = O.sample(st, shots=shots, sequential=sequential)
    assert len(bits) == shots and all(len(b) == n for b in bits)
    obs = np.array([counts.get(format(i, f'0{n}b'), 0) for i in range(2 ** n)], dtype=float)
    exp = shots * p
    chi2 = float(np.sum((obs - exp) ** 2 / exp))
    assert chi2 < 35.0, (chi2, obs, exp)         # 7 degrees of freedom: p-value ~ 1e-5
    assert list(counts) == sorted(counts)


def test_certain_outcomes_draw_nothing():
    """p1 within GLOBAL_MINIMUM of 0 or 1: the batch sampler assigns the whole group without a
    Bernoulli draw (Circuit.py:370-373), so the generator state is untouched."""
    for vec, want in [((1.0, 0.0), '0000'), ((0.0, 1.0), '1111')]:
        st = O.product_state(4, vec)
        torch.manual_seed(7)
        before = torch.get_rng_state().clone()
        bits, counts = O.sample(st, shots=64)
        assert torch.equal(before, torch.get_rng_state())
        assert bits == [want] * 64 and counts == {want: 64}


def test_groups_place_ones_first_and_split_by_history():
    """An X-basis sample of |0>: every qubit is a fair coin; per group the ones come first, so the
    first column is sorted descending and each later column is sorted within each history."""
    st = O.product_state(3)
    torch.manual_seed(3)
    trace = []
    bits = O.conditional_batch_sample(O.sampling_state(st, [0, 0, 0]), [0, 1, 2], 256, trace=trace)
    col0 = [b[0] for b in bits]
    assert col0 == sorted(col0, reverse=True)
    assert 80 < sum(col0) < 176
    assert [h for h, _ in trace[:3]] == [(), (1,), (0,)]
    assert all(abs(p - 0.5) < 1e-6 for _, p in trace)
    for h in [(1,), (0,)]:
        rows = [b for b in bits if b[0] == h[0]]
        assert [r[1] for r in rows] == sorted((r[1] for r in rows), reverse=True)


def test_sample_result_forms_and_errors():
    st = O.product_state(3, (0.6, 0.8))
    torch.manual_seed(0)
    bits = O.sample(st, shots=16, bool_result=True, counts=False)
    assert all(isinstance(v, bool) for row in bits for v in row)
    torch.manual_seed(0)
    ints = O.sample(st, shots=16, sample_string=False, counts=False)
    assert [[int(v) for v in r] for r in bits] == ints
    with pytest.raises(ValueError):
        O.sample(st, shots=4, orientation=[2, 2])
    # reduced qubits are traced; the remaining ones are sampled in increasing index order
    torch.manual_seed(0)
    s2 = O.sample(st, shots=16, reduced=[1], counts=False)
    assert all(len(s) == 2 for s in s2)


def test_conditional_prb_normalises_in_float32_with_the_floor():
    """Circuit.py:328-332: (P + GLOBAL_MINIMUM) normalised in float32, negative entries rejected."""
    assert O.normalise_prb([0.0, 0.0]) == pytest.approx(0.5)
    p = O.normalise_prb([1e-9, 3e-9])
    assert p == pytest.approx((3e-9 + O.GLOBAL_MINIMUM) / (4e-9 + 2 * O.GLOBAL_MINIMUM), rel=1e-6)
    assert O.normalise_prb([0.25, 0.75]) == np.float32(0.75)
    with pytest.raises(RuntimeError):
        O.normalise_prb([0.5, -1e-3])


@pytest.mark.parametrize('reduced', [[], [1], [0, 3], [1, 2, 3]])
def test_trace_rho_rho_with_reduced_qubits_is_the_dense_trace(reduced):
    """dmOperations.py:51-66 on reduced states: Tr(rho0_k rho1_k) of the dense partial traces;
    trace_composited_rho2 as Tr(((rho0 + rho1) / 2)^2)."""
    n = 4
    st0, r0 = _state(n, 21)
    st1, r1 = _state(n, 22)
    keep = [q for q in range(n) if q not in reduced]
    d0, d1 = D.partial_trace(r0, keep, n), D.partial_trace(r1, keep, n)
    assert np.isclose(O.trace_rho_rho(st0, st1, reduced), np.real(np.trace(d0 @ d1)), rtol=1e-12)
    assert np.isclose(O.trace_rho_rho(st1, st1, reduced), np.real(np.trace(d1 @ d1)), rtol=1e-12)
    m = 0.5 * (d0 + d1)
    assert np.isclose(O.trace_composited_rho2([st0, st1], reduced), np.real(np.trace(m @ m)), rtol=1e-12)
