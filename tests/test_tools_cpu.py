"""Tools.cal_fidelity and Tools.sqrt_matrix (Tools.py:512-551) on closed forms (host utilities,
no GPU): pure states give |<psi|phi>|^2, commuting mixed states (sum_i sqrt(p_i q_i))^2, a state
with itself 1, and sqrt_matrix squares back to its argument."""
import numpy as np
import pytest
import torch


def _tools():
    from mpdo_amd import Tools
    return Tools


def _rand_pure(rng, d):
    v = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    return v / np.linalg.norm(v)


def test_fidelity_of_pure_states():
    T = _tools()
    rng = np.random.default_rng(1)
    for d in (2, 4, 16):
        a, b = _rand_pure(rng, d), _rand_pure(rng, d)
        ra, rb = np.outer(a, a.conj()), np.outer(b, b.conj())
        f = float(T.cal_fidelity(torch.from_numpy(ra), torch.from_numpy(rb)))
        assert abs(f - abs(np.vdot(a, b)) ** 2) <= 1e-9


def test_fidelity_of_commuting_mixed_states():
    T = _tools()
    rng = np.random.default_rng(2)
    d = 8
    U, _ = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    p, q = rng.random(d), rng.random(d)
    p, q = p / p.sum(), q / q.sum()
    ra, rb = (U * p) @ U.conj().T, (U * q) @ U.conj().T
    f = float(T.cal_fidelity(torch.from_numpy(ra), torch.from_numpy(rb)))
    assert abs(f - np.sum(np.sqrt(p * q)) ** 2) <= 1e-9
    assert abs(float(T.cal_fidelity(torch.from_numpy(ra), torch.from_numpy(ra))) - 1.0) <= 1e-9


def test_sqrt_matrix_and_shape_checks():
    T = _tools()
    rng = np.random.default_rng(3)
    X = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    A = X @ X.conj().T
    S = T.sqrt_matrix(torch.from_numpy(A)).numpy()
    assert np.abs(S @ S - A).max() <= 1e-10 * np.abs(A).max()
    with pytest.raises(ValueError):
        T.cal_fidelity(torch.eye(2, dtype=torch.complex128), torch.eye(4, dtype=torch.complex128))
