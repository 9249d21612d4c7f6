"""Gradients of noisy expectations through the differentiable evolution (mpdo_amd.autodiff), on the
device (the fp64 Jacobi eigensolver behind GramEigh), against central finite differences of the
oracle (complex128):

* the eigensolver entry point mpdo_eigh against numpy;
* a 5-qubit, 2-layer ansatz with chi and kappa truncating: every one of its 48 parameters;
* the 16-qubit, 8-layer ansatz at chi = 64, kappa = 4 (the differentiation configuration): <H> and
  2 of its 632 gradient components against the committed fixture (tests/golden/make_golden.py
  hea16), all 632 from one reverse sweep with per-layer checkpointing.

Finite differences: h = 1e-4 for angles (at h = 1e-3 one component of the 5-qubit case carries a
truncation error of 4e-4, shown by h = 1e-4), 1e-5 for noise strengths (sqrt(p) curvature at
p ~ 1e-2); complex128, so the round-off is ~1e-11.
Bar: |g - g_fd| <= 1e-4 x max |g_fd| (small case), 1e-4 relative per component (fixture).
"""
import os

import numpy as np
import pytest
import torch

import bridge

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_eigh_entry_point(cuda):
    from mpdo_amd import _lib
    rng = np.random.default_rng(4)
    for n in (7, 64, 100, 300):
        X = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
        A = X @ np.conj(np.swapaxes(X, 1, 2))
        w, V = _lib.eigh(torch.from_numpy(A).to(cuda))
        w, V = w.cpu().numpy(), V.cpu().numpy()
        ref = np.linalg.eigvalsh(A)[:, ::-1]
        assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max()
        R = A @ V - V * w[:, None, :]
        assert np.abs(R).max() <= 1e-11 * np.abs(ref).max(), n
        assert w.shape == (2, n) and np.all(np.diff(w, axis=1) <= 1e-9 * np.abs(ref).max())


def _fd(theta, phi, p1, p2, chi, kappa, name, idx):
    h = 1e-4 if name in ('theta', 'phi') else 1e-5
    vals = []
    for sgn in (1, -1):
        args = dict(theta=theta.copy(), phi=phi.copy(), p1=p1.copy(), p2=p2.copy())
        args[name][idx] += sgn * h
        vals.append(bridge.run_hea_oracle(args['theta'], args['phi'], args['p1'], args['p2'], chi, kappa)[0])
    return (vals[0] - vals[1]) / (2 * h)


def test_hea_small_all_parameters(cuda):
    """5 qubits, 2 layers, chi=4, kappa=2 (both truncate): <H> equals the oracle's, and the
    gradient with respect to every angle and noise strength matches central differences."""
    from mpdo_amd import autodiff as AD
    n, depth, chi, kappa = 5, 2, 4, 2
    rng = np.random.default_rng(12)
    theta = rng.uniform(-np.pi, np.pi, (depth, n))
    phi = rng.uniform(-np.pi, np.pi, (depth, n))
    p1 = rng.uniform(0.005, 0.02, (depth, n, 2))
    p2 = rng.uniform(0.005, 0.02, (depth, n - 1))
    m = AD.HEA(n, depth, chi, kappa, theta, phi, p1, p2, device=str(cuda))
    E = AD.tfim_energy(m())
    E.backward()
    e_ref, _ = bridge.run_hea_oracle(theta, phi, p1, p2, chi, kappa)
    assert abs(E.item() - e_ref) <= 1e-10 * max(abs(e_ref), 1.0)
    assert any(d[0] == 'chi' and d[2] > 1e-8 for d in m.discarded)
    got, want = [], []
    for name, par, arr in [('theta', m.theta, theta), ('phi', m.phi, phi), ('p1', m.p1, p1), ('p2', m.p2, p2)]:
        g = par.grad.detach().cpu().numpy()
        for idx in np.ndindex(g.shape):
            got.append(g[idx])
            want.append(_fd(theta, phi, p1, p2, chi, kappa, name, idx))
    got, want = np.array(got), np.array(want)
    assert got.shape == (48,)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f'max |g - g_fd| / max |g_fd| = {err:.2e}')
    assert err <= 1e-4


@pytest.mark.timeout(900)
def test_hea16_chi64_gradient_fixture(cuda):
    """The 16-qubit differentiation configuration: one forward and one reverse sweep (checkpointed
    per layer) give <H> and its 632 partial derivatives; <H> and 2 of them against the oracle's
    central differences."""
    import importlib.util
    from mpdo_amd import autodiff as AD
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(HERE, 'golden', 'make_golden.py'))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = np.load(os.path.join(HERE, 'golden', 'hea16_chi64_fd_oracle.npz'))
    theta, phi, p1, p2 = mg.hea16_params()
    m = AD.HEA(16, 8, 64, 4, theta, phi, p1, p2, device=str(cuda))
    E = AD.tfim_energy(m())
    assert abs(E.item() - float(g['energy'])) <= 1e-8 * max(abs(float(g['energy'])), 1.0)
    E.backward()
    grads = {'theta': m.theta.grad, 'phi': m.phi.grad, 'p1': m.p1.grad, 'p2': m.p2.grad}
    assert sum(v.numel() for v in grads.values()) == 632
    for (name, idx), fd in zip(mg.HEA16_FD, g['fd']):
        ours = float(grads[name][idx])
        print(name, idx, ours, float(fd))
        assert abs(ours - fd) <= 1e-4 * abs(fd) + 1e-9, (name, idx, ours, float(fd))
