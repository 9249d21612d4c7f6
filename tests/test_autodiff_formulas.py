"""CPU checks of the differentiable evolution's hand-written parts (mpdo_amd.autodiff), with the
device eigensolver replaced by torch.linalg.eigh for the duration of each test (the product path
itself has no CPU eigensolver and refuses tensors off the GPU):

* eigh_backward against torch's autograd through torch.linalg.eigh, for losses of the kept
  eigenvectors' projector and the eigenvalues (a sign error, a transposed skew part or a missing
  diagonal term fails);
* the 4-qubit, 2-layer ansatz with chi and kappa truncating: <H> equals the oracle's and every
  one of its 38 gradient components matches central differences of the oracle (complex128)."""
import numpy as np
import pytest
import torch

import bridge


@pytest.fixture
def cpu_eigh(monkeypatch):
    from mpdo_amd import _lib

    def eigh(A):
        w, V = torch.linalg.eigh(A)
        return w.flip(-1), V.flip(-1)
    monkeypatch.setattr(_lib, 'eigh', eigh)


@pytest.mark.parametrize('n,k,seed', [(6, 3, 0), (9, 2, 1), (5, 5, 2)])
def test_eigh_backward_matches_torch(n, k, seed):
    from mpdo_amd import autodiff as AD
    g = torch.Generator().manual_seed(seed)
    X = torch.randn(n, n, dtype=torch.complex128, generator=g)
    A = (X @ X.mH).requires_grad_(True)
    M = torch.randn(n, n, dtype=torch.complex128, generator=g)
    c = torch.arange(1.0, k + 1, dtype=torch.float64)

    def loss(w, V):
        P = V[:, :k] @ V[:, :k].mH
        return (P * M).sum().real + (w[:k] * c).sum()
    w, V = torch.linalg.eigh(A)
    loss(w.flip(-1), V.flip(-1)).backward()
    with torch.no_grad():
        w0, V0 = torch.linalg.eigh(A)
        w0, V0 = w0.flip(-1), V0.flip(-1)
    wv, Vv = w0.clone().requires_grad_(True), V0.clone().requires_grad_(True)
    gw, gV = torch.autograd.grad(loss(wv, Vv), (wv, Vv))
    got = AD.eigh_backward(w0, V0, gw, gV, torch.tensor(0.0, dtype=torch.float64))
    assert (got - A.grad).abs().max() <= 1e-12 * A.grad.abs().max()


def test_hea_gradient_matches_oracle_differences(cpu_eigh):
    from mpdo_amd import autodiff as AD
    n, depth, chi, kappa = 4, 2, 2, 2
    rng = np.random.default_rng(1)
    theta = rng.uniform(-np.pi, np.pi, (depth, n))
    phi = rng.uniform(-np.pi, np.pi, (depth, n))
    p1 = rng.uniform(0.005, 0.02, (depth, n, 2))
    p2 = rng.uniform(0.005, 0.02, (depth, n - 1))
    m = AD.HEA(n, depth, chi, kappa, theta, phi, p1, p2, device='cpu')
    E = AD.tfim_energy(m())
    E.backward()
    e_ref, _ = bridge.run_hea_oracle(theta, phi, p1, p2, chi, kappa)
    assert abs(E.item() - e_ref) <= 1e-12
    got, want = [], []
    for name, par in [('theta', m.theta), ('phi', m.phi), ('p1', m.p1), ('p2', m.p2)]:
        gr = par.grad.numpy()
        for idx in np.ndindex(gr.shape):
            h = 1e-4 if name in ('theta', 'phi') else 1e-5
            vals = []
            for sgn in (1, -1):
                args = dict(theta=theta.copy(), phi=phi.copy(), p1=p1.copy(), p2=p2.copy())
                args[name][idx] += sgn * h
                vals.append(bridge.run_hea_oracle(args['theta'], args['phi'], args['p1'], args['p2'], chi, kappa)[0])
            got.append(gr[idx])
            want.append((vals[0] - vals[1]) / (2 * h))
    got, want = np.array(got), np.array(want)
    assert got.shape == (38,)
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()
