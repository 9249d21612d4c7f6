"""Parity of each C-ABI entry point (through the Python host) with the oracle function it
replaces, on seeded inputs at sizes the oracle finishes in seconds. complex64 on the GPU vs
complex128 in the oracle; compared on gauge-invariant quantities.

Tolerances: density matrices and expectation values to 1e-5 relative to max|rho| (the
north-star bound); discarded truncation weights to 1e-4 relative (+1e-7 absolute)."""
import numpy as np
import pytest
import torch

from oracle import mpdo_oracle as O

pytestmark = pytest.mark.gpu
RHO_TOL = 1e-5


def _crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def random_mpdo(rng, dims, decay=0.5):
    """Chain of random site tensors T[l, 2, a, r] with geometrically decaying bond spectra,
    normalised to Tr rho = 1. dims: list of (l, a, r)."""
    sites = []
    for (l, a, r) in dims:
        T = _crand(rng, l, 2, a, r)
        T *= (decay ** np.arange(r))[None, None, None, :]
        T *= (decay ** np.arange(l))[:, None, None, None]
        sites.append(T)
    st = O.MPDO(sites, [a > 1 for (_, a, _) in dims], [True] * (len(dims) - 1))
    nrm = O.trace_rho(st)
    f = nrm ** (-0.5 / len(sites))
    st.sites = [t * f for t in st.sites]
    return st


def to_dev(st, dev):
    return [torch.from_numpy(t.astype(np.complex64)).to(dev).contiguous() for t in st.sites]


def from_dev(ts):
    return O.MPDO([t.cpu().numpy().astype(np.complex128) for t in ts], [True] * len(ts), [True] * (len(ts) - 1))


def rho_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_apply_1q(cuda):
    from mpdo_amd import _lib
    rng = np.random.default_rng(1)
    T = _crand(rng, 3, 2, 5, 4)
    E = _crand(rng, 2, 2, 3)
    out = _lib.apply_1q([torch.from_numpy(T.astype(np.complex64)).to(cuda)], [torch.from_numpy(E)])[0]
    ref = O.apply_1q(T, E)
    assert out.shape == ref.shape
    assert np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-6


@pytest.mark.parametrize('dims,K', [(((1, 1, 1), (1, 1, 1)), 1), (((3, 2, 5), (5, 4, 2)), 16),
                                    (((4, 8, 6), (6, 3, 7)), 16), (((16, 4, 32), (32, 4, 16)), 16),
                                    (((2, 1, 3), (3, 1, 1)), 1)])
def test_apply_2q(cuda, dims, K):
    from mpdo_amd import _lib
    rng = np.random.default_rng(sum(sum(d) for d in dims) + K)
    Tmin = _crand(rng, dims[0][0], 2, dims[0][1], dims[0][2])
    Tmax = _crand(rng, dims[1][0], 2, dims[1][1], dims[1][2])
    G = _crand(rng, 2, 2, 2, 2, K) * 0.3
    a, b, disc = _lib.apply_2q(torch.from_numpy(Tmin.astype(np.complex64)).to(cuda),
                               torch.from_numpy(Tmax.astype(np.complex64)).to(cuda),
                               torch.from_numpy(G), O.GLOBAL_MINIMUM)
    ra, rb, rest = O.apply_2q(Tmin, Tmax, G, explicit_theta=True)
    # gauge-invariant: theta = T_min' T_max' and the bond spectrum
    th = np.tensordot(a.cpu().numpy(), b.cpu().numpy(), axes=([3], [0]))
    th_ref = np.tensordot(ra, rb, axes=([3], [0]))
    assert np.abs(th - th_ref).max() / np.abs(th_ref).max() < 2e-6
    # sqrt(S) split: T_min'^H T_min' = diag(s) = T_max' T_max'^H
    l, _, a1, k = a.shape
    A = a.cpu().numpy().reshape(-1, k).astype(np.complex128)
    B = b.cpu().numpy().reshape(k, -1).astype(np.complex128)
    s_ref = np.linalg.svd(th_ref.reshape(A.shape[0], -1), compute_uv=False)
    ga, gb = A.conj().T @ A, B @ B.conj().T
    assert np.abs(ga - np.diag(np.diag(ga))).max() < 1e-5 * s_ref[0]
    assert np.abs(np.diag(ga).real - s_ref[:k]).max() < 1e-5 * s_ref[0]
    assert np.abs(np.diag(gb).real - s_ref[:k]).max() < 1e-5 * s_ref[0]
    assert k == ra.shape[3]


def _chain(rng, n, D, a):
    dims = [(1 if i == 0 else D, a, 1 if i == n - 1 else D) for i in range(n)]
    return random_mpdo(rng, dims)


def test_qr_left2right(cuda):
    from mpdo_amd import _lib
    rng = np.random.default_rng(3)
    st = random_mpdo(rng, [(1, 3, 8), (8, 2, 12), (12, 4, 5), (5, 1, 9), (9, 2, 1)])
    out = _lib.qr_left2right(to_dev(st, cuda))
    ref = st.copy()
    O.qr_left2right(ref)
    for i, (t, r) in enumerate(zip(out, ref.sites)):
        assert t.shape == r.shape, i
        if i < len(out) - 1:
            M = t.cpu().numpy().reshape(-1, t.shape[3]).astype(np.complex128)
            g = M.conj().T @ M
            assert np.abs(g - np.eye(g.shape[0])).max() < 1e-5
    assert rho_err(O.density_matrix(from_dev(out)), O.density_matrix(ref)) < RHO_TOL


@pytest.mark.parametrize('direction', [0, 1])
@pytest.mark.parametrize('chi,err', [(None, None), (6, None), (4, 1e-2), (None, 1e-3)])
def test_svd_sweep(cuda, direction, chi, err):
    from mpdo_amd import _lib
    rng = np.random.default_rng(7 + direction)
    st = random_mpdo(rng, [(1, 2, 8), (8, 3, 12), (12, 2, 10), (10, 2, 6), (6, 2, 1)], decay=0.6)
    out, disc = _lib.svd_sweep(to_dev(st, cuda), direction, chi, err, False)
    ref = st.copy()
    (O.svd_right2left if direction == 0 else O.svd_left2right)(ref, chi, err)
    for t, r in zip(out, ref.sites):
        assert tuple(t.shape) == r.shape
    assert rho_err(O.density_matrix(from_dev(out)), O.density_matrix(ref)) < RHO_TOL
    dref = [d[2] for d in ref.discarded]
    assert np.allclose(disc, dref, rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize('chi,err', [(5, None), (3, 1e-2)])
def test_bond_truncate(cuda, chi, err):
    import mpdo_amd as M
    rng = np.random.default_rng(11)
    st = random_mpdo(rng, [(1, 2, 8), (8, 3, 12), (12, 2, 10), (10, 2, 6), (6, 2, 1)], decay=0.6)
    nodes = [M.SiteNode(t, i, has_I=True, bond_left=i > 0, bond_right=i < 4) for i, t in enumerate(to_dev(st, cuda))]
    disc = M.TNNOptimizer.bondTruncate(nodes, chi, err, return_discarded=True)
    ref = st.copy()
    O.bond_truncate(ref, chi, err)
    for nd, r in zip(nodes, ref.sites):
        assert nd.shape == r.shape
    assert rho_err(O.density_matrix(from_dev([n.tensor for n in nodes])), O.density_matrix(ref)) < RHO_TOL
    assert np.allclose(disc, [d[2] for d in ref.discarded], rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize('kappa,err', [(2, None), (3, 1e-2), (None, 1e-2)])
def test_svd_kappa(cuda, kappa, err):
    from mpdo_amd import _lib
    rng = np.random.default_rng(13)
    st = random_mpdo(rng, [(1, 5, 4), (4, 8, 6), (6, 3, 3), (3, 6, 1)], decay=0.7)
    out, disc = _lib.svd_kappa(to_dev(st, cuda), kappa, err)
    ref = st.copy()
    O.svd_kappa_left2right(ref, kappa, err)
    for t, r in zip(out, ref.sites):
        assert tuple(t.shape) == r.shape
    assert rho_err(O.density_matrix(from_dev(out)), O.density_matrix(ref)) < RHO_TOL
    dref = {d[1]: d[2] for d in ref.discarded}
    for i, d in enumerate(disc):
        if i in dref:
            assert np.isclose(d, dref[i], rtol=1e-4, atol=1e-7)


def test_observables(cuda):
    from mpdo_amd import _lib
    rng = np.random.default_rng(17)
    st = random_mpdo(rng, [(1, 2, 4), (4, 3, 6), (6, 2, 5), (5, 2, 3), (3, 2, 1)])
    ts = to_dev(st, cuda)
    rho = O.density_matrix(st)
    scale = np.abs(rho).max()
    # full and reduced density matrices
    assert rho_err(_lib.reduced_dm(ts, [True] * 5).cpu().numpy(), rho) < RHO_TOL
    red = O.density_matrix(st, [1, 3])
    assert rho_err(_lib.reduced_dm(ts, [True, False, True, False, True]).cpu().numpy(), red) < RHO_TOL
    # Pauli strings, projectors, trace
    codes = [[-1] * 5, [2, -1, -1, -1, -1], [0, 1, 2, 0, 1], [3, 4, 3, -1, 4]]
    got = _lib.pauli_expect(ts, codes)
    assert abs(got[0] - O.trace_rho(st)) < 1e-5
    assert abs(got[1] - O.pauli_expect(st, [2], [0])) < 1e-5
    assert abs(got[2] - O.pauli_expect(st, [0, 1, 2, 0, 1], [0, 1, 2, 3, 4])) < 1e-5
    P0 = np.diag([1.0, 0.0]); P1 = np.diag([0.0, 1.0])
    proj = np.kron(np.kron(np.kron(P0, P1), P0), P1)
    ref_p = O.expect(st, [proj], [[0, 1, 2, 4]])[0]
    assert abs(got[3] - ref_p) < 1e-5
    # Tr(rho0 rho1)
    st2 = random_mpdo(np.random.default_rng(19), [(1, 2, 3), (3, 2, 4), (4, 2, 2), (2, 3, 3), (3, 1, 1)])
    v = _lib.trace_rho_rho(ts, to_dev(st2, cuda))
    assert abs(v - O.trace_rho_rho(st, st2)) < 1e-5 * max(1.0, abs(O.trace_rho_rho(st, st2)))
