"""The gate-time cut decided from the bottom of the spectrum (mpdo_debug_gram_tail_cut), against the
absolute-error rule of decompositions.svd (decompositions.py:121-131 with Circuit.py:120-124's
error 2.718e-8: te = sqrt(cumsum s^2) over the descending singular values, keep idx + 1 for the
first idx with te[-1] - te[idx] <= eps) applied to numpy's full eigendecomposition.

The Gram matrices have random eigenvectors and the spectra measured at the benchmark's gate-time
cuts: a dense bottom (neighbouring eigenvalues 1 - 3 % apart), 5 to 129 eigenvalues below the
largest droppable one, 0 to 23 of them dropped, the trace at the cuts' scale (D* / tr ~ 5e-9).
Bars: kept counts equal; dropped weights within 1e-8 relative (Ritz values against LAPACK's
eigenvalues, ~1e-15 absolute at these scales); P = I - V_d V_d^H within 1e-5 elementwise
(complex64, and a dropped subspace resolved against a 1 % eigenvalue gap); matrices whose bottom is
a large null space are declined (the subspace cannot hold the dropped set).

The chi step's top-k solver (mpdo_debug_gram_top_eig) on the chi-step spectra measured at the
benchmark's pair bonds (k = 256 of 1024, neighbouring values at the cut 1 % apart, the 513th value
~3 % of the 256th): the k leading eigenvalues within 1e-10 of the largest, the remainder (the
chi step's discarded weight) within 1e-9 relative, V_k diag(lambda_k) V_k^H within 1e-9 of the
largest eigenvalue elementwise; a degenerate pair at the cut is declined."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
EPS = 2.718281828459045e-8


def _keep_rule(w_desc, eps=EPS):
    s = np.sqrt(np.maximum(w_desc, 0.0))
    te = np.sqrt(np.cumsum(s ** 2))
    for i in range(len(s) - 1):
        if te[-1] - te[i] <= eps:
            return i + 1
    return len(s)


def _gram(rng, n, lo, c1, c10, c100, dsrel):
    """Hermitian PSD n x n with a log-piecewise-linear ascending spectrum through (0, lo), (c1, D*),
    (c10, 10 D*), (c100, 100 D*), (n-1, 0.07) (relative to the trace), scaled so that D* / tr = dsrel."""
    xs = [0, c1, c10, c100, n - 1]
    ys = np.log([lo, dsrel, 10 * dsrel, 100 * dsrel, 0.07])
    w = np.exp(np.interp(np.arange(n), xs, ys)) * (1 + 0.02 * rng.standard_normal(n))
    w = np.sort(w)
    w = w / w.sum() * (2 * EPS / dsrel) ** 2
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    G = (U * w) @ U.conj().T
    return (G + G.conj().T) / 2


def _run(G, m=128, q=5):
    from mpdo_amd import _lib
    import ctypes
    lib = _lib.load()
    B, n = G.shape[0], G.shape[1]
    Gd = torch.from_numpy(np.ascontiguousarray(G)).to('cuda')
    P = torch.zeros((B, n, n), dtype=torch.complex64, device='cuda')
    kept = (ctypes.c_int * B)()
    disc = (ctypes.c_double * B)()
    info = (ctypes.c_int * 4)()
    st = lib.mpdo_debug_gram_tail_cut(n, B, ctypes.c_void_p(Gd.data_ptr()), EPS, m, q, kept, disc,
                                      ctypes.c_void_p(P.data_ptr()), info, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(st, 'mpdo_debug_gram_tail_cut')
    return list(info), list(kept), list(disc), P.cpu().numpy()


def test_tail_cut_matches_the_full_rule(cuda):
    rng = np.random.default_rng(5)
    n = 1024
    specs = [(7.5e-10, 129, 355, 512, 6.9e-9), (3.5e-9, 14, 189, 399, 5.2e-9), (5.8e-11, 8, 92, 265, 4.8e-9)]
    G = np.stack([_gram(rng, n, *sp) for sp in specs])
    info, kept, disc, P = _run(G)
    print('info', info, 'kept', kept)
    assert info[0] == 1 and info[3] == 0, info
    for b in range(len(specs)):
        w, V = np.linalg.eigh(G[b])
        w_desc = w[::-1]
        k = _keep_rule(w_desc)
        d = n - k
        assert kept[b] == k, (b, kept[b], k)
        ref_disc = float(np.maximum(w[:d], 0).sum())
        assert abs(disc[b] - ref_disc) <= 1e-8 * ref_disc, (b, disc[b], ref_disc)
        Pt = np.eye(n) - V[:, :d] @ V[:, :d].conj().T
        assert np.abs(P[b] - Pt).max() <= 1e-5, (b, np.abs(P[b] - Pt).max())
    assert info[1] == 1 and info[2] == max(kept)


def test_tail_cut_declines_a_null_bottom(cuda):
    """Half the spectrum exactly zero: every subspace vector is droppable, so the subspace cannot
    show where the dropped set ends and the full eigensolver must decide."""
    rng = np.random.default_rng(6)
    n = 512
    w = np.concatenate([np.zeros(n // 2), np.logspace(-6, 0, n // 2)])
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    G = (U * w) @ U.conj().T
    G = ((G + G.conj().T) / 2)[None]
    info, _, _, _ = _run(G)
    assert info[0] == 0 and info[3] != 0, info


CHI_SPEC = [(0, 6.6e-2), (7, 3.1e-2), (31, 8.5e-3), (127, 4.8e-4), (191, 1.2e-4), (255, 3.5e-5), (256, 3.4e-5),
            (383, 5.1e-6), (511, 1.2e-6), (767, 7.6e-8), (1023, 1e-9)]


def _chi_gram(rng, n, tie=False):
    xs, ys = zip(*CHI_SPEC)
    w = np.exp(np.interp(np.arange(n), xs, np.log(ys)))[::-1] * (1 + 0.003 * rng.standard_normal(n))
    w = np.sort(w)
    if tie:
        w[n - 257] = w[n - 256]   # the 256th and 257th largest equal
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    G = (U * w) @ U.conj().T * 40.0
    return (G + G.conj().T) / 2


def _run_top(G, k, m, q=4):
    from mpdo_amd import _lib
    import ctypes
    lib = _lib.load()
    B, n = G.shape[0], G.shape[1]
    Gd = torch.from_numpy(np.ascontiguousarray(G)).to('cuda')
    w = torch.zeros((B, n), dtype=torch.float64, device='cuda')
    V = torch.zeros((B, n, n), dtype=torch.complex128, device='cuda')
    info = (ctypes.c_int * 2)()
    st = lib.mpdo_debug_gram_top_eig(n, B, ctypes.c_void_p(Gd.data_ptr()), k, m, q, ctypes.c_void_p(w.data_ptr()),
                                     ctypes.c_void_p(V.data_ptr()), info,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(st, 'mpdo_debug_gram_top_eig')
    return list(info), w.cpu().numpy(), V.cpu().numpy()


def test_top_solver_matches_the_full_decomposition(cuda):
    rng = np.random.default_rng(8)
    n, k, m = 1024, 256, 512
    G = np.stack([_chi_gram(rng, n) for _ in range(2)])
    info, w, V = _run_top(G, k, m)
    assert info == [1, 0], info
    for b in range(G.shape[0]):
        lt, Vt = np.linalg.eigh(G[b])
        lt, Vt = lt[::-1], Vt[:, ::-1]
        assert np.abs(w[b, :k] - lt[:k]).max() <= 1e-10 * lt[0]
        rem = w[b, k:].sum()
        assert abs(rem - lt[k:].sum()) <= 1e-9 * lt[k:].sum(), (rem, lt[k:].sum())
        A = (V[b][:, :k] * w[b, :k]) @ V[b][:, :k].conj().T
        At = (Vt[:, :k] * lt[:k]) @ Vt[:, :k].conj().T
        assert np.abs(A - At).max() <= 1e-9 * lt[0], np.abs(A - At).max() / lt[0]


def test_top_solver_declines_a_tie_at_the_cut(cuda):
    rng = np.random.default_rng(9)
    G = _chi_gram(rng, 1024, tie=True)[None]
    info, _, _ = _run_top(G, 256, 512)
    assert info[0] == 0 and info[1] != 0, info
