"""Not collected until it has passed on the GPU: the pivoted-Cholesky-preconditioned Gram
eigensolver (mpdo_debug_eigh_gram)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('n,rank', [(100, 100), (256, 256), (512, 300), (1024, 450), (200, 40)])
def test_eigh_gram(cuda, n, rank):
    """The pivoted-Cholesky-preconditioned solver on graded Gram matrices (12 decades) of full and
    deficient rank: eigenvalues above its floor as numpy's, eigenvectors of those (residual and
    orthonormality relative to the largest eigenvalue), zero vectors for the rest."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(n + rank)
    X = _crand(rng, 3, n, rank) * np.logspace(0, -6, rank)[None, None, :]
    H = X @ np.swapaxes(X.conj(), 1, 2)
    H = 0.5 * (H + np.swapaxes(H.conj(), 1, 2))
    w, V, sweeps = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda), gram=True)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert sweeps < 40
    floor = 1e-14 * max(1, ((n + 63) // 64 * 64) // 64)
    for b in range(3):
        ref = np.sort(np.linalg.eigvalsh(H[b]))[::-1]
        top = ref[0]
        big = ref > 100 * floor * top
        assert np.all(np.diff(w[b]) <= 1e-12 * top)
        assert np.abs(w[b][big] - ref[big]).max() <= 1e-12 * top
        assert np.all(np.abs(w[b]) <= ref[0] * (1 + 1e-12))
        keep = w[b] > 1e-9 * top
        Vk = V[b][:, keep]
        assert np.abs(H[b] @ Vk - Vk * w[b][keep][None, :]).max() <= 1e-11 * top
        assert np.abs(Vk.conj().T @ Vk - np.eye(Vk.shape[1])).max() <= 1e-6
        zero = w[b] == 0
        assert np.all(V[b][:, zero] == 0)
        assert zero.sum() >= n - rank

