"""The non-convergence contract (Circuit.py:118-136): an eigensolver that reports MPDO_ENOCONV
surfaces as torch.linalg.LinAlgError; the two-qubit gate split retries up to MAX_ATTEMPTS = 4 times
and then raises the reference's message; a truncation sweep raises at once. The failures are
forced from the host (mpdo_debug_fail_eig: the eigensolver runs to completion and its call then
reports non-convergence), so every kernel exits normally."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _circuit(cuda):
    import mpdo_amd as M
    c = M.TensorCircuit(4, ideal=False, noiseType='idealNoise', chi=4, kappa=2, device=str(cuda),
                        chip=dict(dpc_error_rate=0.02), single_qubit_noise='depolarizing', p1=0.01)
    for q in range(4):
        c.ry(0.3 + 0.2 * q, [q])
    c.cnot(0, 1)
    c.cnot(2, 3)
    c.cnot(1, 2)
    c.truncate()
    return c


def _run(c, cuda):
    import mpdo_amd as M
    st = M.Tools.create_ket0Series(4, device=str(cuda))
    c.evolve(st)
    return c.cal_dm().cpu().numpy()


def test_split_retries_then_succeeds(cuda):
    """Two forced failures: the split's third attempt succeeds. The perturbation (eps64 x norms x
    GLOBAL_MINIMUM x 1e-2) is far below complex64 rounding, as the reference's is of its complex64
    theta, so the result is the unforced one."""
    from mpdo_amd import _lib
    lib = _lib.load()
    ref = _run(_circuit(cuda), cuda)
    try:
        lib.mpdo_debug_fail_eig(2)
        got = _run(_circuit(cuda), cuda)
    finally:
        lib.mpdo_debug_fail_eig(0)
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


def test_split_raises_after_four_attempts(cuda):
    from mpdo_amd import _lib
    lib = _lib.load()
    try:
        lib.mpdo_debug_fail_eig(4)
        with pytest.raises(torch.linalg.LinAlgError, match='Multiple times \\(4\\) of perturbation'):
            _run(_circuit(cuda), cuda)
    finally:
        lib.mpdo_debug_fail_eig(0)


def test_truncation_raises_linalgerror(cuda):
    """bondTruncate's SVD sweep has no retry in the reference (TNNOptimizer.py:111-134): the
    first non-convergence raises."""
    import mpdo_amd as M
    from mpdo_amd import _lib
    lib = _lib.load()
    c = _circuit(cuda)
    st = M.Tools.create_ket0Series(4, device=str(cuda))
    c.evolve(st)
    nodes = c._stateNodes
    try:
        lib.mpdo_debug_fail_eig(1)
        with pytest.raises(torch.linalg.LinAlgError):
            M.TNNOptimizer.svd_right2left(nodes, 2)
    finally:
        lib.mpdo_debug_fail_eig(0)
