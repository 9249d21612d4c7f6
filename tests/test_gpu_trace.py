"""Tr(rho0 rho1) by the transfer sweep (mpdo_trace_rho_rho[_masked]) against the oracle: beyond
the 12 qubits a dense matrix allows, with reduced qubits, and the composited purity."""
import numpy as np
import pytest
import torch

from oracle import mpdo_oracle as O

pytestmark = pytest.mark.gpu


def _evolved(cuda, n, depth, chi, kappa, seed):
    import mpdo_amd as M
    from mpdo_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = W.brickwork(n, depth, rng)
    w = W.Workload('t', n, ops, 'idealNoise', False, chi, kappa, chip=dict(dpc_error_rate=0.03),
                   single_qubit_noise='depolarizing', p1=0.02, depth=depth)
    c = M.TensorCircuit(device=str(cuda), **w.circuit_kwargs())
    W.apply_ops(c, ops)
    c.evolve(M.Tools.create_ket0Series(n, device=str(cuda)))
    return c


def _oracle_state(c):
    st = c.stateNodes
    return O.MPDO([s.tensor.detach().cpu().numpy().astype(np.complex128) for s in st],
                  [s.has_I for s in st], [s.bond_right for s in st[:-1]])


@pytest.mark.parametrize('n,chi,reduced', [(6, 8, []), (16, 16, []), (16, 16, [0, 5, 6, 15]), (20, 32, [3])])
def test_trace_rho_rho_matches_oracle(cuda, n, chi, reduced):
    from mpdo_amd import dmOperations as dm
    c0 = _evolved(cuda, n, 6, chi, 2, 1)
    c1 = _evolved(cuda, n, 6, chi, 2, 2)
    o0, o1 = _oracle_state(c0), _oracle_state(c1)
    for ca, cb, oa, ob in [(c0, c1, o0, o1), (c0, c0, o0, o0)]:
        got = float(dm.trace_rho_rho(ca.cal_dmNodes(reduced), cb.cal_dmNodes(reduced)))
        want = O.trace_rho_rho(oa, ob, reduced)
        assert abs(got - want) <= 1e-9 * abs(want) + 1e-12, (got, want)
    got = float(dm.trace_composited_rho2(c0.cal_dmNodes(reduced), c1.cal_dmNodes(reduced)))
    assert got == pytest.approx(O.trace_composited_rho2([o0, o1], reduced), rel=1e-9)
