"""End-to-end parity: circuits built with the TensorCircuit API and evolved on the GPU vs the
oracle's evolution of the same layer list (complex128), compared on the density matrix
(1e-5 relative to max|rho|) and the discarded weights (1e-4 relative)."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import mpdo_oracle as O

import bridge

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _run_product(w, cuda, **kw):
    import mpdo_amd as M
    from mpdo_amd import workloads as W
    args = w.circuit_kwargs()
    args.update(kw)
    c = M.TensorCircuit(device=str(cuda), **args)
    W.apply_ops(c, w.ops)
    state = M.Tools.create_ket0Series(w.n, device=str(cuda))
    c.evolve(state)
    return c, state


def _rho_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _random_workload(seed, n, depth, noise, chi, kappa, two=('CNOT',)):
    from mpdo_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = W.brickwork(n, depth, rng, two=two, noisy_identity_after_2q=(noise == 'unified'))
    if noise == 'idealNoise':
        return W.Workload('t', n, ops, 'idealNoise', False, chi, kappa, chip=dict(dpc_error_rate=0.03),
                          single_qubit_noise='depolarizing', p1=0.01, depth=depth)
    if noise == 'idealNoise-decay':
        return W.Workload('t', n, ops, 'idealNoise', False, chi, kappa,
                          chip=dict(dpc_error_rate=0.02, decay_param=0.03, dephasing_param=0.02), depth=depth)
    if noise == 'unified':
        return W.Workload('t', n, ops, 'unified', False, chi, kappa,
                          chip=dict(decay_param=0.02, dephasing_param=0.01), depth=depth)
    return W.Workload('t', n, ops, 'no', True, chi, kappa, depth=depth)


@pytest.mark.parametrize('noise', ['ideal', 'idealNoise', 'idealNoise-decay', 'unified'])
@pytest.mark.parametrize('fuse', [True, False])
def test_exact_regime_vs_dense(cuda, noise, fuse):
    """chi and kappa large enough that nothing truncates: the MPDO is exact, so the GPU result
    must equal the dense density-matrix evolution."""
    w = _random_workload(5, 4, 2, noise, None, None, two=('CNOT', 'CZ'))
    c, _ = _run_product(w, cuda, fuse_single_qubit=fuse)
    rho = c.cal_dm().cpu().numpy()
    ref = D.dense_evolve(bridge.oracle_ops(w), w.n, w.noise_type, w.ideal, bridge.oracle_noise(w))
    assert _rho_err(rho, ref) < 1e-5


@pytest.mark.parametrize('noise', ['ideal', 'idealNoise', 'idealNoise-decay', 'unified'])
@pytest.mark.parametrize('chi,kappa', [(4, 2), (6, 3), (8, 4)])
def test_truncating_vs_oracle(cuda, noise, chi, kappa):
    w = _random_workload(7, 6, 4, noise, chi, kappa)
    c, state = _run_product(w, cuda)
    st = bridge.run_oracle(w)
    rho = c.cal_dm().cpu().numpy()
    ref = O.density_matrix(st)
    assert _rho_err(rho, ref) < 1e-5
    _check_discarded(c.discarded, st.discarded)


def _check_discarded(got_all, ref_all):
    """chi cuts: same sequence and weights. kappa cuts: a site the fused Kraus path leaves at
    or below kappa (so it is not cut) must be one where the reference discards ~0."""
    got = [(k, i, v) for (k, i, v) in got_all if k == 'chi']
    ref = [d for d in ref_all if d[0] == 'chi']
    assert [(k, i) for k, i, _ in got] == [(k, i) for k, i, _ in ref]
    assert np.allclose([v for *_, v in got], [v for *_, v in ref], rtol=1e-4, atol=1e-7)
    # kappa entries grouped per truncation (the reference emits them in site order per sweep)
    gk = [(i, v) for (k, i, v) in got_all if k == 'kappa']
    rk = [(i, v) for (k, i, v) in ref_all if k == 'kappa']
    total_g, total_r = sum(v for _, v in gk), sum(v for _, v in rk)
    assert abs(total_g - total_r) <= 1e-4 * abs(total_r) + 1e-6


def test_realnoise_ghz(cuda):
    """The reference's smoke script (test/debug.py): 5-qubit GHZ under realNoise, chip 'best',
    chi = kappa = 4, with the tomographed CZ channel of czDefault.mat."""
    import mpdo_amd as M
    n = 5
    chi_files = {'CZ': {f'{i}{i + 1}': os.path.join(HERE, 'golden', 'czDefault.mat') for i in range(n - 1)}, 'CP': {}}
    c = M.TensorCircuit(n, ideal=False, noiseType='realNoise', chiFileDict=chi_files, chi=4, kappa=4, chip='best',
                        device=str(cuda))
    c.h(0)
    for i in range(n - 1):
        c.cnot(i, i + 1)
    c.truncate()
    state = M.Tools.create_ket0Series(n, device=str(cuda))
    c.evolve(state)
    rho = c.cal_dm().cpu().numpy()
    # oracle: same layer list
    exp = {f'{i}{i + 1}': O.chi_to_kraus(O.load_chi(os.path.join(HERE, 'golden', 'czDefault.mat'))) for i in range(n - 1)}
    b = O.CircuitBuilder(n, 'realNoise', exp)
    b.add('H', [0])
    for i in range(n - 1):
        b.add('CNOT', [i, i + 1])
    b.truncate()
    st = O.product_state(n)
    O.evolve(b.ops, st, 'realNoise', False, O.NoiseSpec(chip='best'), chi=4, kappa=4)
    assert _rho_err(rho, O.density_matrix(st)) < 1e-5


def test_realnoise_builder_rotations(cuda):
    """The rest of the builder under realNoise (AbstractCircuit.py:231-342): rxx/ryy/rzz/cx/cnot
    expanded into CX/CZEXP lists with the tomographed czDefault channel, measured-basis
    rotations, chi and kappa truncating; density matrix vs the oracle evolving the oracle
    builder's list."""
    import mpdo_amd as M
    n = 4
    cz = os.path.join(HERE, 'golden', 'czDefault.mat')
    files = {'CZ': {f'{i}{i + 1}': cz for i in range(n - 1)}, 'CP': {}}
    c = M.TensorCircuit(n, ideal=False, noiseType='realNoise', chiFileDict=files, chi=6, kappa=3, chip='best',
                        device=str(cuda))
    exp = {f'{i}{i + 1}': O.chi_to_kraus(O.load_chi(cz)) for i in range(n - 1)}
    b = O.CircuitBuilder(n, 'realNoise', exp)
    for q in range(n):
        c.h(q); b.add('H', [q])
    c.rxx(torch.tensor(0.4), 0, 1); b.add('RXX', [0, 1], [0.4])
    c.ryy(torch.tensor(-0.7), 2, 3); b.add('RYY', [2, 3], [-0.7])
    c.truncate(); b.truncate()
    c.rzz(torch.tensor(1.3), 1, 2); b.add('RZZ', [1, 2], [1.3])
    c.cx(3, 2); b.add('CX', [3, 2])
    c.truncate(); b.truncate()
    state = M.Tools.create_ket0Series(n, device=str(cuda))
    c.evolve(state)
    rho = c.cal_dm().cpu().numpy()
    st = O.product_state(n)
    O.evolve(b.ops, st, 'realNoise', False, O.NoiseSpec(chip='best'), chi=6, kappa=3)
    assert _rho_err(rho, O.density_matrix(st)) < 1e-5


def test_brickwork8_chi32_golden(cuda):
    """8 qubits, 4 brickwork layers, dep1 p=1e-3, dep2 p=1e-2, chi=32, kappa=4; the
    full 256x256 density matrix against the oracle fixture (tests/golden/make_golden.py)."""
    from mpdo_amd import workloads as W
    g = np.load(os.path.join(HERE, 'golden', 'brickwork8_chi32_oracle.npz'))
    w = W.brickwork8_chi32()
    c, state = _run_product(w, cuda)
    rho = c.cal_dm().cpu().numpy()
    assert _rho_err(rho, g['rho']) < 1e-5
    assert abs(np.trace(rho).real - float(g['trace'])) < 1e-5
    ref = list(zip([str(k) for k in g['disc_kind']], g['disc_index'].tolist(), g['disc_weight'].tolist()))
    _check_discarded(c.discarded, ref)
