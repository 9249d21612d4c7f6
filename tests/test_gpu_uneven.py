"""Uneven batches (scheduler.evolve_many): circuits of different qubit counts, depths, chi, kappa and
noise advanced together on several streams give, family by family, what the same families give run
one after the other on one stream (same kernels, same order per circuit: to the last bit), and
what TensorCircuit.evolve gives per circuit (the per-gate path; the fused layers agree with it to
the parity bar, 1e-5 of Tr rho~); the first 16 circuits of
the 256-circuit batch (workloads.random_brickwork_batch) match the oracle fixture
(tests/golden/make_golden.py batch16): every <Z_i> and Tr rho~ within 1e-5 of Tr rho~."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _circuits(ws, dev):
    import mpdo_amd as M
    from mpdo_amd import workloads as W
    cs = []
    for w in ws:
        c = M.TensorCircuit(device=str(dev), **w.circuit_kwargs())
        W.apply_ops(c, w.ops)
        cs.append(c)
    return cs


def _z_and_trace(state):
    from mpdo_amd import _lib
    n = len(state)
    codes = []
    for i in range(n):
        c = [-1] * n
        c[i] = 2
        codes.append(c)
    codes.append([-1] * n)
    v = np.array(_lib.pauli_expect([nd.tensor.contiguous() for nd in state], codes))
    return v[:-1], v[-1]


def test_evolve_many_matches_one_at_a_time(cuda):
    """A mix: eight circuits of the uneven batch plus a family of three angle-perturbed copies of one
    brickwork circuit (one batched fused call per layer inside the scheduler)."""
    import mpdo_amd as M
    from mpdo_amd import batched, scheduler
    from mpdo_amd import workloads as W
    ws = [w for w in W.random_brickwork_batch()[:24] if w.n <= 14][:8]
    fam = W.brickwork32_chi256_family(copies=3, n=8, depth=5)
    for w in fam:
        w.chi, w.kappa = 16, 4
    ws = ws + fam
    many = _circuits(ws, cuda)
    states = [M.Tools.create_ket0Series(w.n, device=str(cuda)) for w in ws]
    stats = {}
    scheduler.evolve_many(many, states, streams=4, stats=stats)
    assert stats['families'] == len(ws) - 2 and stats['streams'] == 4
    assert stats['fused_layers'] > 0
    # the same families one after the other on the current stream
    seq = _circuits(ws, cuda)
    seq_states = [M.Tools.create_ket0Series(w.n, device=str(cuda)) for w in ws]
    for g in scheduler.families(seq):
        batched.evolve_batched([seq[i] for i in g], [seq_states[i] for i in g])
    for w, a, b in zip(ws, states, seq_states):
        for x, y in zip(a, b):
            assert torch.equal(x.tensor, y.tensor), w.name
    # against the per-gate path of each circuit
    for w, st in zip(ws, states):
        c = _circuits([w], cuda)[0]
        ref = M.Tools.create_ket0Series(w.n, device=str(cuda))
        c.evolve(ref)
        z_a, tr_a = _z_and_trace(st)
        z_b, tr_b = _z_and_trace(ref)
        assert [tuple(a.tensor.shape) for a in st] == [tuple(b.tensor.shape) for b in ref], w.name
        assert abs(tr_a - tr_b) <= 1e-5 * abs(tr_b) and np.abs(z_a - z_b).max() <= 1e-5 * abs(tr_b), w.name


@pytest.mark.timeout(900)
def test_first16_of_the_uneven_batch_vs_oracle(cuda):
    import mpdo_amd as M
    from mpdo_amd import scheduler
    from mpdo_amd import workloads as W
    g = np.load(os.path.join(HERE, 'golden', 'random_brickwork16_oracle.npz'))
    ws = W.random_brickwork_batch()[:16]
    cs = _circuits(ws, cuda)
    states = [M.Tools.create_ket0Series(w.n, device=str(cuda)) for w in ws]
    stats = {}
    scheduler.evolve_many(cs, states, streams=4, stats=stats)
    print(stats)
    off = g['offsets']
    for k, (w, st) in enumerate(zip(ws, states)):
        z, tr = _z_and_trace(st)
        tr_o = float(g['trace'][k])
        z_o = g['z'][off[k]:off[k + 1]]
        assert abs(tr - tr_o) <= 1e-5 * abs(tr_o), (w.name, tr, tr_o)
        assert np.abs(z - z_o).max() <= 1e-5 * abs(tr_o), (w.name, np.abs(z - z_o).max() / abs(tr_o))
