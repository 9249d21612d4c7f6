"""Fused brickwork layers (mpdo_layer_batched through batched.evolve_batched) against the per-gate
path of every copy and against the oracle: density matrix within 1e-5 of max|rho|, the chi
cuts' discarded weights in the same order within 1e-4, kappa totals within 1e-4, gate-time split
weights (at most 2 x 2.7e-8 x |theta| by the split's error bound) within 1e-3 or 1e-12."""
import numpy as np
import pytest
import torch

from oracle import mpdo_oracle as O

import bridge

pytestmark = pytest.mark.gpu


def _family(n, depth, copies, chi, kappa, noise, seed=17):
    """Copies of one brickwork circuit whose angles differ by U[-0.05, 0.05]."""
    from mpdo_amd import workloads as W
    out = []
    for c in range(copies):
        rng = np.random.Generator(np.random.PCG64(seed))
        delta = None if c == 0 else np.random.Generator(np.random.PCG64(seed * 100 + c)).uniform(-0.05, 0.05,
                                                                                                    (depth, n, 2))
        ops = W.brickwork(n, depth, rng, delta=delta)
        if noise == 'dep':
            w = W.Workload(f'f{c}', n, ops, 'idealNoise', False, chi, kappa, chip=dict(dpc_error_rate=0.03),
                           single_qubit_noise='depolarizing', p1=0.01, depth=depth)
        elif noise == 'decay':
            w = W.Workload(f'f{c}', n, ops, 'idealNoise', False, chi, kappa,
                           chip=dict(dpc_error_rate=0.02, decay_param=0.03, dephasing_param=0.02), depth=depth)
        else:
            w = W.Workload(f'f{c}', n, ops, 'no', True, chi, kappa, depth=depth)
        out.append(w)
    return out


def _circuits(ws, cuda):
    import mpdo_amd as M
    from mpdo_amd import workloads as W
    cs = []
    for w in ws:
        c = M.TensorCircuit(device=str(cuda), **w.circuit_kwargs())
        W.apply_ops(c, w.ops)
        cs.append(c)
    return cs


def _rho_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _cuts(disc):
    return [d for d in disc if d[0] == 'chi'], sum(d[2] for d in disc if d[0] == 'kappa')


@pytest.mark.parametrize('n,depth,copies,chi,kappa,noise', [
    (6, 6, 3, 8, 2, 'dep'),
    (7, 5, 2, 16, 4, 'dep'),
    (8, 6, 2, 32, 4, 'dep'),
    (6, 5, 3, 8, 3, 'decay'),
    (9, 4, 2, 4, 2, 'dep'),
])
def test_fused_layers_match_per_gate_path(cuda, n, depth, copies, chi, kappa, noise):
    import mpdo_amd as M
    from mpdo_amd import batched
    ws = _family(n, depth, copies, chi, kappa, noise)
    fused = _circuits(ws, cuda)
    states = [M.Tools.create_ket0Series(n, device=str(cuda)) for _ in ws]
    stats = {}
    batched.evolve_batched(fused, states, stats=stats)
    assert stats['fused_layers'] >= depth - 2, stats
    for w, cf in zip(ws, fused):
        cg = _circuits([w], cuda)[0]
        cg.evolve(M.Tools.create_ket0Series(n, device=str(cuda)))
        rho_f = cf.cal_dm().cpu().numpy()
        rho_g = cg.cal_dm().cpu().numpy()
        assert _rho_err(rho_f, rho_g) < 1e-5
        for a, b in zip(cf._stateNodes, cg._stateNodes):
            assert a.tensor.shape == b.tensor.shape
            assert (a.has_I, a.bond_left, a.bond_right) == (b.has_I, b.bond_left, b.bond_right)
        chi_f, kap_f = _cuts(cf.discarded)
        chi_g, kap_g = _cuts(cg.discarded)
        assert [d[:2] for d in chi_f] == [d[:2] for d in chi_g]
        assert np.allclose([d[2] for d in chi_f], [d[2] for d in chi_g], rtol=1e-4, atol=1e-7)
        assert abs(kap_f - kap_g) <= 1e-4 * abs(kap_g) + 1e-6
        gate_f = [d for d in cf.discarded if d[0] == 'gate']
        gate_g = [d for d in cg.discarded if d[0] == 'gate']
        assert [d[:2] for d in gate_f] == [d[:2] for d in gate_g]
        assert np.allclose([d[2] for d in gate_f], [d[2] for d in gate_g], rtol=1e-3, atol=1e-12)
    # the first copy against the oracle (complex128)
    st = bridge.run_oracle(ws[0])
    assert _rho_err(fused[0].cal_dm().cpu().numpy(), O.density_matrix(st)) < 1e-5


def test_unfusable_layers_fall_back(cuda):
    """Ideal circuits (no Kraus axes) leave the pair max sites too thin for the fused plan:
    every layer runs through the per-gate path and the result is unchanged."""
    import mpdo_amd as M
    from mpdo_amd import batched
    ws = _family(6, 4, 2, 4, None, 'ideal')
    cs = _circuits(ws, cuda)
    states = [M.Tools.create_ket0Series(6, device=str(cuda)) for _ in ws]
    stats = {}
    batched.evolve_batched(cs, states, stats=stats)
    assert stats['fused_layers'] == 0
    for w, c in zip(ws, cs):
        st = bridge.run_oracle(w)
        assert _rho_err(c.cal_dm().cpu().numpy(), O.density_matrix(st)) < 1e-5


def test_layer_batched_matches_entry_point_sequence(cuda):
    """One layer on synthetic steady-state chains (random tensors at the shapes a long circuit
    settles into) against mpdo_apply_1q + mpdo_apply_2q + truncation called per copy."""
    import mpdo_amd as M
    from mpdo_amd import _lib, batched
    from mpdo_amd import workloads as W
    from mpdo_amd.noise import NoiseChannel, choi_compress, compose_kraus
    from mpdo_amd import gates as G
    n, chi, kappa, B = 8, 16, 2, 3
    noise = NoiseChannel(chip=dict(dpc_error_rate=0.02), single_qubit_noise='depolarizing', p1=0.01)
    rng = np.random.Generator(np.random.PCG64(3))
    for parity in (0, 1):
        dims = batched.steady_dims(n, chi, kappa, 4, 16, parity)
        chains = [[torch.from_numpy(t).to(cuda) for t in W.steady_state(dims, seed=10 * parity + b)] for b in range(B)]
        k1 = torch.zeros(B, n, 4, 2, 2, dtype=torch.complex128)
        for b in range(B):
            for q in range(n):
                ry = G.RYGate(float(rng.uniform(0, 6.28)), False).tensor.to(torch.complex128)
                rz = G.RZGate(float(rng.uniform(0, 6.28)), False).tensor.to(torch.complex128)
                E = choi_compress(compose_kraus(noise.single_kraus(ry), noise.single_kraus(rz)))
                k1[b, q] = E.permute(2, 0, 1)
        g2 = noise.double_kraus(G.CNOTGate(False).tensor.to(torch.complex128))
        pairs = list(range(parity, n - 1, 2))
        gates2 = torch.stack([g2] * len(pairs))
        sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
        outs, disc = _lib.layer_batched(sites, k1, parity, gates2, chi, kappa)
        for b in range(B):
            nodes = [M.SiteNode(chains[b][i].clone(), i, has_I=True, bond_left=i > 0, bond_right=i < n - 1)
                     for i in range(n)]
            ts = _lib.apply_1q([nd.tensor for nd in nodes], [k1[b, q].permute(1, 2, 0) for q in range(n)])
            for nd, t in zip(nodes, ts):
                nd.tensor = t
            for p in pairs:
                a, c, _ = _lib.apply_2q(nodes[p].tensor, nodes[p + 1].tensor, g2, 2.718281828459045e-8)
                nodes[p].tensor, nodes[p + 1].tensor = a, c
            d_chi = M.TNNOptimizer.bondTruncate(nodes, chi, return_discarded=True)
            d_kap = M.TNNOptimizer.svdKappa_left2right(nodes, kappa, return_discarded=True)
            rho_f = _lib.reduced_dm([o[b].contiguous() for o in outs], [True] * n).cpu().numpy()
            rho_g = _lib.reduced_dm([nd.tensor.contiguous() for nd in nodes], [True] * n).cpu().numpy()
            assert _rho_err(rho_f, rho_g) < 1e-5, (parity, b)
            assert [tuple(o.shape[1:]) for o in outs] == [tuple(nd.tensor.shape) for nd in nodes]
            got = [disc[b, i] for i in reversed(range(n - 1))]
            assert np.allclose(got, d_chi, rtol=1e-4, atol=1e-9), (parity, b)
            assert abs(sum(disc[b, n + i] for i, _ in d_kap) - sum(v for _, v in d_kap)) <= 1e-4 * sum(
                v for _, v in d_kap) + 1e-9


def test_layers_split_over_streams_match_one_batch(cuda):
    """batched.layers_batched_streams (the chains as sub-batches, each advanced through every layer
    on its own HIP stream and host thread; 7 chains with four streams asked: sub-batches of 3, 3, 1,
    one stream fewer than the four hardware queues) against one call
    over all chains per layer on synthetic steady-state chains, two layers (both pair parities):
    same shapes, observables within 1e-6 of Tr rho~ per chain (GEMM tiles may depend on the batch
    size), discarded weights within 1e-4; one stream reproduces _lib.layer_batched exactly;
    evolve_batched with streams=3 against streams=1 on a noisy family (density matrices within
    1e-6)."""
    import mpdo_amd as M
    from mpdo_amd import _lib, batched
    from mpdo_amd import workloads as W
    n, chi, kappa, B = 8, 32, 4, 7
    ws = W.brickwork32_chi256_family(copies=B, n=n, depth=4)
    cs = _circuits(ws, cuda)
    segs = batched._segments(cs[0])
    codes = [[2 if q in (i, i + 1) else -1 for q in range(n)] for i in range(n - 1)] + [[-1] * n]
    layers = []
    for k in range(2):
        L, why = batched.fused_layer_of(cs, *segs[k + 1])
        assert L is not None, why
        layers.append(L)
    dims = batched.steady_dims(n, chi, kappa, layers[0].kraus1.shape[2], layers[0].gates2.shape[-1], layers[0].parity)
    chains = [[torch.from_numpy(t).to(cuda) for t in W.steady_state(dims, seed=40 + b)] for b in range(B)]
    sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
    one, d1 = sites, []
    for L in layers:
        one, d = _lib.layer_batched(one, L.kraus1, L.parity, L.gates2, chi, kappa)
        d1.append(d)
    same, ds = batched.layers_batched_streams(sites, layers, chi, kappa, streams=1)
    assert all(torch.equal(a, b) for a, b in zip(one, same))
    assert all(np.array_equal(a, b) for a, b in zip(d1, ds))
    split, d4 = batched.layers_batched_streams(sites, layers, chi, kappa, streams=4)
    torch.cuda.synchronize()
    assert [t.shape for t in split] == [t.shape for t in one]
    for b in range(B):
        v1 = np.array(_lib.pauli_expect([t[b].contiguous() for t in one], codes))
        v4 = np.array(_lib.pauli_expect([t[b].contiguous() for t in split], codes))
        assert np.abs(v1 - v4).max() <= 1e-6 * abs(v1[-1]), (b, np.abs(v1 - v4).max())
    for a, b in zip(d4, d1):
        assert np.allclose(a, b, rtol=1e-4, atol=1e-10)
    fam = _family(7, 5, 5, 16, 4, 'dep')
    c1, c3 = _circuits(fam, cuda), _circuits(fam, cuda)
    s1 = [M.Tools.create_ket0Series(7, device=str(cuda)) for _ in fam]
    s3 = [M.Tools.create_ket0Series(7, device=str(cuda)) for _ in fam]
    st = {}
    batched.evolve_batched(c1, s1)
    batched.evolve_batched(c3, s3, streams=3, stats=st)
    assert st['fused_layers'] >= 3, st
    for a, b in zip(c1, c3):
        assert _rho_err(b.cal_dm().cpu().numpy(), a.cal_dm().cpu().numpy()) < 1e-6
    batched.release_stream_pools(cuda)
