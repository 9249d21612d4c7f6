"""Parity of the configurations the benchmark and the 20-qubit damping workload run:

* the fused batched layer at kappa = 8 (chi = 64, small n): every copy against the per-gate path,
  the first copy against the oracle;
* one chi = 256, kappa = 8 layer of the benchmark's circuit family at 10 qubits from a
  numpy-seeded steady-state chain, against the committed oracle fixture
  (tests/golden/make_golden.py layer10) and the per-gate entry points;
* the full benchmark shape (32 qubits, chi = 256, kappa = 8, 32 copies), two layers from the
  synthetic steady state: copy 0 against the per-gate path (every <Z_i Z_j>, shapes, discarded
  weights) and batch invariance (a copy run alone gives what it gives in the batch of 32);
* the 20-qubit depth-20 amplitude-damping + dephasing circuit at chi = 64, kappa = 4 against its
  oracle fixture: 20 <Z_i>, 19 <Z_i Z_i+1>, 1024 bitstring probabilities, discarded weights.

Tolerances: expectation values and probabilities within 1e-5 of Tr rho~ (complex64 data), chi
and kappa discarded weights 1e-4 relative, gate-time split weights (the rule drops tails up to
about 2 x 2.7e-8 x |theta| in total) 1e-3 relative or 1e-12 absolute.
"""
import os

import numpy as np
import pytest
import torch

from oracle import mpdo_oracle as O

import bridge

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


def _rho_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _zz_codes(n, pairs=None):
    """Codes for <Z_i> (n strings), then <Z_i Z_j> for the given pairs (default every i < j),
    then the identity string (Tr rho~)."""
    codes, pairs = [], pairs if pairs is not None else [(i, j) for i in range(n) for j in range(i + 1, n)]
    for i in range(n):
        c = [-1] * n
        c[i] = 2
        codes.append(c)
    for i, j in pairs:
        c = [-1] * n
        c[i] = c[j] = 2
        codes.append(c)
    codes.append([-1] * n)
    return codes, pairs


def _check_layer_weights(disc_b, n, parity, gate, d_chi, d_kap):
    """disc_b: one chain's (3n,) discarded weights of mpdo_layer_batched."""
    got_chi = [disc_b[i] for i in reversed(range(n - 1))]
    assert np.allclose(got_chi, d_chi, rtol=1e-4, atol=1e-10)
    k_f = sum(disc_b[n + i] for i in range(n))
    k_g = sum(d_kap.values())
    assert abs(k_f - k_g) <= 1e-4 * k_g + 1e-10
    for p, v in gate.items():
        assert abs(disc_b[2 * n + p] - v) <= 1e-3 * abs(v) + 1e-12, (p, disc_b[2 * n + p], v)


# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize('n,depth,copies', [(8, 6, 3), (9, 5, 2)])
def test_fused_kappa8_chi64_family(cuda, n, depth, copies):
    """The benchmark's circuit family (RY, RZ + depolarizing p=0.002, CNOT + two-qubit
    depolarizing p=0.002) at chi=64, kappa=8: the kappa cuts of the pair max sites are order-128+
    Gram eigenproblems (pivoted Cholesky + block Jacobi), as at the benchmark shape. Every copy
    against the per-gate path; at 8 qubits copy 0 against the oracle fixture
    (tests/golden/make_golden.py family8)."""
    import mpdo_amd as M
    from mpdo_amd import batched
    from mpdo_amd import workloads as W
    ws = W.brickwork32_chi256_family(copies=copies, n=n, depth=depth)
    for w in ws:
        w.chi = 64
    fused = _circuits(ws, cuda)
    states = [M.Tools.create_ket0Series(n, device=str(cuda)) for _ in ws]
    stats = {}
    batched.evolve_batched(fused, states, stats=stats)
    assert stats['fused_layers'] >= depth - 2, stats
    assert max(s.tensor.shape[2] for s in states[0]) == 8
    for w, cf in zip(ws, fused):
        cg = _circuits([w], cuda)[0]
        cg.evolve(M.Tools.create_ket0Series(n, device=str(cuda)))
        assert _rho_err(cf.cal_dm().cpu().numpy(), cg.cal_dm().cpu().numpy()) < 1e-5
        for a, b in zip(cf._stateNodes, cg._stateNodes):
            assert a.tensor.shape == b.tensor.shape
        chi_f = [d for d in cf.discarded if d[0] == 'chi']
        chi_g = [d for d in cg.discarded if d[0] == 'chi']
        assert [d[:2] for d in chi_f] == [d[:2] for d in chi_g]
        assert np.allclose([d[2] for d in chi_f], [d[2] for d in chi_g], rtol=1e-4, atol=1e-10)
        kf = sum(d[2] for d in cf.discarded if d[0] == 'kappa')
        kg = sum(d[2] for d in cg.discarded if d[0] == 'kappa')
        assert abs(kf - kg) <= 1e-4 * kg + 1e-10
    if n != 8:
        return
    g = np.load(os.path.join(HERE, 'golden', 'family8_chi64_kappa8_oracle.npz'))
    assert _rho_err(fused[0].cal_dm().cpu().numpy(), g['rho']) < 1e-5
    assert [tuple(nd.tensor.shape) for nd in fused[0]._stateNodes] == [tuple(d) for d in g['dims']]
    kinds = [str(k) for k in g['disc_kind']]
    chi_o = [(k, int(i), float(v)) for k, i, v in zip(kinds, g['disc_index'], g['disc_weight']) if k == 'chi']
    chi_f = [d for d in fused[0].discarded if d[0] == 'chi']
    assert [d[:2] for d in chi_f] == [d[:2] for d in chi_o]
    assert np.allclose([d[2] for d in chi_f], [d[2] for d in chi_o], rtol=1e-4, atol=1e-10)
    ko = sum(v for k, v in zip(kinds, g['disc_weight']) if k == 'kappa')
    kf = sum(d[2] for d in fused[0].discarded if d[0] == 'kappa')
    assert abs(kf - ko) <= 1e-4 * ko + 1e-10


# ------------------------------------------------------------------------------------------
def _layer10_inputs(cuda, copies):
    """The 10-qubit steady-state chain of the fixture (copy 0) plus copies from other seeds, and
    the fused operands of the fixture's layer."""
    from mpdo_amd import batched
    from mpdo_amd import workloads as W
    import importlib.util
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(HERE, 'golden', 'make_golden.py'))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert batched.steady_dims(10, 256, 8, 4, 16, 1) == mg.LAYER10_DIMS
    chains = [[torch.from_numpy(t).to(cuda) for t in W.steady_state(mg.LAYER10_DIMS, seed=mg.LAYER10_SEED + 100 * b)]
              for b in range(copies)]
    ws = W.brickwork32_chi256_family(copies=copies, n=10)
    cs = _circuits(ws, cuda)
    start, stop = batched._segments(cs[0])[mg.LAYER10_LAYER]
    layer, why = batched.fused_layer_of(cs, start, stop)
    assert layer is not None, why
    return chains, layer


@pytest.mark.parametrize('mode', [0, 2, 4])
def test_layer10_chi256_kappa8_golden(cuda, mode):
    """One chi=256, kappa=8 layer at 10 qubits from the seeded steady state: the fused layer
    (two chains: the fixture's and one from another seed) against the oracle fixture and
    against the per-gate entry points; the site contractions with fp64 accumulation (mode 0), as
    3M products with fp32 accumulation (mode 2) and as bf16 matrix-core products of operands split
    into three bf16 parts (mode 4)."""
    from mpdo_amd import _lib
    lib = _lib.load()
    old = lib.mpdo_set_math_mode(mode)
    try:
        _layer10_check(cuda)
    finally:
        lib.mpdo_set_math_mode(old)


def _layer10_check(cuda):
    from mpdo_amd import _lib, batched
    g = np.load(os.path.join(HERE, 'golden', 'layer10_chi256_kappa8_oracle.npz'))
    n, B = 10, 2
    chains, L = _layer10_inputs(cuda, B)
    sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
    outs, disc = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
    assert [tuple(o.shape[1:]) for o in outs] == [tuple(d) for d in g['dims']]
    codes, pairs = _zz_codes(n)
    vals = np.array(_lib.pauli_expect([o[0].contiguous() for o in outs], codes))
    tr = float(g['trace'])
    assert abs(vals[-1] - tr) <= 1e-5 * tr
    assert np.abs(vals[:n] - g['z']).max() <= 1e-5 * tr
    zz = np.array([g['zz'][i, j] for i, j in pairs])
    assert np.abs(vals[n:-1] - zz).max() <= 1e-5 * tr
    kinds = [str(k) for k in g['disc_kind']]
    ref_chi = [w for k, w in zip(kinds, g['disc_weight']) if k == 'chi']
    # the chi weights of this layer are 1e-5 .. 1e-8 of Tr: complex64 site data resolves a weight w
    # to about 2 sqrt(w) x 6e-8 per discarded value, so below ~1e-4 Tr the bar is 1e-8 Tr absolute
    assert np.allclose([disc[0, i] for i in reversed(range(n - 1))], ref_chi, rtol=1e-4, atol=1e-8 * tr)
    ref_kap = sum(w for k, w in zip(kinds, g['disc_weight']) if k == 'kappa')
    assert abs(disc[0, n:2 * n].sum() - ref_kap) <= 1e-4 * ref_kap + 1e-8 * tr
    ref_gate = {int(i): w for k, i, w in zip(kinds, g['disc_index'], g['disc_weight']) if k == 'gate'}
    for p, w in ref_gate.items():
        assert abs(disc[0, 2 * n + p] - w) <= 1e-3 * abs(w) + 1e-12, (p, disc[0, 2 * n + p], w)
    # both chains against the per-gate entry points
    for b in range(B):
        res, gate, d_chi, d_kap = batched.per_gate_layer([chains[b][i] for i in range(n)], L, b, 256, 8)
        assert [tuple(t.shape) for t in res] == [tuple(o.shape[1:]) for o in outs]
        vb = np.array(_lib.pauli_expect([o[b].contiguous() for o in outs], codes))
        vg = np.array(_lib.pauli_expect(res, codes))
        assert np.abs(vb - vg).max() <= 1e-5 * abs(vg[-1]), b
        _check_layer_weights(disc[b], n, L.parity, gate, d_chi, d_kap)


# ------------------------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_full_size_n32_chi256_kappa8(cuda):
    """The benchmark shape: 32 copies of the 32-qubit circuit family at chi=256, kappa=8, two
    layers (both pair parities) from the synthetic steady state the benchmark starts from.
    Copy 0 against the per-gate path on the same input (all 496 <Z_i Z_j>, 32 <Z_i>, Tr rho~,
    shapes, discarded weights of every cut); copies 0 and 17 run alone (batch of one) against
    their results in the batch of 32. The fused layers compute their site contractions as the
    benchmark does (bf16 matrix cores, three-part split operands: math mode 4), the per-gate
    reference with fp64 accumulation (mode 0)."""
    from mpdo_amd import _lib
    lib = _lib.load()
    old = lib.mpdo_set_math_mode(0)
    try:
        _full_size_check(cuda, lib)
    finally:
        lib.mpdo_set_math_mode(old)


def _full_size_check(cuda, lib):
    from mpdo_amd import _lib, batched
    from mpdo_amd import workloads as W
    n, B, chi, kappa, first = 32, 32, 256, 8, 10
    ws = W.brickwork32_chi256_family(copies=B, n=n)
    cs = _circuits(ws, cuda)
    segs = batched._segments(cs[0])
    layers = []
    for k in range(2):
        L, why = batched.fused_layer_of(cs, *segs[first + k])
        assert L is not None, why
        layers.append(L)
    dims = batched.steady_dims(n, chi, kappa, layers[0].kraus1.shape[2], layers[0].gates2.shape[-1],
                               layers[0].parity)
    sites = W.steady_chains_device(dims, B, cuda)
    lone = {b: [t[b:b + 1].clone() for t in sites] for b in (0, 17)}
    pg = [t[0].clone() for t in sites]
    disc_all = []
    lib.mpdo_set_math_mode(4)
    for L in layers:
        sites, disc = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, chi, kappa)
        disc_all.append(disc)
        print(f'fused layer parity {L.parity}: kept {_lib.layer_kept()}', flush=True)
    lib.mpdo_set_math_mode(0)
    codes, _ = _zz_codes(n)
    v_f = np.array(_lib.pauli_expect([t[0].contiguous() for t in sites], codes))
    # per-gate path, copy 0
    for k, L in enumerate(layers):
        pg, gate, d_chi, d_kap = batched.per_gate_layer(pg, L, 0, chi, kappa)
        _check_layer_weights(disc_all[k][0], n, L.parity, gate, d_chi, d_kap)
        print(f'per-gate layer {k} done', flush=True)
    assert [tuple(t.shape) for t in pg] == [tuple(t.shape[1:]) for t in sites]
    v_g = np.array(_lib.pauli_expect(pg, codes))
    tr = abs(v_g[-1])
    assert np.abs(v_f - v_g).max() <= 1e-5 * tr, np.abs(v_f - v_g).max() / tr
    # batch invariance
    lib.mpdo_set_math_mode(4)
    for b, ch in lone.items():
        for L in layers:
            k1 = L.kraus1[b:b + 1]
            ch, _ = _lib.layer_batched(ch, k1, L.parity, L.gates2, chi, kappa)
        v_a = np.array(_lib.pauli_expect([t[0].contiguous() for t in ch], codes))
        v_b = np.array(_lib.pauli_expect([t[b].contiguous() for t in sites], codes))
        assert np.abs(v_a - v_b).max() <= 1e-6 * abs(v_b[-1]), (b, np.abs(v_a - v_b).max())


# ------------------------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_brickwork20_damping_chi64_golden(cuda):
    """20 qubits, depth 20, CNOT/CZ brickwork with amplitude damping 0.005 + dephasing 0.005
    after every gate, chi=64, kappa=4 (the reference's composite 1q noise, noisy identities after
    the two-qubit gates), evolved by TensorCircuit.evolve against the oracle fixture."""
    import mpdo_amd as M
    from mpdo_amd import _lib
    from mpdo_amd import workloads as W
    g = np.load(os.path.join(HERE, 'golden', 'brickwork20_damping_chi64_oracle.npz'))
    w = W.brickwork20_damping_chi64()
    c = _circuits([w], cuda)[0]
    state = M.Tools.create_ket0Series(w.n, device=str(cuda))
    c.evolve(state)
    sites = [nd.tensor.contiguous() for nd in state]
    assert [tuple(t.shape) for t in sites] == [tuple(d) for d in g['dims']]
    n = w.n
    codes, _ = _zz_codes(n, [(i, i + 1) for i in range(n - 1)])
    vals = np.array(_lib.pauli_expect(sites, codes))
    tr = float(g['trace'])
    assert abs(vals[-1] - tr) <= 1e-5 * tr
    assert np.abs(vals[:n] - g['z']).max() <= 1e-5 * tr
    assert np.abs(vals[n:-1] - g['zz']).max() <= 1e-5 * tr
    bits = g['bits']
    pcodes = [[3 + int(x) for x in row] for row in bits]
    probs = np.array(_lib.pauli_expect(sites, pcodes))
    assert np.abs(probs - g['probs']).max() <= 1e-5 * tr
    kinds = [str(k) for k in g['disc_kind']]
    ref = [(k, int(i), float(v)) for k, i, v in zip(kinds, g['disc_index'], g['disc_weight']) if k == 'chi']
    got = [d for d in c.discarded if d[0] == 'chi']
    assert [d[:2] for d in got] == [d[:2] for d in ref]
    assert np.allclose([d[2] for d in got], [d[2] for d in ref], rtol=1e-4, atol=1e-10)
    kf = sum(d[2] for d in c.discarded if d[0] == 'kappa')
    ko = sum(v for k, v in zip(kinds, g['disc_weight']) if k == 'kappa')
    assert abs(kf - ko) <= 1e-4 * ko + 1e-10


# ------------------------------------------------------------------------------------------
def test_gate_cut_keep_all_test_matches_eigensolver(cuda):
    """The fused layer decides "the gate-time cut keeps every direction" by a pivoted Cholesky of
    G - D* I instead of the eigensolver: on chains whose gate Grams are well conditioned (flat bond
    weights) the decision is taken for some pairs, and the layer's outputs and discarded weights
    are those of the same layer with the eigensolver forced on every pair, to the bit; on the
    fixture chain (decaying weights: cuts that drop directions) both runs agree as well."""
    from mpdo_amd import _lib
    from mpdo_amd import workloads as W
    lib = _lib.load()
    n, B = 10, 2
    _, L = _layer10_inputs(cuda, B)
    import importlib.util
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(HERE, 'golden', 'make_golden.py'))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for decay in (1.0, 0.97):
        chains = [[torch.from_numpy(t).to(cuda) for t in W.steady_state(mg.LAYER10_DIMS, seed=11 + b, decay=decay)]
                  for b in range(B)]
        sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
        old = lib.mpdo_debug_layer_gate_skip(1)
        old_tail = lib.mpdo_debug_layer_gate_tail(0)   # the other cuts by the eigensolver in both runs
        old_top = lib.mpdo_debug_layer_chi_top(0)
        try:
            outs_a, disc_a = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            skips = lib.mpdo_debug_layer_gate_skips()
            kept_a = _lib.layer_kept()
            lib.mpdo_debug_layer_gate_skip(0)
            outs_b, disc_b = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            assert lib.mpdo_debug_layer_gate_skips() == 0
            kept_b = _lib.layer_kept()
        finally:
            lib.mpdo_debug_layer_gate_skip(old)
            lib.mpdo_debug_layer_gate_tail(old_tail)
            lib.mpdo_debug_layer_chi_top(old_top)
        print(f'decay {decay}: {skips} of {len(kept_a)} gate cuts decided without the eigensolver; kept {kept_b}')
        if decay == 1.0:
            assert skips > 0
        assert kept_a == kept_b
        for a, b in zip(outs_a, outs_b):
            assert torch.equal(a, b)
        assert np.array_equal(disc_a, disc_b)


def test_gate_cut_from_the_bottom_matches_eigensolver(cuda):
    """Where the shifted Cholesky does not settle the gate-time cut, the fused layer decides it from
    the bottom of the spectrum (subspace iteration, Rayleigh-Ritz, certificate) and projects with
    P = I - V_d V_d^H: on decaying-weight chains (cuts that drop a few directions) the kept counts are
    those of the eigensolver, the dropped weights agree to 1e-6 relative, and every <Z_i>, <Z_i Z_j>
    and Tr rho~ to 1e-6 of Tr rho~ (the expectations are read from complex64 data, resolved to
    ~1.2e-7), chi and kappa weights to 1e-6 relative (1e-12 Tr absolute)."""
    from mpdo_amd import _lib
    from mpdo_amd import workloads as W
    lib = _lib.load()
    n, B = 10, 2
    _, L = _layer10_inputs(cuda, B)
    import importlib.util
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(HERE, 'golden', 'make_golden.py'))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    codes, pairs = _zz_codes(n)
    tails_seen = 0
    for decay in (0.995, 0.99, 0.985):
        chains = [[torch.from_numpy(t).to(cuda) for t in W.steady_state(mg.LAYER10_DIMS, seed=21 + b, decay=decay)]
                  for b in range(B)]
        sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
        old = lib.mpdo_debug_layer_gate_tail(1)
        old_top = lib.mpdo_debug_layer_chi_top(0)   # the chi steps by the eigensolver in both runs
        try:
            outs_a, disc_a = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            tails = lib.mpdo_debug_layer_gate_tails()
            kept_a = _lib.layer_kept()
            lib.mpdo_debug_layer_gate_tail(0)
            outs_b, disc_b = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            assert lib.mpdo_debug_layer_gate_tails() == 0
            kept_b = _lib.layer_kept()
        finally:
            lib.mpdo_debug_layer_gate_tail(old)
            lib.mpdo_debug_layer_chi_top(old_top)
        print(f'decay {decay}: {tails} of {len(kept_a)} gate cuts decided from the bottom; kept {kept_a} / {kept_b}')
        tails_seen += tails
        assert kept_a == kept_b
        assert [tuple(a.shape) for a in outs_a] == [tuple(b.shape) for b in outs_b]
        for b in range(B):
            va = np.array(_lib.pauli_expect([o[b].contiguous() for o in outs_a], codes))
            vb = np.array(_lib.pauli_expect([o[b].contiguous() for o in outs_b], codes))
            assert np.abs(va - vb).max() <= 1e-6 * abs(vb[-1]), (decay, b, np.abs(va - vb).max() / abs(vb[-1]))
            tr = abs(vb[-1])
            gate_a, gate_b = disc_a[b, 2 * n:], disc_b[b, 2 * n:]
            assert np.allclose(gate_a, gate_b, rtol=1e-6, atol=1e-15 * tr), (decay, b, gate_a, gate_b)
            assert np.allclose(disc_a[b, :2 * n], disc_b[b, :2 * n], rtol=1e-6, atol=1e-12 * tr), (decay, b)
    assert tails_seen > 0


def test_chi_step_by_top_eigenpairs_matches_eigensolver(cuda):
    """The chi step at a pair bond by its top chi eigenpairs (subspace iteration on 2 chi columns,
    Rayleigh-Ritz, certificate) against the same layer with the full eigensolver: same shapes and
    kept counts, every <Z_i>, <Z_i Z_j> and Tr rho~ within 1e-6 of Tr rho~, discarded weights within
    1e-6 relative (1e-12 Tr absolute), on the fixture chain and a flatter one."""
    from mpdo_amd import _lib
    from mpdo_amd import workloads as W
    lib = _lib.load()
    n, B = 10, 2
    chains0, L = _layer10_inputs(cuda, B)
    import importlib.util
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(HERE, 'golden', 'make_golden.py'))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    codes, _ = _zz_codes(n)
    tops_seen = 0
    for name, chains in (('fixture', chains0),
                         ('decay 0.99', [[torch.from_numpy(t).to(cuda)
                                          for t in W.steady_state(mg.LAYER10_DIMS, seed=31 + b, decay=0.99)]
                                         for b in range(B)])):
        sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
        old = lib.mpdo_debug_layer_chi_top(1)
        try:
            outs_a, disc_a = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            tops = lib.mpdo_debug_layer_chi_tops()
            kept_a = _lib.layer_kept()
            lib.mpdo_debug_layer_chi_top(0)
            outs_b, disc_b = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
            assert lib.mpdo_debug_layer_chi_tops() == 0
            kept_b = _lib.layer_kept()
        finally:
            lib.mpdo_debug_layer_chi_top(old)
        print(f'{name}: {tops} pair-bond chi steps by the top solver; kept {kept_a}')
        tops_seen += tops
        assert kept_a == kept_b
        assert [tuple(a.shape) for a in outs_a] == [tuple(b.shape) for b in outs_b]
        for b in range(B):
            va = np.array(_lib.pauli_expect([o[b].contiguous() for o in outs_a], codes))
            vb = np.array(_lib.pauli_expect([o[b].contiguous() for o in outs_b], codes))
            tr = abs(vb[-1])
            assert np.abs(va - vb).max() <= 1e-6 * tr, (name, b, np.abs(va - vb).max() / tr)
            assert np.allclose(disc_a[b], disc_b[b], rtol=1e-6, atol=1e-12 * tr), (name, b)
    assert tops_seen > 0
