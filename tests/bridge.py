"""Replays a synthetic workload through the oracle (tests only)."""
import numpy as np

from oracle import mpdo_oracle as O


def oracle_noise(w):
    chip = w.chip or {}
    return O.NoiseSpec(chip='best' if chip else 'worst',
                       decay_param=chip.get('decay_param', 0.0 if chip else None),
                       dephasing_param=chip.get('dephasing_param', 0.0 if chip else None),
                       dpc_error_rate=chip.get('dpc_error_rate', 0.0 if chip else None),
                       single=w.single_qubit_noise, p1=w.p1 if w.p1 is not None else 0.0)


def oracle_ops(w, exp_tensors=None):
    b = O.CircuitBuilder(w.n, w.noise_type, exp_tensors)
    for name, q, p in w.ops:
        if name == 'TRUNCATE':
            b.truncate()
        elif name == 'BARRIER':
            b.barrier()
        else:
            b.add(name, q, p)
    return b.ops


def run_oracle(w, chi='default', kappa='default', err=None):
    st = O.product_state(w.n)
    O.evolve(oracle_ops(w), st, w.noise_type, w.ideal, oracle_noise(w),
             chi=w.chi if chi == 'default' else chi, kappa=w.kappa if kappa == 'default' else kappa,
             max_truncation_err=err)
    return st


def run_oracle_fused(w, st=None, layers=None):
    """oracle.evolve's rules (Circuit.py:74-224, 469-491) for a workload of gates and TRUNCATE
    markers, with each qubit's consecutive single-qubit channels composed into one
    Choi-rank-reduced Kraus set (oracle.compose_compressed) applied before the next two-qubit
    gate on that qubit or the next truncation: the sequential sets would multiply the inner
    dimension by up to 36 per layer, and by thousands before the chain is first connected
    (no truncation until then). The composed set is related to the sequential ones by an
    isometry on the inner index, which no kept count, discarded weight or gauge-invariant
    output sees. ``st``: start state (default |0...0>); ``layers``: only the first that many
    circuit-layers. The gate-time split's dropped weight is recorded as ('gate', min, weight)."""
    from mpdo_amd import workloads as W
    noise = oracle_noise(w)
    nt = w.noise_type.lower()
    unified, idealn = (not w.ideal and nt == 'unified'), (not w.ideal and nt == 'idealnoise')
    st = O.product_state(w.n) if st is None else st
    ops = w.ops if layers is None else [op for k in range(layers) for op in W.layer_ops(w.ops, k)]
    pending = {q: [] for q in range(w.n)}

    def flush(qs):
        for q in qs:
            if pending[q]:
                st.sites[q] = O.apply_1q(st.sites[q], O.compose_compressed([E for E, _ in pending[q]]))
                if any(g for _, g in pending[q]):
                    st.has_I[q] = True
                pending[q] = []

    for name, q, p in ops:
        if name == 'TRUNCATE':
            flush(range(w.n))
            if O.check_connectivity(st):
                O.bond_truncate(st, w.chi)
                if not w.ideal:
                    O.svd_kappa_left2right(st, w.kappa)
            continue
        G = O.gate_matrix(name, p)
        if G.ndim == 2:
            gnoise = idealn or unified
            pending[q[0]].append((noise.single_kraus(G) if gnoise else G, gnoise))
            continue
        a, b = sorted(q)
        flush([a, b])
        if idealn:
            G = noise.double_kraus(G)
        if q[0] > q[1]:
            G = np.swapaxes(np.swapaxes(G, 0, 1), 2, 3)
        x, y, s_rest = O.apply_2q(st.sites[a], st.sites[b], G, True)
        st.sites[a], st.sites[b] = x, y
        st.bond[a] = True
        if G.ndim == 5:
            st.has_I[b] = True
        st.discarded.append(('gate', a, float(np.sum(np.abs(s_rest) ** 2))))
    flush(range(w.n))
    if not w.ideal and ops and ops[-1][0] != 'TRUNCATE':
        O.svd_kappa_left2right(st, w.kappa)
    return st


def run_hea_oracle(theta, phi, p1, p2, chi, kappa, h=1.0):
    """The hardware-efficient ansatz of mpdo_amd.autodiff.HEA on the oracle (complex128): per layer
    RY(theta) RZ(phi) on every qubit, each followed by depolarizing of its own strength p1[d, q, 0/1]
    (composed into one Kraus set by oracle.compose_compressed), CNOT(q, q+1) for q = 0..n-2 followed by
    two-qubit depolarizing p2[d, q] (oracle.apply_2q), then bond_truncate(chi) and the kappa cut.
    Returns (<H> unnormalised, H = -sum Z_i Z_i+1 - h sum X_i; the final state)."""
    theta, phi, p1, p2 = (np.asarray(x, dtype=np.float64) for x in (theta, phi, p1, p2))
    depth, n = theta.shape
    st = O.product_state(n)
    for d in range(depth):
        for q in range(n):
            e1 = np.einsum('nlm,li->nim', O.depolarizing_tensor(float(p1[d, q, 0]), 1), O.gate_matrix('RY', [theta[d, q]]))
            e2 = np.einsum('nlm,li->nim', O.depolarizing_tensor(float(p1[d, q, 1]), 1), O.gate_matrix('RZ', [phi[d, q]]))
            st.sites[q] = O.apply_1q(st.sites[q], O.compose_compressed([e1, e2]))
            st.has_I[q] = True
        for q in range(n - 1):
            G = np.einsum('ijklp,klmn->ijmnp', O.depolarizing_tensor(float(p2[d, q]), 2), O.gate_matrix('CNOT'))
            a, b, _ = O.apply_2q(st.sites[q], st.sites[q + 1], G, True)
            st.sites[q], st.sites[q + 1] = a, b
            st.bond[q] = True
            st.has_I[q + 1] = True
        if O.check_connectivity(st):
            O.bond_truncate(st, chi)
            O.svd_kappa_left2right(st, kappa)
    e = -sum(O.pauli_expect(st, [2, 2], [i, i + 1]) for i in range(n - 1))
    e -= h * sum(O.pauli_expect(st, [0], [i]) for i in range(n))
    return e, st
