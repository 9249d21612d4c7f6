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
