"""The GEMM tile autotuner (mpdo_amd.autotune): recording one fused chi=256, kappa=8 layer at 10
qubits and replaying its GEMMs with every fitting tile shape fills the table with every key the
layer launches; the layer run with the tuned table gives the same shapes and, to 1e-6 of Tr rho~,
the same <Z_i>, <Z_i Z_j> and Tr rho~ as with the built-in tile rules (a different tile shape only
changes the order of the fp64 sums)."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_tuned_tiles_give_the_same_layer(cuda):
    from mpdo_amd import _lib, autotune
    from test_gpu_bench_config import _layer10_inputs, _zz_codes
    lib = _lib.load()
    saved = autotune.table()
    n, B = 10, 2
    chains, L = _layer10_inputs(cuda, B)
    sites = [torch.stack([chains[b][i] for b in range(B)]).contiguous() for i in range(n)]
    codes, _ = _zz_codes(n)
    try:
        autotune.set_table([])
        ref, _ = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
        v_ref = [np.array(_lib.pauli_expect([o[b].contiguous() for o in ref], codes)) for b in range(B)]
        entries = autotune.tune(lambda: _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8), reps=2)
        assert len(entries) > 0
        assert all(e['tile'] in autotune.TILES for e in entries)
        print(f'{len(entries)} keys; tiles {sorted({e["tile"] for e in entries})}')
        got, _ = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
        assert [tuple(t.shape) for t in got] == [tuple(t.shape) for t in ref]
        for b in range(B):
            v = np.array(_lib.pauli_expect([o[b].contiguous() for o in got], codes))
            assert np.abs(v - v_ref[b]).max() <= 1e-6 * abs(v_ref[b][-1])
    finally:
        autotune.set_table(saved)
