"""CPU-side checks of libmpdo_gfx950.so: it loads, exports every function include/*.h declares,
and the fused layer's host plan is consistent (no GPU needed: the plan and the operand-extent
checks of every launch run on the host)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ('mpdo.h', 'mpdo_debug.h'):
        txt = open(os.path.join(ROOT, 'include', h)).read()
        txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
        names += re.findall(r'^\s*(?:const\s+)?[a-z_]+\**\s+\**(mpdo_[a-z0-9_]+)\s*\(', txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from mpdo_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 18, names
    for nm in names:
        assert hasattr(lib, nm), nm


def test_version_and_error_text():
    from mpdo_amd import _lib
    lib = _lib.load()
    assert lib.mpdo_version() >= 1
    assert isinstance(lib.mpdo_last_error(), bytes)


@pytest.mark.parametrize('n,chi,kappa,batch', [(32, 256, 8, 32), (32, 256, 8, 1), (12, 64, 4, 4), (9, 4, 2, 2),
                                               (8, 32, 4, 2), (7, 16, 4, 2), (6, 8, 2, 3)])
@pytest.mark.parametrize('parity', [0, 1])
def test_fused_layer_dry_run_steady_state(n, chi, kappa, batch, parity):
    """Every launch of the fused layer at the steady-state shapes passes its extent checks."""
    from mpdo_amd import _lib, batched
    dims = batched.steady_dims(n, chi, kappa, 4, 16, parity)
    ws = _lib.layer_dry_run(batch, dims, 4, parity, 16, chi, kappa)
    assert ws > 0
    out, _ = _lib.layer_dims(dims, 4, parity, 16, chi, kappa)
    nxt = batched.steady_dims(n, chi, kappa, 4, 16, 1 - parity)
    assert out == nxt      # the two parities' shapes form the 2-cycle


def test_fused_layer_dry_run_growth_phase():
    """Shapes on the way to the steady state (bonds growing by 2 kappa K1 per layer)."""
    from mpdo_amd import _lib
    n, chi, kappa = 10, 64, 4
    dims = [(1, 4, 8)] + [(8, 4, 8)] * (n - 2) + [(8, 4, 1)]
    for layer in range(6):
        parity = layer % 2
        out, why = _lib.layer_dims(dims, 4, parity, 16, chi, kappa)
        assert out is not None, why
        _lib.layer_dry_run(3, dims, 4, parity, 16, chi, kappa)
        dims = out


def test_fused_layer_rejects_thin_pairs():
    """Without Kraus axes (ideal gates) the max site cannot carry its bond: the plan refuses and
    the caller uses the per-gate path."""
    from mpdo_amd import _lib
    dims = [(1, 1, 2), (2, 1, 4), (4, 1, 4), (4, 1, 2), (2, 1, 1)]
    out, why = _lib.layer_dims(dims, 1, 1, 1, 4, 0)
    assert out is None and why


def test_layer_workspace_query_matches_dry_run():
    """mpdo_layer_workspace_bytes (the size a caller binds for mpdo_layer_batched) is the dry run's
    workspace at the benchmark shape and at a small one; 0 with the reason for an unplannable layer."""
    import ctypes
    from mpdo_amd import _lib, batched
    lib = _lib.load()
    for n, chi, kappa, B in [(32, 256, 8, 32), (8, 16, 2, 3)]:
        dims = batched.steady_dims(n, chi, kappa, 4, 16, 0)
        ins = (_lib.Site * n)()
        for i, (l, a, r) in enumerate(dims):
            ins[i] = _lib.Site(None, l, a, r)
        got = lib.mpdo_layer_workspace_bytes(B, n, ins, 4, 0, 16, chi, kappa)
        assert got == _lib.layer_dry_run(B, dims, 4, 0, 16, chi, kappa) > 0
    bad = (_lib.Site * 2)(_lib.Site(None, 1, 1, 2), _lib.Site(None, 3, 1, 1))
    assert lib.mpdo_layer_workspace_bytes(1, 2, bad, 4, 0, 16, 4, 2) == 0
    assert b'mismatch' in lib.mpdo_last_error()


def test_gemm_tile_table_round_trip():
    """The tile autotuner's table is host state: set, get and the JSON file round-trip without a
    GPU; out-of-range tiles are refused."""
    from mpdo_amd import _lib, autotune
    saved = autotune.table()
    try:
        entries = [{'key': [256, 16384, 256, 32, 2, 0, 0, 1, 1, 1, 1, 0], 'tile': 2},
                   {'key': [512, 512, 1024, 32, 2, 1, 0, 1, 1, 1, 0, 1], 'tile': 4},
                   {'key': [1, 2, 3, 4, 2, 0, 0, 0, 0, 1, 1, 1], 'tile': 9}]
        assert autotune.set_table(entries) == 2
        got = sorted(autotune.table(), key=lambda e: e['key'])
        assert got == sorted(entries[:2], key=lambda e: e['key'])
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, 't.json')
            autotune.save(path, got)
            autotune.set_table([])
            assert autotune.table() == []
            assert autotune.load(path) == 2
            assert sorted(autotune.table(), key=lambda e: e['key']) == got
    finally:
        autotune.set_table(saved)


def test_sub_batch_split(monkeypatch):
    """The sub-batches of layers_batched_streams / evolve_batched(streams=): contiguous, covering
    every chain once, at most one stream fewer than GPU_MAX_HW_QUEUES (4 when unset)."""
    from mpdo_amd import batched
    monkeypatch.delenv('GPU_MAX_HW_QUEUES', raising=False)
    assert batched.sub_batches(7, 4) == [(0, 3), (3, 6), (6, 7)]
    assert batched.sub_batches(32, 3) == [(0, 11), (11, 22), (22, 32)]
    assert batched.sub_batches(1, 4) == [(0, 1)]
    assert batched.sub_batches(5, 1) == [(0, 5)]
    monkeypatch.setenv('GPU_MAX_HW_QUEUES', '8')
    assert batched.sub_batches(32, 4) == [(0, 8), (8, 16), (16, 24), (24, 32)]
    assert len(batched.sub_batches(32, 16)) == 7
    monkeypatch.setenv('GPU_MAX_HW_QUEUES', '1')
    assert batched.sub_batches(32, 4) == [(0, 32)]
