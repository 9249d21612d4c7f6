"""Tile autotuner of the library's fp64-accumulating complex GEMMs.

The library keeps a table from GEMM keys (shape, precision mode, Hermitian output, triangular
operand, operand storage, reduction slices, operand layout) to tile shapes; keys absent from it use
the built-in rules. ``tune(run)`` records the GEMMs that ``run()`` launches, replays each with every
tile shape that fits it, keeps the fastest, and returns the table; ``save``/``load`` move it to and
from JSON. ``python -m mpdo_amd.autotune`` tunes the fused layer of the 32-qubit chi=256, kappa=8
brickwork family at its steady state (both pair parities) and writes ``gemm_tiles_gfx950.json``,
which the library loads at start-up.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import time
from typing import Callable, Dict, List

import torch

from . import _lib

KEY_FIELDS = ['M', 'N', 'K', 'batch', 'prec', 'herm', 'tri', 'a128', 'b128', 'ksplit', 'a_k_contig', 'b_k_contig']
TILES = {0: '64x64x16', 1: '64x64x32', 2: '128x64x16', 3: '64x128x16', 4: '128x128x16', 5: '64x64x16 two LDS stages',
         6: '32x128x16'}


def table() -> List[Dict]:
    """The library's current table: [{'key': [12 ints], 'tile': t}, ...]."""
    lib = _lib.load()
    n = int(lib.mpdo_gemm_tune_get(None, 0))
    buf = (ctypes.c_int * max(13 * n, 1))()
    n = int(lib.mpdo_gemm_tune_get(buf, n))
    return [{'key': list(buf[13 * i:13 * i + 12]), 'tile': int(buf[13 * i + 12])} for i in range(n)]


def set_table(entries: List[Dict]) -> int:
    lib = _lib.load()
    flat = (ctypes.c_int * max(13 * len(entries), 1))(*[int(v) for e in entries for v in list(e['key']) + [e['tile']]])
    return int(lib.mpdo_gemm_tune_set(flat, len(entries)))


def save(path: str, entries: List[Dict], note: str = '') -> None:
    with open(path, 'w') as f:
        json.dump({'arch': 'gfx950', 'key_fields': KEY_FIELDS, 'tiles': TILES, 'note': note, 'entries': entries}, f,
                  indent=0)


def load(path: str) -> int:
    with open(path) as f:
        return set_table(json.load(f)['entries'])


def tune(run: Callable[[], None], reps: int = 5) -> List[Dict]:
    """Record the GEMMs of run(), replay each with every fitting tile shape on the current stream and
    keep the fastest; returns the resulting table. run()'s results are overwritten by the replays."""
    lib = _lib.load()
    lib.mpdo_gemm_tune_record(1)
    try:
        run()
        torch.cuda.synchronize()
    finally:
        lib.mpdo_gemm_tune_record(0)
    n = int(lib.mpdo_gemm_tune_run(int(reps), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    if n < 0:
        raise RuntimeError('mpdo_gemm_tune_run: HIP error while timing')
    torch.cuda.synchronize()
    return table()


def _fused_layer_workload(copies: int, warmup: int, dev: torch.device, math_mode: int):
    """Steady-state chains of the 32-qubit chi=256, kappa=8 family and the fused operands of the
    layers after the warmup (both parities)."""
    import mpdo_amd as M
    from . import batched
    from . import workloads as W
    lib = _lib.load()
    lib.mpdo_set_math_mode(math_mode)
    ws = W.brickwork32_chi256_family(copies=copies, n=32)
    cs = []
    for w in ws:
        c = M.TensorCircuit(device=str(dev), **w.circuit_kwargs())
        W.apply_ops(c, w.ops)
        cs.append(c)
    segs = batched._segments(cs[0])
    layers = []
    for k in range(warmup + 2):
        start, stop = segs[(10 + k) % len(segs)]
        layer, why = batched.fused_layer_of(cs, start, stop)
        if layer is None:
            raise RuntimeError(f'layer does not fuse: {why}')
        layers.append(layer)
    dims = batched.steady_dims(32, 256, 8, layers[0].kraus1.shape[2], layers[0].gates2.shape[-1], layers[0].parity)
    sites = W.steady_chains_device(dims, copies, dev)
    for k in range(warmup):
        L = layers[k]
        sites, _ = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8)
    torch.cuda.synchronize()
    return sites, layers[warmup:]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--copies', type=int, default=32)
    ap.add_argument('--warmup', type=int, default=6)
    ap.add_argument('--reps', type=int, default=5)
    ap.add_argument('--math', type=int, default=4,
                    help='math mode of the tuned run (the benchmark default: 4, split-bf16 site contractions; the '
                         'tuned fp64 GEMMs are the same in every mode)')
    ap.add_argument('--out', default=_lib.TILE_TABLE)
    args = ap.parse_args()
    if not torch.cuda.is_available():
        raise SystemExit('the autotuner times kernels: it needs the GPU')
    dev = torch.device('cuda:0')
    lib = _lib.load()
    lib.mpdo_gemm_tune_set(None, 0)     # tune from the built-in rules
    sites, layers = _fused_layer_workload(args.copies, args.warmup, dev, args.math)
    entries = []
    for L in layers:   # one layer of each pair parity
        t = time.perf_counter()
        got = tune(lambda: _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, 256, 8), args.reps)
        print(f'parity {L.parity}: {len(got)} keys in the table after {time.perf_counter() - t:.1f} s', flush=True)
        entries = got
    changed = sum(1 for e in entries if e['tile'] not in (0, 5, 6))
    save(args.out, entries, note=f'fused layer, 32 qubits, chi=256, kappa=8, {args.copies} copies, steady state, '
                                 f'math mode {args.math}, {args.reps} timed launches per tile shape')
    print(f'wrote {args.out}: {len(entries)} keys, {changed} on a tile shape other than the 64x64 / 32x128 rules',
          flush=True)


if __name__ == '__main__':
    main()
