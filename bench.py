#!/usr/bin/env python3
"""Throughput of the MPDO noisy-circuit update path on one MI355X (gfx950).

Workload (the 32-qubit precision configuration): depth-30 brickwork of RY(theta) RZ(phi) on every
qubit then CNOT on (0,1),(2,3),... / (1,2),(3,4),... alternately, single- and two-qubit
depolarizing noise p = 0.002, chi = 256, kappa = 8, 32 copies whose angles differ by
U[-0.01, 0.01] (mpdo_amd.workloads.brickwork32_chi256_family). One step = one circuit-layer of
all 32 copies: the composed single-qubit channel of every site, CNOT + two-qubit depolarizing on
every pair of the layer's parity and the truncation (QR sweep, chi sweep, kappa cut), as one call
of mpdo_layer_batched. The chains start from synthetic complex64 site tensors (seeded, decaying bond
and inner weights) at the shapes a long run of this circuit settles into; successive layers
alternate the parity and feed each other. The warmup layers bring the spectra to the circuit's
own steady state as well: the gate-time split keeps ~458 of its 1024 directions on the synthetic
start, then 780, 950, and from the sixth layer on 1017-1024 (the split no longer truncates and
every later eigenproblem has order ~1024, twice the cost of the first layer), so the default
times two layers after six warmup layers. Inputs are resident in HBM when the timed region
starts.

Prints one JSON line: circuit-layers/s (32 x steps / seconds), the roofline of the dominant
kernel class (algorithmic flops per launch / launch duration, HIP events on the launch stream),
and the CPU oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'tomography-assisted-mpdo-qcircuit_amd')
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = 'circuit-layers/sec at χ=256,κ=8 (batched vs plain) @ ≤1e-5 DM error'
PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3}     # MI355X dense peaks
N_QUBITS, CHI, KAPPA, P_NOISE, FIRST_LAYER = 32, 256, 8, 0.002, 10
# HBM bytes per GEMM launch of the timed step, from the two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) of this benchmark, scaled by the copy calibration; the file records the commands
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'gemm_kernel_traffic_fused_layer_n32_chi256_kappa8_b32.md')
CALIB_BYTES = 1 << 30


def _traffic_per_launch():
    """The measured HBM bytes per GEMM launch recorded in TRAFFIC_FILE (line
    `traffic_bytes_per_launch = N`), or None when that measurement is absent."""
    import re
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r'traffic_bytes_per_launch = ([0-9.e+]+)', f.read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def _steady_chains(dims, copies, dev, seed=2024):
    """(copies, l, 2, a, r) complex64 per site: random tensors with geometric weights 0.97^l,
    0.97^r on the bonds and 0.5^a on the inner index, each site scaled so that the running
    transfer-matrix trace of its chain stays 1 (Tr rho = 1, as for a state evolved from |0..0>
    by trace-preserving channels: the gate-time split's absolute error bound then acts at the
    scale it has in a real run)."""
    import torch
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    out = []
    env = torch.ones(copies, 1, 1, dtype=torch.complex128, device=dev)
    for (l, a, r) in dims:
        T = torch.complex(torch.randn(copies, l, 2, a, r, generator=g, device=dev),
                          torch.randn(copies, l, 2, a, r, generator=g, device=dev))
        wl = 0.97 ** torch.arange(l, device=dev, dtype=torch.float32)
        wr = 0.97 ** torch.arange(r, device=dev, dtype=torch.float32)
        wa = 0.5 ** torch.arange(a, device=dev, dtype=torch.float32)
        T = T * (wl[:, None, None, None] * wa[None, None, :, None] * wr[None, None, None, :])
        T64 = T.to(torch.complex128)
        # env'[r, r'] = sum env[l, l'] T[l, s, a, r] conj(T[l', s, a, r'])
        nxt = torch.einsum('bxy,bxsar,bysaz->brz', env, T64, T64.conj())
        f = torch.einsum('brr->b', nxt).real
        T = (T64 / torch.sqrt(f).reshape(copies, 1, 1, 1, 1)).to(torch.complex64)
        env = nxt / f.reshape(copies, 1, 1)
        out.append(T.contiguous())
    return out


def _layers(copies, count, dev):
    """Fused-layer operands of layers FIRST_LAYER .. FIRST_LAYER + count - 1 of the workload."""
    import mpdo_amd as M
    from mpdo_amd import batched
    from mpdo_amd import workloads as W
    ws = W.brickwork32_chi256_family(copies=copies, n=N_QUBITS)
    cs = []
    for w in ws:
        c = M.TensorCircuit(device=str(dev), **w.circuit_kwargs())
        W.apply_ops(c, w.ops)
        cs.append(c)
    segs = batched._segments(cs[0])
    out = []
    for k in range(count):
        start, stop = segs[(FIRST_LAYER + k) % len(segs)]
        layer, why = batched.fused_layer_of(cs, start, stop)
        if layer is None:
            raise RuntimeError(f'layer {FIRST_LAYER + k} does not fuse: {why}')
        out.append(layer)
    return out


def cpu_baseline(Ti, Tj, i, threads: int = 16):
    """The oracle (numpy complex128) on one pair of the benchmark's first circuit-layer (sites i,
    i+1 of copy 0, the same tensors the GPU step starts from): the composed single-qubit channels
    of both sites and the gate-time split of the pair (Circuit.py:74-178); a layer has 16 such
    pairs plus the QR, chi and kappa steps, so 16 x this time is a lower bound on the oracle's
    time per circuit-layer."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import mpdo_oracle as O

    Ti, Tj = Ti.astype(np.complex128), Tj.astype(np.complex128)
    noise = O.NoiseSpec(chip='best', dpc_error_rate=P_NOISE, single='depolarizing', p1=P_NOISE)
    E1 = noise.single_kraus(O.gate_matrix('RY', [1.1]))
    E2 = noise.single_kraus(O.gate_matrix('RZ', [0.3]))
    G2 = noise.double_kraus(O.gate_matrix('CNOT'))
    with threadpool_limits(threads):
        t0 = time.perf_counter()
        a = O.apply_1q(O.apply_1q(Ti, E1), E2)
        b = O.apply_1q(O.apply_1q(Tj, E1), E2)
        O.apply_2q(a, b, G2)
        dt = time.perf_counter() - t0
    pairs = len(range(0, N_QUBITS - 1, 2))
    return {'value': 1.0 / (pairs * dt), 'unit': 'circuit-layers/s', 'cores': threads, 'kind': 'port',
            'sample': (f'oracle (numpy complex128, {threads} BLAS threads) on one middle pair of the first timed '
                       f'circuit-layer (n=32, chi=256, kappa=8, copy 0, sites {i},{i + 1} of shapes '
                       f'{tuple(Ti.shape)}, {tuple(Tj.shape)}): both sites\' RY, RZ + depolarizing channels and '
                       f'the CNOT + depolarizing gate-time split took {dt:.1f} s; value = 1 / ({pairs} x that), an '
                       f'upper bound on the oracle\'s rate since the QR, chi and kappa steps are not timed')}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=6)
    ap.add_argument('--copies', type=int, default=32)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-threads', type=int, default=16)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='before the warmup, run torch.neg over 1 GiB of float32 three times (1 GiB read and '
                         '1 GiB written per launch): the reference launch that scales the HBM counters of a '
                         'rocprofv3 --pmc run of this benchmark')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). '
              'Chains are independent, so more GPUs would each take a share of the copies; run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)

    import torch

    from mpdo_amd import _lib, batched

    if not torch.cuda.is_available():
        raise SystemExit('bench.py needs the MI355X (cuda:0)')
    dev = torch.device('cuda:0')
    if args.copy_calibration:
        src = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        print(f'# copy calibration: 3 launches of torch.neg, {CALIB_BYTES} bytes read and {CALIB_BYTES} '
              'bytes written each', file=sys.stderr, flush=True)
    lib = _lib.load()
    prev_mode = lib.mpdo_set_math_mode(0)
    B = args.copies
    layers = _layers(B, args.warmup + args.steps, dev)
    dims = batched.steady_dims(N_QUBITS, CHI, KAPPA, layers[0].kraus1.shape[2], layers[0].gates2.shape[-1],
                               layers[0].parity)
    sites = _steady_chains(dims, B, dev)
    torch.cuda.synchronize()
    # the CPU sample's pair: sites i, i+1 of copy 0 at the input of the first timed layer
    pair_i = N_QUBITS // 2 - 2 + layers[args.warmup].parity
    print(f'# {B} copies, {N_QUBITS} qubits, chi={CHI}, kappa={KAPPA}; warmup {args.warmup}, steps {args.steps}',
          file=sys.stderr, flush=True)

    def step(k, sites):
        L = layers[k]
        outs, _ = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, CHI, KAPPA)
        return outs

    for k in range(args.warmup):
        t = time.perf_counter()
        sites = step(k, sites)
        torch.cuda.synchronize()
        print(f'# warmup step {k}: {time.perf_counter() - t:.2f} s', file=sys.stderr, flush=True)

    cpu_pair = None
    if not args.no_cpu_baseline:
        cpu_pair = (sites[pair_i][0].cpu().numpy(), sites[pair_i + 1][0].cpu().numpy())
    _lib.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        sites = step(args.warmup + k, sites)
        print(f'# timed step {k} issued', file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    prof = _lib.prof_read()
    _lib.prof_enable(False)
    print(f'# gate-time split kept counts of the last layer: {_lib.layer_kept()}', file=sys.stderr, flush=True)
    eigs = {}
    for (n_e, b_e, sw_e) in _lib.prof_eig_log():
        c = eigs.setdefault(n_e, [0, 0])
        c[0] += 1
        c[1] += sw_e
    print('# blocked eigensolver calls (order: calls, mean sweeps): ' +
          ', '.join(f'{k}: {v[0]}, {v[1] / v[0]:.1f}' for k, v in sorted(eigs.items())), file=sys.stderr, flush=True)
    shapes = sorted(_lib.prof_shapes('gemm'), key=lambda r: -r[2])
    print('# GEMM shapes by time (M, N, K, batch, prec: launches, ms, TFLOP/s): ' +
          '; '.join(f'{sh}: {c}, {t:.0f}, {f / (t * 1e-3) / 1e12:.1f}' for sh, c, t, f in shapes[:14]),
          file=sys.stderr, flush=True)

    # the dominant kernel: mpdo::gemm_kernel, the kernel with the largest share of the steady-state
    # step (30% of kernel time over its operand/tile instantiations; the block-Jacobi rounds are
    # three kernels, bj_update 24%, bj_solve 20.5%, bj_rotate_v_direct 18%, whose per-round tile
    # counts include the identity rotations they skip, so their flop totals are upper bounds and
    # are not used for a roofline)
    name = 'gemm'
    n_l, ms, fl, by = prof[name]
    prec = 'fp64'
    achieved = fl / (ms * 1e-3) / 1e12
    kernels = {'gemm': 'mpdo::gemm_kernel (complex GEMM, fp64 MFMA accumulate)'}
    traffic = _traffic_per_launch() if name == 'gemm' else None
    roofline = {'bound': 'mfma', 'achieved': round(achieved, 3), 'peak': PEAK_TFLOPS[prec], 'unit': 'TFLOP/s',
                'frac': round(achieved / PEAK_TFLOPS[prec], 4), 'traffic': traffic,
                'kernel': kernels.get(name, name), 'launches': n_l, 'avg_launch_ms': round(ms / max(n_l, 1), 4),
                'flops_per_launch': fl / max(n_l, 1), 'operand_bytes_per_launch': by / max(n_l, 1),
                'share_of_kernel_time': round(ms / max(sum(v[1] for v in prof.values()), 1e-9), 3),
                'classes_ms': {k: round(v[1], 1) for k, v in prof.items()}}
    value = B * args.steps / T
    cpu = None
    if not args.no_cpu_baseline:
        print('# cpu baseline (oracle sample)', file=sys.stderr, flush=True)
        cpu = cpu_baseline(cpu_pair[0], cpu_pair[1], pair_i, args.cpu_threads)
    lib.mpdo_set_math_mode(prev_mode)
    line = {'metric': METRIC, 'value': round(value, 6), 'unit': 'circuit-layers/s', 'n_gpus': 1,
            'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(T / args.steps * 1e3, 3),
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'c64 data, fp64 accumulate/eigensolve', 'data': 'synthetic',
            'config': {'workload': ('32-qubit depth-30 RY/RZ+CNOT brickwork, depolarizing p=0.002 (1q, 2q), '
                                    'chi=256, kappa=8, 32 angle-perturbed copies; step = one circuit-layer of '
                                    'every copy from the steady state (fused batched layer)'),
                       'n_qubits': N_QUBITS, 'chi': CHI, 'kappa': KAPPA, 'copies': B, 'global_batch': B,
                       'parallelism': 'none (one GPU)'},
            'roofline': roofline, 'cpu_baseline': cpu}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
