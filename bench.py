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
start, then 780, 950, and from the sixth layer on 1017-1024. Inputs are resident in HBM when the
timed region starts; the timed steps run without the library's launch-timing hooks.

After the timed steps, on the input of the first timed layer (kept aside):
  * one profiled pass of the same layer (HIP events on the launch stream around every launch):
    the roofline of the dominant kernel and of the block-Jacobi kernels, with the work the latter
    executed counted from their skip flags;
  * the plain per-circuit path (the per-gate entry points, one gate at a time, copy 0) in fp32
    accumulation: its circuit-layers/s, and the batched/plain ratio;
  * accuracy: copy 0 of the batched layer against the per-gate path in fp64 accumulation on the
    same input (all 496 <Z_i Z_j>, 32 <Z_i>, Tr rho~, chi/kappa discarded weights), and one
    chi=256, kappa=8 layer at 10 qubits against its complex128 oracle fixture.
Then the uneven batch (64 circuits of mixed size through the multi-stream scheduler, one stream
and eight), one forward + reverse sweep of the 16-qubit differentiable ansatz, and the CPU oracle
on a bounded sample of the same workload.

Prints one JSON line. A ~1 us spin kernel (torch.cuda._sleep) brackets the timed steps and the
profiled pass, so that `bench.py --kernel-trace-summary TRACE.csv` can cut a rocprofv3 kernel trace
of this run into those two windows.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'tomography-assisted-mpdo-qcircuit_amd')
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = 'circuit-layers/sec at χ=256,κ=8 (batched vs plain) @ ≤1e-5 DM error'
PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3, 'bf16': 2500.0}     # MI355X dense peaks
MATH_MODES = {'fp64acc': 0, 'fp32x3m': 2, 'bf16x3': 3, 'bf16x6': 4}
DEFAULT_MATH = 'bf16x6'
# GEMM groups by the precision field of a recorded shape: (group, executed flops per counted flop,
# peak, kernel). A complex product counts 8 M N K flops; 3M executes 6 of them, the split-bf16
# modes 4 real products x 3 or 6 part products (24 or 48 M N K) on the bf16 matrix cores.
GEMM_GROUPS = {5: ('3m', 0.75, 'fp32', 'mpdo::gemm_kernel, 3M products on fp32 MFMA (site contractions)'),
               6: ('bf16x3', 3.0, 'bf16', 'mpdo::gemm_bf16_kernel, 2-part split operands on bf16 MFMA (site contractions)'),
               7: ('bf16x6', 6.0, 'bf16', 'mpdo::gemm_bf16_kernel, 3-part split operands on bf16 MFMA (site contractions)')}
FP64_GROUP = ('fp64', 1.0, 'fp64', 'mpdo::gemm_kernel, fp64 MFMA (Grams, factors, fp64-mode contractions)')
DTYPES = {0: 'c64 data, fp64 accumulate/eigensolve',
          2: 'c64 data; site contractions 3M fp32-accumulate MFMA, Grams/QR/eigensolves fp64',
          3: 'c64 data; site contractions bf16x3 (2-part split, fp32 accumulate) MFMA, Grams/QR/eigensolves fp64',
          4: 'c64 data; site contractions bf16x6 (3-part split, fp32 accumulate) MFMA, Grams/QR/eigensolves fp64'}
N_QUBITS, CHI, KAPPA, P_NOISE, FIRST_LAYER = 32, 256, 8, 0.002, 10
# HBM bytes per GEMM launch of the timed step, from the two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE) of this benchmark, scaled by the copy calibration; the file records the commands
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'gemm_kernel_traffic_fused_layer_n32_chi256_kappa8_b32.md')
LAYER10_FIXTURE = os.path.join(ROOT, 'tests', 'golden', 'layer10_chi256_kappa8_oracle.npz')
CALIB_BYTES = 1 << 30


def _log(msg):
    print(f'# {msg}', file=sys.stderr, flush=True)


def _trace_marker():
    """A spin kernel of about a microsecond: brackets a window of a rocprofv3 kernel trace."""
    import torch
    torch.cuda._sleep(1000)


MARKER = 'spin_kernel'


def kernel_trace_summary(path, top=24):
    """Markdown summary of the two marked windows (timed steps, profiled pass) of a rocprofv3
    kernel trace (CSV: Kernel_Name, Start_Timestamp, End_Timestamp in ns) of a default run."""
    import csv
    rows = []
    with open(path, newline='') as f:
        rd = csv.DictReader(f)
        cols = {c.lower(): c for c in rd.fieldnames}
        kn, ks, ke = cols['kernel_name'], cols['start_timestamp'], cols['end_timestamp']
        for r in rd:
            rows.append((int(r[ks]), int(r[ke]), r[kn]))
    rows.sort()
    marks = [i for i, r in enumerate(rows) if MARKER in r[2]]
    if len(marks) != 4:
        raise SystemExit(f'{path}: expected 4 {MARKER} launches (two windows), found {len(marks)}')
    out = [f'Kernel trace `{os.path.basename(path)}`: {len(rows)} kernels; windows between the '
           f'`{MARKER}` markers.', '']

    def short(name):
        name = name.replace('void ', '').replace('(anonymous namespace)::', '')
        name = name.replace('HIP_vector_type<double, 2u>', 'double2').replace('HIP_vector_type<float, 2u>', 'float2')
        depth, cut = 0, len(name)
        for i, ch in enumerate(name):   # drop the argument list, keep the template arguments
            depth += ch == '<'
            depth -= ch == '>'
            if ch == '(' and depth == 0:
                cut = i
                break
        return name[:cut]

    for title, (a, b) in (('timed steps (no timing hooks)', (marks[0], marks[1])),
                          ('profiled pass (timing hooks on)', (marks[2], marks[3]))):
        win = rows[a + 1:b]
        if not win:
            continue
        agg = {}
        for t0, t1, nm in win:
            v = agg.setdefault(short(nm), [0, 0])
            v[0] += 1
            v[1] += t1 - t0
        tot = sum(v[1] for v in agg.values())
        span = win[-1][1] - win[0][0]
        # GEMM groups: fp64 MFMA (gemm_kernel, last template argument false), 3M on fp32 MFMA
        # (gemm_kernel ... true>), split-bf16 (gemm_bf16_kernel)
        groups = {'fp64 MFMA': lambda k: 'gemm_kernel' in k and not k.rstrip().endswith('true>'),
                  '3M fp32 MFMA': lambda k: 'gemm_kernel' in k and k.rstrip().endswith('true>'),
                  'split-bf16 MFMA': lambda k: 'gemm_bf16_kernel' in k}
        parts = []
        for gname, sel in groups.items():
            g = [v for k, v in agg.items() if sel(k)]
            gl, gt = sum(v[0] for v in g), sum(v[1] for v in g)
            if gl:
                parts.append(f'{gname}: {gl} launches, {gt / 1e6:.1f} ms, average {gt / gl / 1e6:.4f} ms per launch')
        out += [f'### {title}', '',
                f'{len(win)} kernels, {tot / 1e6:.1f} ms inside kernels, {span / 1e6:.1f} ms from the first start '
                f'to the last end. GEMM groups: ' + '; '.join(parts) + '.', '',
                '| kernel | launches | total ms | avg ms | % of kernel time |', '|---|---|---|---|---|']
        items = sorted(agg.items(), key=lambda kv: -kv[1][1])
        for k, (c, t) in items[:top]:
            out.append(f'| `{k}` | {c} | {t / 1e6:.1f} | {t / c / 1e6:.4f} | {100 * t / tot:.1f} |')
        rest = items[top:]
        if rest:
            c = sum(v[0] for _, v in rest)
            t = sum(v[1] for _, v in rest)
            out.append(f'| {len(rest)} other kernels | {c} | {t / 1e6:.1f} | {t / c / 1e6:.4f} | {100 * t / tot:.1f} |')
        out.append('')
    return '\n'.join(out)


GEMM_ROW_FP64 = 'mpdo::gemm_kernel, fp64 MFMA'
GEMM_ROW_3M = 'mpdo::gemm_kernel, 3M on fp32 MFMA'
GEMM_ROW_BF16 = 'mpdo::gemm_bf16_kernel, split-bf16 MFMA'


def counter_summary(path):
    """Markdown summary of a rocprofv3 --pmc run of this benchmark (CSV: Dispatch_Id, Kernel_Name,
    Counter_Name, Counter_Value): for each window between consecutive pairs of marker launches (with
    --copy-calibration the first window holds the calibration copies), the mean value of every
    counter per dispatch of each kernel, and over all mpdo::gemm_kernel dispatches."""
    import csv
    disp = {}
    with open(path, newline='') as f:
        rd = csv.DictReader(f)
        cols = {c.lower(): c for c in rd.fieldnames}
        kd, kn, cn, cv = cols['dispatch_id'], cols['kernel_name'], cols['counter_name'], cols['counter_value']
        for r in rd:
            d = disp.setdefault(int(r[kd]), {'name': r[kn], 'c': {}})
            d['c'][r[cn]] = d['c'].get(r[cn], 0.0) + float(r[cv])
    order = sorted(disp)
    marks = [i for i, k in enumerate(order) if MARKER in disp[k]['name']]
    out = [f'Counter run `{os.path.basename(path)}`: {len(order)} dispatches, {len(marks)} markers.', '']
    for w in range(len(marks) // 2):
        win = [disp[k] for k in order[marks[2 * w] + 1:marks[2 * w + 1]]]
        agg = {}
        for d in win:
            name = d['name'].replace('void ', '').replace('(anonymous namespace)::', '').split('(')[0]
            if 'gemm_bf16_kernel' in name:
                name = GEMM_ROW_BF16
            elif 'gemm_kernel' in name:   # the 3M instantiations end in `true>`
                name = GEMM_ROW_3M if name.rstrip().endswith('true>') else GEMM_ROW_FP64
            a = agg.setdefault(name, [0, {}])
            a[0] += 1
            for c, v in d['c'].items():
                a[1][c] = a[1].get(c, 0.0) + v
        counters = sorted({c for a in agg.values() for c in a[1]})
        out += [f'### window {w + 1}: {len(win)} dispatches', '',
                '| kernel | dispatches | ' + ' | '.join(f'mean {c} per dispatch' for c in counters) + ' |',
                '|---|---|' + '---|' * len(counters)]
        gemm_rows = [k for k in (GEMM_ROW_FP64, GEMM_ROW_3M, GEMM_ROW_BF16) if k in agg]
        rest = [k for k, _ in sorted(agg.items(), key=lambda kv: -kv[1][0]) if k not in gemm_rows][:12]
        for name in gemm_rows + rest:
            n, tot = agg[name]
            out.append(f'| `{name[:90]}` | {n} | ' + ' | '.join(f'{tot.get(c, 0.0) / n:.1f}' for c in counters) + ' |')
        out.append('')
    return '\n'.join(out)


def _traffic_per_launch(group='fp64'):
    """The measured HBM bytes per GEMM launch of one group (fp64, 3m, bf16x3, bf16x6) recorded in
    TRAFFIC_FILE (line `traffic_bytes_per_launch_<group> = N`), or None when absent."""
    import re
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r'traffic_bytes_per_launch_' + group + r' = ([0-9.e+]+)', f.read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def _layers(copies, count, dev, n=N_QUBITS, first=FIRST_LAYER):
    """Fused-layer operands of layers first .. first + count - 1 of the workload (the n-qubit
    member of its circuit family)."""
    import mpdo_amd as M
    from mpdo_amd import batched
    from mpdo_amd import workloads as W
    ws = W.brickwork32_chi256_family(copies=copies, n=n)
    cs = []
    for w in ws:
        c = M.TensorCircuit(device=str(dev), **w.circuit_kwargs())
        W.apply_ops(c, w.ops)
        cs.append(c)
    segs = batched._segments(cs[0])
    out = []
    for k in range(count):
        start, stop = segs[(first + k) % len(segs)]
        layer, why = batched.fused_layer_of(cs, start, stop)
        if layer is None:
            raise RuntimeError(f'layer {first + k} does not fuse: {why}')
        out.append(layer)
    return out


def _zz_codes(n):
    """<Z_i> (n strings), <Z_i Z_j> for every i < j, then the identity (Tr rho~)."""
    codes = []
    for i in range(n):
        c = [-1] * n
        c[i] = 2
        codes.append(c)
    for i in range(n):
        for j in range(i + 1, n):
            c = [-1] * n
            c[i] = c[j] = 2
            codes.append(c)
    codes.append([-1] * n)
    return codes


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or 'unknown'


def cpu_baseline(Ti, Tj, k1i, k1j, g2, threads):
    """The oracle (numpy complex128) on one pair (sites i, i+1 of copy 0, the tensors the first
    timed GPU layer starts from) through every step the layer applies to it: both sites' single-qubit
    channels (RY, RZ + depolarizing, composed into one Choi-reduced Kraus set as the GPU path does),
    the CNOT + two-qubit depolarizing gate-time split (Circuit.py:74-136), the QR step through the
    pair (TNNOptimizer.py:87-108), the chi step at the pair bond (TNNOptimizer.py:111-134,
    decompositions.py:51-146) and the kappa cuts of both sites (TNNOptimizer.py:164-197). A layer
    has 16 such pairs plus 15 chi steps and 15 QR steps between pairs, so 1 / (16 x the pair's
    time) bounds the oracle's rate from above."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import mpdo_oracle as O

    Ti, Tj = Ti.astype(np.complex128), Tj.astype(np.complex128)
    t = {}
    with threadpool_limits(threads):
        t0 = time.perf_counter()
        a = O.apply_1q(Ti, k1i)
        b = O.apply_1q(Tj, k1j)
        t['1q'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        a, b, _ = O.apply_2q(a, b, g2)
        t['split'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        l, _, a1, m = a.shape
        q, R = O.qr(a.reshape(l * 2 * a1, m))
        a = q.reshape(l, 2, a1, q.shape[1])
        b = np.tensordot(R, b, axes=([1], [0]))
        t['qr'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        l, _, a1, m = a.shape
        _, _, a2, r = b.shape
        us, svh, s, _ = O._sweep_split(a.reshape(l * 2 * a1, m), b.reshape(m, 2 * a2 * r), CHI, None, {})
        k = s.shape[0]
        a, b = us.reshape(l, 2, a1, k), svh.reshape(k, 2, a2, r)
        t['chi'] = time.perf_counter() - t0
        t0 = time.perf_counter()
        st = O.MPDO([a, b], [True, True], [True])
        O.svd_kappa_left2right(st, KAPPA)
        t['kappa'] = time.perf_counter() - t0
    total = sum(t.values())
    pairs = len(range(0, N_QUBITS - 1, 2))
    shapes = f'{tuple(Ti.shape)}, {tuple(Tj.shape)}'
    return {'value': 1.0 / (pairs * total), 'unit': 'circuit-layers/s', 'cores': threads, 'kind': 'port',
            'cpu': _cpu_model(), 'dtype': 'complex128 (numpy, LAPACK)',
            'sample': (f'oracle on one middle pair of the first timed circuit-layer (n=32, chi=256, kappa=8, copy 0, '
                       f'input shapes {shapes}): '
                       + ', '.join(f'{k} {v:.1f} s' for k, v in t.items())
                       + f', total {total:.1f} s; value = 1 / ({pairs} x total), an upper bound on the oracle\'s '
                       'rate (the 15 chi and QR steps between pairs are not timed)')}


def _uneven_batch(count, streams, dev):
    """The first `count` circuits of the 256-circuit uneven batch (n 8..24, depth 4..16, chi 16..128,
    kappa 2/4, depolarizing or amplitude damping) from |0..0> through scheduler.evolve_many on
    `streams` streams, and through one stream: aggregate circuit-layers/s."""
    import torch
    import mpdo_amd as M
    from mpdo_amd import scheduler
    from mpdo_amd import workloads as W
    ws = W.random_brickwork_batch()[:count]
    out = {'circuits': count, 'workload': 'workloads.random_brickwork_batch()[:%d] (seed 404)' % count}
    for s in (1, streams):
        cs = []
        for w in ws:
            c = M.TensorCircuit(device=str(dev), **w.circuit_kwargs())
            W.apply_ops(c, w.ops)
            cs.append(c)
        states = [M.Tools.create_ket0Series(w.n, device=str(dev)) for w in ws]
        torch.cuda.synchronize()
        st = {}
        scheduler.evolve_many(cs, states, streams=s, stats=st)
        out[f'streams_{s}'] = {'circuit_layers_per_s': round(st['circuit_layers'] / st['seconds'], 3),
                               'seconds': round(st['seconds'], 3), 'circuit_layers': st['circuit_layers'],
                               'fused_layers': st['fused_layers'], 'per_gate_layers': st['per_gate_layers']}
    return out


def _differentiation(dev):
    """One forward + reverse sweep of the 16-qubit, 8-layer ansatz (chi=64, kappa=4, 256 angles and
    376 per-gate depolarizing strengths, per-layer checkpointing): seconds and the peak device memory it adds."""
    import numpy as np
    import torch
    from mpdo_amd import autodiff as AD
    rng = np.random.Generator(np.random.PCG64(303))
    angles = rng.uniform(-np.pi, np.pi, size=(8, 16, 2))
    m = AD.HEA(16, 8, 64, 4, angles[..., 0], angles[..., 1], np.full((8, 16, 2), 0.005), np.full((8, 15), 0.005),
               device=str(dev))
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    base = torch.cuda.memory_allocated(dev)
    t0 = time.perf_counter()
    E = AD.tfim_energy(m())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    E.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {'workload': '16-qubit 8-layer RY/RZ + CNOT-chain ansatz, per-gate depolarizing p=0.005, chi=64, kappa=4, '
                        'loss <H> (TFIM h=1)', 'parameters': 632, 'energy': E.item(),
            'forward_s': round(t1 - t0, 3), 'backward_s': round(t2 - t1, 3),
            'peak_device_gib': round((torch.cuda.max_memory_allocated(dev) - base) / 2 ** 30, 2)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=6)
    ap.add_argument('--copies', type=int, default=32)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-threads', type=int, default=None,
                    help='threads of the CPU oracle sample (default: the cores this process may run on)')
    ap.add_argument('--math', choices=list(MATH_MODES), default=DEFAULT_MATH,
                    help='contractions whose results are complex64 site data: fp64 MFMA accumulation, three real '
                         'products per complex product accumulated in fp32 (3M), or bf16 matrix-core products of '
                         'operands split into 2 or 3 bf16 parts (fp32 accumulation); Gram matrices, QR/LQ factors '
                         'and eigensolves stay fp64')
    ap.add_argument('--hw-queues', type=int, default=8,
                    help='GPU_MAX_HW_QUEUES for this process (hardware queues HIP maps its streams onto; HIP\'s '
                         'default is 4), set before the GPU is initialised; 0 leaves the environment as it is')
    ap.add_argument('--streams', type=int, default=4,
                    help='the copies run as this many sub-batches, each through every warmup and timed layer on a '
                         'HIP stream of its own (batched.layers_batched_streams); the profiled pass runs one batch '
                         'on one stream')
    ap.add_argument('--uneven', type=int, default=64,
                    help='circuits of the 256-circuit uneven batch (workloads.random_brickwork_batch) to time through '
                         'scheduler.evolve_many after the main measurement (0: skip)')
    ap.add_argument('--uneven-streams', type=int, default=8)
    ap.add_argument('--no-differentiation', action='store_true',
                    help='skip timing one forward + backward of the 16-qubit differentiable ansatz')
    ap.add_argument('--fused-only', action='store_true',
                    help='time the fused steps only (no profiled pass, plain path or accuracy check): the '
                         'command of the rocprofv3 counter passes')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='before the warmup, run torch.neg over 1 GiB of float32 three times (1 GiB read and '
                         '1 GiB written per launch): the reference launch that scales the HBM counters of a '
                         'rocprofv3 --pmc run of this benchmark')
    ap.add_argument('--counter-summary', metavar='CSV',
                    help='print a markdown summary of a rocprofv3 --pmc run (counter_collection CSV) of this '
                         'benchmark, window by window between the markers, then exit (no GPU work)')
    ap.add_argument('--kernel-trace-summary', metavar='CSV',
                    help='print a markdown summary of the timed steps and the profiled pass of a rocprofv3 kernel '
                         'trace (CSV) of this benchmark, then exit (no GPU work)')
    args = ap.parse_args()
    for path in (args.kernel_trace_summary, args.counter_summary):
        if path is not None and not os.path.isfile(path):
            sys.exit(f'bench.py: no such file: {path!r}')
    if args.kernel_trace_summary:
        print(kernel_trace_summary(args.kernel_trace_summary))
        return
    if args.counter_summary:
        print(counter_summary(args.counter_summary))
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). '
              'Chains are independent, so more GPUs would each take a share of the copies; run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)

    if args.hw_queues > 0:   # read by the HIP runtime when it initialises (the first GPU call below)
        os.environ['GPU_MAX_HW_QUEUES'] = str(args.hw_queues)

    import numpy as np
    import torch

    from mpdo_amd import _lib, batched
    from mpdo_amd import workloads as W

    if not torch.cuda.is_available():
        raise SystemExit('bench.py needs the MI355X (cuda:0)')
    dev = torch.device('cuda:0')
    if args.copy_calibration:
        src = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        _trace_marker()
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        _trace_marker()
        del src, dst
        _log(f'copy calibration: 3 launches of torch.neg, {CALIB_BYTES} bytes read and {CALIB_BYTES} bytes '
             'written each')
    lib = _lib.load()
    math_mode = MATH_MODES[args.math]
    prev_mode = lib.mpdo_set_math_mode(math_mode)
    B = args.copies
    layers = _layers(B, args.warmup + args.steps, dev)
    dims = batched.steady_dims(N_QUBITS, CHI, KAPPA, layers[0].kraus1.shape[2], layers[0].gates2.shape[-1],
                               layers[0].parity)
    sites = W.steady_chains_device(dims, B, dev)
    torch.cuda.synchronize()
    _log(f'{B} copies, {N_QUBITS} qubits, chi={CHI}, kappa={KAPPA}; warmup {args.warmup}, steps {args.steps}')

    n_streams = len(batched.sub_batches(B, args.streams))

    def run(first, count, sites):
        outs, _ = batched.layers_batched_streams(sites, layers[first:first + count], CHI, KAPPA, streams=args.streams)
        return outs

    t = time.perf_counter()
    sites = run(0, args.warmup, sites)
    torch.cuda.synchronize()
    _log(f'warmup: {args.warmup} layers in {time.perf_counter() - t:.2f} s')

    saved = None if args.fused_only else [t.clone() for t in sites]   # input of the first timed layer
    _trace_marker()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sites = run(args.warmup, args.steps, sites)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    _trace_marker()
    value = B * args.steps / T
    _log(f'timed: {T:.2f} s, {value:.3f} circuit-layers/s')
    del sites
    fp64 = None
    if math_mode != 0 and not args.fused_only:
        # the same timed layers with every site contraction accumulated in fp64 (math mode 0)
        lib.mpdo_set_math_mode(0)
        s64 = [t.clone() for t in saved]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s64 = run(args.warmup, args.steps, s64)
        torch.cuda.synchronize()
        T64 = time.perf_counter() - t0
        del s64
        lib.mpdo_set_math_mode(math_mode)
        fp64 = {'value': round(B * args.steps / T64, 6), 'ms_per_step': round(T64 / args.steps * 1e3, 3),
                'math': 'every site contraction with fp64 MFMA accumulation (mpdo_set_math_mode(0)), same layers and input'}
        _log(f'fp64 accumulation: {T64:.2f} s, {fp64["value"]:.3f} circuit-layers/s')
    # the sub-batch streams' workspaces (3.6 GiB per chain) go back before the one-stream passes below
    batched.release_stream_pools(dev)
    torch.cuda.empty_cache()

    roofline, others, plain, accuracy, cpu = None, None, None, None, None
    if not args.fused_only:
        L0 = layers[args.warmup]
        # ---- profiled pass: the first timed layer again, on its kept input
        _trace_marker()
        _lib.prof_enable(True)
        outs, disc = _lib.layer_batched(saved, L0.kraus1, L0.parity, L0.gates2, CHI, KAPPA)
        torch.cuda.synchronize()
        _trace_marker()
        prof = _lib.prof_read()
        execd = _lib.prof_read_executed()
        eig_log = _lib.prof_eig_log()
        shapes = sorted(_lib.prof_shapes('gemm'), key=lambda r: -r[2])
        _lib.prof_enable(False)
        _log(f'profiled pass: gate-time split kept counts {_lib.layer_kept()}; '
             f'{lib.mpdo_debug_layer_gate_skips()} cuts decided by the shifted Cholesky (no eigensolver)')
        eigs = {}
        for (n_e, b_e, sw_e) in eig_log:
            c = eigs.setdefault(n_e, [0, 0])
            c[0] += 1
            c[1] += sw_e
        _log('blocked eigensolver calls (order: calls, mean sweeps): ' +
             ', '.join(f'{k}: {v[0]}, {v[1] / v[0]:.1f}' for k, v in sorted(eigs.items())))
        _log('GEMM shapes by time (M, N, K, batch, prec: launches, ms, TFLOP/s): ' +
             '; '.join(f'{sh}: {c}, {t:.0f}, {f / (t * 1e-3) / 1e12:.1f}' for sh, c, t, f in shapes[:14]))
        total_ms = max(sum(v[1] for v in prof.values()), 1e-9)
        n_l, ms, fl, by = prof['gemm']
        # the GEMM launches by arithmetic (GEMM_GROUPS): fp64 MFMA (complex products as four real
        # ones), the 3M products on fp32 MFMA and the split-bf16 products, each at its executed flops
        groups, info = {}, {}
        for sh, c, t, f in shapes:
            key, mult, pk, kern = GEMM_GROUPS.get(sh[4] % 10, FP64_GROUP)
            info[key] = (pk, kern)
            a = groups.setdefault(key, [0, 0.0, 0.0])
            a[0] += c
            a[1] += t
            a[2] += f * mult
        gemm_groups = {}
        for key, (c, t, f) in groups.items():
            peak = PEAK_TFLOPS[info[key][0]]
            a = f / (t * 1e-3) / 1e12
            gemm_groups[key] = {'kernel': info[key][1],
                                'bound': 'mfma', 'achieved': round(a, 3), 'peak': peak, 'unit': 'TFLOP/s',
                                'frac': round(a / peak, 4), 'launches': c, 'avg_launch_ms': round(t / max(c, 1), 4),
                                'executed_flops_per_launch': f / max(c, 1), 'share_of_kernel_time': round(t / total_ms, 3),
                                'traffic': _traffic_per_launch(key)}
        dom = max(gemm_groups, key=lambda k: gemm_groups[k]['share_of_kernel_time'])
        roofline = dict(gemm_groups[dom])
        roofline.update({'traffic_scope': 'HBM bytes per launch of this group over the timed steps (FETCH_SIZE and '
                                          'WRITE_SIZE counter passes scaled by the copy calibration)',
                         'gemm_all_launches': {'launches': n_l, 'avg_launch_ms': round(ms / max(n_l, 1), 4),
                                               'share_of_kernel_time': round(ms / total_ms, 3),
                                               'operand_bytes_per_launch': by / max(n_l, 1)},
                         'gemm_groups': gemm_groups,
                         'classes_ms': {k: round(v[1], 1) for k, v in prof.items()}})
        names = {'eig_update': 'bj_update (two-sided 64x64 block-pair rotation of A, fp64 MFMA)',
                 'eig_rotate': 'bj_rotate_v_direct (eigenvector tile update V W, fp64 MFMA)',
                 'eig_solve': 'bj_solve (64x64 subproblem Jacobi sweep in LDS, fp64 vector ALU)'}
        others = []
        for cls, kname in names.items():
            n_c, ms_c, fl_c, by_c = prof[cls]
            ex_fl, ex_by = execd[cls]
            if n_c == 0 or ms_c <= 0:
                continue
            a_c = ex_fl / (ms_c * 1e-3) / 1e12
            others.append({'kernel': kname, 'bound': 'mfma' if cls != 'eig_solve' else 'fp64 vector ALU',
                           'achieved': round(a_c, 3), 'peak': PEAK_TFLOPS['fp64'], 'unit': 'TFLOP/s',
                           'frac': round(a_c / PEAK_TFLOPS['fp64'], 4), 'launches': n_c,
                           'avg_launch_ms': round(ms_c / n_c, 4), 'executed_flops_per_launch': ex_fl / n_c,
                           'executed_bytes_per_launch': ex_by / n_c,
                           'hbm_gbs_of_executed_bytes': round(ex_by / (ms_c * 1e-3) / 1e9, 1),
                           'share_of_kernel_time': round(ms_c / total_ms, 3)})
        # ---- accuracy of the batched layer: copy 0 against the per-gate path (fp64 accumulation)
        codes = _zz_codes(N_QUBITS)
        v_f = np.array(_lib.pauli_expect([t[0].contiguous() for t in outs], codes))
        chain0 = [t[0] for t in saved]
        del outs
        lib.mpdo_set_math_mode(0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        ref, gate, d_chi, d_kap = batched.per_gate_layer(chain0, L0, 0, CHI, KAPPA)
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t
        v_g = np.array(_lib.pauli_expect(ref, codes))
        tr = abs(v_g[-1])
        got_chi = np.array([disc[0, i] for i in reversed(range(N_QUBITS - 1))])
        chi_rel = float(np.max(np.abs(got_chi - np.array(d_chi)) / np.maximum(np.abs(d_chi), 1e-30)))
        k_f, k_g = float(disc[0, N_QUBITS:2 * N_QUBITS].sum()), float(sum(d_kap.values()))
        del ref
        # ---- the plain per-circuit FP32 GPU path: same input, copy 0, fp32 accumulation
        lib.mpdo_set_math_mode(1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        pl, _, _, _ = batched.per_gate_layer(chain0, L0, 0, CHI, KAPPA)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t
        lib.mpdo_set_math_mode(0)
        v_p = np.array(_lib.pauli_expect(pl, codes))
        del pl
        plain = {'value': round(1.0 / t_plain, 6), 'unit': 'circuit-layers/s', 'seconds_per_circuit_layer': round(t_plain, 3),
                 'path': 'per-gate entry points one gate at a time (mpdo_apply_1q, mpdo_apply_2q, bondTruncate, '
                         'svdKappa_left2right), one chain, fp32 MFMA accumulation (mpdo_set_math_mode(1)), on copy 0 '
                         'of the first timed layer\'s input',
                 'batched_vs_plain': round(value * t_plain, 2),
                 'fp64_per_gate_seconds': round(t_ref, 3)}
        accuracy = {'reference': 'per-gate path, fp64 accumulation, copy 0, same input and layer',
                    'zz_z_trace_max_abs_err_over_trace': float(np.abs(v_f - v_g).max() / tr),
                    'chi_weights_max_rel_err': chi_rel,
                    'kappa_total_rel_err': abs(k_f - k_g) / max(k_g, 1e-30),
                    'plain_fp32_zz_max_abs_err_over_trace': float(np.abs(v_p - v_g).max() / tr)}
        # ---- one chi=256, kappa=8 layer at 10 qubits against its complex128 oracle fixture
        if os.path.exists(LAYER10_FIXTURE):
            g = np.load(LAYER10_FIXTURE)
            dims10 = [tuple(int(x) for x in d) for d in batched.steady_dims(10, CHI, KAPPA, 4, 16, 1)]
            ch = [torch.from_numpy(a).to(dev).unsqueeze(0).contiguous() for a in W.steady_state(dims10, seed=7)]
            L10 = _layers(1, 1, dev, n=10, first=1)[0]    # the fixture's layer: pairs (1,2),(3,4),...
            lib.mpdo_set_math_mode(math_mode)
            o10, _ = _lib.layer_batched(ch, L10.kraus1, L10.parity, L10.gates2, CHI, KAPPA)
            lib.mpdo_set_math_mode(0)
            c10 = _zz_codes(10)
            v10 = np.array(_lib.pauli_expect([t[0].contiguous() for t in o10], c10))
            ref10 = list(g['z']) + [g['zz'][i, j] for i in range(10) for j in range(i + 1, 10)] + [float(g['trace'])]
            accuracy['oracle_complex128_layer10_zz_max_abs_err_over_trace'] = float(
                np.abs(v10 - np.array(ref10)).max() / float(g['trace']))
        _log(f'accuracy: {accuracy}')
        _log(f'plain: {plain}')
        cpu_in = None
        if not args.no_cpu_baseline:
            i = N_QUBITS // 2 - 2 + L0.parity
            k1 = L0.kraus1[0].cpu().numpy()
            cpu_in = (saved[i][0].cpu().numpy(), saved[i + 1][0].cpu().numpy(),
                      np.transpose(k1[i], (1, 2, 0)), np.transpose(k1[i + 1], (1, 2, 0)),
                      L0.gates2[(i - L0.parity) // 2].cpu().numpy())
        del saved
        if cpu_in is not None:
            threads = args.cpu_threads or len(os.sched_getaffinity(0))
            _log(f'cpu baseline (oracle sample, {threads} threads)')
            cpu = cpu_baseline(*cpu_in, threads)
    lib.mpdo_set_math_mode(prev_mode)
    _lib.release_workspaces()    # the fused layer's workspace (3.6 GiB per chain) is not needed below
    torch.cuda.empty_cache()
    uneven = None
    if not args.fused_only and args.uneven > 0:
        uneven = _uneven_batch(args.uneven, args.uneven_streams, dev)
        _log(f'uneven batch: {uneven}')
    diff = None
    if not args.fused_only and not args.no_differentiation:
        diff = _differentiation(dev)
        _log(f'differentiation: {diff}')
    line = {'metric': METRIC, 'value': round(value, 6), 'unit': 'circuit-layers/s', 'n_gpus': 1,
            'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(T / args.steps * 1e3, 3),
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': DTYPES.get(math_mode, DTYPES[0]),
            'data': 'synthetic',
            'config': {'workload': ('32-qubit depth-30 RY/RZ+CNOT brickwork, depolarizing p=0.002 (1q, 2q), '
                                    'chi=256, kappa=8, 32 angle-perturbed copies; step = one circuit-layer of '
                                    'every copy from the steady state (fused batched layers; the copies as '
                                    f'{n_streams} sub-batches, each on a HIP stream of its own)'),
                       'n_qubits': N_QUBITS, 'chi': CHI, 'kappa': KAPPA, 'copies': B, 'global_batch': B,
                       'streams': n_streams, 'hw_queues': os.environ.get('GPU_MAX_HW_QUEUES', '4 (HIP default)'),
                       'parallelism': 'none (one GPU)'},
            'roofline': roofline, 'roofline_block_jacobi': others, 'plain': plain,
            'batched_vs_plain': plain['batched_vs_plain'] if plain else None, 'fp64_accumulate': fp64,
            'accuracy': accuracy,
            'uneven_batch': uneven, 'differentiation': diff, 'cpu_baseline': cpu}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
