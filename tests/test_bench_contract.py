"""The benchmark's contract pieces that need no GPU: it refuses more than one GPU before touching
one, the roofline's HBM traffic is the figure of the committed counter profile, and the workload
family it times is the documented circuit."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 2
    assert 'measures one GPU' in p.stderr
    assert p.stdout == ''   # no JSON line


def test_traffic_is_the_profile_figure():
    """The roofline's HBM traffic per GEMM launch, group by group, is the figure the committed
    counter profile states, and that figure is its table's raw counters times the calibration
    scales (FETCH_SIZE x 2, WRITE_SIZE x 1, KiB)."""
    import bench
    with open(bench.TRAFFIC_FILE) as f:
        text = f.read()
    labels = {'fp64 MFMA': 'fp64', '3M on fp32 MFMA': '3m', 'split-bf16 MFMA, two parts': 'bf16x3',
              'split-bf16 MFMA, three parts': 'bf16x6'}
    rows = {g: (float(fe), float(wr), float(by)) for g, fe, wr, by in re.findall(
        r'\| (' + '|'.join(labels) + r') \| [0-9]+ \| ([0-9.]+) \| ([0-9.]+) \| ([0-9.e+]+) \|', text)}
    # the fp64 group and the group of the benchmark's default site-contraction arithmetic
    prec = {0: 4, 2: 5, 3: 6, 4: 7}[bench.MATH_MODES[bench.DEFAULT_MATH]]
    default = bench.GEMM_GROUPS.get(prec, bench.FP64_GROUP)[0]
    assert {'fp64', default} <= {labels[r] for r in rows}
    for row in rows:
        key = labels[row]
        want = float(re.search(r'traffic_bytes_per_launch_' + key + r' = ([0-9.e+]+)', text).group(1))
        assert bench._traffic_per_launch(key) == want
        fe, wr, by = rows[row]
        assert abs(by - want) <= 1e-4 * want
        assert abs((2.0 * fe + wr) * 1024 - want) <= 1e-4 * want


def test_timed_workload_family():
    """32-qubit-style family at a small size: copies share qubits and gate names, copy 0 keeps the
    base angles, the others differ by at most 0.01 per angle; chi 256, kappa 8, p = 0.002."""
    from mpdo_amd import workloads as W
    ws = W.brickwork32_chi256_family(copies=3, n=6, depth=4)
    assert [w.name for w in ws] == ['brickwork32_chi256_0', 'brickwork32_chi256_1', 'brickwork32_chi256_2']
    for w in ws:
        assert (w.n, w.chi, w.kappa, w.depth, w.noise_type) == (6, 256, 8, 4, 'idealNoise')
        assert w.chip == dict(dpc_error_rate=0.002) and w.p1 == 0.002
        assert [(o[0], o[1]) for o in w.ops] == [(o[0], o[1]) for o in ws[0].ops]
    for w in ws[1:]:
        d = [abs(a - b) for o, o0 in zip(w.ops, ws[0].ops) for a, b in zip(o[2], o0[2])]
        assert d and max(d) <= 0.01 + 1e-12 and max(d) > 0


def _write_csv(path, header, rows):
    import csv
    with open(path, 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_trace_and_counter_summaries(tmp_path):
    """The two rocprofv3 summaries on synthetic CSVs: windows between the marker launches, GEMM
    groups by arithmetic (fp64 gemm_kernel, 3M gemm_kernel ... true>, gemm_bf16_kernel), per-kernel
    averages, and mean counter values per dispatch."""
    import bench
    g64 = 'void mpdo::gemm_kernel<HIP_vector_type<double, 2u>, HIP_vector_type<double, 2u>, double, ' \
          'HIP_vector_type<double, 2u>, 64, 64, 16, 2, false>(mpdo::GemmArgs)'
    g3m = 'void mpdo::gemm_kernel<float2, float2, float, float2, 128, 64, 16, 1, true>(mpdo::GemmArgs)'
    gbf = 'void mpdo::gemm_bf16_kernel<3, 64, 128>(mpdo::GemmArgs)'
    mark = 'void spin_kernel(long)'
    # window 1: two fp64 GEMMs (3 and 5 us) and one bf16 (10 us); window 2: one 3M (4 us)
    trace = [(mark, 0, 1), (g64, 10, 3010), (g64, 4000, 9000), (gbf, 9000, 19000), (mark, 20000, 20001),
             (mark, 30000, 30001), (g3m, 31000, 35000), (mark, 40000, 40001)]
    tp = tmp_path / 'bench_kernel_trace.csv'
    _write_csv(tp, ['Kernel_Name', 'Start_Timestamp', 'End_Timestamp'], trace)
    text = bench.kernel_trace_summary(str(tp))
    assert 'fp64 MFMA: 2 launches, 0.0 ms, average 0.0040 ms per launch' in text
    assert 'split-bf16 MFMA: 1 launches, 0.0 ms, average 0.0100 ms per launch' in text
    assert '3M fp32 MFMA: 1 launches, 0.0 ms, average 0.0040 ms per launch' in text
    assert '`mpdo::gemm_bf16_kernel<3, 64, 128>` | 1 |' in text
    # counters: window 1 (calibration) one neg kernel, window 2 GEMMs of each group
    neg = 'void at::native::vectorized_elementwise_kernel<4, at::native::neg_kernel_cuda>(int)'
    disp = [(1, mark, 0), (2, neg, 100.0), (3, mark, 0), (4, mark, 0), (5, g64, 10.0), (6, g64, 30.0),
            (7, gbf, 50.0), (8, g3m, 7.0), (9, mark, 0)]
    cp = tmp_path / 'fetch_counter_collection.csv'
    _write_csv(cp, ['Dispatch_Id', 'Kernel_Name', 'Counter_Name', 'Counter_Value'],
               [(d, n, 'FETCH_SIZE', v) for d, n, v in disp])
    text = bench.counter_summary(str(cp))
    assert '| `mpdo::gemm_kernel, fp64 MFMA` | 2 | 20.0 |' in text
    assert '| `mpdo::gemm_bf16_kernel, split-bf16 MFMA` | 1 | 50.0 |' in text
    assert '| `mpdo::gemm_kernel, 3M on fp32 MFMA` | 1 | 7.0 |' in text
    assert 'neg_kernel_cuda' in text and '| 1 | 100.0 |' in text
