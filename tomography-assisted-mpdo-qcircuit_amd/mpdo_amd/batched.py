"""Evolution of a family of circuits with one structure over many MPDO chains at once.

A family is a list of ``TensorCircuit`` objects with the same qubit count, truncation settings,
noise model and layer list (same gate types on the same qubits; parameters may differ per copy),
for instance the perturbed copies of one ansatz. ``evolve_batched`` gives the same result per
copy as ``TensorCircuit.evolve`` (Circuit.py:469-491), with one difference in what it runs:

every circuit-layer of the brickwork form

    single-qubit gates ... | one two-qubit gate on every pair (i, i+1), i = p, p+2, ... | truncate

whose truncation finds the chain connected (TNNOptimizer.py:51-61) runs as ONE call of
``mpdo_layer_batched`` over all copies. That entry point applies the composed single-qubit
channel of every site, the two-qubit channel of every pair and the truncation (QR sweep, chi
sweep, kappa cut) without forming the two-site tensors of the gate-time split (their bond is
up to 4 chi), including that split's cut of singular values below ``GLOBAL_MINIMUM``
(Circuit.py:22,110-124), and makes the truncation decisions of the per-gate path (same kept
counts, shapes and discarded weights, ``('gate', ...)`` entries included). Every other layer runs
per copy through the per-gate path.
"""
from __future__ import annotations

import os
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import gates as G
from .circuit import TensorCircuit
from .node import SiteNode
from .noise import choi_compress, compose_kraus

__all__ = ['evolve_batched', 'FusedLayer', 'fused_layer_of', 'steady_dims']


class FusedLayer:
    """One brickwork circuit-layer in the operands of ``mpdo_layer_batched``."""

    __slots__ = ('parity', 'kraus1', 'noisy1', 'gates2', 'noisy2')

    def __init__(self, parity: int, kraus1: torch.Tensor, noisy1: List[bool], gates2: torch.Tensor, noisy2: bool):
        self.parity = parity        # first pair (parity, parity + 1)
        self.kraus1 = kraus1        # (B, n, K1, 2, 2) complex128, operator k as [out][in]
        self.noisy1 = noisy1        # per site: a noisy single-qubit channel (the site gains I_i)
        self.gates2 = gates2        # (npairs, 2, 2, 2, 2, K2) complex128 (o_min, o_max, i_min, i_max, k)
        self.noisy2 = noisy2        # two-qubit tensors carry a Kraus axis (the max site gains I_i)


def _segments(c: TensorCircuit) -> List[Tuple[int, int]]:
    """[start, stop) index ranges of the layer list, each ending after a truncation (the last one
    possibly without)."""
    out, start = [], 0
    for i, layer in enumerate(c.layers):
        if 'truncate' in layer.name.lower():
            out.append((start, i + 1))
            start = i + 1
    if start < len(c.layers):
        out.append((start, len(c.layers)))
    return out


def _same_structure(cs: Sequence[TensorCircuit]) -> Optional[str]:
    c0 = cs[0]
    for c in cs[1:]:
        if (c.qnumber, c.chi, c.kappa, c.max_truncation_err, c.ideal, c.noiseType, str(c.device)) != \
                (c0.qnumber, c0.chi, c0.kappa, c0.max_truncation_err, c0.ideal, c0.noiseType, str(c0.device)):
            return 'circuits differ in qubit count, truncation settings, noise or device'
        if len(c.layers) != len(c0.layers) or c._oqs_list != c0._oqs_list:
            return 'circuits differ in their layer lists'
        for a, b in zip(c.layers, c0.layers):
            if type(a) is not type(b):
                return 'circuits differ in their gate types'
    return None


def fused_layer_of(cs: Sequence[TensorCircuit], start: int, stop: int) -> Tuple[Optional[FusedLayer], str]:
    """The fused operands of layer-list entries [start, stop) of every copy, or (None, reason) when
    the entries are not one brickwork circuit-layer with a truncation at the end."""
    c0 = cs[0]
    n = c0.qnumber
    if c0.chi is None or c0.max_truncation_err is not None:
        return None, 'fused layers need a chi cap and no error threshold'
    if 'truncate' not in c0.layers[stop - 1].name.lower():
        return None, 'no truncation at the end of the layer'
    singles: List[List[int]] = []     # layer indices of single-qubit gates per qubit
    pairs: List[Tuple[int, int]] = []  # (min qubit, layer index)
    singles = [[] for _ in range(n)]
    seen2 = set()
    for i in range(start, stop - 1):
        layer = c0.layers[i]
        name = layer.name.lower()
        if 'barrier' in name:
            continue
        if 'truncate' in name:
            return None, 'truncation inside the layer'
        if not isinstance(layer, G.QuantumGate) or layer.name == 'MeasureZ':
            return None, f'entry {layer.name} is not a gate'
        oqs = c0._oqs_list[i]
        if layer.single:
            for q in oqs:
                if q in seen2:
                    return None, 'single-qubit gate after a two-qubit gate on its qubit'
                singles[q].append(i)
        else:
            if len(oqs) != 2 or abs(oqs[0] - oqs[1]) != 1:
                return None, 'two-qubit gate on non-neighbouring qubits'
            mn = min(oqs)
            if mn in seen2 or mn + 1 in seen2:
                return None, 'two two-qubit gates on one qubit'
            seen2.update((mn, mn + 1))
            pairs.append((mn, i))
    if not pairs:
        return None, 'no two-qubit gates'
    parity = min(p for p, _ in pairs) % 2
    want = list(range(parity, n - 1, 2))
    if sorted(p for p, _ in pairs) != want:
        return None, 'two-qubit gates do not cover every pair of one parity'
    pairs.sort()
    # two-qubit tensors: shared by all copies
    gl, noisy2 = [], None
    for mn, i in pairs:
        Gt = c0._double_kraus(c0.layers[i], c0._oqs_list[i])
        for c in cs[1:]:
            if not torch.equal(c._double_kraus(c.layers[i], c._oqs_list[i]), Gt):
                return None, 'two-qubit channels differ between copies'
        d5 = Gt.dim() == 5
        if noisy2 is not None and d5 != noisy2:
            return None, 'mixed noisy and ideal two-qubit gates'
        noisy2 = d5
        gl.append(Gt if d5 else Gt.unsqueeze(-1))
    K2 = gl[0].shape[-1]
    if any(g.shape[-1] != K2 for g in gl):
        return None, 'two-qubit channels of different Kraus rank'
    gates2 = torch.stack(gl)
    # single-qubit channels per copy and site
    ks, noisy1 = [], [False] * n
    K1 = None
    for c in cs:
        row = []
        for q in range(n):
            E, nz = None, False
            for i in singles[q]:
                Ei, gi = c._single_kraus(c.layers[i])
                E = Ei if E is None else compose_kraus(E, Ei)
                nz = nz or gi
            if E is None:
                E = torch.eye(2, dtype=torch.complex128)
            if E.dim() == 3 and E.shape[2] > 1 and c.fuse:
                E = choi_compress(E)
            if E.dim() == 2:
                E = E.unsqueeze(-1)
            noisy1[q] = noisy1[q] or nz
            if K1 is None:
                K1 = E.shape[2]
            if E.shape[2] != K1:
                return None, 'single-qubit channels of different Kraus rank'
            row.append(E.permute(2, 0, 1))
        ks.append(torch.stack(row))
    return FusedLayer(parity, torch.stack(ks), noisy1, gates2, bool(noisy2)), ''


_POOLS: dict = {}
_POOL_LOCK = threading.Lock()


def _stream_pool(dev: torch.device, k: int) -> List[torch.cuda.Stream]:
    """k streams of `dev` (default priority), the same ones on every call: the library binds one
    workspace per stream, so reusing the streams reuses the workspaces."""
    key = (dev.index if dev.index is not None else torch.cuda.current_device(), k)
    with _POOL_LOCK:
        pool = _POOLS.get(key)
        if pool is None:
            pool = _POOLS[key] = [torch.cuda.Stream(device=dev) for _ in range(k)]
    return pool


def release_stream_pools(dev: torch.device) -> None:
    """Drop the workspaces bound to the streams of ``layers_batched_streams`` and of
    ``evolve_batched(..., streams > 1)`` on ``dev`` (after their
    work): the memory returns to torch's caching allocator."""
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    with _POOL_LOCK:
        pools = [p for (d, _), p in _POOLS.items() if d == idx]
    for pool in pools:
        for s in pool:
            with torch.cuda.device(dev), torch.cuda.stream(s):
                s.synchronize()
                _lib.release_stream_workspace(dev)


def _max_side_streams() -> int:
    """Streams that run concurrently beside the caller's: HIP maps a process's streams onto
    GPU_MAX_HW_QUEUES hardware queues (4 unless set), and streams that share a queue run one after
    the other. Measured at the 32-qubit chi = 256, kappa = 8 benchmark (32 copies): with 4 queues,
    2, 3, 4 and 8 sub-batch streams beside the caller's gave 8.78, 8.96, 6.99 and 5.72
    circuit-layers/s against 8.39 on one stream; with 8 queues, 4 and 6 streams 9.28 and 8.43."""
    try:
        q = int(os.environ.get('GPU_MAX_HW_QUEUES', '4'))
    except ValueError:
        q = 4
    return max(1, q - 1)


def sub_batches(B: int, streams: int) -> List[Tuple[int, int]]:
    """The contiguous (lo, hi) chain ranges that ``layers_batched_streams`` and
    ``evolve_batched(streams=)`` run on separate streams: up to ``streams`` of them, at most one
    fewer than the hardware queues."""
    k = max(1, min(int(streams), B, _max_side_streams()))
    bs = -(-B // k)
    return [(lo, min(B, lo + bs)) for lo in range(0, B, bs)]


def layers_batched_streams(sites: Sequence[torch.Tensor], layers: Sequence['FusedLayer'], chi: Optional[int],
                           kappa: Optional[int], streams: int = 4):
    """Consecutive fused layers (``_lib.layer_batched`` with each layer's kraus1, parity, gates2)
    over B chains split into up to ``streams`` sub-batches, each advanced through every layer by a
    host thread of its own on a HIP stream of its own, with no synchronisation between the
    sub-batches until the end (the caller's stream then continues after all of them). The fused
    layer is partly latency-bound — Cholesky panels and Jacobi rounds run one workgroup per chain,
    each pair reads its kept counts back — so sub-batches that drift apart overlap one's
    latency-bound phases with another's matrix-core work. Per chain the same steps as one call
    over all B chains per layer (results equal up to rounding: GEMM tile choices may depend on the
    batch size). At most one stream fewer than the hardware queues (_max_side_streams).
    Returns (sites after the last layer, [discarded (B, 3n) per layer])."""
    B, n = sites[0].shape[0], len(sites)
    parts = sub_batches(B, streams)
    if len(parts) == 1:
        disc = []
        for L in layers:
            sites, d = _lib.layer_batched(sites, L.kraus1, L.parity, L.gates2, chi, kappa)
            disc.append(d)
        return list(sites), disc
    dev = sites[0].device
    pool = _stream_pool(dev, len(parts))
    caller = torch.cuda.current_stream(dev)
    for s in pool:   # the sub-batches start after the caller's pending work (their inputs)
        s.wait_stream(caller)
    results: List[Optional[List[torch.Tensor]]] = [None] * len(parts)
    discs = [np.zeros((B, 3 * n), dtype=np.float64) for _ in layers]
    errors: List[BaseException] = []

    def work(j: int, lo: int, hi: int):
        try:
            with torch.cuda.device(dev), torch.cuda.stream(pool[j]):
                cur = [t[lo:hi] for t in sites]
                for L, d in zip(layers, discs):
                    cur, dd = _lib.layer_batched(cur, L.kraus1[lo:hi], L.parity, L.gates2, chi, kappa)
                    d[lo:hi] = dd
                results[j] = cur
        except BaseException as e:   # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(j, lo, hi), daemon=True) for j, (lo, hi) in enumerate(parts)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in pool:
        caller.wait_stream(s)
    if errors:
        raise errors[0]
    # the inputs, allocated on the caller's stream, were held until the caller's stream is ordered
    # after every use; the sub-batches' outputs were allocated on their streams and are read below
    # on the caller's
    for j, r in enumerate(results):
        for t in r:
            t.record_stream(caller)
    out = [torch.cat([r[i] for r in results]) for i in range(n)]
    return out, discs


def _stack_sites(states: Sequence[Sequence[SiteNode]], cache) -> List[torch.Tensor]:
    """(B, l, 2, a, r) per site; reuses the batched output of the previous fused layer when every
    copy still holds its views."""
    B, n = len(states), len(states[0])
    if cache is not None:
        outs, views = cache
        if all(states[b][i].tensor is views[b][i] for b in range(B) for i in range(n)):
            return outs
    return [torch.stack([states[b][i].tensor for b in range(B)]).contiguous() for i in range(n)]


def evolve_batched(circuits: Sequence[TensorCircuit], states: Sequence[List[SiteNode]], *, fuse_layers: bool = True,
                   stats: Optional[dict] = None, streams: int = 1) -> None:
    """Evolve ``states[b]`` under ``circuits[b]`` for every copy b, in place (see the module
    docstring). ``stats`` (optional dict) receives the count of fused and per-gate layers.
    ``streams`` > 1 splits the copies into that many groups, each evolved by a host thread of its
    own on a HIP stream of its own (as ``layers_batched_streams`` does for fused layers)."""
    cs = list(circuits)
    if not cs or len(cs) != len(states):
        raise ValueError('one state per circuit is required')
    why = _same_structure(cs)
    if why:
        raise ValueError(why)
    parts = sub_batches(len(cs), streams)
    if len(parts) > 1:
        _evolve_groups(cs, states, parts, fuse_layers, stats)
        return
    for c, st in zip(cs, states):
        c._begin(st)
    n = cs[0].qnumber
    B = len(cs)
    cache = None
    nf = ng = 0
    for start, stop in _segments(cs[0]):
        layer, reason = (fused_layer_of(cs, start, stop) if fuse_layers else (None, 'disabled'))
        if layer is not None:
            ok, reason = _fusable_state(cs, states, layer)
            if not ok:
                layer = None
        if layer is None:
            for c, st in zip(cs, states):
                for i in range(start, stop):
                    c._step(st, i)
            cache = None
            ng += 1
            continue
        c0 = cs[0]
        sites = _stack_sites(states, cache)
        kappa = c0.kappa if (not c0.ideal and c0.kappa is not None) else None
        outs, disc = _lib.layer_batched(sites, layer.kraus1, layer.parity, layer.gates2, c0.chi, kappa)
        a_before = [sites[i].shape[3] * layer.kraus1.shape[2] for i in range(n)]
        for i in range(layer.parity, n - 1, 2):
            a_before[i + 1] *= layer.gates2.shape[-1]
        views = [[outs[i][b] for i in range(n)] for b in range(B)]
        for b, (c, st) in enumerate(zip(cs, states)):
            for i in range(n):
                nd = st[i]
                nd.tensor = views[b][i]
                if layer.noisy1[i]:
                    nd.has_I = True
            for i in range(layer.parity, n - 1, 2):
                st[i].bond_right = True
                st[i + 1].bond_left = True
                if layer.noisy2:
                    st[i + 1].has_I = True
            if c.record_discarded:
                c.discarded.extend(('gate', i, float(disc[b, 2 * n + i])) for i in range(layer.parity, n - 1, 2))
                c.discarded.extend(('chi', i + 1, float(disc[b, i])) for i in reversed(range(n - 1)))
                if kappa is not None:
                    c.discarded.extend(('kappa', i, float(disc[b, n + i])) for i in range(n)
                                       if st[i].has_I and a_before[i] > kappa)
        cache = (outs, views)
        nf += 1
    for c, st in zip(cs, states):
        c._finish(st)
    if stats is not None:
        stats['fused_layers'] = stats.get('fused_layers', 0) + nf
        stats['per_gate_layers'] = stats.get('per_gate_layers', 0) + ng


def _evolve_groups(cs, states, parts, fuse_layers, stats):
    """evolve_batched of each group of copies on a stream and host thread of its own."""
    dev = torch.device(cs[0].device)
    pool = _stream_pool(dev, len(parts))
    caller = torch.cuda.current_stream(dev)
    for s in pool:
        s.wait_stream(caller)
    sts = [dict() for _ in parts]
    errors: List[BaseException] = []

    def work(j: int, lo: int, hi: int):
        try:
            with torch.cuda.device(dev), torch.cuda.stream(pool[j]):
                evolve_batched(cs[lo:hi], states[lo:hi], fuse_layers=fuse_layers, stats=sts[j])
                # the groups' site tensors are read later on other streams
                for st in states[lo:hi]:
                    for nd in st:
                        nd.tensor.record_stream(caller)
        except BaseException as e:   # re-raised in the caller's thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(j, lo, hi), daemon=True) for j, (lo, hi) in enumerate(parts)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in pool:
        caller.wait_stream(s)
    if errors:
        raise errors[0]
    if stats is not None:   # every group runs the same layer list
        for k in ('fused_layers', 'per_gate_layers'):
            stats[k] = stats.get(k, 0) + sts[0].get(k, 0)


def steady_dims(n: int, chi: int, kappa: Optional[int], K1: int, K2: int, parity: int, max_layers: int = 64):
    """Site shapes (l, a, r) that a long brickwork circuit (alternating pair parity, K1/K2 Kraus
    operators per single-/two-qubit channel) has on entry to a layer of the given parity: the
    fixed point of the layer's shape rule (mpdo_layer_dims), iterated from inner dim kappa and
    bonds min(chi, 4^distance to the nearer end)."""
    k = kappa or 1
    bonds = [min(chi, 4 ** (i + 1), 4 ** (n - 1 - i)) for i in range(n - 1)]
    dims = [(1 if i == 0 else bonds[i - 1], k, 1 if i == n - 1 else bonds[i]) for i in range(n)]
    par = 1 - parity
    hist = []
    for _ in range(max_layers):
        out, why = _lib.layer_dims(dims, K1, par, K2, chi, kappa)
        if out is None:
            raise ValueError(f'no fused steady state: {why}')
        dims, par = out, 1 - par
        if par == parity:
            if hist and hist[-1] == dims:
                return dims
            hist.append(dims)
    raise ValueError('shapes did not settle')


def _fusable_state(cs, states, layer: FusedLayer) -> Tuple[bool, str]:
    n = cs[0].qnumber
    ref = states[0]
    for st in states[1:]:
        for a, b in zip(st, ref):
            if a.tensor.shape != b.tensor.shape or (a.has_I, a.bond_left, a.bond_right) != \
                    (b.has_I, b.bond_left, b.bond_right):
                return False, 'copies differ in site shapes or flags'
    pair_min = set(range(layer.parity, n - 1, 2))
    for i in range(n - 1):
        if not (ref[i].bond_right or i in pair_min):
            return False, 'the chain is not connected at the truncation'
    for nd in ref:
        if nd.tensor.shape[2] > 1 and not nd.has_I:
            return False, 'inner index without the I flag'
    dims = [tuple(nd.tensor.shape[i] for i in (0, 2, 3)) for nd in ref]
    kappa = cs[0].kappa if (not cs[0].ideal and cs[0].kappa is not None) else None
    out, why = _lib.layer_dims(dims, layer.kraus1.shape[2], layer.parity, layer.gates2.shape[-1], cs[0].chi, kappa)
    if out is None:
        return False, why
    return True, ''


def per_gate_layer(chain: Sequence[torch.Tensor], layer: FusedLayer, b: int, chi: Optional[int],
                   kappa: Optional[int]):
    """The plain per-circuit path of one fused layer for chain ``b``: the same operands run one
    gate at a time through the per-gate entry points in the reference's order (Circuit.py:469-489):
    mpdo_apply_1q on every site, mpdo_apply_2q on every pair (Circuit.py:74-136, absolute error
    GLOBAL_MINIMUM), then bondTruncate (TNNOptimizer.py:72-84) and svdKappa_left2right
    (TNNOptimizer.py:164-197). ``chain``: the chain's (l, 2, a, r) site tensors (not modified).
    Returns (sites, gate-time split weights by pair min site, chi-step weights in sweep order,
    kappa weights by site)."""
    from . import TNNOptimizer as _opt
    from .circuit import GLOBAL_MINIMUM
    n = len(chain)
    nodes = [SiteNode(t.contiguous().clone(), i, has_I=True, bond_left=i > 0, bond_right=i < n - 1)
             for i, t in enumerate(chain)]
    ts = _lib.apply_1q([nd.tensor for nd in nodes], [layer.kraus1[b, q].permute(1, 2, 0) for q in range(n)])
    for nd, t in zip(nodes, ts):
        nd.tensor = t
    gate = {}
    for k, p in enumerate(range(layer.parity, n - 1, 2)):
        g = layer.gates2[k] if layer.noisy2 else layer.gates2[k][..., 0]
        x, y, d = _lib.apply_2q(nodes[p].tensor, nodes[p + 1].tensor, g, GLOBAL_MINIMUM)
        nodes[p].tensor, nodes[p + 1].tensor = x, y
        gate[p] = d
    d_chi = _opt.bondTruncate(nodes, chi, return_discarded=True)
    d_kap = _opt.svdKappa_left2right(nodes, kappa, return_discarded=True) if kappa is not None else []
    return [nd.tensor.contiguous() for nd in nodes], gate, list(d_chi), dict(d_kap)
