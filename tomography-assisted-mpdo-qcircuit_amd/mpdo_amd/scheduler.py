"""Many independent circuits of uneven qubit counts, depths, bond caps and noise on one GPU.

The reference evolves one circuit at a time (Circuit.py:469-491). Here a list of circuits - any mix
of qubit counts, depths, chi, kappa and noise models - is advanced together:

* circuits of one structure (same qubit count, truncation settings, noise model and layer list,
  parameters free to differ) form a family whose brickwork circuit-layers run as one
  ``mpdo_layer_batched`` call over all its chains (``batched.evolve_batched``);
* the families are dealt, largest estimated cost first, to a pool of host threads, each driving its
  own HIP stream (default priority) with its own workspace (bound from torch's caching allocator):
  the launches of different circuits overlap on the device, and the host-side bookkeeping and the
  kept-count readbacks of one circuit no longer idle the GPU for the others.

Results are those of ``TensorCircuit.evolve`` per circuit (every library call of a circuit is
ordered on its stream; nothing is shared between streams but the read-only inputs).
"""
from __future__ import annotations

import threading
import time
from typing import Dict, List, Optional, Sequence

import torch

from . import _lib, batched
from .circuit import TensorCircuit
from .node import SiteNode

__all__ = ['evolve_many', 'families', 'layer_count']


def layer_count(c: TensorCircuit) -> int:
    """Circuit-layers of a circuit: truncations in its layer list (at least one)."""
    return max(1, sum(1 for g in c.layers if 'truncate' in g.name.lower()))


def _cost(c: TensorCircuit) -> float:
    chi = c.chi or 4 ** (c.qnumber // 2)
    kap = c.kappa or 4
    return float(layer_count(c)) * c.qnumber * (2.0 * kap * chi) ** 3


def families(circuits: Sequence[TensorCircuit]) -> List[List[int]]:
    """Indices of the circuits grouped by structure (batched._same_structure): each group can run its
    fused layers as one batched call."""
    groups: List[List[int]] = []
    for i, c in enumerate(circuits):
        for g in groups:
            if batched._same_structure([circuits[g[0]], c]) is None:
                g.append(i)
                break
        else:
            groups.append([i])
    return groups


def evolve_many(circuits: Sequence[TensorCircuit], states: Sequence[List[SiteNode]], streams: int = 4,
                fuse_layers: bool = True, stats: Optional[Dict] = None) -> None:
    """Evolve ``states[i]`` under ``circuits[i]`` for every i, in place, with up to ``streams``
    families in flight at once. ``stats`` receives the family count, the streams used, the
    circuit-layers run and the wall time."""
    cs = list(circuits)
    if len(cs) != len(states):
        raise ValueError('one state per circuit is required')
    if not cs:
        return
    dev = torch.device(cs[0].device)
    if dev.type != 'cuda':
        raise RuntimeError('evolve_many runs on a ROCm GPU; there is no CPU path')
    groups = sorted(families(cs), key=lambda g: -sum(_cost(cs[i]) for i in g))
    nstreams = max(1, min(int(streams), len(groups)))
    pool = [torch.cuda.Stream(device=dev) for _ in range(nstreams)]
    torch.cuda.synchronize(dev)          # every input is ready before other streams read it
    lock = threading.Lock()
    nxt = [0]
    errors: List[BaseException] = []
    done = {'fused_layers': 0, 'per_gate_layers': 0}

    def worker(stream: torch.cuda.Stream):
        try:
            run(stream)
        finally:   # the stream's workspace goes back to torch's allocator
            with torch.cuda.device(dev), torch.cuda.stream(stream):
                stream.synchronize()
                _lib.release_stream_workspace(dev)

    def run(stream: torch.cuda.Stream):
        while True:
            with lock:
                if nxt[0] >= len(groups) or errors:
                    return
                g = groups[nxt[0]]
                nxt[0] += 1
            try:
                st: Dict = {}
                with torch.cuda.device(dev), torch.cuda.stream(stream):
                    batched.evolve_batched([cs[i] for i in g], [states[i] for i in g], fuse_layers=fuse_layers,
                                           stats=st)
                stream.synchronize()
                with lock:
                    for k in done:
                        done[k] += st.get(k, 0)
            except BaseException as e:   # re-raised in the caller's thread
                with lock:
                    errors.append(e)
                return

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, args=(s,), daemon=True) for s in pool]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize(dev)
    if errors:
        raise errors[0]
    if stats is not None:
        stats.update(done)
        stats['families'] = len(groups)
        stats['streams'] = nstreams
        stats['circuit_layers'] = sum(layer_count(c) for c in cs)
        stats['seconds'] = time.perf_counter() - t0
