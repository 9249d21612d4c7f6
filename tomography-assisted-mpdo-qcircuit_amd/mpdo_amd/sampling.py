"""Shot sampling of an evolved MPDO (Circuit.py:297-467) on the device.

The reference recontracts the whole density-matrix network for every (group, qubit) with the
projectors of the group's history. Here ``mpdo_env_right`` builds the traced tail environments
once and the sweep keeps one left environment per group: a sampled qubit costs one
``mpdo_sample_step`` launch sequence for all groups at once (both outcomes), after which each
child group inherits the environment of its outcome. The host keeps the reference's group
bookkeeping and its torch draws (same calls, same order), so a seeded run reproduces the
reference's bitstrings for the same probabilities.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import _lib

GLOBAL_MINIMUM = 2.718281828459045 * 1e-8     # Circuit.py:22


def normalise_prb(P0: float, P1: float, real_dtype=torch.float32) -> float:
    """Circuit.py:328-332: real parts plus GLOBAL_MINIMUM in the state's real dtype, an error on
    a negative entry, the normalised probability of outcome 1."""
    probs = torch.tensor([float(P0), float(P1)], dtype=real_dtype) + GLOBAL_MINIMUM
    if (probs < 0).sum() > 0:
        raise RuntimeError(f"State is illegal, and your probability distribution is {probs}.")
    return (probs / probs.sum())[-1].item()


class ConditionalSweep:
    """Left-to-right conditional marginals of the sites ``tensors`` for the qubits ``order``
    (increasing; every other qubit traced)."""

    def __init__(self, tensors: Sequence[torch.Tensor], order: Sequence[int]):
        self.sites = [t.contiguous() for t in tensors]
        self.n = len(self.sites)
        self.order = list(order)
        if any(b <= a for a, b in zip(self.order, self.order[1:])):
            raise ValueError('sampled qubits must be in increasing order')
        self.R = _lib.env_right(self.sites)
        self.device = self.sites[0].device
        self._path: List[Tuple[Tuple[int, ...], torch.Tensor]] = []
        self._prob_cache: Dict[Tuple[int, ...], Tuple[float, float]] = {}

    def start(self) -> torch.Tensor:
        return torch.ones(1, 1, 1, dtype=torch.complex128, device=self.device)

    def trace_through(self, envs: torch.Tensor, sel: Optional[Sequence[int]], begin: int, end: int):
        """Apply the traced sites begin..end-1 to the selected environments."""
        for i in range(begin, end):
            envs, _ = _lib.sample_step(self.sites[i], envs, sel, sampled=False)
            sel = None
        return envs, sel

    def project(self, envs: torch.Tensor, sel: Optional[Sequence[int]], q: int):
        """Both outcomes at sampled qubit q: environments (G, 2, r, r) and probabilities (G, 2)."""
        return _lib.sample_step(self.sites[q], envs, sel, sampled=True, R_next=self.R[q + 1])

    def probs_of(self, history: Sequence[int]) -> Tuple[float, float]:
        """Unnormalised (P(history, 0), P(history, 1)) for the qubit ``order[len(history)]``,
        reusing the environments of the longest prefix of the last history asked for."""
        key = tuple(int(h) for h in history)
        if key in self._prob_cache:
            return self._prob_cache[key]
        k = 0
        while k < len(self._path) and k < len(key) and self._path[k][0] == key[:k + 1]:
            k += 1
        del self._path[k:]
        if k == 0:
            envs, pos = self.start(), 0
        else:
            envs, pos = self._path[k - 1][1], self.order[k - 1] + 1
        for j in range(k, len(key)):
            q = self.order[j]
            envs, _ = self.trace_through(envs, None, pos, q)
            out, _ = self.project(envs, None, q)
            envs = out[:, key[j]].contiguous()
            self._path.append((key[:j + 1], envs))
            pos = q + 1
        q = self.order[len(key)]
        envs, _ = self.trace_through(envs, None, pos, q)
        _, P = self.project(envs, None, q)
        res = (float(P[0, 0]), float(P[0, 1]))
        self._prob_cache[key] = res
        return res


def batch_sample(sweep: ConditionalSweep, shots: int, bool_result: bool = False,
                 trace: Optional[list] = None) -> list:
    """Circuit.py:344-391 (_conditional_batch_sample): groups of shots sharing a history, one
    Bernoulli count per group unless p1 is within GLOBAL_MINIMUM of 0 or 1, ones first."""
    order = sweep.order
    L = len(order)
    seq = torch.zeros((shots, L), dtype=torch.bool if bool_result else torch.int)
    val = True if bool_result else 1
    groups = [([], 0, shots)]
    envs, sel, pos = sweep.start(), None, 0
    for j, q in enumerate(order):
        envs, sel = sweep.trace_through(envs, sel, pos, q)
        out, P = sweep.project(envs, sel, q)
        new, new_sel = [], []
        for g, (hist, start, length) in enumerate(groups):
            p1 = normalise_prb(P[g, 0], P[g, 1])
            if trace is not None:
                trace.append((tuple(hist), p1))
            if p1 < GLOBAL_MINIMUM:
                k1, k0 = 0, length
            elif p1 > 1 - GLOBAL_MINIMUM:
                k1, k0 = length, 0
            else:
                k1 = int(torch.bernoulli(torch.full((length,), p1)).sum())
                k0 = length - k1
            seq[start:start + k1, j] = val
            if j < L - 1:
                if k1 > 0:
                    new.append((hist + [1], start, k1))
                    new_sel.append(2 * g + 1)
                if k0 > 0:
                    new.append((hist + [0], start + k1, k0))
                    new_sel.append(2 * g)
        groups = new
        envs = out.view(-1, out.shape[2], out.shape[3])
        sel, pos = new_sel, q + 1
    return seq.tolist()


def sequential_sample(sweep: ConditionalSweep, shots: int) -> list:
    """Circuit.py:334-342 (_conditional_sample): shot by shot, qubit by qubit, one two-outcome
    multinomial draw; probabilities of a history already seen are reused."""
    out = []
    for _ in range(shots):
        ch = [-1] * len(sweep.order)
        for j in range(len(sweep.order)):
            p = normalise_prb(*sweep.probs_of(ch[:j]))
            ch[j] = torch.multinomial(torch.tensor([1 - p, p]), num_samples=1).item()
        out.append(ch)
    return out
