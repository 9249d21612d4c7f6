"""Density-matrix observables with the reference's names (dmOperations.py:11-201).

``dmNodes`` is the ``DMNodes`` list that ``TensorCircuit.cal_dmNodes`` returns: 2n entries (the
ket sites, then their conjugate views) plus the set of traced qubits, the fixed-layout stand-in
for the ket/bra network of Circuit.py:226-245. Every value is the unnormalised Re Tr(O rho~)
of the reference; the contractions run in libmpdo_gfx950.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib
from .node import SiteNode

__all__ = ['PAULI_DICT', 'DMNodes', 'reduce_dmNodes', 'trace_rho', 'trace_rho_rho', 'trace_rho2',
           'trace_composited_rho', 'trace_composited_rho2', 'expect', 'pauli_expect']

PAULI_DICT = {
    0: torch.tensor([[0, 1], [1, 0]]),
    1: torch.tensor([[0, -1j], [1j, 0]]),
    2: torch.tensor([[1, 0], [0, -1]]),
}


class ConjView:
    """The bra copy of one site (the reference's ``con_`` node): a view, never materialised."""
    __slots__ = ('site',)

    def __init__(self, site: SiteNode):
        self.site = site

    @property
    def tensor(self):
        return self.site.tensor.conj()

    @property
    def name(self):
        return f'con_{self.site.name}'


class DMNodes(list):
    """[ket_0 .. ket_{n-1}, bra_0 .. bra_{n-1}] with the shared set of traced qubits."""

    def __init__(self, sites: Sequence[SiteNode], reduced: Sequence[int] = ()):
        super().__init__(list(sites) + [ConjView(s) for s in sites])
        self.reduced = set(int(i) for i in reduced)

    @property
    def qnumber(self) -> int:
        return len(self) // 2

    @property
    def sites(self) -> List[SiteNode]:
        return list(self[:self.qnumber])

    def tensors(self) -> List[torch.Tensor]:
        return [s.tensor.contiguous() for s in self.sites]


def reduce_dmNodes(qubits_nodes, conj_qubits_nodes=None, residual_index=None, reduced_index=None):
    """dmOperations.py:18-37: trace out ``reduced_index`` (int or list of int)."""
    if reduced_index is None:
        return None
    n = len(qubits_nodes) if not isinstance(qubits_nodes, DMNodes) else qubits_nodes.qnumber
    if isinstance(reduced_index, int):
        reduced_index = [reduced_index]
    if not isinstance(reduced_index, list):
        raise TypeError('reduced_index should be int or list[int]')
    if reduced_index and max(reduced_index) >= n:
        raise ValueError(f'Reduced index should not be larger than the qubit number. {max(reduced_index)}-{n}')
    target = qubits_nodes if isinstance(qubits_nodes, DMNodes) else conj_qubits_nodes
    if isinstance(target, DMNodes):
        target.reduced.update(reduced_index)
    return None


def trace_rho(dmNodes: DMNodes) -> torch.Tensor:
    """dmOperations.py:40-48: Re Tr rho~ (unnormalised)."""
    n = dmNodes.qnumber
    v = _lib.pauli_expect(dmNodes.tensors(), [[-1] * n])[0]
    return torch.tensor(v, dtype=torch.float64)


def _dense_rho(dm: DMNodes, keep: Sequence[bool]) -> torch.Tensor:
    return _lib.reduced_dm(dm.tensors(), keep)


def trace_rho_rho(dmNodes_0: DMNodes, dmNodes_1: Optional[DMNodes] = None) -> torch.Tensor:
    """dmOperations.py:51-66: Re Tr(rho0 rho1) over the qubits neither state traces (a transfer
    sweep on the device, any number of qubits). Both states must trace the same qubits: the
    reference leaves the other state's physical edges open there and returns no scalar."""
    dm1 = dmNodes_0 if dmNodes_1 is None else dmNodes_1
    n = dmNodes_0.qnumber
    if dm1.qnumber != n:
        raise ValueError('Density matrices must have the same number of nodes.')
    if dmNodes_0.reduced != dm1.reduced:
        raise ValueError('both density matrices must trace the same qubits')
    keep = None if not dmNodes_0.reduced else [i not in dmNodes_0.reduced for i in range(n)]
    return torch.tensor(_lib.trace_rho_rho(dmNodes_0.tensors(), dm1.tensors(), keep), dtype=torch.float64)


def trace_rho2(dmNodes: DMNodes) -> torch.Tensor:
    """dmOperations.py:69-70 (purity Tr rho~^2)."""
    return trace_rho_rho(dmNodes)


def trace_composited_rho(*dmNodes: DMNodes) -> torch.Tensor:
    """dmOperations.py:73-88: the sum of trace_rho_rho(rho_i) over the slices (the reference sums
    the purities of the slices, whatever its docstring says)."""
    return sum((trace_rho_rho(d) for d in dmNodes), torch.tensor(0.0, dtype=torch.float64))


def trace_composited_rho2(*dmNodes: DMNodes) -> torch.Tensor:
    """dmOperations.py:90-132: Tr(rho^2) of rho = (rho_0 + ... + rho_{S-1}) / S as
    (sum_i Tr rho_i^2 + 2 sum_{i<j} Tr rho_i rho_j) / S^2. The reference takes the same sum from a
    dense 2^k matrix when at most 12 qubits stay untraced; the value is the same."""
    S = len(dmNodes)
    if S == 0:
        raise ValueError('no density matrices given')
    kept = {d.qnumber - len(d.reduced) for d in dmNodes}
    if len(kept) != 1:
        raise ValueError('Density matrices should have same number of un-traced nodes.')
    diag = sum((trace_rho2(d) for d in dmNodes), torch.tensor(0.0, dtype=torch.float64))
    cross = sum((trace_rho_rho(dmNodes[i], dmNodes[j]) for i in range(S) for j in range(i + 1, S)),
                torch.tensor(0.0, dtype=torch.float64))
    return (diag + 2 * cross) / (S ** 2)


def expect(dmNodes: DMNodes, observables: Union[torch.Tensor, List[torch.Tensor]],
           oqs: Union[int, List]) -> List[torch.Tensor]:
    """dmOperations.py:135-171: Re Tr(O rho~) for each observable O on the qubits ``oq`` (first
    listed = most significant), the other qubits traced."""
    n = dmNodes.qnumber
    oqs = [oqs] if isinstance(oqs, int) else oqs
    observables = [observables] if isinstance(observables, torch.Tensor) else observables
    values = []
    for j, (obs, oq) in enumerate(zip(observables, oqs)):
        oq = [oq] if isinstance(oq, int) else list(oq)
        k = len(oq)
        try:
            o = torch.as_tensor(obs).reshape([2] * 2 * k)
        except RuntimeError:
            raise ValueError(f'Shape of the No.{j} obs is not valid, which is: {tuple(obs.shape)}.')
        if k > n:
            raise ValueError(f'Dim of No.{j} - oqs: {k} or obs: {k} exceeds the system size.')
        if any(q in dmNodes.reduced for q in oq) or len(set(oq)) != k:
            raise ValueError(f'No.{j}: observable qubits {oq} are traced or repeated')
        srt = sorted(oq)
        perm = [oq.index(q) for q in srt]
        o = o.permute(perm + [p + k for p in perm]).reshape(2 ** k, 2 ** k).to(torch.complex128)
        rho = _dense_rho(dmNodes, [i in oq for i in range(n)]).to(torch.complex128).cpu()
        values.append(torch.trace(o @ rho).real)
    return values


def pauli_expect(dmNodes: DMNodes, observables: Union[int, List[int], Tuple], oqs: Union[int, List[int], Tuple]):
    """dmOperations.py:174-201: Re Tr((P_1 x ... x P_k) rho~), 0 = X, 1 = Y, 2 = Z."""
    n = dmNodes.qnumber
    oqs = [oqs] if isinstance(oqs, int) else list(oqs)
    observables = [observables] if isinstance(observables, int) else list(observables)
    codes = [-1] * n
    for ob, oq in zip(observables, oqs):
        if ob not in PAULI_DICT:
            raise ValueError(f'Pauli index {ob} not in {{0, 1, 2}}')
        codes[oq] = int(ob)
    return torch.tensor(_lib.pauli_expect(dmNodes.tensors(), [codes])[0], dtype=torch.float64)


def pauli_expect_batch(dmNodes: DMNodes, strings: Sequence[Sequence[int]]) -> torch.Tensor:
    """Extension: many operator strings at once; each string holds one code per qubit
    (-1 identity, 0/1/2 Pauli X/Y/Z, 3 = |0><0|, 4 = |1><1| for bitstring probabilities)."""
    return torch.tensor(_lib.pauli_expect(dmNodes.tensors(), [list(s) for s in strings]), dtype=torch.float64)
