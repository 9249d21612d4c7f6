"""Truncation entry points with the reference's names and signatures (TNNOptimizer.py:51-197).

Each function takes the state as a list of ``SiteNode`` and replaces its elements in place
(``TypeError`` when it is not a list, TNNOptimizer.py:64-69), like the reference mutates its list
of tensornetwork nodes. The arithmetic runs in libmpdo_gfx950 (include/mpdo.h). The discarded
weight sum(s_rest^2) of every cut, which the reference computes and drops
(TNNOptimizer.py:129,186), is returned when ``return_discarded=True``.
"""
from __future__ import annotations

from typing import List, Optional

from . import _lib
from .node import SiteNode

__all__ = ['qr_left2right', 'svd_left2right', 'svd_right2left', 'svdKappa_left2right', 'bondTruncate',
           'checkConnectivity']


def _validate_qubit_list(qubits, func_name: str):
    if not isinstance(qubits, list):
        raise TypeError(f'`{func_name}` expects a list of qubit nodes, got {type(qubits).__name__}')


def _tensors(qubits: List[SiteNode]):
    ts = []
    for q in qubits:
        t = q.tensor
        if not t.is_contiguous():
            t = t.contiguous()
            q.tensor = t
        ts.append(t)
    return ts


def _require_bonds(qubits: List[SiteNode], func_name: str):
    for i in range(len(qubits) - 1):
        if not qubits[i].bond_right:
            raise KeyError(f'`{func_name}`: edge bond_{i}_{i + 1} does not exist')


def checkConnectivity(_qubits: List[SiteNode]) -> bool:
    """TNNOptimizer.py:51-61: a chain is connected when every bond_{i}_{i+1} exists."""
    if len(_qubits) <= 1:
        return True
    return all(q.bond_right for q in _qubits[:-1])


def qr_left2right(_qubits: List[SiteNode]) -> None:
    """TNNOptimizer.py:87-108 / decompositions.py:149-195."""
    _validate_qubit_list(_qubits, 'qr_left2right')
    if len(_qubits) == 0:
        return
    _require_bonds(_qubits, 'qr_left2right')
    outs = _lib.qr_left2right(_tensors(_qubits))
    for q, t in zip(_qubits, outs):
        q.tensor = t


def svd_right2left(_qubits: List[SiteNode], max_singular_values: Optional[int] = None,
                   max_truncation_err: Optional[float] = None, return_discarded: bool = False,
                   _left_canonical: bool = False):
    """TNNOptimizer.py:111-134 (tn.split_node with relative=True: U*sqrt(S) | sqrt(S)*V^H)."""
    _validate_qubit_list(_qubits, 'svd_right2left')
    if len(_qubits) <= 1:
        return [] if return_discarded else None
    _require_bonds(_qubits, 'svd_right2left')
    outs, disc = _lib.svd_sweep(_tensors(_qubits), 0, max_singular_values, max_truncation_err, _left_canonical)
    for q, t in zip(_qubits, outs):
        q.tensor = t
    return disc if return_discarded else None


def svd_left2right(_qubits: List[SiteNode], max_singular_values: int, max_truncation_err: Optional[float] = None,
                   return_discarded: bool = False):
    """TNNOptimizer.py:137-161."""
    _validate_qubit_list(_qubits, 'svd_left2right')
    if len(_qubits) <= 1:
        return [] if return_discarded else None
    _require_bonds(_qubits, 'svd_left2right')
    outs, disc = _lib.svd_sweep(_tensors(_qubits), 1, max_singular_values, max_truncation_err, False)
    for q, t in zip(_qubits, outs):
        q.tensor = t
    return disc if return_discarded else None


def bondTruncate(_qubits: List[SiteNode], max_singular_values: Optional[int] = None,
                 max_truncation_err: Optional[float] = None, regularization: bool = False,
                 return_discarded: bool = False):
    """TNNOptimizer.py:72-84: QR sweep left to right, then the truncating SVD sweep right to
    left. After the QR sweep sites 0..n-2 are left-orthonormal, so each SVD step factors only
    the right tensor of its pair."""
    if max_singular_values is None and max_truncation_err is None and not regularization:
        return [] if return_discarded else None
    qr_left2right(_qubits)
    return svd_right2left(_qubits, max_singular_values, max_truncation_err, return_discarded, _left_canonical=True)


def svdKappa_left2right(_qubits: List[SiteNode], max_singular_values: Optional[int] = None,
                        max_truncation_err: Optional[float] = None, return_discarded: bool = False):
    """TNNOptimizer.py:164-197: inner-index SVD of every site carrying I_i with dim(I_i) > kappa,
    T <- U*S. Sites are independent, so they are processed in one call."""
    if max_singular_values is None and max_truncation_err is None:
        return [] if return_discarded else None
    _validate_qubit_list(_qubits, 'svdKappa_left2right')
    idx = [i for i, q in enumerate(_qubits)
           if q.has_I and (max_singular_values is None or q.tensor.shape[2] > max_singular_values)]
    if not idx:
        return [] if return_discarded else None
    ts = _tensors([_qubits[i] for i in idx])
    outs, disc = _lib.svd_kappa(ts, max_singular_values, max_truncation_err)
    for i, t in zip(idx, outs):
        _qubits[i].tensor = t
    return list(zip(idx, disc)) if return_discarded else None


def cal_entropy(qubits):
    """TNNOptimizer.py:200-211 (a stub in the reference)."""
    raise NotImplementedError
