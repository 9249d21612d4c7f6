"""Initial states and host utilities with the reference's names (Tools.py).

``create_ket*Series`` return the fixed-layout site list the engine evolves (one (1, 2, 1, 1)
tensor per qubit, no bonds, no inner index: Tools.py:110-198). Plotting (Tools.py:276-449) is not
part of this package.
"""
from __future__ import annotations

import warnings
from collections import Counter
from typing import Dict, List, Optional, Union

import numpy as np
import torch

from .node import SiteNode
from .realnoise import gates_list, name2matrix

__all__ = ['create_ket0Series', 'create_ket1Series', 'create_ketPlusSeries', 'create_ketMinusSeries',
           'create_ketRandomSeries', 'create_ketHadamardSeries', 'select_device', 'count_item', 'cal_fidelity',
           'sqrt_matrix', 'random_measurementScheme', 'gates_list', 'name2matrix']

PUBLIC_SQRT2 = 1 / np.sqrt(2)


def _series(vec, qnumber: int, dtype, device) -> List[SiteNode]:
    assert qnumber > 0
    v = torch.as_tensor(vec, dtype=dtype).to(device)
    return [SiteNode(v.reshape(1, 2, 1, 1).clone(), i, f'qubit_{i}') for i in range(qnumber)]


def create_ket0Series(qnumber: int, dtype=torch.complex64, device: Union[str, int] = 'cpu') -> List[SiteNode]:
    return _series([1.0 + 0j, 0.0], qnumber, dtype, device)


def create_ket1Series(qnumber: int, dtype=torch.complex64, device: Union[str, int] = 'cpu') -> List[SiteNode]:
    return _series([0.0, 1.0 + 0j], qnumber, dtype, device)


def create_ketHadamardSeries(qnumber: int, dtype=torch.complex64, device='cpu') -> List[SiteNode]:
    return _series([PUBLIC_SQRT2 + 0j, PUBLIC_SQRT2], qnumber, dtype, device)


def create_ketPlusSeries(qnumber: int, dtype=torch.complex64, device='cpu') -> List[SiteNode]:
    return _series([PUBLIC_SQRT2 + 0j, PUBLIC_SQRT2], qnumber, dtype, device)


def create_ketMinusSeries(qnumber: int, dtype=torch.complex64, device='cpu') -> List[SiteNode]:
    return _series([PUBLIC_SQRT2 + 0j, -PUBLIC_SQRT2], qnumber, dtype, device)


def create_ketRandomSeries(qnumber: int, tensor: torch.Tensor, dtype=torch.complex64, device='cpu') -> List[SiteNode]:
    return _series(tensor.detach().cpu().reshape(-1), qnumber, dtype, device)


def select_device(device: Optional[Union[str, int]] = None):
    """Tools.py:452-463."""
    if isinstance(device, str):
        return device
    if torch.cuda.is_available():
        return 'cuda:0' if device is None else f'cuda:{device}'
    warnings.warn('CUDA is not available, use CPU instead.')
    return 'cpu'


def count_item(data):
    """Tools.py:554-570."""
    def sort_key(x):
        if isinstance(x[0], tuple):
            return int(''.join(str(int(v)) for v in x[0]))
        if isinstance(x[0], str):
            return x[0]
        return int(x[0])
    counted = dict(Counter([tuple(item) if isinstance(item, list) else item for item in data]))
    return dict(sorted(counted.items(), key=sort_key))


def sqrt_matrix(density_matrix: torch.Tensor) -> torch.Tensor:
    """Tools.py:512-525 (host utility)."""
    evs, vecs = torch.linalg.eigh(density_matrix)
    evs = torch.where(evs > 1e-10, evs, torch.zeros_like(evs))
    return vecs @ torch.diag(torch.sqrt(evs)).to(vecs.dtype) @ vecs.conj().transpose(-2, -1)


def cal_fidelity(rho: torch.Tensor, sigma: torch.Tensor) -> torch.Tensor:
    """Tools.py:528-551 (host utility)."""
    if rho.shape != sigma.shape:
        raise ValueError('The shape of rho and sigma should be equal.')
    if rho.shape[0] != rho.shape[1] or rho.shape[0] != sigma.shape[1]:
        raise ValueError('The shape of rho and sigma should be square.')
    sr = sqrt_matrix(rho)
    evs = torch.linalg.eigvalsh(sr @ sigma @ sr)
    evs = torch.where(evs > 1e-12, evs, torch.zeros_like(evs))
    return torch.sum(torch.sqrt(evs)) ** 2


def random_measurementScheme(qnumber: int, amount: int) -> List[List[int]]:
    """Tools.py:573-584."""
    return np.random.randint(0, 3, size=(amount, qnumber)).tolist()
