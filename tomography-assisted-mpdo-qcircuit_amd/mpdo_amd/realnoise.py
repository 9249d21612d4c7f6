"""Tomographed two-qubit noise: chi matrix -> Kraus tensor (the reference's ``RealNoise.py``).

Host-side, once per circuit. The bundled ``chi/czDefault.mat`` and ``chi/ideal_cz.mat`` are the
reference's experimental chi matrices (16 x 16 complex, key 'exp').
"""
from __future__ import annotations

import itertools
import math
import os
from typing import Dict, List, Optional, Union

import numpy as np
import torch

__all__ = ['readExpChi', 'noisyTensor', 'czExp_channel', 'cpExp_channel', 'gates_list', 'name2matrix']

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'chi')


def gates_list(N: int, basis_gates: Optional[List[str]] = None) -> List[str]:
    """Tools.py:466-485: all N-fold products of the basis names, first qubit major."""
    basis_gates = basis_gates or ['I', 'X', 'Y', 'Z']
    return [''.join(g) for g in itertools.product(basis_gates, repeat=N)]


def name2matrix(operation_name: str) -> np.ndarray:
    """Tools.py:488-509; the letter 'Y' stands for -1j*Y = [[0,-1],[1,0]] (Tools.py:504)."""
    ops = {'I': np.eye(2, dtype=np.complex128),
           'X': np.array([[0, 1], [1, 0]], dtype=np.complex128),
           'Y': np.array([[0, -1], [1, 0]], dtype=np.complex128),
           'Z': np.array([[1, 0], [0, -1]], dtype=np.complex128)}
    m = np.ones((1, 1), dtype=np.complex128)
    for ch in operation_name:
        m = np.kron(m, ops[ch])
    return m


def readExpChi(filename: Optional[str] = None) -> np.ndarray:
    """RealNoise.py:22-32."""
    if filename is None:
        raise FileNotFoundError('No file found.')
    if '.mat' in filename:
        from scipy.io import loadmat
        return loadmat(filename)['exp']
    if '.npz' in filename:
        return np.load(filename)['chi']
    raise TypeError('Current file-type is not supported.')


def noisyTensor(chi: np.ndarray, gate_factor: Optional[Dict] = None) -> List[np.ndarray]:
    """RealNoise.py:35-161: E_i = sqrt(d_i) sum_j U_ji P_j over the eigenpairs of chi taken in
    descending (numpy complex) order, as many as there are eigenvalues with |d| > 1e-12, and
    coefficients below 1e-12 dropped. Returns the list of (2,)*2N operator tensors."""
    tol = 1e-12
    if gate_factor is None:
        N = int(round(math.log(chi.shape[0], 4)))
        basis = gates_list(N)
        s, u = np.linalg.eig(chi)
        order = s.argsort()[::-1]
        s, u = s[order], u[:, order]
        count = int(np.sum(np.abs(s) > tol))
        gate_factor = []
        for i in range(count):
            e = u.T[i] * np.sqrt(s[i])
            gate_factor.append({basis[j]: e[j] for j in range(len(basis)) if np.abs(e[j]) >= tol})
    elif not any(isinstance(v, dict) for v in gate_factor.values()):
        gate_factor = [gate_factor]
    else:
        gate_factor = list(gate_factor.values())
    out = []
    for terms in gate_factor:
        N = len(next(iter(terms))) if terms else 2
        m = np.zeros((2 ** N, 2 ** N), dtype=np.complex128)
        for name, coef in terms.items():
            m = m + coef * name2matrix(name)
        out.append(m.reshape([2] * (2 * N)))
    return out


def _exp_channel(filename: Optional[str], default: str, dtype) -> torch.Tensor:
    if filename is None:
        filename = os.path.join(_DATA, default)
    st = np.stack(noisyTensor(readExpChi(filename)))            # (K, 2, 2, 2, 2)
    return torch.from_numpy(np.moveaxis(st, 0, -1).copy()).to(dtype)   # RealNoise.py:169 'ijlmn -> jlmni'


def czExp_channel(filename: Optional[str] = None, dtype=torch.complex128, device: Union[int, str] = 'cpu'):
    """RealNoise.py:164-170 (the bundled czDefault.mat when no file is given)."""
    return _exp_channel(filename, 'czDefault.mat', dtype).to(device)


def cpExp_channel(filename: Optional[str] = None, dtype=torch.complex128, device: Union[int, str] = 'cpu'):
    """RealNoise.py:173-179. The reference ships no cpDefault.mat; a file must be given."""
    if filename is None:
        raise FileNotFoundError('no default CP chi matrix is bundled (the reference has none either)')
    return _exp_channel(filename, 'cpDefault.mat', dtype).to(device)
