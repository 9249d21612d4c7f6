"""Gate objects of the circuit builder (the reference's ``QuantumGates`` package).

Each gate exposes the reference's interface (``name, tensor, rank, dimension, single,
variational, ideal, para``; QuantumGates/AbstractGate.py:31-99). Tensors are host-side
constants (2x2 for one qubit as [out, in]; 2x2x2x2 as [o0, o1, i0, i1] for two qubits, the
first listed qubit the most significant bit) that the engine hands to the device kernels.
"""
from __future__ import annotations

import math
import warnings
from abc import ABC
from typing import Optional, Sequence

import numpy as np
import torch
from torch import nn

__all__ = ['QuantumGate', 'Truncate', 'Barrier', 'IGate', 'HGate', 'XGate', 'YGate', 'ZGate', 'RXGate', 'RYGate',
           'RZGate', 'SGate', 'SDGGate', 'TGate', 'PGate', 'U1Gate', 'U2Gate', 'U3Gate', 'ArbSingleGate',
           'MeasureX', 'MeasureY', 'MeasureZ', 'Reset0', 'Reset1', 'CXGate', 'CNOTGate', 'CYGate', 'CZGate',
           'IIGate', 'SWAPGate', 'ISWAPGate', 'CPGate', 'RXXGate', 'RYYGate', 'RZZGate', 'PSWAPGate',
           'XXPlusYYGate', 'ArbDoubleGate', 'CZEXPGate', 'CPEXPGate']

_S2 = 1.0 / math.sqrt(2.0)


class Truncate(nn.Module):
    """Truncation layer marker (AbstractGate.py:13-19)."""

    def __init__(self):
        super().__init__()
        self.name = 'truncate'

    def forward(self, *args, **kwargs):
        pass


class Barrier(nn.Module):
    """Barrier layer marker (AbstractGate.py:22-28)."""

    def __init__(self):
        super().__init__()
        self.name = 'barrier'

    def forward(self, *args, **kwargs):
        pass


def _val(p) -> float:
    if isinstance(p, torch.Tensor):
        return float(p.detach().real.item()) if p.is_complex() else float(p.detach().item())
    return float(p)


class QuantumGate(ABC, nn.Module):
    """Base class (AbstractGate.py:31-99). Subclasses set NAME/SINGLE/VARIATIONAL and _matrix."""
    NAME = ''
    SINGLE = True
    VARIATIONAL = False

    def __init__(self, ideal: Optional[bool] = True, dtype=torch.complex64, device='cpu', *params):
        super().__init__()
        self._ideal = ideal
        self.dtype, self.device = dtype, device
        self._params = [self._check_param(p) for p in params]
        self.para = (self._params[0] if len(self._params) == 1 else list(self._params)) if self._params else None

    @staticmethod
    def _check_param(p):
        """AbstractGate.py:42-51: tensors and floats are accepted, anything else is an error."""
        if isinstance(p, torch.Tensor) or isinstance(p, (float, int)):
            return p
        raise ValueError(f'Invalid type for gate parameter: {type(p)}')

    @property
    def name(self):
        return self.NAME

    def _matrix(self, p: Sequence[float]) -> np.ndarray:
        raise NotImplementedError

    @property
    def tensor(self) -> torch.Tensor:
        m = self._matrix([_val(x) for x in self._params])
        t = torch.from_numpy(np.asarray(m, dtype=np.complex128))
        return t if self.SINGLE else t.reshape(2, 2, 2, 2)

    @property
    def rank(self):
        return 2 if self.SINGLE else 4

    @property
    def dimension(self):
        return [2, 2] if self.SINGLE else [[2, 2], [2, 2]]

    @property
    def single(self) -> bool:
        return self.SINGLE

    @property
    def variational(self) -> bool:
        return self.VARIATIONAL

    @property
    def para(self):
        return self._para

    @para.setter
    def para(self, para):
        self._para = para

    @property
    def ideal(self) -> Optional[bool]:
        return self._ideal

    @ideal.setter
    def ideal(self, value: Optional[bool]):
        if not isinstance(value, bool):
            raise ValueError('Value must be bool.')
        self._ideal = value


def _fixed(name, mat, single=True):
    m = np.asarray(mat, dtype=np.complex128)

    class _G(QuantumGate):
        NAME = name
        SINGLE = single

        def __init__(self, ideal: Optional[bool] = None, dtype=torch.complex64, device='cpu'):
            super().__init__(ideal, dtype, device)

        def _matrix(self, p):
            return m

    _G.__name__ = name + 'Gate'
    return _G


IGate = _fixed('I', [[1, 0], [0, 1]])                                 # SingleGates.py:17-47
HGate = _fixed('H', [[_S2, _S2], [_S2, -_S2]])                        # SingleGates.py:50-80
XGate = _fixed('X', [[0, 1], [1, 0]])                                 # XGates.py:15-45
YGate = _fixed('Y', [[0, -1j], [1j, 0]])                              # YGates.py:15-45
ZGate = _fixed('Z', [[1, 0], [0, -1]])                                # ZGates.py:15-45
SGate = _fixed('S', [[1, 0], [0, 1j]])                                # PhaseGates.py:16-48
SDGGate = _fixed('Sdg', [[1, 0], [0, -1j]])                           # PhaseGates.py:49-81
TGate = _fixed('T', [[1, 0], [0, (1 + 1j) * _S2]])                    # PhaseGates.py:82-114
CXGate = _fixed('CX', np.eye(4)[[0, 1, 3, 2]], False)                 # XGates.py:88-120
CNOTGate = _fixed('CNOT', np.eye(4)[[0, 1, 3, 2]], False)             # DoubleGates.py:50-82
CYGate = _fixed('CY', [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]], False)   # YGates.py:88-121
CZGate = _fixed('CZ', np.diag([1, 1, 1, -1]), False)                  # ZGates.py:87-119
IIGate = _fixed('II', np.eye(4), False)                               # DoubleGates.py:15-47
SWAPGate = _fixed('SWAP', np.eye(4)[[0, 2, 1, 3]], False)             # DoubleGates.py:120-152
ISWAPGate = _fixed('ISWAP', [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], False)  # DoubleGates.py:85-117


def _param(name, fn, single=True, nparams=1, variational=True):
    class _G(QuantumGate):
        NAME = name
        SINGLE = single
        VARIATIONAL = variational

        def __init__(self, *args, ideal: Optional[bool] = None, dtype=torch.complex64, device='cpu'):
            params, rest = args[:nparams], args[nparams:]
            if rest:
                ideal = rest[0]
            super().__init__(ideal, dtype, device, *params)

        def _matrix(self, p):
            return fn(*p)

    return _G


def _c(t):
    return math.cos(t / 2)


def _s(t):
    return math.sin(t / 2)


RXGate = _param('RX', lambda t: [[_c(t), -1j * _s(t)], [-1j * _s(t), _c(t)]])            # XGates.py:48-85
RYGate = _param('RY', lambda t: [[_c(t), -_s(t)], [_s(t), _c(t)]])                        # YGates.py:48-85
RZGate = _param('RZ', lambda t: [[np.exp(-0.5j * t), 0], [0, np.exp(0.5j * t)]])          # ZGates.py:48-84
PGate = _param('P', lambda t: [[1, 0], [0, np.exp(1j * t)]])                              # PhaseGates.py:115-153
U1Gate = _param('U1', lambda t: [[1, 0], [0, np.exp(1j * t)]])                            # SingleGates.py:83-119
U2Gate = _param('U2', lambda ph, la: np.array([[1, -np.exp(1j * la)],
                                               [np.exp(1j * ph), np.exp(1j * (la + ph))]]) * _S2,
                nparams=2)                                                                # SingleGates.py:122-159
U3Gate = _param('U3', lambda t, ph, la: [[_c(t), -np.exp(1j * la) * _s(t)],
                                         [np.exp(1j * ph) * _s(t), np.exp(1j * (la + ph)) * _c(t)]],
                nparams=3)                                                                # SingleGates.py:162-201
CPGate = _param('CP', lambda t: np.diag([1, 1, 1, np.exp(1j * t)]), single=False)        # PhaseGates.py:154-190
RXXGate = _param('RXX', lambda t: [[_c(t), 0, 0, -1j * _s(t)], [0, _c(t), -1j * _s(t), 0],
                                   [0, -1j * _s(t), _c(t), 0], [-1j * _s(t), 0, 0, _c(t)]],
                 single=False)                                                            # XGates.py:123-160
# YGates.py:124-161: the reference's RYY matrix (:138-143) is exp(-i theta/2 X (x) Y)
RYYGate = _param('RYY', lambda t: [[_c(t), 0, 0, -_s(t)], [0, _c(t), _s(t), 0],
                                   [0, -_s(t), _c(t), 0], [_s(t), 0, 0, _c(t)]],
                 single=False)
RZZGate = _param('RZZ', lambda t: np.diag([np.exp(-0.5j * t), np.exp(0.5j * t), np.exp(0.5j * t),
                                           np.exp(-0.5j * t)]), single=False)              # ZGates.py:122-159
PSWAPGate = _param('SWAP', lambda t: [[1, 0, 0, 0], [0, math.cos(t), math.sin(t), 0],
                                      [0, math.sin(t), math.cos(t), 0], [0, 0, 0, 1]],
                   single=False)                                                          # DoubleGates.py:155-190
XXPlusYYGate = _param('SWAP', lambda t, b: [[1, 0, 0, 0], [0, _c(t), -1j * _s(t) * np.exp(1j * b), 0],
                                            [0, -1j * _s(t) * np.exp(-1j * b), _c(t), 0], [0, 0, 0, 1]],
                      single=False, nparams=2)                                            # DoubleGates.py:229-262


class ArbSingleGate(QuantumGate):
    """SingleGates.py:204-237: an arbitrary 2x2 matrix (not variational)."""
    NAME = 'ArbSingleGate'

    def __init__(self, tensor: torch.Tensor, ideal: Optional[bool] = None, dtype=torch.complex64, device='cpu'):
        super().__init__(ideal, dtype, device)
        self._m = torch.as_tensor(tensor).detach().cpu().to(torch.complex128)

    @property
    def tensor(self):
        if tuple(self._m.shape) != (2, 2):
            warnings.warn(f'You are probably adding a noisy single qubit gate, current shape is {tuple(self._m.shape)}')
        return self._m.reshape(2, 2, -1).squeeze()


class ArbDoubleGate(QuantumGate):
    """DoubleGates.py:193-226: an arbitrary (2,2,2,2[,K]) tensor (variational flag as the reference)."""
    NAME = 'ArbDoubleGate'
    SINGLE = False
    VARIATIONAL = True

    def __init__(self, matrix: torch.Tensor, ideal: Optional[bool] = None, dtype=torch.complex64, device='cpu'):
        super().__init__(ideal, dtype, device)
        self._m = torch.as_tensor(matrix).detach().cpu().to(torch.complex128)

    @property
    def tensor(self):
        if tuple(self._m.shape) != (2, 2, 2, 2):
            warnings.warn(f'You are probably adding a noisy double qubit gate, current shape is {tuple(self._m.shape)}')
        return self._m.reshape(2, 2, 2, 2, -1).squeeze()


class _Fixed1(QuantumGate):
    _M = None

    def __init__(self, dtype=torch.complex64, device='cpu'):
        super().__init__(True, dtype, device)

    def _matrix(self, p):
        return self._M


class MeasureX(_Fixed1):
    NAME = 'MeasureX'                                                                     # SingleGates.py:240-271
    _M = np.array([[_S2, _S2], [_S2, -_S2]])


class MeasureY(_Fixed1):
    NAME = 'MeasureY'                                                                     # SingleGates.py:274-305
    _M = np.array([[_S2, -1j * _S2], [_S2, 1j * _S2]])


class MeasureZ(_Fixed1):
    NAME = 'MeasureZ'                                                                     # SingleGates.py:308-339
    _M = np.eye(2)


class Reset0(_Fixed1):
    NAME = 'Reset0'                                                                       # SingleGates.py:342-373
    _M = np.array([[1, 0], [0, 0]])


class Reset1(_Fixed1):
    NAME = 'Reset1'                                                                       # SingleGates.py:376-407
    _M = np.array([[0, 0], [0, 1]])


class _ExpGate(QuantumGate):
    """Tomographed two-qubit gate carrying its Kraus axis (NoiseGates.py:15-92)."""
    SINGLE = False

    def __init__(self, tensor: Optional[torch.Tensor] = None, dtype=torch.complex64, device='cpu'):
        super().__init__(False, dtype, device)
        self._m = None if tensor is None else torch.as_tensor(tensor).detach().cpu().to(torch.complex128)

    @property
    def rank(self):
        return 5

    @property
    def dimension(self):
        return [[2, 2], [2, 2], [16]]


class CZEXPGate(_ExpGate):
    NAME = 'CZEXP'

    @property
    def tensor(self):
        if self._m is None:
            from .realnoise import czExp_channel
            warnings.warn('No (sufficient) CZ input files, use default tensor.')
            return czExp_channel(dtype=torch.complex128)
        return self._m


class CPEXPGate(_ExpGate):
    NAME = 'CPEXP'

    @property
    def tensor(self):
        if self._m is None:
            from .realnoise import cpExp_channel
            warnings.warn('No (sufficient) CP input files, use default tensor.')
            return cpExp_channel(dtype=torch.complex128)
        return self._m
