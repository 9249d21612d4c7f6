"""Circuit builder and MPDO engine with the reference's API (AbstractCircuit.py, Circuit.py).

``QuantumCircuit`` keeps the builder semantics of AbstractCircuit.py:22-532 (layer list as an
``nn.Sequential`` plus the operating-qubit list, realNoise rewrites of cx/cnot/cz/cp/rxx/ryy/
rzz). ``TensorCircuit.evolve`` runs the layer loop of Circuit.py:469-491 over the fixed-layout
site list; every contraction, SVD, QR and observable is a call into libmpdo_gfx950.

Consecutive single-qubit operations on one qubit with no two-qubit gate or truncation between
them are composed into one channel and reduced to its Choi rank (at most 4 Kraus operators)
before they reach the device (``fuse_single_qubit=True``). The composed Kraus set is related
to the reference's per-gate sets by an isometry on the inner index, so every truncation
decision and every gauge-invariant output (density matrix, expectations, discarded weights)
is unchanged while the inner index stays up to 16x smaller.
"""
from __future__ import annotations

import math
import os
from abc import ABC
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Tuple, Union

import numpy as np
import torch
from torch import nn

from . import TNNOptimizer as _opt
from . import _lib
from . import gates as G
from .dmOperations import DMNodes
from .node import SiteNode, replicate_nodes
from .noise import NoiseChannel, choi_compress, compose_kraus
from .realnoise import cpExp_channel, czExp_channel
from . import sampling as _sampling
from .Tools import count_item, select_device

__all__ = ['QuantumCircuit', 'TensorCircuit', 'GLOBAL_MINIMUM', 'MAX_ATTEMPTS']

GLOBAL_MINIMUM = 2.718281828459045 * 1e-8     # Circuit.py:22
MAX_ATTEMPTS = 4                              # Circuit.py:23

_CHI = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'chi')
CHIFILENAMES = {'CZ': {'01': os.path.join(_CHI, 'czDefault.mat')}, 'CP': {}}   # AbstractCircuit.py:16-19


def _info(p) -> float:
    return float(p.item()) if isinstance(p, torch.Tensor) else float(p)


class QuantumCircuit(ABC, nn.Module):
    """AbstractCircuit.py:22-532."""

    def __init__(self, noiseFiles: Optional[Dict[str, Dict[str, Any]]] = None, chi: Optional[int] = None,
                 kappa: Optional[int] = None, max_truncation_err: Optional[float] = None, dtype=torch.complex64,
                 device: Union[str, int, None] = 'cpu'):
        super().__init__()
        self.device = select_device(device)
        self.dtype = dtype
        self.chi = chi
        self.kappa = kappa
        self.max_truncation_err = max_truncation_err
        self.layers = nn.Sequential()
        self._oqs_list: List[List[Optional[int]]] = []
        self.noiseTensorDict: Dict[str, torch.Tensor] = {}
        self.unified, self.realNoise, self.idealNoise = False, False, False
        self.noiseFiles = noiseFiles if noiseFiles is not None else CHIFILENAMES
        self._initState = None
        self._vector = None
        self._stateNodes, self._dm, self._dmNodes = None, None, None
        self._samples, self._counts = None, None
        self._sequence = 0
        self._nodes4samples, self._indices4samples = None, None

    def _load_exp_tensors(self):
        """AbstractCircuit.py:66-75."""
        self._cz_expTensors, self._cp_expTensors = {}, {}
        for key, val in self.noiseFiles.get('CZ', {}).items():
            self._cz_expTensors[key] = czExp_channel(filename=val)
        for key, val in self.noiseFiles.get('CP', {}).items():
            self._cp_expTensors[key] = cpExp_channel(filename=val)

    @dataclass
    class Group:
        history: List[int]
        start: int
        length: int

    @property
    def dm(self):
        return self._dm

    @property
    def samples(self):
        return self._samples

    @property
    def counts(self):
        return self._counts

    @property
    def initial_state(self):
        return self._initState

    @property
    def stateNodes(self):
        return self._stateNodes

    @property
    def dmNodes(self):
        return self._dmNodes

    @property
    def vector(self):
        return self._vector

    def _add_module(self, _gate: nn.Module, oqs: List, headline: str):
        self._oqs_list.append(oqs)
        self.layers.add_module(headline + f'-S{self._sequence}', _gate)
        self._sequence += 1

    def _iter_add_module(self, _gate_list: List, oqs_list: List, _transpile: bool = False):
        """AbstractCircuit.py:143-154 (its condition ``... or 'CNOT'`` is always true, so every
        gate is added directly)."""
        for _gate, _oq in zip(_gate_list, oqs_list):
            if _gate.para is None:
                headline = f'{_gate.name}{_oq}|None-TRANS'
            else:
                headline = f'{_gate.name}{_oq}|({_info(_gate.para if not isinstance(_gate.para, list) else _gate.para[0]):.3f})-TRANS'.replace('.', ';')
            self._add_module(_gate, _oq, headline)

    # -- one-qubit gates ------------------------------------------------------------------
    def _one(self, cls, oqs, _ideal, headline_name, *params):
        if isinstance(oqs, int):
            oqs = [oqs]
        if params:
            head = f'{headline_name}{oqs}|({_info(params[0]):.3f})'.replace('.', ';')
        else:
            head = f'{headline_name}{oqs}|None'
        self._add_module(cls(*params, _ideal, dtype=self.dtype, device=self.device) if params
                         else cls(_ideal, dtype=self.dtype, device=self.device), oqs, head)

    def i(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.IGate, oqs, _ideal, 'I')

    def h(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.HGate, oqs, _ideal, 'H')

    def x(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.XGate, oqs, _ideal, 'X')

    def y(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.YGate, oqs, _ideal, 'Y')

    def z(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.ZGate, oqs, _ideal, 'Z')

    def rx(self, theta, oqs, _ideal: Optional[bool] = None):
        self._one(G.RXGate, oqs, _ideal, 'RX', theta)

    def ry(self, theta, oqs, _ideal: Optional[bool] = None):
        self._one(G.RYGate, oqs, _ideal, 'RY', theta)

    def rz(self, theta, oqs, _ideal: Optional[bool] = None):
        self._one(G.RZGate, oqs, _ideal, 'RZ', theta)

    def s(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.SGate, oqs, _ideal, 'S')

    def sdg(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.SDGGate, oqs, _ideal, 'SDG')

    def t(self, oqs, _ideal: Optional[bool] = None):
        self._one(G.TGate, oqs, _ideal, 'T')

    def p(self, theta, oqs, _ideal: Optional[bool] = None):
        self._one(G.PGate, oqs, _ideal, 'P', theta)

    def u1(self, theta, oqs, _ideal: Optional[bool] = None):
        self._one(G.U1Gate, oqs, _ideal, 'U1', theta)

    def u2(self, phi, lam, oqs, _ideal: Optional[bool] = None):
        if isinstance(oqs, int):
            oqs = [oqs]
        head = f'U2{oqs}|(P{_info(phi):3f})-(L{_info(lam):3f})'.replace('.', ';')
        self._add_module(G.U2Gate(phi, lam, _ideal, dtype=self.dtype, device=self.device), oqs, head)

    def u3(self, theta, phi, lam, oqs, _ideal: Optional[bool] = None):
        if isinstance(oqs, int):
            oqs = [oqs]
        head = f'U3{oqs}|(T{_info(theta):3f})-(P{_info(phi):3f})-(L{_info(lam):3f})'.replace('.', ';')
        self._add_module(G.U3Gate(theta, phi, lam, _ideal, dtype=self.dtype, device=self.device), oqs, head)

    def arbSingle(self, data: torch.Tensor, oqs, _ideal: Optional[bool] = None):
        if isinstance(oqs, int):
            oqs = [oqs]
        self._add_module(G.ArbSingleGate(data, _ideal, dtype=self.dtype, device=self.device), oqs, f'ArbS{oqs}|None')

    # -- two-qubit gates (AbstractCircuit.py:231-447) -----------------------------------------
    def _two(self, cls, oqs, _ideal, headline_name, *params):
        if params:
            head = f'{headline_name}{oqs}|({_info(params[0]):.3f})'.replace('.', ';')
        else:
            head = f'{headline_name}{oqs}|None'
        self._add_module(cls(*params, _ideal, dtype=self.dtype, device=self.device) if params
                         else cls(_ideal, dtype=self.dtype, device=self.device), oqs, head)

    def rxx(self, theta, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        oqs = [oq0, oq1]
        if not self.realNoise or _ideal:
            self._two(G.RXXGate, oqs, _ideal, 'RXX', theta)
        else:
            self.h(oqs, True); self.cx(oq0, oq1); self.rz(theta, oq1, True); self.cx(oq0, oq1); self.h(oqs, True)

    def ryy(self, theta, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        oqs = [oq0, oq1]
        if not self.realNoise or _ideal:
            self._two(G.RYYGate, oqs, _ideal, 'RYY', theta)
        else:
            self.rx(torch.tensor(math.pi / 2), oqs, True); self.cx(oq0, oq1); self.rz(theta, oq1, True)
            self.cx(oq0, oq1); self.rx(-torch.tensor(math.pi / 2), oqs, True)

    def rzz(self, theta, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        oqs = [oq0, oq1]
        if not self.realNoise or _ideal:
            self._two(G.RZZGate, oqs, _ideal, 'RZZ', theta)
        else:
            self.cx(oq0, oq1); self.rz(theta, oq1, True); self.cx(oq0, oq1)

    def xx_yy(self, theta, beta, control: int, target: int, _ideal: Optional[bool] = None):
        oqs = [control, target]
        head = f'XXPlusYYGate{oqs}|(P{_info(theta):3f})-(L{_info(beta):3f})'.replace('.', ';')
        self._add_module(G.XXPlusYYGate(theta, beta, _ideal, dtype=self.dtype, device=self.device), oqs, head)

    def cx(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        if not self.realNoise or _ideal:
            self._two(G.CXGate, [oq0, oq1], _ideal, 'CX')
        else:
            self.ry(-torch.tensor(math.pi / 2), oq1, True); self.cz(oq0, oq1); self.ry(torch.tensor(math.pi / 2), oq1, True)

    def cy(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        if not self.realNoise or _ideal:
            self._two(G.CYGate, [oq0, oq1], _ideal, 'CY')
        else:
            raise NotImplementedError('EXPCYGate is not implemented yet.')

    def cz(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        oqs = [oq0, oq1]
        if not self.realNoise or _ideal:
            self._two(G.CZGate, oqs, _ideal, 'CZ')
        else:
            t = self._cz_expTensors.get(f'{oq0}{oq1}')
            if t is None:
                t = self._cz_expTensors.get(f'{oq1}{oq0}')
            self._add_module(G.CZEXPGate(t, dtype=self.dtype, device=self.device), oqs, f'CZEXP{oqs}|None')

    def cnot(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        if not self.realNoise or _ideal:
            self._two(G.CNOTGate, [oq0, oq1], _ideal, 'CNOT')
        else:
            self.ry(-torch.tensor(math.pi / 2), oq1, True); self.cz(oq0, oq1); self.ry(torch.tensor(math.pi / 2), oq1, True)

    def swap(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        self._two(G.SWAPGate, [oq0, oq1], _ideal, 'SWAP')

    def pswap(self, theta, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        self._two(G.PSWAPGate, [oq0, oq1], _ideal, 'PSWAP', theta)

    def iswap(self, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        self._two(G.ISWAPGate, [oq0, oq1], _ideal, 'ISWAP')

    def cp(self, theta, oq0: int, oq1: int, _ideal: Optional[bool] = None):
        oqs = [oq0, oq1]
        if not self.realNoise or _ideal:
            self._two(G.CPGate, oqs, _ideal, 'CP', theta)
        else:
            t = self._cp_expTensors.get(f'{oq0}{oq1}')
            if t is None:
                t = self._cp_expTensors.get(f'{oq1}{oq0}')
            self._add_module(G.CPEXPGate(t, dtype=self.dtype, device=self.device), oqs, f'CPEXP{oqs}|None')

    def arbDouble(self, data: torch.Tensor, oq1: int, oq2: int, _ideal: Optional[bool] = None):
        oqs = [oq1, oq2]
        self._add_module(G.ArbDoubleGate(data, _ideal, dtype=self.dtype, device=self.device), oqs, f'ArbD{oqs}|None')

    def ii(self, oq1: int, oq2: int, _ideal: Optional[bool] = None):
        self._two(G.IIGate, [oq1, oq2], _ideal, 'II')

    # -- layer markers and resets (AbstractCircuit.py:483-532) --------------------------------
    def truncate(self):
        self._oqs_list.append([None])
        self.layers.append(G.Truncate())

    def barrier(self):
        self._oqs_list.append([None])
        self.layers.append(G.Barrier())

    def reset0(self, oqs):
        if isinstance(oqs, int):
            oqs = [oqs]
        self._add_module(G.Reset0(dtype=self.dtype, device=self.device), oqs, f'Reset;0{oqs}|None')

    def reset1(self, oqs):
        if isinstance(oqs, int):
            oqs = [oqs]
        self._add_module(G.Reset1(dtype=self.dtype, device=self.device), oqs, f'Reset;1{oqs}|None')

    def measure(self, oqs, orientations=None):
        oqs = [oqs] if isinstance(oqs, int) else oqs
        orientations = orientations or [2] * len(oqs)
        orientations = [orientations] if isinstance(orientations, int) else orientations
        mp = {0: G.MeasureX, 1: G.MeasureY, 2: G.MeasureZ}
        for oq, ori in zip(oqs, orientations):
            if ori not in mp:
                raise ValueError('Orientation beyond the settings.')
            self._add_module(mp[ori](dtype=self.dtype, device=self.device), [oq], f'Measure;{oq}|Orientation;{ori}')


class TensorCircuit(QuantumCircuit):
    """Circuit.py:26-498. Extensions (keyword-only): ``single_qubit_noise`` ('decay_dephasing' as
    the reference, or 'depolarizing' with strength ``p1``), ``fuse_single_qubit`` (see module
    docstring), ``record_discarded`` (keep the discarded weight of every cut in
    ``self.discarded``)."""

    def __init__(self, qn: int, ideal: bool = True, noiseType: str = 'no',
                 chiFileDict: Optional[Dict[str, Dict[str, Any]]] = None, chi: Optional[int] = None,
                 kappa: Optional[int] = None, max_truncation_err: Optional[float] = None,
                 chip: Optional[Union[str, dict]] = None, dtype=torch.complex64,
                 device: Optional[Union[str, int]] = None, *, single_qubit_noise: str = 'decay_dephasing',
                 p1: Optional[float] = None, fuse_single_qubit: bool = True, record_discarded: bool = True):
        super().__init__(noiseFiles=chiFileDict, chi=chi, kappa=kappa, max_truncation_err=max_truncation_err,
                         dtype=dtype, device=device)
        if dtype != torch.complex64:
            raise NotImplementedError('the gfx950 kernels compute in complex64 (the reference default); '
                                      f'dtype {dtype} is not supported')
        self.qnumber = qn
        self.ideal = ideal
        self.noiseType = noiseType.lower()
        self.Noise = None
        self.fuse = fuse_single_qubit
        self.record_discarded = record_discarded
        self.discarded: List[Tuple[str, int, float]] = []
        if not ideal:
            if self.noiseType not in ['unified', 'realnoise', 'idealnoise']:
                raise ValueError(f'Unsupported noise type: {self.noiseType}')
            self.Noise = NoiseChannel(chip=chip, dtype=dtype, device=self.device,
                                      single_qubit_noise=single_qubit_noise, p1=p1)
            self.unified = self.noiseType == 'unified'
            self.realNoise = self.noiseType == 'realnoise'
            self.idealNoise = self.noiseType == 'idealnoise'
            if self.realNoise:
                self._load_exp_tensors()
        self._pending: List[Optional[Tuple[torch.Tensor, bool]]] = []

    # -- gate application ----------------------------------------------------------------------
    def _noisy_single(self, gate) -> torch.Tensor:
        G1 = gate.tensor.to(torch.complex128)
        if not gate.variational:
            return self.noiseTensorDict.setdefault(gate.name, self.Noise.single_kraus(G1))
        return self.Noise.single_kraus(G1)

    def _single_kraus(self, gate) -> Tuple[torch.Tensor, bool]:
        """Kraus tensor (2, 2[, K]) a single-qubit gate applies (Circuit.py:140-155) and whether it
        carries noise (then the site gains the inner index I_i)."""
        gnoise = (self.idealNoise or self.unified) and not gate.ideal
        E = self._noisy_single(gate) if gnoise else gate.tensor.to(torch.complex128)
        if E.dim() not in (2, 3) or tuple(E.shape[:2]) != (2, 2):
            raise ValueError(f'single-qubit gate tensor of shape {tuple(E.shape)}')
        return E, gnoise

    def _apply_single_qubit_gate(self, _qNodes: List[SiteNode], gate, _oqs: List[int]):
        """Circuit.py:138-178."""
        E, gnoise = self._single_kraus(gate)
        if self.fuse:
            for q in _oqs:
                prev = self._pending[q]
                if prev is None:
                    self._pending[q] = (E, gnoise)
                else:
                    self._pending[q] = (compose_kraus(prev[0], E), prev[1] or gnoise)
            return
        outs = _lib.apply_1q([_qNodes[q].tensor for q in _oqs], [E] * len(_oqs))
        for q, t in zip(_oqs, outs):
            _qNodes[q].tensor = t
            if gnoise and E.dim() == 3:
                _qNodes[q].has_I = True

    def _flush(self, _qNodes: List[SiteNode], sites) -> None:
        todo = [q for q in sites if self._pending[q] is not None]
        if not todo:
            return
        ks = []
        for q in todo:
            E, _ = self._pending[q]
            if E.dim() == 3 and E.shape[2] > 1:
                E = choi_compress(E)
            ks.append(E)
        outs = _lib.apply_1q([_qNodes[q].tensor for q in todo], ks)
        for q, t in zip(todo, outs):
            _qNodes[q].tensor = t
            if self._pending[q][1]:
                _qNodes[q].has_I = True
            self._pending[q] = None

    def _double_kraus(self, gate, _oqs: List[int]) -> torch.Tensor:
        """Tensor (o_min, o_max, i_min, i_max[, K]) a two-qubit gate applies to neighbouring qubits
        (Circuit.py:80-106), oriented so that the lower qubit index comes first."""
        if len(_oqs) != 2 or _oqs[0] == _oqs[1]:
            raise ValueError('Invalid operating qubits for a two-qubit gate.')
        if max(_oqs) - min(_oqs) != 1:
            raise NotImplementedError('two-qubit gates act on neighbouring qubits of the chain')
        gnoise = (self.idealNoise and not gate.ideal) or self.realNoise
        Gt = gate.tensor.to(torch.complex128)
        if gnoise and not self.realNoise:
            Gt = (self.noiseTensorDict.setdefault(gate.name, self.Noise.double_kraus(Gt)) if not gate.variational
                  else self.Noise.double_kraus(Gt))
        if self.realNoise and Gt.dim() != 5:
            raise ValueError(f'under realNoise the two-qubit gate {gate.name} must carry a Kraus axis '
                             '(CZEXP/CPEXP or a 5-index ArbDoubleGate)')
        if Gt.dim() not in (4, 5):
            raise ValueError(f'two-qubit gate tensor of shape {tuple(Gt.shape)}')
        if _oqs[0] > _oqs[1]:
            Gt = Gt.transpose(0, 1).transpose(2, 3)
        return Gt

    def _apply_two_qubits_gate(self, _qNodes: List[SiteNode], gate, _oqs: List[int]):
        """Circuit.py:74-136 for adjacent qubits."""
        Gt = self._double_kraus(gate, _oqs)
        mx, mn = max(_oqs), min(_oqs)
        self._flush(_qNodes, [mn, mx])
        tmin, tmax = _qNodes[mn].tensor.contiguous(), _qNodes[mx].tensor.contiguous()
        for attempt in range(MAX_ATTEMPTS):
            try:
                a, b, disc = _lib.apply_2q(tmin, tmax, Gt, GLOBAL_MINIMUM)
                break
            except torch.linalg.LinAlgError:
                # Circuit.py:127-134: perturb by eps64 * max(|theta|, 1) * GLOBAL_MINIMUM * 1e-2
                nrm = float(torch.linalg.vector_norm(tmin)) * float(torch.linalg.vector_norm(tmax))
                mag = torch.finfo(torch.float64).eps * max(nrm, 1.0) * GLOBAL_MINIMUM * 1e-2
                tmin = tmin + mag * (torch.randn_like(tmin.real) + 1j * torch.randn_like(tmin.real))
        else:
            raise torch.linalg.LinAlgError(f'SVD converges failed. Multiple times ({MAX_ATTEMPTS}) of perturbation are tried.')
        _qNodes[mn].tensor, _qNodes[mx].tensor = a, b
        _qNodes[mn].bond_right = True
        _qNodes[mx].bond_left = True
        if Gt.dim() == 5:
            _qNodes[mx].has_I = True
        if self.record_discarded:
            self.discarded.append(('gate', mn, disc))

    def _add_gate(self, _qubits: List[SiteNode], _layer_num: int, _oqs: List[int], _gate=None):
        """Circuit.py:180-224."""
        if not isinstance(_qubits, list):
            raise TypeError('Qubit must be a list of nodes.')
        if not isinstance(_oqs, list):
            raise TypeError('Operating qubits must be a list.')
        if max(_oqs) is None:
            return None
        if max(_oqs) >= self.qnumber:
            raise ValueError(f'Qubit index out of range, max index is Q{max(_oqs)}.')
        gate = _gate or self.layers[_layer_num]
        if not gate or gate.name == 'MeasureZ':
            return None
        if not isinstance(gate, G.QuantumGate):
            raise TypeError(f'Gate must be a QuantumGate, current type is {type(gate)}.')
        if gate.single:
            self._apply_single_qubit_gate(_qubits, gate, _oqs)
        else:
            self._apply_two_qubits_gate(_qubits, gate, _oqs)

    # -- truncation ------------------------------------------------------------------------------
    def _truncate(self, state: List[SiteNode]):
        self._flush(state, range(len(state)))
        if _opt.checkConnectivity(state):
            d = _opt.bondTruncate(state, max_singular_values=self.chi, max_truncation_err=self.max_truncation_err,
                                  return_discarded=True)
            if self.record_discarded and d:
                self.discarded.extend(('chi', len(state) - 1 - j, w) for j, w in enumerate(d))
            if not self.ideal:
                self._kappa(state)

    def _kappa(self, state: List[SiteNode]):
        d = _opt.svdKappa_left2right(state, max_singular_values=self.kappa, max_truncation_err=self.max_truncation_err,
                                     return_discarded=True)
        if self.record_discarded and d:
            self.discarded.extend(('kappa', i, w) for i, w in d)

    def _begin(self, state: List[SiteNode]) -> None:
        if not isinstance(state, list):
            raise TypeError('Qubit must be a list of nodes.')
        if len(state) != self.qnumber:
            raise ValueError(f'state has {len(state)} sites for a {self.qnumber}-qubit circuit')
        dev = torch.device(self.device)
        for s in state:
            s.tensor = s.tensor.to(dtype=self.dtype, device=dev).contiguous()
        self._initState = replicate_nodes(state)
        self._pending = [None] * self.qnumber
        self._dm, self._dmNodes = None, None

    def _step(self, state: List[SiteNode], _i: int) -> None:
        """One entry of the layer list (the body of the loop at Circuit.py:476-485)."""
        name = self.layers[_i].name.lower()
        if 'truncate' in name:
            self._truncate(state)
        elif 'barrier' in name:
            pass
        else:
            self._add_gate(state, _i, _oqs=self._oqs_list[_i])

    def _finish(self, state: List[SiteNode]) -> None:
        self._flush(state, range(len(state)))
        if not self.ideal and len(self.layers) and 'truncate' not in self.layers[-1].name:
            self._kappa(state)
        self._stateNodes = state

    def evolve(self, state: List[SiteNode]):
        """Circuit.py:469-491. The caller's list is evolved in place; ``initial_state`` keeps a
        device copy of what it held on entry."""
        self._begin(state)
        for _i in range(len(self.layers)):
            self._step(state, _i)
        self._finish(state)

    def forward(self, state: List[SiteNode]):
        self.evolve(state)

    # -- density matrix / observables (Circuit.py:226-295) ----------------------------------------
    def _state_or_raise(self):
        if self._stateNodes is None:
            raise ValueError('evolve() the circuit first')
        return self._stateNodes

    def cal_dmNodes(self, reduced_index: Optional[List] = None) -> DMNodes:
        reduced_index = [] if reduced_index is None else reduced_index
        if reduced_index and max(reduced_index) >= self.qnumber:
            raise ValueError(f'Reduced index should not be larger than the qubit number. '
                             f'{max(reduced_index)}-{self.qnumber}')
        self._dmNodes = DMNodes(self._state_or_raise(), reduced_index)
        return self._dmNodes

    def cal_dm(self, reduced_index: Optional[List] = None, _replicate_require: bool = False) -> torch.Tensor:
        reduced_index = [] if reduced_index is None else list(reduced_index)
        st = self._state_or_raise()
        if reduced_index and max(reduced_index) >= self.qnumber:
            raise ValueError(f'Reduced index should not be larger than the qubit number. '
                             f'{max(reduced_index)}-{self.qnumber}')
        keep = [i not in reduced_index for i in range(self.qnumber)]
        self._dm = _lib.reduced_dm([s.tensor.contiguous() for s in st], keep)
        if self._dmNodes is None:
            self._dmNodes = DMNodes(st, reduced_index)
        return self._dm

    def cal_vector(self) -> torch.Tensor:
        if not self.ideal:
            raise ValueError('Noisy circuit cannot be represented by state vector efficiently.')
        st = self._state_or_raise()
        self._vector = _lib.contract_state([s.tensor.contiguous() for s in st]).reshape(2 ** self.qnumber, 1)
        return self._vector

    # -- shot sampling (Circuit.py:297-467) -------------------------------------------------------
    _projectors_string = ['X', 'Y', 'Z']     # AbstractCircuit.py:80

    def _conditional_prb(self, _history: List[Union[int, bool]]) -> float:
        """Circuit.py:297-332: probability that qubit ``_indices4samples[len(_history)]`` reads 1
        given the outcomes ``_history`` of the qubits before it (later qubits traced)."""
        if self._nodes4samples is None:
            raise ValueError('call sample() first')
        return _sampling.normalise_prb(*self._nodes4samples.probs_of(_history))

    def _conditional_sample(self, shots: int, _sampleLength: int, _tqdm_disable: bool = False) -> List[List[int]]:
        """Circuit.py:334-342."""
        return _sampling.sequential_sample(self._nodes4samples, shots)

    def _conditional_batch_sample(self, shots: int, _sampleLength: int, _bool: bool = False,
                                  _tqdm_disable: bool = False) -> List[List[int]]:
        """Circuit.py:344-391."""
        return _sampling.batch_sample(self._nodes4samples, shots, bool_result=_bool)

    def sample(self, shots: Optional[int] = None, orientation: Optional[List[int]] = None,
               reduced: Optional[List[int]] = None, sample_string: bool = True, _tqdm_disable: bool = False,
               _require_sequential_sample: bool = False, _require_bool_result: bool = False,
               _require_counts: bool = True, _stateNodes4Sample: Optional[List[SiteNode]] = None):
        """Circuit.py:393-445. ``orientation[i]`` 0/1/2 = X/Y/Z basis; as in the reference the
        MeasureX/MeasureY rotation of entry i is applied to qubit i (the position in the list,
        which is the i-th sampled qubit only when no earlier qubit is reduced)."""
        shots = shots or 1024
        reduced = reduced or []
        _ori_list = [num for num in range(self.qnumber) if num not in reduced]
        _sampleLength = len(_ori_list)
        orientation = orientation or [2] * _sampleLength
        if len(orientation) != _sampleLength:
            raise ValueError('Length of orientation must match the sample length. Check reduced or unmeasured qubits.')
        src = self._state_or_raise() if _stateNodes4Sample is None else _stateNodes4Sample
        tensors = [s.tensor.contiguous() for s in src]
        for _ori_val, cls in [(0, G.MeasureX), (1, G.MeasureY)]:
            _indices = [i for i, value in enumerate(orientation) if value == _ori_val]
            if _indices:
                if max(_indices) >= self.qnumber:
                    raise ValueError(f'Qubit index out of range, max index is Q{max(_indices)}.')
                M = cls(dtype=self.dtype, device=self.device).tensor.to(torch.complex128)
                outs = _lib.apply_1q([tensors[i] for i in _indices], [M] * len(_indices))
                for i, t in zip(_indices, outs):
                    tensors[i] = t
        self._nodes4samples = _sampling.ConditionalSweep(tensors, _ori_list)
        self._indices4samples = _ori_list
        if not _tqdm_disable:
            print('Sample Direction:'
                  f'\n (scheme-{"".join([self._projectors_string[num] for num in orientation])})')
        if _require_sequential_sample:
            _bitStrings = self._conditional_sample(shots, _sampleLength, _tqdm_disable)
        else:
            _bitStrings = self._conditional_batch_sample(shots, _sampleLength, _require_bool_result, _tqdm_disable)
        if sample_string and not _require_bool_result:
            _bitStrings = [''.join(map(str, _stringList)) for _stringList in _bitStrings]
        self._samples = _bitStrings
        if _require_counts:
            self._counts = count_item(_bitStrings)
            return self._samples, self._counts
        return self._samples

    def randomSample(self, measurement_schemes: List[List[int]], shots_per_scheme: int = 1024,
                     reduced: Optional[List[int]] = None, _tqdm_disable: bool = False,
                     _require_sequential_sample: bool = False, _require_bool_result: bool = False,
                     _stateNodes4Sample: Optional[List[SiteNode]] = None) -> List[List[List[Union[int, bool]]]]:
        """Circuit.py:447-467: one sample() per measurement scheme, bit lists, no counts."""
        return [self.sample(shots=shots_per_scheme, orientation=scheme, reduced=reduced, sample_string=False,
                            _tqdm_disable=True, _require_sequential_sample=_require_sequential_sample,
                            _require_bool_result=_require_bool_result, _require_counts=False,
                            _stateNodes4Sample=_stateNodes4Sample)
                for scheme in measurement_schemes]
