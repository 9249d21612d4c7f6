"""Noise channels and chip presets (the reference's ``NoiseChannel.py`` and ``ChipInfo.py``).

Kraus tensors are host-side complex128 constants built once per circuit; the device kernels
receive them as complex64. Layouts follow the reference: one qubit (out, in, k), two qubits
(o0, o1, i0, i1, k).
"""
from __future__ import annotations

import math
import warnings
from typing import Optional, Union

import numpy as np
import torch

__all__ = ['ChipInformation', 'NoiseChannel', 'choi_compress', 'compose_kraus']


class ChipInformation:
    """ChipInfo.py:8-98. Besides the presets ``best``/``medium``/``worst``, ``custom(...)``
    (extension) sets arbitrary parameters, and ``decay_param``/``dephasing_param`` (extension)
    give the channel strengths gamma and p directly instead of 1 - exp(-bath*rate*t)."""

    def __init__(self, query_time: str = None):
        self.queryTime = query_time
        self.status = None
        self.gateTime = None
        self.bath_rate = None
        self.dephasing_rate = None
        self.decay_rate = None
        self.T1 = None
        self.T2 = None
        self.chipName = None
        self.dpc_errorRate = None
        self.decay_param = None
        self.dephasing_param = None
        self.timeUnit = 'ns'

    def configure_chip(self, name, gate_time, decay_rate, dephasing_rate, bath_rate, T1, T2, dpc_error_rate,
                       status):
        self.chipName = name
        self.gateTime = gate_time
        self.bath_rate = bath_rate
        self.decay_rate = decay_rate
        self.dephasing_rate = dephasing_rate
        self.T1 = T1
        self.T2 = T2
        self.dpc_errorRate = dpc_error_rate
        self.status = status

    def best(self):
        if self.queryTime is None:
            self.configure_chip('best', 30, 0.0, 0.0, 0.0, 2e11, 2e10, 11e-4, True)
        return self

    def medium(self):
        if self.queryTime is None:
            self.configure_chip('medium', 1, 0.98, 0.02, 0.01, 2e11, 2e10, 5e-2, True)
        return self

    def worst(self):
        if self.queryTime is None:
            self.configure_chip('worst', 30, 0.0, 0.0, 0.0, 2e2, 2e1, 11e-2, True)
        return self

    def custom(self, gate_time=30.0, decay_rate=0.0, dephasing_rate=0.0, bath_rate=0.0, T1=2e11, T2=2e10,
               dpc_error_rate=0.0, decay_param=None, dephasing_param=None, name='custom'):
        self.configure_chip(name, gate_time, decay_rate, dephasing_rate, bath_rate, T1, T2, dpc_error_rate, True)
        self.decay_param = decay_param
        self.dephasing_param = dephasing_param
        return self

    def show_property(self):
        print(f'The chip name is: {self.chipName}')
        print(f'The gate time is: {self.gateTime} {self.timeUnit}')
        print(f'The T1 time is: {self.T1} {self.timeUnit}')
        print(f'The T2 time is: {self.T2} {self.timeUnit}')
        print(f'The depolarization error rate is: {self.dpc_errorRate}')
        print(f'The status of the chip is: {self.status}')


_PAULI = [np.eye(2, dtype=np.complex128),
          np.array([[0, 1], [1, 0]], dtype=np.complex128),
          np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
          np.array([[1, 0], [0, -1]], dtype=np.complex128)]


def _t(x: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128))


class NoiseChannel:
    """NoiseChannel.py:19-170. ``chip`` is a preset name, a ChipInformation, or a dict of
    ``ChipInformation.custom`` arguments (extension)."""

    def __init__(self, chip: Optional[Union[str, dict, ChipInformation]] = None, dtype=torch.complex64,
                 device: Union[str, int] = 'cpu', single_qubit_noise: str = 'decay_dephasing',
                 p1: Optional[float] = None):
        self.dtype = dtype
        self.device = device
        if isinstance(chip, ChipInformation):
            self.chip = chip
        elif isinstance(chip, dict):
            self.chip = ChipInformation().custom(**chip)
        else:
            try:
                self.chip = getattr(ChipInformation(), chip or 'worst')()
            except AttributeError:
                raise AttributeError(f'Chip: {chip} is not supported.')
        c = self.chip
        self.bath_rate, self.decay_rate, self.dephasing_rate = c.bath_rate, c.decay_rate, c.dephasing_rate
        self.T1, self.T2, self.GateTime, self.dpc_errorRate = c.T1, c.T2, c.gateTime, c.dpc_errorRate
        if single_qubit_noise not in ('decay_dephasing', 'depolarizing'):
            raise ValueError(f'Unsupported single-qubit noise: {single_qubit_noise}')
        self.single_qubit_noise = single_qubit_noise
        self.p1 = self.dpc_errorRate if p1 is None else p1
        self.decayTensor = self.decay(self.decay_rate, self.GateTime)
        self.dephasingTensor = self.dephasing(self.dephasing_rate, self.GateTime)
        self.dpCTensor = self.depolarization_noise_channel(p=self.dpc_errorRate)
        self.dpCTensor2 = self.depolarization_noise_channel(p=self.dpc_errorRate, qn=2)
        self.apdeCTensor = self.amp_phase_damping_error(time=self.GateTime, T1=self.T1, T2=self.T2)

    def depolarization_noise_channel(self, p: float, qn: int = 1) -> torch.Tensor:
        """NoiseChannel.py:51-91: weights sqrt(1-(4^q-1)p/4^q) on I..I and sqrt(p/4^q) on the other
        Pauli strings (first qubit's Pauli major); shape (2,)*2q + (4^q,)."""
        if not 0 <= p <= 1:
            raise ValueError('Probability p must be in the range [0, 1].')
        d = 4 ** qn
        w = np.array([math.sqrt(1 - (d - 1) * p / d)] + [math.sqrt(p / d)] * (d - 1))
        mats = list(_PAULI)
        for _ in range(qn - 1):
            mats = [np.kron(e, b) for e in mats for b in _PAULI]
        st = np.stack(mats) * w[:, None, None]
        return _t(np.transpose(st, (1, 2, 0)).reshape([2] * (2 * qn) + [d]))

    def amp_phase_damping_error(self, time: float, T1: float, T2: float) -> torch.Tensor:
        """NoiseChannel.py:93-142."""
        if time < 0 or T1 <= 0 or T2 <= 0:
            raise ValueError('The time, T1, T2 must be greater than or equal to 0, '
                             'for some special cases time = 0 is allowed.')
        if time == 0:
            warnings.warn('The time is 0, which means the noise is not applied.')
        T2p = 2 * T1 * T2 / (2 * T1 - T2)
        p1 = 1 - math.exp(-time / T1)
        p2 = 1 - math.exp(-time / T2p)
        k = np.array([[[1, 0], [0, math.sqrt(max(1 - (p1 + p2), 0.0))]],
                      [[0, 0], [0, math.sqrt(p2)]],
                      [[0, math.sqrt(p1)], [0, 0]]], dtype=np.complex128)
        return _t(np.transpose(k, (1, 2, 0)))

    def _param(self, rate, gate_time, override):
        if gate_time <= 0:
            raise ValueError('The gate time must be greater than 0, or set the gate to IDEAL.')
        if override is not None:
            return float(override)
        return 1 - math.exp(-self.bath_rate * rate * gate_time)

    def decay(self, gamma: float, gate_time: float) -> torch.Tensor:
        """NoiseChannel.py:144-155: K1 = diag(1, sqrt(1-g)), K2 = sqrt(g)|0><1|, layout (out, in, k)."""
        g = self._param(gamma, gate_time, self.chip.decay_param)
        k = np.array([[[1, 0], [0, math.sqrt(1 - g)]], [[0, math.sqrt(g)], [0, 0]]], dtype=np.complex128)
        return _t(np.transpose(k, (1, 2, 0)))

    def dephasing(self, gamma: float, gate_time: float) -> torch.Tensor:
        """NoiseChannel.py:157-170: sqrt(1-p) I, sqrt(p)|0><0|, sqrt(p)|1><1|."""
        p = self._param(gamma, gate_time, self.chip.dephasing_param)
        a, b = math.sqrt(p), math.sqrt(1 - p)
        k = np.array([[[b, 0], [0, b]], [[a, 0], [0, 0]], [[0, 0], [0, a]]], dtype=np.complex128)
        return _t(np.transpose(k, (1, 2, 0)))

    # -- the Kraus tensors a noisy gate carries ---------------------------------------------
    def single_kraus(self, G: torch.Tensor) -> torch.Tensor:
        """Circuit.py:144-155: einsum('nlm, ljk, ji -> nimk', decay, dephasing, G).reshape(2, 2, -1);
        with ``single_qubit_noise='depolarizing'`` (extension) the 1q depolarizing channel."""
        G = G.to(torch.complex128)
        if self.single_qubit_noise == 'depolarizing':
            D = self.depolarization_noise_channel(self.p1, 1)
            return torch.einsum('nlm,li->nim', D, G)
        return torch.einsum('nlm,ljk,ji->nimk', self.decayTensor, self.dephasingTensor, G).reshape(2, 2, -1)

    def double_kraus(self, G: torch.Tensor) -> torch.Tensor:
        """Circuit.py:91-99: einsum('ijklp, klmn -> ijmnp', dpCTensor2, G)."""
        return torch.einsum('ijklp,klmn->ijmnp', self.dpCTensor2, G.to(torch.complex128))


def compose_kraus(first: torch.Tensor, then: torch.Tensor) -> torch.Tensor:
    """Kraus set of ``then`` after ``first`` (both (2,2[,K])): E_(i,j) = F2_j F1_i, index i major
    (the old inner index major, as the reference flattens successive Kraus axes)."""
    a = first if first.dim() == 3 else first.unsqueeze(-1)
    b = then if then.dim() == 3 else then.unsqueeze(-1)
    return torch.einsum('abj,bci->acij', b, a).reshape(2, 2, -1)


def choi_compress(E: torch.Tensor, rel_tol: float = 1e-14) -> torch.Tensor:
    """Minimal Kraus set of the 1q channel with Kraus tensor E (2,2,K): eigen-decomposition of
    the Choi matrix sum_k vec(E_k) vec(E_k)^H, operators sqrt(l_i) unvec(v_i) for
    l_i > rel_tol * l_max. The new set is related to the old by an isometry on the Kraus index,
    so every quantity of the path that contracts or splits that index is unchanged."""
    if E.dim() == 2:
        return E.unsqueeze(-1)
    K = E.shape[2]
    vecs = E.permute(2, 0, 1).reshape(K, 4).to(torch.complex128)          # rows: vec(E_k)
    J = vecs.T @ vecs.conj()                                               # 4x4 Choi (up to layout)
    w, V = torch.linalg.eigh(J)
    w, V = w.flip(0), V.flip(1)
    keep = w > rel_tol * max(float(w[0]), 1e-300)
    if int(keep.sum()) >= K:
        return E
    ops = (V[:, keep] * torch.sqrt(w[keep])[None, :]).T.reshape(-1, 2, 2)
    return ops.permute(1, 2, 0).contiguous()
