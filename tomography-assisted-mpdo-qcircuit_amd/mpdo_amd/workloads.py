"""Seeded synthetic circuits of the benchmark configurations.

A circuit is a list of ops ``(name, qubits, params)`` plus ``('TRUNCATE', [], [])`` markers,
replayed onto a ``TensorCircuit`` with ``apply_ops``. Angles come from
``numpy.random.Generator(PCG64(seed))`` drawn in (layer, qubit, [theta, phi]) order.
One circuit-layer = single-qubit sublayer (with noise) + two-qubit sublayer (with noise) +
one truncation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

Op = Tuple[str, List[int], List[float]]


@dataclass
class Workload:
    name: str
    n: int
    ops: List[Op]
    noise_type: str
    ideal: bool
    chi: Optional[int]
    kappa: Optional[int]
    chip: Optional[dict] = None
    single_qubit_noise: str = 'decay_dephasing'
    p1: Optional[float] = None
    depth: int = 0
    extra: Dict = field(default_factory=dict)

    def circuit_kwargs(self) -> dict:
        return dict(qn=self.n, ideal=self.ideal, noiseType=self.noise_type, chi=self.chi, kappa=self.kappa,
                    chip=self.chip, single_qubit_noise=self.single_qubit_noise, p1=self.p1)


def brickwork(n: int, depth: int, rng: np.random.Generator, two: Sequence[str] = ('CNOT',),
              angle_lo: float = 0.0, angle_hi: float = 2 * math.pi, noisy_identity_after_2q: bool = False,
              truncate_each_layer: bool = True, delta: Optional[np.ndarray] = None) -> List[Op]:
    """RY(theta) RZ(phi) on every qubit, then a two-qubit gate on (0,1),(2,3),... in even
    layers and (1,2),(3,4),... in odd layers (``two[layer % len(two)]``), then a truncation."""
    angles = rng.uniform(angle_lo, angle_hi, size=(depth, n, 2))
    if delta is not None:
        angles = angles + delta
    ops: List[Op] = []
    for d in range(depth):
        for q in range(n):
            ops.append(('RY', [q], [float(angles[d, q, 0])]))
            ops.append(('RZ', [q], [float(angles[d, q, 1])]))
        g = two[d % len(two)]
        for a in range(d % 2, n - 1, 2):
            ops.append((g, [a, a + 1], []))
            if noisy_identity_after_2q:
                ops.append(('I', [a], []))
                ops.append(('I', [a + 1], []))
        if truncate_each_layer:
            ops.append(('TRUNCATE', [], []))
    return ops


def brickwork8_chi32(seed: int = 101) -> Workload:
    """8 qubits, 4 layers, CNOT brickwork, dep1 p=1e-3 (1q), dep2 p=1e-2 (2q), chi=32, kappa=4."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return Workload('brickwork8_chi32', 8, brickwork(8, 4, rng), 'idealNoise', False, 32, 4,
                    chip=dict(dpc_error_rate=0.01), single_qubit_noise='depolarizing', p1=1e-3, depth=4)


def brickwork20_damping_chi64(seed: int = 202, n: int = 20, depth: int = 20) -> Workload:
    """20 qubits, depth 20, CNOT (even) / CZ (odd) brickwork, amplitude damping 0.005 plus
    dephasing 0.005 after every gate (1q: the reference composite; 2q: noisy identities on
    both qubits), chi=64, kappa=4."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ops = brickwork(n, depth, rng, two=('CNOT', 'CZ'), noisy_identity_after_2q=True)
    return Workload('brickwork20_damping_chi64', n, ops, 'unified', False, 64, 4,
                    chip=dict(decay_param=0.005, dephasing_param=0.005), depth=depth)


def hea16_chi64(seed: int = 303) -> Workload:
    """16-qubit hardware-efficient ansatz: 8 layers RY, RZ (angles U[-pi, pi)) + CNOT chain,
    depolarizing p=0.005 per gate, chi=64, kappa=4."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n, depth = 16, 8
    angles = rng.uniform(-math.pi, math.pi, size=(depth, n, 2))
    ops: List[Op] = []
    for d in range(depth):
        for q in range(n):
            ops.append(('RY', [q], [float(angles[d, q, 0])]))
            ops.append(('RZ', [q], [float(angles[d, q, 1])]))
        for a in range(n - 1):
            ops.append(('CNOT', [a, a + 1], []))
        ops.append(('TRUNCATE', [], []))
    return Workload('hea16_chi64', n, ops, 'idealNoise', False, 64, 4, chip=dict(dpc_error_rate=0.005),
                    single_qubit_noise='depolarizing', p1=0.005, depth=depth)


def random_brickwork_batch(seed: int = 404, count: int = 256) -> List[Workload]:
    """Independent random brickwork circuits: n ~ U{8..24}, depth ~ U{4..16},
    chi ~ {16,32,64,128}, kappa ~ {2,4}, depolarizing or amplitude damping (Bernoulli 1/2)
    with p ~ U[1e-4, 1e-2]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for c in range(count):
        n = int(rng.integers(8, 25))
        depth = int(rng.integers(4, 17))
        chi = int(rng.choice([16, 32, 64, 128]))
        kappa = int(rng.choice([2, 4]))
        dep = bool(rng.integers(0, 2))
        p = float(rng.uniform(1e-4, 1e-2))
        ops = brickwork(n, depth, rng, noisy_identity_after_2q=not dep)
        if dep:
            w = Workload(f'random_brickwork_{c}', n, ops, 'idealNoise', False, chi, kappa, chip=dict(dpc_error_rate=p),
                         single_qubit_noise='depolarizing', p1=p, depth=depth)
        else:
            w = Workload(f'random_brickwork_{c}', n, ops, 'unified', False, chi, kappa,
                         chip=dict(decay_param=p, dephasing_param=0.0), depth=depth)
        out.append(w)
    return out


def brickwork32_chi256_family(seed: int = 505, copies: int = 32, n: int = 32, depth: int = 30) -> List[Workload]:
    """32 qubits, depth 30 RY/RZ + CNOT brickwork, dep1/dep2 p=0.002, chi=256, kappa=8; copies
    with the same base angles plus delta ~ U[-0.01, 0.01]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    base_state = rng.bit_generator.state
    out = []
    for c in range(copies):
        r = np.random.Generator(np.random.PCG64(seed))
        r.bit_generator.state = base_state
        delta = None
        if c > 0:
            dr = np.random.Generator(np.random.PCG64(seed * 1000 + c))
            delta = dr.uniform(-0.01, 0.01, size=(depth, n, 2))
        ops = brickwork(n, depth, r, delta=delta)
        out.append(Workload(f'brickwork32_chi256_{c}', n, ops, 'idealNoise', False, 256, 8, chip=dict(dpc_error_rate=0.002),
                            single_qubit_noise='depolarizing', p1=0.002, depth=depth))
    return out


def apply_ops(circuit, ops: Sequence[Op]) -> None:
    """Replay ops onto a TensorCircuit through its builder methods."""
    for name, q, p in ops:
        n = name.upper()
        if n == 'TRUNCATE':
            circuit.truncate()
        elif n == 'BARRIER':
            circuit.barrier()
        elif n in ('RX', 'RY', 'RZ', 'P', 'U1'):
            getattr(circuit, n.lower())(p[0], q)
        elif n in ('I', 'H', 'X', 'Y', 'Z', 'S', 'SDG', 'T'):
            getattr(circuit, n.lower())(q)
        elif n in ('CNOT', 'CX', 'CZ', 'CY', 'SWAP', 'ISWAP', 'II'):
            getattr(circuit, n.lower())(q[0], q[1])
        elif n in ('CP', 'RXX', 'RYY', 'RZZ', 'PSWAP'):
            getattr(circuit, n.lower())(p[0], q[0], q[1])
        else:
            raise ValueError(f'unknown op {name}')


def steady_state(dims: Sequence[Tuple[int, int, int]], seed: int = 0, decay: float = 0.97) -> List[np.ndarray]:
    """Synthetic chain with the given site shapes (l, a, r): random complex64 site tensors
    T[l, 2, a, r] whose bond and inner indices carry geometrically decaying weights (a physical,
    non-flat spectrum), each site scaled so that the running transfer-matrix trace stays 1
    (Tr rho = 1, as for a state evolved from |0..0> by trace-preserving channels).
    ``batched.steady_dims`` gives the shapes a long brickwork circuit settles into."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    env = np.ones((1, 1), dtype=np.complex128)
    for (l, a, r) in dims:
        T = rng.standard_normal((l, 2, a, r)) + 1j * rng.standard_normal((l, 2, a, r))
        T *= (decay ** np.arange(l))[:, None, None, None]
        T *= (decay ** np.arange(r))[None, None, None, :]
        T *= (0.5 ** np.arange(a))[None, None, :, None]
        if env.shape[0] != l:      # a segment cut out of a longer chain: start from the identity
            env = np.eye(l, dtype=np.complex128) / l
        nxt = np.einsum('xy,xsar,ysaz->rz', env, T, T.conj(), optimize=True)
        f = float(np.trace(nxt).real)
        T /= np.sqrt(f)
        env = nxt / f
        out.append(T.astype(np.complex64))
    return out


def steady_chains_device(dims: Sequence[Tuple[int, int, int]], copies: int, device, seed: int = 2024,
                         decay: float = 0.97):
    """``copies`` synthetic chains at once on ``device``: one (copies, l, 2, a, r) complex64 tensor
    per site, random with weights decay^l, decay^r on the bonds and 0.5^a on the inner index, each
    site scaled so that every chain's running transfer-matrix trace stays 1 (as ``steady_state``,
    drawn from a torch generator on the device so that 32 chains of 32 sites at chi=256 take
    seconds, not minutes)."""
    import torch
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    out = []
    env = torch.ones(copies, 1, 1, dtype=torch.complex128, device=dev)
    for (l, a, r) in dims:
        T = torch.complex(torch.randn(copies, l, 2, a, r, generator=g, device=dev),
                          torch.randn(copies, l, 2, a, r, generator=g, device=dev))
        wl = decay ** torch.arange(l, device=dev, dtype=torch.float32)
        wr = decay ** torch.arange(r, device=dev, dtype=torch.float32)
        wa = 0.5 ** torch.arange(a, device=dev, dtype=torch.float32)
        T = T * (wl[:, None, None, None] * wa[None, None, :, None] * wr[None, None, None, :])
        T64 = T.to(torch.complex128)
        # env'[r, r'] = sum env[l, l'] T[l, s, a, r] conj(T[l', s, a, r'])
        X = torch.matmul(env.transpose(1, 2), T64.reshape(copies, l, 2 * a * r))      # [b, l', (s a r)]
        nxt = torch.matmul(X.reshape(copies, l * 2 * a, r).transpose(1, 2),
                           T64.conj().reshape(copies, l * 2 * a, r))                  # [b, r, r']
        f = torch.diagonal(nxt, dim1=1, dim2=2).sum(-1).real
        T = (T64 / torch.sqrt(f).reshape(copies, 1, 1, 1, 1)).to(torch.complex64)
        env = nxt / f.reshape(copies, 1, 1)
        out.append(T.contiguous())
    return out


def layer_ops(ops: Sequence[Op], layer: int) -> List[Op]:
    """The ops of one circuit-layer (everything up to and including its truncation)."""
    out, cur = [], 0
    for op in ops:
        if cur == layer:
            out.append(op)
        if op[0] == 'TRUNCATE':
            cur += 1
            if cur > layer:
                break
    return out
