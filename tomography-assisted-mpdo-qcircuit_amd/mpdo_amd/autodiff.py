"""Differentiable noisy-circuit evolution: gradients of noisy expectation values with respect to gate
angles and per-gate noise strengths in one reverse sweep.

The reference's circuit is a torch ``nn.Module`` whose gates carry angle tensors
(AbstractCircuit.py:22, YGates.py:48-85, AbstractGate.py:42-51), but it rebuilds every gate matrix
through ``torch.tensor(list)`` (YGates.py:64-69), which cuts the autograd graph. Here the same
update path - single-qubit Kraus absorption, the two-qubit gate with its noise and the gate-time
sqrt(S) split (Circuit.py:74-178), the left-to-right QR sweep, the right-to-left chi sweep and the
kappa cut (TNNOptimizer.py:72-197) - is written in differentiable form:

* every factorisation (QR, the sqrt(S) splits, the kappa cut) is built on the Hermitian
  eigen-decomposition of a Gram matrix, computed by the library's fp64 Jacobi solver
  (``mpdo_eigh``) inside ``GramEigh``, a ``torch.autograd.Function`` whose backward is written
  out: for G = V diag(w) V^H,
      G_bar = V (diag(w_bar) + F o skew(V^H V_bar)) V^H,   F_ij = (w_j - w_i) / ((w_j - w_i)^2 + eta),
  the Lorentzian-broadened 1 / (s_j^2 - s_i^2) of the singular values s^2 = w. Every eigenpair
  enters it, the discarded ones included, so the truncation residual (the coupling of the kept
  directions to the dropped ones) is part of the gradient;
* the contractions (Kraus absorption, the gate's theta, the factor products) are complex128 tensor
  products whose adjoints are the conjugate-transposed products;
* each circuit-layer is checkpointed (``torch.utils.checkpoint``): the forward keeps only the
  chain between layers and the backward recomputes one layer at a time.

The truncation decisions (kept counts) are those of the per-gate path: count caps chi and kappa,
the gate-time split's absolute error GLOBAL_MINIMUM on the singular values, zero-filled directions
below the rank. They are piecewise constant, so the gradient is that of the map with the kept
counts of the forward pass.

Workload: the hardware-efficient ansatz ``HEA`` (RY, RZ on every qubit then a CNOT chain,
depolarizing noise after every gate with its own learnable strength), loss <H> of the transverse
field Ising Hamiltonian H = -sum Z_i Z_i+1 - h sum X_i, unnormalised as the reference's
dmOperations.expect (Tr(H rho~)).
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import torch
from torch.utils.checkpoint import checkpoint

from . import _lib

__all__ = ['GramEigh', 'eigh_backward', 'HEA', 'tfim_energy', 'GLOBAL_MINIMUM']

GLOBAL_MINIMUM = 2.718281828459045e-8      # Circuit.py:22 (the gate-time split's absolute error)
QR_FLOOR = 1e-13                            # Gram eigenvalues below this x the largest: zero directions
SPLIT_FLOOR = 1e-12                         # the sqrt(S) splits' floor (singular values below 1e-6 s_0)
BROADEN = 1e-12                             # eta = (BROADEN x largest eigenvalue)^2


# ------------------------------------------------------------------------------------------------
def eigh_backward(w: torch.Tensor, V: torch.Tensor, gw: Optional[torch.Tensor], gV: Optional[torch.Tensor],
                  eta: torch.Tensor) -> torch.Tensor:
    """Cotangent of a Hermitian G = V diag(w) V^H from those of (w, V): V (diag(gw) + F o skew(V^H gV)) V^H,
    F_ij = E_ij / (E_ij^2 + eta) with E_ij = w_j - w_i (zero on the diagonal). w: (..., n) real,
    V: (..., n, n), eta: broadcastable to (..., n, n)."""
    n = w.shape[-1]
    inner = torch.zeros(*w.shape[:-1], n, n, dtype=V.dtype, device=V.device)
    if gV is not None:
        VhgV = V.mH @ gV
        S = 0.5 * (VhgV - VhgV.mH)
        E = w.unsqueeze(-2) - w.unsqueeze(-1)
        F = E / (E * E + eta)
        F.diagonal(dim1=-2, dim2=-1).zero_()
        inner = S * F.to(V.dtype)
    if gw is not None:
        inner = inner + torch.diag_embed(gw.to(V.dtype))
    return V @ inner @ V.mH


class GramEigh(torch.autograd.Function):
    """(w, V) of a Hermitian positive semidefinite G (n x n complex128 on the device), eigenvalues
    descending: forward by the library's fp64 Jacobi solver (mpdo_eigh), backward by
    ``eigh_backward`` with broadening eta = (1e-12 w_max)^2."""

    @staticmethod
    def forward(ctx, G: torch.Tensor):
        w, V = _lib.eigh(G.detach().unsqueeze(0))
        w, V = w[0], V[0]
        ctx.save_for_backward(w, V)
        return w, V

    @staticmethod
    def backward(ctx, gw, gV):
        w, V = ctx.saved_tensors
        wmax = float(w.abs().max()) if w.numel() else 0.0
        eta = torch.tensor((BROADEN * wmax) ** 2 + 1e-300, dtype=torch.float64, device=w.device)
        return eigh_backward(w, V, gw, gV, eta)


def _gram_eigh(G: torch.Tensor):
    G = 0.5 * (G + G.mH)        # exactly Hermitian input (the cotangent is Hermitian as well)
    return GramEigh.apply(G)


def _floor_mask(w: torch.Tensor, rel: float) -> torch.Tensor:
    w0 = float(w[0].detach()) if w.numel() else 0.0
    return (w > rel * w0) if w0 > 0 else torch.zeros_like(w, dtype=torch.bool)


def _pow(w: torch.Tensor, p: float, mask: torch.Tensor) -> torch.Tensor:
    safe = torch.where(mask, w, torch.ones_like(w))
    return torch.where(mask, safe.clamp_min(1e-300) ** p, torch.zeros_like(w))


def keep_count(w: Sequence[float], length: int, max_sv: Optional[int], err: Optional[float], relative: bool) -> int:
    """decompositions.py:116-140 on the singular values s = sqrt(w) padded with zeros to `length`:
    keep min(max_sv, k_err), k_err the smallest count whose tail te[-1] - te[k-1] <= err
    (te = sqrt(cumsum s^2), err x s_0 when relative)."""
    n = length
    maxk = n if (max_sv is None or max_sv <= 0 or max_sv > n) else max_sv
    kerr = maxk
    if err is not None and n > 0:
        vals = [max(float(x), 0.0) for x in w][:n]
        total = sum(vals)
        te_last = math.sqrt(total)
        bound = err * math.sqrt(vals[0]) if (relative and vals) else err
        cum = 0.0
        for i in range(n - 1):
            cum += vals[i] if i < len(vals) else 0.0
            if te_last - math.sqrt(cum) <= bound:
                kerr = i + 1
                break
    return min(kerr, maxk)


# ------------------------------------------------------------------------------------------------
def qr_gram(M: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """M (rows x cols) = Q R with Q^H Q = I on M's range (decompositions.py:149-195 up to the gauge
    of Q, which no gauge-invariant output sees): Q = M W w^-1/2, R = w^1/2 W^H from the Gram
    M^H M = W diag(w) W^H; directions below 1e-13 of the largest eigenvalue are zero columns of
    Q and zero rows of R. rows <= cols: Q = I, R = M (the per-gate path's rule)."""
    rows, cols = M.shape
    if rows <= cols:
        return torch.eye(rows, dtype=M.dtype, device=M.device), M
    w, W = _gram_eigh(M.mH @ M)
    mask = _floor_mask(w, QR_FLOOR)
    Q = M @ (W * _pow(w, -0.5, mask).to(M.dtype)[None, :])
    R = _pow(w, 0.5, mask).to(M.dtype)[:, None] * W.mH
    return Q, R


def sqrt_split(A: torch.Tensor, B: torch.Tensor, max_sv: Optional[int], err: Optional[float], relative: bool,
               canonical: bool = False):
    """tn.split_node of theta = A B (A: rows x D, B: D x cols; decompositions.svd's keep rule):
    U sqrt(S) and sqrt(S) V^H. With A = Q_A R_A the left singular vectors of theta are Q_A U_c,
    U_c, s^2 the eigenpairs of R_A (B B^H) R_A^H; sqrt(S) V^H = S^-1/2 U_c^H R_A B. `canonical`:
    A has orthonormal columns (after the QR sweep), R_A = I. Returns (left, right, kept k,
    discarded sum s_rest^2)."""
    rows, cols = A.shape[0], B.shape[1]
    if canonical:
        QA, RB = A, B
    else:
        QA, RA = qr_gram(A)
        RB = RA @ B
    w, U = _gram_eigh(RB @ RB.mH)
    length = min(rows, cols)
    wl = w.detach().cpu().tolist()
    k = keep_count(wl, length, max_sv, err, relative)
    disc = float(sum(max(x, 0.0) for x in wl[k:]))
    kk = min(k, w.shape[0])
    mask = _floor_mask(w, SPLIT_FLOOR)[:kk]
    wk, Uk = w[:kk], U[:, :kk]
    left = QA @ (Uk * _pow(wk, 0.25, mask).to(A.dtype)[None, :])
    right = (_pow(wk, -0.25, mask).to(A.dtype)[:, None] * Uk.mH) @ RB
    if k > kk:        # kept zero singular values (rank below the count): zero vectors
        left = torch.cat([left, left.new_zeros(rows, k - kk)], 1)
        right = torch.cat([right, right.new_zeros(k - kk, cols)], 0)
    return left, right, k, disc


def kappa_cut(T: torch.Tensor, kappa: int):
    """TNNOptimizer.py:164-197 for one site T[l, 2, a, r] with a > kappa: M = [(l s r) x a],
    keep min(kappa, len) directions, T <- U S = M V_k (V^H dropped); V_k the leading eigenvectors
    of M^H M. Returns (T', discarded weight)."""
    l, _, a, r = T.shape
    M = T.permute(0, 1, 3, 2).reshape(l * 2 * r, a)
    w, V = _gram_eigh(M.mH @ M)
    length = min(l * 2 * r, a)
    wl = w.detach().cpu().tolist()
    k = keep_count(wl, length, kappa, None, True)
    kk = min(k, a)
    Tk = M @ V[:, :kk]
    if k > kk:
        Tk = torch.cat([Tk, Tk.new_zeros(Tk.shape[0], k - kk)], 1)
    disc = float(sum(max(x, 0.0) for x in wl[k:]))
    return Tk.reshape(l, 2, r, k).permute(0, 1, 3, 2), disc


# ------------------------------------------------------------------------------------------------
def apply_1q(T: torch.Tensor, E: torch.Tensor) -> torch.Tensor:
    """Circuit.py:138-178: T'[l, s', (a, k), r] = sum_s E[s', s, k] T[l, s, a, r] (old inner index major)."""
    l, _, a, r = T.shape
    K = E.shape[2]
    out = torch.einsum('ijk,ljar->liakr', E, T)
    return out.reshape(l, 2, a * K, r)


def apply_2q(Tmin: torch.Tensor, Tmax: torch.Tensor, G: torch.Tensor):
    """Circuit.py:74-136: theta[(l s1 a1), (s2 a2 k r)] = sum G[s1 s2 i1 i2 k] Tmin[l i1 a1 m] Tmax[m i2 a2 r]
    as A B with A = delta (x) Tmin (rows (l s1 a1), cols (s1' i1 m)) and B = the channel applied to
    Tmax (rows (s1' i1 m), cols (s2 a2 k r)); split with no count cap and absolute error
    GLOBAL_MINIMUM. Returns (new min, new max, discarded weight)."""
    l, _, a1, m = Tmin.shape
    _, _, a2, r = Tmax.shape
    K = G.shape[4]
    eye = torch.eye(2, dtype=Tmin.dtype, device=Tmin.device)
    A = torch.einsum('xz,luam->lxazum', eye, Tmin).reshape(l * 2 * a1, 4 * m)
    B = torch.einsum('xyuvk,mvbr->xumybkr', G, Tmax).reshape(4 * m, 2 * a2 * K * r)
    left, right, k, disc = sqrt_split(A, B, None, GLOBAL_MINIMUM, False)
    return left.reshape(l, 2, a1, k), right.reshape(k, 2, a2 * K, r), disc


def bond_truncate(sites: List[torch.Tensor], chi: int) -> List[float]:
    """TNNOptimizer.py:72-134: QR sweep left to right, then the chi sweep right to left with the
    sqrt(S) split (left tensor orthonormal after the QR sweep). Returns the chi weights."""
    n = len(sites)
    for i in range(n - 1):
        T = sites[i]
        l, _, a, r = T.shape
        Q, R = qr_gram(T.reshape(l * 2 * a, r))
        sites[i] = Q.reshape(l, 2, a, Q.shape[1])
        sites[i + 1] = torch.tensordot(R, sites[i + 1], dims=([1], [0]))
    disc = []
    for idx in range(n - 1, 0, -1):
        A, B = sites[idx - 1], sites[idx]
        l, _, a1, m = A.shape
        _, _, a2, r = B.shape
        left, right, k, d = sqrt_split(A.reshape(l * 2 * a1, m), B.reshape(m, 2 * a2 * r), chi, None, True,
                                       canonical=True)
        sites[idx - 1] = left.reshape(l, 2, a1, k)
        sites[idx] = right.reshape(k, 2, a2, r)
        disc.append(d)
    return disc


# ------------------------------------------------------------------------------------------------
_PAULI = None


def _paulis(dev):
    global _PAULI
    if _PAULI is None or _PAULI.device != dev:
        _PAULI = torch.tensor([[[1, 0], [0, 1]], [[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]],
                              dtype=torch.complex128, device=dev)
    return _PAULI


def ry(t):
    c, s = torch.cos(t / 2), torch.sin(t / 2)
    return torch.stack([torch.stack([c, -s]), torch.stack([s, c])]).to(torch.complex128)


def rz(t):
    e = torch.exp(-0.5j * t.to(torch.complex128))
    z = torch.zeros((), dtype=torch.complex128, device=t.device)
    return torch.stack([torch.stack([e, z]), torch.stack([z, e.conj()])])


def depolarized_1q(U: torch.Tensor, p_first: torch.Tensor, p_second: torch.Tensor) -> torch.Tensor:
    """Kraus tensor (2, 2, 4) of U = G2 G1 with a depolarizing channel (NoiseChannel.py:51-91) of
    strength p_first after G1 and p_second after G2: depolarizing commutes with unitary
    conjugation and two of them compose to one of strength 1 - (1 - p1)(1 - p2), so the channel is
    U followed by that one depolarizing channel: E_k = sqrt(w_k) P_k U with w_0 = 1 - 3 p / 4,
    w_1..3 = p / 4 (the reference's 16 operators of the two channels span the same Choi matrix)."""
    p = 1 - (1 - p_first) * (1 - p_second)
    w = torch.stack([1 - 0.75 * p, 0.25 * p, 0.25 * p, 0.25 * p])
    P = _paulis(U.device)
    E = torch.sqrt(w).to(torch.complex128)[:, None, None] * (P @ U)
    return E.permute(1, 2, 0)


def depolarized_cnot(p: torch.Tensor) -> torch.Tensor:
    """(2, 2, 2, 2, 16) tensor G[o0, o1, i0, i1, k] = sqrt(w_k) (P_k CNOT), P_k the two-qubit Pauli
    strings (first qubit major), w_0 = 1 - 15 p / 16, else p / 16 (NoiseChannel.py:51-99)."""
    dev = p.device
    P = _paulis(dev)
    P2 = torch.einsum('aij,bkl->abikjl', P, P).reshape(16, 4, 4)
    cnot = torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=torch.complex128, device=dev)
    w = torch.cat([(1 - 15 * p / 16).reshape(1), (p / 16).expand(15)])
    E = torch.sqrt(w).to(torch.complex128)[:, None, None] * (P2 @ cnot)
    return E.reshape(16, 2, 2, 2, 2).permute(1, 2, 3, 4, 0)


# ------------------------------------------------------------------------------------------------
class HEA(torch.nn.Module):
    """Hardware-efficient ansatz with learnable angles and per-gate depolarizing strengths:
    `depth` layers of RY(theta) RZ(phi) on every qubit, each followed by a depolarizing channel
    of its own strength, then CNOT(0,1), CNOT(1,2), ..., CNOT(n-2, n-1), each followed by a
    two-qubit depolarizing channel of its own strength, then a truncation (chi, kappa).
    Parameters: theta, phi (depth, n); p1 (depth, n, 2) (RY, RZ); p2 (depth, n - 1)."""

    def __init__(self, n: int, depth: int, chi: int, kappa: int, theta, phi, p1, p2, device='cuda:0',
                 checkpoint_layers: bool = True):
        super().__init__()
        dev = torch.device(device)
        f = lambda x: torch.nn.Parameter(torch.as_tensor(x, dtype=torch.float64, device=dev).clone())
        self.n, self.depth, self.chi, self.kappa = n, depth, chi, kappa
        self.theta, self.phi, self.p1, self.p2 = f(theta), f(phi), f(p1), f(p2)
        self.checkpoint_layers = checkpoint_layers
        self.discarded: List[Tuple[str, int, float]] = []
        self.device = dev

    def _layer(self, d: int, *sites):
        sites = list(sites)
        n = self.n
        for q in range(n):
            U = rz(self.phi[d, q]) @ ry(self.theta[d, q])
            E = depolarized_1q(U, self.p1[d, q, 0], self.p1[d, q, 1])
            sites[q] = apply_1q(sites[q], E)
        disc = []
        for q in range(n - 1):
            G = depolarized_cnot(self.p2[d, q])
            sites[q], sites[q + 1], dg = apply_2q(sites[q], sites[q + 1], G)
            disc.append(('gate', q, dg))
        for i, w in enumerate(bond_truncate(sites, self.chi)):
            disc.append(('chi', n - 1 - i, w))
        for q in range(n):
            if sites[q].shape[2] > self.kappa:
                sites[q], dk = kappa_cut(sites[q], self.kappa)
                disc.append(('kappa', q, dk))
        self._disc_layer = disc
        return tuple(sites)

    def forward(self) -> List[torch.Tensor]:
        """Evolve |0...0> through every layer; returns the final site tensors (complex128)."""
        v = torch.zeros(1, 2, 1, 1, dtype=torch.complex128, device=self.device)
        v[0, 0, 0, 0] = 1
        sites = tuple(v.clone() for _ in range(self.n))
        self.discarded = []
        for d in range(self.depth):
            if self.checkpoint_layers and torch.is_grad_enabled():
                sites = checkpoint(self._layer, d, *sites, use_reentrant=False)
            else:
                sites = self._layer(d, *sites)
            self.discarded.extend(self._disc_layer)
        return list(sites)


def tfim_energy(sites: Sequence[torch.Tensor], h: float = 1.0) -> torch.Tensor:
    """Tr(H rho~) (unnormalised, dmOperations.py:135-171) of H = -sum Z_i Z_i+1 - h sum X_i by transfer
    contractions: left environments L_i of the identity-traced chain, right environments R_i, and
    each term contracted through its one or two sites."""
    n = len(sites)
    dev = sites[0].device
    P = _paulis(dev)
    eye = torch.ones(1, 1, dtype=torch.complex128, device=dev)

    def step(E, T, O=None):
        # E'[r, r'] = sum E[l, l'] T[l, s, a, r] conj(O T)[l', s, a, r']
        X = torch.tensordot(E, T, dims=([0], [0]))                  # [l', s, a, r]
        Tc = T.conj() if O is None else torch.einsum('ts,lsar->ltar', O.conj(), T.conj())
        return torch.tensordot(X, Tc, dims=([0, 1, 2], [0, 1, 2]))   # [r, r']

    Ls = [eye]
    for T in sites:
        Ls.append(step(Ls[-1], T))
    Rs = [None] * (n + 1)
    Rs[n] = eye
    for i in range(n - 1, -1, -1):
        T = sites[i]
        # R_i[l, l'] = sum T[l, s, a, r] conj(T)[l', s, a, r'] R_{i+1}[r, r']
        X = torch.tensordot(T, Rs[i + 1], dims=([3], [0]))           # [l, s, a, r']
        Rs[i] = torch.tensordot(X, T.conj(), dims=([1, 2, 3], [1, 2, 3]))
    energy = torch.zeros((), dtype=torch.float64, device=dev)
    for i in range(n - 1):
        E = step(step(Ls[i], sites[i], P[3]), sites[i + 1], P[3])
        energy = energy - (E * Rs[i + 2]).sum().real
    for i in range(n):
        E = step(Ls[i], sites[i], P[1])
        energy = energy - h * (E * Rs[i + 1]).sum().real
    return energy
