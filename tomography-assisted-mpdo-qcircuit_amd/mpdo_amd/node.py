"""Site node: the fixed-layout replacement of one tensornetwork ``Node`` of the MPDO.

The reference names edges ``bond_{i-1}_{i}``, ``physics_i``, ``I_i``, ``bond_{i}_{i+1}``
(Tools.py:44-72) and finds them by name after every operation. Here every site is one
complex64 tensor ``T[l, 2, a, r]`` on the device with absent edges as dimension 1, plus the two
presence flags on which the reference's behaviour depends: the inner index ``I_i``
(Circuit.py:83,140, TNNOptimizer.py:176-181) and the bonds (TNNOptimizer.py:51-61).
"""
from __future__ import annotations

from typing import List, Optional

import torch


class SiteNode:
    __slots__ = ('tensor', 'name', 'index', 'has_I', 'bond_left', 'bond_right')

    def __init__(self, tensor: torch.Tensor, index: int, name: Optional[str] = None, has_I: bool = False,
                 bond_left: bool = False, bond_right: bool = False):
        if tensor.dim() == 1:
            tensor = tensor.reshape(1, 2, 1, 1)
        if tensor.dim() != 4 or tensor.shape[1] != 2:
            raise ValueError(f'site tensor must be (l, 2, a, r), got {tuple(tensor.shape)}')
        self.tensor = tensor
        self.index = index
        self.name = name or f'qubit_{index}'
        self.has_I = has_I
        self.bond_left = bond_left
        self.bond_right = bond_right

    @property
    def axis_names(self) -> List[str]:
        i = self.index
        names = []
        if self.bond_left:
            names.append(f'bond_{i - 1}_{i}')
        names.append(f'physics_{i}')
        if self.has_I:
            names.append(f'I_{i}')
        if self.bond_right:
            names.append(f'bond_{i}_{i + 1}')
        return names

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    def copy(self) -> 'SiteNode':
        return SiteNode(self.tensor.clone(), self.index, self.name, self.has_I, self.bond_left, self.bond_right)

    def set_tensor(self, t: torch.Tensor) -> None:
        self.tensor = t

    def __repr__(self):
        return f'SiteNode({self.name}, shape={self.shape}, axes={self.axis_names})'


def replicate_nodes(nodes: List[SiteNode]) -> List[SiteNode]:
    return [n.copy() for n in nodes]
