"""MI355X (gfx950) implementation of the MPDO noisy-circuit update path of
WeiguoMa/Tomography-assisted-MPDO-QCircuit (package ``MPDOSimulator``), with the same public
names: ``TensorCircuit``, ``Tools``, ``dmOperations`` and ``TNNOptimizer``.

The arithmetic runs in the hand-written HIP library ``libmpdo_gfx950.so`` (C ABI in
``include/mpdo.h``); this package is the host side that mirrors the reference's API.
"""
from . import Tools
from . import TNNOptimizer
from . import dmOperations
from .circuit import TensorCircuit, QuantumCircuit
from .node import SiteNode

__version__ = '0.1.0'

__all__ = ['TensorCircuit', 'QuantumCircuit', 'Tools', 'dmOperations', 'TNNOptimizer', 'SiteNode']
