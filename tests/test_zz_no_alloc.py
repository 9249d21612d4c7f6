"""mpdo_pauli_expect_noalloc (the Pauli-string expectations through the library's fixed device
scratch, no allocation) against mpdo_pauli_expect through the C ABI: two calls in a row, the
second on a smaller chain (the scratch keeps the first call's contents), Pauli strings and
projectors, within 1e-6 of Tr rho~ (complex64 environments, the same kernels; the two entry
points may group the strings differently); a chain whose strings do not fit the fixed scratch is
refused before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _codes(n, count, seed):
    rng = np.random.default_rng(seed)
    c = rng.integers(-1, 5, size=(count, n)).astype(np.int32)
    c[0] = -1                     # the identity string: Tr rho~
    return c


def _call(fn, sites, codes):
    from mpdo_amd import _lib
    n = len(sites)
    flat = np.ascontiguousarray(codes, dtype=np.int32)
    out = np.zeros(codes.shape[0], dtype=np.float64)
    st = fn(n, _lib._sites(sites), codes.shape[0], flat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _lib._stream(sites[0].device))
    return st, out


def test_noalloc_matches_pauli_expect(cuda):
    from mpdo_amd import _lib, batched
    from mpdo_amd import workloads as W
    lib = _lib.load()
    for n, chi, kappa, count, seed in [(12, 64, 4, 300, 1), (6, 16, 2, 40, 2)]:
        dims = batched.steady_dims(n, chi, kappa, 4, 16, 0)
        sites = [t[0].contiguous() for t in W.steady_chains_device(dims, 1, cuda, seed=seed)]
        codes = _codes(n, count, seed)
        st_a, ref = _call(lib.mpdo_pauli_expect, sites, codes)
        assert st_a == 0
        st_b, got = _call(lib.mpdo_pauli_expect_noalloc, sites, codes)
        assert st_b == 0, lib.mpdo_last_error()
        tr = abs(ref[0])
        assert tr > 0
        assert np.abs(got - ref).max() <= 1e-6 * tr, (n, np.abs(got - ref).max() / tr)


def test_noalloc_refuses_strings_beyond_the_scratch(cuda):
    from mpdo_amd import _lib
    lib = _lib.load()
    # one site of 1 x 2 x 4 x 2^22 complex64 (32 M elements x 2 buffers): beyond 32 MiB per string
    big = torch.zeros(1, 2, 4, 1 << 22, dtype=torch.complex64, device=cuda)
    last = torch.zeros(1 << 22, 2, 1, 1, dtype=torch.complex64, device=cuda)
    st, _ = _call(lib.mpdo_pauli_expect_noalloc, [big, last], _codes(2, 3, 3))
    assert st == _lib.MPDO_ENOMEM
    assert b'fixed scratch' in lib.mpdo_last_error()
