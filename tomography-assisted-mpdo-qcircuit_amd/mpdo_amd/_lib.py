"""ctypes binding of libmpdo_gfx950.so (the C ABI in include/mpdo.h).

This is the only route from the Python host to the device: there is no fallback. Loading fails
loudly when the shared library is missing, and every compute call requires complex64 tensors
that live on a ROCm device.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmpdo_gfx950.so')

MPDO_OK, MPDO_EINVAL, MPDO_ENOCONV, MPDO_ENOMEM, MPDO_EHIP = range(5)


class Site(ctypes.Structure):
    _fields_ = [('data', ctypes.c_void_p), ('l', ctypes.c_int), ('a', ctypes.c_int), ('r', ctypes.c_int)]


_lib = None
_load_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load the HIP library (once). Raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _load_lock:
        if _lib is not None:
            return _lib
        return _load()


def _load() -> ctypes.CDLL:
    global _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is missing: build it with `python -m mpdo_amd._build` '
                          '(hipcc --offload-arch=gfx950); there is no CPU fallback')
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D, VP = ctypes.POINTER, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    SP = P(Site)
    sig = {
        'mpdo_version': (I, []),
        'mpdo_bind_workspace': (I, [VP, VP, ctypes.c_size_t]),
        'mpdo_gemm_tune_record': (I, [I]),
        'mpdo_gemm_tune_run': (I, [I, VP]),
        'mpdo_gemm_tune_get': (I, [P(I), I]),
        'mpdo_gemm_tune_set': (I, [P(I), I]),
        'mpdo_workspace_needed': (ctypes.c_size_t, []),
        'mpdo_layer_workspace_bytes': (ctypes.c_size_t, [I, I, SP, I, I, I, I, I]),
        'mpdo_last_error': (ctypes.c_char_p, []),
        'mpdo_release_workspace': (None, []),
        'mpdo_set_math_mode': (I, [I]),
        'mpdo_apply_1q': (I, [I, SP, SP, P(ctypes.c_float), P(I), VP]),
        'mpdo_apply_2q_capacity': (I, [SP, SP, I]),
        'mpdo_apply_2q': (I, [SP, SP, P(ctypes.c_float), I, D, SP, SP, P(D), VP]),
        'mpdo_qr_left2right_dims': (None, [I, SP, SP]),
        'mpdo_qr_left2right': (I, [I, SP, SP, VP]),
        'mpdo_svd_sweep_dims': (None, [I, SP, I, I, SP]),
        'mpdo_svd_sweep': (I, [I, SP, SP, I, I, D, I, P(D), VP]),
        'mpdo_svd_kappa': (I, [I, SP, SP, I, D, P(D), VP]),
        'mpdo_pauli_expect': (I, [I, SP, I, P(I), P(D), VP]),
        'mpdo_pauli_expect_noalloc': (I, [I, SP, I, P(I), P(D), VP]),
        'mpdo_reduced_dm': (I, [I, SP, P(I), VP, VP]),
        'mpdo_trace_rho_rho': (I, [I, SP, SP, P(D), VP]),
        'mpdo_trace_rho_rho_masked': (I, [I, SP, SP, P(I), P(D), VP]),
        'mpdo_contract_state': (I, [I, SP, VP, VP]),
        'mpdo_eigh': (I, [I, I, VP, P(D), VP, VP]),
        'mpdo_env_right_elems': (ctypes.c_longlong, [I, SP]),
        'mpdo_env_right': (I, [I, SP, VP, VP]),
        'mpdo_sample_step': (I, [SP, I, VP, I, P(I), I, VP, VP, P(D), VP]),
        'mpdo_layer_dims': (I, [I, SP, I, I, I, I, I, SP]),
        'mpdo_layer_batched': (I, [I, I, SP, P(ctypes.c_float), I, I, P(ctypes.c_float), I, I, I, SP, P(D), VP]),
        'mpdo_debug_gemm': (I, [I, I, I, I, I, VP, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, I,
                                VP, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, I,
                                VP, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                D, D, D, D, VP]),
        'mpdo_debug_gram_split': (I, [I, I, I, I, VP, VP, VP, VP]),
        'mpdo_debug_gemm_ex': (I, [I, I, I, I, I, I, I, I, I, VP, ctypes.c_longlong, ctypes.c_longlong,
                                   ctypes.c_longlong, I, VP, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                   I, VP, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, D, D, D, D, VP]),
        'mpdo_debug_eigh': (I, [I, I, VP, P(D), VP, P(I), VP]),
        'mpdo_debug_eigh_gram': (I, [I, I, VP, P(D), VP, P(I), VP]),
        'mpdo_debug_layer_dry_run': (I, [I, I, SP, I, I, I, I, I, SP, P(ctypes.c_ulonglong)]),
        'mpdo_debug_prof_enable': (I, [I]),
        'mpdo_debug_fail_eig': (None, [I]),
        'mpdo_debug_layer_gate_skip': (I, [I]),
        'mpdo_debug_layer_gate_skips': (I, []),
        'mpdo_debug_layer_gate_tail': (I, [I]),
        'mpdo_debug_layer_gate_tails': (I, []),
        'mpdo_debug_layer_tail_shape': (I, [I, I]),
        'mpdo_debug_layer_chi_top': (I, [I]),
        'mpdo_debug_layer_chi_tops': (I, []),
        'mpdo_debug_layer_top_shape': (I, [I, I]),
        'mpdo_debug_gram_top_eig': (I, [I, I, VP, I, I, I, VP, VP, P(I), VP]),
        'mpdo_debug_gram_tail_cut': (I, [I, I, VP, D, I, I, P(I), P(D), VP, P(I), VP]),
        'mpdo_debug_layer_kept': (I, [P(I), I]),
        'mpdo_debug_layer_gate_eigs': (I, [I, P(D), I]),
        'mpdo_debug_layer_chi_eigs': (I, [P(D), I]),
        'mpdo_debug_layer_gemm_log': (I, [P(I), I]),
        'mpdo_debug_gemm_tiles': (I, [I]),
        'mpdo_debug_gemm_bf16_tiles': (I, [I]),
        'mpdo_debug_chol': (I, [I, I, VP, ctypes.c_double, VP, P(I), VP]),
        'mpdo_debug_bj_threads': (I, [I]),
        'mpdo_debug_bj_fused_max': (I, [I]),
        'mpdo_debug_prof_eig_log': (I, [P(I), I]),
        'mpdo_debug_prof_shapes': (I, [I, P(I), P(D), I]),
        'mpdo_debug_prof_read': (I, [I, P(ctypes.c_longlong), P(D), P(D), P(D)]),
        'mpdo_debug_prof_read_executed': (I, [I, P(D), P(D)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _load_tile_table(lib)
    _lib = lib
    return lib


# The GEMM tile autotuner's table for this architecture (mpdo_amd/autotune.py writes it): loaded
# into the library when it is loaded; absent, the library's built-in tile rules apply.
TILE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'gemm_tiles_gfx950.json')


def _load_tile_table(lib) -> int:
    import json
    if not os.path.exists(TILE_TABLE):
        return 0
    with open(TILE_TABLE) as f:
        entries = json.load(f)['entries']
    flat = (ctypes.c_int * max(13 * len(entries), 1))(*[int(v) for e in entries for v in list(e['key']) + [e['tile']]])
    return int(lib.mpdo_gemm_tune_set(flat, len(entries)))


class MPDOError(RuntimeError):
    pass


def check(status: int, what: str) -> None:
    if status == MPDO_OK:
        return
    msg = f'{what}: {load().mpdo_last_error().decode(errors="replace")}'
    if status == MPDO_ENOCONV:
        raise torch.linalg.LinAlgError(msg)
    if status == MPDO_EINVAL:
        raise ValueError(msg)
    if status == MPDO_ENOMEM:
        raise MemoryError(msg)
    raise MPDOError(msg)


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# Workspace: one buffer from torch's caching allocator per (device, stream), bound to the stream
# (mpdo_bind_workspace) so that the library allocates nothing on the compute path and calls on
# different streams use different buffers.
_WS_MIN = 64 << 20
_ws_bufs = {}
_ws_lock = threading.Lock()


def _bind_workspace(device: torch.device, nbytes: int) -> None:
    handle = torch.cuda.current_stream(device).cuda_stream
    key = (device.index if device.index is not None else torch.cuda.current_device(), handle)
    with _ws_lock:
        buf = _ws_bufs.get(key)
        if buf is not None and buf.numel() >= nbytes:
            return
        size = max(int(nbytes * 1.125) + (1 << 20), _WS_MIN)
        buf = torch.empty(size, dtype=torch.uint8, device=device)
        check(load().mpdo_bind_workspace(ctypes.c_void_p(handle), ctypes.c_void_p(buf.data_ptr()), size),
              'mpdo_bind_workspace')
        _ws_bufs[key] = buf


def _run(what: str, device: torch.device, call) -> None:
    """Run one library call on the current stream of `device` with the stream's workspace bound;
    a call that needs a larger workspace (MPDO_ENOMEM before any launch, with the size in
    mpdo_workspace_needed) runs again with one that large."""
    lib = load()
    _bind_workspace(device, 0)
    st = call()
    if st == MPDO_ENOMEM:
        need = int(lib.mpdo_workspace_needed())
        if need > 0:
            _bind_workspace(device, need)
            st = call()
    check(st, what)


def release_workspaces() -> None:
    """Unbind and drop every workspace buffer this module holds."""
    lib = load()
    with _ws_lock:
        for (dev, handle), buf in list(_ws_bufs.items()):
            with torch.cuda.device(dev):
                lib.mpdo_bind_workspace(ctypes.c_void_p(handle), None, 0)
        _ws_bufs.clear()


def release_stream_workspace(device: torch.device) -> None:
    """Unbind and drop the workspace of the current stream of `device` (after the stream's work
    that used it has finished): its memory returns to torch's caching allocator."""
    handle = torch.cuda.current_stream(device).cuda_stream
    key = (device.index if device.index is not None else torch.cuda.current_device(), handle)
    with _ws_lock:
        if _ws_bufs.pop(key, None) is not None:
            load().mpdo_bind_workspace(ctypes.c_void_p(handle), None, 0)


def _require_device(t: torch.Tensor) -> None:
    if t.device.type != 'cuda':
        raise RuntimeError('the MPDO kernels run on a ROCm GPU (gfx950); got a tensor on '
                           f'{t.device} - there is no CPU path')
    if t.dtype != torch.complex64:
        raise TypeError(f'site tensors must be complex64, got {t.dtype}')


def site_desc(t: torch.Tensor) -> Site:
    """Descriptor of a (l, 2, a, r) complex64 device tensor (made contiguous by the caller)."""
    _require_device(t)
    if t.dim() != 4 or t.shape[1] != 2:
        raise ValueError(f'site tensor must have shape (l, 2, a, r), got {tuple(t.shape)}')
    if not t.is_contiguous():
        raise ValueError('site tensor must be contiguous')
    return Site(t.data_ptr(), t.shape[0], t.shape[2], t.shape[3])


def _sites(ts: Sequence[torch.Tensor]):
    arr = (Site * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = site_desc(t)
    return arr


def _empty(n_elems: int, device) -> torch.Tensor:
    return torch.empty(max(int(n_elems), 1), dtype=torch.complex64, device=device)


def _view(buf: torch.Tensor, s: Site) -> torch.Tensor:
    n = s.l * 2 * s.a * s.r
    return buf[:n].view(s.l, 2, s.a, s.r)


MATH_MODES = {'fp64acc': 0, 'fp32acc': 1, 'fp32x3m': 2, 'bf16x3': 3, 'bf16x6': 4}


def set_math_mode(mode: str) -> str:
    """Arithmetic of the contractions whose results are complex64 site data: 'fp64acc' (default,
    fp64 matrix cores), 'fp32acc' (fp32 matrix cores, four real products), 'fp32x3m' (three real
    products), 'bf16x3' / 'bf16x6' (bf16 matrix cores on operands split into two / three bf16
    parts, fp32 accumulation). Returns the previous mode's name."""
    old = load().mpdo_set_math_mode(MATH_MODES[mode])
    return {v: k for k, v in MATH_MODES.items()}[old]


# ---------------------------------------------------------------------------------------------
def apply_1q(sites: Sequence[torch.Tensor], kraus: Sequence[torch.Tensor]) -> List[torch.Tensor]:
    """Kraus/gate absorption (Circuit.py:138-178). kraus[i]: (2, 2) gate or (2, 2, K) tensor."""
    lib = load()
    n = len(sites)
    if n == 0:
        return []
    dev = sites[0].device
    mats, Ks = [], []
    for E in kraus:
        E = E.detach().to(torch.complex64).cpu()
        if E.dim() == 2:
            E = E.unsqueeze(-1)
        if E.shape[:2] != (2, 2):
            raise ValueError(f'Kraus tensor must be (2, 2[, K]), got {tuple(E.shape)}')
        mats.append(E.permute(2, 0, 1).contiguous().view(torch.float32).reshape(-1))
        Ks.append(E.shape[2])
    outs = [torch.empty(t.shape[0], 2, t.shape[2] * K, t.shape[3], dtype=torch.complex64, device=dev)
            for t, K in zip(sites, Ks)]
    flat = torch.cat(mats).contiguous()
    Karr = (ctypes.c_int * n)(*Ks)
    _run('mpdo_apply_1q', dev, lambda: lib.mpdo_apply_1q(n, _sites(sites), _sites(outs),
                                                        ctypes.cast(flat.data_ptr(), ctypes.POINTER(ctypes.c_float)),
                                                        Karr, _stream(dev)))
    return outs


def apply_2q(tmin: torch.Tensor, tmax: torch.Tensor, gate: torch.Tensor, eps_abs: float):
    """Two-qubit gate/channel + sqrt(S) split (Circuit.py:74-136). gate: (2,2,2,2[,K]) in
    (o_min, o_max, i_min, i_max[, k]) order. Returns (new_min, new_max, discarded)."""
    lib = load()
    dev = tmin.device
    G = gate.detach().to(torch.complex64).cpu()
    if G.dim() == 4:
        G = G.unsqueeze(-1)
    if tuple(G.shape[:4]) != (2, 2, 2, 2):
        raise ValueError(f'two-qubit gate must be (2,2,2,2[,K]), got {tuple(G.shape)}')
    K = G.shape[4]
    flat = G.contiguous().view(torch.float32).reshape(-1).contiguous()
    a, b = site_desc(tmin), site_desc(tmax)
    cap = lib.mpdo_apply_2q_capacity(ctypes.byref(a), ctypes.byref(b), K)
    bmin = _empty(tmin.shape[0] * 2 * tmin.shape[2] * cap, dev)
    bmax = _empty(cap * 2 * tmax.shape[2] * K * tmax.shape[3], dev)
    om, oM = Site(bmin.data_ptr(), 0, 0, 0), Site(bmax.data_ptr(), 0, 0, 0)
    disc = ctypes.c_double(0.0)
    _run('mpdo_apply_2q', dev, lambda: lib.mpdo_apply_2q(ctypes.byref(a), ctypes.byref(b),
                                                        ctypes.cast(flat.data_ptr(), ctypes.POINTER(ctypes.c_float)),
                                                        K, float(eps_abs), ctypes.byref(om), ctypes.byref(oM),
                                                        ctypes.byref(disc), _stream(dev)))
    return _view(bmin, om).clone(), _view(bmax, oM).clone(), disc.value


def qr_left2right(sites: Sequence[torch.Tensor]) -> List[torch.Tensor]:
    lib = load()
    n = len(sites)
    dev = sites[0].device
    ins = _sites(sites)
    dims = (Site * n)()
    lib.mpdo_qr_left2right_dims(n, ins, dims)
    bufs = [_empty(d.l * 2 * d.a * d.r, dev) for d in dims]
    outs = (Site * n)()
    for i in range(n):
        outs[i] = Site(bufs[i].data_ptr(), dims[i].l, dims[i].a, dims[i].r)
    _run('mpdo_qr_left2right', dev, lambda: lib.mpdo_qr_left2right(n, ins, outs, _stream(dev)))
    return [_view(bufs[i], outs[i]) for i in range(n)]


def svd_sweep(sites: Sequence[torch.Tensor], direction: int, chi: Optional[int], err: Optional[float],
              left_canonical: bool) -> Tuple[List[torch.Tensor], List[float]]:
    lib = load()
    n = len(sites)
    dev = sites[0].device
    ins = _sites(sites)
    caps = (Site * n)()
    msv = int(chi) if chi is not None else 0
    lib.mpdo_svd_sweep_dims(n, ins, direction, msv, caps)
    bufs = [_empty(c.l * 2 * c.a * c.r, dev) for c in caps]
    outs = (Site * n)()
    for i in range(n):
        outs[i] = Site(bufs[i].data_ptr(), 0, 0, 0)
    disc = (ctypes.c_double * max(n - 1, 1))()
    _run('mpdo_svd_sweep', dev, lambda: lib.mpdo_svd_sweep(n, ins, outs, direction, msv,
                                                          float(err) if err is not None else -1.0,
                                                          1 if left_canonical else 0, disc, _stream(dev)))
    res = [_view(bufs[i], outs[i]) for i in range(n)]
    return res, [disc[i] for i in range(n - 1)]


def svd_kappa(sites: Sequence[torch.Tensor], kappa: Optional[int], err: Optional[float]):
    lib = load()
    n = len(sites)
    if n == 0:
        return [], []
    dev = sites[0].device
    ins = _sites(sites)
    bufs = []
    outs = (Site * n)()
    for i, t in enumerate(sites):
        a = t.shape[2]
        cap = min(kappa, a) if kappa is not None else a
        bufs.append(_empty(t.shape[0] * 2 * cap * t.shape[3], dev))
        outs[i] = Site(bufs[i].data_ptr(), 0, 0, 0)
    disc = (ctypes.c_double * n)()
    _run('mpdo_svd_kappa', dev, lambda: lib.mpdo_svd_kappa(n, ins, outs, int(kappa) if kappa is not None else 0,
                                                          float(err) if err is not None else -1.0, disc,
                                                          _stream(dev)))
    return [_view(bufs[i], outs[i]) for i in range(n)], [disc[i] for i in range(n)]


def pauli_expect(sites: Sequence[torch.Tensor], codes: Sequence[Sequence[int]], noalloc: bool = False) -> List[float]:
    """Re Tr(O rho~) per code string (dmOperations.py:174-201). noalloc: through
    mpdo_pauli_expect_noalloc (the library's fixed device scratch; no workspace)."""
    lib = load()
    n = len(sites)
    ns = len(codes)
    flat = (ctypes.c_int * (n * ns))()
    for t, c in enumerate(codes):
        if len(c) != n:
            raise ValueError('operator code string must cover every qubit')
        for i, v in enumerate(c):
            flat[t * n + i] = int(v)
    out = (ctypes.c_double * ns)()
    dev = sites[0].device
    arr = _sites(sites)
    if noalloc:
        check(lib.mpdo_pauli_expect_noalloc(n, arr, ns, flat, out, _stream(dev)), 'mpdo_pauli_expect_noalloc')
    else:
        _run('mpdo_pauli_expect', dev, lambda: lib.mpdo_pauli_expect(n, arr, ns, flat, out, _stream(dev)))
    return [out[i] for i in range(ns)]


def reduced_dm(sites: Sequence[torch.Tensor], keep: Sequence[bool]) -> torch.Tensor:
    lib = load()
    n = len(sites)
    k = sum(1 for x in keep if x)
    dev = sites[0].device
    out = torch.empty(2 ** k, 2 ** k, dtype=torch.complex64, device=dev)
    kf = (ctypes.c_int * n)(*[1 if x else 0 for x in keep])
    arr = _sites(sites)
    _run('mpdo_reduced_dm', dev, lambda: lib.mpdo_reduced_dm(n, arr, kf, ctypes.c_void_p(out.data_ptr()), _stream(dev)))
    return out


def trace_rho_rho(a: Sequence[torch.Tensor], b: Sequence[torch.Tensor], keep: Optional[Sequence[bool]] = None) -> float:
    """Re Tr(rho0 rho1); qubits with keep[i] False are traced within each state first."""
    lib = load()
    if len(a) != len(b):
        raise ValueError('Density matrices must have the same number of nodes.')
    out = ctypes.c_double(0.0)
    if keep is None:
        sa, sb = _sites(a), _sites(b)
        _run('mpdo_trace_rho_rho', a[0].device,
             lambda: lib.mpdo_trace_rho_rho(len(a), sa, sb, ctypes.byref(out), _stream(a[0].device)))
    else:
        kf = (ctypes.c_int * len(a))(*[1 if k else 0 for k in keep])
        sa, sb = _sites(a), _sites(b)
        _run('mpdo_trace_rho_rho_masked', a[0].device,
             lambda: lib.mpdo_trace_rho_rho_masked(len(a), sa, sb, kf, ctypes.byref(out), _stream(a[0].device)))
    return out.value


def env_right(sites: Sequence[torch.Tensor]) -> List[torch.Tensor]:
    """Traced tail environments R_0..R_n (complex128, R_i of shape (l_i, l_i), R_n = [[1]])."""
    lib = load()
    n = len(sites)
    arr = _sites(sites)
    dev = sites[0].device
    total = lib.mpdo_env_right_elems(n, arr)
    buf = torch.empty(total, dtype=torch.complex128, device=dev)
    _run('mpdo_env_right', dev, lambda: lib.mpdo_env_right(n, arr, ctypes.c_void_p(buf.data_ptr()), _stream(dev)))
    out, off = [], 0
    for t in sites:
        m = t.shape[0]
        out.append(buf[off:off + m * m].view(m, m))
        off += m * m
    out.append(buf[off:off + 1].view(1, 1))
    return out


def sample_step(site: torch.Tensor, prev: torch.Tensor, sel: Optional[Sequence[int]], sampled: bool,
                R_next: Optional[torch.Tensor] = None):
    """One site of the conditional sweep. prev: (m, l, l) complex128 environments; sel picks one
    per group (None: group g = prev[g]). Returns (out, probs): out (G, 2, r, r) and probs (G, 2)
    numpy float64 when sampled, else out (G, r, r) and None."""
    lib = load()
    if prev.dtype != torch.complex128 or prev.dim() != 3 or not prev.is_contiguous():
        raise ValueError('environments must be a contiguous (m, l, l) complex128 tensor')
    l, r = site.shape[0], site.shape[3]
    if prev.shape[1] != l or prev.shape[2] != l:
        raise ValueError(f'environment shape {tuple(prev.shape[1:])} does not match the site bond {l}')
    G = prev.shape[0] if sel is None else len(sel)
    dev = site.device
    if sampled:
        if R_next is None or tuple(R_next.shape) != (r, r) or R_next.dtype != torch.complex128:
            raise ValueError('a sampled site needs its (r, r) complex128 tail environment')
        out = torch.empty(G, 2, r, r, dtype=torch.complex128, device=dev)
    else:
        out = torch.empty(G, r, r, dtype=torch.complex128, device=dev)
    probs = np.zeros((G, 2), dtype=np.float64)
    csel = None if sel is None else (ctypes.c_int * G)(*[int(x) for x in sel])
    sd = site_desc(site)
    Rn = R_next.contiguous() if sampled else None
    _run('mpdo_sample_step', dev,
         lambda: lib.mpdo_sample_step(ctypes.byref(sd), G, ctypes.c_void_p(prev.data_ptr()), prev.shape[0], csel,
                                      1 if sampled else 0, ctypes.c_void_p(Rn.data_ptr()) if sampled else None,
                                      ctypes.c_void_p(out.data_ptr()),
                                      probs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if sampled else None,
                                      _stream(dev)))
    return out, (probs if sampled else None)


def layer_dims(dims: Sequence[Tuple[int, int, int]], K1: int, parity: int, K2: int, chi: Optional[int],
               kappa: Optional[int]):
    """Output (l, a, r) of mpdo_layer_batched, or None with the reason when the layer does not
    qualify for the fused path."""
    lib = load()
    n = len(dims)
    ins = (Site * n)()
    for i, (l, a, r) in enumerate(dims):
        ins[i] = Site(None, l, a, r)
    outs = (Site * n)()
    ok = lib.mpdo_layer_dims(n, ins, K1, parity, K2, int(chi or 0), int(kappa or 0), outs)
    if not ok:
        return None, lib.mpdo_last_error().decode(errors='replace')
    return [(o.l, o.a, o.r) for o in outs], ''


def layer_batched(sites: Sequence[torch.Tensor], kraus1: torch.Tensor, parity: int, gates2: torch.Tensor,
                  chi: Optional[int], kappa: Optional[int], out: Optional[Sequence[torch.Tensor]] = None):
    """One fused brickwork layer. sites[i]: (B, l, 2, a, r) complex64 on the device;
    kraus1: (B, n, K1, 2, 2); gates2: (npairs, 2, 2, 2, 2, K2). Returns (new sites, discarded
    (B, 3n) numpy: chi steps by bond, kappa cuts by site, gate-time splits by pair min site).
    out: optional contiguous complex64 device tensors of the new sites' shapes to write into."""
    lib = load()
    n = len(sites)
    B = sites[0].shape[0]
    dev = sites[0].device
    dims = [(t.shape[1], t.shape[3], t.shape[4]) for t in sites]
    K1, K2 = kraus1.shape[2], gates2.shape[-1]
    out_dims, why = layer_dims(dims, K1, parity, K2, chi, kappa)
    if out_dims is None:
        raise ValueError(f'layer does not qualify for the fused path: {why}')
    ins = (Site * n)()
    for i, t in enumerate(sites):
        _require_device(t)
        if not t.is_contiguous():
            raise ValueError('batched site tensors must be contiguous')
        ins[i] = Site(t.data_ptr(), dims[i][0], dims[i][1], dims[i][2])
    if out is None:
        outs_t = [torch.empty(B, l, 2, a, r, dtype=torch.complex64, device=dev) for (l, a, r) in out_dims]
    else:
        outs_t = list(out)
        for t, (l, a, r) in zip(outs_t, out_dims):
            _require_device(t)
            if tuple(t.shape) != (B, l, 2, a, r) or t.dtype != torch.complex64 or not t.is_contiguous():
                raise ValueError(f'output tensor {tuple(t.shape)} {t.dtype}: expected contiguous complex64 '
                                 f'{(B, l, 2, a, r)}')
        if len(outs_t) != n:
            raise ValueError(f'{len(outs_t)} output tensors for {n} sites')
    outs = (Site * n)()
    for i, (t, (l, a, r)) in enumerate(zip(outs_t, out_dims)):
        outs[i] = Site(t.data_ptr(), l, a, r)
    k1 = kraus1.detach().to(torch.complex64).cpu().contiguous().view(torch.float32).reshape(-1).contiguous()
    g2 = gates2.detach().to(torch.complex64).cpu().contiguous().view(torch.float32).reshape(-1).contiguous()
    disc = np.zeros((B, 3 * n), dtype=np.float64)
    need = int(lib.mpdo_layer_workspace_bytes(B, n, ins, K1, parity, K2, int(chi or 0), int(kappa or 0)))
    _bind_workspace(dev, need)
    st = lib.mpdo_layer_batched(B, n, ins, ctypes.cast(k1.data_ptr(), ctypes.POINTER(ctypes.c_float)), K1, parity,
                                ctypes.cast(g2.data_ptr(), ctypes.POINTER(ctypes.c_float)), K2, int(chi or 0),
                                int(kappa or 0), outs, disc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                _stream(dev))
    check(st, 'mpdo_layer_batched')
    return outs_t, disc


def eigh(A: torch.Tensor):
    """Batched Hermitian eigendecomposition (fp64 Jacobi on the device). A: (batch, n, n)
    complex128 Hermitian (not modified). Returns (w (batch, n) float64 descending, V (batch, n, n)
    complex128, column j the eigenvector of w[:, j])."""
    lib = load()
    if A.device.type != 'cuda' or A.dtype != torch.complex128 or A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError('eigh takes a (batch, n, n) complex128 tensor on the device')
    b, n, _ = A.shape
    work = A.contiguous().clone()
    w = torch.empty(b, n, dtype=torch.float64, device=A.device)
    V = torch.empty(b, n, n, dtype=torch.complex128, device=A.device)
    _run('mpdo_eigh', A.device, lambda: lib.mpdo_eigh(n, b, ctypes.c_void_p(work.data_ptr()),
                                                     ctypes.cast(w.data_ptr(), ctypes.POINTER(ctypes.c_double)),
                                                     ctypes.c_void_p(V.data_ptr()), _stream(A.device)))
    return w, V


def contract_state(sites: Sequence[torch.Tensor]) -> torch.Tensor:
    lib = load()
    n = len(sites)
    out = torch.empty(2 ** n, dtype=torch.complex64, device=sites[0].device)
    arr = _sites(sites)
    _run('mpdo_contract_state', sites[0].device,
         lambda: lib.mpdo_contract_state(n, arr, ctypes.c_void_p(out.data_ptr()), _stream(sites[0].device)))
    return out


# --------------------------------------------------------------------------------- test hooks
def debug_gemm(prec: int, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, M: int, N: int, K: int,
               sa: Tuple[int, int, int], sb: Tuple[int, int, int], sc: Tuple[int, int, int],
               conjA=False, conjB=False, alpha=1.0, beta=0.0, batch=1):
    lib = load()
    st = lib.mpdo_debug_gemm(prec, M, N, K, batch, ctypes.c_void_p(A.data_ptr()), sa[0], sa[1], sa[2], int(conjA),
                             ctypes.c_void_p(B.data_ptr()), sb[0], sb[1], sb[2], int(conjB),
                             ctypes.c_void_p(C.data_ptr()), sc[0], sc[1], sc[2],
                             complex(alpha).real, complex(alpha).imag, complex(beta).real, complex(beta).imag,
                             _stream(A.device))
    check(st, 'mpdo_debug_gemm')


def debug_gram_split(X: torch.Tensor, slices: int) -> torch.Tensor:
    """X (batch, n, K) complex64 on the device -> X X^H (batch, n, n) complex128 through the
    split-reduction Hermitian path."""
    lib = load()
    b, n, K = X.shape
    parts = torch.empty(b * slices * n * n, dtype=torch.complex128, device=X.device)
    G = torch.empty(b, n, n, dtype=torch.complex128, device=X.device)
    check(lib.mpdo_debug_gram_split(n, K, b, slices, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(parts.data_ptr()),
                                    ctypes.c_void_p(G.data_ptr()), _stream(X.device)), 'mpdo_debug_gram_split')
    return G


def debug_gemm_ex(prec: int, herm: bool, inA: int, inB: int, A: torch.Tensor, B: torch.Tensor, C: torch.Tensor,
                  M: int, N: int, K: int, sa: Tuple[int, int, int], sb: Tuple[int, int, int],
                  sc: Tuple[int, int, int], conjA=False, conjB=False, alpha=1.0, beta=0.0, batch=1, tri=0):
    """debug_gemm with a Hermitian-output flag and per-operand storage (0 = as prec, 1 = complex64,
    2 = complex128)."""
    lib = load()
    st = lib.mpdo_debug_gemm_ex(prec, int(herm), inA, inB, int(tri), M, N, K, batch, ctypes.c_void_p(A.data_ptr()), sa[0], sa[1],
                                sa[2], int(conjA), ctypes.c_void_p(B.data_ptr()), sb[0], sb[1], sb[2], int(conjB),
                                ctypes.c_void_p(C.data_ptr()), sc[0], sc[1], sc[2],
                                complex(alpha).real, complex(alpha).imag, complex(beta).real, complex(beta).imag,
                                _stream(A.device))
    check(st, 'mpdo_debug_gemm_ex')


def debug_eigh(A: torch.Tensor, gram: bool = False):
    """A: (batch, n, n) complex128 Hermitian on the device (destroyed). Returns (w, V, sweeps).
    gram=True: the positive-semidefinite solver of the fused layer (zero vectors past the rank)."""
    lib = load()
    b, n, _ = A.shape
    w = torch.empty(b, n, dtype=torch.float64, device=A.device)
    V = torch.empty(b, n, n, dtype=torch.complex128, device=A.device)
    sw = ctypes.c_int(0)
    fn = lib.mpdo_debug_eigh_gram if gram else lib.mpdo_debug_eigh
    _run('mpdo_debug_eigh', A.device,
         lambda: fn(n, b, ctypes.c_void_p(A.data_ptr()), ctypes.cast(w.data_ptr(), ctypes.POINTER(ctypes.c_double)),
                    ctypes.c_void_p(V.data_ptr()), ctypes.byref(sw), _stream(A.device)))
    return w, V, sw.value


def layer_dry_run(batch: int, dims: Sequence[Tuple[int, int, int]], K1: int, parity: int, K2: int, chi: Optional[int],
                  kappa: Optional[int]) -> int:
    """Host-only rehearsal of mpdo_layer_batched at these shapes (every launch's operand extents
    checked against its buffers, nothing launched). Returns the device workspace in bytes;
    raises MPDOError on a failed check or an unplannable layer."""
    lib = load()
    n = len(dims)
    out_dims, why = layer_dims(dims, K1, parity, K2, chi, kappa)
    if out_dims is None:
        raise ValueError(f'layer does not qualify for the fused path: {why}')
    ins, outs = (Site * n)(), (Site * n)()
    addr = 1 << 40
    for i, ((l, a, r), (lo, ao, ro)) in enumerate(zip(dims, out_dims)):
        ins[i] = Site(ctypes.c_void_p(addr), l, a, r)
        addr += ((8 * batch * l * 2 * a * r) | 4095) + 1
        outs[i] = Site(ctypes.c_void_p(addr), lo, ao, ro)
        addr += ((8 * batch * lo * 2 * ao * ro) | 4095) + 1
    ws = ctypes.c_ulonglong(0)
    check(lib.mpdo_debug_layer_dry_run(batch, n, ins, K1, parity, K2, int(chi or 0), int(kappa or 0), outs,
                                       ctypes.byref(ws)), 'mpdo_debug_layer_dry_run')
    return int(ws.value)


PROF_CLASSES = ('gemm', 'eig_solve', 'eig_rotate', 'eig_small', 'permute', 'eig_update')


def prof_enable(on: bool) -> None:
    """Start (reset) or stop recording HIP events around every launch, by kernel class."""
    check(load().mpdo_debug_prof_enable(1 if on else 0), 'mpdo_debug_prof_enable')


def prof_read() -> dict:
    """{class: (launches, ms, flops, bytes)} of the launches recorded since prof_enable(True)."""
    lib = load()
    out = {}
    for i, name in enumerate(PROF_CLASSES):
        n, ms, fl, by = ctypes.c_longlong(0), ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        check(lib.mpdo_debug_prof_read(i, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by)),
              'mpdo_debug_prof_read')
        out[name] = (int(n.value), float(ms.value), float(fl.value), float(by.value))
    return out


def prof_read_executed() -> dict:
    """{class: (flops, bytes)} the launches executed since prof_enable(True), counted from the
    device's skip flags (block-Jacobi classes; zero for the others)."""
    lib = load()
    out = {}
    for i, name in enumerate(PROF_CLASSES):
        fl, by = ctypes.c_double(0), ctypes.c_double(0)
        check(lib.mpdo_debug_prof_read_executed(i, ctypes.byref(fl), ctypes.byref(by)), 'mpdo_debug_prof_read_executed')
        out[name] = (float(fl.value), float(by.value))
    return out


def layer_gemm_log() -> List[Tuple[int, int, int, int, int, int]]:
    """(source line, M, N, K, batch, precision) of every GEMM of the last fused-layer dry run."""
    lib = load()
    n = lib.mpdo_debug_layer_gemm_log(None, 0)
    buf = (ctypes.c_int * max(6 * n, 1))()
    lib.mpdo_debug_layer_gemm_log(buf, 6 * n)
    return [tuple(buf[6 * i:6 * i + 6]) for i in range(n)]


def layer_kept() -> List[int]:
    """Per pair of the last fused layer on this thread: the largest kept count of its gate-time
    split over the chains."""
    buf = (ctypes.c_int * 1024)()
    n = load().mpdo_debug_layer_kept(buf, 1024)
    return [int(buf[i]) for i in range(min(n, 1024))]


def prof_shapes(cls: str = 'gemm') -> List[Tuple[Tuple[int, ...], int, float, float]]:
    """Per-shape totals of one timing class: ((M, N, K, batch, prec), launches, ms, flops)."""
    lib = load()
    c = PROF_CLASSES.index(cls)
    n = lib.mpdo_debug_prof_shapes(c, None, None, 0)
    sh = (ctypes.c_int * max(5 * n, 1))()
    st = (ctypes.c_double * max(3 * n, 1))()
    n = lib.mpdo_debug_prof_shapes(c, sh, st, n)
    return [(tuple(sh[5 * i:5 * i + 5]), int(st[3 * i]), st[3 * i + 1], st[3 * i + 2]) for i in range(n)]


def prof_eig_log() -> List[Tuple[int, int, int]]:
    """(order, batch, sweeps) of every blocked eigensolver call recorded since prof_enable(True)."""
    cap = 3 * 65536
    buf = (ctypes.c_int * cap)()
    n = load().mpdo_debug_prof_eig_log(buf, cap)
    return [(int(buf[3 * i]), int(buf[3 * i + 1]), int(buf[3 * i + 2])) for i in range(min(n, cap // 3))]

