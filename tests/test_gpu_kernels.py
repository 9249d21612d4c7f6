"""Kernel-level checks of libmpdo_gfx950: the strided complex GEMM on the matrix cores and the
fp64 Jacobi eigensolver, against numpy in complex128."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


# Error bars relative to the largest output element. fp32 accumulation: 0 four real products, 5
# three (the imaginary part P3 - P1 - P2 loses up to a few ulp of |A||B| more), 7 four real products
# of operands split into three bf16 parts (six part products: fp32-level); 6: two bf16 parts (three
# part products: each real product within ~3 * 2^-16 of |a||b|).
_TOL = {0: lambda K: 3e-6 * np.sqrt(K), 1: lambda K: 1e-6, 2: lambda K: 1e-13, 3: lambda K: 1e-6,
        4: lambda K: 1e-6, 5: lambda K: 6e-6 * np.sqrt(K), 6: lambda K: 1e-4, 7: lambda K: 3e-6 * np.sqrt(K)}


def _crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize('prec', [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize('M,N,K', [(1, 1, 1), (16, 16, 4), (37, 53, 29), (130, 70, 200), (257, 129, 300)])
@pytest.mark.parametrize('tA,tB,cA,cB', [(0, 0, 0, 0), (1, 0, 1, 0), (0, 1, 0, 1), (1, 1, 0, 0)])
def test_gemm(cuda, prec, M, N, K, tA, tB, cA, cB):
    from mpdo_amd import _lib
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    batch = 2
    din = np.complex64 if prec in (0, 1, 4, 5, 6, 7) else np.complex128
    A = _crand(rng, batch, *((K, M) if tA else (M, K))).astype(din)
    B = _crand(rng, batch, *((N, K) if tB else (K, N))).astype(din)
    C0 = _crand(rng, batch, M, N)
    out_t = torch.complex64 if prec in (0, 3, 4, 5, 6, 7) else torch.complex128
    Cd = torch.from_numpy(C0).to(out_t).to(cuda)
    sa = (1, M, M * K) if tA else (K, 1, M * K)
    sb = (1, K, K * N) if tB else (N, 1, K * N)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
    _lib.debug_gemm(prec, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), Cd, M, N, K, sa, sb,
                    (N, 1, M * N), bool(cA), bool(cB), alpha, beta, batch)
    opA = np.swapaxes(A, 1, 2) if tA else A
    opB = np.swapaxes(B, 1, 2) if tB else B
    opA = opA.conj() if cA else opA
    opB = opB.conj() if cB else opB
    ref = alpha * (opA.astype(np.complex128) @ opB.astype(np.complex128)) + beta * C0.astype(
        np.complex64 if prec in (0, 3, 4, 5, 6, 7) else np.complex128).astype(np.complex128)
    got = Cd.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    tol = _TOL[prec](K)
    assert err < tol, (prec, M, N, K, err)


@pytest.mark.parametrize('prec', [1, 2, 3, 4, 6, 7])
@pytest.mark.parametrize('n,K', [(1, 5), (17, 33), (64, 16), (65, 130), (200, 70), (257, 512)])
def test_gemm_hermitian(cuda, prec, n, K):
    """Hermitian-output mode (prec + 10): X^H X computed from the tiles on and below the
    diagonal, the upper triangle mirrored; vs numpy, and exactly Hermitian."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(n * 7 + K)
    batch = 3
    din = np.complex64 if prec in (1, 4, 6, 7) else np.complex128
    X = _crand(rng, batch, K, n).astype(din)
    out_t = torch.complex64 if prec in (3, 4, 6, 7) else torch.complex128
    Cd = torch.full((batch, n, n), 7.0 + 3.0j, dtype=out_t, device=cuda)
    Xd = torch.from_numpy(X).to(cuda)
    _lib.debug_gemm(prec + 10, Xd, Xd, Cd, n, n, K, (1, n, K * n), (n, 1, K * n), (n, 1, n * n), True, False,
                    1.0, 0.0, batch)
    Xw = X.astype(np.complex128)
    ref = np.conj(np.swapaxes(Xw, 1, 2)) @ Xw
    got = Cd.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < (_TOL[prec](K) if prec in (6, 7) else 1e-13 if prec == 2 else 1e-6), (prec, n, K, err)
    assert np.array_equal(got, np.conj(np.swapaxes(got, 1, 2)) * np.where(np.eye(n, dtype=bool), 0, 1) +
                          got * np.eye(n)), 'not exactly Hermitian off the diagonal'


@pytest.mark.parametrize('prec', [1, 2, 3, 4])
@pytest.mark.parametrize('inA,inB', [(1, 2), (2, 1)])
@pytest.mark.parametrize('M,N,K,tA,tB', [(70, 33, 130, False, False), (64, 128, 16, True, False),
                                         (129, 65, 64, False, True), (5, 300, 17, True, True)])
def test_gemm_mixed_operands(cuda, prec, inA, inB, M, N, K, tA, tB):
    """complex64 x complex128 operands (each in its own storage, widened in the staging) with
    fp64 accumulation, any orientation, conjugation and alpha/beta; vs numpy in complex128."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(M + 7 * N + 13 * K + inA)
    batch = 2
    tya = np.complex64 if inA == 1 else np.complex128
    tyb = np.complex64 if inB == 1 else np.complex128
    A = _crand(rng, batch, *((K, M) if tA else (M, K))).astype(tya)
    B = _crand(rng, batch, *((N, K) if tB else (K, N))).astype(tyb)
    C0 = _crand(rng, batch, M, N)
    out_t = torch.complex64 if prec in (3, 4) else torch.complex128
    Cd = torch.from_numpy(C0).to(out_t).to(cuda)
    sa = (1, M, M * K) if tA else (K, 1, M * K)
    sb = (1, K, K * N) if tB else (N, 1, K * N)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
    _lib.debug_gemm_ex(prec, False, inA, inB, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), Cd, M, N, K,
                       sa, sb, (N, 1, M * N), tA, not tA, alpha, beta, batch)
    opA = np.swapaxes(A, 1, 2) if tA else A
    opB = np.swapaxes(B, 1, 2) if tB else B
    opA = opA.conj() if tA else opA
    opB = opB.conj() if not tA else opB
    Cref = C0.astype(np.complex64 if prec in (3, 4) else np.complex128).astype(np.complex128)
    ref = alpha * (opA.astype(np.complex128) @ opB.astype(np.complex128)) + beta * Cref
    got = Cd.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < (1e-13 if prec in (1, 2) else 1e-6), (prec, inA, inB, M, N, K, err)


@pytest.mark.parametrize('prec', [4, 0, 5, 6, 7])
@pytest.mark.parametrize('M,N,K,tA', [(16, 4096, 256, True), (32, 640, 40, False), (5, 513, 17, True),
                                      (31, 1000, 300, False)])
def test_gemm_few_rows(cuda, prec, M, N, K, tA):
    """At most 32 output rows and a wide output: the 32 x 128 tile shape, with fp64 accumulation
    (mode 4), fp32 accumulation of four (mode 0) and three (mode 5, 3M) real products and the
    split-bf16 products (modes 6, 7); vs numpy with the bars of the general test."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(M * 11 + N + K)
    batch = 3
    A = _crand(rng, batch, *((K, M) if tA else (M, K))).astype(np.complex64)
    B = _crand(rng, batch, K, N).astype(np.complex64)
    C = torch.zeros(batch, M, N, dtype=torch.complex64, device=cuda)
    sa = (1, M, M * K) if tA else (K, 1, M * K)
    _lib.debug_gemm(prec, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), C, M, N, K, sa, (N, 1, K * N),
                    (N, 1, M * N), batch=batch)
    opA = np.swapaxes(A, 1, 2) if tA else A
    ref = opA.astype(np.complex128) @ B.astype(np.complex128)
    got = C.cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < _TOL[prec](K)


@pytest.mark.parametrize('prec', [6, 7])
@pytest.mark.parametrize('tile', [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize('M,N,K,tA', [(700, 8, 40, False), (600, 32, 70, True), (33, 600, 50, False)])
def test_gemm_bf16_tiles(cuda, prec, tile, M, N, K, tA):
    """Every tile shape of the split-bf16 GEMM (the built-in rules, 64x64, 128x64, 64x128, 32x128,
    128x32) on outputs with few columns, few rows and ragged edges; vs numpy with the bars of the
    general test."""
    from mpdo_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(M + 7 * N + K + tile)
    batch = 2
    A = _crand(rng, batch, *((K, M) if tA else (M, K))).astype(np.complex64)
    B = _crand(rng, batch, K, N).astype(np.complex64)
    C = torch.zeros(batch, M, N, dtype=torch.complex64, device=cuda)
    sa = (1, M, M * K) if tA else (K, 1, M * K)
    old = lib.mpdo_debug_gemm_bf16_tiles(tile)
    try:
        _lib.debug_gemm(prec, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), C, M, N, K, sa,
                        (N, 1, K * N), (N, 1, M * N), batch=batch)
    finally:
        lib.mpdo_debug_gemm_bf16_tiles(old)
    opA = np.swapaxes(A, 1, 2) if tA else A
    ref = opA.astype(np.complex128) @ B.astype(np.complex128)
    assert np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max() < _TOL[prec](K)


@pytest.mark.parametrize('n', [1, 37, 64, 100, 200, 513])
def test_unpivoted_cholesky(cuda, n):
    """The unpivoted blocked Cholesky (64-column diagonal blocks factored and inverted in one
    workgroup, the rows below and the trailing matrix as GEMMs): L L^H = A within 1e-12 of max|A| on
    well-conditioned Hermitian matrices, L lower triangular with a real positive diagonal, rank n;
    a matrix whose trailing block is singular stops at the first pivot below the threshold."""
    import ctypes
    from mpdo_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(n)
    batch = 3
    X = _crand(rng, batch, n, n + 8)
    A = X @ np.conj(np.swapaxes(X, 1, 2))
    A = (A + np.conj(np.swapaxes(A, 1, 2))) / 2
    if n >= 100:   # chain 2: rank n - 30, the deficiency in the last 30 rows
        Y = _crand(rng, n - 30, n + 8)
        Y = np.vstack([Y, rng.standard_normal((30, n - 30)) @ Y])
        A[2] = Y @ Y.conj().T
    Ad = torch.from_numpy(A.copy()).to(cuda)
    L = torch.zeros((batch, n, n), dtype=torch.complex128, device=cuda)
    rank = (ctypes.c_int * batch)()
    _lib.check(lib.mpdo_debug_chol(n, batch, ctypes.c_void_p(Ad.data_ptr()), 1e-12, ctypes.c_void_p(L.data_ptr()),
                                   rank, _lib._stream(cuda)), 'mpdo_debug_chol')
    Lh = np.swapaxes(L.cpu().numpy(), 1, 2)   # stored by columns
    for b in range(batch):
        Lb = Lh[b]
        assert np.abs(np.triu(Lb, 1)).max() == 0
        if n >= 100 and b == 2:
            r = rank[b]
            assert n - 30 - 64 < r <= n - 30, r   # stops inside the singular trailing part
            assert np.abs(Lb[:, r:]).max() == 0
            assert np.abs((Lb @ Lb.conj().T - A[b])[:r, :r]).max() <= 1e-10 * np.abs(A[b]).max()
            continue
        assert rank[b] == n
        assert np.all(np.diag(Lb).real > 0) and np.abs(np.diag(Lb).imag).max() == 0
        assert np.abs(Lb @ Lb.conj().T - A[b]).max() <= 1e-12 * np.abs(A[b]).max()


@pytest.mark.parametrize('tri', [1, 2])
@pytest.mark.parametrize('M,N,K', [(300, 130, 130), (130, 257, 130), (64, 64, 64), (17, 40, 40), (40, 17, 40)])
def test_gemm_triangular_operand(cuda, tri, M, N, K):
    """tri = 1: op(B) upper triangular (K x N), tri = 2: op(A) lower triangular (M x K); each output
    tile reduces only over the k range that can be nonzero. Complex64 data x complex128 factor
    with a complex64 output, as in the Cholesky-QR products; vs numpy."""
    from mpdo_amd import _lib
    if (tri == 1 and K != N) or (tri == 2 and K != M):
        pytest.skip('square triangular operand')
    rng = np.random.default_rng(M + 3 * N + 7 * K + tri)
    batch = 2
    A = _crand(rng, batch, M, K)
    B = _crand(rng, batch, K, N)
    if tri == 1:
        B = np.triu(B)
        A = A.astype(np.complex64)
        inA, inB = 1, 2
    else:
        A = np.tril(A)
        B = B.astype(np.complex64)
        inA, inB = 2, 1
    C = torch.zeros(batch, M, N, dtype=torch.complex64, device=cuda)
    _lib.debug_gemm_ex(3, False, inA, inB, torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), C, M, N, K,
                       (K, 1, M * K), (N, 1, K * N), (N, 1, M * N), batch=batch, tri=tri)
    ref = A.astype(np.complex128) @ B.astype(np.complex128)
    got = C.cpu().numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-6


@pytest.mark.parametrize('slices', [1, 2, 4, 8])
@pytest.mark.parametrize('n,K', [(64, 4096), (200, 8192), (256, 2048), (33, 64)])
def test_gram_split_reduction(cuda, n, K, slices):
    """Split-reduction Hermitian Gram: the slices' partial products (one launch, lower tiles
    only), summed in order and mirrored, equal X X^H in complex128 and are exactly Hermitian."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(n + K + slices)
    X = _crand(rng, 3, n, K).astype(np.complex64)
    G = _lib.debug_gram_split(torch.from_numpy(X).to(cuda), slices).cpu().numpy()
    Xw = X.astype(np.complex128)
    ref = Xw @ np.conj(np.swapaxes(Xw, 1, 2))
    assert np.abs(G - ref).max() / np.abs(ref).max() < 1e-13
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(G[:, off], np.conj(np.swapaxes(G, 1, 2))[:, off])


@pytest.mark.parametrize('n', [1, 2, 3, 17, 64, 65, 100, 192])
def test_eigh(cuda, n):
    from mpdo_amd import _lib
    rng = np.random.default_rng(n)
    X = _crand(rng, 3, n, n + 2)
    # graded spectrum: PSD Gram matrices with eigenvalues spanning 12 decades
    scale = np.logspace(0, -6, n)
    X = X * scale[None, :, None]
    H = X @ np.swapaxes(X.conj(), 1, 2)
    w, V, sweeps = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert sweeps < 40
    for b in range(3):
        ref = np.sort(np.linalg.eigvalsh(H[b]))[::-1]
        assert np.all(np.diff(w[b]) <= 1e-12 * ref[0])                   # descending
        assert np.abs(w[b] - ref).max() <= 1e-13 * ref[0]
        assert np.abs(H[b] @ V[b] - V[b] * w[b][None, :]).max() <= 1e-12 * ref[0]
        # off-diagonals below 1e-14 of the largest diagonal are not rotated: absolute error bound
        assert np.all(np.abs(w[b] - ref) <= 1e-13 * ref[0] + 1e-10 * np.abs(ref))
        assert np.abs(V[b].conj().T @ V[b] - np.eye(n)).max() <= 1e-12


@pytest.mark.parametrize('n', [130, 300])
def test_eigh_block_threads(cuda, n):
    """The blocked solver's subproblem kernel strides its rounds, its rotation log and the replay
    of that log into W over any workgroup size: 256, 512 and 1024 threads give the same
    eigenpairs to the bit (each element sees the same operations in the same order), and so does
    the variant that accumulates W in LDS during the sweep."""
    from mpdo_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(n)
    X = _crand(rng, 2, n, n + 3) * np.logspace(0, -5, n)[None, :, None]
    H = X @ np.swapaxes(X.conj(), 1, 2)
    outs = []
    old = lib.mpdo_debug_bj_threads(0)
    old_fused = lib.mpdo_debug_bj_fused_max(0)   # the logged sweep and its replay
    try:
        for t in (256, 512, 1024):
            lib.mpdo_debug_bj_threads(t)
            w, V, _ = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda))
            outs.append((w.cpu().numpy(), V.cpu().numpy()))
        # the rotation product accumulated in LDS next to the subproblem: the same operations on
        # every element in the same order
        lib.mpdo_debug_bj_fused_max(1 << 20)
        w, V, _ = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda))
        fused = (w.cpu().numpy(), V.cpu().numpy())
    finally:
        lib.mpdo_debug_bj_threads(old)
        lib.mpdo_debug_bj_fused_max(old_fused)
    for w, V in outs[:2] + [fused]:
        assert np.array_equal(w, outs[2][0]) and np.array_equal(V, outs[2][1])
    for b in range(2):
        ref = np.sort(np.linalg.eigvalsh(H[b]))[::-1]
        assert np.abs(outs[2][0][b] - ref).max() <= 1e-13 * ref[0]
        assert np.abs(H[b] @ outs[2][1][b] - outs[2][1][b] * outs[2][0][b][None, :]).max() <= 1e-12 * ref[0]


@pytest.mark.parametrize('n,rank', [(40, 5), (100, 10), (256, 30), (200, 0)])
def test_eigh_rank_deficient(cuda, n, rank):
    """Gram matrices of rank << n (zero null space plus rounding noise) must converge."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(n + rank)
    X = _crand(rng, 2, n, max(rank, 1)) * (1.0 if rank else 0.0)
    H = X @ np.swapaxes(X.conj(), 1, 2)
    H = H + 1e-30 * _crand(rng, 2, n, n)
    H = 0.5 * (H + np.swapaxes(H.conj(), 1, 2))
    w, V, sweeps = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert sweeps < 40
    for b in range(2):
        ref = np.sort(np.linalg.eigvalsh(H[b]))[::-1]
        top = max(ref[0], np.abs(H[b]).max())
        assert np.abs(w[b] - ref).max() <= 1e-12 * top
        assert np.abs(V[b].conj().T @ V[b] - np.eye(n)).max() <= 1e-11


@pytest.mark.parametrize('n,rank', [(100, 100), (256, 256), (512, 300), (1024, 450), (200, 40), (640, 600),
                                    (1024, 1024)])
def test_eigh_gram(cuda, n, rank):
    """The pivoted-Cholesky-preconditioned solver on graded Gram matrices (12 decades) of full and
    deficient rank: eigenvalues above its floor as numpy's; eigenvectors with residual within the
    Jacobi floor on L^H L times sqrt(largest / own eigenvalue), orthonormality within it times the
    product of two such factors; zero vectors for zero eigenvalues."""
    from mpdo_amd import _lib
    rng = np.random.default_rng(n + rank)
    X = _crand(rng, 3, n, rank) * np.logspace(0, -6, rank)[None, None, :]
    H = X @ np.swapaxes(X.conj(), 1, 2)
    H = 0.5 * (H + np.swapaxes(H.conj(), 1, 2))
    w, V, sweeps = _lib.debug_eigh(torch.from_numpy(H.copy()).to(cuda), gram=True)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    assert sweeps < 40
    floor = 1e-14 * max(1, ((n + 63) // 64 * 64) // 64)
    for b in range(3):
        ref = np.sort(np.linalg.eigvalsh(H[b]))[::-1]
        top = ref[0]
        big = ref > 100 * floor * top
        assert np.all(np.diff(w[b]) <= 1e-12 * top)
        assert np.abs(w[b][big] - ref[big]).max() <= 1e-12 * top
        assert np.all(np.abs(w[b]) <= ref[0] * (1 + 1e-12))
        keep = w[b] > 1e-9 * top
        Vk, wk = V[b][:, keep], w[b][keep]
        # V = L U Lambda^{-1/2}: the Jacobi residual on L^H L is amplified by sqrt(top / w_j)
        amp = np.sqrt(top / wk)
        scale = max(1, ((n + 63) // 64 * 64) // 64)
        res = np.abs(H[b] @ Vk - Vk * wk[None, :]).max(axis=0)
        assert np.all(res <= 1e-13 * scale * top * amp), (res / (top * amp)).max()
        # V_i^H V_j = (U_i^H M U_j) / sqrt(w_i w_j): the floor times amp_i amp_j
        gram = np.abs(Vk.conj().T @ Vk - np.eye(Vk.shape[1]))
        assert np.all(gram <= 1e-12 * scale * np.outer(amp, amp)), (gram / np.outer(amp, amp)).max()
        zero = w[b] == 0
        assert np.all(V[b][:, zero] == 0)
        assert np.all(w[b][rank:] <= 1e-12 * top)
        # the eigenvalues below the pivot threshold stay as the Schur complement's diagonal
        assert abs(w[b].sum() - np.trace(H[b]).real) <= 1e-12 * np.trace(H[b]).real

