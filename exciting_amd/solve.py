"""The generalized eigensolve that follows the H/O setup, on the GPU.

solvewithlapack (src/src_eigensystem/modfvsystem.f90:585-720) calls
ZHEGVX(1, 'V', 'I', 'U', ...) for the lowest nstfv states of H x = e O x.
Here the same problem is reduced with the Cholesky factor of O and solved
with PyTorch's ROCm linear algebra (rocSOLVER as bundled with PyTorch) on a
given stream, so H and O never leave HBM.  H and O from the setup are exactly
Hermitian off the diagonal (lower = conj(upper)); like ZHEGVX, only one
triangle and the real part of the diagonal are used.

The rocBLAS triangular solves need more workspace than PyTorch's default
hipBLAS workspace at n ~ 10^4; ``configure()`` raises it and must run before
the first BLAS call of the process.
"""
from __future__ import annotations

import os

WORKSPACE = ":1048576:2"   # 2 hipBLAS workspaces of 1 GiB
# PyTorch allocates this much per (hipBLAS handle, stream) pair that runs a
# BLAS call, and keeps it until the process ends
WORKSPACE_BYTES = 2 * 1048576 * 1024


def configure() -> None:
    """Give hipBLAS a workspace large enough for n ~ 10^4 triangular solves.
    Effective only before PyTorch creates its first hipBLAS handle."""
    os.environ.setdefault("HIPBLAS_WORKSPACE_CONFIG", WORKSPACE)


def _lanczos_upper_bound(C, steps: int = 12, seed: int = 7):
    """An upper bound of the spectrum of the Hermitian C: the largest Ritz
    value of a short Lanczos run plus the norm of its last residual."""
    import torch
    n = C.shape[0]
    g = torch.Generator(device=C.device).manual_seed(seed)
    v = torch.randn(n, 1, dtype=C.dtype, device=C.device, generator=g)
    v = v / torch.linalg.vector_norm(v)
    alphas, betas = [], []
    v_prev = torch.zeros_like(v)
    beta = 0.0
    f = None
    for j in range(min(steps, n)):
        f = _cmul(C, v, 1.0, -beta, v_prev)
        alpha = float(_gram(v, f).real)
        f = f - alpha * v
        alphas.append(alpha)
        beta = float(torch.linalg.vector_norm(f))
        if beta == 0.0:
            break
        betas.append(beta)
        v_prev, v = v, f / beta
    k = len(alphas)
    T = torch.zeros((k, k), dtype=torch.float64)
    for i in range(k):
        T[i, i] = alphas[i]
        if i + 1 < k:
            T[i, i + 1] = T[i + 1, i] = betas[i]
    return float(torch.linalg.eigvalsh(T)[-1]) + (betas[k - 1] if len(betas) >= k else 0.0)


# widest block the split-K product takes; wider blocks are compute-bound and
# PyTorch's GEMM is faster there (n = 10364: 0.94 ms against 1.31 ms at 56
# columns, 1.8 against 2.4 ms at 80-128, 8.1 against 7.2 ms at 560)
SPLIT_K_MAX_COLS = 128


def _cmul(C, X, alpha: float = 1.0, beta: float = 0.0, Y0=None, gamma: float = 0.0, Z=None):
    """alpha C X + beta Y0 + gamma Z for the Hermitian C.  On the GPU, for
    blocks of up to SPLIT_K_MAX_COLS columns, one launch of the library's
    split-K product (exh_zgemm_cn_device), which streams C at a rate the
    vendor GEMM does not reach at this shape; otherwise PyTorch."""
    if C.is_cuda and X.shape[1] <= SPLIT_K_MAX_COLS:
        from . import lapw
        return lapw.zgemm_cn(C, X, alpha=alpha, beta=beta, y0=Y0, gamma=gamma, z=Z,
                             hermitian=True)
    Y = alpha * (C @ X)
    if Y0 is not None:
        Y = Y + beta * Y0
    if Z is not None:
        Y = Y + gamma * Z
    return Y


def _gram(A, B):
    """A^H B for two blocks of n rows.  On the GPU, for blocks of up to
    SPLIT_K_MAX_COLS columns, the library's split-K product: PyTorch's GEMM
    gives this tall-K shape (56 x 56 over K = 10364) one workgroup and
    takes about 0.9 ms; split over K into 162 chunks it costs a small
    fraction of one filter product."""
    if A.is_cuda and max(A.shape[1], B.shape[1]) <= SPLIT_K_MAX_COLS:
        from . import lapw
        return lapw.zgemm_cn(A, B)
    return A.conj().transpose(0, 1) @ B


def _chebyshev_filter(C, X, degree: int, a: float, b: float, a0: float):
    """p(C) X with p the degree-d Chebyshev polynomial that damps [a, b] and
    grows below a, scaled to about 1 at a0 (Zhou, Saad, Tiago and
    Chelikowsky, J. Comput. Phys. 219 (2006) 172, Algorithm 3.2)."""
    e = 0.5 * (b - a)
    c = 0.5 * (b + a)
    sigma = e / (a0 - c)
    tau = 2.0 / sigma
    Y = _cmul(C, X, sigma / e, -c * sigma / e, X)
    for _ in range(2, degree + 1):
        sigma_new = 1.0 / (tau - sigma)
        f = 2.0 * sigma_new / e
        Y_new = _cmul(C, Y, f, -c * f, Y, -(sigma * sigma_new), X)
        X, Y, sigma = Y, Y_new, sigma_new
    return Y


def _orthonormalize(X):
    """An orthonormal basis of X's columns: Cholesky QR twice (two Gram
    matrices and triangular solves, far cheaper than Householder QR for a
    block of hundreds of columns), accepted only when the first pass left the
    Gram matrix within 0.5 of the identity (Frobenius), which bounds the
    second pass's loss of orthogonality at a few units of rounding (Yamamoto
    et al., ETNA 44 (2015) 306).  Otherwise, or when a Cholesky factorisation
    fails on a nearly rank-deficient block, Householder QR."""
    import torch
    m = X.shape[1]
    eye = torch.eye(m, dtype=X.dtype, device=X.device)
    R, info = torch.linalg.cholesky_ex(_gram(X, X), upper=True)
    if int(info) == 0:
        X1 = torch.linalg.solve_triangular(R, X, upper=True, left=False)
        G = _gram(X1, X1)
        if float(torch.linalg.matrix_norm(G - eye)) <= 0.5:
            R, info = torch.linalg.cholesky_ex(G, upper=True)
            if int(info) == 0:
                return torch.linalg.solve_triangular(R, X1, upper=True, left=False)
    return torch.linalg.qr(X)[0]


def block_size(nwant: int) -> int:
    """Vectors the subspace iteration carries for nwant states: 40 % more,
    at least 16 more (at n = 10364, 400 states: 560 vectors converge in 5
    filter steps, 500 in 6-8)."""
    return max(nwant + 16, (7 * nwant + 4) // 5)


def filter_degree(nwant: int) -> int:
    """Chebyshev degree per filter step: 32 for small blocks (whose
    orthonormalisation and Rayleigh-Ritz cost little next to the filter), 24
    above 64 states (measured at n = 10364: 40 states 208 ms at degree 32,
    240 ms at 16; 400 states 1.09 s at 24, 1.17 s at 32)."""
    return 32 if nwant <= 64 else 24


# Record of the last lowest_eigvalsh call: filter steps, final largest
# residual norm, the tolerance it met, and whether it fell back to the full
# eigensolve.
last_subspace: dict = {}


def lowest_eigvalsh(C, nwant: int, degree: int = 0, rtol: float = 1e-11, atol: float = 1e-11,
                    maxit: int = 40, seed: int = 1234):
    """The nwant lowest eigenvalues of the Hermitian matrix C (n x n CUDA
    tensor), as ZHEGVX's range 'I' asks, by Chebyshev-filtered subspace
    iteration: a block of block_size(nwant) vectors, filtered with the
    polynomial of degree filter_degree(nwant) (or ``degree``) that damps
    everything above the block's largest Ritz value, orthonormalised and
    Rayleigh-Ritz projected until every wanted Ritz pair has residual norm
    <= max(min(rtol * max|spectrum|, atol), 64 eps max|spectrum|).  The
    residual bounds each eigenvalue's error (Weyl), so atol (Ha) caps the
    error however large the spectrum of C grows with an ill-conditioned
    overlap; the last term is the rounding floor of any FP64 eigensolver on
    C, ZHEGVX's included.  Small problems, and a run that does not meet the
    tolerance in maxit filter steps, use the full Hermitian eigensolve
    instead.  Deterministic: the start block comes from a seeded
    generator."""
    return lowest_eigh(C, nwant, False, degree, rtol, atol, maxit, seed)


def lowest_eigh(C, nwant: int, vectors: bool = True, degree: int = 0, rtol: float = 1e-11,
                atol: float = 1e-11, maxit: int = 40, seed: int = 1234):
    """lowest_eigvalsh with, when vectors is true, the Ritz vectors of the
    nwant lowest pairs: an (n, nwant) orthonormal block, columns in the order
    of the eigenvalues.  A residual r = C x - theta x of norm <= tol puts the
    vector within tol / gap of the exact eigenvector (the invariant subspace
    of a cluster, Davis-Kahan), as ZHEGVX's own vectors are determined only
    up to mixing inside clusters.  Returns w, or (w, X) with vectors."""
    import torch
    n = C.shape[0]
    m = min(n, block_size(nwant))
    degree = degree or filter_degree(nwant)
    last_subspace.clear()

    def full():
        if not vectors:
            return torch.linalg.eigvalsh(C)[:nwant]
        w, V = torch.linalg.eigh(C)
        return w[:nwant], V[:, :nwant]

    if n < 512 or 3 * m > n:
        last_subspace.update(steps=0, fallback="small")
        return full()
    b = _lanczos_upper_bound(C)
    g = torch.Generator(device=C.device).manual_seed(seed)
    X = torch.randn(n, m, dtype=C.dtype, device=C.device, generator=g)
    X = _orthonormalize(X)
    res = float("inf")
    for it in range(maxit):
        CX = _cmul(C, X)
        Hs = _gram(X, CX)
        Hs = 0.5 * (Hs + Hs.conj().transpose(0, 1))
        theta, Q = torch.linalg.eigh(Hs)
        X = X @ Q
        CX = CX @ Q
        R = CX[:, :nwant] - X[:, :nwant] * theta[:nwant].to(X.dtype)[None, :]
        scale = max(abs(b), float(theta.abs().max()))
        tol = max(min(rtol * scale, atol), 64 * 2.0 ** -52 * scale)
        res = float(torch.linalg.vector_norm(R, dim=0).max())
        if res <= tol:
            last_subspace.update(steps=it, residual=res, tol=tol, scale=scale, fallback=None)
            return (theta[:nwant], X[:, :nwant]) if vectors else theta[:nwant]
        a, a0 = float(theta[-1]), float(theta[0])
        if not a < b:
            break
        X = _chebyshev_filter(C, X, degree, a, b, a0)
        X = _orthonormalize(X)
    last_subspace.update(steps=maxit, residual=res, scale=b, fallback="not converged")
    return full()


def geneig(H, O, n: int, nstfv: int, vectors: bool = False, method: str = "full"):
    """Lowest nstfv eigenpairs of H x = e O x.

    H, O: torch complex128 CUDA tensors holding the column-major n x n
    matrices as produced by the setup (flat or (n, n); the row-major view is
    the transpose).  Returns eigenvalues (float64, nstfv) and, if requested,
    eigenvectors as an (n, nstfv) tensor (columns are states, normalised so
    that x^H O x = 1, ZHEGVX's convention).  Raises RuntimeError if O is not
    positive definite (modfvsystem.f90:698-711 stops in that case).
    method: "full" is the full Hermitian eigensolve, "subspace" the
    Chebyshev-filtered subspace iteration (lowest_eigh; with vectors its Ritz
    vectors), "auto" the latter when the block it iterates is at most n/16
    (n = 10364: 40 states in 0.21 s instead of 1.65 s for all eigenvalues;
    400 states in 1.1 s)."""
    import torch
    Ht = H.reshape(-1)[:n * n].view(n, n)   # row-major view = H^T = conj(H)
    Ot = O.reshape(-1)[:n * n].view(n, n)
    # H^T is Hermitian with the same spectrum; use its lower triangle, which
    # is the conjugate of the setup's upper triangle (ZHEGVX 'U' data).
    L, info = torch.linalg.cholesky_ex(Ot)
    if int(info) != 0:
        raise RuntimeError(f"overlap matrix not positive definite (leading minor {int(info)})")
    Y = torch.linalg.solve_triangular(L, Ht, upper=False)
    C = torch.linalg.solve_triangular(L, Y.conj().transpose(0, 1), upper=False)
    C = 0.5 * (C + C.conj().transpose(0, 1))
    few = method == "subspace" or (method == "auto" and 16 * block_size(nstfv) <= n)
    if not vectors:
        if few:
            return lowest_eigvalsh(C, nstfv)
        return torch.linalg.eigvalsh(C)[:nstfv]
    if few:
        w, V = lowest_eigh(C, nstfv, True)
    else:
        w, V = torch.linalg.eigh(C)
        w, V = w[:nstfv], V[:, :nstfv]
    X = torch.linalg.solve_triangular(L.conj().transpose(0, 1), V, upper=True)
    # eigenvectors of the transposed problem are the conjugates of the
    # original ones
    return w[:nstfv], X.conj()


def solvewithlapack(H, O, n: int, nstfv: int, nmatmax: int = 0, method: str = "full"):
    """modfvsystem.f90:585-720 on the GPU: the lowest nstfv eigenpairs of
    H x = e O x with ZHEGVX(1, 'V', 'I', 'U') semantics.

    Returns (evalfv, evecfv): evalfv a float64 tensor (nstfv,), ascending;
    evecfv a complex128 tensor of shape (nstfv, nmatmax) whose memory is the
    Fortran array evecfv(nmatmax, nstfv) (column ist is state ist, rows n+1..
    nmatmax zero).  Normalisation x^H O x = 1 as ZHEGVX's; the phase of each
    vector is the solver's own (the reference's consumers only use
    phase-invariant products).  Raises RuntimeError where ZHEGVX returns
    INFO > n (O not positive definite; the reference stops).  method as in
    geneig ("auto": the subspace iteration for few states of a large
    matrix)."""
    import torch
    nmatmax = max(nmatmax, n)
    w, X = geneig(H, O, n, nstfv, vectors=True, method=method)
    evec = torch.zeros((nstfv, nmatmax), dtype=torch.complex128, device=X.device)
    evec[:, :n] = X.transpose(0, 1)
    return w, evec


def normalise_large_components(evecfv, O0, n: int):
    """seceqnfv.f90:93-148 (ValenceRelativity 'iora*'): after the solve, each
    state is divided by sqrt(|x^H O0 x|), where O0 is the overlap set up
    again without the h1 terms (h1aa = h1loa = h1lolo = 0, h1on = false).
    evecfv as returned by solvewithlapack (updated in place and returned);
    O0 the column-major n x n overlap (flat or (n, n) device tensor)."""
    import torch
    O0t = O0.reshape(-1)[:n * n].view(n, n)       # row-major view = O0^T
    X = evecfv[:, :n]                             # (nstfv, n): row ist = state ist
    # zm = O0 x for every state, zm2(ist, ist) = x^H O0 x
    Z = torch.matmul(X, O0t.contiguous())         # rows: (O0 x)^T
    d = (X.conj() * Z).sum(dim=1).abs()
    evecfv[:, :n] = X / torch.sqrt(d)[:, None]
    return evecfv

