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


def geneig(H, O, n: int, nstfv: int, vectors: bool = False):
    """Lowest nstfv eigenpairs of H x = e O x.

    H, O: torch complex128 CUDA tensors holding the column-major n x n
    matrices as produced by the setup (flat or (n, n); the row-major view is
    the transpose).  Returns eigenvalues (float64, nstfv) and, if requested,
    eigenvectors as an (n, nstfv) tensor (columns are states, normalised so
    that x^H O x = 1, ZHEGVX's convention).  Raises RuntimeError if O is not
    positive definite (modfvsystem.f90:698-711 stops in that case)."""
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
    if not vectors:
        return torch.linalg.eigvalsh(C)[:nstfv]
    w, V = torch.linalg.eigh(C)
    X = torch.linalg.solve_triangular(L.conj().transpose(0, 1), V[:, :nstfv], upper=True)
    # eigenvectors of the transposed problem are the conjugates of the
    # original ones
    return w[:nstfv], X.conj()


def solvewithlapack(H, O, n: int, nstfv: int, nmatmax: int = 0):
    """modfvsystem.f90:585-720 on the GPU: the lowest nstfv eigenpairs of
    H x = e O x with ZHEGVX(1, 'V', 'I', 'U') semantics.

    Returns (evalfv, evecfv): evalfv a float64 tensor (nstfv,), ascending;
    evecfv a complex128 tensor of shape (nstfv, nmatmax) whose memory is the
    Fortran array evecfv(nmatmax, nstfv) (column ist is state ist, rows n+1..
    nmatmax zero).  Normalisation x^H O x = 1 as ZHEGVX's; the phase of each
    vector is the solver's own (the reference's consumers only use
    phase-invariant products).  Raises RuntimeError where ZHEGVX returns
    INFO > n (O not positive definite; the reference stops)."""
    import torch
    nmatmax = max(nmatmax, n)
    w, X = geneig(H, O, n, nstfv, vectors=True)
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

