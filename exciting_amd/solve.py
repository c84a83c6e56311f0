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
