"""The non-local exchange matrix of hybrid functionals in the LAPW basis,
src_hybrids/calc_vnlmat.f90:13-95, on the GPU.

For each k-point the reference runs match (:49-50), overlapsetup alone
(:52-53) and three zgemm (:54-73):

    temp   = X^H O           (nstfv x n)      X = evec(1:n, 1:nstfv)
    temp1  = vxnl temp       (nstfv x n)      vxnl(nstfv, nstfv, ik)
    vnlmat = temp^H temp1    (n x n)

Here match and overlapsetup are libexh's (exh_match_device,
exh_overlapsetup_device: O alone, without the Hamiltonian's products), and
the three products are plain library GEMMs through PyTorch (rocBLAS /
hipBLASLt), all in FP64 with O, X and the result resident in HBM.
"""
from __future__ import annotations

from typing import List


def vnlmat_kpoint(ctx, d, evec, vxnl, O=None):
    """calc_vnlmat.f90:48-73 for one k-point.  ctx: lapw.Context with the
    structure's tables; d: lapw.DeviceKPoint (apwalm is overwritten by
    match); evec: CUDA complex128 tensor (nstfv, nmat), row i holding state i
    (the memory of the Fortran evecfv(nmat, nstfv)); vxnl: (nstfv, nstfv)
    CUDA complex128 tensor.  O: optional rank x rank scratch.  Returns
    vnlmat as an (n, n) tensor T with T[j, i] = vnlmat(i, j) (column-major
    memory, like the setup's matrices)."""
    import torch
    from . import lapw
    n = d.rank
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    if O is None:
        O = torch.zeros(n * n, dtype=torch.complex128, device=evec.device)
    system = lapw.EvSystem(None, lapw.HermitianMatrix(n, O))
    ctx.overlapsetup(system, d.ngk, d.apwalm, d.igpig, d.vgpc, flags=0)
    Om = O.view(n, n).transpose(0, 1)              # O(i, j) as a matrix
    X = evec[:, :n].transpose(0, 1)                # (n, nstfv): column s = state s
    temp = X.conj().transpose(0, 1) @ Om           # X^H O
    temp1 = vxnl @ temp
    vnl = temp.conj().transpose(0, 1).contiguous() @ temp1
    return vnl.transpose(0, 1).contiguous()


def calc_vnlmat(ctx, dks: List, evecs: List, vxnls: List):
    """calc_vnlmat.f90:38-87 over the k-points of one process: a list of
    vnlmat (column-major (n, n) tensors), one per k-point."""
    return [vnlmat_kpoint(ctx, d, x, v) for d, x, v in zip(dks, evecs, vxnls)]
