"""Choice of the split-precision modulus count for a structure.

The split path (EXH_SPLIT(n)) scales every G column of A and aa.A to a
2-norm of 2^c(n) and rounds it to integers, where c grows by about 3.9 bits
per modulus; that rounding is its only error.  How much eigenvalue error a given b causes depends on the structure
(basis size, number of atoms, conditioning of O), so the count is picked by
measurement: on one probe k-point, the lowest states of H x = e O x from the
split setup are compared with the FP64 setup, both solved in FP64 on the
GPU, for n = 6, 7, ... until the largest difference is below the target
divided by a safety margin.  The choice is then checked on a second k-point.
"""
from __future__ import annotations

from typing import Dict, Iterable, Optional

import numpy as np


def eigen_error(ctx, kpt, rank: int, H, O, nmod: int, nstates: int, ref=None):
    """Largest |dE| (Ha) among the lowest nstates eigenvalues of the split
    setup with nmod moduli against FP64, on one k-point of a batch (kpt is
    its exh_kpoint; H and O its device matrices).  ref: the FP64 eigenvalues
    if already known.  Returns (max |dE|, FP64 eigenvalues)."""
    import torch
    from . import lapw, solve
    if ref is None:
        ctx.setup_batch([kpt], flags=lapw.MATCH)
        torch.cuda.synchronize()
        ref = solve.geneig(H, O, rank, nstates).cpu().numpy()
    ctx.setup_batch([kpt], flags=lapw.MATCH | lapw.SPLIT(nmod))
    torch.cuda.synchronize()
    w = solve.geneig(H, O, rank, nstates).cpu().numpy()
    return float(np.abs(w - ref).max()), ref


def tune_moduli(ctx, kpt, rank: int, H, O, nstates: int, target_ha: float = 1e-6,
                margin: float = 3.0, candidates: Iterable[int] = range(6, 13)) -> Dict:
    """Smallest modulus count whose eigenvalue error on the probe k-point is
    at most target_ha / margin.  Returns {"nmod", "errors": {n: |dE|}}; nmod
    is None if no candidate qualifies."""
    errors: Dict[int, float] = {}
    ref: Optional[np.ndarray] = None
    chosen = None
    for n in candidates:
        de, ref = eigen_error(ctx, kpt, rank, H, O, n, nstates, ref)
        errors[n] = de
        if de <= target_ha / margin:
            chosen = n
            break
    return {"nmod": chosen, "errors": errors, "target_ha": target_ha, "margin": margin,
            "nstates": nstates}
