"""Kinetic energy of the first-variational states, kinetic_energy.f90:1-230
('none' and ZORA relativity; the reference stops for iora*), on the GPU.

Per k-point and state x (evecfv(:, ist)) the reference adds
  * a muffin-tin part per atom: zm = A^T x (A = apwalm of the atom,
    kinetic_energy.f90:122-135), then zm^H (t_aa zm + t_alo x_lo) +
    x_lo^H (t_alo^T zm + t_lolo x_lo) with the kinetic radial integrals
    t_aa, t_alo, t_lolo (:94-121, :136-183; diagonal in (l, m)).  As in the
    reference, t_alo and t_lolo carry the ZORA mass factor whatever the
    relativity setting, t_aa only for ZORA;
  * an interstitial part (:208-234): for each Cartesian direction, the
    backward FFT of x_G (G+k)_d placed on the grid, times cfunir (divided by
    1 - alpha^2/2 veffir for ZORA), forward FFT, gathered at the G+k and
    weighted by (G+k)_d; 0.5 x^H of that.

Here the integrals come from libexh (exh_mt_kinetic_device, the same
kernel as hmlint's kinetic block), zm is one library GEMM per atom batch,
the per-atom quadratic forms are batched matrix products, and the
interstitial part is the same cyclic convolution written as one dense
matrix F(i, j) = c((G_i - G_j) mod box) over the G+k set (c the forward
transform of the grid function, one FFT), applied to the three directions'
weighted states in one library GEMM; all in FP64 with the states resident
in HBM.
"""
from __future__ import annotations

import numpy as np

ALPHA = 1.0 / 137.03599911


class KineticPlan:
    """Index maps of one structure: per atom, the rows (lm, io) of zm and the
    LO rows, and the kinetic integral each pair of them takes."""

    def __init__(self, s):
        p, lmaxmat, lmaxapw = s.apwordmax, s.lmaxmat, s.lmaxapw
        lmm = (lmaxapw + 1) ** 2
        nfa = (lmaxmat + 1) * p
        self.lmm, self.p = lmm, p
        self.maxlo = int(max(s.losize)) if len(s.losize) else 0
        dim = lmm * p + self.maxlo
        self.dim = dim
        # per species: f-index of every row of u = [zm (lm, io) ; x_lo (ilo, m)],
        # the row's (l, m), or -1
        self.rowf, self.rowlm = [], []
        for is_ in range(s.nspecies):
            f = np.full(dim, -1, np.int64)
            lmr = np.full(dim, -1, np.int64)
            for l in range(lmaxmat + 1):
                for m in range(-l, l + 1):
                    lm = l * (l + 1) + m
                    for io in range(s.apword[l, is_]):
                        f[lm * p + io] = l * p + io
                        lmr[lm * p + io] = lm
            r = lmm * p
            for ilo in range(s.nlorb[is_]):
                l = s.lorbl[ilo, is_]
                for m in range(-l, l + 1):
                    f[r] = nfa + ilo
                    lmr[r] = l * (l + 1) + m
                    r += 1
            self.rowf.append(f)
            self.rowlm.append(lmr)
        self.species = np.repeat(np.arange(s.nspecies), s.natoms)
        self.lobase = np.concatenate([[0], np.cumsum([s.losize[i] for i in self.species])])[:-1]
        self.losize = [int(s.losize[i]) for i in range(s.nspecies)]
        self._dev = {}

    def on(self, dev):
        """Per species, on device dev (built once): the atoms, the f-index of
        every row (clamped), the mask of row pairs with equal (l, m), and the
        rows of x_lo of each atom's local orbitals."""
        import torch
        key = str(dev)
        if key not in self._dev:
            out = []
            for is_ in range(len(self.rowf)):
                at = np.nonzero(self.species == is_)[0]
                f = torch.as_tensor(self.rowf[is_], device=dev)
                lmr = torch.as_tensor(self.rowlm[is_], device=dev)
                ok = f >= 0
                same = ok[:, None] & ok[None, :] & (lmr[:, None] == lmr[None, :])
                los = self.losize[is_]
                lorows = (self.lobase[at][:, None] + np.arange(los)[None, :]).astype(np.int64)
                out.append((torch.as_tensor(at, device=dev), f.clamp(min=0), same, los,
                            torch.as_tensor(lorows, device=dev)))
            self._dev[key] = out
        return self._dev[key]


def kinetic_energy(ctx, d, evec, cfunir, lofr, valence_rel_none=True, veffmt=None, lmaxvr=0,
                   veffir=None, plan=None):
    """kinetic_energy.f90 for one k-point: engyknst(1:nstfv) as a float64
    CUDA tensor.  d: lapw.DeviceKPoint whose apwalm holds match's result;
    evec: (nstfv, nmat) complex128 CUDA tensor, row = state (the memory of
    evecfv(nmat, nstfv)); cfunir, veffir, lofr, veffmt: CUDA tensors in the
    reference's layouts."""
    import torch
    s = ctx.sys
    if plan is None:
        plan = KineticPlan(s)
    dev = evec.device
    ngp, n = d.ngk, d.rank
    nst = evec.shape[0]
    natm = s.natmtot
    # kinetic_energy.f90 applies the ZORA mass factor to t_alo and t_lolo even
    # for 'none' relativity (:139-167), to t_aa only for ZORA (:94-121)
    from . import lapw
    mass = lapw.KIN_MASS_LO | (0 if valence_rel_none else lapw.KIN_MASS_APW)
    T = ctx.mt_kinetic(lofr, mass, veffmt, lmaxvr)                     # (natm, nf, nf)
    X = evec[:, :ngp].transpose(0, 1).contiguous()                     # (ngp, nst)
    # muffin-tin part
    p, lmm = plan.p, plan.lmm
    A = d.apwalm.view(natm, lmm * p, s.ngkmax)[:, :, :ngp]             # rows (lm, io)
    zm = torch.matmul(A, X)                                            # (natm, lmm p, nst)
    u = torch.zeros((natm, plan.dim, nst), dtype=torch.complex128, device=dev)
    u[:, :lmm * p] = zm
    xlo = evec[:, ngp:].transpose(0, 1)                                # (nlotot, nst)
    M = torch.zeros((natm, plan.dim, plan.dim), dtype=torch.float64, device=dev)
    zero = torch.zeros((), device=dev, dtype=torch.float64)
    for at, fi, same, los, lorows in plan.on(dev):     # one pass per species
        if los:
            u[at, lmm * p:lmm * p + los] = xlo[lorows]
        M[at] = torch.where(same[None], T[at][:, fi[:, None], fi[None, :]], zero)
    Mu = torch.matmul(M.to(torch.complex128), u)
    e_mt = (u.conj() * Mu).sum(dim=(0, 1)).real
    # interstitial part
    ngrid = tuple(int(x) for x in s.intgv[:, 1] - s.intgv[:, 0] + 1)
    from .forces import device_tables
    igp_t = d.igpig[:ngp].long() - 1
    cfun = cfunir
    if not valence_rel_none:
        cfun = cfunir / (1.0 - 0.5 * ALPHA * ALPHA * veffir)
    cf = cfun.view(ngrid[2], ngrid[1], ngrid[0])
    vg = d.vgpc.view(-1, 3)[:ngp]                                      # (ngp, 3)
    # The reference's backward FFT, product with cf on the grid and forward
    # FFT (scaled 1/N) is a cyclic convolution: at G_i it gives
    # sum_j c((G_i - G_j) mod box) z(G_j), with c the forward transform of
    # cf.  On the G+k set that is the dense matrix F(i, j) = c((G_i - G_j)
    # mod box), exact whatever the G differences, so one transform of cf and
    # one matrix product over the three directions replace 2 x 3 x nstfv
    # grid transforms.
    c = torch.fft.fftn(cf.to(torch.complex128), norm="forward").reshape(-1)
    ivk = device_tables(ctx, dev)["ivg"][:, igp_t]                       # (3, ngp)
    box = torch.zeros((ngp, ngp), dtype=torch.int64, device=dev)
    for a in (2, 1, 0):                       # first grid index fastest
        da = torch.remainder(ivk[a][:, None] - ivk[a][None, :], ngrid[a])
        box = box * ngrid[a] + da
    F = c[box]                                                          # (ngp, ngp)
    del box
    W = torch.cat([X * vg[:, dd:dd + 1] for dd in range(3)], dim=1)     # (ngp, 3 nst)
    Y = F @ W
    e_ist = 0.5 * (W.conj() * Y).real.reshape(ngp, 3, nst).sum(dim=(0, 1))
    return e_mt + e_ist
