"""The k-point contribution to the basis-set-correction forces, forcek.f90,
on the GPU: first-variational states, or second-variational ones through
their density matrix in the first-variational basis (tevecsv, normal
spin-polarised or spin-spiral).

For atom alpha and direction l the reference forms dlh = i (g_i - g_j) M_H
and dlo = -i (g_i - g_j) M_O over the basis (g the l-th component of G+k on
the APW rows, 0 on the local-orbital rows, the G-G block through the G
difference), with M_H = H_alpha + dpij zt1 and M_O = O_alpha + zt1 (the
atom's muffin-tin and lo-border parts, and zt1 = -ffacg conj(sfacg) of the
difference G, its step-function term), then ffv = X^H (dlh X + dlo X e)
and adds wkpt sum_j occ_j Re ffv(j, j) (forcek.f90:259-335).

Because the derivative matrices are commutators with diag(g), the diagonal
needs only Y = M X:

    Re ffv(j, j) = -2 sum_i g_i Im(conj(X_ij) Y_H,ij) + 2 e_j sum_i g_i Im(conj(X_ij) Y_O,ij)

summed over the APW rows i.  Y is built without any n x n matrix:
  * muffin-tin and lo-border parts per atom from B = A X (A the atom's
    matching coefficients, l <= lmaxmat): A^H (aa B + loa^H x_lo) and
    A^H B + O_GL x_lo, as batched library GEMMs over all atoms;
  * the step-function term per species: zt1(G_i - G_j) =
    conj(p_i) F_s(i, j) p_j with p = exp(i G.tau) and the real matrix
    F_s(i, j) = -ffacg(|G_i - G_j|) (dpij F_s for H), so
    Y += conj(p) * (F_s (p * X)) for all atoms of a species in one GEMM.
The G differences must lie inside the FFT box (gmaxvr >= 2 gkmax, as
exciting's defaults make it); the reference folds them back into the box
otherwise, which this path refuses.
"""
from __future__ import annotations

import numpy as np


def _rows(s):
    """Per species: indices into apwalm's (lm, io) rows of the atom's APW
    functions with l <= lmaxmat (forcek.f90:71-82 order), and per LO column
    (ilo, m) its lm and ilo."""
    p = s.apwordmax
    out = []
    for is_ in range(s.nspecies):
        r = [(l * (l + 1) + m) * p + io for l in range(s.lmaxmat + 1) for m in range(-l, l + 1)
             for io in range(s.apword[l, is_])]
        lo = [(l * (l + 1) + m, ilo) for ilo in range(s.nlorb[is_])
              for l in [s.lorbl[ilo, is_]] for m in range(-l, l + 1)]
        out.append((np.array(r, np.int64), lo))
    return out


def device_tables(ctx, dev):
    """The per-structure tables forcek and kinetic_energy read, resident on
    ``dev`` and kept on the context: uploaded on first use, again only when
    the context is loaded with another structure or other host muffin-tin
    tables.  ivg (3, ngrtot) and ivgig (flattened box) as int64; the
    context's host aa, loa and salo = oalo + h1loa in [ias, ...] order; per
    species the atom indices, APW row indices and LO columns."""
    import torch
    s = ctx.sys
    key = (str(dev), id(s), id(s.aa), id(s.loa), id(s.oalo), id(s.h1loa))
    c = getattr(ctx, "_device_tables", None)
    if c is not None and c["key"] == key:
        return c
    species = np.repeat(np.arange(s.nspecies), s.natoms)
    salo = np.asarray(s.oalo) + (np.asarray(s.h1loa) if s.h1loa is not None else 0.0)
    c = {
        "key": key,
        "ivg": torch.as_tensor(np.asarray(s.ivg, dtype=np.int64), device=dev),
        "ivgig": torch.as_tensor(np.asarray(s.ivgig, dtype=np.int64).ravel(order="F"),
                                 device=dev),
        "aa": torch.as_tensor(np.ascontiguousarray(np.moveaxis(s.aa, 2, 0)), device=dev),
        "loa": torch.as_tensor(np.ascontiguousarray(np.moveaxis(s.loa, 2, 0)), device=dev),
        "salo": torch.as_tensor(np.ascontiguousarray(np.moveaxis(salo, (0, 1, 2), (2, 1, 0))),
                                device=dev),                        # [ias, ilo, io]
        "lobase": np.concatenate([[0], np.cumsum([s.losize[i] for i in species])])[:-1],
        "species": [],
    }
    for is_, (r_idx, lo_cols) in enumerate(_rows(s)):
        at = np.nonzero(species == is_)[0]
        c["species"].append((at, torch.as_tensor(at, device=dev),
                             torch.as_tensor(r_idx, device=dev), len(r_idx), lo_cols))
    ctx._device_tables = c
    return c


def density_matrix(evecsv, occsv, nstfv, nspinor=2, spiral_ispn=None):
    """The state weights of the second-variational force sum
    (forcek.f90:341-366): dm = sum_j occsv_j sum_b v_bj v_bj^H over the
    nspinor blocks b of nstfv rows of evecsv(nstsv, nstsv), or over the one
    block of first-variational spin spiral_ispn (0-based) for a spin spiral.
    CUDA tensors; returns (nstfv, nstfv) complex128."""
    blocks = [spiral_ispn] if spiral_ispn is not None else range(nspinor)
    dm = None
    for b in blocks:
        v = evecsv[b * nstfv:(b + 1) * nstfv]
        t = (v * occsv[None, :].to(v.dtype)) @ v.conj().transpose(0, 1)
        dm = t if dm is None else dm + t
    return dm


def forcek(ctx, d, evec, evalfv, occ, wkpt, ffacg, sfacg, dm=None, tables=None):
    """forcek.f90 for one k-point: a (3, natmtot) float64 CUDA tensor to add
    to forceibs.  d: lapw.DeviceKPoint (match is run here, as the reference
    does); evec (nstfv, nmat) complex128 (row = state); evalfv, occ (nstfv)
    float64; ffacg (ngvec, nspecies) float64 (genffacg); sfacg
    (ngvec, natmtot) complex128, exp(i G.tau).  CUDA tensors throughout.

    With second-variational states (tevecsv) pass dm = density_matrix(...)
    (occ is then not read): the state sum sum_ab Re ffv_ab dm_ba becomes
    -2 sum_i g_i Im(conj(Z_H) Y_H) + sum_i g_i Im(conj(Z_O) Y_O) summed over
    states, with Z_H = X dm and Z_O = X (dm E + E dm), E = diag(evalfv),
    which for dm = diag(occ) is the first-variational sum.

    tables: optional (aa, loa, salo) CUDA tensors holding the muffin-tin
    tables in the reference's Fortran layouts, flattened (aa(maxaa, maxaa,
    natmtot), loa(maxnlo, maxaa, natmtot), salo = oalo + h1loa
    (apwordmax, nlomax, natmtot)), e.g. the outputs of hmlint on the device;
    without them the context's host tables, kept on the device with the
    other per-structure tables (device_tables), are used."""
    import torch
    s = ctx.sys
    dev = evec.device
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, d.apwalm)
    ngk, natm, p = d.ngk, s.natmtot, s.apwordmax
    lmm = (s.lmaxapw + 1) ** 2
    dt = device_tables(ctx, dev)
    igp_t = d.igpig[:ngk].long() - 1                                   # G index of each row
    # index of G_i - G_j for every pair of rows, on the device (ngk^2 entries)
    ivk = dt["ivg"][:, igp_t]                                          # (3, ngk)
    lo = [int(x) for x in np.asarray(s.intgv)[:, 0]]
    hi = [int(x) for x in np.asarray(s.intgv)[:, 1]]
    nb = [h - l + 1 for l, h in zip(lo, hi)]
    box = torch.zeros((), dtype=torch.int64, device=dev)
    for a in (2, 1, 0):                      # ivgig(i1,i2,i3): first index fastest
        da = ivk[a][:, None] - ivk[a][None, :]
        if bool(((da < lo[a]) | (da > hi[a])).any()):
            raise NotImplementedError("G+k differences outside the FFT box (gmaxvr < 2 gkmax)")
        box = box * nb[a] + (da - lo[a])
    ig_t = dt["ivgig"][box] - 1                                           # (ngk, ngk)
    if int(ig_t.max()) >= s.ngvec:
        raise NotImplementedError("G differences beyond ngvec")
    XG = evec[:, :ngk].transpose(0, 1).contiguous()                     # (ngk, nst)
    nst = XG.shape[1]
    vg = d.vgpc.view(-1, 3)[:ngk]                                        # (ngk, 3) G+k
    dp = 0.5 * (vg @ vg.transpose(0, 1))                                 # dpij
    A_all = d.apwalm.view(natm, lmm * p, s.ngkmax)
    mnlo = max(s.maxnlo, 1)
    if tables is not None:
        aa_t, loa_t, salo_t = tables
        aa = aa_t.view(natm, s.maxaa, s.maxaa).transpose(1, 2)         # [ias, i, j]
        loa = loa_t.view(natm, s.maxaa, mnlo).transpose(1, 2)          # [ias, ilo row, j]
        # [ias, ilo, io]; never read without local orbitals
        salo_d = salo_t.view(natm, s.nlomax, p) if s.nlomax > 0 else None
    else:
        aa, loa, salo_d = dt["aa"], dt["loa"], dt["salo"]
    YH = torch.zeros((natm, ngk, nst), dtype=torch.complex128, device=dev)
    YO = torch.zeros_like(YH)
    lobase = dt["lobase"]
    ngk_lo = ngk
    for is_ in range(s.nspecies):
        at, at_t, r_t, nr, lo_cols = dt["species"][is_]
        A = A_all[at_t][:, r_t, :ngk]                                    # (na, nr, ngk)
        B = torch.matmul(A, XG)                                          # (na, nr, nst)
        Ah = A.conj().transpose(1, 2)
        YH[at_t] = torch.matmul(Ah, torch.matmul(aa[at_t][:, :nr, :nr], B))
        YO[at_t] = torch.matmul(Ah, B)
        los = int(s.losize[is_])
        if los:
            XL = torch.stack([evec[:, ngk_lo + int(lobase[a]):ngk_lo + int(lobase[a]) + los]
                              .transpose(0, 1) for a in at])             # (na, los, nst)
            # H(G, lo) = conj(loa A)^T: Y_H += A^H loa^H x_lo (forcek.f90:112-126)
            YH[at_t] += torch.matmul(Ah, torch.matmul(loa[at_t][:, :los, :nr].conj()
                                                      .transpose(1, 2), XL))
            # O(G, lo) = conj(sum_io apwalm(G, io, lm, alpha) s(io, ilo)) (:170-183)
            OGL = torch.zeros((len(at), ngk, los), dtype=torch.complex128, device=dev)
            for c, (lm, ilo) in enumerate(lo_cols):
                l = s.lorbl[ilo, is_]
                for io in range(s.apword[l, is_]):
                    sv = salo_d[at_t, ilo, io]
                    OGL[:, :, c] += A_all[at_t, lm * p + io, :ngk] * sv[:, None]
            YO[at_t] += torch.matmul(OGL.conj(), XL)
        # step-function term of the species: F(i, j) = -ffacg(G_i - G_j)
        F = -ffacg[ig_t, is_]                                           # (ngk, ngk) real
        ph = sfacg[igp_t][:, at_t].transpose(0, 1)                       # (na, ngk)
        PX = ph[:, :, None] * XG[None]                                   # p * X
        W = PX.transpose(0, 1).reshape(ngk, -1)                          # (ngk, na nst)
        # one real GEMM: [F; dpij F] (2 ngk x ngk) times [Re W, Im W]
        FF = torch.cat([F, dp * F], dim=0)
        R = FF @ torch.cat([W.real, W.imag], dim=1)                     # (2 ngk, 2 na nst)
        m = W.shape[1]
        FW = torch.complex(R[:ngk, :m], R[:ngk, m:]).reshape(ngk, len(at), nst).transpose(0, 1)
        FHW = torch.complex(R[ngk:, :m], R[ngk:, m:]).reshape(ngk, len(at), nst).transpose(0, 1)
        YO[at_t] += ph.conj()[:, :, None] * FW
        YH[at_t] += ph.conj()[:, :, None] * FHW
    if dm is None:
        # Re ffv(j, j) per atom and direction
        im_h = (XG.conj()[None] * YH).imag                               # (natm, ngk, nst)
        im_o = (XG.conj()[None] * YO).imag
        fh = -2.0 * torch.einsum("il,aij->alj", vg, im_h)                # (natm, 3, nst)
        fo = 2.0 * torch.einsum("il,aij->alj", vg, im_o)
        per_state = fh + fo * evalfv[None, None, :]
        return wkpt * (per_state * occ[None, None, :]).sum(dim=2).transpose(0, 1)
    # sum_ab Re ffv_ab dm_ba
    dmc = dm.to(XG.dtype)
    ev = evalfv.to(XG.dtype)
    ZH = XG @ dmc
    ZO = XG @ (dmc * ev[None, :] + ev[:, None] * dmc)
    im_h = (ZH.conj()[None] * YH).imag.sum(dim=2)                         # (natm, ngk)
    im_o = (ZO.conj()[None] * YO).imag.sum(dim=2)
    f = -2.0 * (im_h @ vg) + im_o @ vg                                    # (natm, 3)
    return wkpt * f.transpose(0, 1)
