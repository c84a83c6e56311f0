"""Parity at the workloads the throughput figures are quoted on.

* 8-atom simple cubic cell, k-points 4 and 7 (the others are covered by
  test_gpu_parity and test_gpu_batch): FP64 within 1e-12 of the oracle.
* 16-atom orthorhombic cell, 128 uneven k-points (ngk 3313..3342):
  k-points 0 and 77 through the single-k path and through one batch, FP64
  within 1e-12 of the oracle; split precision (8 moduli) with its lowest 400
  eigenvalues within 1 uHa of those of the oracle's H and O; all 128
  k-points as one batch bitwise equal to the single path, FP64 and split.
* 64-atom supercell, k-point 0 (ngk 9340, n 10364): FP64 H and O within
  1e-12 of the oracle, through the single-k path and a one-k batch (bitwise
  equal); split (8 moduli) lowest 400 eigenvalues within 1 uHa of the
  oracle's.  The subspace eigensolver's 40 lowest eigenvalues within 1e-10 Ha
  of the full solve on the same H and O.

Eigenvalues of both sides come from solve.geneig on the GPU in FP64
(Cholesky reduction + Hermitian eigensolve), which test_gpu_solve pins to
LAPACK's ZHEGVX.  The contract is hamiltonsetup.f90:53-168 +
overlapsetup.f90:51-166 for one k-point; matrices are compared in HBM as
relative Frobenius norms.
"""
import numpy as np
import pytest

import oracle
from exciting_amd import lapw, solve, synth
from tests import _cases

TOL_MAT = 1e-12
TOL_SPLIT_HA = 1e-6
NMOD = 8
NSTATES = 400


def log(msg):
    print(msg, flush=True)


def dev(M):
    """Oracle matrix (Fortran order) -> flat CUDA tensor in column-major memory."""
    import torch
    return torch.from_numpy(np.asfortranarray(M).ravel(order="F")).cuda()


def drel(a, b):
    import torch
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def zeros(n):
    import torch
    return torch.zeros(n * n, dtype=torch.complex128, device="cuda")


def device_setup(ctx, d, flags, H=None, O=None, apwalm=None):
    """match_device + setup_device for one DeviceKPoint (overwrite)."""
    import torch
    H = zeros(d.rank) if H is None else H
    O = zeros(d.rank) if O is None else O
    a = d.apwalm if apwalm is None else apwalm
    ctx.match(d.ngk, d.gpc, d.tpgpc, d.sfacgp, a)
    ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(d.rank, H),
                                              lapw.HermitianMatrix(d.rank, O)),
                                d.ngk, a, d.igpig, d.vgpc, flags=flags)
    torch.cuda.synchronize()
    return H, O


@pytest.mark.gpu
def test_sc8_kpoints_4_and_7():
    s = synth.build(synth.config(1), kpoints=[4, 7])
    S = oracle.System(s)
    ctx = lapw.Context(0).load(s)
    try:
        for ik, k in enumerate(s.kpts):
            A0 = oracle.match(S, ik)
            H0, O0 = oracle.setup(S, ik, A0)
            A = np.zeros_like(A0, order="F")
            ctx.match(k.ngk, k.gpc, k.tpgpc, k.sfacgp, A)
            sysg = lapw.newsystem(k.ngk + s.nlotot)
            ctx.hamiltonandoverlapsetup(sysg, k.ngk, A, k.igpig, k.vgpc)
            ea = np.linalg.norm(A - A0) / np.linalg.norm(A0)
            eh = np.linalg.norm(sysg.hamilton.za - H0) / np.linalg.norm(H0)
            eo = np.linalg.norm(sysg.overlap.za - O0) / np.linalg.norm(O0)
            log(f"  sc8 k{(4, 7)[ik]}: apwalm {ea:.1e} H {eh:.1e} O {eo:.1e}")
            assert ea <= TOL_MAT and eh <= TOL_MAT and eo <= TOL_MAT
    finally:
        ctx.close()


class Ortho16:
    """All 128 k-points of the 16-atom cell resident on the GPU."""

    def __init__(self):
        import torch
        self.s = synth.build(synth.config(2))
        log(f"[gpu-configs] ortho16: {len(self.s.kpts)} k-points, ngk "
            f"{min(k.ngk for k in self.s.kpts)}..{max(k.ngk for k in self.s.kpts)}")
        self.ctx = lapw.Context(0).load(self.s)
        self.dk = [lapw.DeviceKPoint(self.s, ik) for ik in range(len(self.s.kpts))]
        self.H = [zeros(d.rank) for d in self.dk]
        self.O = [zeros(d.rank) for d in self.dk]
        self.kp = [lapw.kpoint_struct(d, h, o) for d, h, o in zip(self.dk, self.H, self.O)]
        torch.cuda.synchronize()

    def batch(self, flags):
        import torch
        for h, o in zip(self.H, self.O):
            h.fill_(7.0)          # the batch overwrites whatever is there
            o.fill_(7.0)
        self.ctx.setup_batch(self.kp, flags=flags | lapw.MATCH)
        torch.cuda.synchronize()

    def close(self):
        self.ctx.close()


@pytest.fixture(scope="module")
def ortho16():
    import torch
    c = Ortho16()
    yield c
    c.close()
    del c
    torch.cuda.empty_cache()


@pytest.mark.gpu
@pytest.mark.parametrize("ik", [0, 77])
def test_ortho16_fp64_and_split_against_the_oracle(ortho16, ik):
    import torch
    c = ortho16
    s = c.s
    S = oracle.System(s)
    A0 = oracle.match(S, ik)
    H0, O0 = oracle.setup(S, ik, A0)
    H0d, O0d = dev(H0), dev(O0)
    d = c.dk[ik]
    # single-k path
    H, O = device_setup(c.ctx, d, 0)
    A = d.apwalm.cpu().numpy().reshape(d.apwalm_shape, order="F")
    ea = np.linalg.norm(A - A0) / np.linalg.norm(A0)
    eh, eo = drel(H, H0d), drel(O, O0d)
    log(f"  ortho16 k{ik} (ngk {d.ngk}) single: apwalm {ea:.1e} H {eh:.1e} O {eo:.1e}")
    assert ea <= TOL_MAT and eh <= TOL_MAT and eo <= TOL_MAT
    # the same k-point inside a batch with its neighbour
    jk = (ik + 1) % len(c.dk)
    hb, ob = zeros(d.rank), zeros(d.rank)
    kp = [lapw.kpoint_struct(d, hb, ob), c.kp[jk]]
    c.ctx.setup_batch(kp, flags=lapw.MATCH)
    torch.cuda.synchronize()
    eh, eo = drel(hb, H0d), drel(ob, O0d)
    log(f"  ortho16 k{ik} batch: H {eh:.1e} O {eo:.1e}")
    assert eh <= TOL_MAT and eo <= TOL_MAT
    assert torch.equal(hb, H) and torch.equal(ob, O)
    # split precision (7 moduli, the tuner's choice, and 8): eigenvalues
    # against the oracle's H and O
    w0 = solve.geneig(H0d, O0d, d.rank, NSTATES).cpu().numpy()
    for nmod in (7, NMOD):
        Hs, Os = device_setup(c.ctx, d, lapw.SPLIT(nmod))
        w = solve.geneig(Hs, Os, d.rank, NSTATES).cpu().numpy()
        de = float(np.abs(w - w0).max())
        log(f"  ortho16 k{ik} split({nmod}): H {drel(Hs, H0d):.1e} O {drel(Os, O0d):.1e} "
            f"lowest {NSTATES} max|dE| {de:.2e} Ha")
        assert de <= TOL_SPLIT_HA


@pytest.mark.gpu
@pytest.mark.parametrize("flags,plan", [(0, -1), (0, 5), (lapw.SPLIT(NMOD), -1)])
def test_ortho16_all_kpoints_one_batch_is_the_single_path(ortho16, flags, plan):
    import torch
    c = ortho16
    c.ctx.batch_plan = plan
    try:
        c.batch(flags)
    finally:
        c.ctx.batch_plan = -1
    apw = torch.empty_like(c.dk[0].apwalm)
    hs, os_ = zeros(max(d.rank for d in c.dk)), zeros(max(d.rank for d in c.dk))
    for ik, d in enumerate(c.dk):
        n2 = d.rank * d.rank
        h, o = hs[:n2], os_[:n2]
        apw.zero_()
        device_setup(c.ctx, d, flags, H=h, O=o, apwalm=apw)
        assert torch.equal(apw, d.apwalm), f"k{ik}: apwalm differs"
        assert torch.equal(h, c.H[ik]), f"k{ik}: H differs"
        assert torch.equal(o, c.O[ik]), f"k{ik}: O differs"
    log(f"  ortho16 flags {flags:#x} plan {plan}: {len(c.dk)} k-points bitwise equal")


@pytest.fixture(scope="module")
def sc64():
    """k-point 0 of the 64-atom cell: the oracle's H and O (on the host, ~1 min
    with OpenBLAS on all cores) and a context holding the cell."""
    import torch
    s = _cases.system("sc64_k0")
    d0 = s.kpts[0]
    log(f"[gpu-configs] sc64 k0: ngk {d0.ngk}, n {d0.ngk + s.nlotot}; oracle running")
    S = oracle.System(s)
    A0 = oracle.match(S, 0)
    H0, O0 = oracle.setup(S, 0, A0)
    log("  oracle done")
    ref = {"A": A0, "H": dev(H0), "O": dev(O0)}
    del H0, O0
    ctx = lapw.Context(0).load(s)
    yield s, ctx, ref
    ctx.close()
    ref.clear()
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_sc64_fp64_against_the_oracle(sc64):
    import torch
    s, ctx, ref = sc64
    d = lapw.DeviceKPoint(s, 0)
    H, O = device_setup(ctx, d, 0)
    A = d.apwalm.cpu().numpy().reshape(d.apwalm_shape, order="F")
    ea = float(np.linalg.norm(A - ref["A"]) / np.linalg.norm(ref["A"]))
    del A
    eh, eo = drel(H, ref["H"]), drel(O, ref["O"])
    log(f"  sc64 k0 single: apwalm {ea:.1e} H {eh:.1e} O {eo:.1e}")
    assert ea <= TOL_MAT and eh <= TOL_MAT and eo <= TOL_MAT
    # the bench's batched path (one k-point, persistent plan) gives the same bits
    hb, ob = zeros(d.rank), zeros(d.rank)
    ctx.batch_plan = 5
    try:
        ctx.setup_batch([lapw.kpoint_struct(d, hb, ob)], flags=lapw.MATCH)
        torch.cuda.synchronize()
    finally:
        ctx.batch_plan = -1
    assert torch.equal(hb, H) and torch.equal(ob, O)
    # few states by subspace iteration against the full eigensolve, same H, O
    wf = solve.geneig(H, O, d.rank, 40, method="full").cpu().numpy()
    ws = solve.geneig(H, O, d.rank, 40, method="subspace").cpu().numpy()
    log(f"  sc64 k0 subspace vs full, 40 states: max|dE| {np.abs(ws - wf).max():.1e} Ha "
        f"({solve.last_subspace})")
    assert np.abs(ws - wf).max() <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("nmod", [7, NMOD])
def test_sc64_split_eigenvalues_against_the_oracle(sc64, nmod):
    """7 moduli: the count the precision tuner picks for this cell."""
    s, ctx, ref = sc64
    d = lapw.DeviceKPoint(s, 0)
    H, O = device_setup(ctx, d, lapw.SPLIT(nmod))
    w0 = solve.geneig(ref["H"], ref["O"], d.rank, NSTATES).cpu().numpy()
    w = solve.geneig(H, O, d.rank, NSTATES).cpu().numpy()
    de = float(np.abs(w - w0).max())
    log(f"  sc64 k0 split({nmod}): H {drel(H, ref['H']):.1e} O {drel(O, ref['O']):.1e} "
        f"lowest {NSTATES} max|dE| {de:.2e} Ha")
    assert de <= TOL_SPLIT_HA


@pytest.mark.gpu
@pytest.mark.parametrize("reuse", [False, True])
def test_ortho16_host_arrays_equal_the_device_path(ortho16, reuse):
    """The host-array entry points at a size where every big copy is staged
    through the pinned buffers in several pieces (apwalm 207 MB as 2-D
    pieces of whole columns, H and O 207 MB each as byte ranges): match into
    a NumPy apwalm, then the accumulating setup into prefilled NumPy H and O,
    bitwise equal to the device entry points on the same inputs (H and O
    prefilled the same way), with and without apwalm reuse."""
    import torch
    c = ortho16
    s, ik = c.s, 77
    k = s.kpts[ik]
    d = c.dk[ik]
    n = d.rank
    lmm = (s.lmaxapw + 1) ** 2
    rng = np.random.default_rng(11)
    pre = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(2)]
    pre = [np.asfortranarray(m + m.conj().T) for m in pre]
    A = np.zeros((s.ngkmax, s.apwordmax, lmm, s.natmtot), np.complex128, order="F")
    Hh, Oh = pre[0].copy(order="F"), pre[1].copy(order="F")
    c.ctx.set_apwalm_reuse(reuse)
    u0, r0 = c.ctx.apwalm_transfers()
    try:
        c.ctx.match(k.ngk, k.gpc, k.tpgpc, k.sfacgp, A)
        c.ctx.hamiltonandoverlapsetup(lapw.EvSystem(lapw.HermitianMatrix(n, Hh),
                                                    lapw.HermitianMatrix(n, Oh)),
                                      k.ngk, A, k.igpig, k.vgpc)
        u1, r1 = c.ctx.apwalm_transfers()
    finally:
        c.ctx.set_apwalm_reuse(False)
    assert (u1 - u0, r1 - r0) == ((0, 1) if reuse else (1, 0))
    Hd, Od = dev(pre[0]), dev(pre[1])
    device_setup(c.ctx, d, lapw.ACCUMULATE, H=Hd, O=Od)
    Ad = d.apwalm.cpu().numpy().reshape(A.shape, order="F")
    assert np.array_equal(A[:k.ngk], Ad[:k.ngk])
    assert not np.any(A[k.ngk:]), "rows beyond ngp stay untouched"
    assert np.array_equal(Hh, Hd.cpu().numpy().reshape(n, n, order="F"))
    assert np.array_equal(Oh, Od.cpu().numpy().reshape(n, n, order="F"))
    log(f"  ortho16 k{ik} host arrays (reuse {reuse}): n {n}, bitwise the device path")
