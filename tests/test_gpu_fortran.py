"""The Fortran ISO_C_BINDING path (exciting_amd/fortran/exh_lapw.f90) end to
end: a flang-built program calls match and hamiltonandoverlapsetup, then the
interstitial producers (gencfun, genveffig, genmeffig), through the binding
module on the GPU; apwalm, H, O and the producers' outputs are checked
against the oracle to the same 1e-12 relative Frobenius bar as the Python
path."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

import oracle
from exciting_amd import build as exhbuild
from exciting_amd import synth
from tests import _cases


def write_input(path, s, ik=0):
    k = s.kpts[ik]
    nbox = int(np.prod(s.ivgig.shape))
    i32 = lambda a: np.asarray(a, dtype=np.int32).ravel(order="F").tobytes()
    f64 = lambda a: np.asarray(a, dtype=np.float64).ravel(order="F").tobytes()
    c128 = lambda a: np.asarray(a, dtype=np.complex128).ravel(order="F").tobytes()
    hdr = [s.nspecies, s.natmtot, s.lmaxapw, s.lmaxmat, s.apwordmax, s.nlomax, s.ngkmax,
           s.nrmtmax, s.maxaa, max(s.maxnlo, 1), s.ngrtot, s.ngvec, k.ngk, nbox, s.nlotot]
    with open(path, "wb") as f:
        f.write(i32(hdr))
        f.write(i32(s.natoms) + i32(s.apword) + i32(s.nlorb) + i32(s.lorbl) + f64([s.omega])
                + f64(s.rmt) + i32(s.nrmt) + f64(s.spr) + f64(s.apwfr))
        f.write(c128(s.aa) + c128(s.loa) + c128(s.lolo) + f64(s.oalo) + f64(s.ololo))
        f.write(i32(s.ivg) + i32(s.intgv) + i32(s.ivgig) + c128(s.cfunig) + c128(s.veffig))
        f.write(f64(k.gpc) + f64(k.tpgpc) + c128(k.sfacgp) + i32(k.igpig) + f64(k.vgpc))
        maxatoms = max(s.natoms)
        atposc = np.zeros((3, maxatoms, s.nspecies), order="F")
        ias = 0
        for is_, na in enumerate(s.natoms):
            atposc[:, :na, is_] = s.atposc[ias:ias + na].T
            ias += na
        f.write(i32([maxatoms]) + f64(s.bvec) + f64(atposc) + f64(veffir_of(s)))
        f.write(i32([s.lmaxvr]) + f64(s.veffmt) + f64(s.lofr))


def veffir_of(s):
    return -0.4 + 0.3 * np.random.default_rng(9).uniform(-1, 1, s.ngrtot)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny", "fcc2"])
def test_fortran_driver(tmp_path, name):
    if not os.path.exists(exhbuild.DRIVER):
        pytest.fail("Fortran driver not built (python -m exciting_amd.build)")
    s = synth.add_mt_potential(dataclasses.replace(_cases.system(name)), lmaxvr=5, seed=3)
    k = s.kpts[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, s)
    print(f"[fortran] case {name}", flush=True)
    r = subprocess.run([exhbuild.DRIVER, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lmm = (s.lmaxapw + 1) ** 2
    n = k.ngk + s.nlotot
    raw = np.fromfile(fout, dtype=np.complex128)
    na = s.ngkmax * s.apwordmax * lmm * s.natmtot
    A = raw[:na].reshape((s.ngkmax, s.apwordmax, lmm, s.natmtot), order="F")
    H = raw[na:na + n * n].reshape((n, n), order="F")
    O = raw[na + n * n:na + 2 * n * n].reshape((n, n), order="F")
    rest = np.fromfile(fout, dtype=np.uint8)[16 * (na + 2 * n * n):]
    ng, nv = s.ngrtot, s.ngvec
    mnlo = max(s.maxnlo, 1)
    nt = [s.maxaa * s.maxaa * s.natmtot, mnlo * s.maxaa * s.natmtot, mnlo * mnlo * s.natmtot]
    cuts = np.cumsum([0, 16 * ng, 8 * ng, 16 * nv, 16 * nv, 16 * nv])
    assert rest.size == cuts[-1] + 16 * sum(nt)
    tabs = rest[cuts[-1]:].view(np.complex128)
    tabs = np.split(tabs, np.cumsum(nt)[:-1])
    cg, cr, vg, mg, m2g = (rest[a:b].view(dt) for a, b, dt in
                           zip(cuts[:-1], cuts[1:], (np.complex128, np.float64, np.complex128,
                                                     np.complex128, np.complex128)))
    S = oracle.System(s)
    A0 = oracle.match(S, 0)
    H0, O0 = oracle.setup(S, 0, A0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  fortran {name}: A {rel(A, A0):.2e} H {rel(H, H0):.2e} O {rel(O, O0):.2e}")
    assert rel(A, A0) <= 1e-12 and rel(H, H0) <= 1e-12 and rel(O, O0) <= 1e-12
    cg0, cr0 = oracle.gencfun(s)
    v = veffir_of(s)
    vg0 = oracle.genveffig(s, v, cr0)
    mg0, m2g0 = oracle.genmeffig(s, v, cr0, True)
    errs = [rel(a, b) for a, b in ((cg, cg0), (cr, cr0), (vg, vg0), (mg, mg0), (m2g, m2g0))]
    print("  fortran producers: " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= 1e-12
    want = [t.ravel(order="F") for t in oracle.hmlint(oracle.System(s))]
    terr = [np.abs(g - w).max() / max(np.abs(w).max(), 1e-300) for g, w in zip(tabs, want)]
    print("  fortran hmlint: " + " ".join(f"{e:.1e}" for e in terr))
    assert max(terr) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tiny", "tiny_lmaxmat", "fcc2"])
def test_reference_name_shims(tmp_path, name):
    """exh_shims.f90 called as exciting calls the reference routines:
    match, then hamiltonandoverlapsetup on a newsystem (seceqn.f90:57,
    seceqnfv.f90:73-77), then overlapsetup alone (calc_vnlmat.f90:59-60),
    then hamiltonandoverlapsetup again (accumulating).  Each result is within
    1e-12 of the oracle, and apwalm crossed PCIe only once (in match's
    download): the three setups reused the copy in HBM."""
    if not os.path.exists(exhbuild.SHIM_DRIVER):
        pytest.fail("shim driver not built (python -m exciting_amd.build)")
    s = synth.add_mt_potential(dataclasses.replace(_cases.system(name)), lmaxvr=5, seed=3)
    k = s.kpts[0]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, s)
    print(f"[fortran-shims] case {name}", flush=True)
    r = subprocess.run([exhbuild.SHIM_DRIVER, str(fin), str(fout)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lmm = (s.lmaxapw + 1) ** 2
    n = k.ngk + s.nlotot
    na = s.ngkmax * s.apwordmax * lmm * s.natmtot
    raw = np.fromfile(fout, dtype=np.uint8)
    z = raw[:16 * (na + 5 * n * n)].view(np.complex128)
    uploads, reused = raw[16 * (na + 5 * n * n):].view(np.int64)
    A = z[:na].reshape((s.ngkmax, s.apwordmax, lmm, s.natmtot), order="F")
    H, O, Oalone, H2, O2 = (z[na + i * n * n:na + (i + 1) * n * n].reshape((n, n), order="F")
                            for i in range(5))
    S = oracle.System(s)
    A0 = oracle.match(S, 0)
    H0, O0 = oracle.setup(S, 0, A0)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    errs = {"apwalm": rel(A, A0), "H": rel(H, H0), "O": rel(O, O0),
            "O alone": rel(Oalone, O0), "H twice": rel(H2, 2 * H0), "O twice": rel(O2, 2 * O0)}
    print(f"  shims {name}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items())
          + f"; apwalm uploads {uploads}, reused {reused}")
    assert max(errs.values()) <= 1e-12
    assert uploads == 0 and reused == 3
