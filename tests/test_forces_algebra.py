"""The algebra the GPU forcek rests on (CPU, NumPy): with the reference's
derivative matrices dlh = i (g_i - g_j) M_H and dlo = -i (g_i - g_j) M_O
(M Hermitian, g zero on the local-orbital rows), the state sum
Re sum_ab ffv_ab dm_ba of ffv = X^H (dlh X + dlo X E) (forcek.f90:286-372)
equals -2 sum_i g_i Im sum_b conj(Z_H) (M_H X) + sum_i g_i Im sum_b
conj(Z_O) (M_O X) with Z_H = X dm and Z_O = X (dm E + E dm); for
dm = diag(occ) it is the first-variational sum of Re ffv(j, j)."""
import numpy as np
import pytest


def herm(rng, n):
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return a + a.conj().T


def explicit(MH, MO, g, X, e, dm):
    G = g[:, None] - g[None, :]
    ffv = X.conj().T @ ((1j * G * MH) @ X + (-1j * G * MO) @ X * e[None, :])
    return np.real(np.sum(ffv * dm.T)), ffv


def commutator_form(MH, MO, g, X, e, dm):
    ZH = X @ dm
    ZO = X @ (dm * e[None, :] + e[:, None] * dm)
    return (-2 * np.sum(g * np.imag(np.sum(ZH.conj() * (MH @ X), axis=1)))
            + np.sum(g * np.imag(np.sum(ZO.conj() * (MO @ X), axis=1))))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_density_matrix_form_equals_the_explicit_state_sum(seed):
    rng = np.random.default_rng(seed)
    n, ng, nst, nsv = 14, 10, 5, 9
    MH, MO = herm(rng, n), herm(rng, n)
    g = np.zeros(n)
    g[:ng] = rng.normal(size=ng)
    X = rng.normal(size=(n, nst)) + 1j * rng.normal(size=(n, nst))
    e = rng.normal(size=nst)
    V = rng.normal(size=(nst, nsv)) + 1j * rng.normal(size=(nst, nsv))
    occ = rng.uniform(0, 1, nsv)
    dm = (V * occ) @ V.conj().T
    want, _ = explicit(MH, MO, g, X, e, dm)
    assert commutator_form(MH, MO, g, X, e, dm) == pytest.approx(want, rel=1e-12)
    # first-variational case: dm = diag(occ) gives sum_j occ_j Re ffv(j, j)
    occ1 = rng.uniform(0, 2, nst)
    _, ffv = explicit(MH, MO, g, X, e, np.eye(nst))
    fv = np.sum(occ1 * np.real(np.diag(ffv)))
    assert commutator_form(MH, MO, g, X, e, np.diag(occ1)) == pytest.approx(fv, rel=1e-12)
