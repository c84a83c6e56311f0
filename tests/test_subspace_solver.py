"""The few-state eigensolver (solve.lowest_eigvalsh: Chebyshev-filtered
subspace iteration) on the CPU with PyTorch tensors, against the full
Hermitian eigensolve: spectra shaped like an LAPW problem (a few dozen low
states below a wide band), a degenerate cluster straddling the block edge,
the small-matrix path and determinism."""
import numpy as np
import pytest
import torch

from exciting_amd import solve


def hermitian_with(ev, seed):
    rng = np.random.default_rng(seed)
    n = len(ev)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    C = torch.tensor((Q * np.asarray(ev)) @ Q.conj().T)
    return 0.5 * (C + C.conj().T)


@pytest.mark.parametrize("nwant", [8, 40])
def test_lowest_eigenvalues_of_an_lapw_like_spectrum(nwant):
    rng = np.random.default_rng(1)
    n = 1200
    ev = np.sort(np.concatenate([rng.uniform(-3, 1, 60), rng.uniform(0.5, 40, n - 60)]))
    C = hermitian_with(ev, 2)
    w = solve.lowest_eigvalsh(C, nwant)
    ref = torch.linalg.eigvalsh(C)[:nwant]
    assert float((w - ref).abs().max()) <= 1e-11


def test_degenerate_cluster_at_the_block_edge():
    """Eigenvalues repeated across the wanted/unwanted boundary and inside
    the block: the residual test still converges to the right values."""
    n = 900
    low = np.concatenate([np.linspace(-2.0, -0.5, 30), np.full(12, -0.25), [0.1] * 6])
    ev = np.sort(np.concatenate([low, np.linspace(0.3, 25.0, n - len(low))]))
    C = hermitian_with(ev, 3)
    nwant = 36                               # cuts through the 12-fold cluster
    w = solve.lowest_eigvalsh(C, nwant)
    assert np.abs(w.numpy() - ev[:nwant]).max() <= 1e-11


def test_small_matrices_take_the_full_solve_and_results_repeat():
    C = hermitian_with(np.linspace(-1, 3, 200), 4)
    w = solve.lowest_eigvalsh(C, 10)
    assert torch.equal(w, torch.linalg.eigvalsh(C)[:10])
    D = hermitian_with(np.linspace(-1, 30, 800), 5)
    assert torch.equal(solve.lowest_eigvalsh(D, 20), solve.lowest_eigvalsh(D, 20))


@pytest.mark.parametrize("nwant", [8, 40])
def test_ritz_vectors_span_the_lowest_eigenspaces(nwant):
    """With vectors: orthonormal, residuals C x - w x within the tolerance,
    and the block spans the exact lowest eigenvectors (singular values of
    V^H X equal 1)."""
    rng = np.random.default_rng(6)
    n = 1200
    ev = np.sort(np.concatenate([rng.uniform(-3, 1, 60), rng.uniform(0.5, 40, n - 60)]))
    C = hermitian_with(ev, 7)
    w, X = solve.lowest_eigh(C, nwant, rtol=1e-12, atol=1e-12)
    assert X.shape == (n, nwant)
    assert float((X.conj().T @ X - torch.eye(nwant, dtype=X.dtype)).abs().max()) <= 1e-12
    R = C @ X - X * w[None, :].to(X.dtype)
    assert float(torch.linalg.vector_norm(R, dim=0).max()) <= 1e-11
    wf, V = torch.linalg.eigh(C)
    assert float((w - wf[:nwant]).abs().max()) <= 1e-12
    sv = torch.linalg.svdvals(V[:, :nwant].conj().T @ X)
    assert float((sv - 1).abs().max()) <= 1e-9


@pytest.mark.parametrize("kappa", [1e2, 1e7, 1e13])
def test_orthonormalize_keeps_span_at_any_conditioning(kappa):
    """Cholesky QR twice where the block is well conditioned, Householder QR
    where it is not: orthonormal to rounding either way, and the same column
    space (Q Q^H X = X) while the block has full numerical rank."""
    rng = np.random.default_rng(3)
    n, m = 900, 60
    U, _ = np.linalg.qr(rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m)))
    V, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    X = torch.from_numpy((U * np.geomspace(1.0, 1.0 / kappa, m)) @ V.conj().T)
    Q = solve._orthonormalize(X)
    assert float((Q.conj().T @ Q - torch.eye(m, dtype=Q.dtype)).abs().max()) <= 1e-13
    if kappa < 1e10:
        P = Q @ (Q.conj().T @ X)
        assert float((P - X).abs().max()) <= 1e-12 * float(X.abs().max())


def test_block_and_degree_rules():
    assert solve.block_size(8) == 24 and solve.block_size(40) == 56
    assert solve.block_size(400) == 560
    assert solve.filter_degree(40) == 32 and solve.filter_degree(400) == 24
