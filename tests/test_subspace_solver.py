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
