"""The identity the kinetic-energy path relies on (CPU, PyTorch): the
reference's interstitial chain (x on the grid at the G+k, backward FFT,
product with the grid function, forward FFT scaled 1/N, gather at the G+k,
kinetic_energy.f90:208-234) equals the dense matrix F(i, j) = c((G_i - G_j)
mod box) applied to x, c being the forward transform of the grid function,
also when the G differences wrap around the box."""
import itertools

import numpy as np
import torch


def test_fft_chain_is_the_dense_cyclic_convolution():
    rng = np.random.default_rng(0)
    n = (6, 7, 5)
    N = n[0] * n[1] * n[2]
    pts = np.array(list(itertools.product(range(-2, 4), range(-3, 4), range(-2, 3)))).T
    iv = pts[:, rng.choice(pts.shape[1], 40, replace=False)]   # distinct box points
    ng, nst = iv.shape[1], 3
    cf = torch.tensor(rng.normal(size=(n[2], n[1], n[0])))
    X = torch.tensor(rng.normal(size=(ng, nst)) + 1j * rng.normal(size=(ng, nst)))
    gi = torch.tensor(((iv[2] % n[2]) * n[1] + (iv[1] % n[1])) * n[0] + (iv[0] % n[0]))
    z = torch.zeros((nst, N), dtype=torch.complex128)
    z[:, gi] = X.T
    z = torch.fft.ifftn(z.view(nst, n[2], n[1], n[0]), dim=(1, 2, 3), norm="forward") * cf
    chain = torch.fft.fftn(z, dim=(1, 2, 3), norm="forward").reshape(nst, N)[:, gi].T
    c = torch.fft.fftn(cf.to(torch.complex128), norm="forward").reshape(-1)
    ivk = torch.tensor(iv)
    box = torch.zeros((ng, ng), dtype=torch.int64)
    for a in (2, 1, 0):
        box = box * n[a] + torch.remainder(ivk[a][:, None] - ivk[a][None, :], n[a])
    dense = c[box] @ X
    d = ivk[:, :, None] - ivk[:, None, :]
    assert bool((d.abs() > torch.tensor(n)[:, None, None] // 2).any())  # some differences wrap
    assert float((dense - chain).abs().max()) <= 1e-13 * float(chain.abs().max())
