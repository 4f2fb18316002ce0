"""The brute-force budget reference of the tests, pinned by closed forms."""
import math

import numpy as np

from .helpers import budgets_numpy, params


def beltrami(p, k, b, b0=1.0, rho=1.0, pr=1.0):
    """B = B0 x_hat + b (sin kz, cos kz, 0): curl B = k (B - B0 x_hat), so
    <A.B> = b^2 / k and <J.B> = k b^2 for the fluctuating part (k = 0 left out)."""
    z = np.arange(p.nz) * p.Lz / p.nz
    shape = (p.nx, p.ny, p.nz)
    uu = np.zeros(shape + (8,))
    prim = np.zeros(shape + (4,))
    uu[..., 0] = rho
    uu[..., 4] = b0 + b * np.sin(k * z)[None, None, :]
    uu[..., 5] = b * np.cos(k * z)[None, None, :]
    prim[..., 3] = pr
    uu[..., 7] = pr / (p.adiabatic_index - 1) + 0.5 * (uu[..., 4:7] ** 2).sum(-1)
    return uu, prim


def spectrum(uu):
    """(nx/2+1, ny, nz, nf) half spectrum FFT/N, x the halved axis"""
    n = uu.shape[0] * uu.shape[1] * uu.shape[2]
    return np.stack([np.fft.rfftn(uu[..., f].T).T / n for f in range(uu.shape[-1])], axis=-1)


def test_beltrami_closed_forms():
    p = params((8, 8, 16))
    k, b = 3, 0.4
    uu, prim = beltrami(p, k, b)
    # a flow along B: u = B / sqrt(rho) gives <u.B> = <B.B>
    prim[..., 0:3] = uu[..., 4:7]
    uu[..., 1:4] = prim[..., 0:3]
    out = budgets_numpy(uu, prim, spectrum(uu), p)
    assert math.isclose(out[2], 0.5 * (1.0 + b * b), rel_tol=1e-13)   # <sin^2 + cos^2> = 1
    assert math.isclose(out[4], 1.0 + b * b, rel_tol=1e-13)
    assert math.isclose(out[5], b * b / k, rel_tol=1e-12)
    assert math.isclose(out[6], k * b * b, rel_tol=1e-12)
    assert math.isclose(out[3], 1.0 / (p.adiabatic_index - 1), rel_tol=1e-13)
    # the opposite helicity flips both signs
    uu2, prim2 = beltrami(p, k, b)
    uu2[..., 5] *= -1
    out2 = budgets_numpy(uu2, prim2, spectrum(uu2), p)
    assert math.isclose(out2[5], -b * b / k, rel_tol=1e-12)
    assert math.isclose(out2[6], -k * b * b, rel_tol=1e-12)


def test_beltrami_along_x_uses_half_spectrum_weights():
    # B = b (0, sin kx, cos kx) lives at kx = k > 0: each stored mode stands for two
    p = params((16, 8, 8))
    k, b = 2, 0.3
    x = np.arange(p.nx) * p.Lx / p.nx
    uu = np.zeros((p.nx, p.ny, p.nz, 8))
    prim = np.zeros((p.nx, p.ny, p.nz, 4))
    uu[..., 0] = 1.0
    uu[..., 5] = b * np.sin(k * x)[:, None, None]
    uu[..., 6] = b * np.cos(k * x)[:, None, None]
    prim[..., 3] = 1.0
    uu[..., 7] = 1.0 / (p.adiabatic_index - 1) + 0.5 * (uu[..., 4:7] ** 2).sum(-1)
    out = budgets_numpy(uu, prim, spectrum(uu), p)
    assert math.isclose(out[5], b * b / k, rel_tol=1e-12)
    assert math.isclose(out[6], k * b * b, rel_tol=1e-12)
