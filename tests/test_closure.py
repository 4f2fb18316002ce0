"""Learned closure, CPU side: the numpy reference of the network is pinned by finite
differences, and the training-data definitions by closed forms (constant state,
block variance, the tensor symmetry)."""
import math

import numpy as np
import pytest

from laps_amd import closure
from tests import closure_np as ref
from tests.helpers import params, seeded_state


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((40, ref.NIN))
    T = rng.standard_normal((40, ref.NOUT))
    P = closure.init_params(1).astype(np.float64)
    P[ref.NIN * ref.NH: ref.NIN * ref.NH + ref.NH] = 0.1  # nonzero b1
    loss, g = ref.loss_grad(X, T, P)
    idx = rng.choice(P.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (ref.loss_grad(X, T, Pp)[0] - ref.loss_grad(X, T, Pm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_parameter_layout():
    assert closure.init_params(0).size == 7186
    W1, b1, W2, b2, W3, b3 = ref.unpack(np.arange(7186.0))
    assert W1.shape == (28, 64) and W2.shape == (64, 64) and W3.shape == (64, 18)
    assert b3[-1] == 7185.0


def test_constant_state_has_no_subgrid_terms():
    p = params(16)
    uu = np.zeros((16, 16, 16, 8), order="F")
    uu[..., 0] = 1.3
    uu[..., 1:4] = (0.2, -0.1, 0.05)
    uu[..., 4:7] = (1.0, 0.3, -0.2)
    uu[..., 7] = 4.0
    prim = np.zeros((16, 16, 16, 4), order="F")
    prim[..., 0:3] = uu[..., 1:4] / 1.3
    # the pressure of the state (update_uu_prim_from_uu, mhdrhs.f90:282-294)
    prim[..., 3] = (p.adiabatic_index - 1) * (4.0 - 0.5 * (uu[..., 1:4] * prim[..., 0:3]).sum(-1)
                                              - 0.5 * (uu[..., 4:7] ** 2).sum(-1))
    uu_c, prim_c, T = closure.filtered_sample(uu, prim, 2, p.adiabatic_index, 0.05, (p.Lx, p.Ly, p.Lz))
    assert uu_c.shape == (8, 8, 8, 8)
    assert np.abs(T).max() < 1e-14


def test_reynolds_and_maxwell_terms_are_block_covariances():
    # rho = 1: R_ij = <m_i m_j> - <m_i><m_j> and M_ij likewise over each 2^3 block
    p = params(16)
    uu, prim = seeded_state(p, seed=3, kmax=8, du=0.3, db=0.3)
    uu[..., 0] = 1.0
    prim[..., 0:3] = uu[..., 1:4]
    L = (p.Lx, p.Ly, p.Lz)
    _, _, T = closure.filtered_sample(uu, prim, 2, p.adiabatic_index, 0.0, L)

    def blocks(f):
        return f.reshape(2, 8, 2, 8, 2, 8, order="F")

    for c, (i, j) in enumerate(closure.PAIRS):
        for off, base in ((0, 1), (6, 4)):
            a, b = blocks(uu[..., base + i]), blocks(uu[..., base + j])
            cov = (a * b).mean(axis=(0, 2, 4)) - a.mean(axis=(0, 2, 4)) * b.mean(axis=(0, 2, 4))
            np.testing.assert_allclose(T[:, off + c], cov.reshape(-1, order="F"), atol=1e-13)


def test_numpy_features_of_a_single_mode():
    # central differences of sin(k x) are sin(k dx)/dx * cos(k x)
    p = params((16, 8, 8))
    n = (16, 8, 8)
    x = np.arange(16) * p.Lx / 16
    uu = np.zeros(n + (8,), order="F")
    prim = np.zeros(n + (4,), order="F")
    uu[..., 0] = 1.0
    uu[..., 4] = np.sin(2 * x)[:, None, None]
    X = ref.features(uu, prim, (p.Lx, p.Ly, p.Lz))
    dx = p.Lx / 16
    expect = (math.sin(2 * dx) / dx * np.cos(2 * x))[:, None, None] * np.ones(n)
    np.testing.assert_allclose(X[:, 7 + 3 * 4], expect.reshape(-1, order="F"), atol=1e-12)
    assert np.abs(X[:, 8 + 3 * 4]).max() == 0.0


def test_energy_flux_target_closed_form():
    # rho = 1, B = 0, p = p0 uniform, u = (U(x,y,z), 0, 0): e = p0/(g-1) + U^2/2, the energy flux
    # (mhdrhs.f90:87-89) is (e + p) U; the coarse pressure carries the filtered kinetic energy
    # of the subgrid motion, p_c = p0 + (g-1)/2 (bar(U^2) - bar(U)^2), so the subgrid flux is
    # bar(U^3)/2 - bar(U^2) bar(U)/2 - (g-1)/2 (bar(U^2) - bar(U)^2) bar(U), and zero across x
    p = params(16)
    g, p0 = p.adiabatic_index, 0.7
    rng = np.random.default_rng(5)
    U = rng.standard_normal((16, 16, 16)) * 0.3
    uu = np.zeros((16, 16, 16, 8), order="F")
    uu[..., 0] = 1.0
    uu[..., 1] = U
    uu[..., 7] = p0 / (g - 1) + 0.5 * U * U
    prim = np.zeros((16, 16, 16, 4), order="F")
    prim[..., 0] = U
    prim[..., 3] = p0
    _, _, T = closure.filtered_sample(uu, prim, 2, g, 0.0, (p.Lx, p.Ly, p.Lz))

    def bar(f):
        return f.reshape(2, 8, 2, 8, 2, 8, order="F").mean(axis=(0, 2, 4)).reshape(-1, order="F")

    ub, u2, u3 = bar(U), bar(U * U), bar(U ** 3)
    want = 0.5 * u3 - 0.5 * u2 * ub - 0.5 * (g - 1) * (u2 - ub * ub) * ub
    np.testing.assert_allclose(T[:, 15], want, atol=1e-14)
    assert np.abs(T[:, 16:18]).max() < 1e-15


def test_network_sample_shapes_are_checked():
    from laps_amd import runtime
    X = np.zeros((10, 28), np.float32)
    with pytest.raises(runtime.LapsError, match="targets"):
        runtime.mlp_grad(X, np.zeros((10, 15), np.float32), np.zeros(runtime.MLP_NPARAM, np.float32))
    with pytest.raises(runtime.LapsError, match="inputs"):
        runtime.mlp_forward(np.zeros((10, 27), np.float32), np.zeros(runtime.MLP_NPARAM, np.float32))


def test_spectral_cut_closed_form():
    """laps_amd.closure.spectral_cut on a sum of plane waves: the waves inside the cut
    (kx < ncx/2, |ky| < ncy/2, |kz| < ncz/2) are sampled on the coarse grid unchanged, the rest
    (here kx = 7 >= 4 and a coarse Nyquist row ky = 2) vanish; back on the fine grid the cut is
    the same sum of the kept waves."""
    from laps_amd import closure
    nf, nc = (16, 8, 32), (8, 4, 8)

    def grid(n):
        return [np.arange(n[a]).reshape([-1 if b == a else 1 for b in range(3)]) * 2 * np.pi / n[a]
                for a in range(3)]

    def waves(x, y, z, kept_only):
        f = 1.0 + np.cos(x) + np.sin(2 * x + y) + 0.5 * np.cos(3 * z - x) + 0 * x * y * z
        if not kept_only:
            f = f + np.cos(7 * x) + 0.25 * np.sin(2 * y + z)
        return f

    f = waves(*grid(nf), kept_only=False)
    assert np.abs(closure.spectral_cut(f, nc) - waves(*grid(nc), kept_only=True)).max() < 1e-13
    assert np.abs(closure.spectral_cut(f, nc, fine_grid=True) - waves(*grid(nf), kept_only=True)).max() < 1e-13
