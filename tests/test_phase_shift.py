"""Phase-shift dealiasing (dealias_option = 3; no reference counterpart).

Pinned by brute force: for rho = 1 every flux of the mass, momentum and
induction equations is a quadratic polynomial in (rho u, B, e) and the Hall
current, so their exact (alias-free) product spectrum is the product on a grid
twice as fine, truncated. The oracle's phase-shifted RHS must equal it inside
the kept ellipsoid r <= sqrt(2)/3, while the unshifted evaluation of the same
state does not.
"""

import numpy as np
import pytest

from .helpers import params, wave_numbers

def signed(n):
    m = np.arange(n)
    return np.where(m <= n // 2, m, m - n)


def ellipsoid(shape):
    # (mx/nx)^2 + (my/ny)^2 + (mz/nz)^2 < 2/9, exactly: on the boundary a mode can alias in two
    # directions at once when an n is a multiple of 3 (m = n/3 in two axes)
    nx, ny, nz = shape
    mx, my, mz = np.meshgrid(signed(nx), signed(ny), signed(nz), indexing="ij")
    X, Y, Z = nx * nx, ny * ny, nz * nz
    return 9 * (mx * mx * Y * Z + my * my * X * Z + mz * mz * X * Y) < 2 * X * Y * Z


def band_limited(shape, rng, scale):
    f = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * ellipsoid(shape)
    return np.real(np.fft.ifftn(f)) * scale / np.sqrt(np.mean(np.real(np.fft.ifftn(f)) ** 2))


def state(p, seed):
    shape = (p.nx, p.ny, p.nz)
    rng = np.random.default_rng(seed)
    uu = np.zeros(shape + (8,))
    uu[..., 0] = 1.0
    for f in (1, 2, 3):
        uu[..., f] = band_limited(shape, rng, 0.3)
    for f in (4, 5, 6):
        uu[..., f] = band_limited(shape, rng, 0.4) + (1.0 if f == 4 else 0.0)
    uu[..., 7] = 3.0 + band_limited(shape, rng, 0.2)
    g = p.adiabatic_index
    pr = np.zeros(shape + (4,))
    pr[..., 0:3] = uu[..., 1:4]
    pr[..., 3] = (uu[..., 7] - 0.5 * ((uu[..., 1:4] ** 2).sum(-1) + (uu[..., 4:7] ** 2).sum(-1))) * (g - 1)
    return uu, pr


def exact_rhs(p, uu):
    """RHS components 1-7 (calc_rhs, mhdrhs.f90:206-233) of the exact products"""
    nx, ny, nz = p.nx, p.ny, p.nz
    N = nx * ny * nz
    big = (2 * nx, 2 * ny, 2 * nz)
    ix = [signed(n) % (2 * n) for n in (nx, ny, nz)]

    def fine(f):
        F = np.fft.fftn(f) / N
        G = np.zeros(big, complex)
        G[np.ix_(*ix)] = F
        return np.real(np.fft.ifftn(G)) * (8 * N)

    wx, wy, wz = (wave_numbers(n, L) for n, L in ((nx, p.Lx), (ny, p.Ly), (nz, p.Lz)))
    KX, KY, KZ = np.meshgrid(wave_numbers(2 * nx, 2 * p.Lx) * 2, wave_numbers(2 * ny, 2 * p.Ly) * 2,
                             wave_numbers(2 * nz, 2 * p.Lz) * 2, indexing="ij")
    m = [fine(uu[..., f]) for f in (1, 2, 3)]
    B = [fine(uu[..., f]) for f in (4, 5, 6)]
    e = fine(uu[..., 7])
    Bh = [np.fft.fftn(b) for b in B]
    J = [np.real(np.fft.ifftn(1j * (KY * Bh[2] - KZ * Bh[1]))), np.real(np.fft.ifftn(1j * (KZ * Bh[0] - KX * Bh[2]))),
         np.real(np.fft.ifftn(1j * (KX * Bh[1] - KY * Bh[0])))]
    g, di = p.adiabatic_index, p.ion_inertial_length
    u = m  # rho = 1
    b2 = sum(x * x for x in B)
    pres = (e - 0.5 * (sum(a * c for a, c in zip(m, u)) + b2)) * (g - 1)
    ptot = pres + 0.5 * b2
    F = [m[0], m[1], m[2]]
    for i in range(3):
        for j in range(3):
            F.append(m[i] * u[j] - B[i] * B[j] + (ptot if i == j else 0.0))
    # reference slots F4..F12: (m_x u_x, m_y u_x, m_z u_x, m_x u_y, ...) = row j of the loop above
    F[3:12] = [m[i] * u[j] - B[i] * B[j] + (ptot if i == j else 0.0) for j in range(3) for i in range(3)]
    E = [u[2] * B[1] - u[1] * B[2], u[0] * B[2] - u[2] * B[0], u[1] * B[0] - u[0] * B[1]]
    if p.if_hall:
        E = [E[0] + di * (J[1] * B[2] - J[2] * B[1]), E[1] + di * (J[2] * B[0] - J[0] * B[2]),
             E[2] + di * (J[0] * B[1] - J[1] * B[0])]
    F += E

    def coarse(f):
        G = np.fft.fftn(f) / (8 * N)
        return G[np.ix_(*ix)][: nx // 2 + 1]

    Fh = [coarse(f) for f in F]
    kx, ky, kz = np.meshgrid(1j * wx[: nx // 2 + 1], 1j * wy, 1j * wz, indexing="ij")
    R = np.zeros((nx // 2 + 1, ny, nz, 7), complex)
    for v in range(4):
        R[..., v] = -(kx * Fh[3 * v] + ky * Fh[3 * v + 1] + kz * Fh[3 * v + 2])
    R[..., 4] = kz * Fh[13] - ky * Fh[14]
    R[..., 5] = kx * Fh[14] - kz * Fh[12]
    R[..., 6] = ky * Fh[12] - kx * Fh[13]
    return R


@pytest.mark.parametrize("shape,hall", [((16, 16, 16), 0), ((12, 12, 12), 1), ((16, 12, 24), 1)])
def test_phase_shift_rhs_is_alias_free(oracle_mod, shape, hall):
    p3 = params(shape, dealias_option=3, if_hall=hall, ion_inertial_length=0.1)
    uu, pr = state(p3, seed=7)
    ref = exact_rhs(p3, uu)
    keep = ellipsoid(shape)[: shape[0] // 2 + 1]
    scale = np.abs(ref[keep]).max()
    o = oracle_mod.Oracle(p3)
    o.set_state(uu, pr)
    got = o.rhs()[..., :7]
    err = np.abs(got[keep] - ref[keep]).max() / scale
    o.close()
    # the same state without the shift aliases
    p1 = params(shape, dealias_option=1, if_hall=hall, ion_inertial_length=0.1)
    o1 = oracle_mod.Oracle(p1)
    o1.set_state(uu, pr)
    err1 = np.abs(o1.rhs()[..., :7][keep] - ref[keep]).max() / scale
    o1.close()
    print(f"phase-shift err {err:.2e}, unshifted err {err1:.2e}")
    assert err < 1e-12
    assert err1 > 1e-3
