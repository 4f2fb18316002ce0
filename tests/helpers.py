"""Shared test helpers: parameter sets and error norms."""
import math

import numpy as np

from laps_amd.ic import cons_from_prim, seeded_primitive
from laps_amd.params import default_params

TWO_PI = 2.0 * math.pi


def params(n=32, **kw):
    if isinstance(n, int):
        n = (n, n, n)
    base = dict(nx=n[0], ny=n[1], nz=n[2], Lx=TWO_PI, Ly=TWO_PI, Lz=TWO_PI, adiabatic_index=5.0 / 3.0,
                cfl=0.4, dealias_option=1, if_visc=1, viscosity=1e-3, if_resis=1, resistivity=1e-3)
    base.update(kw)
    return default_params(**base)


def seeded_state(p, seed=0, kmax=None, du=0.1, db=0.1):
    n = (p.nx, p.ny, p.nz)
    if kmax is None:
        kmax = max(2, min(n) // 8)
    prim = seeded_primitive(n, seed, kmin=1, kmax=kmax, du=du, db=db)
    return cons_from_prim(prim, p.adiabatic_index)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def wave_numbers(n, L):
    """signed wave numbers 2 pi m / L with the +Nyquist convention (mhdinit.f90:83-111)"""
    m = np.arange(n)
    m = np.where(m <= n // 2, m, m - n)
    return TWO_PI * m / L


def budgets_numpy(uu, prim, uhat, p):
    """Brute-force volume averages of the budget quantities (laps_budgets):
    <e>, <(rho u).u>/2, <B.B>/2, <p>/(gamma-1), <u.B>, and over the spectrum
    without k = 0, <A.B> (A_hat = i k x B_hat / |k|^2) and <J.B> (J_hat = i k x B_hat).
    uu (nx,ny,nz,8), prim (nx,ny,nz,4) real; uhat (nx/2+1,ny,nz,8) = FFT/N."""
    m, b, u = uu[..., 1:4], uu[..., 4:7], prim[..., 0:3]
    out = [uu[..., 7].mean(), 0.5 * (m * u).sum(-1).mean(), 0.5 * (b * b).sum(-1).mean(),
           prim[..., 3].mean() / (p.adiabatic_index - 1.0), (u * b).sum(-1).mean()]
    nx = uu.shape[0]
    kx = wave_numbers(p.nx, p.Lx)[: nx // 2 + 1]
    ky = wave_numbers(p.ny, p.Ly)
    kz = wave_numbers(p.nz, p.Lz)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    B = uhat[..., 4:7]
    J = np.stack([1j * (KY * B[..., 2] - KZ * B[..., 1]), 1j * (KZ * B[..., 0] - KX * B[..., 2]),
                  1j * (KX * B[..., 1] - KY * B[..., 0])], axis=-1)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    w = np.full(KX.shape, 2.0)
    w[0] = 1.0
    if nx % 2 == 0:
        w[nx // 2] = 1.0
    jb = np.real(np.conj(J) * B).sum(-1) * w
    nz = k2 > 0
    out += [(jb[nz] / k2[nz]).sum(), jb[nz].sum()]
    return np.array(out)


def rk3_amplification(z):
    """Any 3-stage third-order Runge-Kutta scheme, the low-storage one of rktmod.f90:17-61
    included, multiplies a linear mode dU/dt = z U / dt by 1 + z + z^2/2 + z^3/6 per step."""
    return 1 + z + z * z / 2 + z ** 3 / 6


def circular_wave(p, m, amp, di, b0=1.0):
    """Circularly polarised Alfven / whistler wave of integer wave vector m along a uniform
    field B0 = b0 m/|m|, rho = p = 1, no viscosity or resistivity.

    b = amp (e1 cos th + e2 sin th), th = k.x, (e1, e2, k/|k|) right-handed; u = c b.
    With everything a function of k.x only, |B| and |u| are uniform, so the pressure, the
    energy and the density stay uniform; u x b, J x b and div of the 2k stresses vanish, and
    the Hall MHD equations reduce to a linear mode: d b/dt = b0 d_s u + d_i b0 |k| d_s b,
    d u/dt = b0 d_s b, i.e. propagation along +k at v = b0 (sqrt(1 + h^2) - h), h = d_i|k|/2,
    with c = -b0 / v (v = b0, c = -1 without Hall). Every RK stage stays in this family, so a
    step multiplies the mode's coefficients by rk3_amplification(-i |k| v dt) exactly.

    Returns primitive fields (nx,ny,nz,8) and omega = |k| v (box lengths 2 pi)."""
    n = (p.nx, p.ny, p.nz)
    m = np.asarray(m, dtype=np.float64)
    km = np.linalg.norm(m)
    kh = m / km
    e1 = np.cross(kh, [0.0, 0.0, 1.0]) if abs(kh[2]) < 0.9 else np.cross(kh, [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(kh, e1)
    h = 0.5 * di * km
    v = b0 * (math.sqrt(1.0 + h * h) - h)
    c = -b0 / v
    x = [np.arange(q) * TWO_PI / q for q in n]
    th = m[0] * x[0][:, None, None] + m[1] * x[1][None, :, None] + m[2] * x[2][None, None, :]
    ct, st = np.cos(th), np.sin(th)
    prim = np.zeros(n + (8,), order="F")
    prim[..., 0] = 1.0
    prim[..., 7] = 1.0
    for i in range(3):
        b = amp * (e1[i] * ct + e2[i] * st)
        prim[..., 4 + i] = b0 * kh[i] + b
        prim[..., 1 + i] = c * b
    return prim, km * v


def wave_index(p, m):
    """index of the stored half-spectrum coefficient of wave vector m (mx >= 0)"""
    return (m[0], m[1] % p.ny, m[2] % p.nz)


def kept_modes(p):
    """boolean (nx/2+1, ny, nz) mask of the modes dealiasing keeps: the 2/3 sphere
    sqrt(sum (m_i/n_i)^2) <= 1/3 (dealiasing.f90:8, :87-99) or the open phase-shift ellipsoid
    sum (m_i/n_i)^2 < 2/9, evaluated in integers"""
    X, Y, Z = p.nx, p.ny, p.nz
    mx = np.arange(X // 2 + 1, dtype=np.int64)[:, None, None]
    my = np.fft.fftfreq(Y, 1.0 / Y).astype(np.int64)[None, :, None]
    mz = np.fft.fftfreq(Z, 1.0 / Z).astype(np.int64)[None, None, :]
    if p.dealias_option == 1:
        r = np.sqrt((mx / X) ** 2 + (my / Y) ** 2 + (mz / Z) ** 2)
        return r <= 1.0 / 3.0
    if p.dealias_option == 3:
        lhs = 9 * (mx * mx * Y * Y * Z * Z + my * my * X * X * Z * Z + mz * mz * X * X * Y * Y)
        return lhs < 2 * X * X * Y * Y * Z * Z
    return np.ones((X // 2 + 1, Y, Z), dtype=bool)
