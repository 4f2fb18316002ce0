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
