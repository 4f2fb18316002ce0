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
