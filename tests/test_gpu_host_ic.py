"""The Fortran host's random-phase perturbation hooks ipert = 2..7 (mhdinit.f90:344-829).

Their phases come from the compiler's random_number, as in the reference, so no fixture can
hold their values; every check below is independent of the phases. Each hook adds plane waves
a cos(k.x + ph): the initial output out000.dat (primitive fields, written before the first
step) must hold, for every wave vector the hook uses, the amplitude the reference's formula
gives (|f_hat(k)| = a/2 in the normalised transform), nothing at any other wave vector, and
the structure the formula builds in: b perpendicular to k (div B = 0 mode by mode) and to B0,
u = -b for outward Alfven waves. The run itself takes one step.
"""
import os
import subprocess

import numpy as np
import pytest

from laps_amd.io import read_out

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "host", "_build", "mhd_laps")
N = 32
L = 2 * np.pi

INPUT = """&genr
   tmax = 100.0
   dtout = 100.0
   dtrms = 100.0
/
&numerical
   cfl = 0.4
   dealias_option = 1
/
&prl
   ndim_parallel = 2
/
&grid
   nx = {n}
   ny = {n}
   nz = {n}
   Lx = {L!r}
   Ly = {L!r}
   Lz = {L!r}
/
&field
   ifield = 3
   Bx0 = 1.
   By0 = 0.
   Bz0 = 0.
   press0 = 1.0
/
&pert
{pert}
/
&phys
   adiabatic_index = 1.6666666666666667
   if_resis = T
   resistivity = 1e-3
   if_visc = T
   viscosity = 1e-3
/
&AEB
   if_AEB = F
/
&Hall
   if_Hall = F
/
&laps
   nstep = 1
/
"""


def run_ic(tmp_path, pert):
    if not os.path.exists(HOST):
        pytest.fail("host/_build/mhd_laps is not built (run __graft_entry__.build())")
    (tmp_path / "mhd.input").write_text(INPUT.format(n=N, L=L, pert=pert))
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    t, prim = read_out(tmp_path / "out000.dat", (N, N, N))
    assert t == 0.0
    return prim


def spec(f):
    """normalised 3D transform, indexed [mx, my, mz] with signed modes wrapped (numpy order)"""
    return np.fft.fftn(f) / f.size


def support(mask_modes):
    m = np.zeros((N, N, N), dtype=bool)
    for mx, my, mz in mask_modes:
        m[mx % N, my % N, mz % N] = True
        m[-mx % N, -my % N, -mz % N] = True
    return m


@pytest.mark.parametrize("ipert,field,axis", [(2, 6, 0), (3, 6, 1), (4, 4, 2)])
def test_one_dimensional_random_modes(tmp_path, ipert, field, axis):
    db0, nmode = 0.1, 3
    prim = run_ic(tmp_path, f"   ipert = {ipert}\n   db0 = {db0}\n   nmode = {nmode}")
    bg = 1.0 if field == 4 else 0.0
    f = spec(prim[..., field] - bg)
    modes = []
    for k in range(1, nmode + 1):
        m = [0, 0, 0]
        m[axis] = k
        modes.append(tuple(m))
        assert abs(abs(f[tuple(np.array(m) % N)]) - db0 / 2) < 1e-13
    assert np.abs(f[~support(modes)]).max() < 1e-13
    # every other field is the uniform background
    for v, val in enumerate((1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0)):
        if v != field:
            assert np.abs(prim[..., v] - val).max() < 1e-15


def test_tearing_modes(tmp_path):
    dbx, dbz, nmx, nmz = 0.01, 0.02, 2, 1
    prim = run_ic(tmp_path, f"   ipert = 5\n   dbx = {dbx}\n   dbz = {dbz}\n   nmodex = {nmx}\n"
                            f"   nmodez = {nmz}")
    a0 = 0.05   # the &field default, sheet width
    by = prim[..., 5]
    for iy, yc in ((N // 4, L / 4), (3 * N // 4, 3 * L / 4)):
        # at the sheet centre m = a0, m' = 0: by = sum (kx dbx + kz dbz) a0 sin(kx x + kz z + ph)
        fxz = np.fft.fft2(by[:, iy, :]) / (N * N)
        for ikx in range(1, nmx + 1):
            for ikz in range(0, nmz + 1):
                want = abs(ikx * dbx + ikz * dbz) * a0 / 2
                assert abs(abs(fxz[ikx, ikz]) - want) < 1e-13, (iy, ikx, ikz)
        keep = np.zeros((N, N), dtype=bool)
        for ikx in range(1, nmx + 1):
            for ikz in range(0, nmz + 1):
                keep[ikx, ikz] = keep[-ikx % N, -ikz % N] = True
        assert np.abs(fxz[~keep]).max() < 1e-13
        assert np.abs(prim[:, iy, :, 4] - 1.0).max() < 1e-15   # m' = 0 at the centre: bx, bz untouched
        assert np.abs(prim[:, iy, :, 6]).max() < 1e-15


def alfvenic_amplitude(ikx, iky, ikz, db0):
    a = db0 / np.sqrt(ikx)
    a = a / np.sqrt(abs(iky)) if iky else a
    return a / np.sqrt(abs(ikz)) if ikz else a


def test_outward_alfvenic_modes(tmp_path):
    db0, nm = 0.05, 2
    prim = run_ic(tmp_path, f"   ipert = 6\n   db0 = {db0}\n   nmodex = {nm}\n   nmodey = {nm}\n"
                            f"   nmodez = {nm}\n   in_out_ratio = 0.")
    # outward waves along B0 = +x: u = -b exactly
    assert np.abs(prim[..., 2] + prim[..., 5]).max() < 1e-15
    assert np.abs(prim[..., 3] + prim[..., 6]).max() < 1e-15
    by, bz = spec(prim[..., 5]), spec(prim[..., 6])
    modes = []
    for ikx in range(1, nm + 1):
        for iky in range(-nm, nm + 1):
            for ikz in range(-nm, nm + 1):
                modes.append((ikx, iky, ikz))
                a = alfvenic_amplitude(ikx, iky, ikz, db0)
                e = abs(by[ikx, iky % N, ikz % N]) ** 2 + abs(bz[ikx, iky % N, ikz % N]) ** 2
                want = a * a / 2 if iky == 0 and ikz == 0 else a * a / 4   # circular or linear
                assert abs(e - want) < 1e-14, (ikx, iky, ikz)
                # b perpendicular to k: div B = 0 mode by mode
                assert abs(iky * by[ikx, iky % N, ikz % N] + ikz * bz[ikx, iky % N, ikz % N]) < 1e-14
    keep = support(modes)
    assert max(np.abs(by[~keep]).max(), np.abs(bz[~keep]).max()) < 1e-13
    assert np.abs(prim[..., 4] - 1.0).max() < 1e-15


def test_in_and_outward_alfvenic_modes(tmp_path):
    db0, nm, ratio = 0.05, 2, 0.5
    prim = run_ic(tmp_path, f"   ipert = 6\n   db0 = {db0}\n   nmodex = {nm}\n   nmodey = {nm}\n"
                            f"   nmodez = {nm}\n   in_out_ratio = {ratio}")
    by, bz = spec(prim[..., 5]), spec(prim[..., 6])
    my = np.fft.fftfreq(N, 1.0 / N)[None, :, None]
    mz = np.fft.fftfreq(N, 1.0 / N)[None, None, :]
    assert np.abs(my * by + mz * bz).max() < 1e-13   # div B = 0 mode by mode
    # outward waves carry u = -b, inward ones u = +b: (u - b)/2 is minus the outward part,
    # (u + b)/2 the inward part; an outward wave of k = (kx > 0, ky, kz) has its coefficient at
    # k, an inward one (-kx, ky, kz) at that vector
    out_y, in_y = spec(prim[..., 2] - prim[..., 5]) / -2, spec(prim[..., 2] + prim[..., 5]) / 2
    out_z, in_z = spec(prim[..., 3] - prim[..., 6]) / -2, spec(prim[..., 3] + prim[..., 6]) / 2
    for ikx in range(1, nm + 1):
        for iky in range(-nm, nm + 1):
            for ikz in range(-nm, nm + 1):
                a = alfvenic_amplitude(ikx, iky, ikz, db0)
                f = 2 if iky == 0 and ikz == 0 else 4
                e_out = abs(out_y[ikx, iky % N, ikz % N]) ** 2 + abs(out_z[ikx, iky % N, ikz % N]) ** 2
                e_in = abs(in_y[-ikx % N, iky % N, ikz % N]) ** 2 + abs(in_z[-ikx % N, iky % N, ikz % N]) ** 2
                assert abs(e_out - a * a / f) < 1e-14, (ikx, iky, ikz)
                assert abs(e_in - (ratio * a) ** 2 / f) < 1e-14, (ikx, iky, ikz)


def test_isotropic_modes(tmp_path):
    db0, dv0, drho0, nm = 0.1, 0.08, 0.01, 3
    prim = run_ic(tmp_path, f"   ipert = 7\n   db0 = {db0}\n   dv0 = {dv0}\n   drho0 = {drho0}\n"
                            f"   nmodex = {nm}\n   nmodey = {nm}\n   nmodez = {nm}")
    b = [spec(prim[..., 4] - 1.0), spec(prim[..., 5]), spec(prim[..., 6])]
    u = [spec(prim[..., v]) for v in (1, 2, 3)]
    rho = spec(prim[..., 0] - 1.0)
    m = np.fft.fftfreq(N, 1.0 / N)
    MX, MY, MZ = np.meshgrid(m, m, m, indexing="ij")
    kr = np.sqrt(MX ** 2 + MY ** 2 + MZ ** 2)
    keep = (kr > 0) & (kr <= nm)
    a = np.where(keep, 1.0 / np.maximum(kr, 1e-30) ** 1.5, 0.0)
    bmag = np.sqrt(sum(abs(c) ** 2 for c in b))
    # b along k x B0 / (|k| |B0|), of length sin(k, B0), times db0 |k|^-3/2, in every mode of the
    # sphere, nothing outside
    sin_kb = np.sqrt(MY ** 2 + MZ ** 2) / np.maximum(kr, 1e-30)
    assert np.abs(bmag - db0 * a * sin_kb / 2).max() < 1e-13
    assert np.abs(np.abs(rho) - drho0 * a / 2).max() < 1e-13
    assert np.abs(b[0]).max() < 1e-14                      # b perpendicular to B0 = x
    assert np.abs(MX * b[0] + MY * b[1] + MZ * b[2]).max() < 1e-13   # and to k
    assert np.abs(MX * u[0] + MY * u[1] + MZ * u[2]).max() < 1e-13   # u too
    assert np.abs(u[0]).max() < 1e-14
    assert max(np.abs(c[~keep]).max() for c in u) < 1e-13
