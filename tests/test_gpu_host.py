"""The Fortran host (host/mhd_host.f90) drives the GPU step through the C ABI.

A seeded 32^3 Hall-MHD state is written as out000.dat, the host restarts from
it (if_restart, the reference's own path for external initial conditions),
runs 5 steps and writes out001.dat and rms.dat. The files are compared with
the CPU oracle advanced from the same file contents.
"""
import os
import subprocess

import numpy as np
import pytest

from laps_amd.ic import seeded_primitive
from laps_amd.io import read_out, read_rms, write_out
from tests.helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "host", "_build", "mhd_laps")

MHD_INPUT = """&genr
   tmax = 100.0
   dtout = 100.0
   dtrms = 100.0
   if_restart = T
   n_start = 0
/
&numerical
   cfl = 0.4
   dealias_option = 1
/
&prl
   ndim_parallel = 2
/
&grid
   nx = 32
   ny = 32
   nz = 32
   Lx = 6.283185307179586
   Ly = 6.283185307179586
   Lz = 6.283185307179586
/
&field
   ifield = 3
   Bx0 = 1.
   press0 = 1.0
/
&pert
   ipert = 0
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
   if_Hall = T
   ion_inertial_length = 0.05
/
&laps
   nstep = 5
/
"""


@pytest.mark.gpu
def test_fortran_host_restart_run(tmp_path, oracle_mod):
    if not os.path.exists(HOST):
        pytest.fail("host/_build/mhd_laps is not built (run __graft_entry__.build())")
    n = 32
    prim = seeded_primitive((n, n, n), seed=11, kmin=1, kmax=4, du=0.1, db=0.1)
    write_out(tmp_path / "out000.dat", 0.0, prim)
    (tmp_path / "mhd.input").write_text(MHD_INPUT)
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert "nstep is reached" in r.stdout
    t1, out = read_out(tmp_path / "out001.dat", (n, n, n))

    from tests.helpers import params
    p = params(n, if_hall=1, ion_inertial_length=0.05, cfl=0.4)
    _, prim0 = read_out(tmp_path / "out000.dat", (n, n, n))
    uu, pr = oracle_mod.cons_from_prim(p, prim0)
    o = oracle_mod.Oracle(p)
    o.set_state(uu, pr)
    o.step(5)
    uo, po = o.state()
    ref = np.array(uo)
    ref[..., 1:4] = po[..., 0:3]
    ref[..., 7] = po[..., 3]
    assert abs(t1 - o.scalars()["time"]) < 1e-6 * max(1.0, t1)  # float32 time record
    assert rel_l2(out, ref) < 1e-11
    rows = read_rms(tmp_path / "rms.dat")
    assert rows.shape == (2, 20)
    ave, var, ru2 = o.moments()
    np.testing.assert_allclose(rows[-1, 1:9], ave, rtol=1e-7, atol=1e-12)   # 1pe16.8 text
    assert (tmp_path / "grid.dat").stat().st_size == 4 * (2 + 3) + 4 * (2 + 3 * n)
    assert (tmp_path / "EBM_info.dat").exists()


# The reference's own sample input (src_compressible/mhd.input: ifield = 3 with no field,
# ipert = 8 switchback, 24^3 box, 2/3 rule, cfl 0.5, nu = eta = 1e-4) at 64^3 instead of 512^3.
SWITCHBACK_INPUT = """&genr
   tmax = 20.0
   dtout = 2.0
   dtrms = 0.05
/
&numerical
   cfl = 0.5
   dealias_option = 1
/
&prl
   ndim_parallel = 2
/
&grid
   nx = 64
   ny = 64
   nz = 64
   Lx = 24.0
   Ly = 24.0
   Lz = 24.0
/
&field
   ifield = 3
   Bx0 = 0.
   By0 = 0.
   Bz0 = 0.
   press0 = 1.0
/
&pert
   ipert = 8
   B0_SB = 1.
   dB_SB = -2.
   R0_SB = 2.
   R1_SB = 0.5
   H_SB = 3.
   Rm_SB = 5.
   Rd_SB = 1.5
/
&phys
   adiabatic_index = 1.666667
   if_resis = T
   resistivity = 1e-4
   if_visc = T
   viscosity = 1e-4
/
&AEB
   if_AEB = F
   radius0 = 30.0
   Ur0 = 1.167
/
&Hall
   if_Hall = F
   ion_inertial_length = 0
/
&laps
   nstep = 3
/
"""


def switchback_numpy(n, L, B0, dB, R0, R1, H, Rm, Rd):
    """mhdinit.f90:931-1020 evaluated with numpy: primitive (rho, u, B, p) on an n^3 grid of
    [0, L)^3 over a background of rho = p = 1 and no field."""
    from scipy.special import erf
    g = np.arange(n) * (L / n) - L / 2
    xs, ys, zs = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(ys ** 2 + zs ** 2)
    gring = np.exp(-((r - R0) / R1) ** 2)
    gx = np.exp(-(xs / H) ** 2)
    bx_core = B0 + dB * gring
    bx = B0 + dB * np.exp(-((r - R0) / R1) ** 2 - (xs / H) ** 2)
    c = -(np.exp(-(R0 / R1) ** 2) + np.sqrt(np.pi) * R0 / R1 * erf(R0 / R1))
    rs = np.where(r < 1e-8, 1.0, r)
    sflux = gring + np.sqrt(np.pi) * R0 / R1 * erf((R0 - r) / R1) + c
    br = -dB * (xs * R1 ** 2) / (rs * H ** 2) * gx * sflux
    taper = np.where(r < Rm, 1.0, np.exp(-((r - Rm) / Rd) ** 2))
    dtaper = np.where(r < Rm, 0.0, -2 * (r - Rm) / Rd ** 2 * taper)
    q = np.where(r < Rm, 1.0, bx_core / bx * (1 + 1 / bx_core * (taper * (bx - bx_core) - (
        dtaper * dB * R1 ** 2 / rs * sflux * 0.5 * (gx - 1)))))
    core = r < 1e-8
    br = np.where(core, 0.0, br * taper)
    bx = np.where(core, bx, bx * q)
    fill = bx ** 2 + br ** 2 < B0 ** 2
    bphi = np.where(core | ~fill, 0.0, np.sqrt(np.maximum(B0 ** 2 - br ** 2 - bx ** 2, 0.0)))
    p = np.where(core | fill, 1.0, 1.0 - 0.5 * (bx ** 2 + br ** 2 - B0 ** 2))
    phi = np.where(core, 0.0, np.arctan2(zs, ys))
    prim = np.zeros((n, n, n, 8))
    prim[..., 0] = 1.0
    prim[..., 4] = bx
    prim[..., 5] = br * np.cos(phi) - bphi * np.sin(phi)
    prim[..., 6] = br * np.sin(phi) + bphi * np.cos(phi)
    prim[..., 1:4] = -prim[..., 4:7]
    prim[..., 7] = p
    return prim


@pytest.mark.gpu
def test_fortran_host_switchback_sample_input(tmp_path, oracle_mod):
    if not os.path.exists(HOST):
        pytest.fail("host/_build/mhd_laps is not built (run __graft_entry__.build())")
    n = 64
    (tmp_path / "mhd.input").write_text(SWITCHBACK_INPUT)
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert "nstep is reached" in r.stdout
    t0, prim0 = read_out(tmp_path / "out000.dat", (n, n, n))
    assert t0 == 0.0
    want = switchback_numpy(n, 24.0, 1.0, -2.0, 2.0, 0.5, 3.0, 5.0, 1.5)
    np.testing.assert_allclose(prim0, want, rtol=0, atol=1e-12)
    # the construction keeps the total pressure uniform off the axis r = 0 (where the reference
    # applies no correction): |B| = B0 where the azimuthal field fills it up, and the pressure
    # lowered by the excess elsewhere
    ptot = prim0[..., 7] + 0.5 * (prim0[..., 4:7] ** 2).sum(-1)
    g = np.arange(n) * (24.0 / n) - 12.0
    off_axis = (g[None, :, None] ** 2 + g[None, None, :] ** 2 >= 1e-16) & np.ones((n, 1, 1), bool)
    assert np.abs(ptot - 1.5)[off_axis].max() < 1e-12
    assert np.abs(ptot - 1.5).max() < 1e-6
    assert prim0[..., 4].min() < -0.9          # the reversed field inside the ring
    assert np.array_equal(prim0[..., 1:4], -prim0[..., 4:7])
    # three steps on the GPU against the oracle from the same initial state
    t1, out = read_out(tmp_path / "out001.dat", (n, n, n))
    from tests.helpers import params
    p = params(n, Lx=24.0, Ly=24.0, Lz=24.0, cfl=0.5, adiabatic_index=1.666667, viscosity=1e-4,
               resistivity=1e-4)
    uu, pr = oracle_mod.cons_from_prim(p, prim0)
    o = oracle_mod.Oracle(p)
    o.set_state(uu, pr)
    o.step(3)
    uo, po = o.state()
    ref = np.array(uo)
    ref[..., 1:4] = po[..., 0:3]
    ref[..., 7] = po[..., 3]
    assert abs(t1 - o.scalars()["time"]) < 1e-6 * max(1.0, t1)
    assert rel_l2(out, ref) < 1e-11


RESTART_AEB_INPUT = MHD_INPUT.replace("""&AEB
   if_AEB = F
/""", """&AEB
   if_AEB = T
   radius0 = 1.0
   Ur0 = 0.5
   if_corotating = T
   corotating_angle = 0.3
/""").replace("nstep = 5", "nstep = 3")


@pytest.mark.gpu
def test_fortran_host_restart_at_nonzero_time(tmp_path, oracle_mod):
    # restart.f90:17-63: the time comes from the file's float32 record; with the expanding box
    # the radius, hence every wave number, starts from it (evolve_radius(t_restart))
    if not os.path.exists(HOST):
        pytest.fail("host/_build/mhd_laps is not built (run __graft_entry__.build())")
    n, t_start = 32, 0.75
    prim = seeded_primitive((n, n, n), seed=12, kmin=1, kmax=4, du=0.1, db=0.1)
    write_out(tmp_path / "out000.dat", t_start, prim)
    (tmp_path / "mhd.input").write_text(RESTART_AEB_INPUT)
    r = subprocess.run([HOST], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert "Restart from time" in r.stdout and "nstep is reached" in r.stdout
    t1, out = read_out(tmp_path / "out001.dat", (n, n, n))
    from tests.helpers import params
    p = params(n, if_hall=1, ion_inertial_length=0.05, cfl=0.4, if_aeb=1, radius0=1.0, Ur0=0.5,
               if_corotating=1, corotating_angle=0.3)
    _, prim0 = read_out(tmp_path / "out000.dat", (n, n, n))
    uu, pr = oracle_mod.cons_from_prim(p, prim0)
    o = oracle_mod.Oracle(p)
    o.set_state(uu, pr)
    o.set_time(t_start, 0.0)
    o.step(3)
    sc = o.scalars()
    assert sc["time"] > t_start and sc["radius"] == pytest.approx(1.0 + 0.5 * sc["time"], rel=1e-14)
    assert abs(t1 - sc["time"]) < 1e-6 * t1
    uo, po = o.state()
    ref = np.array(uo)
    ref[..., 1:4] = po[..., 0:3]
    ref[..., 7] = po[..., 3]
    assert rel_l2(out, ref) < 1e-11
    ebm = np.loadtxt(tmp_path / "EBM_info.dat", ndmin=2)
    assert ebm[0, 0] == pytest.approx(t_start) and ebm[0, 1] == pytest.approx(1.0 + 0.5 * t_start, rel=1e-12)
