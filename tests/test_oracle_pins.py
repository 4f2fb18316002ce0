"""Pins the CPU oracle (oracle/laps_oracle.c) before it is trusted.

Each check is chosen so that a plausible restatement mistake (a dropped
term, a wrong sign or index, a transposed operand) fails it: library
routines (numpy.fft), brute-force sums, closed forms of the equations, and
invariants. The reference ships no tests or golden vectors for this path.
"""
import math

import numpy as np
import pytest

from laps_amd.ic import cons_from_prim
from tests.helpers import (circular_wave, params, rel_l2, rk3_amplification, seeded_state, TWO_PI,
                           wave_index)


def brute_dft(x, sign):
    n = x.size
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [2, 4, 8, 12, 16, 24, 48])
def test_fft1d_brute_force(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        got = oracle_mod.fft1d(x, sign)
        assert np.abs(got - brute_dft(x, sign)).max() < 1e-13 * max(1, n)


@pytest.mark.parametrize("n", [16, 32, 64, 96, 128, 256, 512])
def test_fft1d_numpy(oracle_mod, n):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(oracle_mod.fft1d(x, -1), np.fft.fft(x)) < 1e-15
    assert rel_l2(oracle_mod.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-15


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 8), (8, 32, 16), (96, 16, 32)])
def test_fft3d_r2c_c2r_numpy(oracle_mod, shape):
    p = params(shape)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(2)
    f = np.asfortranarray(rng.standard_normal(shape + (2,)))
    s = o.fft_r2c(f)
    n = shape[0] * shape[1] * shape[2]
    for c in range(2):
        ref = np.fft.rfftn(f[..., c].T).T / n      # fftw.f90: each axis divided by its length
        assert rel_l2(s[..., c], ref) < 1e-15
    back = o.fft_c2r(s)                            # unnormalised inverse
    assert rel_l2(back, f) < 1e-14
    # c2r reads only Re of DC / Nyquist coefficients (FFTW c2r convention)
    s2 = s.copy()
    s2[0, 0, 0, :] += 1j * 5.0
    s2[shape[0] // 2, 0, 0, :] += 1j * 7.0
    assert rel_l2(o.fft_c2r(s2), back) < 1e-14


def test_single_mode_transform(oracle_mod):
    # cos(3x + 2y - z) has coefficient 1/2 at (3,2,-1) only (normalised forward)
    p = params(16)
    o = oracle_mod.Oracle(p)
    x = np.arange(16) * TWO_PI / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = o.fft_r2c(np.cos(3 * X + 2 * Y - Z))
    expect = np.zeros_like(s)
    expect[3, 2, 16 - 1] = 0.5
    assert np.abs(s - expect).max() < 1e-15


def test_dealias_sphere_count(oracle_mod):
    # option 1 keeps sqrt((mx/nx)^2+(my/ny)^2+(mz/nz)^2) <= 1/3 (dealiasing.f90:8, :87-99)
    p = params(64)
    o = oracle_mod.Oracle(p)
    u = np.ones((33, 64, 64, 8), dtype=complex, order="F")
    d = o.dealias(u)
    kept = np.count_nonzero(d[..., 0])
    m = np.fft.fftfreq(64, 1 / 64)
    mx = np.arange(33)[:, None, None]
    rr = np.sqrt((mx / 64.0) ** 2 + (m[None, :, None] / 64.0) ** 2 + (m[None, None, :] / 64.0) ** 2)
    assert kept == np.count_nonzero(rr <= 1 / 3) == 21042
    assert np.all(d[..., 0][d[..., 0] != 0] == 1)


def test_compact_filter_closed_form(oracle_mod):
    p = params(32, dealias_option=2)
    o = oracle_mod.Oracle(p)
    fx, fy, fz = o.filters()
    assert fx[0] == 1.0 and fy[0] == 1.0
    # theta = pi: numerator a - b + c = 0 exactly, round-off / (1 - 2a) = 0.01
    assert abs(fx[16]) < 1e-13 and abs(fy[16]) < 1e-13
    a = 0.495
    th = np.arange(17) * TWO_PI / 32
    ref = ((5 + 6 * a) / 8 + (1 + 2 * a) / 2 * np.cos(th) - (1 - 2 * a) / 8 * np.cos(2 * th)) / (1 + 2 * a * np.cos(th))
    assert np.abs(fx - ref).max() < 1e-13
    assert np.all(np.diff(fx) <= 0)


def uniform_prim(n, rho=1.0, u=(0, 0, 0), b=(1, 0, 0), pr=1.0):
    prim = np.zeros((n, n, n, 8), order="F")
    prim[..., 0] = rho
    prim[..., 1:4] = u
    prim[..., 4:7] = b
    prim[..., 7] = pr
    return prim


def test_cfl_closed_form(oracle_mod):
    # rho=1, p=1, B=(1,0,0), u=0: c_fast,x = sqrt(gamma), c_fast,y = c_fast,z = sqrt(gamma+1)
    p = params(16, cfl=0.4)
    o = oracle_mod.Oracle(p)
    uu, pr = oracle_mod.cons_from_prim(p, uniform_prim(16))
    o.set_state(uu, pr)
    raw, dt = o.cfl()
    g = p.adiabatic_index
    dx = TWO_PI / 16
    assert raw == pytest.approx(min(dx / math.sqrt(g), dx / math.sqrt(g + 1)), rel=1e-14)
    assert dt == pytest.approx(0.4 * raw, rel=1e-14)
    # B = 0: c_fast = c_s, c_slow = 0, with |u| added
    uu, pr = oracle_mod.cons_from_prim(p, uniform_prim(16, u=(0.5, 0, 0), b=(0, 0, 0)))
    o.set_state(uu, pr)
    raw, _ = o.cfl()
    assert raw == pytest.approx(dx / (0.5 + math.sqrt(g)), rel=1e-14)


def test_cfl_hall_signed_max(oracle_mod):
    # Hall speed uses the signed max of (Bx,By,Bz) (mhd.f90:396-398): all-negative B gives no Hall limit
    p = params(16, if_hall=1, ion_inertial_length=5.0)
    o = oracle_mod.Oracle(p)
    dx = TWO_PI / 16
    g = p.adiabatic_index
    uu, pr = oracle_mod.cons_from_prim(p, uniform_prim(16, b=(1.0, 0, 0)))
    o.set_state(uu, pr)
    assert o.cfl()[0] == pytest.approx(dx / (5.0 * 1.0 / dx), rel=1e-14)
    uu, pr = oracle_mod.cons_from_prim(p, uniform_prim(16, b=(-1.0, -1e-3, -1e-3)))
    o.set_state(uu, pr)
    assert o.cfl()[0] < dx / math.sqrt(g) * 1.0000001


def rk3_poly(z):
    return 1 + z + z * z / 2 + z ** 3 / 6


@pytest.mark.parametrize("implicit", [False, True])
def test_rk3_viscous_shear_mode(oracle_mod, implicit):
    # u_y = eps sin(x) with uniform rho, p, B = 0 has a linear RHS up to O(eps^2)
    # (viscous heating feeds p, hence u_x, at second order)
    # d(rho u_y)/dt = -nu k^2 rho u_y; explicit: amplification 1+z+z^2/2+z^3/6,
    # implicit: prod 1/(1 + tau_k nu k^2), tau = (8/15, 2/15, 1/3) dt (rktmod.f90:17-61)
    n, nu, dt, eps = 16, 0.5, 0.2, 1e-7
    p = params(n, if_visc=1, viscosity=nu, if_visc_exp=0 if implicit else 1, if_resis=0, dealias_option=1)
    o = oracle_mod.Oracle(p)
    x = np.arange(n) * TWO_PI / n
    prim = uniform_prim(n, b=(0, 0, 0))
    prim[..., 2] = eps * np.sin(x)[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    o.set_fixed_dt(dt)
    u0 = o.spectral()[1, 0, 0, 2]
    o.step(1)
    u1 = o.spectral()[1, 0, 0, 2]
    a = nu * 1.0 * dt
    amp = 1 / ((1 + 8 / 15 * a) * (1 + 2 / 15 * a) * (1 + a / 3)) if implicit else rk3_poly(-a)
    assert abs(u1 / u0 - amp) < 1e-13
    assert abs(rk3_poly(-a) - math.exp(-a)) > 1e-12  # the check distinguishes the scheme


def test_rhs_advection_single_mode(oracle_mod):
    # rho = 1 + eps cos(x), u = (U,0,0): d rho/dt = -U d rho/dx -> fnl1(k=1) = -i U rho_hat
    n, U, eps = 16, 0.3, 1e-2
    p = params(n, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    x = np.arange(n) * TWO_PI / n
    prim = uniform_prim(n, u=(U, 0, 0))
    prim[..., 0] = 1 + eps * np.cos(x)[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    rho_hat = o.spectral()[1, 0, 0, 0]
    assert abs(rho_hat - eps / 2) < 1e-16
    assert abs(f[1, 0, 0, 0] - (-1j * U * rho_hat)) < 1e-15
    mask = np.ones(f.shape[:3], bool)
    mask[1, 0, 0] = False
    assert np.abs(f[..., 0][mask]).max() < 1e-15


def test_rhs_hall_whistler_sign(oracle_mod):
    # B = (b sin z, 0, B0), u = 0, rho = 1: E_Hall = d_i J x B and dB/dt = -curl E
    # gives dBy/dt = d_i B0 b k^2 sin(kz) -> fnl(By) at mz=+1 equals -i d_i B0 b / 2
    n, b, B0, di = 16, 1e-3, 1.0, 0.2
    p = params(n, if_hall=1, ion_inertial_length=di, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    z = np.arange(n) * TWO_PI / n
    prim = uniform_prim(n, b=(0, 0, B0))
    prim[..., 4] = b * np.sin(z)[None, None, :]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    assert abs(f[0, 0, 1, 5] - (-1j * di * B0 * b / 2)) < 1e-16
    assert abs(f[0, 0, n - 1, 5] - (1j * di * B0 * b / 2)) < 1e-16
    assert np.abs(f[..., 4]).max() < 1e-17 and np.abs(f[..., 6]).max() < 1e-17
    p0 = params(n, if_hall=0, if_visc=0, if_resis=0)
    o0 = oracle_mod.Oracle(p0)
    o0.set_state(uu, pr)
    assert np.abs(o0.rhs()[..., 5]).max() < 1e-17    # no Hall: By stays


def test_aeb_uniform_state_decay(oracle_mod):
    # uniform state with expansion: rho ~ R^-2, rho u_x ~ R^-2, rho u_y,z ~ R^-3,
    # Bx ~ R^-2, By,z ~ R^-1 (notes eq. 6-10); tau frozen per step (mhd.f90:248)
    n, R0, Ur, dt = 16, 1.0, 0.5, 0.01
    p = params(n, if_aeb=1, radius0=R0, Ur0=Ur, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    prim = uniform_prim(n, u=(0.1, 0.2, 0.3), b=(1.0, 0.5, 0.25))
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    o.set_fixed_dt(dt)
    u0 = o.spectral()[0, 0, 0, :]
    tau = R0 / Ur
    o.step(1)
    u1 = o.spectral()[0, 0, 0, :]
    rates = np.array([2, 2, 3, 3, 2, 1, 1])
    for v in range(7):
        assert abs(u1[v] / u0[v] - rk3_poly(-rates[v] * dt / tau)) < 1e-13
    # energy: its source term is the adiabatic/kinetic/magnetic decay of each part
    _, prm = o.state()
    g = p.adiabatic_index
    assert prm[0, 0, 0, 3] == pytest.approx(1.0 * math.exp(-2 * g * dt / tau), rel=1e-6)
    assert o.scalars()["radius"] == pytest.approx(R0 + Ur * dt, rel=1e-15)


def test_invariants_seeded(oracle_mod):
    p = params(16, if_hall=1, ion_inertial_length=0.05)
    o = oracle_mod.Oracle(p)
    uu, pr = seeded_state(p, seed=3, kmax=3)
    o.set_state(uu, pr)
    s0 = o.spectral()
    o.step(3)
    s1 = o.spectral()
    assert np.array_equal(s1[0, 0, 0, :], s0[0, 0, 0, :])     # k = 0 modes unchanged without AEB
    assert o.max_divb() < 1e-14                                # solenoidal IC stays solenoidal
    # total energy has no dissipative flux: <e> conserved to round-off
    assert abs(s1[0, 0, 0, 7].real - s0[0, 0, 0, 7].real) < 1e-14
    # kx = 0 plane Hermitian to round-off
    pl = s1[0, :, :, :]
    herm = np.conj(pl[(-np.arange(16)) % 16][:, (-np.arange(16)) % 16])
    assert np.abs(pl - herm).max() < 1e-15


def test_moments_bruteforce(oracle_mod):
    p = params(16)
    o = oracle_mod.Oracle(p)
    uu, pr = seeded_state(p, seed=4, kmax=3)
    o.set_state(uu, pr)
    ave, var, ru2 = o.moments()
    v = np.stack([uu[..., 0], pr[..., 0], pr[..., 1], pr[..., 2], uu[..., 4], uu[..., 5], uu[..., 6], pr[..., 3]], -1)
    assert np.allclose(ave, v.reshape(-1, 8).mean(0), rtol=1e-13, atol=1e-15)
    assert np.allclose(var, (v.reshape(-1, 8) ** 2).mean(0) - v.reshape(-1, 8).mean(0) ** 2, rtol=1e-10, atol=1e-15)
    for i in range(3):
        assert ru2[i] == pytest.approx((uu[..., 0] * (pr[..., i] - ave[1 + i]) ** 2).mean(), rel=1e-12)


def test_divb_bruteforce(oracle_mod):
    p = params(16)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(5)
    prim = uniform_prim(16)
    prim[..., 4:7] += 1e-2 * rng.standard_normal((16, 16, 16, 3))
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    s = o.spectral()
    m = np.fft.fftfreq(16, 1 / 16)
    kx = np.arange(9)[:, None, None]
    d = 1j * (kx * s[..., 4] + m[None, :, None] * s[..., 5] + m[None, None, :] * s[..., 6])
    assert o.max_divb() == pytest.approx(np.abs(d).max(), rel=1e-13)


@pytest.mark.parametrize("di,dealias", [(0.0, 1), (0.3, 1), (0.3, 3), (0.0, 3)])
def test_circular_wave_closed_form(oracle_mod, di, dealias):
    # an exact nonlinear solution of compressible (Hall-)MHD: the whole composed step
    # (transforms, 18 fluxes, Hall EMF, spectral RHS, RK3, dealiasing) must reproduce
    # rk3_amplification(-i omega dt)^n on its one mode and leave every other mode at rest
    n, m, dt, nst = 32, (2, 4, 4), 2e-3, 6
    p = params(n, if_visc=0, if_resis=0, if_hall=1 if di else 0, ion_inertial_length=di,
               dealias_option=dealias)
    prim, omega = circular_wave(p, m, 0.05, di)
    uu, pr = cons_from_prim(prim, p.adiabatic_index)
    o = oracle_mod.Oracle(p)
    o.set_state(uu, pr)
    o.set_fixed_dt(dt)
    s0 = o.spectral()
    o.step(nst)
    s1 = o.spectral()
    g = rk3_amplification(-1j * omega * dt) ** nst
    want = np.zeros_like(s0)
    want[0, 0, 0, :] = s0[0, 0, 0, :]
    w = wave_index(p, m)
    want[w][1:7] = s0[w][1:7] * g
    scale = np.abs(s0[w][1:7]).max()
    assert np.abs(s1 - want).max() < 1e-13 * scale
    # the check tells RK3 from the exact exponential, and the Hall branch from the other one
    assert abs(g - np.exp(-1j * omega * dt * nst)) > 1e-10
    if di:
        h = 0.5 * di * math.sqrt(36.0)
        other = rk3_amplification(-1j * 6.0 * (math.sqrt(1 + h * h) + h) * dt) ** nst
        assert abs(other - g) > 1e-2


# ---------------------------------------------------------------------------------------------
# Closed forms of the individual flux terms (calc_flux, mhdrhs.f90:60-119) on states where the
# density, the pressure or the field vary, so that each term is tested on its own: every field
# depends on x only (or y only), and the RHS at the lowest modes follows from one flux.

def x_profile(n):
    return np.arange(n) * TWO_PI / n


def test_rhs_momentum_flux_nonuniform_density(oracle_mod):
    # rho = 1 + eps cos x, u = (U, V, 0) uniform, p and B uniform: the x-momentum x-flux is
    # rho U^2 + const and the y-momentum x-flux rho U V (mhdrhs.f90:67-79, m_i u_j with u = m/rho),
    # so fnl2(1) = -i U^2 rho_hat and fnl3(1) = -i U V rho_hat, nothing at any other mode
    n, U, V, eps = 16, 0.3, -0.2, 1e-2
    p = params(n, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    x = x_profile(n)
    prim = uniform_prim(n, u=(U, V, 0), b=(0.7, 0, 0), pr=1.3)
    prim[..., 0] = 1 + eps * np.cos(x)[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    rho_hat = eps / 2
    assert abs(f[1, 0, 0, 1] - (-1j * U * U * rho_hat)) < 1e-15
    assert abs(f[1, 0, 0, 2] - (-1j * U * V * rho_hat)) < 1e-15
    mask = np.ones(f.shape[:3], bool)
    mask[1, 0, 0] = False
    for v in (1, 2, 3):
        assert np.abs(f[..., v][mask]).max() < 1e-15
    assert np.abs(f[1, 0, 0, 3]) < 1e-15


def test_rhs_energy_flux_pressure_term(oracle_mod):
    # p = p0 + eps cos x, u = (U, 0, 0), B = 0, rho = 1: the energy flux (e + ptot) u_x with
    # e = p/(gamma-1) + U^2/2 gives fnl8(1) = -i U gamma/(gamma-1) p_hat (mhdrhs.f90:87-89)
    n, U, eps = 16, 0.4, 1e-2
    p = params(n, if_visc=0, if_resis=0)
    g = p.adiabatic_index
    o = oracle_mod.Oracle(p)
    prim = uniform_prim(n, u=(U, 0, 0), b=(0, 0, 0))
    prim[..., 7] = 1 + eps * np.cos(x_profile(n))[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    assert abs(f[1, 0, 0, 7] - (-1j * U * g / (g - 1) * eps / 2)) < 1e-14
    assert np.abs(f[2:, :, :, 7]).max() < 1e-15


def test_rhs_energy_flux_magnetic_pressure_term(oracle_mod):
    # B = (0, beta cos x, 0), u = (U, 0, 0), rho = p = 1: e + ptot = const + beta^2 cos^2 x
    # (the magnetic pressure appears in e and in ptot), u.B = 0, so the energy flux is
    # U beta^2 cos^2 x + const and fnl8(2) = -2i U beta^2 / 4; with e + p in its place it
    # would be half that (mhdrhs.f90:87-89)
    n, U, beta = 16, 0.4, 0.1
    p = params(n, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    prim = uniform_prim(n, u=(U, 0, 0), b=(0, 0, 0))
    prim[..., 5] = beta * np.cos(x_profile(n))[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    assert abs(f[2, 0, 0, 7] - (-1j * U * beta * beta / 2)) < 1e-15
    assert abs(f[1, 0, 0, 7]) < 1e-15 and np.abs(f[3:, :, :, 7]).max() < 1e-15


def test_rhs_energy_flux_udotb_term(oracle_mod):
    # u = (0, V, 0), B = (B0, beta cos x, 0), rho = p = 1: the x energy flux is
    # (e + ptot) u_x - (u.B) B_x = -V beta B0 cos x, so fnl8(1) = +i V beta B0 / 2; the sign
    # of the (u.B) B term decides the sign (mhdrhs.f90:87-89)
    n, V, beta, B0 = 16, 0.3, 0.05, 1.2
    p = params(n, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    prim = uniform_prim(n, u=(0, V, 0), b=(B0, 0, 0))
    prim[..., 5] = beta * np.cos(x_profile(n))[:, None, None]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    assert abs(f[1, 0, 0, 7] - (1j * V * beta * B0 / 2)) < 1e-15


@pytest.mark.parametrize("rho0", [1.0, 2.0])
def test_rhs_hall_density_factor(oracle_mod, rho0):
    # the whistler-sign state at density rho0: the Hall EMF is (d_i/rho) J x B
    # (mhdrhs.f90:106-119), so fnl(By) at mz = +1 is -i d_i B0 b / (2 rho0)
    n, b, B0, di = 16, 1e-3, 1.0, 0.2
    p = params(n, if_hall=1, ion_inertial_length=di, if_visc=0, if_resis=0)
    o = oracle_mod.Oracle(p)
    z = x_profile(n)
    prim = uniform_prim(n, rho=rho0, b=(0, 0, B0))
    prim[..., 4] = b * np.sin(z)[None, None, :]
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    f = o.rhs()
    assert abs(f[0, 0, 1, 5] - (-1j * di * B0 * b / (2 * rho0))) < 1e-17
    assert abs(f[0, 0, n - 1, 5] - (1j * di * B0 * b / (2 * rho0))) < 1e-17


# ---------------------------------------------------------------------------------------------
# Expanding box: the derivative factors k' (mhdrhs.f90:191-204, :313-326) and k'^2
# (AEBmod.f90:87-125) at a radius R != R0, set through the restart time (evolve_radius):
# plain AEB k' = (kx, ky R0/R, kz R0/R); corotating k' = (kx c + ky s, (ky c - kx s) R0/R,
# kz R0/R) with (c, s) the cosine and sine of the corotating angle.

R0_AEB, UR_AEB, T_AEB = 1.0, 0.5, 2.0       # R = R0 + Ur t = 2: R0/R = 1/2, tau = R/Ur = 4


def aeb_oracle(oracle_mod, n, prim, corot=False, angle=0.0, **kw):
    p = params(n, if_aeb=1, radius0=R0_AEB, Ur0=UR_AEB, if_corotating=1 if corot else 0,
               corotating_angle=angle, if_visc=0, if_resis=0, **kw)
    o = oracle_mod.Oracle(p)
    uu, pr = oracle_mod.cons_from_prim(p, prim)
    o.set_state(uu, pr)
    o.set_time(T_AEB, 0.0)
    assert o.scalars()["radius"] == pytest.approx(2.0, rel=1e-15)
    return o


def kprime(kx, ky, kz, corot, angle, r0r):
    if corot:
        c, s = math.cos(angle), math.sin(angle)
        return kx * c + ky * s, (ky * c - kx * s) * r0r, kz * r0r
    return kx, ky * r0r, kz * r0r


def test_aeb_density_mode_along_y(oracle_mod):
    # rho = 1 + eps cos y, u = (0, V, 0): fnl1(0,1,0) = -i (R0/R) V rho_hat - 2 rho_hat / tau
    n, V, eps = 16, 0.3, 1e-2
    prim = uniform_prim(n, u=(0, V, 0))
    prim[..., 0] = 1 + eps * np.cos(x_profile(n))[None, :, None]
    o = aeb_oracle(oracle_mod, n, prim)
    f = o.rhs()
    rho_hat, r0r, tau = eps / 2, R0_AEB / 2.0, 2.0 / UR_AEB
    assert abs(f[0, 1, 0, 0] - (-1j * r0r * V * rho_hat - 2 * rho_hat / tau)) < 1e-15
    assert abs(f[0, n - 1, 0, 0] - (1j * r0r * V * rho_hat - 2 * rho_hat / tau)) < 1e-15


@pytest.mark.parametrize("axis", [0, 1])
def test_aeb_corotating_mode_mixing(oracle_mod, axis):
    # rho = 1 + eps cos(x) (axis 0) or cos(y) (axis 1), u = (U, V, 0): the mass RHS at the mode
    # is -i (kx' U + ky' V) rho_hat - 2 rho_hat / tau with the rotated, scaled k' of that mode
    n, U, V, eps, ang = 16, 0.3, -0.25, 1e-2, 0.4
    prim = uniform_prim(n, u=(U, V, 0))
    prof = np.cos(x_profile(n))
    prim[..., 0] = 1 + eps * (prof[:, None, None] if axis == 0 else prof[None, :, None])
    o = aeb_oracle(oracle_mod, n, prim, corot=True, angle=ang)
    f = o.rhs()
    m = (1, 0, 0) if axis == 0 else (0, 1, 0)
    kx, ky, _ = kprime(*m, True, ang, R0_AEB / 2.0)
    rho_hat, tau = eps / 2, 2.0 / UR_AEB
    want = -1j * (kx * U + ky * V) * rho_hat - 2 * rho_hat / tau
    assert abs(f[m[0], m[1], m[2], 0] - want) < 1e-15
    # the same mode with the opposite sign of the rotation's kx s term differs by O(eps)
    kx2, ky2 = (kx, -ky) if axis == 0 else (kx, ky)
    if axis == 0:
        assert abs(-1j * (kx2 * U + ky2 * V) * rho_hat - 2 * rho_hat / tau - want) > 1e-4


def test_aeb_hall_whistler_scaling(oracle_mod):
    # B = (b sin z, 0, B0), u = 0, rho = 2 at R = 2 R0: J_y = (R0/R) b cos z (the current's
    # kz', mhdrhs.f90:313-326) and dBy/dt = -(R0/R) d_z E_x (the curl's kz', :198-204):
    # fnl(By)(mz=1) = -i (R0/R)^2 (d_i/rho) b B0 / 2; Bx decays at 2/tau (AEB, :235-251)
    n, b, B0, di, rho0 = 16, 1e-3, 1.0, 0.2, 2.0
    prim = uniform_prim(n, rho=rho0, b=(0, 0, B0))
    prim[..., 4] = b * np.sin(x_profile(n))[None, None, :]
    o = aeb_oracle(oracle_mod, n, prim, if_hall=1, ion_inertial_length=di)
    f = o.rhs()
    r0r, tau = R0_AEB / 2.0, 2.0 / UR_AEB
    assert abs(f[0, 0, 1, 5] - (-1j * r0r * r0r * di * B0 * b / (2 * rho0))) < 1e-17
    bx_hat = b / 2j
    assert abs(f[0, 0, 1, 4] - (-2 * bx_hat / tau)) < 1e-17


@pytest.mark.parametrize("corot", [False, True])
def test_aeb_ksquare_is_norm_of_kprime(oracle_mod, corot):
    # update_ksquare (AEBmod.f90:87-125) expands |k'|^2; compare with |k'|^2 formed from the
    # rotation and the scaling of the derivative factors
    n, ang = 16, 0.4
    o = aeb_oracle(oracle_mod, n, uniform_prim(n), corot=corot, angle=ang)
    got = o.ksquare()
    kx = np.arange(n // 2 + 1)[:, None, None].astype(float)
    m = np.arange(n)
    m = np.where(m <= n // 2, m, m - n).astype(float)   # +Nyquist (mhdinit.f90:83-111)
    ky, kz = m[None, :, None], m[None, None, :]
    a, b_, c_ = kprime(kx, ky, kz, corot, ang, R0_AEB / 2.0)
    want = a * a + b_ * b_ + c_ * c_
    assert np.abs(got - want).max() < 1e-12 * want.max()


# ---------------------------------------------------------------------------------------------
# vardt's hysteresis (mhd.f90:422-428): the previous dt is kept while it is within 2 % of the
# new CFL dt and replaced by it otherwise. On a uniform state the CFL dt never changes.

@pytest.mark.parametrize("ratio,kept", [(1.01, True), (0.99, True), (1.03, False), (0.97, False)])
def test_dt_hysteresis(oracle_mod, ratio, kept):
    n = 16
    p = params(n, cfl=0.4)
    o = oracle_mod.Oracle(p)
    uu, pr = oracle_mod.cons_from_prim(p, uniform_prim(n))
    o.set_state(uu, pr)
    raw, _ = o.cfl()
    dtmin = 0.4 * raw
    dt_prev = ratio * dtmin
    o.set_time(0.0, dt_prev)
    o.step(2)
    want = dt_prev if kept else dtmin
    assert o.scalars()["dt"] == want
    assert o.scalars()["time"] == pytest.approx(2 * want, rel=1e-15)
