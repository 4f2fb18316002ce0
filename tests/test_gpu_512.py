"""The 512-point kernels: the x pass, the y/z inverse and forward passes and the
z pass with the RK3 update at n = 512.

* One axis of 512 points and two of 16, with Hall and the corotating expanding
  box (the physics of the 512^3 expanding-box workload), both dealiasing
  masks: FP64 against the oracle (relative L2 <= 1e-11, dt and time to
  1e-13), mixed precision within 1e-4 of the FP64 oracle.
* The full 512^3 expanding-box workload (outward Alfven spectrum k^-1 in
  shells 1-32, db/B0 = 0.5, d_i = 0.02, nu = eta = 5e-5, corotating box with
  R0 = 1, Ur0 = 0.5): after two steps every mode outside the kept set is
  exactly zero, the kx = 0 plane is Hermitian, and mixed precision is within
  1e-4 of FP64 at equal time levels.
* The circularly polarised wave (tests/helpers.py:circular_wave) at 512^3
  without expansion: its one mode follows the RK3 amplification and every
  other mode stays at rest.
"""
import math

import numpy as np
import pytest

from laps_amd.ic import band_limited_state, cons_from_prim
from laps_amd.params import LAPS_FP64, LAPS_MIXED
from tests.helpers import (circular_wave, kept_modes, params, rel_l2, rk3_amplification, seeded_state,
                           wave_index)

pytestmark = pytest.mark.gpu

AXES = [(512, 16, 16), (16, 512, 16), (16, 16, 512)]


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


def expanding(shape, dealias, precision=LAPS_FP64):
    return params(shape, if_hall=1, ion_inertial_length=0.02, viscosity=5e-5, resistivity=5e-5,
                  if_aeb=1, if_corotating=1, corotating_angle=0.3, radius0=1.0, Ur0=0.5,
                  dealias_option=dealias, cfl=0.3 if dealias == 1 else 0.2, precision=precision)


@pytest.mark.parametrize("dealias", [1, 3])
@pytest.mark.parametrize("shape", AXES)
def test_512_axis_vs_oracle(rt, oracle_mod, shape, dealias):
    print(f"512 axis {shape} dealias {dealias}", flush=True)
    p = expanding(shape, dealias)
    uu, pr = seeded_state(p, seed=4, kmax=2, du=0.3, db=0.3)
    o = oracle_mod.Oracle(p)
    try:
        o.set_state(uu, pr)
        o.step(3)
        so, sco = o.spectral(), o.scalars()
    finally:
        o.close()
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.step(3)
        sg, scg = g.spectral(), g.scalars()
    err = rel_l2(sg, so)
    print(f"  fp64 rel l2 {err:.2e}, dt {scg['dt']:.6e} vs {sco['dt']:.6e}", flush=True)
    assert err < 1e-11
    assert scg["dt"] == pytest.approx(sco["dt"], rel=1e-13)
    assert scg["time"] == pytest.approx(sco["time"], rel=1e-13)
    assert scg["radius"] == pytest.approx(sco["radius"], rel=1e-14)
    pm = expanding(shape, dealias, LAPS_MIXED)
    with rt.Laps(pm) as m:
        m.set_state(uu, pr)
        m.step(3)
        sm = m.spectral()
    errm = rel_l2(sm, so)
    print(f"  mixed rel l2 {errm:.2e}", flush=True)
    assert errm < 1e-4


def workload512(precision):
    return params(512, if_hall=1, ion_inertial_length=0.02, viscosity=5e-5, resistivity=5e-5,
                  if_aeb=1, if_corotating=1, radius0=1.0, Ur0=0.5, dealias_option=1, cfl=0.3,
                  precision=precision)


@pytest.mark.timeout(600)
def test_512_cube_expanding_box(rt):
    n = 512
    print("512^3 expanding-box state", flush=True)
    with rt.Laps(workload512(LAPS_FP64)) as g:
        uu, pr = band_limited_state(n, 128, 5.0 / 3.0, g.fft_c2r, seed=2, kmin=1, kmax=32, slope=-1.0,
                                    du=0.5, db=0.5, alfvenic=True)
    keep = kept_modes(workload512(LAPS_FP64))
    spec = {}
    dts = []
    for prec in (LAPS_FP64, LAPS_MIXED):
        name = "fp64" if prec == LAPS_FP64 else "mixed"
        print(f"512^3 {name}: 2 steps", flush=True)
        with rt.Laps(workload512(prec)) as g:
            g.set_state(uu, pr)
            for i in range(2):
                if prec == LAPS_MIXED:
                    g.set_fixed_dt(dts[i])   # the FP64 run's dt sequence: equal time levels
                g.step(1)
                if prec == LAPS_FP64:
                    dts.append(g.scalars()["dt"])
            s1 = g.spectral()
            sc = g.scalars()
        assert np.all(np.isfinite(s1)), name
        # every mode outside the kept set is exactly zero after a step
        drop = ~keep
        assert not any(np.any(s1[..., f][drop]) for f in range(8)), name
        pl = s1[0]
        ineg = (-np.arange(n)) % n
        hdev = np.abs(pl - np.conj(pl[ineg][:, ineg])).max() / np.abs(s1).max()
        print(f"  t {sc['time']:.6e} R {sc['radius']:.6f} Hermitian deviation {hdev:.2e}", flush=True)
        assert hdev < (1e-13 if prec == LAPS_FP64 else 1e-6), name
        assert sc["radius"] == pytest.approx(1.0 + 0.5 * sc["time"], rel=1e-14)
        spec[prec] = s1
        del s1, pl
    err = rel_l2(spec[LAPS_MIXED], spec[LAPS_FP64])
    print(f"512^3 mixed vs fp64 after 2 steps: rel l2 {err:.2e}", flush=True)
    assert err < 1e-4


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,dealias", [(LAPS_FP64, 1), (LAPS_MIXED, 3)])
def test_circular_wave_512(rt, precision, dealias):
    n, m, dt, nst, di = 512, (32, 64, 64), 5e-4, 4, 0.02
    print(f"512^3 circular wave precision {precision} dealias {dealias}", flush=True)
    p = params(n, if_visc=0, if_resis=0, if_hall=1, ion_inertial_length=di, dealias_option=dealias,
               precision=precision)
    prim, omega = circular_wave(p, m, 0.02, di)
    uu, pr = cons_from_prim(prim, p.adiabatic_index)
    del prim
    w = wave_index(p, m)
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        del uu, pr
        g.set_fixed_dt(dt)
        s0 = g.spectral()
        c0 = s0[w][1:7].copy()
        k0 = s0[0, 0, 0, :].copy()
        del s0
        g.step(nst)
        s1 = g.spectral()
    want = c0 * rk3_amplification(-1j * omega * dt) ** nst
    scale = np.abs(c0).max()
    werr = np.abs(s1[w][1:7] - want).max() / scale
    s1[w][1:7] = 0
    s1[0, 0, 0, :] -= k0
    resid = np.abs(s1).max() / scale
    print(f"  omega {omega:.4f}, mode error {werr:.2e}, other modes {resid:.2e}", flush=True)
    tol = 1e-11 if precision == LAPS_FP64 else 1e-5
    assert werr < tol
    assert resid < tol
    assert math.isfinite(omega)
