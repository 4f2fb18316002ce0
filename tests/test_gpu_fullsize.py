"""The HIP path at the full 256^3 size of the Hall-MHD step it is measured on.

* FP64 against the oracle on the same seeded state, two steps with phase-shift
  dealiasing and Hall (the oracle takes about 10 s a step here).
* Properties that hold at any size: k = 0 modes untouched without expansion,
  the dealiased modes exactly zero, the kx = 0 plane Hermitian, div B at
  round-off, mixed precision within 1e-4 of FP64.
* The circularly polarised Alfven / whistler wave, an exact nonlinear solution
  (tests/helpers.py:circular_wave): its one mode must be multiplied by the RK3
  amplification of -i omega dt per step and every other mode stay at rest.
"""
import numpy as np
import pytest

from laps_amd.ic import cons_from_prim
from laps_amd.params import LAPS_FP64, LAPS_MIXED
from tests.helpers import (circular_wave, kept_modes, params, rel_l2, rk3_amplification,
                           seeded_state, wave_index)

pytestmark = pytest.mark.gpu

N = 256


def hall256(precision=LAPS_FP64, dealias=3):
    return params(N, if_hall=1, ion_inertial_length=0.05, viscosity=2e-4, resistivity=2e-4,
                  dealias_option=dealias, cfl=0.2 if dealias == 3 else 0.4, precision=precision)


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


@pytest.fixture(scope="module")
def state256():
    print("seeded 256^3 state", flush=True)
    return seeded_state(hall256(), seed=1, kmax=16, du=0.3, db=0.3)


@pytest.mark.timeout(400)
def test_fp64_vs_oracle_256(rt, oracle_mod, state256):
    uu, pr = state256
    p = hall256()
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.step(2)
        sg, scg = g.spectral(), g.scalars()
    o = oracle_mod.Oracle(p)
    try:
        o.set_state(uu, pr)
        o.step(2)
        so, sco = o.spectral(), o.scalars()
    finally:
        o.close()
    err = rel_l2(sg, so)
    print(f"256^3 Hall phase shift, 2 steps: rel l2 {err:.2e}, dt {scg['dt']} vs {sco['dt']}")
    assert err < 1e-11
    assert scg["dt"] == pytest.approx(sco["dt"], rel=1e-13)
    assert scg["time"] == pytest.approx(sco["time"], rel=1e-13)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("dealias", [3, 1])
def test_invariants_and_mixed_256(rt, state256, dealias):
    uu, pr = state256
    out = {}
    for prec in (LAPS_FP64, LAPS_MIXED):
        p = hall256(prec, dealias)
        with rt.Laps(p) as g:
            g.set_state(uu, pr)
            s0 = g.spectral()
            g.step(3)
            s1 = g.spectral()
            divb, sc = g.max_divb(), g.scalars()
        out[prec] = s1
        name = "fp64" if prec == LAPS_FP64 else "mixed"
        scale = np.abs(s1).max()
        # no flux divergence and no diffusion at k = 0
        assert np.array_equal(s1[0, 0, 0, :], s0[0, 0, 0, :]), name
        # every mode outside the kept set is exactly zero after a step
        drop = ~kept_modes(p)
        assert not np.any(s1[drop]), name
        pl = s1[0]
        herm = np.conj(pl[(-np.arange(N)) % N][:, (-np.arange(N)) % N])
        hdev = np.abs(pl - herm).max() / scale
        print(f"{name} dealias {dealias}: div B {divb:.2e}, Hermitian deviation {hdev:.2e}, "
              f"dt {sc['dt']:.6e}", flush=True)
        assert hdev < (1e-13 if prec == LAPS_FP64 else 1e-6), name
        assert divb < (1e-11 if prec == LAPS_FP64 else 1e-5), name
        assert np.all(np.isfinite(s1)), name
    err = rel_l2(out[LAPS_MIXED], out[LAPS_FP64])
    print(f"mixed vs fp64 after 3 steps: rel l2 {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("precision,dealias,di", [(LAPS_FP64, 3, 0.05), (LAPS_FP64, 1, 0.05),
                                                  (LAPS_FP64, 3, 0.0), (LAPS_MIXED, 3, 0.05)])
def test_circular_wave_256(rt, precision, dealias, di):
    m, dt, nst = (16, 32, 32), 1e-3, 8
    p = params(N, if_visc=0, if_resis=0, if_hall=1 if di else 0, ion_inertial_length=di,
               dealias_option=dealias, precision=precision)
    prim, omega = circular_wave(p, m, 0.02, di)
    uu, pr = cons_from_prim(prim, p.adiabatic_index)
    del prim
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        g.set_fixed_dt(dt)
        s0 = g.spectral()
        g.step(nst)
        s1 = g.spectral()
    amp = rk3_amplification(-1j * omega * dt) ** nst
    w = wave_index(p, m)
    want = s0[w][1:7] * amp
    scale = np.abs(s0[w][1:7]).max()
    werr = np.abs(s1[w][1:7] - want).max() / scale
    rest = s1.copy()
    rest[w][1:7] = 0
    rest[0, 0, 0, :] -= s0[0, 0, 0, :]
    resid = np.abs(rest).max() / scale
    print(f"precision {precision} dealias {dealias} d_i {di}: omega {omega:.4f}, mode error {werr:.2e}, "
          f"other modes {resid:.2e}", flush=True)
    tol = 1e-11 if precision == LAPS_FP64 else 1e-5
    assert werr < tol
    assert resid < tol
    # the bound separates RK3 from the exact exponential
    assert abs(amp - np.exp(-1j * omega * dt * nst)) > 100 * tol or precision != LAPS_FP64
