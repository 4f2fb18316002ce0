"""Every FFT variant the plan autotuner can pick for the inverse passes (two plans x two tile
widths, LAPS_FFT_PLAN=z,y), at 256 points along y and z where the low-radix variant is the
8·8·4 plan with two last-stage butterflies per work-item, against the oracle.

The autotuner picks by timing, so on a given GPU some variants never run in the other tests;
each is forced here. FP64 with phase shift and Hall (both chains of a z-inverse slab in one
work-group) must match the oracle at the 1e-11 bar, and the mixed step stay within 1e-5 of it.
"""
import os

import pytest

from laps_amd.params import LAPS_FP64, LAPS_MIXED
from tests.helpers import params, rel_l2, seeded_state

pytestmark = pytest.mark.gpu

SHAPE = (32, 256, 256)


def case(precision=LAPS_FP64):
    return params(SHAPE, if_hall=1, ion_inertial_length=0.05, viscosity=2e-4, resistivity=2e-4,
                  dealias_option=3, cfl=0.2, precision=precision)


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


@pytest.fixture(scope="module")
def reference(oracle_mod):
    p = case()
    uu, pr = seeded_state(p, seed=4, kmax=12, du=0.3, db=0.3)
    o = oracle_mod.Oracle(p)
    try:
        o.set_state(uu, pr)
        o.step(2)
        return uu, pr, o.spectral(), o.scalars()
    finally:
        o.close()


def run(rt, p, uu, pr, variants):
    old = os.environ.get("LAPS_FFT_PLAN")
    os.environ["LAPS_FFT_PLAN"] = variants
    try:
        with rt.Laps(p) as g:
            g.set_state(uu, pr)
            g.step(2)
            return g.spectral(), g.scalars()
    finally:
        if old is None:
            os.environ.pop("LAPS_FFT_PLAN", None)
        else:
            os.environ["LAPS_FFT_PLAN"] = old


@pytest.mark.timeout(300)
@pytest.mark.parametrize("variants", ["0,0", "1,1", "2,2", "3,3", "1,0", "0,3"])
def test_inverse_variants_fp64(rt, reference, variants):
    uu, pr, so, sco = reference
    sg, scg = run(rt, case(), uu, pr, variants)
    err = rel_l2(sg, so)
    print(f"{SHAPE} FP64 variants z,y = {variants}: rel l2 {err:.2e}", flush=True)
    assert err < 1e-11
    assert scg["dt"] == pytest.approx(sco["dt"], rel=1e-13)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("variants", ["0,0", "1,1", "2,2", "3,3"])
def test_inverse_variants_mixed(rt, reference, variants):
    uu, pr, so, _ = reference
    sg, _ = run(rt, case(LAPS_MIXED), uu, pr, variants)
    err = rel_l2(sg, so)
    print(f"{SHAPE} mixed variants z,y = {variants}: rel l2 {err:.2e}", flush=True)
    assert err < 1e-5
