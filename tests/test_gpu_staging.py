"""Output staging (laps_stage_output / laps_staged_output): the outNNN.dat payload formed
on the device and copied to pinned memory while later steps run.

The payload must equal what output_uu writes from laps_get_state (rho, rho u or u, B, e or p;
mhdoutput.f90:72-131) bit for bit, and it must be the state of the step it was staged at even
when more steps are queued and run before it is read.
"""
import numpy as np
import pytest

from tests.helpers import params, seeded_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


def payload(uu, pr, primitive):
    w = np.array(uu, order="F", copy=True)
    if primitive:
        w[..., 1:4] = pr[..., 0:3]
        w[..., 7] = pr[..., 3]
    return w


@pytest.mark.parametrize("case", [dict(), dict(dealias_option=3, if_hall=1, ion_inertial_length=0.05),
                                  dict(precision=1, if_hall=1, ion_inertial_length=0.05)])
def test_staged_payload_is_the_state_of_its_step(rt, case):
    p = params((64, 32, 96), **case)
    uu, pr = seeded_state(p, seed=5, kmax=5, du=0.2, db=0.2)
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        for primitive in (True, False):
            g.step(2)
            u2, p2 = g.state()
            sc2 = g.scalars()
            g.stage_output(primitive)
            g.step(3)              # queued behind the staging copy; runs while it proceeds
            w, t, dt, ist = g.staged_output()
            print(f"case {case} primitive {primitive}: staged at istep {ist}, t {t:.6e}", flush=True)
            assert np.array_equal(w, payload(u2, p2, primitive))
            assert (t, dt, ist) == (sc2["time"], sc2["dt"], sc2["istep"])
            # the steps after the staging point went on: the state has moved
            u5, _ = g.state()
            assert g.scalars()["istep"] == ist + 3 and not np.array_equal(u5, u2)
            # staging again right away gives the current step
            g.stage_output(primitive)
            w5, _, _, ist5 = g.staged_output()
            assert ist5 == ist + 3
            assert np.array_equal(w5, payload(*g.state(), primitive))


def test_staged_output_needs_a_staging_call(rt):
    p = params(32)
    uu, pr = seeded_state(p, seed=6)
    with rt.Laps(p) as g:
        g.set_state(uu, pr)
        with pytest.raises(rt.LapsError, match="no output staged"):
            g.staged_output()
