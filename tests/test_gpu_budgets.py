"""Budgets (laps_budgets) on the GPU: against the brute-force numpy reference
on the same state, against the oracle's state after steps, and the closed-form
resistive decay of a force-free (Beltrami) field's magnetic helicity."""
import math

import numpy as np
import pytest

from tests.helpers import budgets_numpy, params, seeded_state
from tests.test_budgets import beltrami

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def laps_mod():
    from laps_amd import runtime
    runtime.lib()
    return runtime


def close(g, ref, rtol):
    # component-wise, on the scale of the largest quadratic budget (energies)
    scale = max(abs(ref[0]), abs(ref[1]), abs(ref[2]))
    for a, b in zip(g, ref):
        assert abs(a - b) <= rtol * max(abs(b), 1e-6 * scale), (a, b)


def test_budgets_of_uploaded_state(laps_mod):
    p = params(32, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=11, du=0.2, db=0.3)
    g = laps_mod.Laps(p)
    g.set_state(uu, pr)
    got = np.array(list(g.budgets().values()))
    ref = budgets_numpy(uu, pr, g.spectral(), p)
    print("budgets", got)
    close(got, ref, 1e-12)
    g.close()


def test_budgets_after_steps_vs_oracle(laps_mod, oracle_mod):
    p = params(32, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=12, du=0.2, db=0.3)
    g = laps_mod.Laps(p)
    o = oracle_mod.Oracle(p)
    g.set_state(uu, pr)
    o.set_state(uu, pr)
    g.step(5)
    o.step(5)
    uo, po = o.state()
    ref = budgets_numpy(uo, po, o.spectral(), p)
    got = np.array(list(g.budgets().values()))
    close(got, ref, 1e-10)
    g.close()
    o.close()


def test_resistive_decay_of_beltrami_helicity(laps_mod):
    # B = b (sin kz, cos kz, 0) with u = 0 and uniform rho, p is force-free (J = k B), also
    # for the Hall EMF (J x B = 0): only the implicit resistive divisor acts, so each RK3
    # substage divides B_hat by 1 + tau_s eta k^2 (rktmod.f90:40-61) and <A.B>, <J.B> fall
    # by the square of the product over the three substages.
    eta, k, b, dt = 2e-2, 3, 0.4, 1e-2
    p = params(32, if_hall=1, ion_inertial_length=0.05, resistivity=eta, viscosity=1e-3)
    uu, pr = beltrami(p, k, b, b0=0.0)
    g = laps_mod.Laps(p)
    g.set_state(uu, pr)
    g.set_fixed_dt(dt)
    b0 = g.budgets()
    g.step(1)
    b1 = g.budgets()
    fac = 1.0
    for tau in (8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0):
        fac /= (1.0 + tau * dt * eta * k * k) ** 2
    print("helicity ratio", b1["magnetic_helicity"] / b0["magnetic_helicity"], "expected", fac)
    assert math.isclose(b0["magnetic_helicity"], b * b / k, rel_tol=1e-12)
    assert math.isclose(b1["magnetic_helicity"] / b0["magnetic_helicity"], fac, rel_tol=1e-10)
    assert math.isclose(b1["current_helicity"] / b0["current_helicity"], fac, rel_tol=1e-10)
    assert abs(b1["cross_helicity"]) < 1e-12
    g.close()
