"""Ensembles (laps_ens_*): boxes of different sizes, physics and precision
advanced together, one hipGraph launch per step after the first.

Every box of the ensemble must hold exactly (bit for bit) the state of the
same box stepped on its own with laps_step: the graph replays the very
launches of one eager step. Each replay is checked, with host reads and
synchronisation between replays, and a box given a new state or a fixed dt
makes the next step capture again. One box is also held against the oracle.
"""
import numpy as np
import pytest

from laps_amd.params import LAPS_MIXED
from tests.helpers import params, rel_l2, seeded_state

pytestmark = pytest.mark.gpu

CASES = [
    dict(n=32, seed=11),
    dict(n=64, seed=12, precision=LAPS_MIXED, if_hall=1, ion_inertial_length=0.05),
    dict(n=32, seed=13, dealias_option=3, if_hall=1, ion_inertial_length=0.05, cfl=0.2),
    dict(n=(64, 32, 32), seed=14, if_aeb=1, radius0=1.0, Ur0=0.5, Lx=9.0),
    dict(n=32, seed=15, precision=LAPS_MIXED, dealias_option=2),
    dict(n=(96, 32, 32), seed=16, precision=LAPS_MIXED, if_hall=1, ion_inertial_length=0.05),  # radix 3
]


@pytest.fixture(scope="module")
def rt():
    from laps_amd import runtime
    runtime.lib()
    return runtime


@pytest.fixture(autouse=True)
def fixed_plan(monkeypatch):
    # the autotuner times both FFT plans per context and may keep different ones for
    # two copies of a box; bit-for-bit comparisons need the same plan in both
    monkeypatch.setenv("LAPS_FFT_PLAN", "0")


def setup(rt):
    ps, states = [], []
    for c in CASES:
        c = dict(c)
        n, seed = c.pop("n"), c.pop("seed")
        p = params(n, **c)
        ps.append(p)
        states.append(seeded_state(p, seed=seed, kmax=6, du=0.2, db=0.2))
    ens = rt.Ensemble(ps)
    solo = [rt.Laps(p) for p in ps]
    for i, (uu, pr) in enumerate(states):
        ens.box(i).set_state(uu, pr)
        solo[i].set_state(uu, pr)
    return ps, states, ens, solo


def check_equal(ens, solo, tag):
    for i, s in enumerate(solo):
        a, b = ens.box(i).spectral(), s.spectral()
        assert np.array_equal(a, b), f"{tag}: box {i} differs (rel l2 {rel_l2(a, b):.3e})"
        assert ens.box(i).scalars() == s.scalars(), f"{tag}: box {i} scalars"


def test_graph_replays_match_eager_steps(rt):
    ps, states, ens, solo = setup(rt)
    try:
        # first step eager, then one graph launch per step; every replay checked
        for k in range(4):
            ens.step(1)
            for s in solo:
                s.step(1)
            print(f"step {k + 1}", flush=True)
            check_equal(ens, solo, f"step {k + 1}")
        ens.step(5)
        for s in solo:
            s.step(5)
        check_equal(ens, solo, "5 more steps")
        # a new state in one box and a fixed dt in another: captured again
        uu, pr = seeded_state(ps[1], seed=99, kmax=6, du=0.2, db=0.2)
        ens.box(1).set_state(uu, pr)
        solo[1].set_state(uu, pr)
        ens.box(4).set_fixed_dt(1e-3)
        solo[4].set_fixed_dt(1e-3)
        ens.step(3)
        for s in solo:
            s.step(3)
        check_equal(ens, solo, "after re-capture")
        steps = [ens.box(i).scalars()["istep"] for i in range(len(ens))]
        assert steps == [12, 3, 12, 12, 12, 12]
    finally:
        ens.close()
        for s in solo:
            s.close()


def test_ensemble_box_vs_oracle(rt, oracle_mod):
    p = params(32, if_hall=1, ion_inertial_length=0.05)
    uu, pr = seeded_state(p, seed=21, kmax=6, du=0.2, db=0.2)
    ens = rt.Ensemble([p, params(64)])
    try:
        ens.box(0).set_state(uu, pr)
        uu2, pr2 = seeded_state(params(64), seed=22, kmax=6)
        ens.box(1).set_state(uu2, pr2)
        o = oracle_mod.Oracle(p)
        o.set_state(uu, pr)
        ens.step(6)
        o.step(6)
        assert rel_l2(ens.box(0).spectral(), o.spectral()) < 1e-11
        assert ens.box(0).scalars()["dt"] == pytest.approx(o.scalars()["dt"], rel=1e-13)
    finally:
        ens.close()


GROUPED = [
    (dict(n=32, if_hall=1, ion_inertial_length=0.05), (31, 32, 33)),
    (dict(n=32, precision=LAPS_MIXED, dealias_option=3, if_hall=1, ion_inertial_length=0.05, cfl=0.2), (34, 35)),
    (dict(n=(64, 32, 32), if_aeb=1, if_corotating=1, corotating_angle=0.3, radius0=1.0, Ur0=0.5), (36, 37)),
    (dict(n=(96, 32, 32), precision=LAPS_MIXED, if_hall=1, ion_inertial_length=0.05), (38,)),
]


def test_batched_groups_match_eager_steps(rt):
    """Boxes of identical parameters are stepped by shared launches (one launch per kernel for
    the group, the boxes as grid rows): every box must still equal the same box stepped alone,
    also after one box of a group gets a new state and another a fixed dt (that group is then
    captured box by box)."""
    ps, states = [], []
    for kw, seeds in GROUPED:
        kw = dict(kw)
        n = kw.pop("n")
        for sd in seeds:
            p = params(n, **kw)
            ps.append(p)
            states.append(seeded_state(p, seed=sd, kmax=6, du=0.2, db=0.2))
    ens = rt.Ensemble(ps)
    solo = [rt.Laps(p) for p in ps]
    try:
        assert ens.groups() == len(GROUPED)
        for i, (uu, pr) in enumerate(states):
            ens.box(i).set_state(uu, pr)
            solo[i].set_state(uu, pr)
        for k in range(3):
            ens.step(1)
            for s in solo:
                s.step(1)
            print(f"batched step {k + 1}", flush=True)
            check_equal(ens, solo, f"batched step {k + 1}")
        ens.step(6)
        for s in solo:
            s.step(6)
        check_equal(ens, solo, "6 more batched steps")
        ens.step(17)   # two launches of the 8-step graph and one of the 1-step graph
        for s in solo:
            s.step(17)
        check_equal(ens, solo, "17 more steps")
        uu, pr = seeded_state(ps[1], seed=90, kmax=6, du=0.2, db=0.2)
        ens.box(1).set_state(uu, pr)
        solo[1].set_state(uu, pr)
        ens.box(3).set_fixed_dt(2e-3)
        solo[3].set_fixed_dt(2e-3)
        ens.step(4)
        for s in solo:
            s.step(4)
        check_equal(ens, solo, "after a new state and a fixed dt")
        times = [ens.box(i).scalars()["time"] for i in range(len(ens))]
        assert len(set(times)) > 3   # every box keeps its own CFL dt
    finally:
        ens.close()
        for s in solo:
            s.close()
