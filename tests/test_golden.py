"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_fixtures.py).

CPU: the oracle reproduces them bit for bit. GPU: the HIP path through the C
ABI matches them to the FP64 parity bound.
"""
import os

import numpy as np
import pytest

from tests.golden.make_fixtures import CASES, run_case

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, f"{name}.npz")))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    g = load(name)
    r = run_case(CASES[name])
    assert np.array_equal(r["idx"], g["idx"])
    assert np.array_equal(r["dts"], g["dts"])
    np.testing.assert_allclose(r["uhat"], g["uhat"], rtol=0, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_golden(name):
    from laps_amd.ic import cons_from_prim, seeded_primitive
    from laps_amd.runtime import Laps
    from tests.helpers import params, rel_l2
    spec = CASES[name]
    g = load(name)
    n = spec["n"]
    p = params(n, **spec["p"])
    uu, pr = cons_from_prim(seeded_primitive((n, n, n), seed=spec["seed"], kmin=1, kmax=spec["kmax"]),
                           p.adiabatic_index)
    with Laps(p) as box:
        box.set_state(uu, pr)
        dts = []
        for _ in range(spec["steps"]):
            box.step(1)
            dts.append(box.scalars()["dt"])
        uhat = box.spectral()
        i = g["idx"].astype(np.int64)
        got = uhat[i[:, 0], i[:, 1], i[:, 2], :]
        assert rel_l2(got, g["uhat"]) < 1e-11
        np.testing.assert_allclose(dts, g["dts"], rtol=1e-13)
        ave, var, ru2 = box.moments()
        np.testing.assert_allclose(ave, g["ave"], rtol=1e-12, atol=1e-15)
