"""Writes tests/golden/*.npz: small regression vectors of the LAPS step.

These are frozen output of the CPU oracle (oracle/laps_oracle.c), not an
independent pin of it. The reference ships no tests or fixtures, and its own
step cannot be built here without a stand-in: its transforms call FFTW3
(fftw.f90:8, not vendored, absent from this image) and it needs MPI. Most of
the reference's sources compile with amdflang; fftw.f90 and its dependants
(mhdrhs.f90, mhd.f90) do not, and building them against a stand-in FFTW
would not be the reference. What pins the oracle is tests/test_oracle_pins.py
and tests/test_phase_shift.py (library routines, brute force, closed forms,
invariants). These vectors only catch a change that moves a result: a change
to the oracle or to the initial-condition generator fails the CPU test, and
the GPU path is checked against the same numbers.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from laps_amd.ic import cons_from_prim, seeded_primitive  # noqa: E402
from oracle import binding as ob  # noqa: E402
from tests.helpers import params  # noqa: E402

CASES = {
    "hall_aeb_corot_16": dict(n=16, seed=21, kmax=3, steps=3,
                              p=dict(if_hall=1, ion_inertial_length=0.05, if_aeb=1, if_corotating=1,
                                     corotating_angle=0.4, radius0=1.0, Ur0=0.5)),
    "filter_explicit_16": dict(n=16, seed=22, kmax=3, steps=3, store_kmax=4,
                               p=dict(dealias_option=2, if_visc_exp=1, if_resis_exp=1)),
}


def run_case(spec):
    n = spec["n"]
    p = params(n, **spec["p"])
    prim = seeded_primitive((n, n, n), seed=spec["seed"], kmin=1, kmax=spec["kmax"])
    uu, pr = cons_from_prim(prim, p.adiabatic_index)
    o = ob.Oracle(p)
    o.set_state(uu, pr)
    dts = []
    for _ in range(spec["steps"]):
        o.step(1)
        dts.append(o.scalars()["dt"])
    uhat = o.spectral()
    ave, var, ru2 = o.moments()
    keep = np.abs(uhat).sum(axis=3) > 0
    if "store_kmax" in spec:   # the filter keeps every mode: store the low-k block only
        m = np.fft.fftfreq(n, 1.0 / n)
        km = spec["store_kmax"]
        keep &= (np.arange(n // 2 + 1)[:, None, None] <= km) & (np.abs(m)[None, :, None] <= km) & \
                (np.abs(m)[None, None, :] <= km)
    idx = np.argwhere(keep)
    return dict(idx=idx.astype(np.int16), uhat=uhat[keep], dts=np.array(dts), time=o.scalars()["time"],
                ave=ave, var=var, rho_u2=ru2, max_divb=o.max_divb())


def main():
    out = os.path.dirname(os.path.abspath(__file__))
    for name, spec in CASES.items():
        r = run_case(spec)
        np.savez_compressed(os.path.join(out, f"{name}.npz"), **r)
        print(name, r["uhat"].shape, r["dts"])


if __name__ == "__main__":
    main()
