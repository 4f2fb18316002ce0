"""The oracle's closed-form pins catch plausible mistakes.

Each case plants one mistake in a scratch copy of oracle/laps_oracle.c (on the
CPU, in a temporary directory), builds it, and runs the pins of
tests/test_oracle_pins.py aimed at that term against the mutated library
(LAPS_ORACLE_LIB): they must fail. The unmodified copy must pass the same
selection, so a failure means the pin saw the planted mistake and not a broken
harness. The mistakes are those a restatement of mhdrhs.f90 / mhd.f90 could
plausibly make:

  energy flux      (e + p) for (e + ptot)                 mhdrhs.f90:87-89
  energy flux      +u.B for -u.B in the (u.B) B term      mhdrhs.f90:87-89
  momentum flux    rho u_i rho u_j for rho u_i u_j        mhdrhs.f90:67-79
  Hall EMF         d_i without 1/rho                      mhdrhs.f90:106-119
  AEB derivative   k_y without R0/R                       mhdrhs.f90:191-204
  corotating box   k_x' = k_x cos - k_y sin (sign)        mhdrhs.f90:198-204
  dt hysteresis    dt reset at every step                 mhd.f90:422-428
"""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "laps_oracle.c")


def _momentum_rho_rho(s):
    # F(3..11) = m_i u_j - B_i B_j (+ ptot): u_j -> m_j
    def fix(m):
        return re.sub(r"\* u([xyz])", r"* m\1", m.group(0))
    return re.sub(r"F\[(?:3|4|5|6|7|8|9|10|11) \* N \+ i\] = m[xyz] \* u[xyz]", fix, s)


MUTATIONS = {
    "energy_e_plus_p": (lambda s: s.replace("(e + ptot) * u", "(e + p) * u"), 3, "energy_flux"),
    "energy_udotb_sign": (lambda s: s.replace("- udotb * B", "+ udotb * B"), 3, "energy_flux"),
    "momentum_rho_rho": (_momentum_rho_rho, 9, "momentum_flux_nonuniform_density"),
    "hall_no_inv_rho": (lambda s: s.replace("di / rho * (J", "di * (J"), 3, "hall_density_factor"),
    "aeb_ky_unscaled": (lambda s: s.replace("*ky = cx_divr(cx_scale(z, o->p.radius0), o->radius);", "*ky = z;"), 1,
                        "aeb_density_mode_along_y"),
    "corotating_sign": (lambda s: s.replace("cx_add(cx_scale(ikx, o->cos_c), cx_scale(iky, o->sin_c))",
                                            "cx_add(cx_scale(ikx, o->cos_c), cx_scale(iky, -o->sin_c))"), 1,
                        "aeb_corotating_mode_mixing"),
    "no_hysteresis": (lambda s: s.replace("if (o->dt < 0.98 * dtmin || o->dt > 1.02 * dtmin) o->dt = dtmin;",
                                          "o->dt = dtmin;")
                      .replace("if (dt < 0.98 * dtmin || dt > 1.02 * dtmin) dt = dtmin;", "dt = dtmin;"), 2,
                      "dt_hysteresis"),
}


def _changed_lines(a, b):
    return sum(x != y for x, y in zip(a.splitlines(), b.splitlines()))


def _build(tmp, text):
    src = os.path.join(tmp, "laps_oracle.c")
    with open(src, "w") as f:
        f.write(text)
    lib = os.path.join(tmp, "liblaps_oracle_mut.so")
    # the copy includes oracle/laps_oracle.h (and through it include/laps.h) from the tree
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-std=c11", "-D_GNU_SOURCE", "-I", os.path.join(ROOT, "oracle"),
                    "-shared", "-o", lib, src, "-lm"], check=True, capture_output=True)
    return lib


def _pins(lib, select):
    env = dict(os.environ, LAPS_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-k", select],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode, r.stdout[-2000:]


@pytest.fixture(scope="module")
def source():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    with open(ORACLE_SRC) as f:
        return f.read()


def test_unmutated_copy_passes_every_selection(source, tmp_path):
    lib = _build(str(tmp_path), source)
    sel = " or ".join(sorted({m[2] for m in MUTATIONS.values()}))
    rc, out = _pins(lib, sel)
    assert rc == 0, out


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_fails_a_pin(source, tmp_path, name):
    mutate, nlines, select = MUTATIONS[name]
    text = mutate(source)
    assert _changed_lines(source, text) == nlines, "the mutation no longer matches the oracle source"
    lib = _build(str(tmp_path), text)
    rc, out = _pins(lib, select)
    assert rc == 1, f"pins '{select}' did not catch {name}:\n{out}"
