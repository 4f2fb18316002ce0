"""The C-ABI library builds, loads and exports every entry point of include/laps.h.

Runs without a GPU: only symbol lookup, parameter defaults and the
error path of laps_create are exercised (no compute call).
"""
import ctypes
import os
import re

import pytest

from laps_amd import runtime
from laps_amd.params import LapsParams, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "laps.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(laps_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding():
    assert declared_symbols() == sorted(runtime.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("liblaps_amd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(runtime.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_params_layout_and_defaults():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("liblaps_amd.so not built")
    p = LapsParams()
    runtime.lib().laps_default_params(ctypes.byref(p))
    q = default_params()
    for name, _ in LapsParams._fields_:
        assert getattr(p, name) == pytest.approx(getattr(q, name)), name
    assert ctypes.sizeof(LapsParams) == 168   # static_assert in laps_api.hip


def test_create_rejects_bad_params_without_device_work():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.skip("liblaps_amd.so not built")
    p = default_params(nx=48)
    with pytest.raises(runtime.LapsError, match="grid lengths"):
        runtime.Laps(p)
    p = default_params(nx=32, ny=32, nz=32, dealias_option=7)
    with pytest.raises(runtime.LapsError, match="dealias_option"):
        runtime.Laps(p)
