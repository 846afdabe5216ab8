"""The C ABI library: it loads, exports every symbol include/ekg.h declares,
and its host-side FP64 helpers round like the reference.  No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

import ekgsim_amd as E
from ekgsim_amd import _capi
from oracle import oracle as O
from tests.conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "ekg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ekg_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_capi.EXPORTS)


def test_version_string():
    assert b"gfx950" in E.lib().ekg_version()


def test_lead_basis_and_move_match_oracle_bitwise():
    rng = np.random.default_rng(4)
    pts = np.vstack([rng.normal(size=(30, 3)) * 80, [[0, 0, 7.0], [1e-3, 0, 0], [87.4, 104.4, -82.4]]])
    for p in pts:
        u1, v1 = E.lead_basis(p)
        u2, v2 = O.lead_basis(p)
        assert np.array_equal(u1, u2) and np.array_equal(v1, v2)
    disp = rng.uniform(-50, 50, size=(len(pts), 2))
    assert np.array_equal(E.move_leads(pts, disp), O.move_leads(pts, disp))


def test_num_steps_matches_reference_rule():
    for length, dt in [(600, 1.0), (400, 1 / 3), (600, 0.25), (37, 1.0), (5, 10.0), (0, 1.0)]:
        assert E.num_steps(length, dt) == O.num_steps(length, dt)


def test_error_reporting_without_device():
    h = ctypes.c_void_p()
    rc = E.lib().ekg_ctx_create(0, None)
    assert rc == 1 and b"NULL" in E.lib().ekg_last_error()
    rc = E.lib().ekg_ctx_create(-1, ctypes.byref(h))
    assert rc != 0 and E.lib().ekg_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "none.so"))
    monkeypatch.setattr(_capi, "_lib", None)
    with pytest.raises(ImportError):
        _capi.lib()
