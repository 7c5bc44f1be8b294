"""Argument checks that need no GPU: the Python array checks in front of the
C-ABI and the host-side checks of tmrx_problem_create (boundary-condition
names, material), which run before any device call."""
import ctypes as C

import numpy as np
import pytest

from tmr_amd import TMR, _lib, solver


def test_ptr_rejects_wrong_dtype_length_layout():
    good = np.zeros(12)
    p, dev = solver._ptr(good, 12, "x")
    assert dev == 0 and p.value == good.ctypes.data
    with pytest.raises(ValueError, match="float64"):
        solver._ptr(np.zeros(12, dtype=np.float32), 12, "x")
    with pytest.raises(ValueError, match="expected 12 entries"):
        solver._ptr(np.zeros(11), 12, "x")
    with pytest.raises(ValueError, match="contiguous"):
        solver._ptr(np.zeros(24)[::2], 12, "x")
    ro = np.zeros(12)
    ro.flags.writeable = False
    with pytest.raises(ValueError, match="read-only"):
        solver._ptr(ro, 12, "u", out=True)
    with pytest.raises(ValueError):
        solver._ptr([0.0] * 12, 12, "x")


def test_ptr_rejects_float32_tensor():
    torch = pytest.importorskip("torch")
    with pytest.raises(ValueError, match="float64"):
        solver._ptr(torch.zeros(12), 12, "f")
    with pytest.raises(ValueError, match="expected 12 entries"):
        solver._ptr(torch.zeros(13, dtype=torch.float64), 12, "f")
    p, dev = solver._ptr(torch.zeros(12, dtype=torch.float64), 12, "f")
    assert dev == 0
    with pytest.raises(ValueError, match="both"):
        solver._same_side(0, 1, "f and u")


def _create(forest, names, material=None):
    lib = _lib.lib()
    m = material or _lib.Material(1.0, 0.3, 1.0, 3.0, 1e-6, 0.3, 0)
    carr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    marr = (C.c_int * max(1, len(names)))(*([7] * len(names)))
    farr = (C.c_void_p * 1)(forest.ptr.value)
    h = C.c_void_p()
    r = lib.tmrx_problem_create(1, farr, C.byref(m), len(names), carr, marr, 0, C.byref(h))
    return r, lib.tmrx_last_error().decode()


@pytest.fixture(scope="module")
def cube():
    f = TMR.createBoxForest(1, 1, 1, face_names={"x-": "fixed", "x+": "load"})
    f.createTrees(2)
    f.createNodes()
    return f


def test_create_rejects_unknown_boundary_condition(cube):
    r, msg = _create(cube, ["fixed", "nowhere"])
    assert r < 0
    assert "'nowhere' selects no node" in msg


def test_create_rejects_bad_material(cube):
    for bad in [_lib.Material(-1.0, 0.3, 1.0, 3.0, 1e-6, 0.3, 0),
                _lib.Material(1.0, 0.5, 1.0, 3.0, 1e-6, 0.3, 0),
                _lib.Material(1.0, 0.3, 1.0, 3.0, float("nan"), 0.3, 0),
                _lib.Material(1.0, 0.3, 1.0, 3.0, 1e-6, 0.3, 7)]:
        r, msg = _create(cube, ["fixed"], bad)
        assert r < 0 and "material" in msg
