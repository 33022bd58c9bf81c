"""CPU: the C-ABI library loads, exports every declared symbol, and its host
helpers (projection matrices of Renderer.cpp:40-83, znear of :324-326) match
their closed forms.  No compute call is made without a GPU."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import _capi, scenes

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"^\s*int\s+(mr_\w+)\s*\(", h.read_text(), re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = _capi.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes mirror binds every one of them
    assert names <= set(_capi.SIGNATURES), sorted(names - set(_capi.SIGNATURES))


def _mat(fn, *args):
    out = (C.c_float * 16)()
    assert fn(*args, out) == 0
    return np.array(out, np.float64).reshape(4, 4)


def test_projection_frustum_matches_glu_perspective():
    lib = _capi.load_library()
    fovy, aspect, n, f = math.radians(45), 4 / 3, 0.1, 100.0
    P = _mat(lib.mr_projection_frustum, fovy, aspect, n, f)
    c = 1 / math.tan(fovy / 2)
    ref = np.array([[c / aspect, 0, 0, 0], [0, c, 0, 0],
                    [0, 0, (f + n) / (n - f), 2 * f * n / (n - f)], [0, 0, -1, 0]])
    assert np.allclose(P, ref, rtol=2e-6, atol=1e-7)
    # python host helper agrees bit for bit with the library
    assert np.array_equal(scenes.projection_frustum(fovy, aspect, n, f).astype(np.float64), P)


def test_projection_frustum_h_and_ortho():
    lib = _capi.load_library()
    fovx, aspect, n, f = math.radians(60), 16 / 9, 0.5, 50.0
    P = _mat(lib.mr_projection_frustum_h, fovx, aspect, n, f)
    c = 1 / math.tan(fovx / 2)
    assert math.isclose(P[0, 0], c, rel_tol=2e-6)
    assert math.isclose(P[1, 1], c * aspect, rel_tol=2e-6)
    l, r, b, t = -2.0, 6.0, -1.0, 3.0
    O = _mat(lib.mr_projection_ortho, l, r, b, t, n, f)
    ref = np.array([[2 / (r - l), 0, 0, -(r + l) / (r - l)], [0, 2 / (t - b), 0, -(t + b) / (t - b)],
                    [0, 0, -2 / (f - n), -(f + n) / (f - n)], [0, 0, 0, 1]])
    assert np.allclose(O, ref, rtol=1e-6, atol=1e-7)
    assert np.array_equal(scenes.projection_ortho(l, r, b, t, n, f).astype(np.float64), O)
    Of = _mat(lib.mr_projection_ortho_fov, 4.0, 2.0, n, f)  # fov = height
    assert np.allclose(Of, _mat(lib.mr_projection_ortho, -4.0, 4.0, -2.0, 2.0, n, f))


def test_projection_cv_principal_point():
    """A pinhole with principal point at the image centre maps the optical
    axis to NDC (0, 0) and focal length fx to 2 fx / w."""
    lib = _capi.load_library()
    w, h, fx, fy = 640.0, 480.0, 500.0, 520.0
    K = np.zeros(16, np.float32)
    K[[0, 5, 2, 6, 10]] = [fx, fy, w / 2, h / 2, 1]
    P = _mat(lib.mr_projection_cv, K.ctypes.data_as(C.POINTER(C.c_float)), w, h, 0.1, 10.0)
    assert math.isclose(P[0, 0], 2 * fx / w, rel_tol=1e-6)
    assert math.isclose(P[1, 1], 2 * fy / h, rel_tol=1e-6)
    assert abs(P[0, 2]) < 1e-6 and abs(P[1, 2]) < 1e-6


@pytest.mark.parametrize("persp", [True, False])
def test_znear_recovers_near_plane(persp):
    lib = _capi.load_library()
    n, f = 0.37, 80.0
    P = (scenes.projection_frustum(1.0, 1.5, n, f) if persp
         else scenes.projection_ortho(-1, 1, -1, 1, n, f))
    zn = C.c_float()
    assert lib.mr_znear(np.ascontiguousarray(P).ctypes.data_as(C.POINTER(C.c_float)), C.byref(zn)) == 0
    # ortho: (P23 + 1) / P22 cancels; its fp32 error grows like eps * f / n
    tol = 1e-5 if persp else 4 * 2.0 ** -24 * (f + n) / n
    assert math.isclose(zn.value, -n, rel_tol=tol)


def test_no_gpu_is_an_error_not_a_fallback():
    lib = _capi.load_library()
    n = C.c_int32(-1)
    assert lib.mr_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        assert lib.mr_ctx_create(0, 0, C.byref(h)) == -6  # MR_ERR_NOGPU
        with pytest.raises(mra.MrError):
            mra.Context(0)


def test_null_arguments_rejected():
    lib = _capi.load_library()
    assert lib.mr_ctx_create(0, 0, None) == -1
    assert lib.mr_render(None, None, None) == -1
    assert lib.mr_projection_ortho(0, 1, 0, 1, 0, 1, None) == -1
