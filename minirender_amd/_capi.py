"""ctypes mirror of include/minirender_amd.h and the library loader.

The HIP library is loaded from this package directory only; a missing library
raises -- there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libminirender_amd.so"

MR_OK = 0
ERRORS = {
    -1: "MR_ERR_ARG", -2: "MR_ERR_HIP", -3: "MR_ERR_NOMEM", -4: "MR_ERR_BUDGET",
    -5: "MR_ERR_HANDLE", -6: "MR_ERR_NOGPU", -7: "MR_ERR_INDEX", -8: "MR_ERR_DEVICE",
    -9: "MR_ERR_TIMEOUT", -10: "MR_ERR_CLOSED", -11: "MR_ERR_LIMIT",
}
MR_ERR_TIMEOUT = -9
MR_DRAW_NO_NEAR_CLIP = 1
MR_TEX_RGB8 = 0
MR_TEX_RGB32F = 1
NO_WINNER = 0xFFFFFFFF


class MrError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed: {ERRORS.get(code, code)}")
        self.code = code


class mr_draw(C.Structure):
    _fields_ = [
        ("mesh", C.c_int32), ("texture", C.c_int32),
        ("modelview", C.c_float * 16), ("normalmat", C.c_float * 16),
        ("diffuse", C.c_float * 3), ("specular", C.c_float * 3),
        ("emissive", C.c_float * 3), ("shininess", C.c_float), ("flags", C.c_int32),
    ]


class mr_frame(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("projection", C.c_float * 16), ("background", C.c_float * 3),
        ("light", C.c_float * 3), ("light_is_point", C.c_int32),
        ("ambient", C.c_float), ("lighting", C.c_int32),
        ("texturing", C.c_int32), ("save_normals", C.c_int32),
        ("ndraws", C.c_int32), ("draws", C.POINTER(mr_draw)),
    ]


class mr_outputs(C.Structure):
    _fields_ = [
        ("rgb8", C.c_void_p), ("depth", C.c_void_p), ("winner", C.c_void_p),
        ("rgb", C.c_void_p), ("normals", C.c_void_p), ("points", C.c_void_p),
        ("on_device", C.c_int32),
    ]


class mr_stats(C.Structure):
    _fields_ = [
        ("triangles_in", C.c_int64), ("triangles_clipped", C.c_int64),
        ("records", C.c_int64), ("bin_entries", C.c_int64), ("tiles", C.c_int64),
        ("retries", C.c_int64), ("stage_ms", C.c_double * 8), ("frame_ms", C.c_double),
        ("launches", C.c_int64),
        ("graph_replays", C.c_int64), ("graph_updates", C.c_int64), ("graph_captures", C.c_int64),
        ("split_tiles", C.c_int64), ("raster_parts", C.c_int64),
    ]


class mr_tuning(C.Structure):
    _fields_ = [
        ("tile", C.c_int32), ("raster_threads", C.c_int32), ("bin_factor", C.c_float),
        ("frag_size", C.c_int32), ("ms", C.c_double), ("from_cache", C.c_int32),
        ("candidates", C.c_int32),
    ]


class mr_server_config(C.Structure):
    _fields_ = [
        ("max_batch", C.c_int32), ("slots", C.c_int32),
        ("hbm_budget", C.c_uint64), ("max_wait_us", C.c_int32),
    ]


class mr_request_info(C.Structure):
    _fields_ = [
        ("submit_us", C.c_double), ("launch_us", C.c_double),
        ("done_us", C.c_double), ("batch_id", C.c_int64), ("batch_size", C.c_int32),
        ("status", C.c_int32),
    ]


_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)

# name -> (restype, argtypes).  Every symbol the header declares.
SIGNATURES = {
    "mr_device_count": (C.c_int, [_IP]),
    "mr_ctx_create": (C.c_int, [_I, C.c_uint64, C.POINTER(_P)]),
    "mr_ctx_destroy": (C.c_int, [_P]),
    "mr_ctx_set_stream": (C.c_int, [_P, _P]),
    "mr_ctx_get_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "mr_ctx_set_tile": (C.c_int, [_P, _I, _I]),
    "mr_ctx_set_bin_factor": (C.c_int, [_P, _F]),
    "mr_ctx_set_raster_threads": (C.c_int, [_P, _I]),
    "mr_ctx_set_frag_size": (C.c_int, [_P, _I]),
    "mr_ctx_set_profiling": (C.c_int, [_P, _I]),
    "mr_ctx_set_graphs": (C.c_int, [_P, _I]),
    "mr_ctx_set_persistent": (C.c_int, [_P, _I]),
    "mr_ctx_set_wait_limit": (C.c_int, [_P, C.c_int64]),
    "mr_ctx_get_stats": (C.c_int, [_P, C.POINTER(mr_stats)]),
    "mr_ctx_memory": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mr_mesh_upload": (C.c_int, [_P, _P, _I, _P, _I, _P, _I, _P, _P, _P, _I, _IP]),
    "mr_mesh_upload_shared": (C.c_int, [_P, _I, _P, _P, _P, _I, _IP]),
    "mr_mesh_release": (C.c_int, [_P, _I]),
    "mr_texture_upload": (C.c_int, [_P, _P, _I, _I, _I, _IP]),
    "mr_texture_release": (C.c_int, [_P, _I]),
    "mr_render": (C.c_int, [_P, C.POINTER(mr_frame), C.POINTER(mr_outputs)]),
    "mr_render_noalloc": (C.c_int, [_P, C.POINTER(mr_frame), C.POINTER(mr_outputs)]),
    "mr_ctx_set_band_work": (C.c_int, [_P, _I]),
    "mr_render_batch": (C.c_int, [_P, C.POINTER(mr_frame), _I, C.POINTER(mr_outputs)]),
    "mr_ctx_autotune": (C.c_int, [_P, C.POINTER(mr_frame), _I, C.POINTER(mr_outputs), _I,
                                  C.c_char_p, C.POINTER(mr_tuning)]),
    "mr_server_create": (C.c_int, [_P, C.POINTER(mr_server_config), C.POINTER(_P)]),
    "mr_server_destroy": (C.c_int, [_P]),
    "mr_server_submit": (C.c_int, [_P, C.POINTER(mr_frame), C.POINTER(mr_outputs),
                                   C.POINTER(C.c_uint64)]),
    "mr_server_poll": (C.c_int, [_P, C.c_uint64, _IP]),
    "mr_server_wait": (C.c_int, [_P, C.c_uint64, _I, C.POINTER(mr_request_info)]),
    "mr_server_run_trace": (C.c_int, [_P, C.POINTER(mr_frame), _I, C.POINTER(C.c_double),
                                      C.POINTER(mr_outputs), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "mr_projection_ortho": (C.c_int, [_F, _F, _F, _F, _F, _F, _FP]),
    "mr_projection_perspective": (C.c_int, [_F, _F, _F, _F, _F, _F, _FP]),
    "mr_projection_frustum": (C.c_int, [_F, _F, _F, _F, _FP]),
    "mr_projection_frustum_h": (C.c_int, [_F, _F, _F, _F, _FP]),
    "mr_projection_ortho_fov": (C.c_int, [_F, _F, _F, _F, _FP]),
    "mr_projection_cv": (C.c_int, [_FP, _F, _F, _F, _F, _FP]),
    "mr_znear": (C.c_int, [_FP, _FP]),
}

_lib = None


def load_library() -> C.CDLL:
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m minirender_amd.build`"
            )
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str) -> None:
    if rc != MR_OK:
        raise MrError(rc, what)
