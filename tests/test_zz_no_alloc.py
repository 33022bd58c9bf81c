"""GPU: mr_render_noalloc, the entry point that allocates nothing, against
mr_render on the same frames, through the C-ABI.

Two calls in a row share the static scratch array: a larger frame (near
clipping, several draws, a texture) and then a smaller one; each result must
equal mr_render's bit for bit (RGB8, depth, winner, float colour, normals,
range image).  A frame beyond the fixed scratch is refused with
MR_ERR_LIMIT, host outputs with MR_ERR_ARG, and a call after a refusal still
renders correctly.  Device outputs come from the HIP runtime the library
itself links (no second runtime in the process).
"""
import ctypes as C

import numpy as np
import pytest

from minirender_amd import scenes
from minirender_amd._capi import mr_outputs
from tests.parity import fuzz_frame

pytestmark = pytest.mark.gpu

KINDS = ("rgb8", "depth", "winner", "rgb", "normals", "points")
BYTES = {"rgb8": 3, "depth": 4, "winner": 4, "rgb": 12, "normals": 12, "points": 12}
DTYPE = {"rgb8": np.uint8, "depth": np.float32, "winner": np.uint32, "rgb": np.float32,
         "normals": np.float32, "points": np.float32}
MR_ERR_ARG, MR_ERR_LIMIT = -1, -11


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library loaded."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        assert self.lib.hipMemset(p, 0, n) == 0
        return p.value

    def to_host(self, ptr, arr):
        assert self.lib.hipMemcpy(arr.ctypes.data, ptr, arr.nbytes, 2) == 0  # device to host
        return arr

    def free(self, ptr):
        self.lib.hipFree(ptr)


def _device_outputs(frame, hip):
    n = frame.width * frame.height
    ptrs = {k: hip.malloc(n * BYTES[k]) for k in KINDS}
    o = mr_outputs()
    for k, p in ptrs.items():
        setattr(o, k, p)
    o.on_device = 1
    return o, ptrs


def _noalloc(ctx, frame, hip):
    f, keep = ctx.frame(frame)
    o, ptrs = _device_outputs(frame, hip)
    rc = ctx.lib.mr_render_noalloc(ctx.h, C.byref(f), C.byref(o))
    h, w = frame.height, frame.width
    out = {}
    for k, p in ptrs.items():
        shape = (h, w) if BYTES[k] == 4 else (h, w, 3)
        out[k] = hip.to_host(p, np.empty(shape, DTYPE[k]))
        hip.free(p)
    return rc, out


def _same(frame, got, ref):
    for k in KINDS:
        if k == "normals" and not frame.save_normals:
            continue
        g = got[k].view(np.uint8).reshape(-1)
        r = np.ascontiguousarray(ref[k]).view(np.uint8).reshape(-1)
        assert np.array_equal(g, r), f"{k} differs from mr_render"


def test_noalloc_matches_render(gpu_ctx):
    hip = _Hip()
    ctx = gpu_ctx
    large = scenes.config_textured(n=200, width=640, height=360, ntris=400, tex_size=256)
    large.save_normals = True
    small = fuzz_frame(77, 97, 61)
    assert large.ntri > 50_000 and small.ntri < large.ntri
    for fr in (large, small):
        ref = ctx.render(fr, KINDS)
        rc, got = _noalloc(ctx, fr, hip)
        assert rc == 0, f"mr_render_noalloc returned {rc}"
        _same(fr, got, ref)
        assert (got["winner"] != 0xFFFFFFFF).sum() > 100
    # beyond the fixed scratch: refused, nothing allocated
    big = scenes.config_terrain(n=1024, width=1920, height=1080, shapes=False)
    f, keep = ctx.frame(big)
    o, ptrs = _device_outputs(big, hip)
    assert ctx.lib.mr_render_noalloc(ctx.h, C.byref(f), C.byref(o)) == MR_ERR_LIMIT
    for p in ptrs.values():
        hip.free(p)
    # host outputs: refused
    f, keep = ctx.frame(small)
    host = np.zeros((small.height, small.width, 3), np.uint8)
    oh = mr_outputs()
    oh.rgb8 = host.ctypes.data
    oh.on_device = 0
    assert ctx.lib.mr_render_noalloc(ctx.h, C.byref(f), C.byref(oh)) == MR_ERR_ARG
    # and a call after the refusals is still right
    ref = ctx.render(small, KINDS)
    rc, got = _noalloc(ctx, small, hip)
    assert rc == 0
    _same(small, got, ref)
