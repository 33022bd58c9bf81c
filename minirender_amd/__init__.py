"""minirender_amd -- minirender's rasterization path on MI355X (gfx950).

Python host side of the C-ABI in include/minirender_amd.h.  `Context` keeps
meshes and textures resident on the GPU and renders FrameSpecs
(minirender_amd.scenes) with the HIP pipeline.  The HIP library must be built
(`python -m minirender_amd.build`); nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import hashlib
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import scenes
from ._capi import (MR_DRAW_NO_NEAR_CLIP, MR_ERR_TIMEOUT, MR_TEX_RGB8, MR_TEX_RGB32F,
                    NO_WINNER, MrError, check, load_library, mr_draw, mr_frame, mr_outputs,
                    mr_stats)
from .scenes import Draw, FrameSpec, Material, Mesh, Texture

__all__ = ["Context", "FrameSpec", "Draw", "Material", "Mesh", "Texture", "MrError",
           "NO_WINNER", "scenes", "frame_struct", "load_library", "OUTPUT_KINDS"]

OUTPUT_KINDS = ("rgb8", "depth", "winner", "rgb", "normals", "points")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _attr_key(m: Mesh):
    """Identity of a mesh's attribute arrays: the memory they live in."""
    def one(a):
        return None if a is None else (a.__array_interface__["data"][0], a.shape, a.dtype.str)
    return one(m.positions), one(m.normals), one(m.uvs if m.has_uv else None)


def _attr_digest(m: Mesh) -> bytes:
    """Digest of a mesh's attribute contents (at the time of the call)."""
    h = hashlib.blake2b(digest_size=16)
    h.update(m.positions)
    h.update(m.normals)
    if m.has_uv:
        h.update(m.uvs)
    return h.digest()


def _f(seq, n):
    return (C.c_float * n)(*[float(v) for v in np.asarray(seq, np.float32).ravel()[:n]])


def frame_struct(frame: FrameSpec, mesh_handle, tex_handle):
    """mr_frame for `frame`; draws map meshes/textures through the handle
    callbacks.  Returns (mr_frame, keepalive)."""
    n = len(frame.draws)
    draws = (mr_draw * max(n, 1))()
    for i, d in enumerate(frame.draws):
        mv = scenes.modelview(frame, d)
        nm = scenes.normal_matrix(mv)
        s = draws[i]
        s.mesh = mesh_handle(d.mesh)
        s.texture = tex_handle(d.material.texture) if d.material.texture is not None else -1
        s.modelview = _f(mv, 16)
        s.normalmat = _f(nm, 16)
        s.diffuse = _f(d.material.diffuse, 3)
        s.specular = _f(d.material.specular, 3)
        s.emissive = _f(d.material.emissive, 3)
        s.shininess = float(np.float32(d.material.shininess))
        s.flags = MR_DRAW_NO_NEAR_CLIP if d.no_clip else 0
    f = mr_frame()
    f.width, f.height = int(frame.width), int(frame.height)
    f.projection = _f(frame.projection, 16)
    f.background = _f(frame.background, 3)
    f.light = _f(scenes.eye_light(frame), 3)
    f.light_is_point = int(bool(frame.light_is_point))
    f.ambient = float(np.float32(frame.ambient))
    f.lighting = int(bool(frame.lighting))
    f.texturing = int(bool(frame.texturing))
    f.save_normals = int(bool(frame.save_normals))
    f.ndraws = n
    f.draws = C.cast(draws, C.POINTER(mr_draw))
    return f, draws


def alloc_outputs(frame: FrameSpec, kinds: Iterable[str]) -> Dict[str, np.ndarray]:
    h, w = frame.height, frame.width
    shapes = {"rgb8": ((h, w, 3), np.uint8), "depth": ((h, w), np.float32),
              "winner": ((h, w), np.uint32), "rgb": ((h, w, 3), np.float32),
              "normals": ((h, w, 3), np.float32), "points": ((h, w, 3), np.float32)}
    return {k: np.empty(*shapes[k]) for k in kinds}


def outputs_struct(out: Dict[str, np.ndarray]) -> mr_outputs:
    o = mr_outputs()
    for k in OUTPUT_KINDS:
        setattr(o, k, _ptr(out.get(k)))
    o.on_device = 0
    return o


class Context:
    """A device context with a resident mesh/texture pool."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib.mr_ctx_create(device, hbm_budget, C.byref(h)), "mr_ctx_create")
        self.h = h
        self._meshes: Dict[int, tuple] = {}
        self._textures: Dict[int, tuple] = {}
        # attribute-array identity -> (handle of a resident mesh holding them, digest)
        self._attrs: Dict[tuple, tuple] = {}

    def close(self) -> None:
        if self.h:
            self.lib.mr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- configuration -------------------------------------------------------
    def set_tile(self, tw: int, th: int) -> None:
        check(self.lib.mr_ctx_set_tile(self.h, tw, th), "mr_ctx_set_tile")

    def set_raster_threads(self, n: int) -> None:
        check(self.lib.mr_ctx_set_raster_threads(self.h, n), "mr_ctx_set_raster_threads")

    def set_frag_size(self, pairs: int) -> None:
        check(self.lib.mr_ctx_set_frag_size(self.h, pairs), "mr_ctx_set_frag_size")

    def set_band_work(self, units: int) -> None:
        check(self.lib.mr_ctx_set_band_work(self.h, units), "mr_ctx_set_band_work")

    def set_bin_factor(self, f: float) -> None:
        check(self.lib.mr_ctx_set_bin_factor(self.h, f), "mr_ctx_set_bin_factor")

    def set_profiling(self, on: bool) -> None:
        check(self.lib.mr_ctx_set_profiling(self.h, int(on)), "mr_ctx_set_profiling")

    def set_graphs(self, on: bool) -> None:
        check(self.lib.mr_ctx_set_graphs(self.h, int(on)), "mr_ctx_set_graphs")

    def set_persistent(self, on: bool) -> None:
        """Synchronous batches as one launch of the persistent kernel."""
        check(self.lib.mr_ctx_set_persistent(self.h, int(on)), "mr_ctx_set_persistent")

    def set_wait_limit(self, us: int) -> None:
        """Microseconds one wait of the persistent kernel may take (negative:
        default; 0: every wait gives up, the batch fails with MR_ERR_DEVICE)."""
        check(self.lib.mr_ctx_set_wait_limit(self.h, int(us)), "mr_ctx_set_wait_limit")

    def set_stream(self, stream: int) -> None:
        check(self.lib.mr_ctx_set_stream(self.h, C.c_void_p(stream)), "mr_ctx_set_stream")

    def stats(self) -> mr_stats:
        s = mr_stats()
        check(self.lib.mr_ctx_get_stats(self.h, C.byref(s)), "mr_ctx_get_stats")
        return s

    def memory(self):
        r, s = C.c_uint64(), C.c_uint64()
        check(self.lib.mr_ctx_memory(self.h, C.byref(r), C.byref(s)), "mr_ctx_memory")
        return r.value, s.value

    # -- resident pool -------------------------------------------------------
    def mesh_handle(self, m: Mesh) -> int:
        ent = self._meshes.get(id(m))
        if ent is not None and ent[1] is m:
            return ent[0]
        h = C.c_int32()
        uvs = m.uvs if m.has_uv else None
        tix = m.uv_indices if m.has_uv else None
        nix = m.nidx
        # Meshes viewing the same attribute arrays (the per-material parts of
        # one model) share one resident copy of them, and with it one vertex
        # transform per frame; the digest guards against arrays changed in
        # place since the first upload.
        key, digest = _attr_key(m), _attr_digest(m)
        base = self._attrs.get(key)
        if base is not None and base[1] == digest:
            check(self.lib.mr_mesh_upload_shared(
                self.h, base[0], _ptr(m.indices), _ptr(nix), _ptr(tix), m.ntri, C.byref(h)),
                "mr_mesh_upload_shared")
        else:
            check(self.lib.mr_mesh_upload(
                self.h, _ptr(m.positions), len(m.positions), _ptr(m.normals), len(m.normals),
                _ptr(uvs), 0 if uvs is None else len(uvs), _ptr(m.indices), _ptr(nix),
                _ptr(tix), m.ntri, C.byref(h)), "mr_mesh_upload")
            self._attrs[key] = (h.value, digest)
        self._meshes[id(m)] = (h.value, m)
        return h.value

    def release_mesh(self, m: Mesh) -> None:
        ent = self._meshes.pop(id(m), None)
        if ent is not None:
            self._attrs = {k: v for k, v in self._attrs.items() if v[0] != ent[0]}
            check(self.lib.mr_mesh_release(self.h, ent[0]), "mr_mesh_release")

    def texture_handle(self, t: Texture) -> int:
        ent = self._textures.get(id(t))
        if ent is not None and ent[1] is t:
            return ent[0]
        tex = np.ascontiguousarray(t.texels)
        fmt = MR_TEX_RGB8 if tex.dtype == np.uint8 else MR_TEX_RGB32F
        if fmt == MR_TEX_RGB32F:
            tex = np.ascontiguousarray(tex, dtype=np.float32)
        h = C.c_int32()
        check(self.lib.mr_texture_upload(self.h, _ptr(tex), fmt, t.rows, t.cols, C.byref(h)),
              "mr_texture_upload")
        self._textures[id(t)] = (h.value, t)
        return h.value

    def frame(self, frame: FrameSpec):
        return frame_struct(frame, self.mesh_handle, self.texture_handle)

    # -- rendering -----------------------------------------------------------
    def render(self, frame: FrameSpec, kinds: Sequence[str] = ("rgb8", "depth", "winner")):
        f, keep = self.frame(frame)
        out = alloc_outputs(frame, kinds)
        o = outputs_struct(out)
        check(self.lib.mr_render(self.h, C.byref(f), C.byref(o)), "mr_render")
        return out

    def render_device(self, frame: FrameSpec, ptrs: Dict[str, int]) -> None:
        """Render into device buffers (e.g. torch tensors' data_ptr())."""
        f, keep = self.frame(frame)
        o = mr_outputs()
        for k in OUTPUT_KINDS:
            setattr(o, k, C.c_void_p(ptrs[k]) if ptrs.get(k) else None)
        o.on_device = 1
        check(self.lib.mr_render(self.h, C.byref(f), C.byref(o)), "mr_render")

    def render_batch(self, frames: List[FrameSpec],
                     kinds: Sequence[str] = ("rgb8", "depth", "winner")):
        structs, keep, outs = [], [], []
        for fr in frames:
            f, k = self.frame(fr)
            structs.append(f)
            keep.append(k)
            outs.append(alloc_outputs(fr, kinds))
        farr = (mr_frame * len(frames))(*structs)
        oarr = (mr_outputs * len(frames))(*[outputs_struct(o) for o in outs])
        check(self.lib.mr_render_batch(self.h, farr, len(frames), oarr), "mr_render_batch")
        return outs


    def autotune(self, frames: List[FrameSpec], outs=None, reps: int = 3,
                 cache_path: Optional[str] = None) -> dict:
        """Time the launch configurations on `frames` (outputs: `outs`, a list
        of mr_outputs structs, or host RGB8 + depth buffers) and apply the
        fastest; see mr_ctx_autotune."""
        from ._capi import mr_tuning
        structs, keep = [], []
        for fr in frames:
            f, k = self.frame(fr)
            structs.append(f)
            keep.append(k)
        host = None
        if outs is None:
            host = [alloc_outputs(fr, ("rgb8", "depth")) for fr in frames]
            outs = [outputs_struct(o) for o in host]
        farr = (mr_frame * len(frames))(*structs)
        oarr = (mr_outputs * len(frames))(*outs)
        t = mr_tuning()
        check(self.lib.mr_ctx_autotune(self.h, farr, len(frames), oarr, reps,
                                       cache_path.encode() if cache_path else None, C.byref(t)),
              "mr_ctx_autotune")
        return {"tile": t.tile, "raster_threads": t.raster_threads,
                "bin_factor": round(t.bin_factor, 3), "frag_size": t.frag_size,
                "ms": t.ms, "from_cache": bool(t.from_cache), "candidates": t.candidates}


class Server:
    """Render service with continuous batching over a Context's resident
    meshes (C-ABI mr_server_*).  max_batch=1, slots=1 is the plain
    one-request-at-a-time GPU path."""

    def __init__(self, ctx: Context, max_batch: int = 64, slots: int = 2,
                 hbm_budget: int = 0, max_wait_us: int = 0):
        from ._capi import mr_server_config
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = mr_server_config(max_batch, slots, hbm_budget, max_wait_us)
        h = C.c_void_p()
        check(self.lib.mr_server_create(ctx.h, C.byref(cfg), C.byref(h)), "mr_server_create")
        self.h = h
        self._pending = {}

    def close(self) -> None:
        if self.h:
            self.lib.mr_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, frame: FrameSpec, kinds: Sequence[str] = ("rgb8", "depth", "winner")) -> int:
        f, keep = self.ctx.frame(frame)
        out = alloc_outputs(frame, kinds)
        o = outputs_struct(out)
        t = C.c_uint64()
        check(self.lib.mr_server_submit(self.h, C.byref(f), C.byref(o), C.byref(t)),
              "mr_server_submit")
        self._pending[t.value] = (out, keep, o)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        from ._capi import mr_request_info
        info = mr_request_info()
        rc = self.lib.mr_server_wait(self.h, ticket, timeout_ms, C.byref(info))
        if rc == MR_ERR_TIMEOUT:
            # Still live in the server: its worker will write into these
            # output arrays later, so they stay referenced until a wait
            # that completes.
            check(rc, "mr_server_wait")
        out = self._pending.pop(ticket)[0]
        check(rc, "mr_server_wait")
        return out, info

    def run_trace(self, structs, arrivals: Sequence[float], outs_structs):
        """Replay prebuilt mr_frame / mr_outputs arrays at the given arrival
        times (seconds); returns (latency_ms array, makespan_s)."""
        n = len(arrivals)
        arr = (C.c_double * n)(*arrivals)
        lat = (C.c_double * n)()
        mk = C.c_double()
        check(self.lib.mr_server_run_trace(self.h, structs, n, arr, outs_structs, lat,
                                           C.byref(mk)), "mr_server_run_trace")
        return np.array(lat[:], dtype=np.float64), mk.value
