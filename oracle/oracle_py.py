"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Dict, Sequence

import numpy as np

from minirender_amd import alloc_outputs, frame_struct, outputs_struct
from minirender_amd._capi import MR_TEX_RGB8, MR_TEX_RGB32F, check, mr_draw, mr_frame, mr_outputs
from minirender_amd.scenes import FrameSpec

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"


class mro_mesh(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("npositions", C.c_int32),
        ("normals", C.c_void_p), ("nnormals", C.c_int32),
        ("uvs", C.c_void_p), ("nuvs", C.c_int32),
        ("indices", C.c_void_p), ("normal_indices", C.c_void_p),
        ("uv_indices", C.c_void_p), ("ntriangles", C.c_int32),
    ]


class mro_texture(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("format", C.c_int32),
                ("rows", C.c_int32), ("cols", C.c_int32)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.mro_render.restype = C.c_int
        L.mro_render.argtypes = [C.POINTER(mr_frame), C.POINTER(mro_mesh), C.c_int32,
                                 C.POINTER(mro_texture), C.c_int32, C.POINTER(mr_outputs)]
        L.mro_last_clipped.restype = C.c_int64
        L.mro_znear.restype = C.c_float
        L.mro_znear.argtypes = [C.POINTER(C.c_float)]
        L.mro_paint_triangle.restype = C.c_int
        L.mro_paint_triangle.argtypes = [C.POINTER(mr_frame), C.POINTER(mr_draw),
                                         C.POINTER(mro_texture), C.POINTER(C.c_float),
                                         C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
        L.mro_clip_triangle.restype = C.c_int
        L.mro_clip_triangle.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        _lib = L
    return _lib


class _Tables:
    """Mesh / texture tables indexed by the handles written into mr_frame."""

    def __init__(self):
        self.meshes, self.mesh_ids, self.textures, self.tex_ids, self.keep = [], {}, [], {}, []

    def mesh(self, m) -> int:
        if id(m) in self.mesh_ids:
            return self.mesh_ids[id(m)]
        s = mro_mesh()
        s.positions = m.positions.ctypes.data
        s.npositions = len(m.positions)
        s.normals = m.normals.ctypes.data
        s.nnormals = len(m.normals)
        s.indices = m.indices.ctypes.data
        s.normal_indices = m.nidx.ctypes.data
        if m.has_uv:
            s.uvs = m.uvs.ctypes.data
            s.nuvs = len(m.uvs)
            s.uv_indices = m.uv_indices.ctypes.data
        s.ntriangles = m.ntri
        self.keep.append(m)
        self.meshes.append(s)
        self.mesh_ids[id(m)] = len(self.meshes) - 1
        return len(self.meshes) - 1

    def texture(self, t) -> int:
        if id(t) in self.tex_ids:
            return self.tex_ids[id(t)]
        tex = np.ascontiguousarray(t.texels)
        fmt = MR_TEX_RGB8 if tex.dtype == np.uint8 else MR_TEX_RGB32F
        if fmt == MR_TEX_RGB32F:
            tex = np.ascontiguousarray(tex, dtype=np.float32)
        s = mro_texture(tex.ctypes.data, fmt, t.rows, t.cols)
        self.keep.append(tex)
        self.textures.append(s)
        self.tex_ids[id(t)] = len(self.textures) - 1
        return len(self.textures) - 1


def render(frame: FrameSpec, kinds: Sequence[str] = ("rgb8", "depth", "winner")) -> Dict[str, np.ndarray]:
    """Renderer::render() on the CPU (the oracle), same outputs as the GPU."""
    L = lib()
    tb = _Tables()
    f, keep = frame_struct(frame, tb.mesh, tb.texture)
    meshes = (mro_mesh * max(len(tb.meshes), 1))(*tb.meshes)
    texes = (mro_texture * max(len(tb.textures), 1))(*tb.textures)
    out = alloc_outputs(frame, kinds)
    o = outputs_struct(out)
    check(L.mro_render(C.byref(f), meshes, len(tb.meshes), texes, len(tb.textures), C.byref(o)),
          "mro_render")
    return out


def last_clipped() -> int:
    return int(lib().mro_last_clipped())
