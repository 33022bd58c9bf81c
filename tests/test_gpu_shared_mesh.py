"""Meshes that share attribute arrays (loadOBJ's per-material meshes,
io.cpp:313-315) through mr_mesh_upload_shared: rendering is bit-identical to
uploading each mesh with its own copy of the arrays, matches the oracle, and
releasing meshes in any order keeps the shared arrays for the rest."""
import ctypes as C
import math

import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import scenes
from minirender_amd.scenes import Draw, FrameSpec, Material, Mesh
from oracle import oracle_py
from tests.parity import assert_parity, fuzz_frame

pytestmark = pytest.mark.gpu

KINDS = ("rgb8", "depth", "winner", "rgb", "normals", "points")


def parts_frame(parts, width=240, height=180):
    """Every part drawn twice: at the origin and moved and turned (a second
    modelview, so a second set of vertex records of the same arrays)."""
    proj = scenes.projection_frustum(math.radians(50), width / height, 0.1, 40.0)
    view = scenes.translate(0.0, -0.2, -2.6) @ scenes.rotate((1.0, 0.0, 0.0), 0.5)
    moved = scenes.translate(0.4, 0.1, -0.3) @ scenes.rotate((0.0, 1.0, 0.0), 0.7)
    draws = [Draw(p.mesh, p.material) for p in parts]
    draws += [Draw(p.mesh, p.material, moved.astype(np.float32)) for p in parts]
    return FrameSpec(width, height, proj, view=view, draws=draws, light=(1.0, 2.0, 2.5),
                     light_is_point=True, ambient=0.1, save_normals=True)


def own_copies(parts):
    """The same parts, each with attribute arrays of its own."""
    out = []
    for p in parts:
        m = p.mesh
        out.append(Draw(Mesh(m.positions.copy(), m.normals.copy(), m.indices,
                             m.normal_indices, None if m.uvs is None else m.uvs.copy(),
                             m.uv_indices), p.material))
    return out


def same(a, b, what):
    for k in KINDS:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (what, k)


def test_shared_attributes_bit_identical():
    rng = np.random.default_rng(21)
    terrain = scenes.height_field(48, 5, 0.3, uv=True)
    parts = scenes.split_materials(terrain, 4, rng)
    tex = scenes.checker_texture(64, 2, 8)
    parts[1].material.texture = tex
    copies = own_copies(parts)
    with mra.Context(0) as shared, mra.Context(0) as separate:
        fr_s, fr_c = parts_frame(parts), parts_frame(copies)
        g_s = shared.render(fr_s, KINDS)
        g_c = separate.render(fr_c, KINDS)
        same(g_s, g_c, "shared vs own arrays")
        assert_parity(g_s, oracle_py.render(fr_s, KINDS), "shared attributes", normals=True)
        assert (g_s["winner"] != mra.NO_WINNER).sum() > 5000
        # One resident copy of the arrays: the first part holds them, the
        # other three only their index arrays.
        r_s, r_c = shared.memory()[0], separate.memory()[0]
        # (device allocations are rounded up to 256 bytes)
        attr = sum((a.nbytes + 255) // 256 * 256
                   for a in (terrain.positions, terrain.normals, terrain.uvs))
        assert r_c - r_s == 3 * attr, (r_s, r_c, attr)
        # Batched with a frame of other meshes: the same images.
        other = fuzz_frame(3, 96, 64)
        batch = shared.render_batch([other, fr_s, fr_s], KINDS)
        same(batch[1], g_s, "batched shared")
        same(batch[2], g_s, "batched shared, second job")


def test_shared_release_order():
    rng = np.random.default_rng(22)
    sphere = scenes.create_sphere(0.8, 20, 40)
    parts = scenes.split_materials(sphere, 3, rng)
    copies = own_copies(parts)
    with mra.Context(0) as shared, mra.Context(0) as separate:
        base0 = shared.memory()[0]
        shared.render(parts_frame(parts), ("rgb8",))
        # Release the part that holds the arrays: the others still render.
        shared.release_mesh(parts[0].mesh)
        rest_s = shared.render(parts_frame(parts[1:]), KINDS)
        rest_c = separate.render(parts_frame(copies[1:]), KINDS)
        same(rest_s, rest_c, "after releasing the owner")
        # A new mesh takes a free slot, never the one still holding arrays.
        extra = scenes.create_sphere(0.3, 6, 12)
        shared.render(parts_frame([Draw(extra, Material())]), ("rgb8",))
        rest_s = shared.render(parts_frame(parts[1:]), KINDS)
        same(rest_s, rest_c, "after a new upload")
        # Upload the first part again (its own arrays again): all three render.
        all_s = shared.render(parts_frame(parts), KINDS)
        all_c = separate.render(parts_frame(copies), KINDS)
        same(all_s, all_c, "after re-upload")
        for p in parts:
            shared.release_mesh(p.mesh)
        shared.release_mesh(extra)
        assert shared.memory()[0] == base0


def test_shared_upload_errors():
    with mra.Context(0) as ctx:
        lib = ctx.lib
        m = scenes.create_sphere(1.0, 4, 8)
        base = ctx.mesh_handle(m)
        idx = np.ascontiguousarray(m.indices)
        h = C.c_int32(-1)
        p = idx.ctypes.data_as(C.c_void_p)
        assert lib.mr_mesh_upload_shared(ctx.h, 999, p, p, None, len(idx), C.byref(h)) == -5
        bad = idx.copy()
        bad[0, 0] = len(m.positions)
        pb = bad.ctypes.data_as(C.c_void_p)
        assert lib.mr_mesh_upload_shared(ctx.h, base, pb, pb, None, len(bad), C.byref(h)) == -7
        assert lib.mr_mesh_upload_shared(ctx.h, base, p, p, None, len(idx), C.byref(h)) == 0
        # A mesh shared from a sharer uses the same arrays.
        h2 = C.c_int32(-1)
        assert lib.mr_mesh_upload_shared(ctx.h, h.value, p, p, None, len(idx), C.byref(h2)) == 0
        assert lib.mr_mesh_release(ctx.h, base) == 0
        assert lib.mr_mesh_release(ctx.h, base) == -5
        assert lib.mr_mesh_release(ctx.h, h.value) == 0
        assert lib.mr_mesh_release(ctx.h, h2.value) == 0


def test_served_pool_requests():
    """Requests over a resident pool of split models (the service's
    workload, small): served in batches, each result equals rendering the
    request alone; the first matches the oracle."""
    kinds = ("rgb8", "depth", "winner", "rgb")
    pool = scenes.serving_pool(nmeshes=6, seed=3, tri_lo=2000, tri_hi=20000)
    reqs = scenes.serving_requests(pool, 10, seed=4)
    with mra.Context(0) as ctx:
        alone = [ctx.render(f, kinds) for f in reqs]
        with mra.Server(ctx, max_batch=4, slots=2) as srv:
            tickets = [srv.submit(f, kinds) for f in reqs]
            results = [srv.wait(t) for t in tickets]
        for i, (out, info) in enumerate(results):
            assert info.status == 0
            for k in kinds:
                assert np.array_equal(out[k].view(np.uint8), alone[i][k].view(np.uint8)), (i, k)
        assert_parity(alone[0], oracle_py.render(reqs[0], kinds), "pool request 0")
        assert (alone[0]["winner"] != mra.NO_WINNER).sum() > 1000
