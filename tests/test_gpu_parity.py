"""GPU parity: the HIP rasterization path against the CPU oracle.

Every test renders the same seeded frame through mr_render (C-ABI, HIP
kernels on the MI355X) and through the oracle's restatement of
Renderer::render, and requires winner and depth bit-equal and RGB8 within
1 LSB.
"""
import math

import numpy as np
import pytest

from minirender_amd import scenes
from minirender_amd.scenes import Draw, FrameSpec, Material
from oracle import oracle_py
from tests.parity import assert_parity, fuzz_frame, stack_frame

pytestmark = pytest.mark.gpu

ALL = ("rgb8", "depth", "winner", "rgb", "normals", "points")


def check(ctx, frame, what, kinds=ALL):
    g = ctx.render(frame, kinds)
    o = oracle_py.render(frame, kinds)
    assert_parity(g, o, what, normals=frame.save_normals)
    return g, o


def test_dyadic_golden_images(gpu_ctx):
    """GPU against the exact-arithmetic golden images directly."""
    from tests.test_oracle_kat import GOLDEN, ortho16, pixel_mesh
    import json
    data = json.loads(GOLDEN.read_text())
    for k, sc in enumerate(data["scenes"]):
        fr = FrameSpec(16, 16, ortho16(), draws=[Draw(pixel_mesh(sc["triangles"]))], lighting=False)
        g = gpu_ctx.render(fr, ("depth", "winner"))
        win = np.array(sc["winner"], np.int64)
        exp_w = np.where(win < 0, 0xFFFFFFFF, win).astype(np.uint32)
        exp_d = np.array([[1e11 if d is None else d for d in row] for row in sc["depth"]], np.float32)
        assert np.array_equal(g["winner"], exp_w), f"scene {k}: winner"
        assert np.array_equal(g["depth"].view(np.uint32), exp_d.view(np.uint32)), f"scene {k}: depth"


def test_sphere_config1(gpu_ctx):
    g, o = check(gpu_ctx, scenes.config_sphere(), "sphere 640x480")
    assert (g["winner"] != 0xFFFFFFFF).sum() > 100000


@pytest.mark.parametrize("seed", range(40))
def test_fuzz_scenes(gpu_ctx, seed):
    check(gpu_ctx, fuzz_frame(seed), f"fuzz seed {seed}")


@pytest.mark.parametrize("tile,threads", [(8, 64), (8, 128), (16, 64), (16, 128), (16, 256),
                                          (16, 512), (32, 64), (32, 128), (32, 256), (32, 512)])
def test_tile_sizes(gpu_ctx, tile, threads):
    gpu_ctx.set_tile(tile, tile)
    gpu_ctx.set_raster_threads(threads)
    try:
        what = f"tile {tile}/{threads}"
        for seed in (100, 101, 102):
            check(gpu_ctx, fuzz_frame(seed, 133, 71), f"{what} seed {seed}")
        # split tiles (more pairs than one raster work item) with this shape
        check(gpu_ctx, stack_frame(1500, 80, 56, seed=tile + threads), f"{what} split")
    finally:
        gpu_ctx.set_tile(16, 16)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_edge_checkpoints_large_triangles(gpu_ctx, tile):
    """Shading walks from the raster's edge checkpoints: large and clipped
    triangles (long rows, many tile columns per triangle), textures, tiles of
    several raster rounds, split tiles, and a frame whose checkpoints do not
    all fit (those rows are walked from the bbox)."""
    gpu_ctx.set_tile(tile, tile)
    try:
        rng_tris = scenes.random_triangles(40, 21, lo=0.2, hi=1.5, center_lo=(-1, -1, -3),
                                           center_hi=(1, 1, -2))
        fr = FrameSpec(640, 360, scenes.projection_frustum(math.radians(60), 16 / 9, 0.1, 50.0),
                       draws=[Draw(rng_tris, Material(shininess=40.0))], light=(0.3, 0.5, 1.0))
        check(gpu_ctx, fr, f"checkpoints large triangles t{tile}")
        check(gpu_ctx, scenes.config_textured(n=80, width=320, height=180, ntris=300, tex_size=256),
              f"checkpoints textured t{tile}")
        check(gpu_ctx, scenes.config_terrain(n=96, width=320, height=180, shapes=True),
              f"checkpoints terrain t{tile}")
        # tiles of several rounds in one workgroup: later rounds take pixels
        # over from earlier ones
        gpu_ctx.set_raster_threads(64)
        check(gpu_ctx, stack_frame(300, 64, 48, seed=4), f"checkpoints rounds t{tile}")
        # crowded tiles split into parts; at 800 screen-covering layers the
        # checkpoint buffer (2 per pixel + 64 Ki) overflows
        gpu_ctx.set_frag_size(64)
        check(gpu_ctx, stack_frame(800, 64, 48, seed=5), f"checkpoints split t{tile}")
    finally:
        gpu_ctx.set_frag_size(1024)
        gpu_ctx.set_tile(16, 16)


def test_terrain_clipped(gpu_ctx):
    fr = scenes.config_terrain(n=96, width=320, height=180, shapes=True)
    g, o = check(gpu_ctx, fr, "terrain 96")
    assert gpu_ctx.stats().triangles_clipped == oracle_py.last_clipped() > 0


def test_textured_terrain(gpu_ctx):
    fr = scenes.config_textured(n=80, width=320, height=180, ntris=300, tex_size=256)
    check(gpu_ctx, fr, "textured terrain")


def test_fullscreen_triangles(gpu_ctx):
    """Triangles much larger than the screen: long edge walks (exact
    fast-forward of e += n.x) and many tiles per triangle."""
    rng = np.random.default_rng(3)
    tris = scenes.random_triangles(12, 9, lo=2.0, hi=6.0, center_lo=(-1, -1, -3), center_hi=(1, 1, -2))
    fr = FrameSpec(1920, 1080, scenes.projection_frustum(math.radians(60), 16 / 9, 0.1, 50.0),
                   draws=[Draw(tris, Material(shininess=20.0))], light=(0.3, 0.5, 1.0))
    check(gpu_ctx, fr, "fullscreen triangles", ("rgb8", "depth", "winner"))


def test_separate_index_arrays(gpu_ctx):
    """Normal and texcoord index arrays of their own (face normals, permuted
    texcoords), as OBJ files give them: shading reads them instead of the
    position indices."""
    from minirender_amd.scenes import Mesh
    rng = np.random.default_rng(17)
    sph = scenes.create_sphere(1.0, 12, 24)
    pos, idx = sph.positions, sph.indices
    a, b, c = pos[idx[:, 0]], pos[idx[:, 1]], pos[idx[:, 2]]
    fn = np.cross(b - a, c - a).astype(np.float32)
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-12).astype(np.float32)
    nidx = np.repeat(np.arange(len(idx), dtype=np.int32)[:, None], 3, axis=1)
    uvs = rng.uniform(0.0, 1.0, (97, 2)).astype(np.float32)
    uidx = rng.integers(0, len(uvs), idx.shape).astype(np.int32)
    mesh = Mesh(pos, fn, idx, normal_indices=nidx, uvs=uvs, uv_indices=uidx)
    mat = Material(shininess=20.0, texture=scenes.checker_texture(64, 3, 8))
    fr = FrameSpec(200, 150, scenes.projection_frustum(math.radians(45), 4 / 3, 0.1, 50.0),
                   view=scenes.translate(0.0, 0.0, -3.0), draws=[Draw(mesh, mat)],
                   light=(1.0, 2.0, 3.0), light_is_point=True)
    g, o = check(gpu_ctx, fr, "separate index arrays")
    assert (g["winner"] != 0xFFFFFFFF).sum() > 5000


def test_bin_overflow_retry(gpu_ctx):
    fr = fuzz_frame(7, 640, 480)
    gpu_ctx.set_bin_factor(1.0)
    try:
        check(gpu_ctx, fr, "bin overflow")
    finally:
        gpu_ctx.set_bin_factor(3.0)


def test_batch_equals_single(gpu_ctx):
    frames = [fuzz_frame(200 + i, 64 + 13 * i, 48 + 7 * i) for i in range(6)]
    batch = gpu_ctx.render_batch(frames, ALL)
    for i, fr in enumerate(frames):
        single = gpu_ctx.render(fr, ALL)
        for k in ALL:
            if k == "normals" and not fr.save_normals:
                continue
            assert np.array_equal(batch[i][k].view(np.uint8), single[k].view(np.uint8)), (i, k)
        assert_parity(batch[i], oracle_py.render(fr, ALL), f"batch frame {i}",
                      normals=fr.save_normals)


def test_large_batch_equals_single(gpu_ctx):
    """150 jobs in one pass (job lookups past one wavefront of jobs, many
    draws, mixed sizes and tiles): every job bit-identical to rendering it
    alone; a sample of them against the oracle."""
    frames = [fuzz_frame(400 + i, 24 + (i * 37) % 120, 16 + (i * 23) % 90) for i in range(150)]
    kinds = ("rgb8", "depth", "winner")
    gpu_ctx.set_tile(8, 8)
    try:
        batch = gpu_ctx.render_batch(frames, kinds)
        for i, fr in enumerate(frames):
            single = gpu_ctx.render(fr, kinds)
            for k in kinds:
                assert np.array_equal(batch[i][k].view(np.uint8), single[k].view(np.uint8)), (i, k)
        for i in (0, 63, 64, 65, 149):
            assert_parity(batch[i], oracle_py.render(frames[i], kinds), f"large batch frame {i}")
    finally:
        gpu_ctx.set_tile(16, 16)


def test_empty_and_degenerate(gpu_ctx):
    # No draws: background only.
    fr = FrameSpec(40, 30, scenes.projection_frustum(1.0, 4 / 3, 0.1, 10), background=(0.2, 0.4, 0.6))
    check(gpu_ctx, fr, "empty")
    # Degenerate triangles (repeated vertices, collinear) and a zero-area mesh.
    pos = np.array([[0, 0, -2], [0, 0, -2], [1, 0, -2], [0, 1, -2], [2, 2, -2], [1, 1, -2]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (6, 1))
    idx = np.array([[0, 1, 2], [0, 5, 4], [0, 2, 3], [3, 2, 0]], np.int32)
    m = scenes.Mesh(pos, nrm, idx)
    fr.draws = [Draw(m, Material())]
    check(gpu_ctx, fr, "degenerate")


def test_empty_frame_after_wide_triangles(gpu_ctx):
    """A frame without triangles right after one whose wide triangles get
    edge checkpoints, on the same context, then the wide frame again: the
    empty frame is background (its row pre-pass finds no row tasks of the
    frame before) and the repeat equals the first render bit for bit."""
    kinds = ("rgb8", "depth", "winner")
    tris = scenes.random_triangles(12, 9, lo=2.0, hi=6.0, center_lo=(-1, -1, -3), center_hi=(1, 1, -2))
    wide = FrameSpec(640, 360, scenes.projection_frustum(math.radians(60), 16 / 9, 0.1, 50.0),
                     draws=[Draw(tris, Material(shininess=20.0))], light=(0.3, 0.5, 1.0))
    g1, _ = check(gpu_ctx, wide, "wide", kinds)
    empty = FrameSpec(640, 360, wide.projection, background=(0.2, 0.4, 0.6))
    check(gpu_ctx, empty, "empty after wide", kinds)
    g2 = gpu_ctx.render(wide, kinds)
    for k in kinds:
        assert np.array_equal(g1[k], g2[k]), k


def test_shared_edges_first_wins(gpu_ctx):
    """A pixel-aligned quad split along its diagonal, drawn twice with
    different materials: ties go to the first-drawn triangle (strict z test)
    and to sub-triangle order along shared edges."""
    proj = scenes.projection_ortho(0, 16, 0, 16, -1, 1)
    pos = np.array([[2, 2, 0], [14, 2, 0], [14, 14, 0], [2, 14, 0]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    m = scenes.Mesh(pos, nrm, idx)
    fr = FrameSpec(16, 16, proj, draws=[Draw(m, Material(diffuse=(1, 0, 0))),
                                        Draw(m, Material(diffuse=(0, 1, 0)))], lighting=False)
    g, o = check(gpu_ctx, fr, "shared edges")
    assert (g["winner"][g["winner"] != 0xFFFFFFFF] < 4).all()


def test_graph_replay_matches_eager(gpu_ctx):
    """Steady-state HIP graphs: frames of one shape with different cameras
    replay the captured graph; every replay equals the eager render, with
    eager work (other shapes, synchronisations) in between."""
    import minirender_amd as mra
    base = scenes.config_sphere(160, 120)
    views = []
    for k in range(4):
        fr = scenes.config_sphere(160, 120)
        fr.view = scenes.translate(0.1 * k, -0.05 * k, -3.0 + 0.2 * k) @ scenes.rotate((0, 1, 0), 0.3 * k)
        fr.light = (2.0 - k, 3.0, 4.0)
        views.append(fr)
    other = fuzz_frame(31, 97, 61)
    eager = [gpu_ctx.render(fr, ALL) for fr in views]
    eager_other = gpu_ctx.render(other, ALL)
    ctx = mra.Context(0)
    try:
        ctx.set_graphs(True)
        for rep in range(3):
            for i, fr in enumerate(views):
                g = ctx.render(fr, ALL)
                for k in ("rgb8", "depth", "winner", "rgb"):
                    assert np.array_equal(g[k].view(np.uint8), eager[i][k].view(np.uint8)), (rep, i, k)
            g = ctx.render(other, ALL)  # a different shape: re-capture
            for k in ("rgb8", "depth", "winner", "rgb"):
                assert np.array_equal(g[k].view(np.uint8), eager_other[k].view(np.uint8)), (rep, k)
            gpu_ctx.render(base, ("rgb8",))  # eager work on another context
    finally:
        ctx.close()
    assert_parity(eager[2], oracle_py.render(views[2], ALL), "graph view 2", normals=False)


def test_graph_update_mixed_shapes(gpu_ctx):
    """Captured batches of different sizes with the same kernels update one
    instantiated graph in place and replay it: every result equals the
    eager render, across three rounds with eager work in between, and the
    context counts updates and replays rather than new instantiations."""
    import minirender_amd as mra
    frames = [fuzz_frame(600 + i, 64 + 24 * i, 48 + 10 * i) for i in range(5)]
    batches = [frames[:2], frames[2:5], frames[1:4]]
    kinds = ("rgb8", "depth", "winner", "rgb")
    eager = [gpu_ctx.render_batch(b, kinds) for b in batches]
    ctx = mra.Context(0)
    try:
        ctx.set_graphs(True)
        for rep in range(3):
            for bi, b in enumerate(batches):
                out = ctx.render_batch(b, kinds)
                for j in range(len(b)):
                    for k in kinds:
                        assert np.array_equal(out[j][k].view(np.uint8),
                                              eager[bi][j][k].view(np.uint8)), (rep, bi, j, k)
            gpu_ctx.render(frames[0], ("rgb8",))  # eager work on another context
        st = ctx.stats()
    finally:
        ctx.close()
    counts = (st.graph_replays, st.graph_updates, st.graph_captures)
    assert sum(counts) == 9, counts  # every batch went through a graph
    assert st.graph_captures <= 3, counts


def test_autotune_and_cache(tmp_path):
    """The autotuner picks a built configuration, results stay exact under
    it, and a second call with the same cache file reads the choice back."""
    import minirender_amd as mra
    frames = [fuzz_frame(300 + i, 160, 120) for i in range(3)]
    cache = str(tmp_path / "tune.txt")
    ctx = mra.Context(0)
    try:
        t1 = ctx.autotune(frames, reps=1, cache_path=cache)
        assert t1["tile"] in (8, 16, 32) and not t1["from_cache"] and t1["candidates"] >= 10
        assert t1["bin_factor"] >= 1.0 and t1["ms"] > 0
        for i, fr in enumerate(frames):
            check(ctx, fr, f"autotuned frame {i}")
        t2 = ctx.autotune(frames, reps=1, cache_path=cache)
        assert t2["from_cache"]
        for k in ("tile", "raster_threads", "frag_size"):
            assert t2[k] == t1[k]
    finally:
        ctx.close()


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_bbox_end_bound_rounds_up(gpu_ctx, tile):
    """The last column is the last pixel centre of the reference's fp32 loop
    x <= pmax.x + 0.5f (Renderer.cpp:243): with pmax.x = 1024 - 2^-14 that
    is column 1024, which this scene covers (binning, raster and shading
    all use the same bound)."""
    from tests.parity import bound_frame
    gpu_ctx.set_tile(tile, tile)
    try:
        g, o = check(gpu_ctx, bound_frame(), f"bbox end bound t{tile}")
    finally:
        gpu_ctx.set_tile(16, 16)
    assert g["winner"][8, 1024] == 0


def test_no_near_clip_draws(gpu_ctx):
    """Draws flagged MR_DRAW_NO_NEAR_CLIP (Renderer::paintTriangle with
    world = false, Renderer.cpp:169) skip the near test: a triangle with a
    vertex between the eye and the near plane is drawn whole, and a sphere
    around the camera is drawn with vertices behind the eye projected as
    they are (no clipping, no near rejection)."""
    F = np.float32
    pos = np.array([[-1.0, -0.8, -2.0], [1.5, -0.6, -3.0], [0.2, 1.2, -0.3]], F)
    tri = scenes.Mesh(pos, np.tile(np.array([[0, 0, 1]], F), (3, 1)), np.array([[0, 1, 2]], np.int32))
    mat = Material(emissive=(0.9, 0.5, 0.1))
    proj = scenes.projection_frustum(1.0, 96 / 64, 0.5, 50.0)
    clipped = FrameSpec(96, 64, proj, draws=[Draw(tri, mat)], lighting=False)
    whole = FrameSpec(96, 64, proj, draws=[Draw(tri, mat, no_clip=True)], lighting=False)
    gc, _ = check(gpu_ctx, clipped, "near-clipped triangle")
    gw, _ = check(gpu_ctx, whole, "unclipped triangle")
    assert (gw["winner"] != gc["winner"]).sum() > 0
    sph = scenes.create_sphere(2.0, 10, 20)
    fr = FrameSpec(160, 120, scenes.projection_frustum(1.2, 4 / 3, 0.1, 20.0),
                   view=scenes.translate(0.3, -0.2, 0.5),
                   draws=[Draw(sph, Material(shininess=20.0), no_clip=True),
                          Draw(scenes.create_cube(0.7), Material(diffuse=(0.2, 0.8, 0.3)),
                               scenes.translate(0.0, 0.0, -1.5))],
                   light=(0.5, 1.0, 2.0), light_is_point=True)
    check(gpu_ctx, fr, "unclipped sphere around the eye")


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_row_bands(gpu_ctx, tile):
    """Tiles with much walk work are rastered as row bands, one workgroup per
    band with all of the tile's entries: a tiny threshold makes nearly every
    tile a banded one (large and clipped triangles, checkpoints, textures,
    crowded tiles split by entries next to banded ones)."""
    gpu_ctx.set_tile(tile, tile)
    gpu_ctx.set_band_work(64)
    try:
        for seed in (110, 111):
            check(gpu_ctx, fuzz_frame(seed, 160, 100), f"bands t{tile} seed {seed}")
        check(gpu_ctx, scenes.config_terrain(n=96, width=320, height=180, shapes=True),
              f"bands terrain t{tile}")
        check(gpu_ctx, stack_frame(1500, 80, 56, seed=tile), f"bands split t{tile}")
        rng_tris = scenes.random_triangles(40, 21, lo=0.2, hi=1.5, center_lo=(-1, -1, -3),
                                           center_hi=(1, 1, -2))
        fr = FrameSpec(640, 360, scenes.projection_frustum(math.radians(60), 16 / 9, 0.1, 50.0),
                       draws=[Draw(rng_tris, Material(shininess=40.0))], light=(0.3, 0.5, 1.0))
        check(gpu_ctx, fr, f"bands large triangles t{tile}")
    finally:
        gpu_ctx.set_band_work(32768)
        gpu_ctx.set_tile(16, 16)
