"""GPU parity at the full size of every benchmark workload.

The bench times these workloads; here their outputs are checked:
- the terrain frame (1024^2 height-field + 100 shapes, 2.1 M triangles,
  1920x1080, near clipping) under every tile size, the textured frame
  (2048^2 texture, 5 k random triangles) and the 1 M-triangle grid, each
  against the oracle (winner and depth bit-equal, RGB8 within 1 LSB);
- all 256 jobs of the batched workload (1 k-500 k triangles, 1-8
  materials, 320x240 to 1920x1080) rendered in one mr_render_batch, each
  bit-identical to rendering it alone, and a stratified sample of them
  (every resolution, the largest job) against the oracle;
- 200 requests over the full 64-mesh serving pool through the continuous-
  batching server (64 per batch, 4 slots): each bit-identical to rendering it
  alone, and a sample against the oracle.
The reference renders one frame per call (Renderer.cpp:311-332); these are
its frames, whatever the path that produced them.
"""
import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import scenes
from oracle import oracle_py
from tests.parity import assert_parity

pytestmark = pytest.mark.gpu

KINDS = ("rgb8", "depth", "winner")


def _same(a, b):
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)) for k in a)


@pytest.fixture(scope="module")
def terrain_oracle():
    fr = scenes.config_terrain()
    o = oracle_py.render(fr, KINDS)
    return fr, o, oracle_py.last_clipped()


@pytest.mark.parametrize("tile,threads", [(8, 64), (16, 256), (32, 256)])
def test_terrain_full(gpu_ctx, terrain_oracle, tile, threads):
    fr, o, clipped = terrain_oracle
    gpu_ctx.set_tile(tile, tile)
    gpu_ctx.set_raster_threads(threads)
    try:
        g = gpu_ctx.render(fr, KINDS)
        st = gpu_ctx.stats()
    finally:
        gpu_ctx.set_tile(16, 16)
    assert st.triangles_in == fr.ntri > 2_000_000
    assert st.triangles_clipped == clipped > 0
    assert_parity(g, o, f"terrain 1080p tile {tile}/{threads}")
    assert (g["winner"] != mra.NO_WINNER).sum() > 1_000_000


def test_textured_full(gpu_ctx):
    fr = scenes.config_textured()
    g = gpu_ctx.render(fr, KINDS)
    assert_parity(g, oracle_py.render(fr, KINDS), "textured 1080p")


def test_megatri_full(gpu_ctx):
    fr = scenes.config_megatri()
    assert fr.ntri == 999_698
    g = gpu_ctx.render(fr, KINDS)
    assert_parity(g, oracle_py.render(fr, KINDS), "1M-triangle grid 1080p")
    assert (g["winner"] != mra.NO_WINNER).sum() > 500_000


def _stratified(frames, per_res=4):
    """Job indices: up to per_res of every resolution (spread over triangle
    counts), plus the job with the most triangles."""
    pick = set()
    for w, h in scenes.RESOLUTIONS:
        ids = sorted((i for i, f in enumerate(frames) if (f.width, f.height) == (w, h)),
                     key=lambda i: frames[i].ntri)
        if ids:
            for q in np.linspace(0, len(ids) - 1, min(per_res, len(ids))):
                pick.add(ids[int(round(q))])
    pick.add(max(range(len(frames)), key=lambda i: frames[i].ntri))
    return sorted(pick)


def test_batch_256_jobs(gpu_ctx):
    frames = scenes.batch_jobs()
    assert len(frames) == 256
    batch = gpu_ctx.render_batch(frames, KINDS)
    for i, fr in enumerate(frames):
        assert _same(batch[i], gpu_ctx.render(fr, KINDS)), f"job {i} differs batched vs alone"
    sample = _stratified(frames)
    assert len(sample) >= 16
    for i in sample:
        assert_parity(batch[i], oracle_py.render(frames[i], KINDS), f"batch job {i}")


def test_serving_pool_full(gpu_ctx):
    pool = scenes.serving_pool()
    assert len(pool) == 64
    reqs = scenes.serving_requests(pool, 200)
    alone = [gpu_ctx.render(f, KINDS) for f in reqs]
    with mra.Server(gpu_ctx, max_batch=64, slots=4) as srv:
        tickets = [srv.submit(f, KINDS) for f in reqs]
        results = [srv.wait(t) for t in tickets]
    assert max(info.batch_size for _, info in results) > 1, "requests were never batched"
    for i, (out, info) in enumerate(results):
        assert info.status == 0
        assert _same(out, alone[i]), f"request {i} differs alone vs in the mix"
    for i in _stratified(reqs, per_res=2):
        assert_parity(results[i][0], oracle_py.render(reqs[i], KINDS), f"served request {i}")


def test_target_stream_served(gpu_ctx):
    """The 1M-triangle 1080p request stream of the service comparison: 96
    requests (per-request orbit camera and light) through the batching
    server, each bit-identical to rendering it alone, a sample against the
    oracle."""
    reqs = scenes.megatri_requests(96)
    alone = [gpu_ctx.render(f, KINDS) for f in reqs]
    with mra.Server(gpu_ctx, max_batch=64, slots=4) as srv:
        tickets = [srv.submit(f, KINDS) for f in reqs]
        results = [srv.wait(t) for t in tickets]
    for i, (out, info) in enumerate(results):
        assert info.status == 0
        assert _same(out, alone[i]), f"request {i} differs alone vs in the mix"
    for i in (0, 47, 95):
        assert_parity(results[i][0], oracle_py.render(reqs[i], KINDS), f"target request {i}")
