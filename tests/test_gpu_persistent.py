"""The persistent kernel (mr_ctx_set_persistent): every stage of a batch in
one launch, its workgroups fed by device work queues.  Its outputs must be
bit-identical to the same batch run as separate launches, and match the
oracle, with the control words left clear for the next launch (batches run
back to back), through overflow re-runs, and for frames without geometry.
The reference renders one frame per call (Renderer.cpp:311-332, in the loop
of samples/bench.cpp:128-135); these are its frames.
"""
import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import scenes
from minirender_amd.scenes import Draw, FrameSpec, Material
from oracle import oracle_py
from tests.parity import assert_parity, fuzz_frame

pytestmark = pytest.mark.gpu

KINDS = ("rgb8", "depth", "winner")
MR_ERR_DEVICE = -8  # include/minirender_amd.h
ALL = ("rgb8", "depth", "winner", "rgb", "normals", "points")


def _same(a, b, kinds):
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)) for k in kinds)


@pytest.fixture
def pctx(gpu_ctx):
    gpu_ctx.set_tile(16, 16)
    gpu_ctx.set_persistent(True)
    try:
        yield gpu_ctx
    finally:
        gpu_ctx.set_persistent(False)


def test_persistent_batch_256_jobs(gpu_ctx):
    """All 256 jobs of the batched workload (1 k-500 k triangles, 1-8
    materials, 320x240 to 1920x1080): one launch, twice in a row, each job
    bit-identical to the separate launches; a sample against the oracle."""
    frames = scenes.batch_jobs()
    gpu_ctx.set_tile(16, 16)
    ref = gpu_ctx.render_batch(frames, KINDS)
    assert gpu_ctx.stats().launches > 1
    gpu_ctx.set_persistent(True)
    try:
        for rep in range(2):
            out = gpu_ctx.render_batch(frames, KINDS)
            assert gpu_ctx.stats().launches == 1
            for i in range(len(frames)):
                assert _same(out[i], ref[i], KINDS), f"rep {rep} job {i}: persistent != launches"
    finally:
        gpu_ctx.set_persistent(False)
    big = max(range(len(frames)), key=lambda i: frames[i].ntri)
    for i in sorted({0, 100, 255, big}):
        assert_parity(out[i], oracle_py.render(frames[i], KINDS), f"persistent batch job {i}")


def test_persistent_terrain_and_textured(pctx):
    """The 2.1 M-triangle terrain frame (near clipping, crowded and split
    tiles) and the textured frame, one launch each, against the oracle."""
    fr = scenes.config_terrain()
    g = pctx.render(fr, KINDS)
    st = pctx.stats()
    assert st.launches == 1 and st.triangles_in == fr.ntri
    assert_parity(g, oracle_py.render(fr, KINDS), "persistent terrain")
    fr = scenes.config_textured()
    assert_parity(pctx.render(fr, KINDS), oracle_py.render(fr, KINDS), "persistent textured")


def test_persistent_small_frames(pctx):
    """Small and odd frames: every output bit-identical to the separate
    launches, and the oracle's bar (fuzz frames as the parity tests draw
    them); an empty frame (no geometry: the geometry stages do not run),
    degenerate triangles, row bands."""
    for seed in (501, 502, 503):
        fr = fuzz_frame(seed, 97 + 17 * (seed - 500), 61 + 9 * (seed - 500))
        g = pctx.render(fr, ALL)
        assert pctx.stats().launches == 1
        pctx.set_persistent(False)
        try:
            ref = pctx.render(fr, ALL)
        finally:
            pctx.set_persistent(True)
        kinds = ALL if fr.save_normals else tuple(k for k in ALL if k != "normals")
        assert _same(g, ref, kinds), f"fuzz {seed}: persistent != launches"
    for seed in (1, 2, 3):
        fr = fuzz_frame(seed)
        assert_parity(pctx.render(fr, ALL), oracle_py.render(fr, ALL), f"persistent fuzz {seed}",
                      normals=fr.save_normals)
    empty = FrameSpec(40, 30, scenes.projection_frustum(1.0, 4 / 3, 0.1, 10),
                      background=(0.2, 0.4, 0.6))
    assert_parity(pctx.render(empty, ALL), oracle_py.render(empty, ALL), "persistent empty",
                  normals=False)
    pos = np.array([[0, 0, -2], [0, 0, -2], [1, 0, -2], [0, 1, -2]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    empty.draws = [Draw(scenes.Mesh(pos, nrm, idx), Material())]
    assert_parity(pctx.render(empty, ALL), oracle_py.render(empty, ALL), "persistent degenerate",
                  normals=False)
    pctx.set_band_work(64)
    try:
        fr = scenes.config_terrain(n=96, width=320, height=180, shapes=True)
        assert_parity(pctx.render(fr, KINDS), oracle_py.render(fr, KINDS), "persistent bands")
    finally:
        pctx.set_band_work(32768)


def test_persistent_overflow_rerun(pctx):
    """A bin list too small for the frame: the launch flags the overflow,
    the batch re-runs with larger buffers, still as one launch each."""
    fr = fuzz_frame(7, 640, 480)
    pctx.set_bin_factor(1.0)
    try:
        g = pctx.render(fr, ALL)
        st = pctx.stats()
    finally:
        pctx.set_bin_factor(3.0)
    assert st.launches == 1
    assert_parity(g, oracle_py.render(fr, ALL), "persistent bin overflow",
                  normals=fr.save_normals)


def test_persistent_needs_16_tiles(gpu_ctx):
    gpu_ctx.set_tile(8, 8)
    try:
        with pytest.raises(mra.MrError):
            gpu_ctx.set_persistent(True)
    finally:
        gpu_ctx.set_tile(16, 16)


def test_persistent_wait_gives_up(gpu_ctx):
    """A wait limit of 0 makes every wait of the persistent kernel give up:
    the launch drains (every workgroup exits), the batch fails with
    MR_ERR_DEVICE, and the next batches on the same context are right again
    (its control words and the counters of the stages that did not run are
    cleared first)."""
    fr = scenes.config_terrain(n=96, width=320, height=180, shapes=True)
    other = fuzz_frame(11)
    gpu_ctx.set_tile(16, 16)
    ref = gpu_ctx.render(fr, ALL)
    ref_other = gpu_ctx.render(other, ALL)
    gpu_ctx.set_persistent(True)
    try:
        gpu_ctx.set_wait_limit(0)
        for _ in range(2):
            with pytest.raises(mra.MrError) as e:
                gpu_ctx.render(fr, ALL)
            assert e.value.code == MR_ERR_DEVICE
        gpu_ctx.set_wait_limit(-1)
        for f, r in ((fr, ref), (other, ref_other), (fr, ref)):
            g = gpu_ctx.render(f, ALL)
            assert gpu_ctx.stats().launches == 1
            kinds = ALL if f.save_normals else tuple(k for k in ALL if k != "normals")
            assert _same(g, r, kinds)
    finally:
        gpu_ctx.set_wait_limit(-1)
        gpu_ctx.set_persistent(False)


def test_persistent_graph_replay(gpu_ctx):
    """Persistent batches captured as HIP graphs: replays (and in-place
    updates for other sizes) equal the eager separate launches, round after
    round, with the control words cleared by each replayed launch."""
    frames = [fuzz_frame(700 + i, 80 + 16 * i, 60 + 8 * i) for i in range(3)]
    kinds = ("rgb8", "depth", "winner", "rgb")
    gpu_ctx.set_tile(16, 16)
    eager = [gpu_ctx.render(f, kinds) for f in frames]
    ctx = mra.Context(0)
    try:
        ctx.set_tile(16, 16)
        ctx.set_persistent(True)
        ctx.set_graphs(True)
        for rep in range(3):
            for i, f in enumerate(frames):
                g = ctx.render(f, kinds)
                assert ctx.stats().launches == 1
                assert _same(g, eager[i], kinds), (rep, i)
        st = ctx.stats()
    finally:
        ctx.close()
    assert st.graph_replays + st.graph_updates + st.graph_captures == 9
    assert st.graph_replays + st.graph_updates >= 6
