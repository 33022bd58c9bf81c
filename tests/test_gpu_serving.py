"""GPU: the render service (continuous batching, mr_server_*).

- every request's result in a mixed stream (different meshes, resolutions,
  cameras, batched with others, several batches in flight) is bit-identical
  to rendering that frame alone with mr_render, and matches the oracle;
- admission by HBM budget: requests larger than the budget are refused,
  a budget that holds one request at a time still completes everything;
- a replayed Poisson arrival trace completes with finite latencies.
"""
import numpy as np
import pytest

import minirender_amd as mra
from minirender_amd import MrError, scenes
from oracle import oracle_py
from tests.parity import assert_parity, fuzz_frame

pytestmark = pytest.mark.gpu

KINDS = ("rgb8", "depth", "winner", "rgb")


def mixed_frames(n, seed=0):
    rng = np.random.default_rng(seed)
    frames = []
    for i in range(n):
        w, h = [(97, 61), (160, 120), (64, 48), (211, 97)][int(rng.integers(0, 4))]
        frames.append(fuzz_frame(1000 + i, w, h))
    return frames


def same(a, b):
    return all(np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)) for k in a)


@pytest.mark.parametrize("max_batch,slots", [(8, 2), (3, 3), (1, 1)])
def test_alone_equals_in_mix(gpu_ctx, max_batch, slots):
    frames = mixed_frames(24, seed=max_batch)
    alone = [gpu_ctx.render(f, KINDS) for f in frames]
    with mra.Server(gpu_ctx, max_batch=max_batch, slots=slots) as srv:
        tickets = [srv.submit(f, KINDS) for f in frames]
        results = [srv.wait(t) for t in tickets]
    sizes = [info.batch_size for _, info in results]
    if max_batch > 1:
        assert max(sizes) > 1, "requests were never batched together"
    for i, (out, info) in enumerate(results):
        assert info.status == 0
        assert same(out, alone[i]), f"request {i} differs alone vs in the mix"
    for i in (0, 5, 11):
        assert_parity(results[i][0], oracle_py.render(frames[i], KINDS), f"served frame {i}")


def test_budget_admission(gpu_ctx):
    frames = mixed_frames(6, seed=7)
    alone = [gpu_ctx.render(f, KINDS) for f in frames]
    # a budget that fits about one request: everything still completes
    big = max(f.width * f.height for f in frames) * (4 + 3 + 4 + 4 + 12) + 10_000_000
    with mra.Server(gpu_ctx, max_batch=8, slots=2, hbm_budget=big) as srv:
        tickets = [srv.submit(f, KINDS) for f in frames]
        for i, t in enumerate(tickets):
            out, info = srv.wait(t)
            assert same(out, alone[i])
    # a request alone above the budget is refused
    with mra.Server(gpu_ctx, max_batch=8, slots=1, hbm_budget=1000) as srv:
        with pytest.raises(MrError):
            srv.submit(frames[0], KINDS)


def test_poisson_trace(gpu_ctx):
    from minirender_amd._capi import mr_frame, mr_outputs
    frames = mixed_frames(40, seed=3)
    structs, keep, outs = [], [], []
    for f in frames:
        s, k = gpu_ctx.frame(f)
        structs.append(s)
        keep.append(k)
        outs.append(mra.alloc_outputs(f, ("rgb8", "depth")))
    farr = (mr_frame * len(frames))(*structs)
    oarr = (mr_outputs * len(frames))(*[mra.outputs_struct(o) for o in outs])
    rng = np.random.default_rng(0)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, len(frames)))
    with mra.Server(gpu_ctx, max_batch=16, slots=2) as srv:
        lat, makespan = srv.run_trace(farr, arrivals.tolist(), oarr)
    assert (lat > 0).all() and np.isfinite(lat).all()
    assert makespan > 0
    ref = gpu_ctx.render(frames[17], ("rgb8", "depth"))
    assert same(outs[17], ref)
