"""Benchmark of the rasterization path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload terrain|megatri|sphere]

A step is one Renderer::render() of the workload (inputs resident in HBM,
outputs to device buffers).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="terrain", choices=["terrain", "megatri", "sphere", "textured"])
    p.add_argument("--tile", type=int, default=16)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def make_frame(name):
    from minirender_amd import scenes
    if name == "terrain":
        return scenes.config_terrain(), "terrain 1024^2 + 10x10 shapes, 1920x1080"
    if name == "megatri":
        return scenes.config_megatri(), "1M-tri grid 708^2, 1920x1080"
    if name == "textured":
        return scenes.config_textured(), "textured terrain 1024^2 + 2048^2 tex + 5k random tris, 1920x1080"
    return scenes.config_sphere(), "UV sphere 64x32, 640x480"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import minirender_amd as mra

    frame, wl = make_frame(args.workload)
    ctx = mra.Context(0)
    ctx.set_tile(args.tile, args.tile)
    dev = torch.device("cuda:0")
    H, W = frame.height, frame.width
    rgb8 = torch.empty((H, W, 3), dtype=torch.uint8, device=dev)
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    import ctypes as C
    from minirender_amd._capi import mr_outputs
    fstruct, keep = ctx.frame(frame)  # host matrices built once, as a caller would
    outs = mr_outputs()
    outs.rgb8, outs.depth, outs.on_device = rgb8.data_ptr(), depth.data_ptr(), 1

    def step():
        rc = ctx.lib.mr_render(ctx.h, C.byref(fstruct), C.byref(outs))
        if rc:
            raise RuntimeError(f"mr_render failed: {rc}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(True)
    stage = np.zeros(6)
    fms = []
    for _ in range(10):
        step()
        s = ctx.stats()
        stage += np.array(s.stage_ms[:6])
        fms.append(s.frame_ms)
    ctx.set_profiling(False)
    st = ctx.stats()
    T = frame.ntri
    out = {
        "metric": "Mtris/s (plain per-request GPU path)", "value": T / dt / 1e6, "unit": "Mtris/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "frames_per_s": 1.0 / dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl, "triangles": T, "tile": args.tile},
        "stage_ms": dict(zip(["clear", "geometry", "scan", "bin", "raster", "shade"], (stage / 10).round(4).tolist())),
        "frame_ms_gpu": float(np.median(fms)),
        "records": st.records, "bin_entries": st.bin_entries, "clipped": st.triangles_clipped,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
