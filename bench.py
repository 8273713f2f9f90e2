#!/usr/bin/env python3
"""Benchmark: SoftRast's sort-middle path on one MI355X.

Workload (default): 1920x1080, 1,000,000 triangles = 4,000 seeded UV spheres
(250 tris each, log-normal radii), 8 draw calls each bound to its own
2048^2 RGBA8 value-noise texture with a full mip chain, UnlitDiffuse
shading, reverse-Z, clear + draw + EndFrame per step. Inputs (vertex, index
and texture data) are resident in HBM before timing starts; a step is one
whole frame (clear, 8 DrawIndexed, synchronous EndFrame).

Prints one JSON line. Also reported:
  roofline      HBM roofline of the dominant kernel (k_backend), timed with
                HIP events on the library's stream over the timed frames
  cpu_baseline  the CPU oracle (a scalar C port of the reference path,
                compiled -O3 -mavx2 -mfma) on the host cores, same frame
  parity        differing-pixel counts of the benchmarked frame vs the oracle
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "frames/s + Gfrags/s at 1080p/1M tris; batched+learned vs plain GPU, same accuracy"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(host cores, 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--tile-order", type=int, default=0)
    return ap.parse_args()


def build_frame(sr, ctx, fb, scene, texs):
    calls = []
    for d in scene.draws:
        c = sr.DrawCall()
        c.SetFrameBuffer(fb).SetMVP(scene.mvp)
        c.SetIndexBuffer(d.indices, 4, len(d.indices))
        c.SetPositionBuffer(d.vertices, 32, len(d.vertices))
        c.SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6)
        c.SetPixelShader(sr.UnlitDiffuseShader, texs[d.texture])
        calls.append(c)
    return calls


def algorithmic_bytes(scene, covered_px):
    """B = s_idx*3*T + 32*V + 8*W*H + 16*P_cov per frame; back-end share = 8*W*H + 16*P_cov."""
    s_idx = scene.draws[0].indices.dtype.itemsize
    T = scene.num_tris
    V = sum(len(d.vertices) for d in scene.draws)
    fe = s_idx * 3 * T + 32 * V
    be = 8 * scene.width * scene.height + 16 * covered_px
    return fe, be


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              f"(views shard across GPUs with no exchange; see the multi-GPU design notes)", file=sys.stderr)
        sys.exit(2)
    import softrast_amd as sr
    from softrast_amd import scenes

    t0 = time.time()
    scene = scenes.spheres_scene()
    texs = [sr.TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in scene.textures]
    print(f"[bench] scene ready: {scene.num_tris} tris, {len(scene.draws)} draws, {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    ctx = sr.RenderContext(flags=sr.CFG_TIMING | sr.CFG_STATS, tile_order=args.tile_order)
    fb = sr.FrameBuffer(ctx, scene.width, scene.height)
    calls = build_frame(sr, ctx, fb, scene, texs)

    def frame():
        ctx.BeginFrame()
        ctx.ClearFrameBuffer(fb, 0)
        for c in calls:
            c.m_frameBuffer = (fb, fb.m_writePlane)
            ctx.DrawIndexed(c)
        ctx.EndFrame()

    for _ in range(args.warmup):  # also uploads and caches every buffer and texture
        frame()
    be_ms, fe_ms, dev_ms = [], [], []
    t_start = time.perf_counter()  # EndFrame is synchronous: the device is idle here
    for _ in range(args.steps):
        frame()
        st = ctx.frame_stats()
        be_ms.append(st["ms_backend"])
        fe_ms.append(st["ms_frontend"])
        dev_ms.append(st["ms_total"])
    elapsed = time.perf_counter() - t_start  # and idle again here
    stats = ctx.frame_stats()
    ms_per_step = 1000.0 * elapsed / args.steps
    fps = args.steps / elapsed

    colour, depth = fb.read_tiles()
    covered = int(np.count_nonzero(depth.reshape(-1).view(np.uint32) != 0))
    fe_bytes, be_bytes = algorithmic_bytes(scene, covered)
    be_avg = statistics.mean(be_ms)
    achieved = be_bytes / (be_avg * 1e-3) / 1e9
    traffic = load_traffic("k_backend_spheres_1080p_1m_traffic.json")

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded spheres 0x5EED0002, value-noise textures)",
        "config": {"workload": "spheres_1080p_1m", "width": scene.width, "height": scene.height,
                   "triangles": scene.num_tris, "draws": len(scene.draws), "textures": "8x2048^2 RGBA8 + mips",
                   "path": "plain single-view", "tile_order": args.tile_order},
        "gfrags_per_s": round(stats["frags_raster"] * fps / 1e9, 3),
        "mtris_per_s": round(scene.num_tris * fps / 1e6, 1),
        "device_ms": {"frontend": round(statistics.mean(fe_ms), 4), "backend": round(be_avg, 4),
                      "total": round(statistics.mean(dev_ms), 4)},
        "frags_raster": stats["frags_raster"],
        "bin_entries": stats["bin_entries"],
        "roofline": {"bound": "hbm", "kernel": "k_backend", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "algorithmic_bytes": be_bytes,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None},
        "frame_algorithmic_bytes": fe_bytes + be_bytes,
    }

    if not args.no_parity or not args.no_cpu_baseline:
        from oracle.oracle import Oracle, Plane, make_texture
        stex = [make_texture(t) for t in texs]
        odraws = []
        for d in scene.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            odraws.append(dict(shader=0, tex=stex[d.texture][0], idx=d.indices, idx_stride=4, pos=vf, pos_stride=32,
                               pos_num=len(vf), attr=vf, attr_stride=32, attr_num=len(vf), uv_offset=6, plane=0,
                               mvp=scene.mvp))
        nthreads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        orc = Oracle(nthreads=nthreads, fast=True)
        plane = Plane(scene.width, scene.height)
        if not args.no_parity:
            orc.clear(plane, 0)
            orc.render(odraws, [plane])
            dd = int(np.count_nonzero(plane.depth.view(np.uint32) != depth.reshape(-1).view(np.uint32)))
            dc = int(np.count_nonzero(plane.colour != colour.reshape(-1)))
            out["parity"] = {"depth_px_diff": dd, "rgba8_px_diff": dc, "vs": "oracle"}
        if not args.no_cpu_baseline:
            times = []
            for i in range(5 + args.cpu_frames):
                t = time.perf_counter()
                orc.clear(plane, 0)
                orc.render(odraws, [plane])
                times.append(time.perf_counter() - t)
            med = statistics.median(times[5:])
            out["cpu_baseline"] = {"value": round(1.0 / med, 3), "unit": "frames/s", "cores": nthreads,
                                   "kind": "port",
                                   "sample": f"same 1080p/1M-tri frame (clear+render), median of {args.cpu_frames} "
                                             f"frames after 5 warm-up, {nthreads} threads, oracle -O3 AVX2/FMA build"}
            out["speedup_vs_cpu"] = round(fps * med, 1)
        orc.close()

    print(json.dumps(out))
    fb.destroy()
    ctx.Shutdown()


if __name__ == "__main__":
    main()
