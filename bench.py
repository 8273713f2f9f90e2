#!/usr/bin/env python3
"""Benchmark: SoftRast's sort-middle path on one MI355X.

Workload: 1920x1080, 1,000,000 triangles = 4,000 seeded UV spheres (250
tris each, log-normal radii), 8 draw calls each bound to its own 2048^2
RGBA8 value-noise texture with a full mip chain, UnlitDiffuse shading,
reverse-Z. Vertex, index and texture data are resident in HBM before timing
starts. Two paths are timed:

  plain    one view per EndFrame: clear + 8 DrawIndexed + synchronous EndFrame
  batched  --views V (default 64) views of the same scene (cameras turned by
           up to +-2 degrees, one framebuffer each) in one EndFrame, launched as a replayed
           hipGraph, work-bin layout chosen by the library's autotuner, tile
           queue ordered by --tile-order (default: the learned cost model,
           trained on the chosen layout before timing); the same path with
           the tri-count order is timed beside it

A step is one EndFrame; frames/s counts views completed per second
(value = the batched path). Also reported:
  roofline      HBM roofline of the dominant kernel of the batched launch
                (k_backend on the chosen layout): the algorithmic bytes of all
                its views over the kernel's time, from HIP events around it on
                the library's stream in launches of the same layout and order
                without a graph; HBM traffic and VALU issue counts from the
                rocprofv3 summaries in profiles/ (the plain path's figures
                are under "plain")
  cpu_baseline  the CPU oracle (a C port of the reference path, -O3
                AVX2/FMA build, reference task model) on the host cores
  parity        differing-pixel counts vs the oracle: the plain frame, and
                the first, middle and last view of the timed batched launch
"""
import argparse
import json
import math
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
    ap.add_argument("--views", type=int, default=64, help="views per EndFrame in the batched path")
    ap.add_argument("--tile-order", type=int, default=2, help="batched path: 0 natural, 1 tri-count, 2 learned")
    ap.add_argument("--no-autotune", action="store_true", help="batched path: library default layout")
    ap.add_argument("--bin-size", type=int, default=0, help="work-bin edge 16/32/64 (0 = library default)")
    ap.add_argument("--cpu-frames", type=int, default=20)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(host cores, 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--sched-study", action="store_true",
                    help="tile-order study on the clustered 3840x2160 / 4M-triangle workload instead of the default")
    ap.add_argument("--no-plain", action="store_true",
                    help="profiling aid: time only the batched path (--steps launches) and print its rate")
    ap.add_argument("--workload", choices=("spheres", "mixed", "combined", "sponza"), default="spheres",
                    help="mixed: 256 mixed-resolution views per EndFrame under the autotuner; combined: 64 views "
                         "of 2560x1440 x 2M triangles, learned vs tri-count order; sponza: the 1080p/1M frame "
                         "drawn with the SponzaShader (16 point lights) against UnlitDiffuse (the default is the "
                         "1080p/1M workload)")
    ap.add_argument("--nviews", type=int, default=0, help="mixed/combined: views per batch (0 = the workload's)")
    ap.add_argument("--present", action="store_true",
                    help="the plain 1080p/1M frame with its pixels brought to the host each frame: a synchronous "
                         "Blit, and an asynchronous one overlapped with the next frame (PCIe-inclusive rates)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def view_mvp(scene, yaw_deg):
    from softrast_amd import scenes
    t = math.radians(yaw_deg)
    return scenes.camera_mvp((0.0, 0.0, 0.0), (math.sin(t), 0.0, math.cos(t)), 60.0, scene.width / scene.height)


def make_calls(sr, fb, scene, texs, mvp, shader=None):
    calls = []
    for d in scene.draws:
        c = sr.DrawCall()
        c.SetFrameBuffer(fb).SetMVP(mvp)
        c.SetIndexBuffer(d.indices, 4, len(d.indices))
        c.SetPositionBuffer(d.vertices, 32, len(d.vertices))
        c.SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6)
        c.SetPixelShader(shader or sr.UnlitDiffuseShader, texs[d.texture])
        calls.append(c)
    return calls


def time_path(sr, scene, texs, views, flags, tile_order, bin_size, steps, warmup, shader=None, sponza=None,
              wg_per_tile=0):
    """Returns (frames/s, ms per EndFrame, per-frame stats, ctx, framebuffers).
    Untimed frames first: `warmup`, and with the autotuner or the learned
    order, until the layout is chosen and the model trained on it. shader:
    the pixel shader of every draw (UnlitDiffuse by default); sponza: the
    SponzaShader constants."""
    ctx = sr.RenderContext(flags=flags, tile_order=tile_order, bin_size=bin_size, wg_per_tile=wg_per_tile,
                           autotune_frames=3)
    if sponza is not None:
        ctx.set_sponza_constants(sponza)
    fbs, calls = [], []
    for v in range(views):
        fb = sr.FrameBuffer(ctx, scene.width, scene.height)
        # cameras turned by up to +-2 degrees: every view sees nearly the plain view's scene
        mvp = batched_view_mvp(scene, views, v)
        fbs.append(fb)
        calls.append(make_calls(sr, fb, scene, texs, mvp, shader))

    draws = ctx.PrepareDrawList([c for cs in calls for c in cs])  # uploads every buffer and texture
    clears = ctx.PrepareClearList(fbs)

    def frame():
        ctx.BeginFrame()
        ctx.ClearFrameBuffers(clears, 0)
        ctx.DrawIndexedList(draws)
        ctx.EndFrame()

    for _ in range(warmup):
        frame()
    for _ in range(200):
        tuned = not (flags & sr.CFG_AUTOTUNE) or ctx.tuning()["state"] == "chosen"
        trained = tile_order != sr.ORDER_LEARNED or ctx.frame_stats()["sched_trained"]
        if tuned and trained:
            break
        frame()
    # the frames after the layout is chosen or the model trained capture
    # their graph (a few ms, once): outside the timed frames
    for _ in range(2):
        frame()
    stats = []
    t0 = time.perf_counter()  # EndFrame is synchronous: the device is idle here
    for _ in range(steps):
        frame()
        stats.append(ctx.frame_stats())
    elapsed = time.perf_counter() - t0  # and idle again here
    return views * steps / elapsed, 1000.0 * elapsed / steps, stats, ctx, fbs


def batched_view_mvp(scene, views, v):
    """Camera of view v of a batch: turned by up to +-2 degrees, so every view
    sees nearly the plain view's scene."""
    return scene.mvp if views == 1 else view_mvp(scene, (v - (views - 1) / 2.0) * 4.0 / (views - 1))


def batched_stats(sr, scene, texs, views, bin_size):
    """Rasterised fragments per view and covered pixels of each view of the
    batched views, from one untimed statistics frame (the views' cameras
    differ from the plain view's)."""
    ctx = sr.RenderContext(flags=sr.CFG_STATS, bin_size=bin_size)
    fbs = []
    ctx.BeginFrame()
    for v in range(views):
        fb = sr.FrameBuffer(ctx, scene.width, scene.height)
        fbs.append(fb)
        ctx.ClearFrameBuffer(fb, 0)
        for c in make_calls(sr, fb, scene, texs, batched_view_mvp(scene, views, v)):
            ctx.DrawIndexed(c)
    ctx.EndFrame()
    frags = ctx.frame_stats()["frags_raster"]
    covered = []
    for fb in fbs:
        covered.append(int(np.count_nonzero(fb.read_tiles()[1].reshape(-1).view(np.uint32) != 0)))
        fb.destroy()
    ctx.Shutdown()
    return frags / views, covered


def load_profile_value(name, key):
    """An integer 'key = N' line of a committed rocprofv3 summary in profiles/."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            for line in f:
                if line.strip().startswith(key):
                    return int(line.split("=", 1)[1].strip())
    except (OSError, ValueError):
        pass
    return None


def load_traffic(name):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    return load_profile_value(name, "traffic_bytes_per_launch")


def profile_batched(sr, scene, texs, args):
    """--no-plain: the batched path alone (for a profiler); prints its rate only."""
    flags = sr.CFG_TIMING | sr.CFG_GRAPH | (0 if args.no_autotune else sr.CFG_AUTOTUNE)
    fps, ms, _, ctx, fbs = time_path(sr, scene, texs, args.views, flags, args.tile_order, args.bin_size, args.steps,
                                     args.warmup)
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()
    print(json.dumps({"metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": round(ms, 4), "path": f"batched x{args.views} only"}))


def sched_study(sr, args):
    """--sched-study: one view of the clustered workload (3840x2160, 4,000,000
    triangles, 90% in 3 clusters, 16 textures of 4096^2 built on the device),
    rendered with each tile order; the learned order trains on the tile costs
    of 64 warm-up frames first."""
    from softrast_amd import scenes
    t0 = time.time()
    scene = scenes.clustered_scene()
    print(f"[bench] clustered scene ready: {scene.num_tris} tris, {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    out = {"metric": "frames/s per tile order, clustered 3840x2160 / 4M tris (one view per EndFrame)",
           "unit": "frames/s", "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "data": "synthetic",
           "config": {"workload": "clustered_4k_4m", "width": scene.width, "height": scene.height,
                      "triangles": scene.num_tris, "textures": "16x4096^2 RGBA8 + mips (device-built)"},
           "orders": {}}
    runs = ((sr.ORDER_NATURAL, "natural", 0), (sr.ORDER_TRICOUNT, "tricount", 0), (sr.ORDER_LEARNED, "learned", 0),
            (sr.ORDER_TRICOUNT, "tricount_autotuned", sr.CFG_AUTOTUNE))
    for order, name, extra in runs:
        ctx = sr.RenderContext(flags=sr.CFG_TIMING | extra, tile_order=order, bin_size=args.bin_size,
                               sched_warmup_frames=64, autotune_frames=2)
        texs = [ctx.CreateTexture(t, t.shape[1], t.shape[0], True) for t in scene.textures]
        fb = sr.FrameBuffer(ctx, scene.width, scene.height)
        draws = ctx.PrepareDrawList(make_calls(sr, fb, scene, texs, scene.mvp))

        def frame():
            ctx.BeginFrame()
            ctx.ClearFrameBuffer(fb, 0)
            ctx.DrawIndexedList(draws)
            ctx.EndFrame()
        for _ in range(max(args.warmup, 70)):
            frame()
        st = []
        t = time.perf_counter()
        for _ in range(args.steps):
            frame()
            st.append(ctx.frame_stats())
        el = time.perf_counter() - t
        out["orders"][name] = {"frames_per_s": round(args.steps / el, 2),
                               "device_ms": round(statistics.mean(s["ms_total"] for s in st), 4),
                               "backend_ms": round(statistics.mean(s["ms_backend"] for s in st), 4),
                               "layout": {"bin_size": st[-1]["bin_size"], "wg_per_tile": st[-1]["wg_per_tile"]}}
        if extra & sr.CFG_AUTOTUNE:
            t = ctx.tuning()
            out["orders"][name]["autotune_device_ms"] = {f"bin{b}_wg{w}": round(ms, 3) for (b, w), ms in t["ms"].items()}
        if order == sr.ORDER_LEARNED:
            out["orders"][name]["sched_err_log2"] = round(statistics.mean(s["sched_err"] for s in st), 4)
            out["orders"][name]["trained"] = bool(st[-1]["sched_trained"])
        print(f"[bench] {name}: {out['orders'][name]}", file=sys.stderr, flush=True)
        fb.destroy()
        ctx.Shutdown()
    if not args.no_cpu_baseline:
        view = scenes.Scene(scene.width, scene.height, scene.mvp, scene.draws, [])
        out["cpu_baseline"] = cpu_baseline_views([view], host_textures(sr, scene.textures), args, budget_s=15.0,
                                                 what=": the same 4K frame")
        best = max(v["frames_per_s"] for v in out["orders"].values())
        out["speedup_vs_cpu"] = round(best / out["cpu_baseline"]["value"], 1)
        print(f"[bench] cpu baseline: {out['cpu_baseline']}", file=sys.stderr, flush=True)
    print(json.dumps(out))


def viewset_calls(sr, vs, fbs, texs):
    calls = []
    shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader, 2: sr.VisualizeUVsShader}
    for v, fb in zip(vs.views, fbs):
        for d in v.draws:
            c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(v.mvp)
            c.SetIndexBuffer(d.indices, d.indices.dtype.itemsize, len(d.indices))
            c.SetPositionBuffer(d.vertices, 32, len(d.vertices))
            c.SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6)
            c.SetPixelShader(shader[d.shader], texs[d.texture])
            calls.append(c)
    return calls


def viewset_frame(sr, ctx, fbs, draws):
    ctx.BeginFrame()
    for fb in fbs:
        ctx.ClearFrameBuffer(fb, 0)
    ctx.DrawIndexedList(draws)
    ctx.EndFrame()


def host_textures(sr, images):
    """TextureData (tiled, with mips) of each image, built on the GPU and read
    back: the same bytes as the host build, for the CPU baseline."""
    ctx = sr.RenderContext()
    out = [ctx.CreateTexture(t, t.shape[1], t.shape[0], True).read() for t in images]
    ctx.Shutdown()
    return out


def cpu_baseline_views(views, texdata, args, budget_s=20.0, min_views=2, what=""):
    """The CPU oracle (-O3 AVX2/FMA build, the reference's task model) on the
    host cores, one view per frame (clear + render), views taken evenly
    spaced from `views` (cycling for fewer views than frames) until about
    budget_s of rendering: a bounded sample of the workload."""
    from oracle.oracle import Oracle, Plane, make_texture
    nthreads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    orc = Oracle(nthreads=nthreads, fast=True)
    stex = [make_texture(t) for t in texdata]
    n = len(views)
    stride = max(1, n // 8)
    order = [i for off in range(stride) for i in range(off, n, stride)]  # every stride-th view first
    times, done, k = [], [], 0
    while k < min_views or (sum(times) < budget_s and k < max(n, min_views) * 4):
        vi = order[k % n]
        v = views[vi]
        plane = Plane(v.width, v.height)
        draws = []
        for d in v.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            draws.append(dict(shader=d.shader, tex=stex[d.texture][0] if d.texture >= 0 else None, idx=d.indices,
                              idx_stride=d.indices.dtype.itemsize, pos=vf, pos_stride=32, pos_num=len(vf), attr=vf,
                              attr_stride=32, attr_num=len(vf), uv_offset=d.uv_offset, plane=0, mvp=v.mvp))
        t = time.perf_counter()
        orc.clear(plane, 0)
        orc.render(draws, [plane])
        times.append(time.perf_counter() - t)
        done.append(vi)
        k += 1
    orc.close()
    return {"value": round(len(times) / sum(times), 3), "unit": "views/s", "cores": nthreads, "kind": "port",
            "sample": f"{len(times)} view renders ({len(set(done))} distinct views{what}), clear + render each, "
                      f"{sum(times):.1f} s, oracle -O3 AVX2/FMA build, reference task model"}


def batch_workload(sr, args):
    """--workload mixed | combined: every view of the batch in one EndFrame
    (one hipGraph-captured launch sequence); frames/s counts views.

    mixed     256 views, resolutions uniform over {256^2, 512^2, 1024x768,
              1920x1080}, 1k-500k triangles each (log-uniform), sphere/cube
              mixes, 8 x 1024^2 textures; tri-count order; the autotuner sweeps
              bin size {8,16,32,64} x workgroups per bin {1,2,4} first
    combined  64 views of 2560x1440 x 2,000,000 triangles, alternating uniform
              and 90/10 clustered layouts, 8 x 2048^2 textures; tri-count
              order, then the learned order (trained after 16 frames,
              retrained every 16)
    """
    from softrast_amd import scenes
    t0 = time.time()
    if args.workload == "mixed":
        vs = scenes.mixed_views(nviews=args.nviews or 256)
        runs = [("autotuned", dict(flags=sr.CFG_TIMING | sr.CFG_GRAPH | sr.CFG_AUTOTUNE, tile_order=sr.ORDER_TRICOUNT,
                                   autotune_frames=2))]
        config = {"workload": f"mixed_views_{len(vs.views)}", "views": len(vs.views),
                  "resolutions": "uniform over 256x256, 512x512, 1024x768, 1920x1080",
                  "triangles_per_view": "log-uniform 1k-500k", "textures": "8x1024^2 RGBA8 + mips (device-built)"}
    else:
        vs = scenes.combined_views(nviews=args.nviews or 64)
        runs = [("tricount", dict(flags=sr.CFG_TIMING | sr.CFG_GRAPH, tile_order=sr.ORDER_TRICOUNT)),
                ("learned", dict(flags=sr.CFG_TIMING | sr.CFG_GRAPH, tile_order=sr.ORDER_LEARNED,
                                 sched_warmup_frames=16, sched_retrain_interval=16))]
        config = {"workload": f"combined_{len(vs.views)}x2560x1440_2m", "views": len(vs.views),
                  "width": 2560, "height": 1440, "triangles_per_view": vs.views[0].num_tris,
                  "layouts": "even views uniform, odd views 90% in 3 clusters",
                  "textures": "8x2048^2 RGBA8 + mips (device-built)"}
    ntris = vs.num_tris
    print(f"[bench] {args.workload} views ready: {len(vs.views)} views, {ntris} tris, {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    out = {"metric": f"views/s, {config['workload']} (all views in one EndFrame)", "unit": "views/s",
           "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32+int32", "data": "synthetic (seeded, see scenes.py)",
           "config": config, "runs": {}}
    best = None
    for name, kw in runs:
        ctx = sr.RenderContext(bin_size=args.bin_size, **kw)
        texs = [ctx.CreateTexture(t, t.shape[1], t.shape[0], True) for t in vs.textures]
        fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
        calls = ctx.PrepareDrawList(viewset_calls(sr, vs, fbs, texs))
        t1 = time.time()
        viewset_frame(sr, ctx, fbs, calls)  # uploads every buffer, grows the arenas
        print(f"[bench] {name}: first batch {time.time() - t1:.1f}s", file=sys.stderr, flush=True)
        run = {}
        if kw["flags"] & sr.CFG_AUTOTUNE:
            sweep = 1
            while ctx.tuning()["state"] == "sweeping":
                viewset_frame(sr, ctx, fbs, calls)
                sweep += 1
            t = ctx.tuning()
            run["autotune"] = {"frames": sweep, "chosen": {"bin_size": t["layout"][0], "wg_per_tile": t["layout"][1]},
                               "device_ms": {f"bin{b}_wg{w}": round(ms, 3) for (b, w), ms in t["ms"].items()}}
            print(f"[bench] {name}: autotuner chose {t['layout']} after {sweep} batches", file=sys.stderr, flush=True)
        warm = args.warmup
        if kw.get("tile_order") == sr.ORDER_LEARNED:
            warm = max(warm, kw["sched_warmup_frames"] + 2)
        for _ in range(warm):
            viewset_frame(sr, ctx, fbs, calls)
        st = []
        t1 = time.perf_counter()  # EndFrame is synchronous: the device is idle here
        for _ in range(args.steps):
            viewset_frame(sr, ctx, fbs, calls)
            st.append(ctx.frame_stats())
        el = time.perf_counter() - t1
        run.update({"views_per_s": round(len(vs.views) * args.steps / el, 2), "ms_per_batch": round(1e3 * el / args.steps, 3),
                    "device_ms_per_batch": round(statistics.mean(s["ms_total"] for s in st), 3),
                    "layout": {"bin_size": st[-1]["bin_size"], "wg_per_tile": st[-1]["wg_per_tile"]}})
        if kw.get("tile_order") == sr.ORDER_LEARNED:
            run["sched_err_log2"] = round(statistics.mean(s["sched_err"] for s in st), 4)
            run["trainings_in_timed_batches"] = sum(s["sched_trainings"] for s in st)
        out["runs"][name] = run
        print(f"[bench] {name}: {run}", file=sys.stderr, flush=True)
        if best is None or run["views_per_s"] > out["runs"][best]["views_per_s"]:
            best = name
        layout = (st[-1]["bin_size"], st[-1]["wg_per_tile"])
        for fb in fbs:
            fb.destroy()
        ctx.Shutdown()

    # untimed frame with statistics on the best run's layout: fragments and covered pixels
    sctx = sr.RenderContext(flags=sr.CFG_STATS, bin_size=layout[0], wg_per_tile=layout[1])
    texs = [sctx.CreateTexture(t, t.shape[1], t.shape[0], True) for t in vs.textures]
    fbs = [sr.FrameBuffer(sctx, v.width, v.height) for v in vs.views]
    viewset_frame(sr, sctx, fbs, sctx.PrepareDrawList(viewset_calls(sr, vs, fbs, texs)))
    sst = sctx.frame_stats()
    covered = 0
    for fb in fbs:
        covered += int(np.count_nonzero(fb.read_tiles()[1].reshape(-1).view(np.uint32) != 0))
        fb.destroy()
    sctx.Shutdown()
    value = out["runs"][best]["views_per_s"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_views(vs.views, host_textures(sr, vs.textures), args,
                                                 what=f" of {len(vs.views)}, evenly spaced")
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        print(f"[bench] cpu baseline: {out['cpu_baseline']}", file=sys.stderr, flush=True)
    batches = value / len(vs.views)
    algo = sum(d.indices.dtype.itemsize * len(d.indices) + 32 * len(d.vertices) for v in vs.views for d in v.draws)
    algo += sum(8 * v.width * v.height for v in vs.views) + 16 * covered
    out.update({"value": value, "ms_per_step": out["runs"][best]["ms_per_batch"], "best_run": best,
                "mtris_per_s": round(ntris * batches / 1e6, 1), "frags_raster": sst["frags_raster"],
                "frags_source": "gpu stats frame", "gfrags_per_s": round(sst["frags_raster"] * batches / 1e9, 3),
                "roofline": {"bound": "hbm", "kernel": "whole batch", "achieved": round(algo * batches / 1e9, 2),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(algo * batches / 1e9 / HBM_PEAK_GBS, 5), "algorithmic_bytes": algo,
                             "traffic": None}})
    print(json.dumps(out))


def sponza_workload(sr, args):
    """--workload sponza: the 1080p / 1M-triangle frame with every draw on
    the SponzaShader (sun, 16 point lights placed as SponzaScene::Init
    places them, ambient; eight planes per primitive), timed as the plain
    path and as 64-view batches, beside the same frame on UnlitDiffuse; the
    plain frame is compared with the oracle and the oracle is timed on the
    host as the CPU baseline."""
    from softrast_amd import scenes
    t0 = time.time()
    scene = scenes.spheres_scene()
    texs = [sr.TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in scene.textures]
    k = scenes.sponza_constants(0)
    print(f"[bench] sponza scene ready: {scene.num_tris} tris, {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    res = {}
    for name, shader in (("unlit", sr.UnlitDiffuseShader), ("sponza", sr.SponzaShader)):
        # device times from an instrumented pass, frames/s from an uninstrumented one
        _, _, st, ctx, fbs = time_path(sr, scene, texs, 1, sr.CFG_TIMING, sr.ORDER_NATURAL, args.bin_size,
                                       min(args.steps, 100), args.warmup, shader, k)
        for fb in fbs:
            fb.destroy()
        ctx.Shutdown()
        fps, ms, _, ctx, fbs = time_path(sr, scene, texs, 1, 0, sr.ORDER_NATURAL, args.bin_size, args.steps,
                                         args.warmup, shader, k)
        res[name] = {"frames_per_s": round(fps, 2), "ms_per_frame": round(ms, 4),
                     "device_ms": {"frontend": round(statistics.mean(s["ms_frontend"] for s in st), 4),
                                   "backend": round(statistics.mean(s["ms_backend"] for s in st), 4),
                                   "total": round(statistics.mean(s["ms_total"] for s in st), 4)}}
        if name == "sponza":
            colour, depth = fbs[0].read_tiles()
        for fb in fbs:
            fb.destroy()
        ctx.Shutdown()
        print(f"[bench] plain {name}: {fps:.1f} frames/s", file=sys.stderr, flush=True)
    bflags = sr.CFG_GRAPH | (0 if args.no_autotune else sr.CFG_AUTOTUNE)
    _, _, st_t, tctx, tfbs = time_path(sr, scene, texs, args.views, bflags | sr.CFG_TIMING, args.tile_order,
                                       args.bin_size, 10, args.warmup, sr.SponzaShader, k)
    dev_b = statistics.mean(s["ms_total"] for s in st_t)
    for fb in tfbs:
        fb.destroy()
    tctx.Shutdown()
    fps_b, ms_b, st_b, bctx, bfbs = time_path(sr, scene, texs, args.views, bflags, args.tile_order, args.bin_size,
                                              max(20, args.steps // args.views), args.warmup, sr.SponzaShader, k)
    layout = {"bin_size": st_b[-1]["bin_size"], "wg_per_tile": st_b[-1]["wg_per_tile"]}
    # views of the batched launch to check against the oracle
    sampled = {v: bfbs[v].read_tiles() for v in sorted({0, (args.views - 1) // 2, args.views - 1})}
    for fb in bfbs:
        fb.destroy()
    bctx.Shutdown()
    print(f"[bench] batched x{args.views} sponza: {fps_b:.1f} frames/s, layout {layout}", file=sys.stderr, flush=True)
    out = {
        "metric": "frames/s at 1080p/1M tris, SponzaShader (16 point lights), batched+learned",
        "value": round(fps_b, 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_b, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+int32", "data": "synthetic (seeded spheres 0x5EED0002, value-noise textures, lights as "
                                       "SponzaScene::Init ranges)",
        "config": {"workload": "sponza_spheres_1080p_1m", "width": scene.width, "height": scene.height,
                   "triangles_per_view": scene.num_tris, "shader": "SponzaShader",
                   "path": f"batched x{args.views} views, hipGraph, "
                           f"{('natural', 'tri-count', 'learned')[args.tile_order]} tile order"},
        "batched": {"views": args.views, "frames_per_s": round(fps_b, 2), "device_ms_per_launch": round(dev_b, 4),
                    "layout": layout},
        "plain": res,
        "sponza_vs_unlit_plain": round(res["sponza"]["frames_per_s"] / res["unlit"]["frames_per_s"], 3),
    }
    if not args.no_parity or not args.no_cpu_baseline:
        from oracle.oracle import Oracle, Plane, make_texture
        stex = [make_texture(t) for t in texs]
        odraws = []
        for d in scene.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            odraws.append(dict(shader=sr.SHADER_SPONZA, tex=stex[d.texture][0], idx=d.indices, idx_stride=4, pos=vf,
                               pos_stride=32, pos_num=len(vf), attr=vf, attr_stride=32, attr_num=len(vf), uv_offset=6,
                               plane=0, mvp=scene.mvp))
        nthreads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        orc = Oracle(nthreads=nthreads, fast=True)
        orc.set_sponza_constants(k.to_bytes())
        plane = Plane(scene.width, scene.height)
        if not args.no_parity:
            orc.clear(plane, 0)
            orc.render(odraws, [plane])
            dd = int(np.count_nonzero(plane.depth.view(np.uint32) != depth.reshape(-1).view(np.uint32)))
            dc = int(np.count_nonzero(plane.colour != colour.reshape(-1)))
            out["parity"] = {"depth_px_diff": dd, "rgba8_px_diff": dc, "vs": "oracle", "frame": "plain path"}
            # sampled views of the batched launch `value` measures, against the
            # oracle's render of each view's camera
            bp = {"views": sorted(sampled), "depth_px_diff": 0, "rgba8_px_diff": 0, "vs": "oracle"}
            for v, (bcol, bdep) in sorted(sampled.items()):
                vdraws = [dict(d, mvp=batched_view_mvp(scene, args.views, v)) for d in odraws]
                vplane = Plane(scene.width, scene.height)
                orc.clear(vplane, 0)
                orc.render(vdraws, [vplane])
                bp["depth_px_diff"] += int(np.count_nonzero(vplane.depth.view(np.uint32) !=
                                                            bdep.reshape(-1).view(np.uint32)))
                bp["rgba8_px_diff"] += int(np.count_nonzero(vplane.colour != bcol.reshape(-1)))
            if sampled:
                out["batched"]["parity"] = bp
        if not args.no_cpu_baseline:
            times = []
            for _ in range(3 + args.cpu_frames):
                t = time.perf_counter()
                orc.clear(plane, 0)
                orc.render(odraws, [plane])
                times.append(time.perf_counter() - t)
            med = statistics.median(times[3:])
            out["cpu_baseline"] = {"value": round(1.0 / med, 3), "unit": "frames/s", "cores": nthreads,
                                   "kind": "port",
                                   "sample": f"same 1080p/1M-tri SponzaShader frame (clear+render), median of "
                                             f"{args.cpu_frames} frames after 3 warm-up, {nthreads} threads, "
                                             f"oracle -O3 AVX2/FMA build"}
            out["speedup_vs_cpu"] = round(fps_b * med, 1)
        orc.close()
    print(json.dumps(out))


def present_rates(sr, args):
    """--present: frames/s of the plain path when every frame's pixels go to
    host memory (linear RGBA8), by the synchronous Blit and by BlitAsync
    overlapped with rendering the next frame; beside it the plain rate
    without a blit, in the same process."""
    from softrast_amd import scenes
    scene = scenes.spheres_scene()
    texs = [sr.TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in scene.textures]
    out = {"metric": "frames/s at 1080p/1M tris with the frame's pixels in host memory", "unit": "frames/s",
           "config": {"workload": "spheres_1080p_1m", "bytes_per_frame": scene.width * scene.height * 4}}
    for mode in ("none", "sync", "async"):
        ctx = sr.RenderContext(flags=0)
        fb = sr.FrameBuffer(ctx, scene.width, scene.height)
        draws = ctx.PrepareDrawList(make_calls(sr, fb, scene, texs, scene.mvp))
        host = [np.empty((scene.height, scene.width), np.uint32) for _ in range(2)]  # pageable, as a caller's

        def frame(i):
            ctx.BeginFrame()
            ctx.ClearFrameBuffer(fb, 0)
            ctx.DrawIndexedList(draws)
            ctx.EndFrame()
            if mode == "sync":
                ctx.Blit(fb, host[i & 1])
            elif mode == "async":
                ctx.BlitWait(fb)  # the previous frame's copy, overlapped with this frame's rendering
                ctx.BlitAsync(fb, host[i & 1])

        for i in range(args.warmup):
            frame(i)
        t0 = time.perf_counter()
        for i in range(args.steps):
            frame(i)
        if mode == "async":
            ctx.BlitWait(fb)
        out[mode] = round(args.steps / (time.perf_counter() - t0), 2)
        print(f"[bench] present {mode}: {out[mode]} frames/s", file=sys.stderr, flush=True)
        fb.destroy()
        ctx.Shutdown()
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              f"(views shard across GPUs with no exchange; the multi-GPU design is described, not built)",
              file=sys.stderr)
        sys.exit(2)
    import softrast_amd as sr
    from softrast_amd import scenes
    if args.sched_study:
        return sched_study(sr, args)
    if args.present:
        return present_rates(sr, args)
    if args.workload == "sponza":
        return sponza_workload(sr, args)
    if args.workload != "spheres":
        return batch_workload(sr, args)

    t0 = time.time()
    scene = scenes.spheres_scene()
    texs = [sr.TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in scene.textures]
    print(f"[bench] scene ready: {scene.num_tris} tris, {len(scene.draws)} draws, {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    if args.no_plain:
        return profile_batched(sr, scene, texs, args)
    # ---- plain path: one view per EndFrame, no graph, natural tile order ----
    # Frames/s of every path come from uninstrumented contexts; the device
    # times (HIP events, SRG_CFG_TIMING) from a second, instrumented pass.
    fps_plain, ms_plain, _, ctx, fbs = time_path(sr, scene, texs, 1, 0, sr.ORDER_NATURAL, args.bin_size, args.steps,
                                                 args.warmup)
    print(f"[bench] plain: {fps_plain:.1f} frames/s", file=sys.stderr, flush=True)
    colour, depth = fbs[0].read_tiles()
    covered = int(np.count_nonzero(depth.reshape(-1).view(np.uint32) != 0))
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()
    _, _, st_plain, ctx, fbs = time_path(sr, scene, texs, 1, sr.CFG_TIMING, sr.ORDER_NATURAL, args.bin_size,
                                         min(args.steps, 100), args.warmup)
    be_ms = statistics.mean(s["ms_backend"] for s in st_plain)
    fe_ms = statistics.mean(s["ms_frontend"] for s in st_plain)
    dev_ms = statistics.mean(s["ms_total"] for s in st_plain)
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()

    # algorithmic bytes per frame: s_idx*3*T + 32*V + 8*W*H + 16*P_cov; the back end's share
    # is the colour + depth written once and one bilinear RGBA8 quad per covered pixel
    fe_bytes = 4 * 3 * scene.num_tris + 32 * sum(len(d.vertices) for d in scene.draws)
    be_bytes = 8 * scene.width * scene.height + 16 * covered
    achieved = be_bytes / (be_ms * 1e-3) / 1e9

    # ---- batched path ----
    fps_batched = ms_batched = dev_batched = None
    batched_runs = {}
    sampled = {}  # view -> (colour, depth) tiles of the headline batched path, for parity
    be_ms_batched = None
    if not args.no_batched:
        bflags = sr.CFG_GRAPH | (0 if args.no_autotune else sr.CFG_AUTOTUNE)
        orders = [args.tile_order] + ([sr.ORDER_TRICOUNT] if args.tile_order != sr.ORDER_TRICOUNT else [])
        for order in orders:
            # device time per launch from an instrumented context first
            _, _, st_t, tctx, tfbs = time_path(sr, scene, texs, args.views, bflags | sr.CFG_TIMING, order,
                                               args.bin_size, 10, args.warmup)
            dev_ms_launch = statistics.mean(s["ms_total"] for s in st_t)
            for fb in tfbs:
                fb.destroy()
            tctx.Shutdown()
            fps, ms, st_b, bctx, bfbs = time_path(sr, scene, texs, args.views, bflags, order, args.bin_size,
                                                  max(20, args.steps // args.views), args.warmup)
            run = {"tile_order": ("natural", "tricount", "learned")[order], "frames_per_s": round(fps, 2),
                   "ms_per_launch": round(ms, 4), "device_ms_per_launch": round(dev_ms_launch, 4),
                   "layout": {"bin_size": st_b[-1]["bin_size"], "wg_per_tile": st_b[-1]["wg_per_tile"]},
                   "graph_captures_in_timed_launches": sum(s["graph_captures"] for s in st_b)}
            if bflags & sr.CFG_AUTOTUNE:
                t = bctx.tuning()
                run["autotune_device_ms"] = {f"bin{b}_wg{w}": round(v, 4) for (b, w), v in t["ms"].items()}
            if order == sr.ORDER_LEARNED:
                run["sched_err_log2"] = round(statistics.mean(s["sched_err"] for s in st_b), 4)
            batched_runs[run["tile_order"]] = run
            if order == args.tile_order:
                fps_batched, ms_batched, dev_batched = fps, ms, run["device_ms_per_launch"]
                for v in sorted({0, (args.views - 1) // 2, args.views - 1}):
                    sampled[v] = bfbs[v].read_tiles()
            for fb in bfbs:
                fb.destroy()
            bctx.Shutdown()
            if order == args.tile_order:
                # the back-end kernel of this path alone: HIP events around it on
                # the library's stream, in launches without a graph on the same
                # layout and order
                lay = run["layout"]
                _, _, st_k, kctx, kfbs = time_path(sr, scene, texs, args.views, sr.CFG_TIMING, order, lay["bin_size"],
                                                   10, args.warmup, wg_per_tile=lay["wg_per_tile"])
                be_ms_batched = statistics.mean(s["ms_backend"] for s in st_k)
                run["backend_ms_per_launch"] = round(be_ms_batched, 4)
                for fb in kfbs:
                    fb.destroy()
                kctx.Shutdown()
            print(f"[bench] batched x{args.views} {run['tile_order']}: {fps:.1f} frames/s, layout {run['layout']}",
                  file=sys.stderr, flush=True)

    value = fps_batched if fps_batched else fps_plain
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batched if ms_batched else ms_plain, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int32",
        "data": "synthetic (seeded spheres 0x5EED0002, value-noise textures)",
        "config": {"workload": "spheres_1080p_1m", "width": scene.width, "height": scene.height,
                   "triangles_per_view": scene.num_tris, "draws_per_view": len(scene.draws),
                   "textures": "8x2048^2 RGBA8 + mips",
                   "path": (f"batched x{args.views} views, hipGraph, "
                            f"{'autotuned layout, ' if not args.no_autotune else ''}"
                            f"{('natural', 'tri-count', 'learned')[args.tile_order]} tile order"
                            if fps_batched else "plain single-view")},
        "plain": {"frames_per_s": round(fps_plain, 2), "ms_per_frame": round(ms_plain, 4),
                  "device_ms": {"frontend": round(fe_ms, 4), "backend": round(be_ms, 4), "total": round(dev_ms, 4)}},
        "mtris_per_s": round(scene.num_tris * value / 1e6, 1),
        "frame_algorithmic_bytes": fe_bytes + be_bytes,
    }

    def valu_issue(kernel, profile, measured_ms):
        # the back end is bound by instruction issue, not HBM: its VALU wave
        # instructions per launch (SQ_INSTS_VALU, profiles/) at 4 cycles each
        # on 1,024 SIMDs at 2.4 GHz, against the live kernel time
        valu = load_profile_value(profile, "valu_insts_per_launch")
        if not valu:
            return None
        floor_us = valu * 4.0 / 1024 / 2.4e3
        return {"kernel": kernel, "valu_insts_per_launch": valu, "issue_floor_us": round(floor_us, 1),
                "measured_us": round(measured_ms * 1e3, 1), "frac": round(floor_us / (measured_ms * 1e3), 3),
                "source": f"profiles/{profile}"}

    plain_kernel = f"k_backend<{st_plain[-1]['bin_size']}, {st_plain[-1]['bin_size'] // st_plain[-1]['wg_per_tile']}>"
    out["plain"]["roofline"] = {
        "bound": "hbm", "kernel": plain_kernel + " (one view)", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "algorithmic_bytes": be_bytes,
        "traffic": load_traffic("k_backend_spheres_1080p_1m_traffic.md")}
    vi = valu_issue(plain_kernel, "k_backend_spheres_1080p_1m_sq_counters.md", be_ms)
    if vi:
        out["plain"]["valu_issue"] = vi
    if fps_batched:
        out["batched"] = {"views": args.views, "frames_per_s": round(fps_batched, 2),
                          "ms_per_launch": round(ms_batched, 4), "device_ms_per_launch": round(dev_batched, 4),
                          "vs_plain": round(fps_batched / fps_plain, 3), "runs": batched_runs}
        per_view, covered_views = batched_stats(sr, scene, texs, args.views, args.bin_size)
        out["batched"]["frags_raster_per_view"] = int(per_view)
        # the dominant kernel of the path `value` measures: the batched back end,
        # its algorithmic bytes summed over the launch's views
        lay = batched_runs[("natural", "tricount", "learned")[args.tile_order]]["layout"]
        kname = f"k_backend<{lay['bin_size']}, {lay['bin_size'] // lay['wg_per_tile']}>"
        bb = sum(8 * scene.width * scene.height + 16 * c for c in covered_views)
        ach_b = bb / (be_ms_batched * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": f"{kname} (batched x{args.views})", "achieved": round(ach_b, 2),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach_b / HBM_PEAK_GBS, 5),
                           "algorithmic_bytes": bb, "launch_us": round(be_ms_batched * 1e3, 1),
                           "traffic": load_traffic("k_backend_spheres_1080p_1m_x64_traffic.md"),
                           "traffic_source": "profiles/k_backend_spheres_1080p_1m_x64_traffic.md"}
        vi = valu_issue(kname, "k_backend_spheres_1080p_1m_x64_sq_counters.md", be_ms_batched)
        if vi:
            out["valu_issue"] = vi
    else:
        out["roofline"] = dict(out["plain"]["roofline"])
        if "valu_issue" in out["plain"]:
            out["valu_issue"] = out["plain"]["valu_issue"]

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
            ostats = orc.render(odraws, [plane])
            out["frags_raster"] = ostats["frags_raster"]
            out["frags_source"] = "oracle"
            dd = int(np.count_nonzero(plane.depth.view(np.uint32) != depth.reshape(-1).view(np.uint32)))
            dc = int(np.count_nonzero(plane.colour != colour.reshape(-1)))
            out["parity"] = {"depth_px_diff": dd, "rgba8_px_diff": dc, "vs": "oracle", "frame": "plain path"}
        if not args.no_cpu_baseline:
            times = []
            for _ in range(5 + args.cpu_frames):
                t = time.perf_counter()
                orc.clear(plane, 0)
                orc.render(odraws, [plane])
                times.append(time.perf_counter() - t)
            med = statistics.median(times[5:])
            out["cpu_baseline"] = {"value": round(1.0 / med, 3), "unit": "frames/s", "cores": nthreads,
                                   "kind": "port",
                                   "sample": f"same 1080p/1M-tri frame (clear+render), median of {args.cpu_frames} "
                                             f"frames after 5 warm-up, {nthreads} threads, oracle -O3 AVX2/FMA build"}
            out["speedup_vs_cpu"] = round(value * med, 1)
            # the host the baseline ran on, and the same frame on one thread
            one = Oracle(nthreads=1, fast=True)
            t1 = []
            for _ in range(3):
                t = time.perf_counter()
                one.clear(plane, 0)
                one.render(odraws, [plane])
                t1.append(time.perf_counter() - t)
            one.close()
            out["cpu_host"] = {"model": cpu_model(), "nproc": os.cpu_count(), "threads_used": nthreads,
                               "one_thread_frames_per_s": round(1.0 / statistics.median(t1), 3)}
        orc.close()

    if "frags_raster" not in out:  # no oracle frame: count on the GPU in an untimed frame
        sctx = sr.RenderContext(flags=sr.CFG_STATS, bin_size=args.bin_size)
        sfb = sr.FrameBuffer(sctx, scene.width, scene.height)
        sctx.BeginFrame()
        sctx.ClearFrameBuffer(sfb, 0)
        for c in make_calls(sr, sfb, scene, texs, scene.mvp):
            sctx.DrawIndexed(c)
        sctx.EndFrame()
        out["frags_raster"] = sctx.frame_stats()["frags_raster"]
        out["frags_source"] = "gpu stats frame"
        sfb.destroy()
        sctx.Shutdown()
    out["plain"]["gfrags_per_s"] = round(out["frags_raster"] * fps_plain / 1e9, 3)
    per_view = out.get("batched", {}).get("frags_raster_per_view", out["frags_raster"])
    out["gfrags_per_s"] = round(per_view * value / 1e9, 3)  # of the path `value` measures
    print(json.dumps(out))


if __name__ == "__main__":
    main()
