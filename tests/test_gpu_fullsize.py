"""GPU <-> oracle parity at the workloads' full stated sizes.

* the benchmark's batched launch: 64 views of the 1920x1080 / 1M-triangle
  sphere scene (cameras turned by up to +-2 degrees) in one hipGraph-captured
  EndFrame, autotuned layout, learned tile order once trained;
* the clustered 3840x2160 / 4M-triangle frame with its 16 device-built
  4096^2 textures (13 mips), autotuned layout, learned order;
* the 256 mixed-resolution views (20.3M triangles, 1024^2 textures) in one
  EndFrame under the autotuner; more than 64 views exercise the back end's
  binary search for a bin's view;
* the combined batch: 64 views of 2560x1440 x 2M triangles (128M in one
  EndFrame), learned order retrained while rendering.

Every view is compared over its whole tile-padded plane: coverage, primitive
order id, depth bits and RGBA8 identical, and the rasterised-fragment count
equal. The oracle renders one view at a time (tests.helpers.
oracle_render_views_each); textures built on the device are read back for it
(their bytes equal the host builder's, tests/test_gpu_texture.py)."""
import math
import sys
import time

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_all(vs, planes, fbs, what):
    bad = []
    for vi, (fb, p) in enumerate(zip(fbs, planes)):
        col, dep, pid = fb.read_tiles(primid=True)
        d = helpers.compare(p, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
        if any(d.values()):
            bad.append((vi, vs.views[vi].width, vs.views[vi].height, d))
    _say(f"[fullsize] {what}: {len(fbs)} views compared, {len(bad)} differ")
    assert not bad, (what, bad[:4])


def _device_textures(sr, ctx, images):
    """Textures built on the device for the GPU frame, and their bytes read
    back as TextureData for the oracle."""
    dev = [ctx.CreateTexture(t, t.shape[1], t.shape[0], True) for t in images]
    return dev, [t.read() for t in dev]


def _frames(sr, ctx, vs, dev, fbs, n, what):
    st = None
    for k in range(n):
        t = time.perf_counter()
        st = helpers.gpu_render_views(vs, dev, ctx, fbs)
        _say(f"[fullsize] {what} frame={k} layout={st['bin_size']}x{st['wg_per_tile']} "
             f"trained={st['sched_trained']} {1e3 * (time.perf_counter() - t):.1f} ms")
    return st


def test_batched_spheres_64_views_1080p_1m():
    import softrast_amd as sr
    from softrast_amd import scenes
    sc = scenes.spheres_scene()
    views = []
    for v in range(64):
        t = math.radians((v - 31.5) * 4.0 / 63)
        mvp = scenes.camera_mvp((0.0, 0.0, 0.0), (math.sin(t), 0.0, math.cos(t)), 60.0, sc.width / sc.height)
        views.append(scenes.Scene(sc.width, sc.height, mvp, sc.draws, [], {}))
    vs = scenes.ViewSet(views, sc.textures)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_GRAPH | sr.CFG_AUTOTUNE,
                           tile_order=sr.ORDER_LEARNED, autotune_frames=1, sched_warmup_frames=2)
    dev, host = _device_textures(sr, ctx, vs.textures)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    _say(f"[fullsize] case=batched_spheres views=64 tris={vs.num_tris}")
    st = _frames(sr, ctx, vs, dev, fbs, 8 + 5, "batched_spheres")
    assert st["sched_trained"] and ctx.tuning()["state"] == "chosen"
    planes, rstats = helpers.oracle_render_views_each(vs, host)
    _compare_all(vs, planes, fbs, "batched_spheres")
    assert st["frags_raster"] == rstats["frags_raster"]
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


def test_clustered_4k_4m_device_textures():
    import softrast_amd as sr
    from softrast_amd import scenes
    sc = scenes.clustered_scene()
    assert sc.num_tris == 4_000_000 and len(sc.textures) == 16 and sc.textures[0].shape[:2] == (4096, 4096)
    vs = scenes.ViewSet([scenes.Scene(sc.width, sc.height, sc.mvp, sc.draws, [], {})], sc.textures)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_AUTOTUNE, tile_order=sr.ORDER_LEARNED,
                           autotune_frames=1, sched_warmup_frames=2)
    dev, host = _device_textures(sr, ctx, vs.textures)
    assert host[0].m_numMips == 13
    fb = sr.FrameBuffer(ctx, sc.width, sc.height)
    _say(f"[fullsize] case=clustered_4k tris={sc.num_tris}")
    st = _frames(sr, ctx, vs, dev, [fb], 8 + 5, "clustered_4k")
    assert st["sched_trained"] and ctx.tuning()["state"] == "chosen"
    planes, rstats = helpers.oracle_render_views_each(vs, host)
    _compare_all(vs, planes, [fb], "clustered_4k")
    assert st["frags_raster"] == rstats["frags_raster"]
    fb.destroy()
    ctx.Shutdown()


def test_mixed_views_256_autotuned():
    import softrast_amd as sr
    from softrast_amd import scenes
    vs = scenes.mixed_views()
    assert len(vs.views) == 256 and vs.num_tris > 20_000_000
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_GRAPH | sr.CFG_AUTOTUNE,
                           tile_order=sr.ORDER_TRICOUNT, autotune_frames=1)
    dev, host = _device_textures(sr, ctx, vs.textures)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    _say(f"[fullsize] case=mixed_views views=256 tris={vs.num_tris}")
    st = _frames(sr, ctx, vs, dev, fbs, 8 + 1, "mixed_views")
    assert ctx.tuning()["state"] == "chosen"
    planes, rstats = helpers.oracle_render_views_each(vs, host)
    _compare_all(vs, planes, fbs, "mixed_views")
    assert st["frags_raster"] == rstats["frags_raster"]
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


def test_combined_64_views_1440p_2m_learned():
    import softrast_amd as sr
    from softrast_amd import scenes
    vs = scenes.combined_views()
    assert len(vs.views) == 64 and vs.num_tris == 128_000_000
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_GRAPH, tile_order=sr.ORDER_LEARNED,
                           sched_warmup_frames=2, sched_retrain_interval=2)
    dev, host = _device_textures(sr, ctx, vs.textures)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    _say(f"[fullsize] case=combined views=64 tris={vs.num_tris}")
    st = _frames(sr, ctx, vs, dev, fbs, 5, "combined")
    assert st["sched_trained"]
    planes, rstats = helpers.oracle_render_views_each(vs, host)
    _compare_all(vs, planes, fbs, "combined")
    assert st["frags_raster"] == rstats["frags_raster"]
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()
