"""srg_end_frame_noalloc, the EndFrame that allocates nothing (no
counterpart in the reference): through the C ABI its frames equal
srg_end_frame's bit for bit (coverage, primitive order, depth, RGBA8,
rasterised-fragment count) on consecutive calls, the second with a smaller
frame, in contexts that also end frames the ordinary way; a frame larger
than the fixed scratch is refused with SRG_ERR_OVERFLOW before any work and
leaves the context usable."""
import sys

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu

SRG_ERR_OVERFLOW = -5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _frame(sc, tex, ctx, fb, noalloc):
    st = helpers.gpu_render(sc, tex, ctx=ctx, fb=fb, noalloc=noalloc)
    col, dep, pid = fb.read_tiles(primid=True)
    return col.reshape(-1), dep.reshape(-1), pid.reshape(-1), st


def _same(a, b, name):
    for k, (x, y) in enumerate(zip(a[:3], b[:3])):
        n = int(np.count_nonzero(x.view(np.uint32) != y.view(np.uint32)))
        assert n == 0, f"{name}: plane {k} differs in {n} pixels"
    assert a[3]["frags_raster"] == b[3]["frags_raster"], name
    assert a[3]["prims"] == b[3]["prims"], name


@pytest.mark.parametrize("tile_order", [0, 2])
def test_noalloc_frames_equal_end_frame(tile_order):
    import softrast_amd as sr
    big = helpers.random_scene(301, 120000, 640, 360, clip_heavy=True, ndraws=3)
    small = helpers.random_scene(302, 9000, 320, 200, ndraws=2)
    flags = sr.CFG_KEEP_PRIMID | sr.CFG_STATS
    ref_ctx = sr.RenderContext(flags=flags)
    na_ctx = sr.RenderContext(flags=flags, tile_order=tile_order)
    fbs = {}
    for step, (name, sc, noalloc) in enumerate((("big", big, True), ("small", small, True),
                                                ("small", small, False), ("big", big, True))):
        _say(f"[noalloc] order={tile_order} step={step} frame={name} tris={sc.num_tris} noalloc={noalloc}")
        tex = helpers.build_textures(sc)
        if name not in fbs:
            fbs[name] = (sr.FrameBuffer(ref_ctx, sc.width, sc.height), sr.FrameBuffer(na_ctx, sc.width, sc.height))
        ref = _frame(sc, tex, ref_ctx, fbs[name][0], False)
        got = _frame(sc, tex, na_ctx, fbs[name][1], noalloc)
        assert ref[3]["prims"] > 0
        _same(ref, got, f"{name} step {step}")
    for a, b in fbs.values():
        a.destroy()
        b.destroy()
    ref_ctx.Shutdown()
    na_ctx.Shutdown()


def test_noalloc_refuses_a_frame_beyond_the_scratch():
    import softrast_amd as sr
    huge = helpers.random_scene(303, 600000, 640, 360)
    small = helpers.random_scene(304, 5000, 256, 160)
    flags = sr.CFG_KEEP_PRIMID | sr.CFG_STATS
    ctx = sr.RenderContext(flags=flags)
    _say(f"[noalloc] case=too_big tris={huge.num_tris}")
    tex = helpers.build_textures(huge)
    fb = sr.FrameBuffer(ctx, huge.width, huge.height)
    with pytest.raises(sr.SrgError) as e:
        helpers.gpu_render(huge, tex, ctx=ctx, fb=fb, noalloc=True)
    assert e.value.status == SRG_ERR_OVERFLOW
    fb.destroy()
    _say("[noalloc] case=after_refusal")
    tex = helpers.build_textures(small)
    fb = sr.FrameBuffer(ctx, small.width, small.height)
    got = _frame(small, tex, ctx, fb, True)
    ref_plane, rstats = helpers.oracle_render(small, tex, nthreads=8)
    d = helpers.compare(ref_plane, got[0], got[1], got[2])
    assert not any(d.values()), d
    assert got[3]["frags_raster"] == rstats["frags_raster"]
    fb.destroy()
    ctx.Shutdown()
