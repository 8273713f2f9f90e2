"""Edge cases through the C ABI: empty frames, zero and ragged index counts,
indices past the vertex buffer, non-finite and huge vertex coordinates, and
framebuffers that are not a whole number of tiles. Images are compared with
the oracle under the same bar as test_gpu_parity (bit-exact)."""
import sys

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(sc, name, **kw):
    _say(f"[edge] case={name} size={sc.width}x{sc.height} tris={sc.num_tris}")
    tex = helpers.build_textures(sc)
    ref, rstats = helpers.oracle_render(sc, tex, nthreads=8)
    col, dep, pid, st = helpers.gpu_render(sc, tex, **kw)
    d = helpers.compare(ref, col, dep, pid)
    _say(f"[edge] case={name} diffs={d}")
    assert not any(d.values()), d
    assert st["frags_raster"] == rstats["frags_raster"]
    return ref, st


def test_empty_frames():
    """BeginFrame/EndFrame with nothing drawn, then with only a clear: the
    clear lands (colour = low byte replicated, depth 0), nothing else runs."""
    import softrast_amd as sr
    ctx = sr.RenderContext(flags=sr.CFG_STATS)
    fb = sr.FrameBuffer(ctx, 200, 100)
    ctx.BeginFrame()
    ctx.EndFrame()
    ctx.BeginFrame()
    ctx.ClearFrameBuffer(fb, 0xA1B2C3D4)
    ctx.EndFrame()
    st = ctx.frame_stats()
    assert st["tris_in"] == 0 and st["frags_raster"] == 0
    col, dep = fb.read_tiles()
    assert np.all(col == 0xD4D4D4D4) and np.all(dep == 0.0)
    fb.destroy()
    ctx.Shutdown()


def test_zero_and_ragged_index_counts():
    """Draws of 0, 1, 2 and 3k+1 / 3k+2 indices: the trailing indices that do
    not make a triangle are ignored, as num / 3 triangles are drawn."""
    base = helpers.random_scene(90, 1200, 256, 192, ndraws=4)  # 900 indices per draw
    for k, n in enumerate((0, 2, 301, 599)):
        base.draws[k].indices = base.draws[k].indices[:n].copy()
    _check(base, "ragged_counts")


def test_index_out_of_range_is_an_error():
    """An index at or past the vertex count makes EndFrame fail with
    SRG_ERR_INDEX_RANGE (the front end checks every index before it reads a
    vertex); the context renders the next, valid frame correctly."""
    import softrast_amd as sr
    sc = helpers.random_scene(91, 1500, 256, 192)
    bad = helpers.random_scene(91, 1500, 256, 192)
    bad.draws[0].indices = bad.draws[0].indices.copy()
    bad.draws[0].indices[777] = len(bad.draws[0].vertices)
    tex = helpers.build_textures(sc)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    fb = sr.FrameBuffer(ctx, sc.width, sc.height)
    with pytest.raises(sr.SrgError) as e:
        helpers.gpu_render(bad, tex, ctx=ctx, fb=fb)
    assert e.value.status == -4
    ref, _ = helpers.oracle_render(sc, tex, nthreads=8)
    helpers.gpu_render(sc, tex, ctx=ctx, fb=fb)
    col, dep, pid = fb.read_tiles(primid=True)
    d = helpers.compare(ref, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
    assert not any(d.values()), d
    fb.destroy()
    ctx.Shutdown()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nonfinite_and_huge_vertices(seed):
    """5% of the vertex coordinates replaced by NaN, +-Inf, +-1e30, 3e38 or
    1e-30: the reference's conversions (INT_MIN on NaN or overflow), clip
    codes and comparisons decide what is drawn, on both sides alike."""
    sc = helpers.random_scene(80 + seed, 3000, 320, 240, clip_heavy=True)
    v = sc.draws[0].vertices.view(np.float32).reshape(-1, 8)
    rng = np.random.default_rng(seed)
    bad = np.nonzero(rng.random(len(v)) < 0.05)[0]
    vals = np.array([np.nan, np.inf, -np.inf, 1e30, -1e30, 3e38, 1e-30], np.float32)
    v[bad, rng.integers(0, 3, len(bad))] = rng.choice(vals, len(bad))
    _check(sc, f"nonfinite_{seed}")


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (65, 65), (130, 61), (64, 1)])
def test_partial_tile_framebuffers(w, h):
    """Framebuffers smaller than a tile or with partial edge tiles: the
    bbox clamps and the tile padding behave as the reference's."""
    sc = helpers.random_scene(92, 6000, w, h, clip_heavy=True)
    _check(sc, f"fb_{w}x{h}")


def test_clear_list_clears_every_view():
    """ClearFrameBuffers (srg_clear_n) over a batch of framebuffers equals a
    ClearFrameBuffer per framebuffer: after an empty frame each holds the
    low byte replicated and depth 0; a colour-only clear keeps the depth."""
    import softrast_amd as sr
    ctx = sr.RenderContext(flags=sr.CFG_STATS)
    fbs = [sr.FrameBuffer(ctx, w, h) for w, h in ((64, 64), (130, 61), (200, 100))]
    cl = ctx.PrepareClearList(fbs)
    ctx.BeginFrame()
    ctx.ClearFrameBuffers(cl, 0x11223344)
    ctx.EndFrame()
    for fb in fbs:
        col, dep = fb.read_tiles()
        assert np.all(col == 0x44444444) and np.all(dep == 0.0)
        fb.write_tiles(col, np.full_like(dep, 0.5))
    ctx.BeginFrame()
    ctx.ClearFrameBuffers(cl, 0x000000AB, clearDepth=False)
    ctx.EndFrame()
    for fb in fbs:
        col, dep = fb.read_tiles()
        assert np.all(col == 0xABABABAB) and np.all(dep == 0.5)
        fb.destroy()
    ctx.Shutdown()


SRG_ERR_HANDLE = -9


def _draw_scene(ctx, sc, fb, texdata):
    import softrast_amd as sr
    for d in sc.draws:
        call = sr.DrawCall().SetFrameBuffer(fb).SetMVP(sc.mvp)
        call.SetIndexBuffer(d.indices, d.indices.dtype.itemsize, len(d.indices))
        call.SetPositionBuffer(d.vertices, helpers.ATTR_STRIDE, len(d.vertices))
        call.SetAttributeBuffer(d.vertices, helpers.ATTR_STRIDE, len(d.vertices), d.uv_offset)
        call.SetPixelShader(sr.UnlitDiffuseShader, texdata[d.texture] if d.texture >= 0 else None)
        ctx.DrawIndexed(call)


def test_released_handles_fail_the_frame():
    """A vertex buffer released, a texture released or a framebuffer destroyed
    between DrawIndexed and EndFrame makes EndFrame fail with SRG_ERR_HANDLE
    before any kernel reads the released memory; a clear list holding a
    destroyed framebuffer fails the same way and clears none of its other
    framebuffers. The context then renders a valid frame bit-exactly."""
    import softrast_amd as sr
    sc = helpers.random_scene(93, 1500, 256, 192)
    tex = helpers.build_textures(sc)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    fb = sr.FrameBuffer(ctx, sc.width, sc.height)

    _say("[edge] case=released_buffer")
    ctx.BeginFrame()
    _draw_scene(ctx, sc, fb, tex)
    ctx.invalidate(sc.draws[0].vertices)
    with pytest.raises(sr.SrgError) as e:
        ctx.EndFrame()
    assert e.value.status == SRG_ERR_HANDLE

    _say("[edge] case=released_texture")
    dtex = [sr.DeviceTexture(ctx, t, t.shape[1], t.shape[0]) for t in sc.textures]
    ctx.BeginFrame()
    _draw_scene(ctx, sc, fb, dtex)
    dtex[0].release()
    with pytest.raises(sr.SrgError) as e:
        ctx.EndFrame()
    assert e.value.status == SRG_ERR_HANDLE

    _say("[edge] case=destroyed_framebuffer")
    fb2 = sr.FrameBuffer(ctx, sc.width, sc.height)
    ctx.BeginFrame()
    _draw_scene(ctx, sc, fb2, tex)
    fb2.destroy()
    with pytest.raises(sr.SrgError) as e:
        ctx.EndFrame()
    assert e.value.status == SRG_ERR_HANDLE

    _say("[edge] case=clear_list_with_destroyed_framebuffer")
    keep, gone = sr.FrameBuffer(ctx, 130, 61), sr.FrameBuffer(ctx, 64, 64)
    cl = ctx.PrepareClearList([keep, gone])
    col, dep = keep.read_tiles()
    keep.write_tiles(np.full_like(col, 0x01020304), np.full_like(dep, 0.5))
    gone.destroy()
    ctx.BeginFrame()
    with pytest.raises(sr.SrgError) as e:
        ctx.ClearFrameBuffers(cl, 0x11223344)
    assert e.value.status == SRG_ERR_HANDLE
    ctx.EndFrame()
    col, dep = keep.read_tiles()
    assert np.all(col == 0x01020304) and np.all(dep == 0.5)

    _say("[edge] case=valid_frame_after_errors")
    ref, _ = helpers.oracle_render(sc, tex, nthreads=8)
    helpers.gpu_render(sc, tex, ctx=ctx, fb=fb)
    col, dep, pid = fb.read_tiles(primid=True)
    d = helpers.compare(ref, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
    assert not any(d.values()), d
    for f in (fb, keep):
        f.destroy()
    ctx.Shutdown()


@pytest.mark.parametrize("seed", [0, 1])
def test_tiny_lattice_triangles(seed):
    """20,000 triangles of up to a few px^2 with vertices on a half-pixel
    lattice (identity MVP, 256x256, so edges pass exactly through pixel
    centres and edge values hit 0): many survivors cover one pixel or none.
    The front end's no-coverage skip and the back end's exact-zero handling
    give the oracle's image and rasterised-fragment count."""
    from softrast_amd import scenes
    rng = np.random.default_rng(300 + seed)
    n, W = 20000, 256
    centre = rng.integers(16, 2 * W - 16, (n, 1, 2)) * 0.5
    pts = centre + rng.integers(-3, 4, (n, 3, 2)) * 0.5
    v = np.zeros((n * 3, 8), np.float32)
    v[:, 0] = ((pts[..., 0] - 128.0) / 128.0).reshape(-1)
    v[:, 1] = ((128.0 - pts[..., 1]) / 128.0).reshape(-1)
    v[:, 2] = np.repeat(rng.uniform(0.1, 0.9, n), 3)
    v[:, 3:6] = rng.normal(0, 1, (n * 3, 3))
    v[:, 6:8] = rng.uniform(-2, 2, (n * 3, 2))
    assert np.array_equal(v[:, 0] * 128.0 + 128.0, pts[..., 0].reshape(-1).astype(np.float32))
    draw = scenes.Draw(v.view(scenes.VERTEX_DTYPE).reshape(-1), np.arange(n * 3, dtype=np.uint32), 0, 0, 6)
    sc = scenes.Scene(W, W, np.eye(4, dtype=np.float32).reshape(-1), [draw],
                      [rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)], {"config": "lattice"})
    _check(sc, f"lattice_{seed}")
