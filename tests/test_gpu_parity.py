"""GPU <-> oracle parity through the C ABI (libsrgpu.so).

Bar (per pixel, whole tile-padded plane): coverage, primitive order id and
depth bits identical; RGBA8 identical (the contract allows 1 LSB; the test
asserts exact and reports the count beyond 1 LSB separately).
Differing-pixel counts are printed for every case.
"""
import sys

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(scene, name, clear=0, plane_init=None, nthreads=8, **kw):
    _say(f"[parity] case={name} size={scene.width}x{scene.height} tris={scene.num_tris}")
    tex = helpers.build_textures(scene)
    ref, rstats = helpers.oracle_render(scene, tex, nthreads=nthreads, clear=clear, plane_init=plane_init)
    col, dep, pid, gstats = helpers.gpu_render(scene, tex, clear=clear, plane_init=plane_init, **kw)
    d = helpers.compare(ref, col, dep, pid)
    _say(f"[parity] case={name} diffs={d} frags oracle={rstats['frags_raster']} gpu={gstats['frags_raster']} "
         f"entries oracle={rstats['bin_entries']} gpu={gstats['bin_entries']}")
    assert d["coverage"] == 0
    assert d["primid"] == 0
    assert d["depth"] == 0
    assert d["colour_gt1"] == 0
    assert d["colour"] == 0
    assert gstats["frags_raster"] == rstats["frags_raster"]
    assert gstats["bin_entries"] == rstats["bin_entries"]
    return ref, gstats


def test_cube_grid_1280x720():
    from softrast_amd import scenes
    _check(scenes.cube_grid_scene(), "cube_grid")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_clip_heavy(seed):
    sc = helpers.random_scene(seed, 3000, 333, 211, clip_heavy=True)
    _check(sc, f"clip_heavy_{seed}")


def test_many_fan_clips():
    """Clipped polygons of 6 to 8 vertices (4 to 6 fan triangles per
    submitted triangle; random clip-heavy scenes stop at 5 vertices): the
    fan numbering, the fan records and their order keys match the oracle."""
    sc, hist = helpers.many_fan_scene(5, 333, 211)
    _say(f"[parity] case=many_fan polygon sizes {hist}")
    assert max(hist) >= 8
    _check(sc, "many_fan")


@pytest.mark.parametrize("stride", [1, 2, 4])
def test_index_strides(stride):
    n = 84 if stride == 1 else 1500
    sc = helpers.random_scene(10 + stride, n, 256, 192, index_stride=stride)
    _check(sc, f"index_stride_{stride}")


@pytest.mark.parametrize("shader", [1, 2])
def test_other_shaders(shader):
    sc = helpers.random_scene(20 + shader, 2000, 320, 240, shaders=(shader,), clip_heavy=True)
    _check(sc, f"shader_{shader}")


def test_no_texture_is_white():
    sc = helpers.random_scene(30, 1000, 200, 120, with_texture=False)
    ref, _ = _check(sc, "no_texture")
    written = ref.primid != 0xFFFFFFFF
    assert written.any() and np.all(ref.colour[written] == 0xFFFFFFFF)


def test_uv_offset_not_6():
    sc = helpers.random_scene(31, 2000, 300, 200, uv_offset=3)
    _check(sc, "uv_offset_3")


def test_non_square_texture():
    sc = helpers.random_scene(32, 2000, 300, 200, tex_size=(128, 32))
    _check(sc, "tex_128x32")


def test_depth_ties_first_wins():
    sc = helpers.random_scene(33, 1500, 256, 256, coplanar_dups=400)
    _check(sc, "coplanar_dups")


def test_multi_draw_mixed_shaders():
    sc = helpers.random_scene(34, 3000, 400, 300, ndraws=3, shaders=(0, 1, 2), clip_heavy=True)
    _check(sc, "multi_draw")


def test_not_cleared_depth_and_colour():
    sc = helpers.random_scene(35, 2000, 256, 128)
    rng = np.random.default_rng(35)
    tiles = 4 * 2 * 4096
    depth = rng.uniform(-0.2, 0.2, tiles).astype(np.float32)
    depth[rng.random(tiles) < 0.05] = np.nan
    depth[rng.random(tiles) < 0.05] = -0.0
    colour = rng.integers(0, 2**32, tiles, dtype=np.uint64).astype(np.uint32)
    _check(sc, "no_clear", clear=None, plane_init=(colour, depth))


def test_clear_colour_low_byte():
    sc = helpers.random_scene(36, 300, 128, 128)
    ref, _ = _check(sc, "clear_colour", clear=0x12345678)
    untouched = ref.primid == 0xFFFFFFFF
    assert np.all(ref.colour[untouched] == 0x78787878)


def test_large_triangles_full_blocks():
    from softrast_amd import scenes
    rng = np.random.default_rng(37)
    v = np.zeros((6, 8), np.float32)
    v[:, 0:3] = [[-50, -30, 20], [50, -30, 25], [0, 40, 22], [-40, 35, 30], [45, 30, 18], [0, -40, 26]]
    v[:, 6:8] = rng.uniform(-1, 1, (6, 2))
    idx = np.array([0, 2, 1, 3, 5, 4], np.uint16)
    sc = scenes.Scene(640, 480, scenes.camera_mvp((0, 0, 0), (0, 0, 1), 60.0, 640 / 480),
                      [scenes.Draw(v.view(scenes.VERTEX_DTYPE).reshape(-1), idx, 0)],
                      [rng.integers(0, 256, (256, 256, 4), dtype=np.uint8)])
    _check(sc, "large_tris")


def test_two_frames_same_context():
    import softrast_amd as sr
    sc = helpers.random_scene(38, 2000, 300, 300, clip_heavy=True)
    tex = helpers.build_textures(sc)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    a = helpers.gpu_render(sc, tex, ctx=ctx)
    b = helpers.gpu_render(sc, tex, ctx=ctx)
    ctx.Shutdown()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[2], b[2])


def test_spheres_1080p_1m_tris():
    """Full benchmark workload (1920x1080, 1,000,000 triangles, 8 x 2048^2 textures)."""
    from softrast_amd import scenes
    sc = scenes.spheres_scene()
    assert sc.num_tris == 1_000_000
    _check(sc, "spheres_1m", nthreads=16)


# (work-bin edge, workgroups per bin): every layout the library accepts
LAYOUTS = [(8, 1), (16, 1), (32, 1), (32, 2), (32, 4), (64, 1), (64, 2), (64, 4)]


@pytest.mark.parametrize("bin_size,wg_per_tile", LAYOUTS)
def test_layouts_same_image(bin_size, wg_per_tile):
    """Work-bin size and workgroups per bin are scheduling choices: every
    layout reproduces the oracle."""
    sc = helpers.random_scene(40, 4000, 517, 333, clip_heavy=True, ndraws=2, shaders=(0, 1))
    _check(sc, f"layout_{bin_size}x{wg_per_tile}", bin_size=bin_size, wg_per_tile=wg_per_tile)


@pytest.mark.parametrize("bin_size,wg_per_tile", [(8, 1), (32, 4), (64, 2), (64, 4)])
def test_layouts_clustered_learned(bin_size, wg_per_tile):
    """Strips and one-wave bins under the learned order: the image is the
    oracle's, and on every frame that records (each frame until trained, the
    frame after training, then one in sixteen) every non-empty bin has a
    recorded cost (summed over its strips)."""
    import softrast_amd as sr
    from softrast_amd import scenes
    sc = scenes.clustered_scene(ntris=60000, width=960, height=544, ndraws=2, tex_size=128)
    tex = helpers.build_textures(sc)
    ref, rstats = helpers.oracle_render(sc, tex, nthreads=8)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=sr.ORDER_LEARNED,
                           bin_size=bin_size, wg_per_tile=wg_per_tile, sched_warmup_frames=1)
    recorded = []
    for frame in range(6):
        _say(f"[parity] case=layout_clustered {bin_size}x{wg_per_tile} frame={frame}")
        col, dep, pid, st = helpers.gpu_render(sc, tex, ctx=ctx)
        d = helpers.compare(ref, col, dep, pid)
        assert not any(d.values()), (frame, d)
        assert (st["bin_size"], st["wg_per_tile"]) == (bin_size, wg_per_tile)
        counts, costs = ctx.bin_stats()
        assert len(counts) == ((sc.width + 63) // 64) * ((sc.height + 63) // 64) * (64 // bin_size) ** 2
        if costs.any():
            assert np.all(costs[counts > 0] > 0), frame
            recorded.append(frame)
    assert st["frags_raster"] == rstats["frags_raster"] and st["sched_trained"]
    # frames 0 (untrained) and 1 (right after training) record, then one in sixteen
    assert recorded[:2] == [0, 1] and 2 <= len(recorded) <= 4, recorded
    ctx.Shutdown()


def test_autotune_sweeps_every_layout():
    """SRG_CFG_AUTOTUNE times each layout over autotune_frames frames of one
    frame signature, keeps the fastest, and starts a new sweep for a new
    signature; every frame of the sweep is the oracle's image. The learned
    order records nothing during the sweep and trains on the chosen layout."""
    import softrast_amd as sr
    a = helpers.random_scene(40, 4000, 517, 333, clip_heavy=True, ndraws=2, shaders=(0, 1))
    b = helpers.random_scene(41, 2000, 300, 200, clip_heavy=True)
    refs = {}
    for name, sc in (("a", a), ("b", b)):
        tex = helpers.build_textures(sc)
        refs[name] = (sc, tex, helpers.oracle_render(sc, tex, nthreads=8)[0])
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_AUTOTUNE, tile_order=sr.ORDER_LEARNED,
                           autotune_frames=2, sched_warmup_frames=2)

    def frame(name):
        sc, tex, ref = refs[name]
        col, dep, pid, st = helpers.gpu_render(sc, tex, ctx=ctx)
        d = helpers.compare(ref, col, dep, pid)
        assert not any(d.values()), (name, d)
        return st

    used = []
    for i in range(2 * len(LAYOUTS)):
        _say(f"[parity] case=autotune sweep frame={i}")
        st = frame("a")
        used.append((st["bin_size"], st["wg_per_tile"]))
        assert not st["sched_trained"]
    assert used == [l for l in LAYOUTS for _ in range(2)]
    t = ctx.tuning()
    assert t["state"] == "chosen" and all(v > 0 for v in t["ms"].values())
    chosen = min(t["ms"], key=t["ms"].get)
    assert t["layout"] == chosen
    for i in range(3):
        st = frame("a")
        assert (st["bin_size"], st["wg_per_tile"]) == chosen
    assert st["sched_trained"]
    st = frame("b")  # a new signature starts its own sweep
    assert (st["bin_size"], st["wg_per_tile"]) == LAYOUTS[0] and ctx.tuning()["state"] == "sweeping"
    st = frame("a")  # the first signature keeps its choice
    assert (st["bin_size"], st["wg_per_tile"]) == chosen and ctx.tuning()["layout"] == chosen
    ctx.Shutdown()


def _multi_view_scene(seed):
    """Three views of different sizes sharing vertex data, one draw list each."""
    from softrast_amd import scenes
    base = helpers.random_scene(seed, 3000, 64, 64, clip_heavy=True, ndraws=2, shaders=(0, 1))
    sizes = [(333, 211), (640, 360), (128, 500)]
    views = []
    for i, (w, h) in enumerate(sizes):
        mvp = scenes.camera_mvp((0.3 * i, 0.0, 0.0), (0.1 * i, 0.0, 1.0), 60.0 + 10 * i, w / h)
        views.append((w, h, mvp))
    return base, views


@pytest.mark.parametrize("tile_order,graph", [(0, False), (1, False), (1, True), (2, False), (2, True)])
def test_batched_views_schedules_same_image(tile_order, graph):
    import softrast_amd as sr
    from oracle.oracle import Oracle, Plane, make_texture
    base, views = _multi_view_scene(41)
    tex = helpers.build_textures(base)
    _say(f"[parity] case=batched_views order={tile_order} graph={graph}")
    # oracle: all views in one frame, one plane per view
    orc = Oracle(nthreads=8)
    planes = [Plane(w, h) for (w, h, _) in views]
    for p in planes:
        orc.clear(p, 0)
    stex = [make_texture(t) for t in tex]
    odraws = []
    for vi, (w, h, mvp) in enumerate(views):
        for d in base.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            odraws.append(dict(shader=d.shader, tex=stex[0][0], idx=d.indices, idx_stride=4, pos=vf, pos_stride=32,
                               pos_num=len(vf), attr=vf, attr_stride=32, attr_num=len(vf), uv_offset=6, plane=vi,
                               mvp=mvp))
    orc.render(odraws, planes)
    orc.close()
    # GPU: same frame, three consecutive times (graph capture then replays)
    flags = sr.CFG_KEEP_PRIMID | (sr.CFG_GRAPH if graph else 0)
    # the learned order trains its model after the first frame
    ctx = sr.RenderContext(flags=flags, tile_order=tile_order, sched_warmup_frames=1)
    fbs = [sr.FrameBuffer(ctx, w, h) for (w, h, _) in views]
    shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader}
    for it in range(3):
        ctx.BeginFrame()
        for fb, (w, h, mvp) in zip(fbs, views):
            ctx.ClearFrameBuffer(fb, 0)
            for d in base.draws:
                c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(mvp)
                c.SetIndexBuffer(d.indices, 4, len(d.indices)).SetPositionBuffer(d.vertices, 32, len(d.vertices))
                c.SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6).SetPixelShader(shader[d.shader], tex[0])
                ctx.DrawIndexed(c)
        ctx.EndFrame()
        for vi, (fb, p) in enumerate(zip(fbs, planes)):
            col, dep, pid = fb.read_tiles(primid=True)
            d = helpers.compare(p, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
            _say(f"[parity] batched_views frame={it} view={vi} diffs={d}")
            assert not any(d.values()), d
    st = ctx.frame_stats()
    if graph:
        assert st["graph_captures"] == 0  # the last frame replayed the captured graph
    if tile_order == 2:
        assert st["sched_trained"] == 1 and st["sched_trainings"] == 0
        assert np.isfinite(st["sched_err"]) and st["sched_err"] > 0.0
    ctx.Shutdown()


def test_shared_geometry_groups():
    """Draws that share their geometry are set up from one fetch per
    triangle (setup groups of up to 16 draws), draws of their own one by one:
    20 views draw scene A's two draws (groups of 16 and 4 per draw), views 3
    and 7 also scene B (a group of 2), view 5 scene C (a group of one), and
    view 9 scene A's first draw twice more (exact depth ties within a view).
    Every view against the oracle's render of the same frame."""
    import softrast_amd as sr
    from oracle.oracle import Oracle, Plane, make_texture
    from softrast_amd import scenes
    a = helpers.random_scene(61, 2000, 64, 64, clip_heavy=True, ndraws=2, shaders=(0, 1))
    b = helpers.random_scene(62, 900, 64, 64, clip_heavy=True, ndraws=1, shaders=(0,))
    c = helpers.random_scene(63, 700, 64, 64, ndraws=1, shaders=(0,))
    tex = helpers.build_textures(a)
    views = []
    for i in range(20):
        w, h = (96, 64) if i % 2 else (160, 90)
        mvp = scenes.camera_mvp((0.2 * (i % 5), 0.1 * (i // 5), 0.0), (0.05 * i, 0.0, 1.0), 60.0, w / h)
        draws = list(a.draws) + (list(b.draws) if i in (3, 7) else []) + (list(c.draws) if i == 5 else [])
        if i == 9:
            draws += [a.draws[0], a.draws[0]]
        views.append((w, h, mvp, draws))
    _say(f"[parity] case=shared_geometry views={len(views)}")
    orc = Oracle(nthreads=8)
    planes = [Plane(w, h) for (w, h, _, _) in views]
    for p in planes:
        orc.clear(p, 0)
    stex = [make_texture(t) for t in tex]
    odraws = []
    for vi, (w, h, mvp, draws) in enumerate(views):
        for d in draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            odraws.append(dict(shader=d.shader, tex=stex[0][0], idx=d.indices, idx_stride=4, pos=vf, pos_stride=32,
                               pos_num=len(vf), attr=vf, attr_stride=32, attr_num=len(vf), uv_offset=6, plane=vi,
                               mvp=mvp))
    orc.render(odraws, planes)
    orc.close()
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    fbs = [sr.FrameBuffer(ctx, w, h) for (w, h, _, _) in views]
    shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader}
    ctx.BeginFrame()
    for fb, (w, h, mvp, draws) in zip(fbs, views):
        ctx.ClearFrameBuffer(fb, 0)
        for d in draws:
            cl = sr.DrawCall().SetFrameBuffer(fb).SetMVP(mvp)
            cl.SetIndexBuffer(d.indices, 4, len(d.indices)).SetPositionBuffer(d.vertices, 32, len(d.vertices))
            cl.SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6).SetPixelShader(shader[d.shader], tex[0])
            ctx.DrawIndexed(cl)
    ctx.EndFrame()
    st = ctx.frame_stats()
    assert st["tris_clipped"] > 0
    for vi, (fb, p) in enumerate(zip(fbs, planes)):
        col, dep, pid = fb.read_tiles(primid=True)
        d = helpers.compare(p, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
        assert not any(d.values()), (vi, d)
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


@pytest.mark.gpu
def test_sched_model_roundtrip_and_order_only():
    """A model set through the C ABI is read back unchanged, and an arbitrary
    (even random) model changes only the bin order, never the image."""
    import softrast_amd as sr
    base = helpers.random_scene(43, 2500, 400, 300, clip_heavy=True, ndraws=1, shaders=(0,))
    tex = helpers.build_textures(base)
    _say("[parity] case=sched_model_roundtrip")
    ref = helpers.gpu_render(base, tex, tile_order=0)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=2, sched_warmup_frames=1000)
    rng = np.random.default_rng(5)
    model = (0.3 * rng.standard_normal(sr.RenderContext.SCHED_PARAMS)).astype(np.float32)
    ctx.set_sched_model(model)
    assert np.array_equal(ctx.sched_model(), model)
    for _ in range(2):
        out = helpers.gpu_render(base, tex, ctx=ctx)
        for a, b in zip(ref[:3], out[:3]):
            assert np.array_equal(a, b)
        assert out[3]["sched_trained"] == 1
    ctx.Shutdown()


@pytest.mark.gpu
def test_async_blit_overlaps_next_frame_and_matches_sync_blit():
    """Blit with the reference's asynchrony (Renderer.cpp:350-372): the
    callback fires with the rows in place while the next frame renders into
    the other plane; the rows equal the synchronous blit of the same frame;
    BGRA exchanges red and blue (Window_Win32.cpp:85-112)."""
    import threading
    import softrast_amd as sr
    base = helpers.random_scene(51, 2000, 300, 170, ndraws=1, shaders=(1,), with_texture=False)
    other = helpers.random_scene(52, 2000, 300, 170, ndraws=1, shaders=(1,), with_texture=False)
    _say("[parity] case=async_blit")
    ctx = sr.RenderContext()
    fb = sr.FrameBuffer(ctx, base.width, base.height)
    fired = threading.Event()
    rows = []
    for bgra in (False, True):
        for sc in (base, other):
            helpers.gpu_render(sc, [], ctx=ctx, fb=fb, flags=0)
            rows.append(ctx.BlitAsync(fb, onFinish=lambda u: fired.set(), bgra=bgra))
        ctx.BlitWait(fb)
    assert fired.is_set()
    sync = []
    for sc in (base, other):
        helpers.gpu_render(sc, [], ctx=ctx, fb=fb, flags=0)
        sync.append(ctx.Blit(fb).copy())
    for i in range(2):
        assert np.array_equal(rows[i], sync[i])
        a, b = rows[2 + i].view(np.uint8).reshape(-1, 4), sync[i].view(np.uint8).reshape(-1, 4)
        assert np.array_equal(a[:, [2, 1, 0, 3]], b)
    assert not np.array_equal(sync[0], sync[1])
    ctx.Shutdown()


@pytest.mark.gpu
@pytest.mark.parametrize("w,h", [(130, 61), (65, 65), (300, 170), (1, 1)])
def test_blit_rows_equal_oracle_blit(w, h):
    """Blit, BlitAsync and BlitAsync with the BGRA present swizzle give the
    oracle's blit (BlitJobFn: per tile, min(64, W - 64*tx) pixels of each of
    min(64, H - 64*ty) rows, Renderer.cpp:319-347) of the oracle's frame, on
    framebuffers with partial edge tiles."""
    import softrast_amd as sr
    from oracle.oracle import Oracle
    sc = helpers.random_scene(60 + w, 3000, w, h, clip_heavy=True)
    tex = helpers.build_textures(sc)
    _say(f"[parity] case=blit_vs_oracle size={w}x{h}")
    ref, _ = helpers.oracle_render(sc, tex, nthreads=8)
    o = Oracle(nthreads=1)
    want = o.blit(ref)
    o.close()
    assert want.shape == (h, w)
    ctx = sr.RenderContext()
    fb = sr.FrameBuffer(ctx, w, h)
    helpers.gpu_render(sc, tex, ctx=ctx, fb=fb, flags=0)
    got = ctx.Blit(fb).copy()
    assert np.array_equal(got, want), int(np.count_nonzero(got != want))
    swapped = want.view(np.uint8).reshape(h, w, 4)[:, :, [2, 1, 0, 3]].copy().view(np.uint32).reshape(h, w)
    for bgra, exp in ((False, want), (True, swapped)):
        helpers.gpu_render(sc, tex, ctx=ctx, fb=fb, flags=0)
        rows = ctx.BlitAsync(fb, bgra=bgra)
        ctx.BlitWait(fb)
        assert np.array_equal(rows, exp), (bgra, int(np.count_nonzero(rows != exp)))
    fb.destroy()
    ctx.Shutdown()


@pytest.mark.gpu
@pytest.mark.parametrize("tile_order", [0, 1, 2])
def test_clustered_scene_every_order(tile_order):
    """The clustered layout of the learned-scheduler workload (90% of the
    triangles in 3 clusters, areas 0.5-2000 px^2) at reduced size: every queue
    order gives the oracle's image."""
    import softrast_amd as sr
    from softrast_amd import scenes
    sc = scenes.clustered_scene(ntris=120000, width=1280, height=720, ndraws=4, tex_size=256)
    tex = helpers.build_textures(sc)
    _say(f"[parity] case=clustered order={tile_order} tris={sc.num_tris}")
    ref, rstats = helpers.oracle_render(sc, tex, nthreads=8)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=tile_order, sched_warmup_frames=1)
    for frame in range(3):
        col, dep, pid, st = helpers.gpu_render(sc, tex, ctx=ctx)
        d = helpers.compare(ref, col, dep, pid)
        _say(f"[parity] clustered order={tile_order} frame={frame} diffs={d}")
        assert not any(d.values()), d
    assert st["frags_raster"] == rstats["frags_raster"]
    ctx.Shutdown()


@pytest.mark.gpu
@pytest.mark.parametrize("tile_order", [0, 1, 2])
def test_arena_overflow_retry(tile_order):
    """A frame whose clip-fan or entry arena overflows is rendered again with
    larger arenas, in every queue order. The overflowed attempt has bin starts
    past the entry arena and no entry lists, so every kernel that reads them
    (emit, back end, the learned order's features) must skip it. One-element
    arenas overflow the first frame; a larger second scene overflows the grown
    arenas again after the learned model has trained."""
    import softrast_amd as sr
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=tile_order, max_prims=1,
                           max_entries=1, sched_warmup_frames=1, sched_retrain_interval=1)
    for name, sc, frames in (("small", helpers.random_scene(71, 3000, 320, 240, clip_heavy=True), 3),
                             ("large", helpers.random_scene(72, 40000, 800, 600, clip_heavy=True), 2)):
        tex = helpers.build_textures(sc)
        ref, rstats = helpers.oracle_render(sc, tex, nthreads=8)
        for frame in range(frames):
            _say(f"[parity] case=overflow_retry order={tile_order} scene={name} frame={frame}")
            col, dep, pid, st = helpers.gpu_render(sc, tex, ctx=ctx)
            d = helpers.compare(ref, col, dep, pid)
            assert not any(d.values()), (name, frame, d)
            # re-runs of this frame: the first overflows, the grown arenas are kept
            assert (st["retries"] > 0) == (frame == 0), (name, frame, st)
        assert st["tris_clipped"] > 0 and st["frags_raster"] == rstats["frags_raster"]
    if tile_order == sr.ORDER_LEARNED:
        assert st["sched_trained"]
    ctx.Shutdown()


def _sliver_scene(seed, n, width, height):
    """Needle-thin triangles (0.0005-0.008 px wide, 2-40 px long) in clip
    space, both windings: mostly primitives without a tight pixel bound,
    whose accepted pixels the fill bias can push outside their bbox."""
    from softrast_amd import scenes
    rng = np.random.default_rng(seed)
    p0 = rng.uniform([0, 0], [width, height], (n, 2))
    ang = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.cos(ang), np.sin(ang)], 1) * rng.uniform(2, 40, n)[:, None]
    nrm = np.stack([-d[:, 1], d[:, 0]], 1) / np.linalg.norm(d, axis=1)[:, None]
    w = rng.uniform(0.0005, 0.008, n)[:, None]
    t = rng.uniform(0.2, 0.8, n)[:, None]
    p1 = p0 + d
    p2 = p0 + d * t + nrm * w
    flip = rng.random(n) < 0.5
    p1[flip], p2[flip] = p2[flip].copy(), p1[flip].copy()
    pts = np.stack([p0, p1, p2], 1).reshape(-1, 2)
    v = np.zeros(n * 3, scenes.VERTEX_DTYPE)
    v["pos"][:, 0] = 2 * pts[:, 0] / width - 1
    v["pos"][:, 1] = 1 - 2 * pts[:, 1] / height
    v["pos"][:, 2] = np.repeat(rng.uniform(0.05, 0.95, n), 3)
    v["nrm"] = rng.normal(0, 1, (n * 3, 3))
    v["uv"] = rng.uniform(0, 2, (n * 3, 2))
    draw = scenes.Draw(v, np.arange(n * 3, dtype=np.uint32), -1, 1)
    return scenes.Scene(width, height, np.eye(4, dtype=np.float32).reshape(-1), [draw], [], {"config": "slivers"})


@pytest.mark.gpu
@pytest.mark.parametrize("bin_size,wg_per_tile", LAYOUTS)
def test_needle_slivers(bin_size, wg_per_tile):
    sc = _sliver_scene(61, 20000, 400, 300)
    _check(sc, f"slivers_{bin_size}x{wg_per_tile}", bin_size=bin_size, wg_per_tile=wg_per_tile)


def _viewset_textures(vs):
    from softrast_amd import TextureData
    return [TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in vs.textures]


def _compare_views(vs, planes, fbs, what):
    for vi, (fb, p) in enumerate(zip(fbs, planes)):
        col, dep, pid = fb.read_tiles(primid=True)
        d = helpers.compare(p, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
        assert not any(d.values()), (what, vi, vs.views[vi].width, vs.views[vi].height, d)


def test_mixed_views_batch_autotuned():
    """The mixed-views batch at reduced triangle counts: 12 views of the four
    resolutions (sphere/cube mixes, u16 and u32 indices) in one
    hipGraph-captured EndFrame while the autotuner sweeps every layout; every
    view of every frame is the oracle's image."""
    import softrast_amd as sr
    from softrast_amd import scenes
    vs = scenes.mixed_views(nviews=12, tri_max=20000, tex_size=128)
    tex = _viewset_textures(vs)
    planes, rstats = helpers.oracle_render_views(vs, tex)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_GRAPH | sr.CFG_AUTOTUNE,
                           tile_order=sr.ORDER_TRICOUNT, autotune_frames=1)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    for frame in range(len(LAYOUTS) + 2):
        _say(f"[parity] case=mixed_views frame={frame} tris={vs.num_tris}")
        st = helpers.gpu_render_views(vs, tex, ctx, fbs)
        _compare_views(vs, planes, fbs, frame)
        assert st["frags_raster"] == rstats["frags_raster"]
    t = ctx.tuning()
    assert t["state"] == "chosen" and (st["bin_size"], st["wg_per_tile"]) == t["layout"]
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


def test_combined_views_learned_retrain():
    """The combined workload at reduced size: 4 views alternating uniform and
    90/10 clustered layouts in one EndFrame, learned order trained after 2
    frames and retrained every 2; every view of every frame is the oracle's
    image."""
    import softrast_amd as sr
    from softrast_amd import scenes
    vs = scenes.combined_views(nviews=4, width=640, height=360, ntris=30000, tex_size=128, threads=4)
    tex = _viewset_textures(vs)
    planes, rstats = helpers.oracle_render_views(vs, tex)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=sr.ORDER_LEARNED,
                           sched_warmup_frames=2, sched_retrain_interval=2)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    trainings = 0
    for frame in range(7):
        _say(f"[parity] case=combined_views frame={frame}")
        st = helpers.gpu_render_views(vs, tex, ctx, fbs)
        _compare_views(vs, planes, fbs, frame)
        assert st["frags_raster"] == rstats["frags_raster"]
        trainings += st["sched_trainings"]
    assert st["sched_trained"] and trainings >= 3
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


def test_learned_order_falls_back_on_large_error():
    """A model whose measured error exceeds one log2 unit hands the queue
    back to the entry-count order (the fallback); the image never changes."""
    import softrast_amd as sr
    sc = helpers.random_scene(42, 3000, 400, 300, clip_heavy=True)
    tex = helpers.build_textures(sc)
    ref, _ = helpers.oracle_render(sc, tex, nthreads=8)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=sr.ORDER_LEARNED)
    bad = np.zeros(ctx.SCHED_PARAMS, np.float32)
    bad[-1] = 40.0  # output bias: every bin predicted at 2^40 ticks
    ctx.set_sched_model(bad)
    for frame in range(3):
        _say(f"[parity] case=sched_fallback frame={frame}")
        col, dep, pid, st = helpers.gpu_render(sc, tex, ctx=ctx)
        d = helpers.compare(ref, col, dep, pid)
        assert not any(d.values()), (frame, d)
        assert st["sched_trained"] and st["sched_err"] > 1.0 and st["sched_fallback"] == 1, st
    ctx.Shutdown()


def test_draw_list_follows_write_plane():
    """A DrawList (one srg_draw_indexed_n call per frame) renders the oracle's
    image, and after Blit swaps the planes it draws into the new write plane
    (SRG_PLANE_WRITE), leaving the blitted plane as it was."""
    import softrast_amd as sr
    sc = helpers.random_scene(43, 3000, 320, 240, clip_heavy=True, ndraws=3, shaders=(0, 1, 2))
    tex = helpers.build_textures(sc)
    refs = [helpers.oracle_render(sc, tex, nthreads=8, clear=c)[0] for c in (0, 0x40)]
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    fb = sr.FrameBuffer(ctx, sc.width, sc.height)
    shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader, 2: sr.VisualizeUVsShader}
    calls = []
    for d in sc.draws:
        c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(sc.mvp)
        c.SetIndexBuffer(d.indices, d.indices.dtype.itemsize, len(d.indices))
        c.SetPositionBuffer(d.vertices, 32, len(d.vertices)).SetAttributeBuffer(d.vertices, 32, len(d.vertices), 6)
        c.SetPixelShader(shader[d.shader], tex[d.texture] if d.texture >= 0 else None)
        calls.append(c)
    draws = ctx.PrepareDrawList(calls)
    assert len(draws) == len(sc.draws)
    shown = None
    for frame in range(2):
        _say(f"[parity] case=draw_list frame={frame}")
        plane = fb.m_writePlane
        ctx.BeginFrame()
        ctx.ClearFrameBuffer(fb, (0, 0x40)[frame])  # the two frames differ in the background
        ctx.DrawIndexedList(draws)
        ctx.EndFrame()
        col, dep, pid = fb.read_tiles(primid=True)
        d = helpers.compare(refs[frame], col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
        assert not any(d.values()), (frame, d)
        if shown is not None:  # the plane blitted last frame was not drawn into
            assert np.array_equal(fb.read_tiles(plane=plane ^ 1)[0], shown)
            assert not np.array_equal(col, shown)
        shown = col.copy()
        ctx.Blit(fb)
        assert fb.m_writePlane == plane ^ 1
    fb.destroy()
    ctx.Shutdown()


def test_learned_order_view_count_changes():
    """A trained learned-order context whose frames change the number and
    sizes of views (so every bin's number changes): the frame after each
    change predicts afresh, and every view of every frame is the oracle's
    image."""
    import softrast_amd as sr
    from softrast_amd import scenes
    pool = [helpers.random_scene(700 + k, 2500, w, h, clip_heavy=(k % 2 == 1), tex_size=(64, 64))
            for k, (w, h) in enumerate(((320, 200), (256, 256), (400, 180), (200, 300), (333, 222)))]
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS, tile_order=sr.ORDER_LEARNED,
                           sched_warmup_frames=2)
    fbs = {}
    for frame, pick in enumerate(((0, 1, 2), (0, 1, 2), (0, 1, 2), (3, 4, 0, 1, 2), (2, 4), (3, 4, 0, 1, 2), (1,))):
        vs = scenes.ViewSet([pool[k] for k in pick], pool[0].textures)
        for v in vs.views:
            for d in v.draws:
                d.texture = 0
        tex = _viewset_textures(vs)
        views = [fbs.setdefault(k, sr.FrameBuffer(ctx, pool[k].width, pool[k].height)) for k in pick]
        _say(f"[parity] case=learned_view_changes frame={frame} views={pick}")
        st = helpers.gpu_render_views(vs, tex, ctx, views)
        planes, rstats = helpers.oracle_render_views(vs, tex)
        _compare_views(vs, planes, views, frame)
        assert st["frags_raster"] == rstats["frags_raster"]
        if frame >= 3:
            assert st["sched_trained"]
    for fb in fbs.values():
        fb.destroy()
    ctx.Shutdown()
