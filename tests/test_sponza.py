"""SponzaShader (Viewer/SponzaScene.cpp:13-104): the oracle pinned by closed
forms on the CPU, then the GPU against the oracle.

The closed forms draw a viewport-filling quad whose attributes are the same
at every vertex (position P, normal N, uv), so every pixel sees the same
varyings and the expected colour follows from the shader's formula in
float64: sun term max(0.1, N.sun), each light's N.L * intensity /
(1 + 0.1 d + 0.01 d^2) * colour, ambient, times the texel. The inputs are
chosen so that each term moves the result by many LSBs (a dropped or
swapped term fails); the tolerance is 1 LSB, from RSQRTPS / RCPPS
(|rel err| <= 1.5 * 2^-12) and the texel lerp's rounding."""
import sys

import numpy as np
import pytest

import softrast_amd as sr
from softrast_amd import scenes
from tests import helpers

TEXEL = (200, 150, 100, 255)


def _consts(lights=(), **kw):
    """SponzaConstants whose unused lights sit far away with zero intensity
    (a light at the surface point itself would make the result NaN)."""
    far = [((1e3, 1e3 + i, 1e3), (0.0, 0.0, 0.0), 0.0, 0.0) for i in range(16 - len(lights))]
    return sr.SponzaConstants(lights=list(lights) + far, **kw)


def _const_texture(rgba=TEXEL, size=32):
    img = np.empty((size, size, 4), np.uint8)
    img[:] = rgba
    return sr.TextureData().CreateFromRGBA8(img, size, size, False)


def _quad_draw(tex, P, N, uv=(0.3, 0.6), uv_offset=6):
    from oracle.oracle import make_texture
    pos = np.array([[-1, -1, 0.5], [1, -1, 0.5], [1, 1, 0.5], [-1, 1, 0.5]], np.float32)
    attr = np.zeros((4, 8), np.float32)
    attr[:, 0:3] = P
    attr[:, 3:6] = N
    attr[:, uv_offset:uv_offset + 2] = uv
    idx = np.array([0, 1, 2, 0, 2, 3, 0, 2, 1, 0, 3, 2], np.uint32)  # both windings: one of each pair is drawn
    st = make_texture(tex)[0] if tex is not None else None
    return dict(shader=3, tex=st, idx=idx, idx_stride=4, pos=pos, pos_stride=12, pos_num=4, attr=attr,
                attr_stride=32, attr_num=4, uv_offset=uv_offset, plane=0, mvp=np.eye(4, dtype=np.float32)), tex


def _render_const(consts, P, N, tex=TEXEL, uv_offset=6):
    """Every pixel of a 64x64 frame with one set of varyings; returns the
    distinct colours (must be one)."""
    from oracle.oracle import Oracle, Plane
    o = Oracle(nthreads=1)
    o.set_sponza_constants(consts.to_bytes())
    p = Plane(64, 64)
    o.clear(p, 0)
    t = _const_texture(tex) if tex is not None else None
    draw, keep = _quad_draw(t, P, N, uv_offset=uv_offset)
    o.render([draw], [p])
    o.close()
    vals = np.unique(p.colour)
    assert len(vals) == 1, vals[:8]
    return int(vals[0])


def _expect(rad, tex=TEXEL):
    """Packed RGBA8 of rad * texel / 255 (RNE(fma(c, 255, 0.5)) per channel)."""
    out = []
    for c in range(3):
        out.append(int(np.rint(rad[c] * tex[c] / 255.0 * 255.0 + 0.5)))
    out.append(int(np.rint(tex[3] / 255.0 * 255.0 + 0.5)) if tex[3] < 255 else 255)
    return out


def _bytes(u):
    return [(u >> (8 * k)) & 255 for k in range(4)]


def _close(got, want):
    g = _bytes(got)
    assert all(abs(a - b) <= 1 for a, b in zip(g, want)), (g, want)


def test_sun_and_ambient_only():
    """No light intensity: rad = max(0.1, N.sun) + ambient per channel."""
    n = np.array([0.6, 0.0, 0.8])
    k = _consts(sun_dir=(0.5, 0.5, 0.5), ambient=(0.05, 0.1, 0.2))
    got = _render_const(k, (1, 2, 3), n)
    ndl = max(0.1, float(n @ np.array([0.5, 0.5, 0.5])))   # 0.7
    _close(got, _expect([ndl + 0.05, ndl + 0.1, ndl + 0.2]))
    # a normal facing away: the 0.1 bias holds the sun term up
    got = _render_const(k, (1, 2, 3), -n)
    _close(got, _expect([0.1 + 0.05, 0.1 + 0.1, 0.1 + 0.2]))


def test_one_light_attenuation():
    """Light at distance 10 (d^2 = 100): atten = 1 / (1 + 1 + 1) = 1/3,
    N.L = 0.8; every term of the attenuation moves the red byte by > 20."""
    k = _consts(ambient=(0.05, 0.1, 0.2), lights=[((0.0, 6.0, 8.0), (1.0, 0.5, 0.25), 0.9, 1000.0)])
    got = _render_const(k, (0, 0, 0), (0, 0, 1))
    lr = 0.8 * 0.9 / 3.0
    _close(got, _expect([0.1 + lr * 1.0 + 0.05, 0.1 + lr * 0.5 + 0.1, 0.1 + lr * 0.25 + 0.2]))


def test_sixteen_lights_sum():
    rng = np.random.default_rng(5)
    P = np.array([1.5, -2.0, 4.0])
    N = np.array([0.0, 0.6, 0.8])
    lights = []
    rad = np.full(3, max(0.1, float(N @ np.array([0.3, 0.3, 0.3]))))
    for i in range(16):
        d = rng.normal(0, 1, 3)
        d /= np.linalg.norm(d)
        dist = 2.0 + 3.0 * i
        L = (P + d * dist).astype(np.float32)
        col = rng.uniform(0, 1, 3).astype(np.float32)
        inten = np.float32(rng.uniform(0.05, 0.4))
        lights.append((tuple(L), tuple(col), float(inten), 0.0))
        t = L.astype(np.float64) - P
        dd = np.linalg.norm(t)
        nl = max(0.0, float((t / dd) @ N))
        rad += nl * inten / (1.0 + 0.1 * dd + 0.01 * dd * dd) * col
    rad += np.array([0.02, 0.03, 0.04])
    k = sr.SponzaConstants(sun_dir=(0.3, 0.3, 0.3), ambient=(0.02, 0.03, 0.04), lights=lights)
    _close(_render_const(k, P, N), _expect(list(rad)))


def test_light_at_the_surface_point_is_nan():
    """d^2 = 0: RSQRTPS(0) = inf, L = 0 * inf = NaN; max(0, NaN) returns the
    NaN (second operand), so the colour channels pack to 0; alpha is the
    texel's."""
    k = _consts(lights=[((1.0, 2.0, 3.0), (1.0, 1.0, 1.0), 1.0, 0.0)])
    assert _render_const(k, (1, 2, 3), (0, 0, 1)) == 0xFF000000


def test_no_texture_is_white():
    k = _consts(ambient=(0.5, 0.5, 0.5))
    assert _render_const(k, (0, 0, 0), (0, 0, 1), tex=None) == 0xFFFFFFFF


def test_uv_offset_moves_the_sampled_pair():
    """uv at attributes 2-3 instead of 6-7 (positions then overlap): only
    the sample and its derivatives move; a constant texture hides that, so
    this pins that the shader still reads the eight varyings as the
    reference lays them out."""
    k = _consts(ambient=(0.1, 0.1, 0.1))
    got = _render_const(k, (0, 0, 0), (0, 0, 1), uv_offset=2)
    _close(got, _expect([0.2, 0.2, 0.2]))


def test_attribute_stride_must_be_32():
    from oracle.oracle import Oracle, Plane
    o = Oracle(nthreads=1)
    p = Plane(64, 64)
    draw, keep = _quad_draw(None, (0, 0, 0), (0, 0, 1))
    draw["attr"] = draw["attr"][:, :6].copy()
    draw["attr_stride"] = 24
    with pytest.raises(RuntimeError):
        o.render([draw], [p])
    o.close()


# ---------------------------------------------------------------- GPU parity

def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sponza_scene(seed, ntris, w, h, clip_heavy=False, uv_offset=6, mix=False):
    shaders = (3, 0) if mix else (3,)
    return helpers.random_scene(seed, ntris, w, h, clip_heavy=clip_heavy, shaders=shaders, uv_offset=uv_offset,
                                ndraws=2 if mix else 1)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(seed=300, ntris=20000, w=640, h=360, near=False),
    dict(seed=301, ntris=20000, w=640, h=360, near=True),
    dict(seed=302, ntris=12000, w=320, h=240, near=True, clip_heavy=True),
    dict(seed=303, ntris=12000, w=320, h=240, near=False, mix=True),
    dict(seed=304, ntris=8000, w=320, h=240, near=True, uv_offset=2),
])
def test_gpu_sponza_matches_oracle(case):
    """SponzaShader frames, bit-exact against the oracle: the reference's
    own light set-up (far lights, SponzaScene::Init ranges), near lights
    that saturate and hit the packus quirks, clipped triangles (the wide
    clipper), a frame mixing SponzaShader and UnlitDiffuse draws, and uv at
    attributes 2-3."""
    sc = _sponza_scene(case["seed"], case["ntris"], case["w"], case["h"], case.get("clip_heavy", False),
                       case.get("uv_offset", 6), case.get("mix", False))
    k = scenes.sponza_constants(case["seed"], near=case["near"])
    _say(f"[sponza] case={case}")
    tex = helpers.build_textures(sc)
    ref, rstats = helpers.oracle_render(sc, tex, nthreads=8, sponza=k)
    col, dep, pid, st = helpers.gpu_render(sc, tex, sponza=k)
    d = helpers.compare(ref, col, dep, pid)
    _say(f"[sponza] diffs={d}")
    assert not any(d.values()), d
    assert st["frags_raster"] == rstats["frags_raster"]
    lit = np.count_nonzero(ref.primid != 0xFFFFFFFF)
    assert lit > 1000


@pytest.mark.gpu
def test_gpu_sponza_many_fan_clips():
    """The eight-plane clipper (k_clip<8>) on clipped polygons of 6 to 8
    vertices, every attribute lerped through each plane, bit-exact."""
    sc, hist = helpers.many_fan_scene(7, 320, 240, shader=3)
    _say(f"[sponza] case=many_fan polygon sizes {hist}")
    assert max(hist) >= 8
    k = scenes.sponza_constants(7, near=True)
    tex = helpers.build_textures(sc)
    ref, rstats = helpers.oracle_render(sc, tex, nthreads=8, sponza=k)
    col, dep, pid, st = helpers.gpu_render(sc, tex, sponza=k)
    d = helpers.compare(ref, col, dep, pid)
    _say(f"[sponza] case=many_fan diffs={d}")
    assert not any(d.values()), d
    assert st["frags_raster"] == rstats["frags_raster"]


@pytest.mark.gpu
def test_gpu_sponza_then_narrow_frame():
    """A context that renders a SponzaShader frame (eight planes per record)
    and then an UnlitDiffuse frame (four) gets both right: the record
    layout follows each frame's draws."""
    k = scenes.sponza_constants(7, near=False)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    for sc in (_sponza_scene(310, 6000, 320, 240), helpers.random_scene(311, 6000, 320, 240),
               _sponza_scene(312, 6000, 320, 240, clip_heavy=True)):
        tex = helpers.build_textures(sc)
        ref, _ = helpers.oracle_render(sc, tex, nthreads=8, sponza=k)
        fb = sr.FrameBuffer(ctx, sc.width, sc.height)
        helpers.gpu_render(sc, tex, ctx=ctx, fb=fb, sponza=k)
        col, dep, pid = fb.read_tiles(primid=True)
        d = helpers.compare(ref, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
        _say(f"[sponza] sequence diffs={d}")
        assert not any(d.values()), d
        fb.destroy()
    ctx.Shutdown()


def _strip_draw(x0, x1, P, N, W):
    """Quad over pixel columns [x0, x1) of a W-wide frame (identity MVP, w = 1),
    every vertex with position attribute P, normal N."""
    a, b = 2.0 * x0 / W - 1.0, 2.0 * x1 / W - 1.0
    pos = np.array([[a, -1, 0.5], [b, -1, 0.5], [b, 1, 0.5], [a, 1, 0.5]], np.float32)
    attr = np.zeros((4, 8), np.float32)
    attr[:, 0:3] = P
    attr[:, 3:6] = N
    attr[:, 6:8] = (0.3, 0.6)
    idx = np.array([0, 1, 2, 0, 2, 3, 0, 2, 1, 0, 3, 2], np.uint32)
    return pos, attr, idx


@pytest.mark.gpu
def test_gpu_sponza_special_distances():
    """Waves that mix pixels whose squared light distance is a positive normal
    number (the fast pass) with pixels where it is 0, denormal or NaN (redone
    with the exact RSQRTPS / RCPPS): strips of 37, 29 and 51 columns so that
    waves straddle their borders; bit-exact against the oracle."""
    from oracle.oracle import Oracle, Plane, make_texture
    W, H = 128, 64
    k = _consts(ambient=(0.1, 0.1, 0.1), lights=[((0.0, 0.0, 0.0), (1.0, 0.5, 0.25), 0.9, 0.0),
                                                 ((4.0, 6.0, 3.0), (0.2, 0.9, 0.4), 0.7, 0.0)])
    # d^2 to the first light: 0, 1e-40 (denormal), NaN, then positive normal
    strips = [(0, 37, np.zeros(3, np.float32)), (37, 66, np.array([1e-20, 0, 0], np.float32)),
              (66, 117, np.array([np.nan, 0, 0], np.float32)), (117, 128, np.array([0.5, -1.0, 2.0], np.float32))]
    tex = _const_texture((180, 90, 60, 200))
    st = make_texture(tex)
    o = Oracle(nthreads=1)
    o.set_sponza_constants(k.to_bytes())
    plane = Plane(W, H)
    o.clear(plane, 0)
    odraws, bufs = [], []
    for x0, x1, P in strips:
        pos, attr, idx = _strip_draw(x0, x1, P, (0.0, 0.0, 1.0), W)
        bufs.append((pos, attr, idx))
        odraws.append(dict(shader=3, tex=st[0], idx=idx, idx_stride=4, pos=pos, pos_stride=12, pos_num=4, attr=attr,
                           attr_stride=32, attr_num=4, uv_offset=6, plane=0, mvp=np.eye(4, dtype=np.float32)))
    o.render(odraws, [plane])
    o.close()
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    ctx.set_sponza_constants(k)
    fb = sr.FrameBuffer(ctx, W, H)
    ctx.BeginFrame()
    ctx.ClearFrameBuffer(fb, 0)
    for pos, attr, idx in bufs:
        c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(np.eye(4, dtype=np.float32))
        c.SetIndexBuffer(idx, 4, len(idx)).SetPositionBuffer(pos, 12, 4).SetAttributeBuffer(attr, 32, 4, 6)
        c.SetPixelShader(sr.SponzaShader, tex)
        ctx.DrawIndexed(c)
    ctx.EndFrame()
    col, dep, pid = fb.read_tiles(primid=True)
    d = helpers.compare(plane, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
    _say(f"[sponza] special distances diffs={d}")
    assert not any(d.values()), d
    fb.destroy()
    ctx.Shutdown()
    vals = set(np.unique(plane.colour).tolist())
    assert 0xC9000000 in vals                     # the NaN strips: colour 0, the texel's alpha
    assert any(v & 0xFFFFFF for v in vals)        # the last strip is lit


@pytest.mark.gpu
def test_gpu_sponza_views_batch_graph_autotuned():
    """Five views mixing SponzaShader and UnlitDiffuse draws in one
    hipGraph-captured EndFrame while the autotuner sweeps every layout, in
    the learned order; the constants change half-way (the captured frame
    must be re-captured with them). Every view of every frame is the
    oracle's image."""
    sizes = [(320, 240), (256, 256), (640, 360), (200, 130), (512, 288)]
    views = []
    for k, (w, h) in enumerate(sizes):
        sc = helpers.random_scene(320 + k, 5000, w, h, clip_heavy=(k % 2 == 1), shaders=(3, 0), ndraws=2)
        views.append(sc)
    vs = scenes.ViewSet([scenes.Scene(v.width, v.height, v.mvp, v.draws, []) for v in views], views[0].textures)
    tex = [sr.TextureData().CreateFromRGBA8(t, t.shape[1], t.shape[0], True) for t in vs.textures]
    consts = [scenes.sponza_constants(11, near=False), scenes.sponza_constants(12, near=True)]
    refs = [helpers.oracle_render_views(vs, tex, sponza=k) for k in consts]
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS | sr.CFG_GRAPH | sr.CFG_AUTOTUNE,
                           tile_order=sr.ORDER_LEARNED, autotune_frames=1, sched_warmup_frames=2)
    fbs = [sr.FrameBuffer(ctx, v.width, v.height) for v in vs.views]
    nframes = len(sr.LAYOUTS) + 4
    for frame in range(nframes):
        which = 0 if frame < nframes // 2 else 1
        _say(f"[sponza] case=views_batch frame={frame} constants={which}")
        st = helpers.gpu_render_views(vs, tex, ctx, fbs, sponza=consts[which])
        planes, rstats = refs[which]
        for vi, (fb, p) in enumerate(zip(fbs, planes)):
            col, dep, pid = fb.read_tiles(primid=True)
            d = helpers.compare(p, col.reshape(-1), dep.reshape(-1), pid.reshape(-1))
            assert not any(d.values()), (frame, vi, d)
        assert st["frags_raster"] == rstats["frags_raster"]
    assert ctx.tuning()["state"] == "chosen"
    for fb in fbs:
        fb.destroy()
    ctx.Shutdown()


@pytest.mark.gpu
def test_gpu_sponza_draw_validation():
    """A SponzaShader draw reads eight varyings: an attribute stride other
    than 32 bytes, or a uv offset past the seventh float, is rejected at
    DrawIndexed with SRG_ERR_INVALID; the context then renders normally."""
    ctx = sr.RenderContext(flags=sr.CFG_STATS)
    fb = sr.FrameBuffer(ctx, 64, 64)
    pos, attr, idx = _strip_draw(0, 64, np.zeros(3, np.float32), (0.0, 0.0, 1.0), 64)
    ctx.BeginFrame()
    for stride, attrs, uv in ((24, attr[:, :6].copy(), 4), (32, attr, 7)):
        c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(np.eye(4, dtype=np.float32))
        c.SetIndexBuffer(idx, 4, len(idx)).SetPositionBuffer(pos, 12, 4).SetAttributeBuffer(attrs, stride, 4, uv)
        c.SetPixelShader(sr.SponzaShader, None)
        with pytest.raises(sr.SrgError) as e:
            ctx.DrawIndexed(c)
        assert e.value.status == -1
    c = sr.DrawCall().SetFrameBuffer(fb).SetMVP(np.eye(4, dtype=np.float32))
    c.SetIndexBuffer(idx, 4, len(idx)).SetPositionBuffer(pos, 12, 4).SetAttributeBuffer(attr, 32, 4, 6)
    c.SetPixelShader(sr.SponzaShader, None)
    ctx.DrawIndexed(c)
    ctx.EndFrame()
    col, _ = fb.read_tiles()
    assert np.mean(col.reshape(-1)[:64 * 64] == 0xFFFFFFFF) > 0.9  # no texture: white
    fb.destroy()
    ctx.Shutdown()
