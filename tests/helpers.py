"""Shared test scaffolding: seeded scenes, and the same frame rendered by the
CPU oracle and by the GPU library through the C ABI."""
import numpy as np

from softrast_amd import scenes

ATTR_STRIDE = 32


def random_scene(seed, ntris, width, height, clip_heavy=False, index_stride=4, shaders=(0,), uv_offset=6,
                 ndraws=1, tex_size=(64, 64), with_texture=True, coplanar_dups=0):
    """Random triangles in front of a reverse-Z camera. With clip_heavy, vertices
    extend outside the frustum (sides and the z < 0.1 near region) so many
    triangles go through the clipper. coplanar_dups re-submits that many
    earlier triangles unchanged (exact depth ties)."""
    rng = np.random.default_rng(seed)
    draws = []
    for k in range(ndraws):
        n = ntris // ndraws
        centres = np.stack([rng.uniform(-12, 12, n), rng.uniform(-7, 7, n), rng.uniform(6, 40, n)], 1)
        if clip_heavy:
            centres[:, 0] *= 2.2
            centres[:, 1] *= 2.2
            centres[:, 2] = rng.uniform(-4, 30, n)
        size = np.exp(rng.normal(-0.5, 1.0, n))[:, None, None]
        offs = rng.normal(0, 1, (n, 3, 3)) * size
        pos = (centres[:, None, :] + offs).reshape(-1, 3)
        if coplanar_dups and k == 0:
            pick = rng.integers(0, n, coplanar_dups)
            pos = np.concatenate([pos, pos.reshape(-1, 3, 3)[pick].reshape(-1, 3)])
        nv = len(pos)
        v = np.zeros((nv, 8), np.float32)
        v[:, 0:3] = pos
        nrm = rng.normal(0, 1, (nv, 3))
        v[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
        v[:, 6:8] = rng.uniform(-3, 3, (nv, 2))
        if uv_offset != 6:
            v[:, uv_offset:uv_offset + 2] = rng.uniform(-2, 2, (nv, 2))
        # shuffle vertex storage so indices are not trivial
        perm = rng.permutation(nv)
        inv = np.empty(nv, np.int64)
        inv[perm] = np.arange(nv)
        v = v[perm]
        idx = inv[np.arange(nv)]
        # both windings so about half the triangles face the camera
        t = idx.reshape(-1, 3)
        flip = rng.random(len(t)) < 0.5
        t[flip] = t[flip][:, ::-1]
        idx = t.reshape(-1)
        if index_stride == 1:
            assert nv <= 256
        dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[index_stride]
        draws.append(scenes.Draw(v.view(scenes.VERTEX_DTYPE).reshape(-1), idx.astype(dt),
                                 0 if with_texture else -1, shaders[k % len(shaders)], uv_offset))
    mvp = scenes.camera_mvp((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 60.0, width / height)
    texs = []
    if with_texture:
        tw, th = tex_size
        texs = [rng.integers(0, 256, (th, tw, 4), dtype=np.uint8)]
    return scenes.Scene(width, height, mvp, draws, texs, {"config": "random"})


def many_fan_scene(seed, width, height, candidates=60000, min_verts=6, shader=0):
    """Triangles whose clipped polygons have min_verts to 9 vertices (4 to 7
    fan triangles), chosen with the oracle's clipper from candidates built in
    clip space -- an edge through a line where two frustum planes meet, the
    third vertex anywhere -- and mapped back through the inverse camera
    matrix. Returns (scene, histogram of polygon sizes kept)."""
    from oracle import oracle as O
    lib = O.load()
    rng = np.random.default_rng(seed)
    mvp = np.asarray(scenes.camera_mvp((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 60.0, width / height), np.float32)
    m = mvp.reshape(4, 4)  # clip[r] = ((m[0,r] x + m[1,r] y) + m[2,r] z) + m[3,r], in float32
    inv = np.linalg.inv(m.astype(np.float64).T)
    kept, hist = [], {}
    for _ in range(candidates):
        w0 = rng.uniform(0.2, 5.0)
        corner = np.array([rng.choice([-1, 1]) * w0, rng.choice([-1, 1]) * w0, rng.uniform(-1, 1) * w0, w0])
        d = rng.normal(size=4)
        third = rng.uniform(-3, 3, 4)
        third[3] = abs(third[3]) + 0.1
        clip64 = [corner + rng.uniform(0.2, 3) * d, corner - rng.uniform(0.2, 3) * d, third]
        if min(c[3] for c in clip64) <= 0.01:
            continue
        obj = np.array([(inv @ c)[:3] / (inv @ c)[3] for c in clip64], np.float32)
        cl = np.array([[((m[0, r] * x + m[1, r] * y) + m[2, r] * z) + m[3, r] for r in range(4)] for x, y, z in obj],
                      np.float32)
        codes = [lib.sro_clip_codes(O.fptr(np.ascontiguousarray(c))) for c in cl]
        mor = codes[0] | codes[1] | codes[2]
        if mor == 0 or (codes[0] & codes[1] & codes[2]):
            continue
        ov, oa, at = np.zeros(36, np.float32), np.zeros(72, np.float32), np.zeros(24, np.float32)
        n = lib.sro_clip_triangle(O.fptr(np.ascontiguousarray(cl.reshape(-1))), O.fptr(at), mor, O.fptr(ov), O.fptr(oa))
        if n >= min_verts:
            kept.append(obj)
            hist[int(n)] = hist.get(int(n), 0) + 1
    pos = np.concatenate(kept)
    v = np.zeros((len(pos), 8), np.float32)
    v[:, 0:3] = pos
    v[:, 3:6] = (0.0, 0.0, 1.0)
    v[:, 6:8] = rng.uniform(-3, 3, (len(pos), 2))
    idx = np.arange(len(pos), dtype=np.uint32).reshape(-1, 3)
    flip = rng.random(len(idx)) < 0.5
    idx[flip] = idx[flip][:, ::-1]
    draw = scenes.Draw(v.view(scenes.VERTEX_DTYPE).reshape(-1), idx.reshape(-1).copy(), 0, shader, 6)
    tex = [rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)]
    return scenes.Scene(width, height, mvp, [draw], tex, {"config": "many_fan"}), dict(sorted(hist.items()))


def build_textures(scene):
    from softrast_amd import TextureData
    out = []
    for t in scene.textures:
        h, w = t.shape[:2]
        out.append(TextureData().CreateFromRGBA8(t, w, h, True))
    return out


def oracle_render(scene, texdata, nthreads=1, clear=0, plane_init=None, fast=False, tile_entries=False, sponza=None):
    """Render one frame of `scene` with the oracle; returns (plane, stats).
    sponza: SponzaConstants for SponzaShader draws."""
    from oracle.oracle import Oracle, Plane, make_texture
    o = Oracle(nthreads=nthreads, fast=fast)
    if sponza is not None:
        o.set_sponza_constants(sponza.to_bytes())
    p = Plane(scene.width, scene.height)
    if plane_init is not None:
        p.colour[:] = plane_init[0].reshape(-1)
        p.depth[:] = plane_init[1].reshape(-1)
    if clear is not None:
        o.clear(p, clear)
    stex = [make_texture(t) for t in texdata]
    draws = []
    for d in scene.draws:
        vf = d.vertices.view(np.float32).reshape(-1, 8)
        draws.append(dict(shader=d.shader, tex=stex[d.texture][0] if d.texture >= 0 else None, idx=d.indices,
                          idx_stride=d.indices.dtype.itemsize, pos=vf, pos_stride=ATTR_STRIDE, pos_num=len(vf),
                          attr=vf, attr_stride=ATTR_STRIDE, attr_num=len(vf), uv_offset=d.uv_offset, plane=0,
                          mvp=scene.mvp))
    stats = o.render(draws, [p], tile_entries=tile_entries)
    o.close()
    return p, stats


def gpu_render(scene, texdata, clear=0, plane_init=None, ctx=None, flags=None, fb=None, sponza=None, noalloc=False,
               **ctx_kw):
    """Render one frame on the GPU through the C ABI; returns (colour, depth, primid, stats).
    With fb given, renders into it and returns the stats only. sponza:
    SponzaConstants set before the frame. noalloc: end the frame with
    srg_end_frame_noalloc."""
    import softrast_amd as sr
    if flags is None:
        flags = sr.CFG_KEEP_PRIMID | sr.CFG_STATS
    own = ctx is None
    if own:
        ctx = sr.RenderContext(flags=flags, **ctx_kw)
    own_fb = fb is None
    if own_fb:
        fb = sr.FrameBuffer(ctx, scene.width, scene.height)
    if plane_init is not None:
        fb.write_tiles(plane_init[0], plane_init[1])
    if sponza is not None:
        ctx.set_sponza_constants(sponza)
    ctx.BeginFrame()
    if clear is not None:
        ctx.ClearFrameBuffer(fb, clear)
    for d in scene.draws:
        call = sr.DrawCall()
        call.SetFrameBuffer(fb).SetMVP(scene.mvp)
        call.SetIndexBuffer(d.indices, d.indices.dtype.itemsize, len(d.indices))
        call.SetPositionBuffer(d.vertices, ATTR_STRIDE, len(d.vertices))
        call.SetAttributeBuffer(d.vertices, ATTR_STRIDE, len(d.vertices), d.uv_offset)
        shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader, 2: sr.VisualizeUVsShader,
                  3: sr.SponzaShader}[d.shader]
        call.SetPixelShader(shader, texdata[d.texture] if d.texture >= 0 else None)
        ctx.DrawIndexed(call)
    ctx.EndFrame(noalloc=noalloc)
    stats = ctx.frame_stats()
    if not own_fb:
        return stats
    res = fb.read_tiles(primid=True) if flags & sr.CFG_KEEP_PRIMID else fb.read_tiles() + (None,)
    fb.destroy()
    if own:
        ctx.Shutdown()
    return res[0].reshape(-1), res[1].reshape(-1), (res[2].reshape(-1) if res[2] is not None else None), stats


def compare(oracle_plane, colour, depth, primid=None):
    """Differing-pixel counts: coverage (primid != none), primitive id, depth bits, RGBA8 (exact / >1 LSB)."""
    out = {}
    if primid is not None:
        out["coverage"] = int(np.count_nonzero((oracle_plane.primid != 0xFFFFFFFF) != (primid != 0xFFFFFFFF)))
        out["primid"] = int(np.count_nonzero(oracle_plane.primid != primid))
    out["depth"] = int(np.count_nonzero(oracle_plane.depth.view(np.uint32) != depth.view(np.uint32)))
    a = oracle_plane.colour.view(np.uint8).reshape(-1, 4).astype(np.int16)
    b = colour.view(np.uint8).reshape(-1, 4).astype(np.int16)
    diff = np.abs(a - b).max(axis=1)
    out["colour"] = int(np.count_nonzero(diff))
    out["colour_gt1"] = int(np.count_nonzero(diff > 1))
    return out


def oracle_render_views(viewset, texdata, nthreads=8, clear=0, sponza=None):
    """Every view of a scenes.ViewSet in one oracle frame, one plane per view
    (the order key runs over all views' draws, as in one EndFrame)."""
    from oracle.oracle import Oracle, Plane, make_texture
    o = Oracle(nthreads=nthreads)
    if sponza is not None:
        o.set_sponza_constants(sponza.to_bytes())
    planes = [Plane(v.width, v.height) for v in viewset.views]
    for p in planes:
        o.clear(p, clear)
    stex = [make_texture(t) for t in texdata]
    draws = []
    for vi, v in enumerate(viewset.views):
        for d in v.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            draws.append(dict(shader=d.shader, tex=stex[d.texture][0] if d.texture >= 0 else None, idx=d.indices,
                              idx_stride=d.indices.dtype.itemsize, pos=vf, pos_stride=ATTR_STRIDE, pos_num=len(vf),
                              attr=vf, attr_stride=ATTR_STRIDE, attr_num=len(vf), uv_offset=d.uv_offset, plane=vi,
                              mvp=v.mvp))
    stats = o.render(draws, planes)
    o.close()
    return planes, stats


def oracle_render_views_each(viewset, texdata, nthreads=16, clear=0, fast=False):
    """The oracle's image of every view of a scenes.ViewSet as one EndFrame
    draws them, rendered one view per oracle frame so that memory holds one
    view's bins at a time. A view rendered alone numbers its triangles from
    0; its primitive ids are shifted to the batch's numbering, (index of the
    view's first triangle among all views' draws) << 3, the order key's
    triangle field. Returns (planes, summed stats)."""
    from oracle.oracle import Oracle, Plane, make_texture
    o = Oracle(nthreads=nthreads, fast=fast)
    stex = [make_texture(t) for t in texdata]
    planes, total, base = [], {}, 0
    for v in viewset.views:
        p = Plane(v.width, v.height)
        o.clear(p, clear)
        draws = []
        for d in v.draws:
            vf = d.vertices.view(np.float32).reshape(-1, 8)
            draws.append(dict(shader=d.shader, tex=stex[d.texture][0] if d.texture >= 0 else None, idx=d.indices,
                              idx_stride=d.indices.dtype.itemsize, pos=vf, pos_stride=ATTR_STRIDE, pos_num=len(vf),
                              attr=vf, attr_stride=ATTR_STRIDE, attr_num=len(vf), uv_offset=d.uv_offset, plane=0,
                              mvp=v.mvp))
        st = o.render(draws, [p])
        covered = p.primid != 0xFFFFFFFF
        p.primid[covered] += np.uint32(base << 3)
        base += sum(len(d.indices) // 3 for d in v.draws)
        for k, val in st.items():
            total[k] = total.get(k, 0) + val
        planes.append(p)
    o.close()
    return planes, total


def gpu_render_views(viewset, texdata, ctx, fbs, clear=0, sponza=None):
    """Every view of a scenes.ViewSet in one EndFrame, view k into fbs[k];
    `texdata` holds TextureData or device textures by ViewSet texture index."""
    import softrast_amd as sr
    shader = {0: sr.UnlitDiffuseShader, 1: sr.VisualizeNormalsShader, 2: sr.VisualizeUVsShader,
              3: sr.SponzaShader}
    if sponza is not None:
        ctx.set_sponza_constants(sponza)
    ctx.BeginFrame()
    for v, fb in zip(viewset.views, fbs):
        ctx.ClearFrameBuffer(fb, clear)
        for d in v.draws:
            call = sr.DrawCall().SetFrameBuffer(fb).SetMVP(v.mvp)
            call.SetIndexBuffer(d.indices, d.indices.dtype.itemsize, len(d.indices))
            call.SetPositionBuffer(d.vertices, ATTR_STRIDE, len(d.vertices))
            call.SetAttributeBuffer(d.vertices, ATTR_STRIDE, len(d.vertices), d.uv_offset)
            call.SetPixelShader(shader[d.shader], texdata[d.texture] if d.texture >= 0 else None)
            ctx.DrawIndexed(call)
    ctx.EndFrame()
    return ctx.frame_stats()
