"""softrast_amd -- MI355X-native sort-middle rasteriser (SoftRast's hot path).

The compute path is libsrgpu.so (hand-written HIP for gfx950 behind the C ABI
in include/srgpu.h). This module binds that ABI with ctypes and mirrors the
reference's host interface (SoftRast/Renderer.h:110-177, Texture.h:21-41):

    ctx = RenderContext()
    fb = FrameBuffer(ctx, 1280, 720)
    ctx.BeginFrame()
    ctx.ClearFrameBuffer(fb, 0)
    call = DrawCall().SetFrameBuffer(fb).SetMVP(mvp) \\
        .SetIndexBuffer(indices, 2, n).SetPositionBuffer(verts, 32, nv) \\
        .SetAttributeBuffer(verts, 32, nv, 6).SetPixelShader(UnlitDiffuseShader, tex)
    ctx.DrawIndexed(call)
    ctx.EndFrame()
    ctx.Blit(fb, out_rgba8)

Errors raise SrgError carrying the library status. There is no CPU fallback:
if the library is missing or no GPU is present, construction fails.
"""
import ctypes
import os
import pathlib

import numpy as np

_PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsrgpu.so"

SRG_OK = 0
SHADER_UNLIT_DIFFUSE = 0
SHADER_VISUALIZE_NORMALS = 1
SHADER_VISUALIZE_UVS = 2
SHADER_SPONZA = 3
SPONZA_LIGHTS = 16
ORDER_NATURAL, ORDER_TRICOUNT, ORDER_LEARNED = 0, 1, 2
CFG_KEEP_PRIMID, CFG_STATS, CFG_TIMING, CFG_TILE_COSTS, CFG_GRAPH, CFG_AUTOTUNE = 1, 2, 4, 8, 16, 32
# back-end layouts (work-bin edge, workgroups per bin) the library accepts and the autotuner sweeps
LAYOUTS = ((8, 1), (16, 1), (32, 1), (32, 2), (32, 4), (64, 1), (64, 2), (64, 4))
FB_COLOUR, FB_DEPTH = 1, 2

SRG_SYMBOLS = (
    "srg_create", "srg_destroy", "srg_status_string", "srg_register_buffer", "srg_register_device_buffer",
    "srg_update_buffer", "srg_release_buffer", "srg_register_texture", "srg_release_texture",
    "srg_fb_create", "srg_fb_destroy", "srg_fb_write_plane", "srg_fb_swap_planes", "srg_fb_read_tiles",
    "srg_fb_write_tiles", "srg_begin_frame", "srg_clear", "srg_draw_indexed", "srg_end_frame", "srg_blit",
    "srg_get_frame_stats", "srg_get_bin_stats", "srg_sched_get_model", "srg_sched_set_model",
    "srg_texture_build_rgba8", "srg_texture_create_rgba8", "srg_texture_read", "srg_blit_async", "srg_blit_wait",
    "srg_get_tuning", "srg_draw_indexed_n", "srg_obj_load", "srg_obj_parse", "srg_obj_mesh_count",
    "srg_obj_get_mesh", "srg_obj_material_count", "srg_obj_material_name", "srg_obj_material_diffuse",
    "srg_obj_free", "srg_set_sponza_constants", "srg_clear_n", "srg_end_frame_noalloc", "srg_sched_eval",
    "srg_sched_train",
)
OBJ_FLIP_WINDING, OBJ_FLIP_UVS = 1, 4  # Obj::LoadFlags (Viewer/Obj.h:54-60)
PLANE_WRITE = 0xFFFFFFFF


class SrgError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


class SrgConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("device", ctypes.c_int32), ("flags", ctypes.c_uint32),
                ("tile_order", ctypes.c_uint32), ("max_prims", ctypes.c_uint32), ("max_entries", ctypes.c_uint32),
                ("wg_per_cu", ctypes.c_uint32), ("bin_size", ctypes.c_uint32),
                ("sched_warmup_frames", ctypes.c_uint32), ("sched_retrain_interval", ctypes.c_uint32),
                ("wg_per_tile", ctypes.c_uint32), ("autotune_frames", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 4)]


class SrgDraw(ctypes.Structure):
    _fields_ = [("shader", ctypes.c_uint32), ("texture", ctypes.c_uint64),
                ("index_buffer", ctypes.c_uint64), ("index_offset", ctypes.c_uint64),
                ("index_stride", ctypes.c_uint32), ("index_num", ctypes.c_uint32),
                ("position_buffer", ctypes.c_uint64), ("position_offset", ctypes.c_uint64),
                ("position_stride", ctypes.c_uint32), ("position_num", ctypes.c_uint32),
                ("attribute_buffer", ctypes.c_uint64), ("attribute_offset", ctypes.c_uint64),
                ("attribute_stride", ctypes.c_uint32), ("attribute_num", ctypes.c_uint32),
                ("uv_offset", ctypes.c_uint32), ("fb", ctypes.c_void_p), ("plane", ctypes.c_uint32),
                ("mvp", ctypes.c_float * 16)]


class SrgFrameStats(ctypes.Structure):
    _fields_ = [("tris_in", ctypes.c_uint64), ("tris_clipped", ctypes.c_uint64), ("prims", ctypes.c_uint64),
                ("bin_entries", ctypes.c_uint64), ("work_entries", ctypes.c_uint64), ("frags_raster", ctypes.c_uint64), ("tiles_queued", ctypes.c_uint32), ("retries", ctypes.c_uint32), ("graph_captures", ctypes.c_uint32),
                ("ms_frontend", ctypes.c_float), ("ms_backend", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("sched_err", ctypes.c_float), ("sched_trained", ctypes.c_uint32), ("sched_trainings", ctypes.c_uint32),
                ("bin_size", ctypes.c_uint32), ("wg_per_tile", ctypes.c_uint32), ("sched_fallback", ctypes.c_uint32)]


class SrgObjMesh(ctypes.Structure):
    _fields_ = [("index_stride", ctypes.c_uint32), ("num_indices", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("num_vertices", ctypes.c_uint32), ("vertices", ctypes.c_void_p), ("material", ctypes.c_uint32)]


TUNE_LAYOUTS = 8


class SrgTuneLayout(ctypes.Structure):
    _fields_ = [("bin_size", ctypes.c_uint32), ("wg_per_tile", ctypes.c_uint32), ("ms", ctypes.c_float)]


class SrgTuning(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint32), ("bin_size", ctypes.c_uint32), ("wg_per_tile", ctypes.c_uint32),
                ("nlayouts", ctypes.c_uint32), ("layouts", SrgTuneLayout * TUNE_LAYOUTS)]


_lib = None


def load_library(path=None):
    """Load libsrgpu.so and declare its prototypes. Raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run softrast_amd/build.py (no CPU fallback exists)")
    lib = ctypes.CDLL(str(p))
    vp, u32, i32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_size_t
    pu64 = ctypes.POINTER(ctypes.c_uint64)
    sigs = {
        "srg_create": (i32, [ctypes.POINTER(SrgConfig), ctypes.POINTER(vp)]),
        "srg_destroy": (None, [vp]),
        "srg_status_string": (ctypes.c_char_p, [i32]),
        "srg_register_buffer": (i32, [vp, vp, sz, pu64]),
        "srg_register_device_buffer": (i32, [vp, vp, sz, pu64]),
        "srg_update_buffer": (i32, [vp, u64, sz, vp, sz]),
        "srg_release_buffer": (i32, [vp, u64]),
        "srg_register_texture": (i32, [vp, vp, u64, ctypes.POINTER(u32), u32, u32, u32, pu64]),
        "srg_texture_create_rgba8": (i32, [vp, vp, u32, u32, i32, pu64]),
        "srg_blit_async": (i32, [vp, vp, vp, u32, vp, vp]),
        "srg_blit_wait": (i32, [vp, vp]),
        "srg_texture_read": (i32, [vp, u64, vp, u64, pu64, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                   ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "srg_release_texture": (i32, [vp, u64]),
        "srg_fb_create": (i32, [vp, u32, u32, u32, ctypes.POINTER(vp)]),
        "srg_fb_destroy": (None, [vp, vp]),
        "srg_fb_write_plane": (u32, [vp]),
        "srg_fb_swap_planes": (None, [vp]),
        "srg_fb_read_tiles": (i32, [vp, vp, u32, vp, vp, vp]),
        "srg_fb_write_tiles": (i32, [vp, vp, u32, vp, vp]),
        "srg_begin_frame": (i32, [vp]),
        "srg_clear": (i32, [vp, vp, u32, i32, i32]),
        "srg_draw_indexed": (i32, [vp, ctypes.POINTER(SrgDraw)]),
        "srg_end_frame": (i32, [vp]),
        "srg_end_frame_noalloc": (i32, [vp]),
        "srg_blit": (i32, [vp, vp, vp, vp, vp]),
        "srg_get_frame_stats": (i32, [vp, ctypes.POINTER(SrgFrameStats)]),
        "srg_get_bin_stats": (i32, [vp, vp, vp, u32, ctypes.POINTER(u32)]),
        "srg_sched_get_model": (i32, [vp, vp, u32]),
        "srg_sched_set_model": (i32, [vp, vp, u32]),
        "srg_sched_eval": (i32, [vp, vp, u32, vp]),
        "srg_sched_train": (i32, [vp, vp, u32, u32, ctypes.c_float, u32]),
        "srg_get_tuning": (i32, [vp, ctypes.POINTER(SrgTuning)]),
        "srg_draw_indexed_n": (i32, [vp, vp, u32]),
        "srg_obj_load": (i32, [ctypes.c_char_p, u32, ctypes.POINTER(vp)]),
        "srg_obj_parse": (i32, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, u32, ctypes.POINTER(vp)]),
        "srg_obj_mesh_count": (u32, [vp]),
        "srg_obj_get_mesh": (i32, [vp, u32, ctypes.POINTER(SrgObjMesh)]),
        "srg_obj_material_count": (u32, [vp]),
        "srg_obj_material_name": (ctypes.c_char_p, [vp, u32]),
        "srg_obj_material_diffuse": (ctypes.c_char_p, [vp, u32]),
        "srg_obj_free": (None, [vp]),
        "srg_set_sponza_constants": (i32, [vp, vp]),
        "srg_clear_n": (i32, [vp, vp, u32, u32, i32, i32]),
        "srg_texture_build_rgba8": (i32, [vp, u32, u32, i32, vp, pu64, ctypes.POINTER(u32),
                                          ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def status_string(status):
    try:
        return load_library().srg_status_string(status).decode()
    except Exception:
        return "unknown"


def _check(status, what):
    if status != SRG_OK:
        raise SrgError(status, what)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


# ---------------------------------------------------------------------------
# textures: Tex::TextureData (Texture.h:21-41)
# ---------------------------------------------------------------------------

class TextureData:
    """Tiled, Morton-swizzled RGBA8 mip chain (Texture.cpp:119-199)."""

    def __init__(self):
        self.m_texels = np.zeros(0, np.uint8)
        self.m_mipOffsets = np.zeros(14, np.uint32)
        self.m_widthLog2 = 0
        self.m_heightLog2 = 0
        self.m_numMips = 0
        self.m_bytesPerPixel = 0

    def CreateFromRGBA8(self, texels, width, height, calcMips=False):
        lib = load_library()
        src = np.ascontiguousarray(texels, dtype=np.uint8).reshape(-1)
        if src.size != width * height * 4:
            raise ValueError("texel array does not match width*height*4")
        nbytes = ctypes.c_uint64()
        offs = (ctypes.c_uint32 * 14)()
        wl, hl, nm = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.srg_texture_build_rgba8(_ptr(src), width, height, int(calcMips), None, ctypes.byref(nbytes), offs,
                                           ctypes.byref(wl), ctypes.byref(hl), ctypes.byref(nm)), "CreateFromRGBA8")
        out = np.zeros(nbytes.value, np.uint8)
        _check(lib.srg_texture_build_rgba8(_ptr(src), width, height, int(calcMips), _ptr(out), ctypes.byref(nbytes),
                                           offs, ctypes.byref(wl), ctypes.byref(hl), ctypes.byref(nm)),
               "CreateFromRGBA8")
        self.m_texels = out
        self.m_mipOffsets = np.array(list(offs), np.uint32)
        self.m_widthLog2, self.m_heightLog2, self.m_numMips = wl.value, hl.value, nm.value
        self.m_bytesPerPixel = 4
        return self

    def Clear(self):
        self.m_texels = np.zeros(0, np.uint8)


class DeviceTexture:
    """A texture built and kept on the device: TextureData::CreateFromRGBA8
    run on the GPU (srg_texture_create_rgba8). Usable wherever a TextureData
    is, as the UnlitDiffuse uniforms of a draw of the same context."""

    def __init__(self, ctx, texels, width, height, calcMips=True):
        src = np.ascontiguousarray(texels, dtype=np.uint8).reshape(-1)
        if src.size != width * height * 4:
            raise ValueError("texel array does not match width*height*4")
        out = ctypes.c_uint64()
        _check(ctx._lib.srg_texture_create_rgba8(ctx._ctx, _ptr(src), width, height, int(calcMips), ctypes.byref(out)),
               "texture_create_rgba8")
        self._ctx, self.handle = ctx, out.value

    def release(self):
        """Frees the device copy (srg_release_texture). A frame that drew with
        it and has not ended fails at EndFrame with SRG_ERR_HANDLE."""
        if self.handle:
            _check(self._ctx._lib.srg_release_texture(self._ctx._ctx, self.handle), "release_texture")
            self.handle = 0

    def read(self):
        """Host copy as a TextureData (tiled bytes and layout)."""
        lib, c = self._ctx._lib, self._ctx._ctx
        nbytes = ctypes.c_uint64()
        offs = (ctypes.c_uint32 * 14)()
        wl, hl, nm = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.srg_texture_read(c, self.handle, None, 0, ctypes.byref(nbytes), offs, ctypes.byref(wl),
                                    ctypes.byref(hl), ctypes.byref(nm)), "texture_read")
        data = np.zeros(nbytes.value, np.uint8)
        _check(lib.srg_texture_read(c, self.handle, _ptr(data), data.nbytes, ctypes.byref(nbytes), offs,
                                    ctypes.byref(wl), ctypes.byref(hl), ctypes.byref(nm)), "texture_read")
        t = TextureData()
        t.m_texels, t.m_mipOffsets = data, np.array(list(offs), np.uint32)
        t.m_widthLog2, t.m_heightLog2, t.m_numMips, t.m_bytesPerPixel = wl.value, hl.value, nm.value, 4
        return t


# ---------------------------------------------------------------------------
# shaders: Viewer/Shaders.h. Host shader callbacks cannot run on the device;
# these names map to the device implementations.
# ---------------------------------------------------------------------------

class _Shader:
    def __init__(self, name, sid):
        self.name, self.id = name, sid

    def __repr__(self):
        return self.name


UnlitDiffuseShader = _Shader("UnlitDiffuseShader", SHADER_UNLIT_DIFFUSE)
VisualizeNormalsShader = _Shader("VisualizeNormalsShader", SHADER_VISUALIZE_NORMALS)
VisualizeUVsShader = _Shader("VisualizeUVsShader", SHADER_VISUALIZE_UVS)
SponzaShader = _Shader("SponzaShader", SHADER_SPONZA)
_SAMPLING_SHADERS = (SHADER_UNLIT_DIFFUSE, SHADER_SPONZA)


class SponzaConstants:
    """SponzaScene::Constants (Viewer/SponzaScene.h:15-47) as the SponzaShader
    reads it: sun_dir and ambient (3 floats each), 16 point lights of
    (pos[3], colour[3], intensity, falloff). to_bytes() is the
    srg_sponza_constants / sro_sponza_constants layout."""

    def __init__(self, sun_dir=(0.0, 0.0, 0.0), ambient=(0.0, 0.0, 0.0), lights=()):
        self.sun_dir = np.asarray(sun_dir, np.float32).reshape(3)
        self.ambient = np.asarray(ambient, np.float32).reshape(3)
        self.lights = np.zeros((SPONZA_LIGHTS, 8), np.float32)
        for i, light in enumerate(lights):
            pos, colour, intensity = light[0], light[1], light[2]
            falloff = light[3] if len(light) > 3 else 0.0
            self.lights[i] = [*pos, *colour, intensity, falloff]

    def to_bytes(self):
        return np.concatenate([self.sun_dir, self.ambient, self.lights.reshape(-1)]).astype(np.float32).tobytes()


# ---------------------------------------------------------------------------
# frame API: RenderContext / FrameBuffer / DrawCall (Renderer.h)
# ---------------------------------------------------------------------------

class GenericDrawBuffer:
    def __init__(self):
        self.m_ptr = None
        self.m_num = 0
        self.m_stride = 0


class DrawCall:
    """Per-draw state (Renderer.h:119-150). m_ptr holds a numpy array."""

    UV_OFFSET_INVALID = 0xFFFFFFFF

    def __init__(self):
        self.m_pixelShader = None
        self.m_pixelUniforms = None
        self.m_indexBuffer = GenericDrawBuffer()
        self.m_positionBuffer = GenericDrawBuffer()
        self.m_attributeBuffer = GenericDrawBuffer()
        self.m_uvOffset = 0
        self.m_frameBuffer = None     # (FrameBuffer, plane)
        self.m_mvp = np.eye(4, dtype=np.float32)  # column-major as m_mvp[col][row]
        self.m_drawCallIdx = 0

    def SetPixelShader(self, fn, uniforms):
        self.m_pixelShader, self.m_pixelUniforms = fn, uniforms
        return self

    def SetIndexBuffer(self, buf, stride, num):
        self.m_indexBuffer.m_ptr, self.m_indexBuffer.m_stride, self.m_indexBuffer.m_num = buf, stride, num
        return self

    def SetPositionBuffer(self, buf, stride, num):
        self.m_positionBuffer.m_ptr, self.m_positionBuffer.m_stride, self.m_positionBuffer.m_num = buf, stride, num
        return self

    def SetAttributeBuffer(self, buf, stride, num, uvOffset=0):
        self.m_uvOffset = uvOffset
        b = self.m_attributeBuffer
        b.m_ptr, b.m_stride, b.m_num = buf, stride, num
        return self

    def SetFrameBuffer(self, fb):
        self.m_frameBuffer = (fb, fb.m_writePlane)
        return self

    def SetMVP(self, mvp):
        """mvp: 16 floats column-major, or a (4, 4) array indexed [col][row]."""
        self.m_mvp = np.asarray(mvp, dtype=np.float32).reshape(4, 4)
        return self


class FrameBuffer:
    """Double-buffered tiled RGBA8 + fp32 depth (Renderer.h:75-108)."""

    def __init__(self, ctx, width, height, colour=True, depth=True):
        self._ctx = ctx
        self.m_width, self.m_height = width, height
        self.tiles_x, self.tiles_y = (width + 63) // 64, (height + 63) // 64
        h = ctypes.c_void_p()
        flags = (FB_COLOUR if colour else 0) | (FB_DEPTH if depth else 0)
        _check(ctx._lib.srg_fb_create(ctx._ctx, width, height, flags, ctypes.byref(h)), "FrameBuffer")
        self._h = h

    @property
    def m_writePlane(self):
        return self._ctx._lib.srg_fb_write_plane(self._h)

    def SwapPlanes(self):
        self._ctx._lib.srg_fb_swap_planes(self._h)

    def read_tiles(self, plane=None, primid=False):
        """Tiled (colour u32, depth f32[, primitive order u32]) of a plane, shape (tiles, 64, 64)."""
        plane = self.m_writePlane if plane is None else plane
        n = self.tiles_x * self.tiles_y * 4096
        col = np.empty(n, np.uint32)
        dep = np.empty(n, np.float32)
        pid = np.empty(n, np.uint32) if primid else None
        _check(self._ctx._lib.srg_fb_read_tiles(self._ctx._ctx, self._h, plane, _ptr(col), _ptr(dep), _ptr(pid)),
               "read_tiles")
        shp = (self.tiles_x * self.tiles_y, 64, 64)
        out = (col.reshape(shp), dep.reshape(shp))
        return out + (pid.reshape(shp),) if primid else out

    def write_tiles(self, colour=None, depth=None, plane=None):
        plane = self.m_writePlane if plane is None else plane
        c = None if colour is None else np.ascontiguousarray(colour, np.uint32).reshape(-1)
        d = None if depth is None else np.ascontiguousarray(depth, np.float32).reshape(-1)
        _check(self._ctx._lib.srg_fb_write_tiles(self._ctx._ctx, self._h, plane, _ptr(c), _ptr(d)), "write_tiles")

    def destroy(self):
        if self._h:
            self._ctx._lib.srg_fb_destroy(self._ctx._ctx, self._h)
            self._h = None


class ObjMesh:
    """Obj::Mesh (Viewer/Obj.h:24-42): `vertices` in Obj::Vertex layout
    (pos3, norm3, uv2; scenes.VERTEX_DTYPE), `indices` uint16 or uint32."""

    def __init__(self, vertices, indices, material):
        self.vertices, self.indices, self.material = vertices, indices, material


class ObjModel:
    """Obj::Model::Load (Viewer/Obj.cpp:374-567) through srg_obj_load /
    srg_obj_parse: `meshes` (ObjMesh) and `materials` ((name, map_Kd path))."""

    def __init__(self, handle, lib):
        from . import scenes
        try:
            self.meshes = []
            for i in range(lib.srg_obj_mesh_count(handle)):
                m = SrgObjMesh()
                _check(lib.srg_obj_get_mesh(handle, i, ctypes.byref(m)), "obj mesh")
                dt = np.uint16 if m.index_stride == 2 else np.uint32
                idx = np.ctypeslib.as_array((ctypes.c_uint8 * (m.num_indices * m.index_stride)).from_address(m.indices)) \
                    .view(dt).copy() if m.num_indices else np.zeros(0, dt)
                vtx = np.ctypeslib.as_array((ctypes.c_uint8 * (m.num_vertices * 32)).from_address(m.vertices)) \
                    .view(scenes.VERTEX_DTYPE).copy() if m.num_vertices else np.zeros(0, scenes.VERTEX_DTYPE)
                self.meshes.append(ObjMesh(vtx, idx, m.material))
            self.materials = [(lib.srg_obj_material_name(handle, i).decode(),
                               lib.srg_obj_material_diffuse(handle, i).decode())
                              for i in range(lib.srg_obj_material_count(handle))]
        finally:
            lib.srg_obj_free(handle)


def obj_load(path, flags=0):
    """Loads a Wavefront OBJ file as the reference's viewer does (flags:
    OBJ_FLIP_WINDING, OBJ_FLIP_UVS). Raises SrgError on a parse error."""
    lib = load_library()
    h = ctypes.c_void_p()
    _check(lib.srg_obj_load(str(path).encode(), flags, ctypes.byref(h)), "obj_load")
    return ObjModel(h, lib)


def obj_parse(text, flags=0, directory=None):
    """obj_load on OBJ text in memory; mtllib files are read from `directory`."""
    lib = load_library()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    _check(lib.srg_obj_parse(data, len(data), str(directory).encode() if directory else None, flags,
                             ctypes.byref(h)), "obj_parse")
    return ObjModel(h, lib)


class DrawList:
    """DrawCalls prepared once (buffers registered, srg_draw records built)
    and submitted each frame with one library call. Each call draws to its
    framebuffer's write plane at submission (SRG_PLANE_WRITE), so the list
    stays valid across SwapPlanes / Blit. Build a new list after changing a
    call or invalidating one of its buffers."""

    def __init__(self, ctx, calls):
        self._calls = list(calls)  # keeps the calls' arrays alive
        self._n = len(self._calls)
        self._arr = (SrgDraw * max(1, self._n))()
        for k, c in enumerate(self._calls):
            if c.m_indexBuffer.m_ptr is None:
                raise SrgError(-1, "DrawList: no index buffer bound")
            d = ctx._prepare_draw(c, np.asarray(c.m_mvp, np.float32).reshape(-1))
            d.fb, d.plane = c.m_frameBuffer[0]._h, PLANE_WRITE
            self._arr[k] = d

    def __len__(self):
        return self._n


class RenderContext:
    """RenderContext (Renderer.h:153-177) over libsrgpu.

    Host arrays passed in DrawCalls are uploaded once and cached by
    (address, size, dtype); call invalidate(array) after changing one.
    """

    def __init__(self, device=0, flags=0, tile_order=ORDER_NATURAL, max_prims=0, max_entries=0, wg_per_cu=0,
                 bin_size=0, sched_warmup_frames=0, sched_retrain_interval=0, wg_per_tile=0, autotune_frames=0,
                 lib_path=None):
        self._lib = load_library(lib_path)
        cfg = SrgConfig()
        cfg.struct_size = ctypes.sizeof(SrgConfig)
        cfg.device, cfg.flags, cfg.tile_order = device, flags, tile_order
        cfg.max_prims, cfg.max_entries, cfg.wg_per_cu = max_prims, max_entries, wg_per_cu
        cfg.bin_size = bin_size
        cfg.sched_warmup_frames, cfg.sched_retrain_interval = sched_warmup_frames, sched_retrain_interval
        cfg.wg_per_tile, cfg.autotune_frames = wg_per_tile, autotune_frames
        h = ctypes.c_void_p()
        _check(self._lib.srg_create(ctypes.byref(cfg), ctypes.byref(h)), "RenderContext")
        self._ctx = h
        self._buffers = {}
        self._textures = {}
        self._keep = []
        self._blit_keep = {}
        self._gen = 0
        self._draws = []

    # -- resources ---------------------------------------------------------
    def buffer_handle(self, arr):
        arr = np.ascontiguousarray(arr)
        key = (arr.ctypes.data, arr.nbytes, arr.dtype.str)
        h = self._buffers.get(key)
        if h is None:
            out = ctypes.c_uint64()
            _check(self._lib.srg_register_buffer(self._ctx, _ptr(arr), arr.nbytes, ctypes.byref(out)), "register_buffer")
            h = out.value
            self._buffers[key] = h
            self._keep.append(arr)
        return h

    def invalidate(self, arr):
        key = (arr.ctypes.data, arr.nbytes, arr.dtype.str)
        h = self._buffers.pop(key, None)
        if h is not None:
            self._lib.srg_release_buffer(self._ctx, h)
        self._gen += 1  # draws prepared with the old handle are rebuilt

    def texture_handle(self, tex):
        if tex is None:
            return 0
        if isinstance(tex, DeviceTexture):
            if tex._ctx is not self:
                raise SrgError(-1, "DeviceTexture belongs to another RenderContext")
            return tex.handle
        key = id(tex)
        h = self._textures.get(key)
        if h is None:
            out = ctypes.c_uint64()
            offs = (ctypes.c_uint32 * 14)(*[int(x) for x in tex.m_mipOffsets])
            data = np.ascontiguousarray(tex.m_texels, np.uint8)
            _check(self._lib.srg_register_texture(self._ctx, _ptr(data), data.nbytes, offs, tex.m_widthLog2,
                                                  tex.m_heightLog2, tex.m_numMips, ctypes.byref(out)),
                   "register_texture")
            h = out.value
            self._textures[key] = h
            self._keep.append(tex)
        return h

    # -- frame -------------------------------------------------------------
    def BeginFrame(self):
        self._draws = []
        _check(self._lib.srg_begin_frame(self._ctx), "BeginFrame")

    def ClearFrameBuffer(self, fb, color=0, clearColour=True, clearDepth=True):
        _check(self._lib.srg_clear(self._ctx, fb._h, color & 0xFFFFFFFF, int(clearColour), int(clearDepth)),
               "ClearFrameBuffer")

    def PrepareClearList(self, fbs):
        """The framebuffers of a batch of views, for ClearFrameBuffers."""
        fbs = list(fbs)
        arr = (ctypes.c_void_p * max(1, len(fbs)))(*[fb._h for fb in fbs])
        return (arr, len(fbs), fbs)

    def ClearFrameBuffers(self, clear_list, color=0, clearColour=True, clearDepth=True):
        """ClearFrameBuffer for every framebuffer of a PrepareClearList, in one
        library call (srg_clear_n)."""
        arr, n, _ = clear_list
        _check(self._lib.srg_clear_n(self._ctx, arr, n, color & 0xFFFFFFFF, int(clearColour), int(clearDepth)),
               "ClearFrameBuffers")

    def DrawIndexed(self, call):
        ib, pb, ab = call.m_indexBuffer, call.m_positionBuffer, call.m_attributeBuffer
        if ib.m_ptr is None:
            raise SrgError(-1, "DrawIndexed: no index buffer bound")
        mvp = np.asarray(call.m_mvp, np.float32).reshape(-1)
        # the srg_draw of a DrawCall is rebuilt only when its state changed
        key = (id(ib.m_ptr), ib.m_stride, ib.m_num, id(pb.m_ptr), pb.m_stride, pb.m_num, id(ab.m_ptr), ab.m_stride,
               ab.m_num, call.m_uvOffset, id(call.m_pixelShader), id(call.m_pixelUniforms), mvp.tobytes(), self._gen)
        cached = call.__dict__.get("_srg_draw")
        if cached is not None and cached[0] is self and cached[1] == key:
            d = cached[2]
        else:
            d = self._prepare_draw(call, mvp)
            call._srg_draw = (self, key, d)
        fb, plane = call.m_frameBuffer
        d.fb, d.plane = fb._h, plane
        _check(self._lib.srg_draw_indexed(self._ctx, ctypes.byref(d)), "DrawIndexed")
        call.m_drawCallIdx = len(self._draws)
        self._draws.append(call)

    def PrepareDrawList(self, calls):
        """A DrawList of these DrawCalls (see DrawList)."""
        return DrawList(self, calls)

    def DrawIndexedList(self, draws):
        """DrawIndexed for every call of a DrawList, in order, with one
        library call (srg_draw_indexed_n)."""
        _check(self._lib.srg_draw_indexed_n(self._ctx, ctypes.byref(draws._arr), draws._n), "DrawIndexedList")
        self._draws.extend(draws._calls)

    def _prepare_draw(self, call, mvp):
        d = SrgDraw()
        shader = call.m_pixelShader if call.m_pixelShader is not None else VisualizeNormalsShader
        d.shader = shader.id
        d.texture = self.texture_handle(call.m_pixelUniforms) if shader.id in _SAMPLING_SHADERS else 0
        if isinstance(call.m_pixelUniforms, TextureData) and shader.id in _SAMPLING_SHADERS and \
                call.m_pixelUniforms.m_texels.size == 0:
            d.texture = 0
        for name, buf in (("index", call.m_indexBuffer), ("position", call.m_positionBuffer),
                          ("attribute", call.m_attributeBuffer)):
            setattr(d, name + "_buffer", self.buffer_handle(buf.m_ptr))
            setattr(d, name + "_offset", 0)
            setattr(d, name + "_stride", buf.m_stride)
            setattr(d, name + "_num", buf.m_num)
        d.uv_offset = call.m_uvOffset
        d.mvp[:] = [float(x) for x in mvp]
        return d

    def EndFrame(self, noalloc=False):
        """EndFrame; with noalloc, srg_end_frame_noalloc: the frame allocates
        nothing and works in the library's fixed per-GPU scratch (entry-count
        order, no graph, no autotuner sweep; SrgError with SRG_ERR_OVERFLOW
        when the frame does not fit)."""
        if noalloc:
            _check(self._lib.srg_end_frame_noalloc(self._ctx), "EndFrame(noalloc)")
        else:
            _check(self._lib.srg_end_frame(self._ctx), "EndFrame")

    def Blit(self, fb, linear=None, onFinish=None, user=None):
        out = linear if linear is not None else np.empty((fb.m_height, fb.m_width), np.uint32)
        if out.nbytes < fb.m_width * fb.m_height * 4 or not out.flags.c_contiguous:
            raise ValueError("Blit target too small or not contiguous")
        _check(self._lib.srg_blit(self._ctx, fb._h, _ptr(out), None, None), "Blit")
        if onFinish is not None:
            onFinish(user)
        return out

    _BLIT_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p)

    def BlitAsync(self, fb, linear=None, onFinish=None, user=None, bgra=False):
        """Blit with the reference's asynchrony (srg_blit_async): returns the
        target array at once; onFinish(user) runs from a runtime thread when
        the rows are in it. The planes are swapped at once."""
        out = linear if linear is not None else np.empty((fb.m_height, fb.m_width), np.uint32)
        if out.nbytes < fb.m_width * fb.m_height * 4 or not out.flags.c_contiguous:
            raise ValueError("Blit target too small or not contiguous")
        cb = self._BLIT_CB((lambda _p: onFinish(user)) if onFinish is not None else (lambda _p: None))
        self._blit_keep.setdefault(id(fb), []).append((out, cb))  # alive until BlitWait
        _check(self._lib.srg_blit_async(self._ctx, fb._h, _ptr(out), 1 if bgra else 0, cb, None), "BlitAsync")
        return out

    def BlitWait(self, fb):
        _check(self._lib.srg_blit_wait(self._ctx, fb._h), "BlitWait")
        self._blit_keep.pop(id(fb), None)

    def frame_stats(self):
        s = SrgFrameStats()
        _check(self._lib.srg_get_frame_stats(self._ctx, ctypes.byref(s)), "frame_stats")
        return {k: getattr(s, k) for k, _ in SrgFrameStats._fields_}

    def bin_stats(self):
        """(counts, costs) per work bin of the last frame."""
        n = ctypes.c_uint32()
        _check(self._lib.srg_get_bin_stats(self._ctx, None, None, 0, ctypes.byref(n)), "bin_stats")
        counts = np.zeros(n.value, np.uint32)
        costs = np.zeros(n.value, np.uint64)
        _check(self._lib.srg_get_bin_stats(self._ctx, _ptr(counts), _ptr(costs), n.value, ctypes.byref(n)),
               "bin_stats")
        return counts, costs

    SCHED_PARAMS = 1249

    def CreateTexture(self, texels, width, height, calcMips=True):
        """TextureData::CreateFromRGBA8 on the device; see DeviceTexture."""
        return DeviceTexture(self, texels, width, height, calcMips)

    def sched_model(self):
        """Learned tile-cost model parameters (the scheduler's checkpoint)."""
        p = np.zeros(self.SCHED_PARAMS, np.float32)
        _check(self._lib.srg_sched_get_model(self._ctx, _ptr(p), p.size), "sched_model")
        return p

    def set_sched_model(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(self._lib.srg_sched_set_model(self._ctx, _ptr(p), p.size), "set_sched_model")

    def set_sponza_constants(self, consts):
        """SponzaShader constants for frames ended after this call
        (SponzaConstants, or bytes in its layout)."""
        raw = consts.to_bytes() if isinstance(consts, SponzaConstants) else bytes(consts)
        buf = np.frombuffer(raw, np.float32).copy()
        if buf.size != 6 + SPONZA_LIGHTS * 8:
            raise ValueError("SponzaScene::Constants is 134 floats")
        _check(self._lib.srg_set_sponza_constants(self._ctx, _ptr(buf)), "set_sponza_constants")

    def sched_eval(self, features):
        """The model's predicted log2 cost of rows of 4 features (srg_sched_eval)."""
        x = np.ascontiguousarray(features, np.float32).reshape(-1, 4)
        out = np.zeros(len(x), np.float32)
        _check(self._lib.srg_sched_eval(self._ctx, _ptr(x), len(x), _ptr(out)), "sched_eval")
        return out

    def sched_train(self, samples, steps, lr, seed):
        """Runs the device's Adam on rows of (4 features, log2 cost) from a
        cleared optimiser state (srg_sched_train)."""
        x = np.ascontiguousarray(samples, np.float32).reshape(-1, 5)
        _check(self._lib.srg_sched_train(self._ctx, _ptr(x), len(x), steps, lr, seed), "sched_train")

    def tuning(self):
        """Autotuner state for the last frame's signature: {"state": "off" |
        "sweeping" | "chosen", "layout": (bin_size, wg_per_tile) or None,
        "ms": {(bin_size, wg_per_tile): best device ms, 0 = not measured}}."""
        t = SrgTuning()
        _check(self._lib.srg_get_tuning(self._ctx, ctypes.byref(t)), "tuning")
        return {"state": ("off", "sweeping", "chosen")[t.state],
                "layout": (t.bin_size, t.wg_per_tile) if t.state == 2 else None,
                "ms": {(l.bin_size, l.wg_per_tile): l.ms for l in t.layouts[:t.nlayouts]}}

    def Shutdown(self):
        if self._ctx:
            self._lib.srg_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.Shutdown()
        except Exception:
            pass
