"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg. See sr_oracle.c for the
reference functions each part restates and how each is pinned.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REFERENCE_ROOT = "/root/reference"  # reference sources, present only where oracle/_ref is built
MAX_MIPS = 14


class SroTexture(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("size", ctypes.c_uint64), ("mip_offsets", ctypes.c_uint32 * MAX_MIPS),
                ("wlog2", ctypes.c_uint32), ("hlog2", ctypes.c_uint32), ("nmips", ctypes.c_uint32)]


class SroDraw(ctypes.Structure):
    _fields_ = [("shader", ctypes.c_uint32), ("tex", ctypes.POINTER(SroTexture)),
                ("idx", ctypes.c_void_p), ("idx_num", ctypes.c_uint32), ("idx_stride", ctypes.c_uint32),
                ("pos", ctypes.c_void_p), ("pos_num", ctypes.c_uint32), ("pos_stride", ctypes.c_uint32),
                ("attr", ctypes.c_void_p), ("attr_num", ctypes.c_uint32), ("attr_stride", ctypes.c_uint32),
                ("uv_offset", ctypes.c_uint32), ("plane", ctypes.c_uint32), ("mvp", ctypes.c_float * 16)]


class SroPlane(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("tiles_x", ctypes.c_uint32),
                ("tiles_y", ctypes.c_uint32), ("colour", ctypes.c_void_p), ("depth", ctypes.c_void_p),
                ("primid", ctypes.c_void_p)]


class SroStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("tris_in", "tris_clipped", "prims_binned", "bin_entries", "frags_raster", "frags_passed",
                 "blocks_tested")]


_libs = {}


def build(quiet=True):
    """Compile both oracle builds with the committed Makefile."""
    res = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


def load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = HERE / name
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    vp, u32, i32, f = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
    lib.sro_create.restype = vp
    lib.sro_create.argtypes = [u32]
    lib.sro_destroy.argtypes = [vp]
    lib.sro_destroy.restype = None
    lib.sro_clear.argtypes = [ctypes.POINTER(SroPlane), u32, i32, i32]
    lib.sro_clear.restype = None
    lib.sro_render.argtypes = [vp, ctypes.POINTER(SroDraw), u32, ctypes.POINTER(SroPlane), u32,
                               ctypes.POINTER(SroStats), vp]
    lib.sro_render.restype = i32
    lib.sro_blit.argtypes = [ctypes.POINTER(SroPlane), vp]
    lib.sro_blit.restype = None
    lib.sro_rcpps.argtypes = [f]
    lib.sro_rcpps.restype = f
    lib.sro_rsqrtps.argtypes = [f]
    lib.sro_rsqrtps.restype = f
    lib.sro_set_sponza_constants.argtypes = [vp, vp]
    lib.sro_pack_rgba.argtypes = [f, f, f, f]
    lib.sro_pack_rgba.restype = u32
    lib.sro_morton5.argtypes = [u32, u32]
    lib.sro_morton5.restype = u32
    lib.sro_calc_mip.argtypes = [ctypes.POINTER(SroTexture), f, f, f, f]
    lib.sro_calc_mip.restype = ctypes.c_int32
    lib.sro_sample_wrap.argtypes = [ctypes.POINTER(SroTexture), f, f, f, f, f, f, ctypes.POINTER(f)]
    lib.sro_sample_wrap.restype = u32
    lib.sro_transform.argtypes = [ctypes.POINTER(f), ctypes.POINTER(f), ctypes.POINTER(f)]
    lib.sro_transform.restype = None
    lib.sro_clip_codes.argtypes = [ctypes.POINTER(f)]
    lib.sro_clip_codes.restype = u32
    lib.sro_clip_triangle.argtypes = [ctypes.POINTER(f)] * 2 + [u32] + [ctypes.POINTER(f)] * 2
    lib.sro_clip_triangle.restype = u32
    lib.sro_setup_triangle.argtypes = [ctypes.POINTER(f), ctypes.POINTER(f), u32, u32, u32,
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(f)]
    lib.sro_setup_triangle.restype = i32
    lib.sro_mip_resize_rgba8.argtypes = [vp, u32, u32, vp, u32, u32]
    lib.sro_mip_resize_rgba8.restype = i32
    _libs[name] = lib
    return lib


def mip_resize(img, ow, oh):
    """One mip level of an (h, w, 4) uint8 image as the reference makes it
    (stbir_resize_uint8, Mitchell); returns (oh, ow, 4) uint8."""
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros((oh, ow, 4), np.uint8)
    if load().sro_mip_resize_rgba8(img.ctypes.data, img.shape[1], img.shape[0], out.ctypes.data, ow, oh) != 0:
        raise ValueError("sro_mip_resize_rgba8: unsupported size")
    return out


def load_reference_stbir():
    """The reference's stb_image_resize compiled as is (oracle/_ref), or None
    where it was not built (no reference sources on this machine)."""
    path = HERE / "_ref" / "libstbir.so"
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.stbir_resize_uint8.argtypes = [vp, i, i, i, vp, i, i, i, i]
    lib.stbir_resize_uint8.restype = i
    return lib


def build_reference(quiet=True):
    """Compiles oracle/_ref from the reference sources when they are present."""
    if not pathlib.Path(REFERENCE_ROOT, "SoftRast", "stb_image_resize.h").exists():
        return False
    res = subprocess.run(["make", "-C", str(HERE), "ref", f"REF={REFERENCE_ROOT}"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("reference stbir build failed:\n" + res.stdout + res.stderr)
    return True


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def make_texture(tex):
    """SroTexture from anything with TextureData fields (m_texels, m_mipOffsets, ...)."""
    t = SroTexture()
    data = np.ascontiguousarray(tex.m_texels, np.uint8)
    t.texels = data.ctypes.data if data.size else None
    t.size = data.size
    for k in range(MAX_MIPS):
        t.mip_offsets[k] = int(tex.m_mipOffsets[k]) if k < len(tex.m_mipOffsets) else 0
    t.wlog2, t.hlog2, t.nmips = tex.m_widthLog2, tex.m_heightLog2, tex.m_numMips
    return t, data


class Plane:
    """Tiled plane with caller-owned numpy storage."""

    def __init__(self, width, height, primid=True):
        self.width, self.height = width, height
        self.tiles_x, self.tiles_y = (width + 63) // 64, (height + 63) // 64
        n = self.tiles_x * self.tiles_y * 4096
        self.colour = np.zeros(n, np.uint32)
        self.depth = np.zeros(n, np.float32)
        self.primid = np.full(n, 0xFFFFFFFF, np.uint32) if primid else None

    def struct(self):
        p = SroPlane()
        p.width, p.height, p.tiles_x, p.tiles_y = self.width, self.height, self.tiles_x, self.tiles_y
        p.colour, p.depth = self.colour.ctypes.data, self.depth.ctypes.data
        p.primid = self.primid.ctypes.data if self.primid is not None else None
        return p


class Oracle:
    """sro_ctx: nthreads = total threads including the caller."""

    def __init__(self, nthreads=1, fast=False):
        self.lib = load(fast)
        self.ctx = self.lib.sro_create(nthreads)

    def set_sponza_constants(self, consts):
        """consts: bytes-like sro_sponza_constants (6 + 16 * 8 float32)."""
        buf = np.ascontiguousarray(np.frombuffer(bytes(consts), np.float32))
        assert buf.size == 6 + 16 * 8
        self.lib.sro_set_sponza_constants(self.ctx, buf.ctypes.data)

    def clear(self, plane, colour=0, clear_colour=True, clear_depth=True):
        p = plane.struct()
        self.lib.sro_clear(ctypes.byref(p), colour & 0xFFFFFFFF, int(clear_colour), int(clear_depth))

    def render(self, draws, planes, tile_entries=False):
        """draws: list of dicts with keys shader, tex (SroTexture or None), idx (array), idx_stride,
        pos (array), pos_stride, pos_num, attr (array), attr_stride, attr_num, uv_offset, plane, mvp."""
        arr = (SroDraw * max(1, len(draws)))()
        keep = []
        for k, d in enumerate(draws):
            s = arr[k]
            s.shader = d["shader"]
            if d.get("tex") is not None:
                keep.append(d["tex"])
                s.tex = ctypes.pointer(d["tex"])
            idx = np.ascontiguousarray(d["idx"])
            pos = np.ascontiguousarray(d["pos"])
            att = np.ascontiguousarray(d["attr"])
            keep += [idx, pos, att]
            s.idx, s.idx_num, s.idx_stride = idx.ctypes.data, idx.size, d["idx_stride"]
            s.pos, s.pos_num, s.pos_stride = pos.ctypes.data, d["pos_num"], d["pos_stride"]
            s.attr, s.attr_num, s.attr_stride = att.ctypes.data, d["attr_num"], d["attr_stride"]
            s.uv_offset, s.plane = d.get("uv_offset", 6), d.get("plane", 0)
            s.mvp[:] = [float(x) for x in np.asarray(d["mvp"], np.float32).reshape(-1)]
        pls = (SroPlane * len(planes))(*[p.struct() for p in planes])
        stats = SroStats()
        te = None
        if tile_entries:
            te = np.zeros(sum(p.tiles_x * p.tiles_y for p in planes), np.uint32)
        st = self.lib.sro_render(self.ctx, arr, len(draws), pls, len(planes), ctypes.byref(stats),
                                 te.ctypes.data if te is not None else None)
        if st != 0:
            raise RuntimeError(f"oracle render failed: {st}")
        out = {n: getattr(stats, n) for n, _ in SroStats._fields_}
        if te is not None:
            out["tile_entries"] = te
        return out

    def blit(self, plane):
        out = np.empty((plane.height, plane.width), np.uint32)
        p = plane.struct()
        self.lib.sro_blit(ctypes.byref(p), out.ctypes.data)
        return out

    def close(self):
        if self.ctx:
            self.lib.sro_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
