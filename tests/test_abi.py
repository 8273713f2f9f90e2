"""C ABI checks that run without a GPU: the library loads, exports every
entry point include/srgpu.h declares, fails cleanly without a device, and its
host-side texture builder (TextureData::CreateFromRGBA8) lays texels out as
Texture.cpp:73-101,159-175 specify."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import softrast_amd as sr
from oracle import oracle as O

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include/srgpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(srg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(sr.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(sr.SRG_SYMBOLS) <= set(names)


def test_python_binding_declares_every_symbol():
    lib = sr.load_library()
    for n in declared_symbols():
        assert getattr(lib, n).argtypes is not None, n


def test_status_strings():
    lib = sr.load_library()
    for code in range(0, -10, -1):
        s = lib.srg_status_string(code).decode()
        assert s and s != "unknown status", code


def test_no_device_is_an_error_not_a_fallback():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = sr.load_library()
    cfg = sr.SrgConfig()
    cfg.struct_size = ctypes.sizeof(cfg)
    h = ctypes.c_void_p()
    assert lib.srg_create(ctypes.byref(cfg), ctypes.byref(h)) == -8  # SRG_ERR_NO_DEVICE
    with pytest.raises(sr.SrgError):
        sr.RenderContext()


@pytest.mark.parametrize("bin_size,wg_per_tile", [(16, 2), (8, 2), (8, 4), (16, 4), (48, 1), (32, 3), (128, 1)])
def test_invalid_layout_rejected_before_device(bin_size, wg_per_tile):
    """A work bin split into strips must keep whole 8-pixel block rows and at
    least one pixel per lane; other layouts are refused at creation."""
    lib = sr.load_library()
    cfg = sr.SrgConfig()
    cfg.struct_size = ctypes.sizeof(cfg)
    cfg.bin_size, cfg.wg_per_tile = bin_size, wg_per_tile
    h = ctypes.c_void_p()
    assert lib.srg_create(ctypes.byref(cfg), ctypes.byref(h)) == -1  # SRG_ERR_INVALID
    assert not h.value


def test_valid_layouts_pass_validation():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    lib = sr.load_library()
    for bin_size, wg_per_tile in sr.LAYOUTS:
        cfg = sr.SrgConfig()
        cfg.struct_size = ctypes.sizeof(cfg)
        cfg.bin_size, cfg.wg_per_tile = bin_size, wg_per_tile
        h = ctypes.c_void_p()
        assert lib.srg_create(ctypes.byref(cfg), ctypes.byref(h)) == -8  # past validation: SRG_ERR_NO_DEVICE


def test_null_arguments_rejected():
    lib = sr.load_library()
    assert lib.srg_begin_frame(None) == -1
    assert lib.srg_end_frame(None) == -1
    assert lib.srg_draw_indexed(None, None) == -1


@pytest.mark.parametrize("w,h", [(31, 64), (48, 64), (64, 16), (0, 64), (16384, 32)])
def test_texture_builder_rejects_bad_sizes(w, h):
    with pytest.raises((sr.SrgError, ValueError)):
        sr.TextureData().CreateFromRGBA8(np.zeros(max(1, w * h * 4), np.uint8), w, h, True)


def test_texture_builder_layout():
    rng = np.random.default_rng(0)
    w, h = 128, 64
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    t = sr.TextureData().CreateFromRGBA8(img, w, h, True)
    assert (t.m_widthLog2, t.m_heightLog2, t.m_numMips) == (7, 6, 8)
    # mip offsets: each mip padded to a multiple of 32 texels in both dimensions
    off = 0
    for m in range(t.m_numMips):
        mw, mh = max(1, w >> m), max(1, h >> m)
        assert t.m_mipOffsets[m] == off
        off += ((mw + 31) // 32 * 32) * ((mh + 31) // 32 * 32) * 4
    assert t.m_texels.size == off
    # base level: texel (x, y) at 4 * (1024 * (tile row * tiles per row + tile col) + morton(x & 31, y & 31))
    lib = O.load()
    tex = t.m_texels.view(np.uint32)
    lin = img.view(np.uint32).reshape(h, w)
    for y in range(0, h, 7):
        for x in range(0, w, 5):
            o = 1024 * ((y >> 5) * (w // 32) + (x >> 5)) + lib.sro_morton5(x & 31, y & 31)
            assert tex[o] == lin[y, x]
    # without mips: a single level
    t1 = sr.TextureData().CreateFromRGBA8(img, w, h, False)
    assert t1.m_numMips == 1 and t1.m_texels.size == w * h * 4
