"""TextureData::CreateFromRGBA8 on the device (srg_texture_create_rgba8):
byte-identical to the host build (itself pinned to the reference's
stb_image_resize in test_mips.py), and interchangeable with it in a draw."""
import sys

import numpy as np
import pytest

from tests import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("w,h,seed", [(32, 32, 1), (64, 32, 2), (128, 512, 3), (2048, 2048, 4)])
def test_device_texture_matches_host_build(w, h, seed):
    import softrast_amd as sr
    _say(f"[texture] case=device_build size={w}x{h}")
    rng = np.random.default_rng(seed)
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    host = sr.TextureData().CreateFromRGBA8(img, w, h, True)
    ctx = sr.RenderContext()
    dev = ctx.CreateTexture(img, w, h, True).read()
    assert dev.m_numMips == host.m_numMips and dev.m_widthLog2 == host.m_widthLog2
    assert np.array_equal(dev.m_mipOffsets, host.m_mipOffsets)
    assert np.array_equal(dev.m_texels, host.m_texels)
    base_only = ctx.CreateTexture(img, w, h, False).read()
    assert np.array_equal(base_only.m_texels, sr.TextureData().CreateFromRGBA8(img, w, h, False).m_texels)
    ctx.Shutdown()


def test_device_texture_matches_reference_golden():
    import softrast_amd as sr
    from tests.test_mips import GOLDEN, CASES, _untile
    ctx = sr.RenderContext()
    for case in CASES:
        img = GOLDEN[f"{case}/in"]
        h, w, _ = img.shape
        _say(f"[texture] case=golden {case}")
        t = ctx.CreateTexture(img, w, h, True).read()
        for m in range(1, t.m_numMips):
            ow, oh = max(1, w >> m), max(1, h >> m)
            assert np.array_equal(_untile(t.m_texels, int(t.m_mipOffsets[m]), ow, oh), GOLDEN[f"{case}/mip{m}"]), (case, m)
    ctx.Shutdown()


def test_draw_with_device_texture_same_image():
    import softrast_amd as sr
    sc = helpers.random_scene(21, 3000, 320, 200, tex_size=(128, 64))
    tex = helpers.build_textures(sc)
    _say("[texture] case=draw_with_device_texture")
    ref = helpers.gpu_render(sc, tex)
    ctx = sr.RenderContext(flags=sr.CFG_KEEP_PRIMID | sr.CFG_STATS)
    dtex = [ctx.CreateTexture(img, img.shape[1], img.shape[0], True) for img in sc.textures]
    out = helpers.gpu_render(sc, dtex, ctx=ctx)
    for a, b in zip(ref[:3], out[:3]):
        assert np.array_equal(a, b)
    ctx.Shutdown()
