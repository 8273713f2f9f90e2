#!/usr/bin/env python3
"""Writes tests/golden/stbir_mips.npz: input images and every mip level the
reference's mip filter makes of them.

The reference builds mip m of a W x H texture as
stbir_resize_uint8(base, W, H, 0, out, max(1, W >> m), max(1, H >> m), 0, 4)
(SoftRast/Texture.cpp:196) with its vendored stb_image_resize v0.96. That
header is compiled as is into oracle/_ref/libstbir.so by `make -C oracle ref`
(it needs only libc and libm); this script calls it and stores the results.
Inputs are seeded, so the file is reproducible.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libstbir.so")


def images():
    rng = np.random.default_rng(0x57B1)
    out = {}
    out["random_32x32"] = rng.integers(0, 256, (32, 32, 4), dtype=np.uint8)
    out["random_64x32"] = rng.integers(0, 256, (32, 64, 4), dtype=np.uint8)
    yy, xx = np.indices((64, 32))
    out["checker_32x64"] = np.repeat((((xx + yy) % 2) * 255).astype(np.uint8)[..., None], 4, 2)
    out["sparse_64x64"] = ((rng.random((64, 64, 4)) < 0.05) * 255).astype(np.uint8)
    yy, xx = np.indices((32, 128))
    grad = np.stack([xx * 2, yy * 8, (xx + yy) % 256, 255 - xx], -1)
    out["gradient_128x32"] = grad.astype(np.uint8)
    return out


def main():
    if not os.path.exists(LIB):
        sys.exit(f"{LIB} not built: run make -C oracle ref")
    lib = ctypes.CDLL(LIB)
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.stbir_resize_uint8.argtypes = [vp, i, i, i, vp, i, i, i, i]
    lib.stbir_resize_uint8.restype = i
    arrays = {}
    for name, img in images().items():
        img = np.ascontiguousarray(img)
        h, w, _ = img.shape
        arrays[f"{name}/in"] = img
        levels = int(np.log2(max(w, h)))
        for m in range(1, levels + 1):
            ow, oh = max(1, w >> m), max(1, h >> m)
            res = np.zeros((oh, ow, 4), np.uint8)
            if lib.stbir_resize_uint8(img.ctypes.data, w, h, 0, res.ctypes.data, ow, oh, 0, 4) != 1:
                sys.exit(f"stbir_resize_uint8 failed for {name} mip {m}")
            arrays[f"{name}/mip{m}"] = res
    np.savez_compressed(os.path.join(HERE, "stbir_mips.npz"), **arrays)
    print(f"wrote {len(arrays)} arrays")


if __name__ == "__main__":
    main()
