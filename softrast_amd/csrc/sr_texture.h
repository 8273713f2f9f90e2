// sr_texture.h -- TextureData::CreateFromRGBA8 (Texture.cpp:119-199) on the
// device: the 32x32 Morton tiling of every level and the mip chain, each mip
// resized from the base level with stb_image_resize's Mitchell filter as
// stbir_resize_uint8 computes it (Texture.cpp:196).
//
// The resize is separable and rounds after each pass, so it runs as two
// kernels per level with the same float operations in the same order as the
// library: k_mip_rows weights each input row into an output-width row of
// float4 (the library's horizontal pass), k_mip_cols weights those rows into
// each output texel, encodes it to RGBA8 and writes it at its tiled address.
// The per-output tap lists (input index, weight) are built on the host from
// the library's contributor and normalisation rules; the device only
// accumulates, so weights are identical on both sides.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srg {

// x (even bits) or y (odd bits) of a 10-bit Morton index inside a 32x32 tile.
__device__ __forceinline__ uint32_t morton_compact5(uint32_t m)
{
    m &= 0x155u;
    m = (m | (m >> 1)) & 0x133u;
    m = (m | (m >> 2)) & 0x10Fu;
    m = (m | (m >> 4)) & 0x01Fu;
    return m;
}

// Tiled texel index t of a level with tiles_w tiles per row -> (x, y).
__device__ __forceinline__ void tiled_xy(uint32_t t, uint32_t tiles_w, uint32_t& x, uint32_t& y)
{
    const uint32_t tile = t >> 10, m = t & 1023u;
    x = ((tile % tiles_w) << 5) | morton_compact5(m);
    y = ((tile / tiles_w) << 5) | morton_compact5(m >> 1);
}

struct MipTaps {
    const uint32_t* first;  // taps of output i: [first[i], first[i + 1])
    const uint32_t* src;    // clamped input index
    const float* weight;
};

// TileTexture (Texture.cpp:73-101) of a linear RGBA8 level; texels of the
// 32-aligned padding are zero, as in the zero-initialised texel array.
__global__ void k_tex_tile(const uint32_t* __restrict__ src, uint32_t w, uint32_t h, uint32_t* __restrict__ dst,
                           uint32_t tiles_w, uint32_t ntexels)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntexels) return;
    uint32_t x, y;
    tiled_xy(t, tiles_w, x, y);
    dst[t] = (x < w && y < h) ? src[(size_t)y * w + x] : 0u;
}

__device__ __forceinline__ float decode_u8(uint32_t v) { return (float)v / 255.0f; }

// Horizontal pass for every input row: rows[y][x] = sum over the taps of x of
// decode(texel) * weight, accumulated from 0 in tap order
// (stbir__resample_horizontal_downsample, stb_image_resize.h:1526-1655).
__global__ void k_mip_rows(const uint32_t* __restrict__ src, uint32_t w, uint32_t h, uint32_t ow, MipTaps tx,
                           float4* __restrict__ rows)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * ow) return;
    const uint32_t y = i / ow, x = i - y * ow;
    const uint32_t* row = src + (size_t)y * w;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (uint32_t t = tx.first[x]; t < tx.first[x + 1]; ++t) {
        const uint32_t p = row[tx.src[t]];
        const float k = tx.weight[t];
        a0 = a0 + decode_u8(p & 0xFFu) * k;
        a1 = a1 + decode_u8((p >> 8) & 0xFFu) * k;
        a2 = a2 + decode_u8((p >> 16) & 0xFFu) * k;
        a3 = a3 + decode_u8(p >> 24) * k;
    }
    rows[i] = make_float4(a0, a1, a2, a3);
}

// STBIR__ENCODE_LINEAR8 (stb_image_resize.h:1743): saturate, * 255 in float,
// + 0.5 in double, truncate.
__device__ __forceinline__ uint32_t encode_u8(float f)
{
    f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
    return (uint32_t)(int)((double)(f * 255.0f) + 0.5);
}

// Vertical pass, encode and tiling of one level (stbir__resample_vertical_
// downsample, :1987-2066, and the encode of :1695-1760).
__global__ void k_mip_cols(const float4* __restrict__ rows, uint32_t ow, uint32_t oh, MipTaps ty,
                           uint32_t* __restrict__ dst, uint32_t tiles_w, uint32_t ntexels)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntexels) return;
    uint32_t x, y;
    tiled_xy(t, tiles_w, x, y);
    if (x >= ow || y >= oh) {
        dst[t] = 0u;
        return;
    }
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (uint32_t i = ty.first[y]; i < ty.first[y + 1]; ++i) {
        const float4 v = rows[(size_t)ty.src[i] * ow + x];
        const float k = ty.weight[i];
        a0 = a0 + v.x * k;
        a1 = a1 + v.y * k;
        a2 = a2 + v.z * k;
        a3 = a3 + v.w * k;
    }
    dst[t] = encode_u8(a0) | (encode_u8(a1) << 8) | (encode_u8(a2) << 16) | (encode_u8(a3) << 24);
}

}  // namespace srg
