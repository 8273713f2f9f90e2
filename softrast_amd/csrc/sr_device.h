// sr_device.h -- device-side arithmetic shared by the gfx950 kernels.
//
// Every function here reproduces one step of SoftRast's CPU arithmetic bit
// for bit. The library is compiled with -ffp-contract=off, correctly rounded
// fp32 divide/sqrt and denormals preserved, so a*b+c stays two roundings and
// __builtin_fmaf appears exactly where the reference uses an FMA intrinsic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srg {

constexpr uint32_t kBin = 64;             // Config.h:18-21 (the semantic tile)
constexpr uint32_t kTilePixels = kBin * kBin;
constexpr uint32_t kMaxSlots = 4;         // attribute planes kept per primitive
constexpr uint32_t kWideSlots = 8;        // ... by frames with a draw that reads eight varyings
constexpr uint32_t kNoTex = 0xFFFFFFFFu;

__device__ __constant__ uint32_t c_rcpps_table[2048] = {
#include "rcpps_table.inc"
};
__device__ __constant__ uint32_t c_rsqrtps_table[4096] = {
#include "rsqrtps_table.inc"
};

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float ffrom(uint32_t u) { return __uint_as_float(u); }

// C float -> int32 cast as x86 compiles it (cvttss2si): truncate, NaN and
// out-of-range -> INT32_MIN (Binning.cpp:301-303).
__device__ __forceinline__ int32_t cvtt_i32(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
}

// _mm256_cvtps_epi32 with round-to-nearest-even; NaN / out-of-range ->
// INT32_MIN (Texture.cpp:436-437, SIMDUtil.h:92-95).
__device__ __forceinline__ int32_t cvtn_i32(float f)
{
    float r = __builtin_rintf(f);
    return (r >= -2147483648.0f && r < 2147483648.0f) ? (int32_t)r : INT32_MIN;
}

// _mm256_max_ps(a, b) = a > b ? a : b (second operand on NaN).
__device__ __forceinline__ float maxps(float a, float b) { return a > b ? a : b; }

// _mm256_rcp_ps as recorded on the reference's Intel host (Rasterizer.cpp:375-376):
// table of the top 11 mantissa bits, exact exponent scaling, subnormal
// results flushed to zero, zero/denormal -> inf, inf -> 0, NaN quieted.
__device__ __forceinline__ float rcpps(float x)
{
    const uint32_t u = fbits(x);
    const uint32_t sign = u & 0x80000000u;
    const uint32_t e = (u >> 23) & 255u;
    if (e == 255u) return ffrom((u & 0x7FFFFFu) ? (u | 0x00400000u) : sign);
    if (e == 0u) return ffrom(sign | 0x7F800000u);
    const uint32_t t = c_rcpps_table[(u >> 12) & 2047u];
    const int ne = (int)((t >> 23) & 255u) - ((int)e - 127);
    return ffrom(ne <= 0 ? sign : (sign | ((uint32_t)ne << 23) | (t & 0x7FFFFFu)));
}

// _mm256_rsqrt_ps as recorded on the reference's Intel host (SponzaScene.cpp:67):
// for x = 2^E * f, table T[E & 1][top 11 mantissa bits] scaled by
// 2^-(E - (E & 1)) / 2 (always a normal result); +-0 and denormals -> +-inf,
// +inf -> +0, other negatives -> the default NaN, NaN quieted.
__device__ __forceinline__ float rsqrtps(float x)
{
    const uint32_t u = fbits(x);
    const uint32_t e = (u >> 23) & 255u;
    if (e == 255u && (u & 0x7FFFFFu)) return ffrom(u | 0x00400000u);
    if (e == 0u) return ffrom((u & 0x80000000u) | 0x7F800000u);
    if (u & 0x80000000u) return ffrom(0xFFC00000u);
    if (e == 255u) return 0.0f;
    const int ub = (int)e - 127, p = ub & 1;
    const uint32_t t = c_rsqrtps_table[(uint32_t)p * 2048u + ((u >> 12) & 2047u)];
    const int re = (int)((t >> 23) & 255u) - (ub - p) / 2;
    return ffrom(((uint32_t)re << 23) | (t & 0x7FFFFFu));
}

// rsqrtps / rcpps for inputs known to be positive normal numbers (rcpps: biased
// exponent <= 251), with no special cases: for those inputs they return the
// same bits as the functions above (checked over every such input). The
// table index of rsqrtps is the top 12 bits below the sign with the exponent's
// low bit inverted (= T[E & 1]), the scale 2^-(E - (E & 1)) / 2 an integer
// subtraction from the exponent field.
__device__ __forceinline__ float rsqrtps_pos(float x)
{
    const uint32_t u = fbits(x);
    const uint32_t t = c_rsqrtps_table[((u >> 12) & 0xFFFu) ^ 0x800u];
    return ffrom(t - (((u + 0x00800000u) & 0xFF000000u) >> 1) + 0x20000000u);
}
__device__ __forceinline__ float rcpps_pos(float x)
{
    const uint32_t u = fbits(x);
    const uint32_t t = c_rcpps_table[(u >> 12) & 2047u];
    return ffrom(t - (u & 0x7F800000u) + 0x3F800000u);
}

// One channel of RGBA32SoA_To_RGBA8AoS (SIMDUtil.h:87-121): RNE(fma(c,255,.5)),
// packus_epi32 then packus_epi16 (signed 16-bit input).
__device__ __forceinline__ uint32_t pack_channel(float c)
{
    const int32_t i = cvtn_i32(__builtin_fmaf(c, 255.0f, 0.5f));
    return i < 0 ? 0u : (i <= 255 ? (uint32_t)i : (i <= 32767 ? 255u : 0u));
}

__device__ __forceinline__ uint32_t pack_rgba(float r, float g, float b, float a)
{
    return pack_channel(r) | (pack_channel(g) << 8) | (pack_channel(b) << 16) | (pack_channel(a) << 24);
}

// 5-bit Morton interleave (Texture.cpp:36-67): the low 5 bits of a
// coordinate spread to the even bit positions (x; y takes the odd ones, one
// shift up), for two coordinates at once, one in each 16-bit half (a spread
// value stays below bit 9 of its half)
__device__ __forceinline__ uint32_t spread5x2(uint32_t a, uint32_t b)
{
    uint32_t v = (a & 31u) | ((b & 31u) << 16);
    v = (v | (v << 4)) & 0x010F010Fu;  // bit 4 -> bit 8
    v = (v | (v << 2)) & 0x01330133u;  // bits 2, 3 -> 4, 5
    v = (v | (v << 1)) & 0x01550155u;  // odd bits -> even positions
    return v;
}

// Device-memory view of a pointer read from a table: accesses through it
// compile to global memory instructions. Through a generic pointer they are
// flat instructions, which the wait counters track together with LDS
// operations, so a wait for an LDS result also waits for them.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global_ptr(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

struct TexDev {
    const uint8_t* texels;
    uint32_t mip_offsets[14];
    uint32_t wlog2, hlog2, nmips, empty;
};

// simdutil::Lerp (SIMDUtil.h:136-145): fma(t, b, fnmadd(t, a, a)).
__device__ __forceinline__ float lerp_fma(float a, float b, float t)
{
    return __builtin_fmaf(t, b, __builtin_fmaf(-t, a, a));
}

__device__ __forceinline__ float texel_channel(uint32_t byte) { return (float)byte * (1.0f / 255.0f); }

// CalcMipLevels (Texture.cpp:212-233), SampleWrap (:381-452) and
// GatherQuadsAndInterpolate (:243-379) for one pixel: RGBA in [0, 1].
__device__ __forceinline__ void sample_wrap(const TexDev& T, float u, float v, float dudx, float dudy, float dvdx,
                                            float dvdy, float o[4])
{
    const float height = (float)(1u << T.hlog2), width = (float)(1u << T.wlog2);
    const float p = dudx * width, q = dudy * height, r = dvdx * width, s = dvdy * height;
    const float du2 = __builtin_fmaf(p, p, q * q);
    const float dv2 = __builtin_fmaf(r, r, s * s);
    // ExtractExponent(sqrt(m)) = floor((E(m) - 127) / 2) for every normal m
    // (checked over all of them); zero, denormals, inf and NaN take the sqrt
    const float m2 = maxps(du2, dv2);
    const uint32_t me = (fbits(m2) >> 23) & 0xFFu;
    int32_t mip = ((int32_t)me - 127) >> 1;
    if (me == 0u || me == 255u) mip = (int32_t)((((int32_t)fbits(__builtin_sqrtf(m2))) >> 23) & 0xFF) - 127;
    mip = mip > 0 ? mip : 0;
    mip = mip < (int32_t)T.nmips - 1 ? mip : (int32_t)T.nmips - 1;

    const int32_t wl = (int32_t)T.wlog2, hl = (int32_t)T.hlog2;
    const int32_t mw = 1 << (wl - (wl < mip ? wl : mip));
    const int32_t mh = 1 << (hl - (hl < mip ? hl : mip));

    const uint32_t us = fbits(u) & 0x80000000u, vs = fbits(v) & 0x80000000u;
    const float au = ffrom(fbits(u) ^ us), av = ffrom(fbits(v) ^ vs);
    const float fu = au - __builtin_floorf(au), fv = av - __builtin_floorf(av);
    const float fuw = us ? 1.0f - fu : fu;
    const float fvw = vs ? 1.0f - fv : fv;
    const float ut = (float)mw * fuw, vt = (float)mh * fvw;
    const float uf = __builtin_floorf(ut), vf = __builtin_floorf(vt);
    const float ui = ut - uf, vi = vt - vf;
    const uint32_t x0 = (uint32_t)(cvtn_i32(uf) & (mw - 1)), y0 = (uint32_t)(cvtn_i32(vf) & (mh - 1));
    const uint32_t x1 = (x0 + 1u) & (uint32_t)(mw - 1), y1 = (y0 + 1u) & (uint32_t)(mh - 1);

    const uint32_t tw = (uint32_t)(mw > 32 ? mw : 32) >> 5;
    const auto base = global_ptr((const uint32_t*)(T.texels + T.mip_offsets[mip]));
    const uint32_t ty0 = (y0 >> 5) * tw, ty1 = (y1 >> 5) * tw;
    const uint32_t mx = spread5x2(x0, x1), my = spread5x2(y0, y1) << 1;
    const uint32_t mx0 = mx & 0xFFFFu, mx1 = mx >> 16, my0 = my & 0xFFFFu, my1 = my >> 16;
    const uint32_t t00 = base[1024u * (ty0 + (x0 >> 5)) + (mx0 | my0)];
    const uint32_t t01 = base[1024u * (ty1 + (x0 >> 5)) + (mx0 | my1)];
    const uint32_t t10 = base[1024u * (ty0 + (x1 >> 5)) + (mx1 | my0)];
    const uint32_t t11 = base[1024u * (ty1 + (x1 >> 5)) + (mx1 | my1)];

#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t sh = 8u * (uint32_t)c;
        const float left = lerp_fma(texel_channel((t00 >> sh) & 255u), texel_channel((t01 >> sh) & 255u), vi);
        const float right = lerp_fma(texel_channel((t10 >> sh) & 255u), texel_channel((t11 >> sh) & 255u), vi);
        o[c] = lerp_fma(left, right, ui);
    }
}

// UnlitDiffuseShader's sample (Viewer/Shaders.h:71-104), packed to RGBA8.
__device__ __forceinline__ uint32_t sample_unlit(const TexDev& T, float u, float v, float dudx, float dudy,
                                                 float dvdx, float dvdy)
{
    float o[4];
    sample_wrap(T, u, v, dudx, dudy, dvdx, dvdy, o);
    return pack_rgba(o[0], o[1], o[2], o[3]);
}

}  // namespace srg
