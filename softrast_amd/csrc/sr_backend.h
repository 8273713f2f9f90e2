// sr_backend.h -- back end: per-bin coverage, early depth and resolve,
// then interpolation and shading of the visible fragment of each pixel.
//
// k_backend<S, H> is persistent: each workgroup takes an S x H strip of an
// S x S bin from the frame's queue (H = S: the whole bin; H = S/2, S/4: two
// or four workgroups per bin, sharing its entry list). A 64-bit visibility
// key per pixel lives in LDS. Waves
// stride over the bin's entries 64 at a time, each lane loading one
// primitive record so the loads overlap; the entries' column pairs of 8x8
// block rows then form a queue the wave takes 64 at a time, a lane per pair
// walking its eight rows (RasterizeTrisInBin_OutputFragments and
// ComputeBlockMask8x8, Rasterizer.cpp:97-304). Then each lane resolves its pixels (the key names
// the fragment the reference would have shaded last) and shades them
// (ComputeInterpolantsDrawCallImpl :306-419, Shaders.h, Texture.cpp),
// writing RGBA8 and depth once.
#pragma once
#include <type_traits>

#include "sr_frontend.h"

namespace srg {

// Workgroup size for an S x H strip: 4 waves, or one wave for an 8 x 8 bin.
__host__ __device__ constexpr uint32_t backend_threads(uint32_t S, uint32_t H) { return S * H >= 256u ? 256u : 64u; }

// The reference keeps a fragment iff z > 0 and z > D (ordered compare) and
// shades every kept fragment in submission order, so the final pixel is the
// earliest-submitted fragment of largest depth. Key = depth bits << 32 |
// (0xFFFFFFFE - order); an atomic max over the keys selects it. Initial key
// from the depth already in the plane: D <= 0 -> 0, D NaN -> all ones
// (nothing passes), D > 0 -> D's bits above every order field.
__device__ __forceinline__ unsigned long long init_key(float d0)
{
    if (d0 != d0) return ~0ull;
    if (!(d0 > 0.0f)) return 0ull;
    return ((unsigned long long)fbits(d0) << 32) | 0xFFFFFFFFull;
}

__device__ __forceinline__ uint32_t record_of_order(const FrameDev& F, uint32_t order)
{
    const uint32_t g = order >> 3, fan = order & 7u;
    return fan == kFanUnclipped ? g : F.ntris + F.clip_base[g] + fan;
}

// SponzaShader (Viewer/SponzaScene.cpp:13-104) for the visible fragment of
// one pixel: all eight varyings (0-2 position, 3-5 normal, 6-7 uv), the uv
// derivatives from the slots at uv_offset, sun, 16 point lights by
// RSQRTPS / RCPPS, ambient, times the texture sample.
__device__ __forceinline__ uint32_t shade_sponza(const FrameDev& F, const DrawDev& D, uint32_t rec, int32_t X0,
                                                 int32_t Y0, uint32_t px, uint32_t py)
{
    if (D.tex == kNoTex) return 0xFFFFFFFFu;
    const TexDev& T = F.texs[D.tex];
    if (T.empty) return 0xFFFFFFFFu;
    const float* Sp = F.rec_shade + (size_t)rec * F.shade_stride;
    const float4* Ep = F.rec_ext + (size_t)rec * 3;
    const float4 s0 = make_float4(Sp[0], Sp[1], Sp[2], Sp[3]), s1 = make_float4(Sp[4], Sp[5], Sp[6], Sp[7]);
    const float4 s2 = make_float4(Sp[8], Sp[9], Sp[10], Sp[11]), s3 = make_float4(Sp[12], Sp[13], Sp[14], 0.0f);
    const float4 e0 = Ep[0], e1 = Ep[1], e2 = Ep[2];
    const uint4 q2 = F.rec_raster[(size_t)rec * 4 + 2];
    const uint4 q3 = F.rec_raster[(size_t)rec * 4 + 3];
    const float sx = (float)X0 - ffrom(q2.w), sy = (float)Y0 - ffrom(q3.x);
    const float wdx = s0.x, wdy = s0.y;
    const float wc = wdx * sx + wdy * sy + s0.z;
    const float pdx[kWideSlots] = { s0.w, s1.z, s2.y, s3.x, e0.x, e0.w, e1.z, e2.y };
    const float pdy[kWideSlots] = { s1.x, s1.w, s2.z, s3.y, e0.y, e1.x, e1.w, e2.z };
    const float pf0[kWideSlots] = { s1.y, s2.x, s2.w, s3.z, e0.z, e1.y, e2.x, e2.w };
    const float X = (float)px, Y = (float)py;
    const float iq = 1.0f / __builtin_fmaf(X, wdx, __builtin_fmaf(Y, wdy, wc));
    float pc[kWideSlots], var[kWideSlots];
#pragma unroll
    for (uint32_t k = 0; k < kWideSlots; ++k) {
        pc[k] = pdx[k] * sx + pdy[k] * sy + pf0[k];
        var[k] = iq * __builtin_fmaf(pdy[k], Y, __builtin_fmaf(pdx[k], X, pc[k]));
    }
    // (u, v) planes at uv_offset: selected without indexing the arrays by a
    // run-time value (which would move them to scratch memory)
    float tdx[2] = { 0.0f, 0.0f }, tdy[2] = { 0.0f, 0.0f }, tc[2] = { 0.0f, 0.0f }, tv[2] = { 0.0f, 0.0f };
    const uint32_t ds = D.deriv_slot;
#pragma unroll
    for (uint32_t k = 0; k + 1 < kWideSlots; ++k)
        if (k == ds) {
            tdx[0] = pdx[k]; tdy[0] = pdy[k]; tc[0] = pc[k]; tv[0] = var[k];
            tdx[1] = pdx[k + 1]; tdy[1] = pdy[k + 1]; tc[1] = pc[k + 1]; tv[1] = var[k + 1];
        }
    const float X1 = 1.0f + X, Y1 = 1.0f + Y;
    const float i10 = rcpps(__builtin_fmaf(wdx, X1, __builtin_fmaf(wdy, Y, wc)));
    const float i01 = rcpps(__builtin_fmaf(wdx, X, __builtin_fmaf(wdy, Y1, wc)));
    float der[4];
#pragma unroll
    for (uint32_t uv = 0; uv < 2; ++uv) {
        der[2 * uv] = i10 * __builtin_fmaf(tdx[uv], X1, __builtin_fmaf(tdy[uv], Y, tc[uv])) - tv[uv];
        der[2 * uv + 1] = i01 * __builtin_fmaf(tdx[uv], X, __builtin_fmaf(tdy[uv], Y1, tc[uv])) - tv[uv];
    }

    const SponzaDev& K = *F.sponza;
    const float nx = var[3], ny = var[4], nz = var[5];
    // Dot3SoA (SIMDUtil.h:123-126): fma(x0, x1, fma(y0, y1, z0 * z1))
    const float ndl = maxps(0.1f, __builtin_fmaf(nx, K.sun_dir[0], __builtin_fmaf(ny, K.sun_dir[1], nz * K.sun_dir[2])));
    float r0 = ndl, r1 = ndl, r2 = ndl;
    // Fast pass: while d^2 is a positive normal number, RSQRTPS's result and
    // RCPPS's two inputs (1/d <= 2^63, 1 + 0.1 d + 0.01 d^2 < 2^122) are
    // positive normal too, so the special-case-free forms give the same bits.
    // A lane with any other d^2 (0, denormal, inf, NaN) is redone exactly.
    bool special = false;
#pragma unroll 4
    for (uint32_t i = 0; i < kSponzaLights; ++i) {
        const SponzaLightDev& L = K.lights[i];
        const float tx = L.pos[0] - var[0], ty = L.pos[1] - var[1], tz = L.pos[2] - var[2];
        const float dsq = __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz));
        special = special || (fbits(dsq) - 0x00800000u) >= 0x7F000000u;
        const float rd = rsqrtps_pos(dsq);
        const float dist = rcpps_pos(rd);
        const float lx = tx * rd, ly = ty * rd, lz = tz * rd;
        const float nl = maxps(0.0f, __builtin_fmaf(lx, nx, __builtin_fmaf(ly, ny, lz * nz)));
        const float atten = rcpps_pos(1.0f + __builtin_fmaf(0.1f, dist, dsq * 0.01f));
        const float lr = nl * (L.intensity * atten);
        r0 = r0 + lr * L.colour[0];
        r1 = r1 + lr * L.colour[1];
        r2 = r2 + lr * L.colour[2];
    }
    if (special) {
        r0 = r1 = r2 = ndl;
        for (uint32_t i = 0; i < kSponzaLights; ++i) {
            const SponzaLightDev& L = K.lights[i];
            const float tx = L.pos[0] - var[0], ty = L.pos[1] - var[1], tz = L.pos[2] - var[2];
            const float dsq = __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz));
            const float rd = rsqrtps(dsq);
            const float dist = rcpps(rd);
            const float lx = tx * rd, ly = ty * rd, lz = tz * rd;
            const float nl = maxps(0.0f, __builtin_fmaf(lx, nx, __builtin_fmaf(ly, ny, lz * nz)));
            const float atten = rcpps(1.0f + __builtin_fmaf(0.1f, dist, dsq * 0.01f));
            const float lr = nl * (L.intensity * atten);
            r0 = r0 + lr * L.colour[0];
            r1 = r1 + lr * L.colour[1];
            r2 = r2 + lr * L.colour[2];
        }
    }
    r0 = r0 + K.ambient[0];
    r1 = r1 + K.ambient[1];
    r2 = r2 + K.ambient[2];
    float o[4];
    sample_wrap(T, var[6], var[7], der[0], der[1], der[2], der[3], o);
    return pack_rgba(r0 * o[0], r1 * o[1], r2 * o[2], o[3]);
}

// ComputeInterpolants + pixel shader for the visible fragment of one pixel
// (x, y tile-relative) (Rasterizer.cpp:356-400, Viewer/Shaders.h:71-130).
// WIDE: the frame keeps eight attribute planes per record and may hold
// SponzaShader draws.
template <bool WIDE>
__device__ __forceinline__ uint32_t shade_pixel(const FrameDev& F, uint32_t rec, int32_t X0, int32_t Y0, uint32_t px,
                                                uint32_t py)
{
    const uint32_t d = F.rec_meta[rec];
    const DrawDev& D = F.draws[d];
    if (WIDE && D.shader == SRG_SHADER_SPONZA) return shade_sponza(F, D, rec, X0, Y0, px, py);
    // shade record: (wdx, wdy, w0), then (dx, dy, value) of each kept slot
    const float* Sp = F.rec_shade + (size_t)rec * F.shade_stride;
    const float4 s0 = make_float4(Sp[0], Sp[1], Sp[2], Sp[3]), s1 = make_float4(Sp[4], Sp[5], Sp[6], Sp[7]);
    const float pf01 = Sp[8];
    const uint32_t* Rw = (const uint32_t*)(F.rec_raster + (size_t)rec * 4);
    // plane constants re-based to the tile origin (Binning.cpp:435-452)
    const float sx = (float)X0 - ffrom(Rw[11]), sy = (float)Y0 - ffrom(Rw[12]);
    const float wdx = s0.x, wdy = s0.y;
    const float wc = wdx * sx + wdy * sy + s0.z;
    const float X = (float)px, Y = (float)py;
    const float iq = 1.0f / __builtin_fmaf(X, wdx, __builtin_fmaf(Y, wdy, wc));
    // slots 0 and 1: the sampled (u, v), the normal's x, y, or the shown uv
    float tdx[2] = { s0.w, s1.z }, tdy[2] = { s1.x, s1.w }, tc[2], tv[2];
    tc[0] = tdx[0] * sx + tdy[0] * sy + s1.y;
    tc[1] = tdx[1] * sx + tdy[1] * sy + pf01;
    tv[0] = iq * __builtin_fmaf(tdy[0], Y, __builtin_fmaf(tdx[0], X, tc[0]));
    tv[1] = iq * __builtin_fmaf(tdy[1], Y, __builtin_fmaf(tdx[1], X, tc[1]));
    const float u = tv[0], v = tv[1];
    switch (D.shader) {
    case SRG_SHADER_VISUALIZE_NORMALS: {
        const float4 s2 = make_float4(Sp[8], Sp[9], Sp[10], Sp[11]);
        const float c2 = s2.y * sx + s2.z * sy + s2.w;
        const float n2 = iq * __builtin_fmaf(s2.z, Y, __builtin_fmaf(s2.y, X, c2));
        return pack_rgba(__builtin_fmaf(u, 0.5f, 0.5f), __builtin_fmaf(v, 0.5f, 0.5f), __builtin_fmaf(n2, 0.5f, 0.5f),
                         1.0f);
    }
    case SRG_SHADER_VISUALIZE_UVS:
        return pack_rgba(u, v, 0.0f, 0.0f);
    default: break;
    }
    if (D.tex == kNoTex) return 0xFFFFFFFFu;
    const TexDev& T = F.texs[D.tex];
    if (T.empty) return 0xFFFFFFFFu;
    if (D.deriv_slot != 0) {  // uv_offset != 6: derivatives from slots 2 and 3
        const float4 s2 = make_float4(Sp[8], Sp[9], Sp[10], Sp[11]), s3 = make_float4(Sp[12], Sp[13], Sp[14], 0.0f);
        tdx[0] = s2.y; tdy[0] = s2.z; tdx[1] = s3.x; tdy[1] = s3.y;
        tc[0] = tdx[0] * sx + tdy[0] * sy + s2.w;
        tc[1] = tdx[1] * sx + tdy[1] * sy + s3.z;
        tv[0] = iq * __builtin_fmaf(tdy[0], Y, __builtin_fmaf(tdx[0], X, tc[0]));
        tv[1] = iq * __builtin_fmaf(tdy[1], Y, __builtin_fmaf(tdx[1], X, tc[1]));
    }
    // neighbour 1/w by RCPPS, forward differences (Rasterizer.cpp:373-400);
    // the special-case-free RCPPS while both arguments are positive normal
    // numbers below 2^125, the exact one otherwise
    const float X1 = 1.0f + X, Y1 = 1.0f + Y;
    const float a10 = __builtin_fmaf(wdx, X1, __builtin_fmaf(wdy, Y, wc));
    const float a01 = __builtin_fmaf(wdx, X, __builtin_fmaf(wdy, Y1, wc));
    float i10 = rcpps_pos(a10), i01 = rcpps_pos(a01);
    if ((fbits(a10) - 0x00800000u) >= 0x7D800000u || (fbits(a01) - 0x00800000u) >= 0x7D800000u) {
        i10 = rcpps(a10);
        i01 = rcpps(a01);
    }
    float der[4];
#pragma unroll
    for (uint32_t k = 0; k < 2; ++k) {
        const float t10 = i10 * __builtin_fmaf(tdx[k], X1, __builtin_fmaf(tdy[k], Y, tc[k]));
        const float t01 = i01 * __builtin_fmaf(tdx[k], X, __builtin_fmaf(tdy[k], Y1, tc[k]));
        der[2 * k] = t10 - tv[k];
        der[2 * k + 1] = t01 - tv[k];
    }
    return sample_unlit(T, u, v, der[0], der[1], der[2], der[3]);
}

// The view a bin belongs to: the last view whose first bin is <= gb. gb must
// be the same in every lane of the wave: with at most 64 views each lane
// reads one view's first bin and a ballot counts them (one load instead of
// a chain of dependent loads); more views are binary-searched.
__device__ __forceinline__ uint32_t view_of_bin(const FrameDev& F, uint32_t gb)
{
    if (F.nviews <= 64u) {
        const uint32_t lane = lane_id();
        const bool le = lane < F.nviews && F.views[lane].bin_base <= gb;
        return (uint32_t)__popcll(__ballot(le)) - 1u;
    }
    uint32_t lo = 0, hi = F.nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (F.views[mid].bin_base <= gb) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t bcast(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// 64-wide block reject / full-accept test of RasterizeTrisInBin_OutputFragments
// (Rasterizer.cpp:230-272): corners {xb, xb+64} x {yb, yb+64} of the block,
// tile-relative, each edge value in wrapping int32 as the reference computes
// it. The block is rejected when some edge has no corner > 0 (the largest of
// its four corner values is <= 0) and fully covered when every corner of
// every edge is > 0 (the smallest of all twelve is > 0).
// Returns 0 = reject, 1 = test pixels, 2 = every pixel covered.
__device__ __forceinline__ uint32_t block_class(const uint32_t ct[3], const uint32_t dxu[3], const uint32_t dyu[3],
                                                uint32_t xb, uint32_t yb)
{
    int32_t rej = INT32_MAX, full = INT32_MAX;  // min over edges of the largest corner; min over all corners
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t a = ct[k] + dyu[k] * xb + dxu[k] * yb;  // (xb, yb)
        const uint32_t b = a + dxu[k] * kBin;                  // (xb, yb + 64)
        const uint32_t c = a + dyu[k] * kBin;                  // (xb + 64, yb)
        const uint32_t d = c + dxu[k] * kBin;                  // (xb + 64, yb + 64)
        rej = min(rej, max(max((int32_t)a, (int32_t)b), max((int32_t)c, (int32_t)d)));
        full = min(full, min(min((int32_t)a, (int32_t)b), min((int32_t)c, (int32_t)d)));
    }
    return rej <= 0 ? 0u : (full > 0 ? 2u : 1u);
}

// LDS slot of the key of strip-local pixel (x, y): rows of S keys, and inside
// each group of G = min(S, 16) keys of a row the slot is rotated by
// 2*(y>>3) + (x>>4) + 8*(y&1) (mod G). Keys are 8 bytes, so 16 consecutive
// keys cover the LDS banks once.
// Without the rotation the 64 lanes of a wave hit the same few banks: one
// block per wave puts its 8 rows on the same banks, and one block per lane
// puts the same pixel of every lane's block on the same bank. With it, a
// block's 64 pixels fall on 16 bank pairs 4 apiece and the same pixel of the
// 16 blocks of a 32x32 bin on 16 different bank pairs.
template <uint32_t S>
__device__ __forceinline__ uint32_t key_slot(uint32_t x, uint32_t y)
{
    constexpr uint32_t G = S < 16u ? S : 16u;
    const uint32_t rot = 2u * (y >> 3) + (x >> 4) + 8u * (y & 1u);
    return y * S + (x & ~(G - 1u)) + ((x + rot) & (G - 1u));
}

// One primitive's tile-relative setup (Binning.cpp:423-451 re-basing) and its
// block range inside the S x H strip at (sx0, sy0). Strips start on 8-pixel
// rows, so the blocks walked over a bin's strips are the blocks walked over
// the bin.
struct EntryRaster {
    uint32_t ct[3], dxu[3], dyu[3];
    uint32_t xs, xe, ys, ye;      // the reference's walk: blocks from (xs, ys) while xb < xe, yb < ye
    uint32_t cx0, cx1, cy0, cy1;  // pixels that can be covered (tile-relative, ends exclusive): the walked
                                  // pixels, cut to the tight bound where the primitive has one
    uint32_t bx0, by0, nbx, nby;  // the 8x8 blocks holding those pixels
    bool tight;
    float zdx, zdy, zc0;
    unsigned long long low;
};

__device__ __forceinline__ EntryRaster entry_raster(uint4 q0, uint4 q1, uint4 q2, uint4 q3, uint32_t order, int32_t X0,
                                                    int32_t Y0, uint32_t sx0, uint32_t sy0, uint32_t S, uint32_t H)
{
    EntryRaster e;
    e.dxu[0] = q0.w; e.dxu[1] = q1.x; e.dxu[2] = q1.y;
    e.dyu[0] = q1.z; e.dyu[1] = q1.w; e.dyu[2] = q2.x;
    const uint32_t c[3] = { q0.x, q0.y, q0.z };
#pragma unroll
    for (int k = 0; k < 3; ++k) e.ct[k] = c[k] + (e.dxu[k] * (uint32_t)Y0 + e.dyu[k] * (uint32_t)X0);
    const int32_t xmin = q2.y & 0x3FFF, xmax = (q2.y >> 16) & 0x3FFF, ymin = q2.z & 0x3FFF, ymax = (q2.z >> 16) & 0x3FFF;
    const uint32_t bminx = (uint32_t)clampi(xmin - X0, 0, 64), bmaxx = (uint32_t)clampi(xmax - X0, 0, 64);
    const uint32_t bminy = (uint32_t)clampi(ymin - Y0, 0, 64), bmaxy = (uint32_t)clampi(ymax - Y0, 0, 64);
    e.xs = max(bminx & ~7u, sx0);
    e.xe = min(bmaxx, sx0 + S);
    e.ys = max(bminy & ~7u, sy0);
    e.ye = min(bmaxy, sy0 + H);
    // walked pixels: whole blocks from xs while xb < xe (xs, ys are multiples of 8)
    e.tight = ((q2.y >> 14) & 1u) != 0;
    const int32_t wx1 = e.xe > e.xs ? (int32_t)(e.xs + ((e.xe - e.xs + 7u) & ~7u)) : (int32_t)e.xs;
    const int32_t wy1 = e.ye > e.ys ? (int32_t)(e.ys + ((e.ye - e.ys + 7u) & ~7u)) : (int32_t)e.ys;
    if (e.tight) {
        e.cx0 = (uint32_t)max((int32_t)e.xs, xmin - (int32_t)((q2.y >> 15) & 1u) - X0);
        e.cy0 = (uint32_t)max((int32_t)e.ys, ymin - (int32_t)((q2.z >> 14) & 1u) - Y0);
        const int32_t bx1 = ((q2.y >> 30) & 1u) ? wx1 : xmax - X0 + 1, by1 = ((q2.z >> 30) & 1u) ? wy1 : ymax - Y0 + 1;
        e.cx1 = (uint32_t)max((int32_t)e.cx0, min(wx1, bx1));
        e.cy1 = (uint32_t)max((int32_t)e.cy0, min(wy1, by1));
    } else {
        e.cx0 = e.xs;
        e.cy0 = e.ys;
        e.cx1 = (uint32_t)wx1;
        e.cy1 = (uint32_t)wy1;
    }
    e.bx0 = e.cx0 & ~7u;
    e.by0 = e.cy0 & ~7u;
    e.nbx = e.cx1 > e.cx0 ? ((e.cx1 - 1u) >> 3) - (e.cx0 >> 3) + 1u : 0u;
    e.nby = e.cy1 > e.cy0 ? ((e.cy1 - 1u) >> 3) - (e.cy0 >> 3) + 1u : 0u;
    const float r0x = ffrom(q2.w), r0y = ffrom(q3.x);
    e.zdx = ffrom(q3.y);
    e.zdy = ffrom(q3.z);
    const float sxf = (float)X0 - r0x, syf = (float)Y0 - r0y;
    e.zc0 = e.zdx * sxf + e.zdy * syf + ffrom(q3.w);
    e.low = 0xFFFFFFFEull - order;
    return e;
}

// Rasterisation of a wave's 64 entries (one per lane, `e`). Every entry's
// candidate pixels (cx0..cx1 x cy0..cy1) split into column pairs of 8x8 block
// rows: the even-aligned pairs (2m, 2m + 1) that overlap cx0..cx1, in each of
// its block rows (a pair never straddles a block). These (entry, block row,
// pair) items form a queue the wave takes 64 at a time, one lane per item;
// the lane walks both columns of its pair down the block row, carrying the
// reference's per-column depth (ComputeBlockMask8x8, Rasterizer.cpp:134-192):
// z = fma(yb, dy, fma(l, dx, c0)) + xb*dx at the block's top row, then + dy
// per row, and the edge values + dx_k per row.
//
// A block the reference rejects by its 64-wide corner test is skipped and a
// fully accepted one is covered without per-pixel tests (Rasterizer.cpp:
// 230-272); otherwise a pixel is covered when its three edge values are
// >= 0. Per item this is one threshold on min(E0, E1, E2): INT_MIN (full),
// 0 (test pixels). For a tight primitive (tight_bound in sr_frontend.h) only
// the rows and columns of its bound are visited and the corner test matters
// only where some edge value is exactly 0 (with exact edge values a rejected
// block has no pixel whose three values are > 0, a fully accepted block no
// pixel with a negative one): its threshold is 1, and pixels whose minimum
// is exactly 0 are collected and decided by the corner test after the rows,
// with the depth the row walk gives them.
// Words of a wave's LDS staging area: 15 fields of each of its 64 entries,
// field-major (field f of lane j at f * 64 + j), then the 64 owner marks.
constexpr uint32_t kStageFields = 15;
constexpr uint32_t kStageWords = (kStageFields + 1u) * 64u;

template <uint32_t S, bool STATS>
__device__ __forceinline__ void raster_wave(unsigned long long* keys, const EntryRaster& e, bool has, uint32_t sx0,
                                            uint32_t sy0, unsigned long long& frags, uint32_t* stage)
{
    const uint32_t lane = lane_id();
    const uint32_t cw = e.cx1 > e.cx0 ? e.cx1 - e.cx0 : 0u;
    const uint32_t np = cw ? ((e.cx1 - 1u) >> 1) - (e.cx0 >> 1) + 1u : 0u;  // column pairs per block row
    const uint32_t n = has ? e.nby * np : 0u;
    const uint32_t incl = wave_incl_sum(n);
    const uint32_t total = bcast(incl, 63);
    // per-entry values the column workers need, packed; the first block row
    // is cy0 rounded down to 8
    // k / np == (k * magic) >> 16 for k < 512 when magic is ceil(2^16 / np)
    // or one more (np <= 64); the 1-ulp reciprocal gives one of the two
    const uint32_t npp = max(np, 1u);
    const uint32_t magic = (uint32_t)(65536.0f * __builtin_amdgcn_rcpf((float)npp)) + 1u;
    const uint32_t pk_cand = e.cx0 | (cw << 7) | (e.cy0 << 14) | (e.cy1 << 21) | ((e.tight ? 1u : 0u) << 28);
    const uint32_t pk_div = magic | (npp << 17);
    const uint32_t first = incl - n;
    // each lane's entry into the wave's staging area: items read their
    // owner's fields from there (a broadcast where lanes share an owner)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the previous entries' reads are issued
    {
        const uint32_t v[kStageFields] = { e.ct[0], e.ct[1], e.ct[2], e.dxu[0], e.dxu[1], e.dxu[2], e.dyu[0], e.dyu[1],
                                           e.dyu[2], pk_cand, pk_div, fbits(e.zdx), fbits(e.zdy), fbits(e.zc0),
                                           (uint32_t)e.low };
#pragma unroll
        for (uint32_t f = 0; f < kStageFields; ++f) stage[f * 64u + lane] = v[f];
    }
    uint32_t* mark = stage + kStageFields * 64u;
    uint32_t carry = 0;  // mark of the previous pass's last item
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t it = base + lane;
        const bool valid = it < total;
        const uint32_t m = wave_owner_first(mark, n, first, base, carry);
        const uint32_t mo = m & 127u;
        const int src = (int)(valid && mo ? min(mo - 1u, 63u) : 0u);
        const uint32_t* o = stage + src;  // the owner's fields (wave_owner_first fenced the stores)
        uint32_t ct[3], dxu[3], dyu[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ct[k] = o[k * 64];
            dxu[k] = o[(3 + k) * 64];
            dyu[k] = o[(6 + k) * 64];
        }
        const uint32_t pc = o[9 * 64], pd = o[10 * 64];
        const float zdx = ffrom(o[11 * 64]), zdy = ffrom(o[12 * 64]), zc0 = ffrom(o[13 * 64]);
        const unsigned long long lw = (unsigned long long)pin(o[14 * 64]);
        const uint32_t k = it - (m >> 7);
        const uint32_t cx0 = pc & 127u, cy0 = (pc >> 14) & 127u;
        const bool tight = ((pc >> 28) & 1u) != 0;
        const uint32_t snp = pd >> 17;
        const uint32_t brow = __umul24(k, pd & 0x1FFFFu) >> 16;  // k < 512, magic < 2^17
        const uint32_t x = 2u * ((cx0 >> 1) + (k - brow * snp)), yb = (cy0 & ~7u) + 8u * brow;
        const uint32_t xb = x & ~7u, l = x & 7u;
        // A valid item lies in its owner's candidate rectangle by its
        // numbering: the np pairs counted are those overlapping cx0..cx1 and
        // the nby block rows those overlapping cy0..cy1.
        const bool inside = valid;
        int32_t thr = 1;  // tight: all three edge values > 0; an exact 0 is decided below
        bool walk_rows = inside;
        if (inside && !tight) {
            const uint32_t cls = block_class(ct, dxu, dyu, xb, yb);
            thr = cls == 2u ? INT32_MIN : 0;
            walk_rows = cls != 0u;  // a rejected block draws nothing
        }
        // key_slot of (x - sx0, yb + r - sy0): rows of S keys, the in-group
        // rotation alternating with the row's parity (yb is a multiple of 8),
        // so each column has one slot base for even rows and one for odd
        // rows, and row r adds r * S keys (a constant offset). x is even, so
        // x + 1 is in the same group of G keys.
        constexpr uint32_t G = S < 16u ? S : 16u;
        const uint32_t sxl = x - sx0, ybl = yb - sy0;
        const uint32_t t = sxl + 2u * (ybl >> 3) + (sxl >> 4);
        const uint32_t grp = ybl * S + (sxl & ~(G - 1u));
        // byte offsets into keys of the four slot bases; grp is a multiple
        // of G, so the odd rows' rotation by 8 flips bit 3 of the slot (G =
        // 16) or changes nothing (G = 8)
        const uint32_t b0e = pin(8u * (grp + (t & (G - 1u)))), b1e = pin(8u * (grp + ((t + 1u) & (G - 1u))));
        const uint32_t b0o = pin(G == 16u ? b0e ^ 64u : b0e), b1o = pin(G == 16u ? b1e ^ 64u : b1e);
        char* const kb = reinterpret_cast<char*>(keys);
        // One walk down all eight rows of the block row. Rows and columns
        // outside cy0..cy1 x cx0..cx1 are walked too: for a primitive without
        // a tight bound the rectangle is whole blocks, so there are none, and
        // for a tight one every pixel of the walked blocks outside its bound
        // has a negative edge value (the bound holds every pixel whose three
        // values are >= 0), so they cover nothing and have no exact zero.
        // The first walk covers the pixels whose minimum edge value reaches
        // thr and returns the smallest minimum as unsigned (0 iff some pixel
        // has a minimum of exactly 0); the second, taken only where the corner
        // test accepts the block, covers those pixels.
        auto walk = [&](auto second_walk) -> uint32_t {
            constexpr bool second = decltype(second_walk)::value;
            const float z0 = __builtin_fmaf((float)yb, zdy, __builtin_fmaf((float)l, zdx, zc0)) + (float)xb * zdx;
            const float z1 = __builtin_fmaf((float)yb, zdy, __builtin_fmaf((float)(l + 1u), zdx, zc0)) + (float)xb * zdx;
            // the keys of the pair's pixels in the current row: depth bits
            // above the order word, the depth advanced in place
            unsigned long long q0 = pin(((unsigned long long)fbits(z0) << 32) | lw);
            unsigned long long q1 = pin(((unsigned long long)fbits(z1) << 32) | lw);
            // edge values of the pair's two pixels in the current row
            uint32_t E[3], F[3];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                E[kk] = ct[kk] + dyu[kk] * x + dxu[kk] * yb;
                F[kk] = E[kk] + dyu[kk];
            }
            uint32_t zmin = 0xFFFFFFFFu;
#pragma unroll
            for (uint32_t r = 0; r < 8u; ++r) {
                const int32_t m0 = min(min((int32_t)E[0], (int32_t)E[1]), (int32_t)E[2]);
                const int32_t m1 = min(min((int32_t)F[0], (int32_t)F[1]), (int32_t)F[2]);
                const bool cov0 = second ? m0 == 0 : m0 >= thr;
                const bool cov1 = second ? m1 == 0 : m1 >= thr;
                if (!second) zmin = min(zmin, min((uint32_t)m0, (uint32_t)m1));
                if (STATS) frags += (cov0 ? 1u : 0u) + (cov1 ? 1u : 0u);
                const uint32_t ro = 8u * S * r;
                auto* const k0 = reinterpret_cast<unsigned long long*>(kb + ((r & 1u) ? b0o : b0e) + ro);
                auto* const k1 = reinterpret_cast<unsigned long long*>(kb + ((r & 1u) ? b1o : b1e) + ro);
                if (cov0 & (ffrom((uint32_t)(q0 >> 32)) > 0.0f)) atomicMax(k0, q0);
                if (cov1 & (ffrom((uint32_t)(q1 >> 32)) > 0.0f)) atomicMax(k1, q1);
                q0 = pin(((unsigned long long)fbits(ffrom((uint32_t)(q0 >> 32)) + zdy) << 32) | (uint32_t)q0);
                q1 = pin(((unsigned long long)fbits(ffrom((uint32_t)(q1 >> 32)) + zdy) << 32) | (uint32_t)q1);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    E[kk] += dxu[kk];
                    F[kk] += dxu[kk];
                }
            }
            return zmin;
        };
        bool zero = false;
        if (walk_rows) zero = walk(std::false_type{}) == 0u && tight;
        // an edge exactly through a pixel of a tight primitive: the corner
        // test decides it (rare; a wave-uniform branch)
        if (__ballot(zero))
            if (zero && block_class(ct, dxu, dyu, xb, yb) != 0u) (void)walk(std::true_type{});
    }
}

// Waves per SIMD the compiler is asked to keep room for (registers): six
// for the frame kernels whose LDS allows six (strips of at most 1,024 keys
// with four waves' staging: 24 KB or less per workgroup); measured 5% faster
// than the five that 88 registers give on 32 x 32 bins. The statistics and
// eight-plane kernels keep their own register count (a bound would spill).
__host__ __device__ constexpr uint32_t backend_min_waves(uint32_t S, uint32_t H, bool stats, bool wide)
{
    return (!stats && !wide && backend_threads(S, H) == 256u && S * H <= 1024u) ? 6u : 1u;
}

template <uint32_t S, uint32_t H, bool STATS, bool WIDE>
__global__ __launch_bounds__(backend_threads(S, H), backend_min_waves(S, H, STATS, WIDE)) void k_backend(FrameDev F)
{
    constexpr uint32_t T = backend_threads(S, H);
    constexpr uint32_t W = T / 64u;
    constexpr uint32_t K = kBin / S;
    constexpr uint32_t PARTS = S / H;  // workgroups per bin
    constexpr uint32_t NPIX = S * H;
    constexpr uint32_t PPT = NPIX / T;
    constexpr uint32_t BATCH = PPT < 4 ? PPT : 4;
    static_assert(H % 8u == 0 && S % H == 0, "strips are whole 8-pixel block rows of the bin");
    static_assert(NPIX % T == 0 && PPT % BATCH == 0, "strip must hold a whole number of pixels per thread");
    __shared__ unsigned long long keys[NPIX];
    __shared__ uint32_t s_stage[W * kStageWords];
    // the next queue item, handed to the workgroup through the first key slot
    // (free between the end of one item's shading and its key initialisation)
    uint32_t& s_q = *(uint32_t*)&keys[0];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6));
    unsigned long long frags = 0;
    if (F.counters[C_FLAGS] & (FLAG_ENTRY_OVERFLOW | FLAG_CLIP_OVERFLOW)) return;
    const uint32_t qlen = F.counters[C_QLEN] * PARTS;  // strips of the queued bins, a bin's strips adjacent
    // per-bin costs: requested, and (for the learned order) a recording frame
    const bool timed = (F.flags & FF_TILE_COSTS) && (!F.sched_gate || F.sched_gate[1]);

    for (;;) {
        if (tid == 0) s_q = atomicAdd(&F.counters[C_QHEAD], 1u);
        __syncthreads();
        const uint32_t qi = s_q;
        __syncthreads();
        if (qi >= qlen) break;
        const uint64_t t_start = timed ? clock64() : 0;
        const uint32_t gb = F.queue[qi / PARTS];
        const ViewDev V = F.views[view_of_bin(F, gb)];
        const uint32_t lb = gb - V.bin_base;
        const uint32_t sbx = lb % V.bins_x, sby = lb / V.bins_x;
        const uint32_t tx = sbx / K, ty = sby / K;
        const uint32_t sx0 = (sbx % K) * S, sy0 = (sby % K) * S + (qi % PARTS) * H;
        const int32_t X0 = (int32_t)(tx * kBin), Y0 = (int32_t)(ty * kBin);
        const size_t tofs = (size_t)(ty * V.tiles_x + tx) * kTilePixels;
        const bool clear_d = (V.flags & VF_CLEAR_DEPTH) != 0;
        const bool clear_c = (V.flags & VF_CLEAR_COLOUR) != 0;

        for (uint32_t p = tid; p < NPIX; p += T) {
            const uint32_t x = sx0 + p % S, y = sy0 + p / S;
            keys[key_slot<S>(p % S, p / S)] = clear_d ? 0ull : init_key(global_ptr(V.depth)[tofs + y * kBin + x]);
        }
        __syncthreads();

        const uint32_t beg = F.bin_start[gb], cnt = F.bin_count[gb];
        for (uint32_t b0 = 0; b0 + wave < cnt; b0 += T) {
            // lane j of wave w holds entry b0 + w + W*j
            const uint32_t my = b0 + wave + W * lane;
            const bool has = my < cnt;
            uint4 q0 = make_uint4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
            uint32_t order = 0;
            if (has) {
                const uint32_t r = F.entries[beg + my];
                const uint4* R = F.rec_raster + (size_t)r * 4;
                q0 = R[0]; q1 = R[1]; q2 = R[2]; q3 = R[3];
                order = r < F.ntris ? ((r << 3) | kFanUnclipped) : F.clip_order[r - F.ntris];
            }
            const EntryRaster e = entry_raster(q0, q1, q2, q3, order, X0, Y0, sx0, sy0, S, H);
            raster_wave<S, STATS>(keys, e, has, sx0, sy0, frags, s_stage + wave * kStageWords);
        }
        __syncthreads();

        for (uint32_t p0 = 0; p0 < PPT; p0 += BATCH) {
            uint32_t rec[BATCH], col[BATCH];
            unsigned long long kk[BATCH];
            bool win[BATCH];
#pragma unroll
            for (uint32_t i = 0; i < BATCH; ++i) {
                const uint32_t p = tid + (p0 + i) * T;
                kk[i] = keys[key_slot<S>(p % S, p / S)];
                const uint32_t lo = (uint32_t)kk[i];
                win[i] = lo != 0xFFFFFFFFu && kk[i] != 0ull;
                rec[i] = win[i] ? record_of_order(F, 0xFFFFFFFEu - lo) : 0u;
            }
#pragma unroll
            for (uint32_t i = 0; i < BATCH; ++i) {
                const uint32_t p = tid + (p0 + i) * T;
                col[i] = (win[i] && (V.flags & VF_HAS_COLOUR)) ? shade_pixel<WIDE>(F, rec[i], X0, Y0, sx0 + p % S, sy0 + p / S)
                                                               : V.clear_colour;
            }
#pragma unroll
            for (uint32_t i = 0; i < BATCH; ++i) {
                const uint32_t p = tid + (p0 + i) * T;
                const size_t o = tofs + (sy0 + p / S) * kBin + sx0 + p % S;
                const auto colour = global_ptr(V.colour);
                const auto depth = global_ptr(V.depth);
                if (win[i]) {
                    if (V.flags & VF_HAS_COLOUR) colour[o] = col[i];
                    depth[o] = ffrom((uint32_t)(kk[i] >> 32));
                    if (V.primid) global_ptr(V.primid)[o] = 0xFFFFFFFEu - (uint32_t)kk[i];
                } else {
                    if (clear_c) colour[o] = col[i];
                    if (clear_d) depth[o] = 0.0f;
                }
            }
        }
        // a bin's cost is the sum over its strips (zeroed at frame start)
        if (timed && tid == 0)
            atomicAdd((unsigned long long*)&F.bin_cost[gb], (unsigned long long)(clock64() - t_start));
        __syncthreads();
    }
    if (STATS) {
        // per-lane counts -> wave total -> one atomic per wave
        uint32_t lo = (uint32_t)frags, hi = (uint32_t)(frags >> 32);
        lo = wave_sum(lo);
        hi = wave_sum(hi);
        const unsigned long long total = ((unsigned long long)hi << 32) + lo;
        if (lane == 0 && total) atomicAdd((unsigned long long*)&F.counters[C_FRAGS], total);
    }
}

}  // namespace srg
