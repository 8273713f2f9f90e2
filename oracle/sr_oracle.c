/*
 * sr_oracle.c -- CPU oracle for SoftRast's sort-middle rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg. The product library never
 * links or calls it.
 *
 * Function map (reference file:line it restates) and how each is pinned:
 *   fetch / transform      Binning.cpp:167-212,474-479   KAT vs float64 closed form (tests)
 *   clip_codes             Binning.cpp:56-68             KAT on hand-built vertices
 *   clip_polygon / fan     Binning.cpp:85-165,502-533    invariants: every output vertex inside
 *                                                       all planes (to rounding), fan count n-2,
 *                                                       unclipped triangles pass unchanged
 *   setup (snap, cull,     Binning.cpp:279-350           brute force on tiny triangles against
 *     bbox, edges, planes)                              exact rational edge functions; planes
 *                                                       reproduce f at the three vertices
 *   bin loop               Binning.cpp:352-455           brute force: every covered pixel lies in a
 *                                                       binned tile; the numXbins quirk (:359) kept
 *   raster (block walk,    Rasterizer.cpp:97-304         brute force per pixel against the
 *     coverage, depth)                                  closed-form predicate; depth vs float64 plane
 *   interpolants           Rasterizer.cpp:306-419        closed form on an affine (w = 1) triangle
 *   rcpps                  Rasterizer.cpp:375-376        recorded host table (tests/golden)
 *   rsqrtps                SponzaScene.cpp:67            recorded host table (tests/golden)
 *   SponzaShader           SponzaScene.cpp:13-104        closed forms (one light, no light, NaN)
 *   calc_mip               Texture.cpp:212-233           known-gradient cases (exact powers of two)
 *   sample_wrap / gather   Texture.cpp:236-452           constant texture, 0/1 weights, wrap symmetry,
 *                                                       Morton = bit-interleave bijection
 *   pack                   SIMDUtil.h:87-121             table of packus saturation cases
 *   shaders                Viewer/Shaders.h:71-130       closed forms
 *   clear / blit           Renderer.cpp:168-194,319-347  closed forms
 *   mip_resize_rgba8       Texture.cpp:188-198 ->        the vendored stb_image_resize.h compiled
 *                          stb_image_resize.h:2464-2471  as is (oracle/_ref) and golden vectors made
 *                                                        from it (tests/golden/stbir_mips.npz)
 * Semantics the reference leaves to its absent `kt` library (op order of
 * Mat4*Vec4, Dot, Lerp; radix-sort stability) are fixed here and marked
 * "defined here"; they have no reference test to pin them.
 */
#include "sr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define TRIS_PER_CHUNK 32u   /* Binning.h:16 c_trisPerBinChunk */
#define PACKET_TRIS 2048u    /* Renderer.cpp:247 front-end task granularity */
#define SUBPIXEL_BITS 8
#define SUBPIXEL_STEP 256
#define SUBPIXEL_MASK 255
#define CLIP_VERTS 9         /* Binning.cpp:71 */

static const uint32_t g_rcpps_table[2048] = {
#include "rcpps_table.inc"
};
static const uint32_t g_rsqrtps_table[4096] = {
#include "rsqrtps_table.inc"
};

/* ------------------------------------------------------------------ */
/* scalar helpers                                                      */
/* ------------------------------------------------------------------ */

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float ffrom(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* C float->int32 cast as compiled for x86 (cvttss2si): truncation; NaN and
 * out-of-range give INT32_MIN. Binning.cpp:301-303. */
static inline int32_t cvtt_i32(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int32_t)f;
    return INT32_MIN;
}

/* _mm256_cvtps_epi32 under the default rounding mode: round to nearest even;
 * NaN and out-of-range give INT32_MIN. Texture.cpp:436-437, SIMDUtil.h:92-95. */
static inline int32_t cvtn_i32(float f)
{
    float r = nearbyintf(f);
    if (r >= -2147483648.0f && r < 2147483648.0f) return (int32_t)r;
    return INT32_MIN;
}

/* _mm256_max_ps(a, b): a > b ? a : b (b when either is NaN). */
static inline float maxps(float a, float b) { return a > b ? a : b; }

/* _mm256_rcp_ps, Rasterizer.cpp:375-376: table of the top 11 mantissa bits,
 * exact exponent scaling, subnormal results flushed to zero. */
/* _mm256_rsqrt_ps on the recorded host (tests/golden/gen_rsqrtps_table.c):
 * T[E & 1][top 11 mantissa bits] scaled by 2^-(E - (E & 1)) / 2 for a
 * positive normal x = 2^E * f; +-0 and denormals -> +-inf, +inf -> +0,
 * other negatives -> the default NaN, a NaN -> that NaN quieted. */
float sro_rsqrtps(float x)
{
    uint32_t u = fbits(x);
    uint32_t e = (u >> 23) & 255u;
    if (e == 255u && (u & 0x7FFFFFu)) return ffrom(u | 0x00400000u);
    if (e == 0u) return ffrom((u & 0x80000000u) | 0x7F800000u);
    if (u & 0x80000000u) return ffrom(0xFFC00000u);
    if (e == 255u) return 0.0f;
    int ub = (int)e - 127, p = ub & 1;
    uint32_t t = g_rsqrtps_table[(uint32_t)p * 2048u + ((u >> 12) & 2047u)];
    int re = (int)((t >> 23) & 255u) - (ub - p) / 2;
    return ffrom(((uint32_t)re << 23) | (t & 0x7FFFFFu));
}

float sro_rcpps(float x)
{
    uint32_t u = fbits(x);
    uint32_t sign = u & 0x80000000u;
    uint32_t e = (u >> 23) & 255u;
    if (e == 255u) {
        if (u & 0x7FFFFFu) return ffrom(u | 0x00400000u); /* NaN: quieted input */
        return ffrom(sign);                                /* +-inf -> +-0 */
    }
    if (e == 0u) return ffrom(sign | 0x7F800000u);         /* zero / denormal -> +-inf */
    uint32_t t = g_rcpps_table[(u >> 12) & 2047u];
    int ne = (int)((t >> 23) & 255u) - ((int)e - 127);
    if (ne <= 0) return ffrom(sign);
    return ffrom(sign | ((uint32_t)ne << 23) | (t & 0x7FFFFFu));
}

/* SIMDUtil.h:87-121 per channel: cvt_RNE(fma(c, 255, 0.5)) then packus_epi32
 * (unsigned 16-bit saturation) then packus_epi16 on signed 16-bit input. */
static inline uint32_t pack_channel(float c)
{
    int32_t i = cvtn_i32(fmaf(c, 255.0f, 0.5f));
    if (i < 0) return 0;
    if (i <= 255) return (uint32_t)i;
    if (i <= 32767) return 255;
    return 0;
}

uint32_t sro_pack_rgba(float r, float g, float b, float a)
{
    return pack_channel(r) | (pack_channel(g) << 8) | (pack_channel(b) << 16) | (pack_channel(a) << 24);
}

/* Texture.cpp:36-41: x bits at even positions, y bits at odd positions. */
uint32_t sro_morton5(uint32_t x, uint32_t y)
{
    uint32_t m = 0;
    for (uint32_t b = 0; b < 16; ++b) {
        m |= ((x >> b) & 1u) << (2 * b);
        m |= ((y >> b) & 1u) << (2 * b + 1);
    }
    return m;
}

/* ------------------------------------------------------------------ */
/* sampling: Texture.cpp:212-452                                       */
/* ------------------------------------------------------------------ */

/* CalcMipLevels, Texture.cpp:212-233 (note dudy*height and dvdx*width). */
int32_t sro_calc_mip(const sro_texture* tex, float dudx, float dudy, float dvdx, float dvdy)
{
    float height = (float)(1u << tex->hlog2);
    float width = (float)(1u << tex->wlog2);
    float p = dudx * width, q = dudy * height, r = dvdx * width, s = dvdy * height;
    float du2 = fmaf(p, p, q * q);
    float dv2 = fmaf(r, r, s * s);
    float m = sqrtf(maxps(du2, dv2));
    int32_t e = (int32_t)((((int32_t)fbits(m)) >> 23) & 0xFF) - 127;
    int32_t lo = e > 0 ? e : 0;
    int32_t hi = (int32_t)tex->nmips - 1;
    return hi < lo ? hi : lo;
}

static inline float texel_channel(uint8_t b) { return (float)b * (1.0f / 255.0f); }

/* simdutil::Lerp, SIMDUtil.h:136-145: fma(t, b, fnmadd(t, a, a)). */
static inline float lerp_fma(float a, float b, float t) { return fmaf(t, b, fmaf(-t, a, a)); }

/* SampleWrap + GatherQuadsAndInterpolate for one lane, Texture.cpp:381-452,243-379.
 * Returns the mip level used; writes filtered RGBA in [0,1]. */
uint32_t sro_sample_wrap(const sro_texture* tex, float u, float v, float dudx, float dudy,
                         float dvdx, float dvdy, float out[4])
{
    int32_t mip = sro_calc_mip(tex, dudx, dudy, dvdx, dvdy);
    int32_t wl = (int32_t)tex->wlog2, hl = (int32_t)tex->hlog2;
    int32_t mw = 1 << (wl - (wl < mip ? wl : mip));
    int32_t mh = 1 << (hl - (hl < mip ? hl : mip));

    uint32_t us = fbits(u) & 0x80000000u, vs = fbits(v) & 0x80000000u;
    float au = ffrom(fbits(u) ^ us), av = ffrom(fbits(v) ^ vs);
    float fu = au - floorf(au), fv = av - floorf(av);
    float fuw = us ? 1.0f - fu : fu;
    float fvw = vs ? 1.0f - fv : fv;
    float ut = (float)mw * fuw, vt = (float)mh * fvw;
    float uf = floorf(ut), vf = floorf(vt);
    float ui = ut - uf, vi = vt - vf;
    int32_t x0 = cvtn_i32(uf) & (mw - 1), y0 = cvtn_i32(vf) & (mh - 1);
    int32_t x1 = (x0 + 1) & (mw - 1), y1 = (y0 + 1) & (mh - 1);

    int32_t tile_w = (mw > 32 ? mw : 32) >> 5;
    const uint8_t* base = tex->texels + tex->mip_offsets[mip];
    int32_t xs[2] = { x0, x1 }, ys[2] = { y0, y1 };
    float t[2][2][4]; /* [x][y][channel] */
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            uint32_t x = (uint32_t)xs[i], y = (uint32_t)ys[j];
            uint32_t offs = 1024u * ((y >> 5) * (uint32_t)tile_w + (x >> 5)) + sro_morton5(x & 31u, y & 31u);
            const uint8_t* px = base + 4u * offs;
            for (int c = 0; c < 4; ++c) t[i][j][c] = texel_channel(px[c]);
        }
    for (int c = 0; c < 4; ++c) {
        float left = lerp_fma(t[0][0][c], t[0][1][c], vi);
        float right = lerp_fma(t[1][0][c], t[1][1][c], vi);
        out[c] = lerp_fma(left, right, ui);
    }
    return (uint32_t)mip;
}

/* ------------------------------------------------------------------ */
/* front end: Binning.cpp                                              */
/* ------------------------------------------------------------------ */

typedef struct { float x, y, z, w; } vec4;

/* kt::Mat4 * kt::Vec4 (Binning.cpp:477-479); op order defined here:
 * ((c0*x + c1*y) + c2*z) + c3*w, every op rounded. */
static inline vec4 transform(const float m[16], float x, float y, float z)
{
    vec4 r;
    const float w = 1.0f;
    r.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
    r.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
    r.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
    r.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
    return r;
}

void sro_transform(const float mvp[16], const float p[3], float out[4])
{
    vec4 r = transform(mvp, p[0], p[1], p[2]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

/* ComputeClipMask, Binning.cpp:56-68. */
static inline uint32_t clip_codes(vec4 v)
{
    uint32_t m = 0;
    if (v.x + v.w < 0.0f) m |= 0x01;
    if (v.x - v.w > 0.0f) m |= 0x02;
    if (v.y + v.w < 0.0f) m |= 0x04;
    if (v.y - v.w > 0.0f) m |= 0x08;
    if (v.z < 0.0f) m |= 0x10;
    if (v.z - v.w > 0.0f) m |= 0x20;
    return m;
}

uint32_t sro_clip_codes(const float v[4])
{
    vec4 t = { v[0], v[1], v[2], v[3] };
    return clip_codes(t);
}

static const float k_clip_planes[6][4] = {
    { 1.0f, 0.0f, 0.0f, 1.0f }, { -1.0f, 0.0f, 0.0f, 1.0f },
    { 0.0f, 1.0f, 0.0f, 1.0f }, { 0.0f, -1.0f, 0.0f, 1.0f },
    { 0.0f, 0.0f, 1.0f, 1.0f }, { 0.0f, 0.0f, -1.0f, 1.0f },
};

/* kt::Dot (Binning.cpp:111,116); op order defined here. */
static inline float dot4(const float p[4], vec4 v) { return ((p[0] * v.x + p[1] * v.y) + p[2] * v.z) + p[3] * v.w; }
/* kt::Lerp (Binning.cpp:136-153); defined here as a + (b - a) * t. */
static inline float lerp_kt(float a, float b, float t) { return a + (b - a) * t; }

typedef struct {
    vec4 verts[2][CLIP_VERTS];
    float attribs[2][CLIP_VERTS][SRO_MAX_VARYINGS];
    uint32_t n_in, n_out, in_idx;
} clip_buffer;

/* ClipPlane, Binning.cpp:85-165 (Sutherland-Hodgman against one plane). */
static void clip_plane(clip_buffer* b, uint32_t plane)
{
    const float* cp = k_clip_planes[plane];
    uint32_t i0 = b->n_in - 1;
    vec4* in_v = b->verts[b->in_idx];
    float (*in_a)[SRO_MAX_VARYINGS] = b->attribs[b->in_idx];
    vec4* out_v = b->verts[b->in_idx ^ 1];
    float (*out_a)[SRO_MAX_VARYINGS] = b->attribs[b->in_idx ^ 1];
    float d0 = dot4(cp, in_v[i0]);
    for (uint32_t i1 = 0; i1 < b->n_in; ++i1) {
        float d1 = dot4(cp, in_v[i1]);
        int in0 = d0 >= 0.0f, in1 = d1 >= 0.0f;
        /* the reference asserts room for every emission (Binning.cpp:122,
         * :133, :145); one that would overflow CLIP_VERTS is dropped */
        if (in0 && b->n_out < CLIP_VERTS) {
            uint32_t o = b->n_out++;
            out_v[o] = in_v[i0];
            memcpy(out_a[o], in_a[i0], sizeof(float) * SRO_MAX_VARYINGS);
        }
        if ((in0 ^ in1) && b->n_out < CLIP_VERTS) {
            uint32_t o = b->n_out++;
            uint32_t a, c; float t;
            if (in1) { t = d1 / (d1 - d0); a = i1; c = i0; }
            else     { t = d0 / (d0 - d1); a = i0; c = i1; }
            out_v[o].x = lerp_kt(in_v[a].x, in_v[c].x, t);
            out_v[o].y = lerp_kt(in_v[a].y, in_v[c].y, t);
            out_v[o].z = lerp_kt(in_v[a].z, in_v[c].z, t);
            out_v[o].w = lerp_kt(in_v[a].w, in_v[c].w, t);
            for (uint32_t k = 0; k < SRO_MAX_VARYINGS; ++k) out_a[o][k] = lerp_kt(in_a[a][k], in_a[c][k], t);
        }
        d0 = d1;
        i0 = i1;
    }
    b->n_in = b->n_out;
    b->n_out = 0;
    b->in_idx ^= 1;
}

/* Clip loop, Binning.cpp:505-523: set bits of the OR mask, lowest first,
 * stopping early when the polygon empties. */
static void clip_polygon(clip_buffer* b, uint32_t mask)
{
    do {
        uint32_t plane = (uint32_t)__builtin_ctz(mask);
        mask ^= 1u << plane;
        clip_plane(b, plane);
    } while (mask && b->n_in);
}

uint32_t sro_clip_triangle(const float verts[12], const float attribs[24], uint32_t mask,
                           float out_verts[36], float out_attribs[72])
{
    clip_buffer b;
    memset(&b, 0, sizeof(b));
    b.n_in = 3;
    for (int i = 0; i < 3; ++i) {
        b.verts[0][i].x = verts[4 * i]; b.verts[0][i].y = verts[4 * i + 1];
        b.verts[0][i].z = verts[4 * i + 2]; b.verts[0][i].w = verts[4 * i + 3];
        memcpy(b.attribs[0][i], attribs + 8 * i, 32);
    }
    if (mask) clip_polygon(&b, mask);
    for (uint32_t i = 0; i < b.n_in; ++i) {
        vec4 v = b.verts[b.in_idx][i];
        out_verts[4 * i] = v.x; out_verts[4 * i + 1] = v.y; out_verts[4 * i + 2] = v.z; out_verts[4 * i + 3] = v.w;
        memcpy(out_attribs + 8 * i, b.attribs[b.in_idx][i], 32);
    }
    return b.n_in;
}

typedef struct { int32_t c[3], dx[3], dy[3]; uint8_t bminx, bmaxx, bminy, bmaxy; } edge_eq;
typedef struct { float c0, dx, dy; } plane_eq;

/* Per-primitive setup results (the values BinTransformedAndClippedTri holds
 * before its bin loop). */
typedef struct {
    int32_t fp[3][2];
    int32_t c[3], dx[3], dy[3];
    int32_t xmin, xmax, ymin, ymax;
    float r0x, r0y;
    float z_dx, z_dy, z_f0;
    float w_dx, w_dy, w_f0;
    float a_dx[SRO_MAX_VARYINGS], a_dy[SRO_MAX_VARYINGS], a_f0[SRO_MAX_VARYINGS];
} tri_setup;

/* SetupEdge, Binning.cpp:242-259. */
static inline void setup_edge(tri_setup* s, int k, const int32_t a[2], const int32_t b[2])
{
    int32_t dy = (int32_t)((uint32_t)b[1] - (uint32_t)a[1]);
    int32_t dx = (int32_t)((uint32_t)a[0] - (uint32_t)b[0]);
    int64_t c = (int64_t)a[1] * (int64_t)(int32_t)((uint32_t)b[0] - (uint32_t)a[0]) -
                (int64_t)a[0] * (int64_t)(int32_t)((uint32_t)b[1] - (uint32_t)a[1]);
    if (dy < 0 || (dy == 0 && dx > 0)) c += SUBPIXEL_STEP;
    s->dy[k] = dy;
    s->dx[k] = dx;
    s->c[k] = (int32_t)(uint32_t)(c >> SUBPIXEL_BITS);
}

/* SetupPlane, Binning.cpp:261-277. */
static inline void setup_plane(float cterm, float d10x, float d10y, float d20x, float d20y, float f10,
                               float f20, float* odx, float* ody)
{
    float A = d10y * f20 - f10 * d20y;
    float B = d20x * f10 - d10x * f20;
    *odx = -A / cterm;
    *ody = -B / cterm;
}

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int32_t min3i(int32_t a, int32_t b, int32_t c) { int32_t m = a < b ? a : b; return m < c ? m : c; }
static inline int32_t max3i(int32_t a, int32_t b, int32_t c) { int32_t m = a > b ? a : b; return m > c ? m : c; }

/* Binning.cpp:291-350: viewport, snap, cull, bbox, edges, planes.
 * Returns 0 when culled. */
static int setup_tri(tri_setup* s, vec4 v0, vec4 v1, vec4 v2, const float* a0, const float* a1,
                     const float* a2, uint32_t nattr, uint32_t width, uint32_t height)
{
    float hw = (float)width * 0.5f, hh = (float)height * 0.5f;
    float iw0 = 1.0f / v0.w, iw1 = 1.0f / v1.w, iw2 = 1.0f / v2.w;
    float r0x = iw0 * v0.x * hw + hw, r0y = iw0 * v0.y * -hh + hh;
    float r1x = iw1 * v1.x * hw + hw, r1y = iw1 * v1.y * -hh + hh;
    float r2x = iw2 * v2.x * hw + hw, r2y = iw2 * v2.y * -hh + hh;
    const float step = (float)SUBPIXEL_STEP;
    s->fp[0][0] = cvtt_i32(r0x * step + 0.5f); s->fp[0][1] = cvtt_i32(r0y * step + 0.5f);
    s->fp[1][0] = cvtt_i32(r1x * step + 0.5f); s->fp[1][1] = cvtt_i32(r1y * step + 0.5f);
    s->fp[2][0] = cvtt_i32(r2x * step + 0.5f); s->fp[2][1] = cvtt_i32(r2y * step + 0.5f);

    int64_t area = (int64_t)(int32_t)((uint32_t)s->fp[2][0] - (uint32_t)s->fp[0][0]) *
                       (int64_t)(int32_t)((uint32_t)s->fp[1][1] - (uint32_t)s->fp[0][1]) -
                   (int64_t)(int32_t)((uint32_t)s->fp[2][1] - (uint32_t)s->fp[0][1]) *
                       (int64_t)(int32_t)((uint32_t)s->fp[1][0] - (uint32_t)s->fp[0][0]);
    area >>= SUBPIXEL_BITS;
    if (area <= 0) return 0;

    int32_t mnx = min3i(s->fp[0][0], s->fp[1][0], s->fp[2][0]);
    int32_t mny = min3i(s->fp[0][1], s->fp[1][1], s->fp[2][1]);
    int32_t mxx = max3i(s->fp[0][0], s->fp[1][0], s->fp[2][0]);
    int32_t mxy = max3i(s->fp[0][1], s->fp[1][1], s->fp[2][1]);
    s->xmin = (uint16_t)clampi((int32_t)((uint32_t)mnx + SUBPIXEL_MASK) >> SUBPIXEL_BITS, 0, (int32_t)width - 1);
    s->ymin = (uint16_t)clampi((int32_t)((uint32_t)mny + SUBPIXEL_MASK) >> SUBPIXEL_BITS, 0, (int32_t)height - 1);
    s->xmax = (uint16_t)clampi((int32_t)((uint32_t)mxx + SUBPIXEL_MASK) >> SUBPIXEL_BITS, 0, (int32_t)width - 1);
    s->ymax = (uint16_t)clampi((int32_t)((uint32_t)mxy + SUBPIXEL_MASK) >> SUBPIXEL_BITS, 0, (int32_t)height - 1);

    setup_edge(s, 0, s->fp[0], s->fp[1]);
    setup_edge(s, 1, s->fp[1], s->fp[2]);
    setup_edge(s, 2, s->fp[2], s->fp[0]);

    float d10x = r1x - r0x, d10y = r1y - r0y;
    float d20x = r2x - r0x, d20y = r2y - r0y;
    float cterm = d10x * d20y - d10y * d20x;

    float z0 = v0.z * iw0;
    setup_plane(cterm, d10x, d10y, d20x, d20y, v1.z * iw1 - v0.z * iw0, v2.z * iw2 - v0.z * iw0, &s->z_dx, &s->z_dy);
    setup_plane(cterm, d10x, d10y, d20x, d20y, iw1 - iw0, iw2 - iw0, &s->w_dx, &s->w_dy);
    for (uint32_t k = 0; k < nattr; ++k)
        setup_plane(cterm, d10x, d10y, d20x, d20y, a1[k] * iw1 - a0[k] * iw0, a2[k] * iw2 - a0[k] * iw0,
                    &s->a_dx[k], &s->a_dy[k]);
    s->r0x = r0x; s->r0y = r0y;
    s->z_f0 = z0; s->w_f0 = iw0;
    for (uint32_t k = 0; k < nattr; ++k) s->a_f0[k] = a0[k] * iw0;
    return 1;
}

int sro_setup_triangle(const float v[12], const float attribs[24], uint32_t nattr, uint32_t width,
                       uint32_t height, int32_t ints[19], float floats[32])
{
    tri_setup s;
    vec4 a = { v[0], v[1], v[2], v[3] }, b = { v[4], v[5], v[6], v[7] }, c = { v[8], v[9], v[10], v[11] };
    if (nattr > SRO_MAX_VARYINGS) return -1;
    if (!setup_tri(&s, a, b, c, attribs, attribs + 8, attribs + 16, nattr, width, height)) return 0;
    for (int k = 0; k < 3; ++k) { ints[k] = s.c[k]; ints[3 + k] = s.dx[k]; ints[6 + k] = s.dy[k]; }
    ints[9] = s.xmin; ints[10] = s.xmax; ints[11] = s.ymin; ints[12] = s.ymax;
    for (int k = 0; k < 3; ++k) { ints[13 + 2 * k] = s.fp[k][0]; ints[14 + 2 * k] = s.fp[k][1]; }
    floats[0] = s.r0x; floats[1] = s.r0y;
    floats[2] = s.z_dx; floats[3] = s.z_dy; floats[4] = s.z_f0;
    floats[5] = s.w_dx; floats[6] = s.w_dy; floats[7] = s.w_f0;
    for (uint32_t k = 0; k < nattr; ++k) { floats[8 + 3 * k] = s.a_dx[k]; floats[9 + 3 * k] = s.a_dy[k]; floats[10 + 3 * k] = s.a_f0[k]; }
    return 1;
}

/* ------------------------------------------------------------------ */
/* bins, chunks, thread state                                          */
/* ------------------------------------------------------------------ */

/* BinChunk, Binning.h:20-57, plus the order key of each triangle. */
typedef struct {
    edge_eq edge[TRIS_PER_CHUNK];
    plane_eq recipw[TRIS_PER_CHUNK];
    plane_eq zoverw[TRIS_PER_CHUNK];
    float adx[TRIS_PER_CHUNK * SRO_MAX_VARYINGS];
    float ady[TRIS_PER_CHUNK * SRO_MAX_VARYINGS];
    float ac[TRIS_PER_CHUNK * SRO_MAX_VARYINGS];
    uint32_t primid[TRIS_PER_CHUNK];
    uint32_t attribs_per_tri, num_tris, draw_idx, packet_idx;
} chunk_t;

typedef struct { uint32_t* items; uint32_t n, cap; } u32_vec;

static int u32_push(u32_vec* v, uint32_t x)
{
    if (v->n == v->cap) {
        uint32_t nc = v->cap ? v->cap * 2 : 16;
        uint32_t* p = (uint32_t*)realloc(v->items, sizeof(uint32_t) * nc);
        if (!p) return -1;
        v->items = p;
        v->cap = nc;
    }
    v->items[v->n++] = x;
    return 0;
}

/* Rasterizer.cpp:19-33 Frag, with a 32-bit chunk index. */
typedef struct { uint8_t x, y, tri, pad; uint32_t chunk; } frag_t;

typedef struct { uint32_t thread, chunk, draw, packet, pos; } chunk_ref;

typedef struct {
    /* chunk arena (ThreadScratchAllocator stand-in), reset per frame */
    chunk_t** blocks;
    uint32_t nblocks, nchunks;
    /* bins: one per (plane, tile) */
    u32_vec* bins;
    uint32_t nbins;
    /* back-end scratch */
    frag_t* frags;
    uint32_t frag_cap;
    chunk_ref* refs;
    uint32_t ref_cap;
    /* statistics */
    uint64_t tris_clipped, prims_binned, bin_entries, frags_raster, frags_passed, blocks_tested;
    int error;
} thread_state;

#define CHUNKS_PER_BLOCK 256u

static chunk_t* chunk_at(thread_state* ts, uint32_t i) { return &ts->blocks[i / CHUNKS_PER_BLOCK][i % CHUNKS_PER_BLOCK]; }

static int chunk_alloc(thread_state* ts, uint32_t* idx)
{
    if (ts->nchunks == ts->nblocks * CHUNKS_PER_BLOCK) {
        chunk_t** nb = (chunk_t**)realloc(ts->blocks, sizeof(chunk_t*) * (ts->nblocks + 1));
        if (!nb) return -1;
        ts->blocks = nb;
        ts->blocks[ts->nblocks] = (chunk_t*)malloc(sizeof(chunk_t) * CHUNKS_PER_BLOCK);
        if (!ts->blocks[ts->nblocks]) return -1;
        ts->nblocks++;
    }
    *idx = ts->nchunks++;
    return 0;
}

struct sro_ctx {
    sro_sponza_constants sponza;
    uint32_t nthreads;
    pthread_t* threads;
    pthread_mutex_t mu;
    pthread_cond_t cv_start, cv_done;
    uint64_t generation;
    uint32_t running;
    int quit;
    void (*job)(sro_ctx*, uint32_t);
    thread_state* ts;

    /* current frame */
    const sro_draw* draws;
    uint32_t ndraws;
    sro_plane* planes;
    uint32_t nplanes;
    uint32_t* tile_base;   /* per plane: first bin index */
    uint32_t total_tiles;
    uint32_t* tri_base;    /* per draw: first global triangle index */
    uint32_t* packet_draw; /* per packet */
    uint32_t* packet_begin;
    uint32_t npackets;
    atomic_uint next_packet;
    uint32_t* tile_list;   /* non-empty bins, in (plane, row-major) order */
    uint32_t ntile_list;
    atomic_uint next_tile;
};

/* ------------------------------------------------------------------ */
/* binning                                                             */
/* ------------------------------------------------------------------ */

/* GetOrCreateBinForDrawCall, Binning.cpp:223-240. A chunk is also closed when
 * the packet changes so that chunks can be ordered by (draw, packet) when
 * several threads bin one draw call. */
static chunk_t* get_chunk(thread_state* ts, u32_vec* bin, uint32_t draw, uint32_t packet, uint32_t apt)
{
    if (bin->n) {
        chunk_t* last = chunk_at(ts, bin->items[bin->n - 1]);
        if (last->draw_idx == draw && last->packet_idx == packet && last->num_tris < TRIS_PER_CHUNK) return last;
    }
    uint32_t idx;
    if (chunk_alloc(ts, &idx) || u32_push(bin, idx)) { ts->error = -2; return NULL; }
    chunk_t* c = chunk_at(ts, idx);
    c->num_tris = 0;
    c->attribs_per_tri = apt;
    c->draw_idx = draw;
    c->packet_idx = packet;
    return c;
}

/* BinTransformedAndClippedTri, Binning.cpp:279-456. */
static void bin_tri(sro_ctx* ctx, thread_state* ts, uint32_t draw_idx, uint32_t packet, vec4 v0, vec4 v1,
                    vec4 v2, const float* a0, const float* a1, const float* a2, uint32_t primid)
{
    const sro_draw* d = &ctx->draws[draw_idx];
    const sro_plane* pl = &ctx->planes[d->plane];
    uint32_t apt = d->attr_stride / 4u;
    tri_setup s;
    if (!setup_tri(&s, v0, v1, v2, a0, a1, a2, apt, pl->width, pl->height)) return;
    ts->prims_binned++;

    uint32_t bymin = (uint32_t)s.ymin >> 6, bymax = (uint32_t)s.ymax >> 6;
    uint32_t bxmin = (uint32_t)s.xmin >> 6, bxmax = (uint32_t)s.xmax >> 6;
    uint32_t nybins = bymax - bymin + 1;
    uint32_t nxbins = bymax - bymin + 1; /* sic: Binning.cpp:359 computes it from Y */
    int check = !(nybins <= 2 && nxbins <= 2);

    for (uint32_t by = bymin; by <= bymax; ++by)
        for (uint32_t bx = bxmin; bx <= bxmax; ++bx) {
            uint32_t X0 = bx * SRO_BIN_SIZE, X1 = X0 + SRO_BIN_SIZE;
            uint32_t Y0 = by * SRO_BIN_SIZE, Y1 = Y0 + SRO_BIN_SIZE;
            if (check) {
                int reject = 0;
                for (int k = 0; k < 3 && !reject; ++k) {
                    uint32_t c = (uint32_t)s.c[k], dx = (uint32_t)s.dx[k], dy = (uint32_t)s.dy[k];
                    uint32_t m = ((int32_t)(c + dy * X0 + dx * Y0) > 0) | (((int32_t)(c + dy * X0 + dx * Y1) > 0) << 1) |
                                 (((int32_t)(c + dy * X1 + dx * Y0) > 0) << 2) | (((int32_t)(c + dy * X1 + dx * Y1) > 0) << 3);
                    if (!m) reject = 1;
                }
                if (reject) continue;
            }
            u32_vec* bin = &ts->bins[ctx->tile_base[d->plane] + by * pl->tiles_x + bx];
            chunk_t* ch = get_chunk(ts, bin, draw_idx, packet, apt);
            if (!ch) return;
            uint32_t t = ch->num_tris++;
            ts->bin_entries++;
            edge_eq* e = &ch->edge[t];
            for (int k = 0; k < 3; ++k) {
                e->dx[k] = s.dx[k];
                e->dy[k] = s.dy[k];
                e->c[k] = (int32_t)((uint32_t)s.c[k] + ((uint32_t)s.dx[k] * Y0 + (uint32_t)s.dy[k] * X0));
            }
            e->bminx = (uint8_t)clampi(s.xmin - (int32_t)X0, 0, SRO_BIN_SIZE);
            e->bmaxx = (uint8_t)clampi(s.xmax - (int32_t)X0, 0, SRO_BIN_SIZE);
            e->bminy = (uint8_t)clampi(s.ymin - (int32_t)Y0, 0, SRO_BIN_SIZE);
            e->bmaxy = (uint8_t)clampi(s.ymax - (int32_t)Y0, 0, SRO_BIN_SIZE);
            float sx = (float)(int32_t)X0 - s.r0x;
            float sy = (float)(int32_t)Y0 - s.r0y;
            ch->recipw[t].dx = s.w_dx; ch->recipw[t].dy = s.w_dy;
            ch->recipw[t].c0 = s.w_dx * sx + s.w_dy * sy + s.w_f0;
            ch->zoverw[t].dx = s.z_dx; ch->zoverw[t].dy = s.z_dy;
            ch->zoverw[t].c0 = s.z_dx * sx + s.z_dy * sy + s.z_f0;
            for (uint32_t k = 0; k < apt; ++k) {
                ch->adx[t * apt + k] = s.a_dx[k];
                ch->ady[t * apt + k] = s.a_dy[k];
                ch->ac[t * apt + k] = s.a_dx[k] * sx + s.a_dy[k] * sy + s.a_f0[k];
            }
            ch->primid[t] = primid;
        }
}

static inline uint32_t fetch_index(const sro_draw* d, uint32_t i)
{
    switch (d->idx_stride) {
    case 1: return ((const uint8_t*)d->idx)[i];
    case 2: return ((const uint16_t*)d->idx)[i];
    default: return ((const uint32_t*)d->idx)[i];
    }
}

/* BinTrisEntry, Binning.cpp:464-535. */
static void bin_packet(sro_ctx* ctx, thread_state* ts, uint32_t packet)
{
    uint32_t di = ctx->packet_draw[packet];
    const sro_draw* d = &ctx->draws[di];
    uint32_t ntris = d->idx_num / 3u;
    uint32_t begin = ctx->packet_begin[packet];
    uint32_t end = begin + PACKET_TRIS < ntris ? begin + PACKET_TRIS : ntris;
    const uint8_t* pos = (const uint8_t*)d->pos;
    const uint8_t* attr = (const uint8_t*)d->attr;
    for (uint32_t tri = begin; tri < end; ++tri) {
        uint32_t idx[3];
        vec4 v[3];
        for (int k = 0; k < 3; ++k) {
            idx[k] = fetch_index(d, tri * 3u + (uint32_t)k);
            if (idx[k] >= d->pos_num || idx[k] >= d->attr_num) { ts->error = -3; return; }
            const float* p = (const float*)(pos + (size_t)idx[k] * d->pos_stride);
            v[k] = transform(d->mvp, p[0], p[1], p[2]);
        }
        uint32_t primbase = (ctx->tri_base[di] + tri) << 3;
        uint32_t c0 = clip_codes(v[0]), c1 = clip_codes(v[1]), c2 = clip_codes(v[2]);
        uint32_t mor = c0 | c1 | c2;
        const float* ap[3];
        for (int k = 0; k < 3; ++k) ap[k] = (const float*)(attr + (size_t)idx[k] * d->attr_stride);
        if (mor == 0) {
            bin_tri(ctx, ts, di, packet, v[0], v[1], v[2], ap[0], ap[1], ap[2], primbase | 7u);
            continue;
        }
        if (c0 & c1 & c2) continue;
        clip_buffer b;
        b.n_in = 3; b.n_out = 0; b.in_idx = 0;
        for (int k = 0; k < 3; ++k) {
            memset(b.attribs[0][k], 0, sizeof(b.attribs[0][k]));
            memcpy(b.attribs[0][k], ap[k], d->attr_stride);
            b.verts[0][k] = v[k];
        }
        ts->tris_clipped++;
        clip_polygon(&b, mor);
        vec4* iv = b.verts[b.in_idx];
        float (*ia)[SRO_MAX_VARYINGS] = b.attribs[b.in_idx];
        for (uint32_t i = 2; i < b.n_in; ++i)
            bin_tri(ctx, ts, di, packet, iv[0], iv[i - 1], iv[i], ia[0], ia[i - 1], ia[i], primbase | (i - 2));
    }
}

/* ------------------------------------------------------------------ */
/* back end: Rasterizer.cpp                                            */
/* ------------------------------------------------------------------ */

static int frag_push(thread_state* ts, uint32_t* n, uint8_t x, uint8_t y, uint32_t chunk, uint8_t tri)
{
    if (*n == ts->frag_cap) {
        uint32_t nc = ts->frag_cap ? ts->frag_cap * 2 : 4096;
        frag_t* p = (frag_t*)realloc(ts->frags, sizeof(frag_t) * nc);
        if (!p) return -1;
        ts->frags = p;
        ts->frag_cap = nc;
    }
    frag_t* f = &ts->frags[(*n)++];
    f->x = x; f->y = y; f->tri = tri; f->pad = 0; f->chunk = chunk;
    return 0;
}

/* RasterizeTrisInBin_OutputFragments + ComputeBlockMask8x8(_DepthOnly),
 * Rasterizer.cpp:97-304. */
static void raster_chunk(thread_state* ts, const chunk_t* ch, uint32_t chunk_pos, float* depth, uint32_t* nfrag)
{
    for (uint32_t t = 0; t < ch->num_tris; ++t) {
        const edge_eq* e = &ch->edge[t];
        uint32_t xbeg = e->bminx & ~7u, ybeg = e->bminy & ~7u;
        uint32_t xend = e->bmaxx, yend = e->bmaxy;
        float zdx = ch->zoverw[t].dx, zdy = ch->zoverw[t].dy, zc0 = ch->zoverw[t].c0;
        float top[8];
        for (int l = 0; l < 8; ++l) top[l] = fmaf((float)l, zdx, zc0);
        for (uint32_t yb = ybeg; yb < yend; yb += 8)
            for (uint32_t xb = xbeg; xb < xend; xb += 8) {
                uint32_t X0 = xb, X1 = xb + SRO_BIN_SIZE, Y0 = yb, Y1 = yb + SRO_BIN_SIZE;
                uint32_t m[3];
                for (int k = 0; k < 3; ++k) {
                    uint32_t c = (uint32_t)e->c[k], dx = (uint32_t)e->dx[k], dy = (uint32_t)e->dy[k];
                    m[k] = ((int32_t)(c + dy * X0 + dx * Y0) > 0) | (((int32_t)(c + dy * X0 + dx * Y1) > 0) << 1) |
                           (((int32_t)(c + dy * X1 + dx * Y0) > 0) << 2) | (((int32_t)(c + dy * X1 + dx * Y1) > 0) << 3);
                }
                if (!m[0] || !m[1] || !m[2]) continue;
                ts->blocks_tested++;
                int depth_only = m[0] == 0xF && m[1] == 0xF && m[2] == 0xF;
                float z[8];
                for (int l = 0; l < 8; ++l) z[l] = fmaf((float)yb, zdy, top[l]) + (float)xb * zdx;
                uint32_t ev[3][8];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 8; ++l)
                        ev[k][l] = ((uint32_t)e->dy[k] * (uint32_t)l + (uint32_t)e->c[k]) +
                                   ((uint32_t)yb * (uint32_t)e->dx[k] + (uint32_t)xb * (uint32_t)e->dy[k]);
                uint64_t mask = 0;
                for (uint32_t r = 0; r < 8; ++r) {
                    float* drow = depth + (yb + r) * SRO_BIN_SIZE + xb;
                    for (int l = 0; l < 8; ++l) {
                        int covered = depth_only || !((ev[0][l] | ev[1][l] | ev[2][l]) & 0x80000000u);
                        if (covered) ts->frags_raster++;
                        if (covered && z[l] > 0.0f && z[l] > drow[l]) {
                            drow[l] = z[l];
                            mask |= 1ull << (r * 8 + (uint32_t)l);
                        }
                    }
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 8; ++l) ev[k][l] += (uint32_t)e->dx[k];
                    for (int l = 0; l < 8; ++l) z[l] = z[l] + zdy;
                }
                while (mask) {
                    uint32_t bit = (uint32_t)__builtin_ctzll(mask);
                    if (frag_push(ts, nfrag, (uint8_t)(xb + (bit & 7)), (uint8_t)(yb + bit / 8), chunk_pos, (uint8_t)t)) {
                        ts->error = -2;
                        return;
                    }
                    mask &= mask - 1;
                }
            }
    }
}

/* SponzaShader (Viewer/SponzaScene.cpp:13-104) on one fragment: varyings 0-2
 * position, 3-5 normal, 6-7 uv. Dot3SoA (SIMDUtil.h:123-126) is
 * fma(x0, x1, fma(y0, y1, z0 * z1)); _mm256_max_ps(a, b) is a > b ? a : b. */
static uint32_t shade_sponza(const sro_sponza_constants* k, const sro_draw* d, const float* var,
                             const float der[4])
{
    if (!d->tex || !d->tex->texels || d->tex->size == 0) return 0xFFFFFFFFu;
    const float px = var[0], py = var[1], pz = var[2], nx = var[3], ny = var[4], nz = var[5];
    float ndl = maxps(0.1f, fmaf(nx, k->sun_dir[0], fmaf(ny, k->sun_dir[1], nz * k->sun_dir[2])));
    float rad[3] = { ndl, ndl, ndl };
    for (int i = 0; i < SRO_SPONZA_LIGHTS; ++i) {
        const sro_sponza_light* L = &k->lights[i];
        const float tx = L->pos[0] - px, ty = L->pos[1] - py, tz = L->pos[2] - pz;
        const float dsq = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
        const float rd = sro_rsqrtps(dsq);
        const float dist = sro_rcpps(rd);
        const float lx = tx * rd, ly = ty * rd, lz = tz * rd;
        const float nl = maxps(0.0f, fmaf(lx, nx, fmaf(ly, ny, lz * nz)));
        const float atten = sro_rcpps(1.0f + fmaf(0.1f, dist, dsq * 0.01f));
        const float lr = nl * (L->intensity * atten);
        for (int c = 0; c < 3; ++c) rad[c] = rad[c] + lr * L->colour[c];
    }
    for (int c = 0; c < 3; ++c) rad[c] = rad[c] + k->ambient[c];
    float rgba[4];
    sro_sample_wrap(d->tex, var[6], var[7], der[0], der[1], der[2], der[3], rgba);
    return sro_pack_rgba(rad[0] * rgba[0], rad[1] * rgba[1], rad[2] * rgba[2], rgba[3]);
}

/* Pixel shaders, Viewer/Shaders.h:71-130, on one fragment's interpolants. */
static uint32_t shade(const sro_sponza_constants* sponza, const sro_draw* d, const float* var, const float der[4])
{
    switch (d->shader) {
    case SRO_SHADER_SPONZA:
        return shade_sponza(sponza, d, var, der);
    case SRO_SHADER_VISUALIZE_NORMALS:
        return sro_pack_rgba(fmaf(var[3], 0.5f, 0.5f), fmaf(var[4], 0.5f, 0.5f), fmaf(var[5], 0.5f, 0.5f), 1.0f);
    case SRO_SHADER_VISUALIZE_UVS:
        return sro_pack_rgba(var[6], var[7], 0.0f, 0.0f);
    default: {
        if (!d->tex || !d->tex->texels || d->tex->size == 0) return 0xFFFFFFFFu;
        float rgba[4];
        sro_sample_wrap(d->tex, var[6], var[7], der[0], der[1], der[2], der[3], rgba);
        return sro_pack_rgba(rgba[0], rgba[1], rgba[2], rgba[3]);
    }
    }
}

static int cmp_chunk_ref(const void* a, const void* b)
{
    const chunk_ref* x = (const chunk_ref*)a;
    const chunk_ref* y = (const chunk_ref*)b;
    if (x->draw != y->draw) return x->draw < y->draw ? -1 : 1;
    if (x->packet != y->packet) return x->packet < y->packet ? -1 : 1;
    if (x->pos != y->pos) return x->pos < y->pos ? -1 : 1;
    return 0;
}

/* RasterAndShadeBin, Rasterizer.cpp:525-577, with ShadeFragmentBuffer
 * (:460-523) and ComputeInterpolants (:306-458) per fragment. */
static void raster_tile(sro_ctx* ctx, thread_state* ts, uint32_t bin)
{
    uint32_t p = 0;
    while (p + 1 < ctx->nplanes && ctx->tile_base[p + 1] <= bin) ++p;
    sro_plane* pl = &ctx->planes[p];
    uint32_t tile = bin - ctx->tile_base[p];

    uint32_t nref = 0;
    for (uint32_t th = 0; th < ctx->nthreads; ++th) nref += ctx->ts[th].bins[bin].n;
    if (!nref) return;
    if (nref > ts->ref_cap) {
        chunk_ref* r = (chunk_ref*)realloc(ts->refs, sizeof(chunk_ref) * nref);
        if (!r) { ts->error = -2; return; }
        ts->refs = r;
        ts->ref_cap = nref;
    }
    uint32_t n = 0;
    for (uint32_t th = 0; th < ctx->nthreads; ++th) {
        u32_vec* b = &ctx->ts[th].bins[bin];
        for (uint32_t i = 0; i < b->n; ++i) {
            chunk_t* c = chunk_at(&ctx->ts[th], b->items[i]);
            chunk_ref r = { th, b->items[i], c->draw_idx, c->packet_idx, i };
            ts->refs[n++] = r;
        }
    }
    if (ctx->nthreads > 1) qsort(ts->refs, n, sizeof(chunk_ref), cmp_chunk_ref);

    float* depth = pl->depth + (size_t)tile * 4096u;
    uint32_t* colour = pl->colour + (size_t)tile * 4096u;
    uint32_t nfrag = 0;
    for (uint32_t i = 0; i < n; ++i) {
        raster_chunk(ts, chunk_at(&ctx->ts[ts->refs[i].thread], ts->refs[i].chunk), i, depth, &nfrag);
        if (ts->error) return;
    }
    ts->frags_passed += nfrag;

    for (uint32_t f = 0; f < nfrag; ++f) {
        const frag_t* fr = &ts->frags[f];
        const chunk_t* ch = chunk_at(&ctx->ts[ts->refs[fr->chunk].thread], ts->refs[fr->chunk].chunk);
        const sro_draw* d = &ctx->draws[ch->draw_idx];
        uint32_t t = fr->tri, apt = ch->attribs_per_tri;
        const float* adx = &ch->adx[apt * t];
        const float* ady = &ch->ady[apt * t];
        const float* ac = &ch->ac[apt * t];
        float wdx = ch->recipw[t].dx, wdy = ch->recipw[t].dy, wc = ch->recipw[t].c0;
        float X = (float)fr->x, Y = (float)fr->y;
        float iq = 1.0f / fmaf(X, wdx, fmaf(Y, wdy, wc));
        float var[SRO_MAX_VARYINGS] = { 0 };
        for (uint32_t k = 0; k < apt; ++k) var[k] = iq * fmaf(ady[k], Y, fmaf(adx[k], X, ac[k]));
        float X1 = 1.0f + X, Y1 = 1.0f + Y;
        float i10 = sro_rcpps(fmaf(wdx, X1, fmaf(wdy, Y, wc)));
        float i01 = sro_rcpps(fmaf(wdx, X, fmaf(wdy, Y1, wc)));
        float der[4];
        for (uint32_t uv = 0; uv < 2; ++uv) {
            uint32_t k = d->uv_offset + uv;
            float t10 = i10 * fmaf(adx[k], X1, fmaf(ady[k], Y, ac[k]));
            float t01 = i01 * fmaf(adx[k], X, fmaf(ady[k], Y1, ac[k]));
            der[2 * uv] = t10 - var[k];
            der[2 * uv + 1] = t01 - var[k];
        }
        uint32_t px = (uint32_t)fr->y * SRO_BIN_SIZE + fr->x;
        colour[px] = shade(&ctx->sponza, d, var, der);
        if (pl->primid) pl->primid[(size_t)tile * 4096u + px] = ch->primid[t];
    }
}

/* ------------------------------------------------------------------ */
/* thread pool (TaskSystem stand-in)                                   */
/* ------------------------------------------------------------------ */

static void job_frontend(sro_ctx* ctx, uint32_t tid)
{
    thread_state* ts = &ctx->ts[tid];
    for (;;) {
        uint32_t p = atomic_fetch_add(&ctx->next_packet, 1u);
        if (p >= ctx->npackets || ts->error) break;
        bin_packet(ctx, ts, p);
    }
}

static void job_backend(sro_ctx* ctx, uint32_t tid)
{
    thread_state* ts = &ctx->ts[tid];
    for (;;) {
        uint32_t i = atomic_fetch_add(&ctx->next_tile, 1u);
        if (i >= ctx->ntile_list || ts->error) break;
        raster_tile(ctx, ts, ctx->tile_list[i]);
    }
}

typedef struct { sro_ctx* ctx; uint32_t tid; } worker_arg;

static void* worker_main(void* arg)
{
    worker_arg* wa = (worker_arg*)arg;
    sro_ctx* ctx = wa->ctx;
    uint32_t tid = wa->tid;
    free(wa);
    uint64_t seen = 0;
    pthread_mutex_lock(&ctx->mu);
    for (;;) {
        while (!ctx->quit && ctx->generation == seen) pthread_cond_wait(&ctx->cv_start, &ctx->mu);
        if (ctx->quit) break;
        seen = ctx->generation;
        void (*job)(sro_ctx*, uint32_t) = ctx->job;
        pthread_mutex_unlock(&ctx->mu);
        job(ctx, tid);
        pthread_mutex_lock(&ctx->mu);
        if (--ctx->running == 0) pthread_cond_signal(&ctx->cv_done);
    }
    pthread_mutex_unlock(&ctx->mu);
    return NULL;
}

static void run_parallel(sro_ctx* ctx, void (*job)(sro_ctx*, uint32_t))
{
    if (ctx->nthreads > 1) {
        pthread_mutex_lock(&ctx->mu);
        ctx->job = job;
        ctx->running = ctx->nthreads - 1;
        ctx->generation++;
        pthread_cond_broadcast(&ctx->cv_start);
        pthread_mutex_unlock(&ctx->mu);
    }
    job(ctx, 0);
    if (ctx->nthreads > 1) {
        pthread_mutex_lock(&ctx->mu);
        while (ctx->running) pthread_cond_wait(&ctx->cv_done, &ctx->mu);
        pthread_mutex_unlock(&ctx->mu);
    }
}

sro_ctx* sro_create(uint32_t nthreads)
{
    if (nthreads == 0) nthreads = 1;
    sro_ctx* ctx = (sro_ctx*)calloc(1, sizeof(sro_ctx));
    if (!ctx) return NULL;
    ctx->nthreads = nthreads;
    ctx->ts = (thread_state*)calloc(nthreads, sizeof(thread_state));
    pthread_mutex_init(&ctx->mu, NULL);
    pthread_cond_init(&ctx->cv_start, NULL);
    pthread_cond_init(&ctx->cv_done, NULL);
    if (nthreads > 1) {
        ctx->threads = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
        for (uint32_t i = 1; i < nthreads; ++i) {
            worker_arg* wa = (worker_arg*)malloc(sizeof(worker_arg));
            wa->ctx = ctx;
            wa->tid = i;
            pthread_create(&ctx->threads[i], NULL, worker_main, wa);
        }
    }
    return ctx;
}

void sro_destroy(sro_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->nthreads > 1) {
        pthread_mutex_lock(&ctx->mu);
        ctx->quit = 1;
        pthread_cond_broadcast(&ctx->cv_start);
        pthread_mutex_unlock(&ctx->mu);
        for (uint32_t i = 1; i < ctx->nthreads; ++i) pthread_join(ctx->threads[i], NULL);
        free(ctx->threads);
    }
    for (uint32_t i = 0; i < ctx->nthreads; ++i) {
        thread_state* ts = &ctx->ts[i];
        for (uint32_t b = 0; b < ts->nblocks; ++b) free(ts->blocks[b]);
        free(ts->blocks);
        for (uint32_t b = 0; b < ts->nbins; ++b) free(ts->bins[b].items);
        free(ts->bins);
        free(ts->frags);
        free(ts->refs);
    }
    free(ctx->ts);
    free(ctx->tile_base);
    free(ctx->tri_base);
    free(ctx->packet_draw);
    free(ctx->packet_begin);
    free(ctx->tile_list);
    pthread_mutex_destroy(&ctx->mu);
    pthread_cond_destroy(&ctx->cv_start);
    pthread_cond_destroy(&ctx->cv_done);
    free(ctx);
}

/* ------------------------------------------------------------------ */
/* frame API                                                           */
/* ------------------------------------------------------------------ */

void sro_plane_dims(uint32_t width, uint32_t height, uint32_t* tiles_x, uint32_t* tiles_y)
{
    *tiles_x = (width + SRO_BIN_SIZE - 1) / SRO_BIN_SIZE;
    *tiles_y = (height + SRO_BIN_SIZE - 1) / SRO_BIN_SIZE;
}

void sro_clear(sro_plane* pl, uint32_t colour, int clear_colour, int clear_depth)
{
    size_t n = (size_t)pl->tiles_x * pl->tiles_y * 4096u;
    if (clear_depth)
        for (size_t i = 0; i < n; ++i) pl->depth[i] = 0.0f;
    if (clear_colour) memset(pl->colour, (int)(uint8_t)colour, n * 4u);
}

void sro_set_sponza_constants(sro_ctx* ctx, const sro_sponza_constants* k)
{
    if (ctx && k) ctx->sponza = *k;
}

void sro_blit(const sro_plane* pl, uint32_t* linear)
{
    for (uint32_t ty = 0; ty < pl->tiles_y; ++ty)
        for (uint32_t tx = 0; tx < pl->tiles_x; ++tx) {
            const uint32_t* tile = pl->colour + (size_t)(ty * pl->tiles_x + tx) * 4096u;
            uint32_t yend = pl->height - ty * 64u < 64u ? pl->height - ty * 64u : 64u;
            uint32_t w = pl->width - tx * 64u < 64u ? pl->width - tx * 64u : 64u;
            for (uint32_t y = 0; y < yend; ++y)
                memcpy(linear + (size_t)(ty * 64u + y) * pl->width + tx * 64u, tile + y * 64u, 4u * w);
        }
}

static int validate_draw(const sro_draw* d, uint32_t nplanes)
{
    if (!d->idx || (d->idx_stride != 1 && d->idx_stride != 2 && d->idx_stride != 4)) return -1;
    if (!d->pos || d->pos_stride < 12) return -1;
    if (!d->attr || d->attr_stride == 0 || d->attr_stride > 32 || (d->attr_stride & 3)) return -1;
    if (d->plane >= nplanes) return -1;
    uint32_t apt = d->attr_stride / 4u;
    switch (d->shader) {
    case SRO_SHADER_UNLIT_DIFFUSE: if (apt < 8 || d->uv_offset + 1 >= apt) return -1; break;
    case SRO_SHADER_VISUALIZE_NORMALS: if (apt < 6 || d->uv_offset + 1 >= apt) return -1; break;
    case SRO_SHADER_VISUALIZE_UVS: if (apt < 8 || d->uv_offset + 1 >= apt) return -1; break;
    case SRO_SHADER_SPONZA: if (apt != 8 || d->uv_offset + 1 >= apt) return -1; break;
    default: return -1;
    }
    return 0;
}

int sro_render(sro_ctx* ctx, const sro_draw* draws, uint32_t ndraws, sro_plane* planes, uint32_t nplanes,
               sro_stats* stats, uint32_t* tile_entries)
{
    for (uint32_t i = 0; i < ndraws; ++i)
        if (validate_draw(&draws[i], nplanes)) return -1;
    ctx->draws = draws;
    ctx->ndraws = ndraws;
    ctx->planes = planes;
    ctx->nplanes = nplanes;

    free(ctx->tile_base);
    ctx->tile_base = (uint32_t*)malloc(sizeof(uint32_t) * (nplanes + 1));
    uint32_t total = 0;
    for (uint32_t p = 0; p < nplanes; ++p) {
        sro_plane_dims(planes[p].width, planes[p].height, &planes[p].tiles_x, &planes[p].tiles_y);
        ctx->tile_base[p] = total;
        total += planes[p].tiles_x * planes[p].tiles_y;
        if (planes[p].primid)
            memset(planes[p].primid, 0xFF, sizeof(uint32_t) * planes[p].tiles_x * planes[p].tiles_y * 4096u);
    }
    ctx->tile_base[nplanes] = total;
    ctx->total_tiles = total;

    free(ctx->tri_base);
    ctx->tri_base = (uint32_t*)malloc(sizeof(uint32_t) * (ndraws + 1));
    uint32_t npk = 0, tb = 0;
    for (uint32_t i = 0; i < ndraws; ++i) {
        ctx->tri_base[i] = tb;
        uint32_t nt = draws[i].idx_num / 3u;
        tb += nt;
        npk += (nt + PACKET_TRIS - 1) / PACKET_TRIS;
    }
    free(ctx->packet_draw);
    free(ctx->packet_begin);
    ctx->packet_draw = (uint32_t*)malloc(sizeof(uint32_t) * (npk + 1));
    ctx->packet_begin = (uint32_t*)malloc(sizeof(uint32_t) * (npk + 1));
    npk = 0;
    for (uint32_t i = 0; i < ndraws; ++i) {
        uint32_t nt = draws[i].idx_num / 3u;
        for (uint32_t b = 0; b < nt; b += PACKET_TRIS) {
            ctx->packet_draw[npk] = i;
            ctx->packet_begin[npk] = b;
            ++npk;
        }
    }
    ctx->npackets = npk;

    for (uint32_t th = 0; th < ctx->nthreads; ++th) {
        thread_state* ts = &ctx->ts[th];
        if (ts->nbins < total) {
            u32_vec* nb = (u32_vec*)realloc(ts->bins, sizeof(u32_vec) * total);
            if (!nb) return -2;
            memset(nb + ts->nbins, 0, sizeof(u32_vec) * (total - ts->nbins));
            ts->bins = nb;
            ts->nbins = total;
        }
        for (uint32_t b = 0; b < ts->nbins; ++b) ts->bins[b].n = 0;
        ts->nchunks = 0;
        ts->tris_clipped = ts->prims_binned = ts->bin_entries = ts->frags_raster = ts->frags_passed = 0;
        ts->blocks_tested = 0;
        ts->error = 0;
    }

    atomic_store(&ctx->next_packet, 0u);
    run_parallel(ctx, job_frontend);
    for (uint32_t th = 0; th < ctx->nthreads; ++th)
        if (ctx->ts[th].error) return ctx->ts[th].error;

    free(ctx->tile_list);
    ctx->tile_list = (uint32_t*)malloc(sizeof(uint32_t) * (total + 1));
    ctx->ntile_list = 0;
    for (uint32_t b = 0; b < total; ++b) {
        uint32_t any = 0, entries = 0;
        for (uint32_t th = 0; th < ctx->nthreads; ++th) {
            u32_vec* v = &ctx->ts[th].bins[b];
            any |= v->n;
            for (uint32_t i = 0; i < v->n; ++i) entries += chunk_at(&ctx->ts[th], v->items[i])->num_tris;
        }
        if (tile_entries) tile_entries[b] = entries;
        if (any) ctx->tile_list[ctx->ntile_list++] = b;
    }
    atomic_store(&ctx->next_tile, 0u);
    run_parallel(ctx, job_backend);

    sro_stats st = { 0 };
    for (uint32_t i = 0; i < ndraws; ++i) st.tris_in += draws[i].idx_num / 3u;
    for (uint32_t th = 0; th < ctx->nthreads; ++th) {
        thread_state* ts = &ctx->ts[th];
        if (ts->error) return ts->error;
        st.tris_clipped += ts->tris_clipped;
        st.prims_binned += ts->prims_binned;
        st.bin_entries += ts->bin_entries;
        st.frags_raster += ts->frags_raster;
        st.frags_passed += ts->frags_passed;
        st.blocks_tested += ts->blocks_tested;
    }
    if (stats) *stats = st;
    return 0;
}

/* ------------------------------------------------------------------ */
/* mip generation: stb_image_resize v0.96 as Texture.cpp:196 calls it  */
/* ------------------------------------------------------------------ */

/* stbir__filter_mitchell, stb_image_resize.h:824-836: each operation in float. */
static float stbir_mitchell(float x)
{
    x = fabsf(x);
    if (x < 1.0f) {
        const float t = 21.0f * x - 36.0f;
        return (16.0f + (x * x) * t) / 18.0f;
    }
    if (x < 2.0f) {
        float t = 36.0f - 7.0f * x;
        t = -60.0f + x * t;
        return (32.0f + x * t) / 18.0f;
    }
    return 0.0f;
}

/* One axis of a downsampling resize: for every input position n of the
 * margin-extended range [-margin, in + margin) the output range [n0, n1] it
 * scatters into and its weights (stb_image_resize.h:1023-1035, 1087-1115,
 * 1117-1192, 1194-1232 for scale <= 1, Mitchell support 2). */
typedef struct {
    int margin, count;
    int* n0;
    int* n1;
    float* w; /* 4 per contributor */
} stbir_axis;

static int stbir_axis_build(stbir_axis* a, uint32_t in, uint32_t out)
{
    const float scale = ((float)out / (float)in) / 1.0f;
    const float radius = 2.0f / scale;
    const int width = (int)ceil((double)(4.0f / scale));
    a->margin = width / 2;
    a->count = (int)in + 2 * a->margin;
    a->n0 = (int*)malloc(sizeof(int) * a->count);
    a->n1 = (int*)malloc(sizeof(int) * a->count);
    a->w = (float*)calloc((size_t)a->count * 4 + 4, sizeof(float));
    if (!a->n0 || !a->n1 || !a->w) return -1;
    for (int n = 0; n < a->count; ++n) {
        const float c = (float)(n - a->margin) + 0.5f;
        const float lo = (c - radius) * scale - 0.0f;
        const float hi = (c + radius) * scale - 0.0f;
        const float oc = c * scale - 0.0f;
        const int first = (int)floor((double)lo + 0.5);
        const int last = (int)floor((double)hi - 0.5);
        if (last - first > 4) return -1;
        a->n0[n] = first;
        a->n1[n] = last;
        float* w = a->w + (size_t)n * 4;
        /* a fifth weight (last - first == 4) lands in the next group and is
         * overwritten by it, as in the reference's packed coefficient array */
        for (int i = 0; i <= last - first; ++i) w[i] = stbir_mitchell(((float)(first + i) + 0.5f) - oc) * scale;
        for (int i = last - first; i >= 0; --i) {
            if (w[i] != 0.0f) break;
            a->n1[n] = first + i - 1;
        }
    }
    /* normalise, per output pixel, over every contributor that reaches it */
    for (int o = 0; o < (int)out; ++o) {
        float total = 0.0f;
        for (int n = 0; n < a->count && o >= a->n0[n]; ++n)
            if (o <= a->n1[n]) total += a->w[(size_t)n * 4 + (o - a->n0[n])];
        const float k = 1.0f / total;
        for (int n = 0; n < a->count && o >= a->n0[n]; ++n)
            if (o <= a->n1[n]) a->w[(size_t)n * 4 + (o - a->n0[n])] *= k;
    }
    /* drop leading zero weights and outputs before 0; clamp to the last output */
    for (int n = 0; n < a->count; ++n) {
        float* w = a->w + (size_t)n * 4;
        int skip = 0;
        while (skip < 4 && w[skip] == 0.0f) ++skip;
        a->n0[n] += skip;
        while (a->n0[n] < 0) {
            a->n0[n]++;
            skip++;
        }
        const int range = a->n1[n] - a->n0[n] + 1;
        const int m = range < 4 ? range : 4;
        for (int i = 0; i < m && i + skip < 4; ++i) w[i] = w[i + skip];
        if (a->n1[n] > (int)out - 1) a->n1[n] = (int)out - 1;
    }
    return 0;
}

static void stbir_axis_free(stbir_axis* a)
{
    free(a->n0);
    free(a->n1);
    free(a->w);
}

int sro_mip_resize_rgba8(const uint8_t* in, uint32_t w, uint32_t h, uint8_t* out, uint32_t ow, uint32_t oh)
{
    if (!in || !out || w == 0 || h == 0 || ow == 0 || oh == 0 || ow > w || oh > h) return -1;
    stbir_axis ax = { 0 }, ay = { 0 };
    float* hrow = (float*)malloc(sizeof(float) * ow * 4);
    float* acc = (float*)calloc((size_t)ow * oh * 4, sizeof(float));
    int rc = (hrow && acc) ? 0 : -1;
    if (!rc) rc = stbir_axis_build(&ax, w, ow);
    if (!rc) rc = stbir_axis_build(&ay, h, oh);
    /* rows in order (stb_image_resize.h:2166-2205): decode with clamped
     * edges, scatter horizontally into one row, then scatter that row into
     * the output rows it reaches */
    for (int yc = 0; !rc && yc < ay.count; ++yc) {
        if (ay.n0[yc] > ay.n1[yc]) continue;
        const uint8_t* src = in + (size_t)clampi(yc - ay.margin, 0, (int)h - 1) * w * 4;
        memset(hrow, 0, sizeof(float) * ow * 4);
        for (int xc = 0; xc < ax.count; ++xc) {
            const uint8_t* p = src + (size_t)clampi(xc - ax.margin, 0, (int)w - 1) * 4;
            float d[4];
            for (int c = 0; c < 4; ++c) d[c] = (float)p[c] / 255.0f;
            for (int k = ax.n0[xc]; k <= ax.n1[xc]; ++k) {
                const float cw = ax.w[(size_t)xc * 4 + (k - ax.n0[xc])];
                for (int c = 0; c < 4; ++c) hrow[k * 4 + c] += d[c] * cw;
            }
        }
        for (int k = ay.n0[yc]; k <= ay.n1[yc]; ++k) {
            const float cw = ay.w[(size_t)yc * 4 + (k - ay.n0[yc])];
            float* dst = acc + (size_t)k * ow * 4;
            for (uint32_t i = 0; i < ow * 4; ++i) dst[i] += hrow[i] * cw;
        }
    }
    /* encode, stb_image_resize.h:1736-1757 (STBIR__SATURATE_INT undefined) */
    for (size_t i = 0; !rc && i < (size_t)ow * oh * 4; ++i) {
        float f = acc[i];
        if (f < 0.0f) f = 0.0f;
        if (f > 1.0f) f = 1.0f;
        out[i] = (uint8_t)(int)((double)(f * 255.0f) + 0.5);
    }
    stbir_axis_free(&ax);
    stbir_axis_free(&ay);
    free(hrow);
    free(acc);
    return rc;
}
