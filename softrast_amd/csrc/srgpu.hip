// srgpu.hip -- MI355X (gfx950) sort-middle rasteriser behind the C ABI of
// include/srgpu.h.
//
// Frame = five launches on one stream:
//   k_frame_init  clear counters and per-tile arrays (fill kernel, no memset)
//   k_setup       one lane per submitted triangle: fetch, MVP, clip, snap,
//                 cull, edge/plane setup (Binning.cpp:464-535, 279-350);
//                 writes primitive records and counts tile entries (LDS
//                 histogram per workgroup, one global add per touched tile)
//   k_scan        one workgroup: exclusive scan of tile counts, tile queue
//   k_emit        one lane per record: re-walks its bins (Binning.cpp:352-455)
//                 and appends the record to each tile list
//   k_backend     persistent: a workgroup owns one 64x64 tile at a time with a
//                 64-bit visibility key per pixel in LDS; waves take a tile
//                 entry each and evaluate an 8x8 block per wave (one pixel per
//                 lane, Rasterizer.cpp:97-304); then each lane resolves and
//                 shades its pixels (Rasterizer.cpp:306-523, Texture.cpp,
//                 Shaders.h) and writes RGBA8 + depth once.
//
// Visibility: the reference rasterises a tile's triangles in submission order
// with a strict reverse-Z test and shades every passing fragment, so the
// final pixel is the earliest-submitted fragment among those with the
// largest depth. The back end computes exactly that, order-independently, as
// max over (depth bits << 32 | (0xFFFFFFFE - primitive order)), which lets
// tile entries arrive in any order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "srgpu.h"
#include "sr_device.h"

namespace srg {

// ------------------------------------------------------------------------
// device tables
// ------------------------------------------------------------------------

struct DrawDev {
    const uint8_t* idx;
    const uint8_t* pos;
    const uint8_t* attr;
    uint32_t idx_stride, pos_stride, attr_stride, ntris;
    uint32_t pos_num, attr_num, tri_base, view;
    uint32_t shader, tex, nslots, deriv_slot;  // deriv_slot: slot of attribute uv_offset
    uint32_t slot_attr[kMaxSlots];
    float mvp[16];
};

enum : uint32_t {
    VF_CLEAR_COLOUR = 1u,
    VF_CLEAR_DEPTH = 2u,
    VF_HAS_COLOUR = 4u,
};

struct ViewDev {
    uint32_t* colour;
    float* depth;
    uint32_t* primid;
    uint32_t width, height, tiles_x, tiles_y;
    uint32_t tile_base, flags, clear_colour, pad;
};

enum : uint32_t {
    C_REC = 0,         // records allocated
    C_FLAGS = 1,       // error / overflow bits
    C_QLEN = 2,        // tile queue length
    C_QHEAD = 3,       // tile queue head (persistent back end)
    C_FRAGS = 4,       // 64-bit: rasterised fragments (words 4,5)
    C_ENTRIES = 6,     // tile entries in this frame
    C_NCOUNTERS = 16,
};

enum : uint32_t {
    FLAG_REC_OVERFLOW = 1u,
    FLAG_ENTRY_OVERFLOW = 2u,
    FLAG_INDEX_RANGE = 4u,
};

enum : uint32_t { FF_STATS = 1u, FF_PRIMID = 2u, FF_TILE_COSTS = 4u };

struct FrameDev {
    const DrawDev* draws;
    const ViewDev* views;
    const TexDev* texs;
    const uint32_t* block_draw;
    const uint32_t* block_tri0;
    uint32_t ndraws, nviews, total_tiles, total_tris;
    uint4* rec_raster;   // 4 x uint4 per record
    float4* rec_shade;   // 4 x float4 per record
    uint32_t* rec_order; // primitive order key ((global tri << 3) | fan)
    uint32_t* rec_meta;  // draw index, 0xFFFFFFFF = culled fan
    uint32_t* tri_rec;   // per submitted triangle: its first record
    uint32_t rec_cap, entry_cap;
    uint32_t* tile_count;
    uint32_t* tile_start;  // total_tiles + 1
    uint32_t* tile_cursor;
    uint32_t* entries;
    uint32_t* queue;
    uint64_t* tile_cost;   // optional per global tile
    uint32_t* counters;
    uint32_t flags;
    uint32_t lds_hist_tiles;
};

// ------------------------------------------------------------------------
// front end
// ------------------------------------------------------------------------

struct Prim {
    int32_t c[3], dx[3], dy[3];
    int32_t xmin, xmax, ymin, ymax;
    float r0x, r0y, zdx, zdy, zf0, wdx, wdy, wf0;
    float adx[kMaxSlots], ady[kMaxSlots], af0[kMaxSlots];
};

// kt::Mat4 * kt::Vec4 (Binning.cpp:477-479) with the order
// ((c0*x + c1*y) + c2*z) + c3*w, w = 1, every op rounded.
__device__ __forceinline__ float4 transform(const float* m, float x, float y, float z)
{
    const float w = 1.0f;
    float4 r;
    r.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
    r.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
    r.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
    r.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
    return r;
}

// ComputeClipMask (Binning.cpp:56-68).
__device__ __forceinline__ uint32_t clip_codes(float4 v)
{
    uint32_t m = 0;
    m |= (v.x + v.w < 0.0f) ? 0x01u : 0u;
    m |= (v.x - v.w > 0.0f) ? 0x02u : 0u;
    m |= (v.y + v.w < 0.0f) ? 0x04u : 0u;
    m |= (v.y - v.w > 0.0f) ? 0x08u : 0u;
    m |= (v.z < 0.0f) ? 0x10u : 0u;
    m |= (v.z - v.w > 0.0f) ? 0x20u : 0u;
    return m;
}

// SetupEdge (Binning.cpp:242-259).
__device__ __forceinline__ void setup_edge(Prim& s, int k, int32_t ax, int32_t ay, int32_t bx, int32_t by)
{
    const int32_t dy = (int32_t)((uint32_t)by - (uint32_t)ay);
    const int32_t dx = (int32_t)((uint32_t)ax - (uint32_t)bx);
    int64_t c = (int64_t)ay * (int64_t)(int32_t)((uint32_t)bx - (uint32_t)ax) -
                (int64_t)ax * (int64_t)(int32_t)((uint32_t)by - (uint32_t)ay);
    if (dy < 0 || (dy == 0 && dx > 0)) c += 256;
    s.dy[k] = dy;
    s.dx[k] = dx;
    s.c[k] = (int32_t)(uint32_t)(uint64_t)(c >> 8);
}

// SetupPlane (Binning.cpp:261-277).
__device__ __forceinline__ void setup_plane(float cterm, float d10x, float d10y, float d20x, float d20y, float f10,
                                            float f20, float& odx, float& ody)
{
    const float A = d10y * f20 - f10 * d20y;
    const float B = d20x * f10 - d10x * f20;
    odx = -A / cterm;
    ody = -B / cterm;
}

__device__ __forceinline__ int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Viewport, snap, cull, bbox, edges, planes (Binning.cpp:291-350).
// a[k][slot] are the vertex attributes of the kept slots. Returns false if culled.
__device__ __forceinline__ bool setup_prim(Prim& s, const float4 v[3], const float a[3][kMaxSlots], uint32_t nslots,
                                           uint32_t width, uint32_t height)
{
    const float hw = (float)width * 0.5f, hh = (float)height * 0.5f;
    const float iw0 = 1.0f / v[0].w, iw1 = 1.0f / v[1].w, iw2 = 1.0f / v[2].w;
    const float r0x = iw0 * v[0].x * hw + hw, r0y = iw0 * v[0].y * -hh + hh;
    const float r1x = iw1 * v[1].x * hw + hw, r1y = iw1 * v[1].y * -hh + hh;
    const float r2x = iw2 * v[2].x * hw + hw, r2y = iw2 * v[2].y * -hh + hh;
    const int32_t f0x = cvtt_i32(r0x * 256.0f + 0.5f), f0y = cvtt_i32(r0y * 256.0f + 0.5f);
    const int32_t f1x = cvtt_i32(r1x * 256.0f + 0.5f), f1y = cvtt_i32(r1y * 256.0f + 0.5f);
    const int32_t f2x = cvtt_i32(r2x * 256.0f + 0.5f), f2y = cvtt_i32(r2y * 256.0f + 0.5f);

    int64_t area = (int64_t)(int32_t)((uint32_t)f2x - (uint32_t)f0x) * (int64_t)(int32_t)((uint32_t)f1y - (uint32_t)f0y) -
                   (int64_t)(int32_t)((uint32_t)f2y - (uint32_t)f0y) * (int64_t)(int32_t)((uint32_t)f1x - (uint32_t)f0x);
    area >>= 8;
    if (area <= 0) return false;

    const int32_t mnx = min(min(f0x, f1x), f2x), mny = min(min(f0y, f1y), f2y);
    const int32_t mxx = max(max(f0x, f1x), f2x), mxy = max(max(f0y, f1y), f2y);
    s.xmin = (uint16_t)clampi((int32_t)((uint32_t)mnx + 255u) >> 8, 0, (int32_t)width - 1);
    s.ymin = (uint16_t)clampi((int32_t)((uint32_t)mny + 255u) >> 8, 0, (int32_t)height - 1);
    s.xmax = (uint16_t)clampi((int32_t)((uint32_t)mxx + 255u) >> 8, 0, (int32_t)width - 1);
    s.ymax = (uint16_t)clampi((int32_t)((uint32_t)mxy + 255u) >> 8, 0, (int32_t)height - 1);

    setup_edge(s, 0, f0x, f0y, f1x, f1y);
    setup_edge(s, 1, f1x, f1y, f2x, f2y);
    setup_edge(s, 2, f2x, f2y, f0x, f0y);

    const float d10x = r1x - r0x, d10y = r1y - r0y, d20x = r2x - r0x, d20y = r2y - r0y;
    const float cterm = d10x * d20y - d10y * d20x;
    setup_plane(cterm, d10x, d10y, d20x, d20y, v[1].z * iw1 - v[0].z * iw0, v[2].z * iw2 - v[0].z * iw0, s.zdx, s.zdy);
    setup_plane(cterm, d10x, d10y, d20x, d20y, iw1 - iw0, iw2 - iw0, s.wdx, s.wdy);
    for (uint32_t k = 0; k < nslots; ++k)
        setup_plane(cterm, d10x, d10y, d20x, d20y, a[1][k] * iw1 - a[0][k] * iw0, a[2][k] * iw2 - a[0][k] * iw0,
                    s.adx[k], s.ady[k]);
    s.r0x = r0x;
    s.r0y = r0y;
    s.zf0 = v[0].z * iw0;
    s.wf0 = iw0;
    for (uint32_t k = 0; k < nslots; ++k) s.af0[k] = a[0][k] * iw0;
    return true;
}

// Bin-level corner test (Binning.cpp:376-409): false if some edge has no
// corner of the 64x64 bin strictly inside.
__device__ __forceinline__ bool bin_accepts(const int32_t c[3], const int32_t dx[3], const int32_t dy[3], uint32_t X0,
                                            uint32_t Y0)
{
    const uint32_t X1 = X0 + kBin, Y1 = Y0 + kBin;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t cc = (uint32_t)c[k], ex = (uint32_t)dx[k], ey = (uint32_t)dy[k];
        const bool any = ((int32_t)(cc + ey * X0 + ex * Y0) > 0) | ((int32_t)(cc + ey * X0 + ex * Y1) > 0) |
                         ((int32_t)(cc + ey * X1 + ex * Y0) > 0) | ((int32_t)(cc + ey * X1 + ex * Y1) > 0);
        ok = ok && any;
    }
    return ok;
}

// Bin walk of BinTransformedAndClippedTri (Binning.cpp:352-411), including
// the numXbins-from-Y quirk of :359.
template <typename Fn>
__device__ __forceinline__ void for_each_bin(const int32_t c[3], const int32_t dx[3], const int32_t dy[3], int32_t xmin,
                                             int32_t xmax, int32_t ymin, int32_t ymax, Fn&& fn)
{
    const uint32_t bymin = (uint32_t)ymin >> 6, bymax = (uint32_t)ymax >> 6;
    const uint32_t bxmin = (uint32_t)xmin >> 6, bxmax = (uint32_t)xmax >> 6;
    const uint32_t nyb = bymax - bymin + 1;
    const bool check = !(nyb <= 2);  // numXbins == numYbins in the reference
    for (uint32_t by = bymin; by <= bymax; ++by)
        for (uint32_t bx = bxmin; bx <= bxmax; ++bx)
            if (!check || bin_accepts(c, dx, dy, bx * kBin, by * kBin)) fn(bx, by);
}

__device__ __forceinline__ void write_record(const FrameDev& F, uint32_t r, const Prim& s, uint32_t order, uint32_t draw)
{
    uint4* R = F.rec_raster + (size_t)r * 4;
    R[0] = make_uint4((uint32_t)s.c[0], (uint32_t)s.c[1], (uint32_t)s.c[2], (uint32_t)s.dx[0]);
    R[1] = make_uint4((uint32_t)s.dx[1], (uint32_t)s.dx[2], (uint32_t)s.dy[0], (uint32_t)s.dy[1]);
    R[2] = make_uint4((uint32_t)s.dy[2], (uint32_t)s.xmin | ((uint32_t)s.xmax << 16),
                      (uint32_t)s.ymin | ((uint32_t)s.ymax << 16), fbits(s.r0x));
    R[3] = make_uint4(fbits(s.r0y), fbits(s.zdx), fbits(s.zdy), fbits(s.zf0));
    float4* S = F.rec_shade + (size_t)r * 4;
    S[0] = make_float4(s.wdx, s.wdy, s.wf0, s.adx[0]);
    S[1] = make_float4(s.ady[0], s.af0[0], s.adx[1], s.ady[1]);
    S[2] = make_float4(s.af0[1], s.adx[2], s.ady[2], s.af0[2]);
    S[3] = make_float4(s.adx[3], s.ady[3], s.af0[3], 0.0f);
    F.rec_order[r] = order;
    F.rec_meta[r] = draw;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t& total)
{
    const uint32_t lane = __lane_id();
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

__device__ __forceinline__ void count_bins(const FrameDev& F, uint32_t* hist, bool use_lds, const ViewDev& V,
                                           const Prim& s)
{
    for_each_bin(s.c, s.dx, s.dy, s.xmin, s.xmax, s.ymin, s.ymax, [&](uint32_t bx, uint32_t by) {
        const uint32_t lt = by * V.tiles_x + bx;
        if (use_lds) atomicAdd(&hist[lt], 1u);
        else atomicAdd(&F.tile_count[V.tile_base + lt], 1u);
    });
}

constexpr uint32_t kSetupThreads = 256;
constexpr uint32_t kLdsHistTiles = 8192;

// Sutherland-Hodgman clip polygon (Binning.cpp:70-165): positions plus the
// kept attribute slots (the other attributes never reach the output).
struct ClipPoly {
    float4 v[2][9];
    float a[2][9][kMaxSlots];
};

// kt::Dot / kt::Lerp with the op order the oracle defines.
__device__ __forceinline__ float clip_dot(int plane, float4 v)
{
    const float px = plane == 0 ? 1.0f : (plane == 1 ? -1.0f : 0.0f);
    const float py = plane == 2 ? 1.0f : (plane == 3 ? -1.0f : 0.0f);
    const float pz = plane == 4 ? 1.0f : (plane == 5 ? -1.0f : 0.0f);
    const float pw = 1.0f;
    return ((px * v.x + py * v.y) + pz * v.z) + pw * v.w;
}
__device__ __forceinline__ float lerp_kt(float a, float b, float t) { return a + (b - a) * t; }

__global__ __launch_bounds__(kSetupThreads) void k_setup(FrameDev F)
{
    __shared__ uint32_t hist[kLdsHistTiles];
    const uint32_t d = F.block_draw[blockIdx.x];
    const DrawDev& D = F.draws[d];
    const ViewDev& V = F.views[D.view];
    const uint32_t ntile_v = V.tiles_x * V.tiles_y;
    const bool use_lds = ntile_v <= F.lds_hist_tiles;
    if (use_lds)
        for (uint32_t i = threadIdx.x; i < ntile_v; i += kSetupThreads) hist[i] = 0;
    __syncthreads();

    const uint32_t tri = F.block_tri0[blockIdx.x] + threadIdx.x;
    const bool valid = tri < D.ntris;
    const uint32_t nslots = D.nslots;

    float4 v[3];
    float a[3][kMaxSlots];
    uint32_t mask_or = 0, mask_and = 0;
    bool ok = false;
    if (valid) {
        uint32_t idx[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t i = (size_t)tri * 3u + (size_t)k;
            idx[k] = D.idx_stride == 2 ? ((const uint16_t*)D.idx)[i]
                   : D.idx_stride == 4 ? ((const uint32_t*)D.idx)[i]
                                       : (uint32_t)D.idx[i];
        }
        ok = idx[0] < D.pos_num && idx[1] < D.pos_num && idx[2] < D.pos_num && idx[0] < D.attr_num &&
             idx[1] < D.attr_num && idx[2] < D.attr_num;
        if (!ok) {
            atomicOr(&F.counters[C_FLAGS], FLAG_INDEX_RANGE);
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float* p = (const float*)(D.pos + (size_t)idx[k] * D.pos_stride);
                v[k] = transform(D.mvp, p[0], p[1], p[2]);
                const float* at = (const float*)(D.attr + (size_t)idx[k] * D.attr_stride);
                for (uint32_t s = 0; s < kMaxSlots; ++s) a[k][s] = s < nslots ? at[D.slot_attr[s]] : 0.0f;
            }
            const uint32_t c0 = clip_codes(v[0]), c1 = clip_codes(v[1]), c2 = clip_codes(v[2]);
            mask_or = c0 | c1 | c2;
            mask_and = c0 & c1 & c2;
        }
    }

    // Unclipped triangles: set up now to know whether they survive the cull.
    Prim s;
    uint32_t nprim = 0;
    bool clipped = false;
    if (ok && mask_and == 0) {
        if (mask_or == 0) {
            nprim = setup_prim(s, v, a, nslots, V.width, V.height) ? 1u : 0u;
        } else {
            clipped = true;
        }
    }

    // Clipped triangles: run the clipper; every fan triangle gets a record.
    ClipPoly poly;
    uint32_t pin = 0, npoly = 0;
    if (clipped) {
        for (int k = 0; k < 3; ++k) {
            poly.v[0][k] = v[k];
            for (uint32_t t = 0; t < kMaxSlots; ++t) poly.a[0][k][t] = a[k][t];
        }
        npoly = 3;
        uint32_t m = mask_or;
        do {
            const int plane = __builtin_ctz(m);
            m ^= 1u << plane;
            uint32_t nout = 0;
            uint32_t i0 = npoly - 1;
            float d0 = clip_dot(plane, poly.v[pin][i0]);
            for (uint32_t i1 = 0; i1 < npoly; ++i1) {
                const float d1 = clip_dot(plane, poly.v[pin][i1]);
                const bool in0 = d0 >= 0.0f, in1 = d1 >= 0.0f;
                if (in0) {
                    poly.v[pin ^ 1][nout] = poly.v[pin][i0];
                    for (uint32_t t = 0; t < kMaxSlots; ++t) poly.a[pin ^ 1][nout][t] = poly.a[pin][i0][t];
                    ++nout;
                }
                if (in0 != in1) {
                    uint32_t ia, ib;
                    float tt;
                    if (in1) { tt = d1 / (d1 - d0); ia = i1; ib = i0; }
                    else     { tt = d0 / (d0 - d1); ia = i0; ib = i1; }
                    const float4 pa = poly.v[pin][ia], pb = poly.v[pin][ib];
                    poly.v[pin ^ 1][nout] = make_float4(lerp_kt(pa.x, pb.x, tt), lerp_kt(pa.y, pb.y, tt),
                                                        lerp_kt(pa.z, pb.z, tt), lerp_kt(pa.w, pb.w, tt));
                    for (uint32_t t = 0; t < kMaxSlots; ++t)
                        poly.a[pin ^ 1][nout][t] = lerp_kt(poly.a[pin][ia][t], poly.a[pin][ib][t], tt);
                    ++nout;
                }
                d0 = d1;
                i0 = i1;
            }
            npoly = nout;
            pin ^= 1;
        } while (m && npoly);
        nprim = npoly >= 3 ? npoly - 2 : 0;
    }

    // Allocate records: one atomic per wave.
    uint32_t wave_total;
    const uint32_t excl = wave_excl_scan(nprim, wave_total);
    uint32_t base = 0;
    if (__lane_id() == 0 && wave_total) base = atomicAdd(&F.counters[C_REC], wave_total);
    base = __shfl(base, 0, 64) + excl;
    const bool fits = base + nprim <= F.rec_cap;
    if (nprim && !fits) atomicOr(&F.counters[C_FLAGS], FLAG_REC_OVERFLOW);

    if (nprim && fits) {
        const uint32_t gtri = D.tri_base + tri;
        F.tri_rec[gtri] = base;
        if (!clipped) {
            write_record(F, base, s, gtri << 3, d);
            count_bins(F, hist, use_lds, V, s);
        } else {
            for (uint32_t i = 2; i < npoly; ++i) {
                float4 fv[3] = { poly.v[pin][0], poly.v[pin][i - 1], poly.v[pin][i] };
                float fa[3][kMaxSlots];
                for (uint32_t t = 0; t < kMaxSlots; ++t) {
                    fa[0][t] = poly.a[pin][0][t];
                    fa[1][t] = poly.a[pin][i - 1][t];
                    fa[2][t] = poly.a[pin][i][t];
                }
                const uint32_t r = base + (i - 2);
                if (setup_prim(s, fv, fa, nslots, V.width, V.height)) {
                    write_record(F, r, s, (gtri << 3) | (i - 2), d);
                    count_bins(F, hist, use_lds, V, s);
                } else {
                    F.rec_meta[r] = 0xFFFFFFFFu;
                }
            }
        }
    }

    __syncthreads();
    if (use_lds)
        for (uint32_t i = threadIdx.x; i < ntile_v; i += kSetupThreads) {
            const uint32_t n = hist[i];
            if (n) atomicAdd(&F.tile_count[V.tile_base + i], n);
        }
}

// Exclusive scan of tile counts and tile-queue construction (one workgroup).
constexpr uint32_t kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void k_scan(FrameDev F)
{
    __shared__ uint32_t s_cnt[kScanThreads];
    __shared__ uint32_t s_q[kScanThreads];
    const uint32_t T = F.total_tiles;
    const uint32_t per = (T + kScanThreads - 1) / kScanThreads;
    const uint32_t b = threadIdx.x * per, e = min(T, b + per);
    uint32_t v = 0;
    while (v + 1 < F.nviews && F.views[v + 1].tile_base <= b) ++v;
    uint32_t sum = 0, nq = 0;
    for (uint32_t t = b, vv = v; t < e; ++t) {
        while (vv + 1 < F.nviews && F.views[vv + 1].tile_base <= t) ++vv;
        const uint32_t c = F.tile_count[t];
        sum += c;
        nq += (c || (F.views[vv].flags & (VF_CLEAR_COLOUR | VF_CLEAR_DEPTH))) ? 1u : 0u;
    }
    s_cnt[threadIdx.x] = sum;
    s_q[threadIdx.x] = nq;
    __syncthreads();
    for (uint32_t o = 1; o < kScanThreads; o <<= 1) {
        const uint32_t a1 = threadIdx.x >= o ? s_cnt[threadIdx.x - o] : 0u;
        const uint32_t a2 = threadIdx.x >= o ? s_q[threadIdx.x - o] : 0u;
        __syncthreads();
        s_cnt[threadIdx.x] += a1;
        s_q[threadIdx.x] += a2;
        __syncthreads();
    }
    uint32_t run = s_cnt[threadIdx.x] - sum, qrun = s_q[threadIdx.x] - nq;
    for (uint32_t t = b, vv = v; t < e; ++t) {
        while (vv + 1 < F.nviews && F.views[vv + 1].tile_base <= t) ++vv;
        const uint32_t c = F.tile_count[t];
        F.tile_start[t] = run;
        run += c;
        if (c || (F.views[vv].flags & (VF_CLEAR_COLOUR | VF_CLEAR_DEPTH))) F.queue[qrun++] = t;
    }
    if (threadIdx.x == kScanThreads - 1) {
        F.tile_start[T] = s_cnt[kScanThreads - 1];
        F.counters[C_ENTRIES] = s_cnt[kScanThreads - 1];
        F.counters[C_QLEN] = s_q[kScanThreads - 1];
        if (s_cnt[kScanThreads - 1] > F.entry_cap) atomicOr(&F.counters[C_FLAGS], FLAG_ENTRY_OVERFLOW);
    }
}

__global__ __launch_bounds__(256) void k_emit(FrameDev F)
{
    if (F.counters[C_FLAGS] & (FLAG_ENTRY_OVERFLOW | FLAG_REC_OVERFLOW)) return;
    const uint32_t nrec = F.counters[C_REC];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
        const uint32_t d = F.rec_meta[r];
        if (d == 0xFFFFFFFFu) continue;
        const ViewDev& V = F.views[F.draws[d].view];
        const uint4* R = F.rec_raster + (size_t)r * 4;
        const uint4 q0 = R[0], q1 = R[1], q2 = R[2];
        const int32_t c[3] = { (int32_t)q0.x, (int32_t)q0.y, (int32_t)q0.z };
        const int32_t dx[3] = { (int32_t)q0.w, (int32_t)q1.x, (int32_t)q1.y };
        const int32_t dy[3] = { (int32_t)q1.z, (int32_t)q1.w, (int32_t)q2.x };
        const int32_t xmin = q2.y & 0xFFFF, xmax = q2.y >> 16, ymin = q2.z & 0xFFFF, ymax = q2.z >> 16;
        for_each_bin(c, dx, dy, xmin, xmax, ymin, ymax, [&](uint32_t bx, uint32_t by) {
            const uint32_t gt = V.tile_base + by * V.tiles_x + bx;
            const uint32_t slot = atomicAdd(&F.tile_cursor[gt], 1u);
            F.entries[F.tile_start[gt] + slot] = r;
        });
    }
}

// ------------------------------------------------------------------------
// back end
// ------------------------------------------------------------------------

constexpr uint32_t kBackendThreads = 256;
constexpr uint32_t kBackendWaves = kBackendThreads / 64;

// Depth loaded from the framebuffer -> initial visibility key. A fragment
// passes the reference test iff z > 0 and z > D (ordered), which the key
// order reproduces: D <= 0 -> 0; D NaN -> all ones (nothing passes);
// D > 0 -> D's bits with an order field above any primitive's.
__device__ __forceinline__ unsigned long long init_key(float d0)
{
    if (d0 != d0) return ~0ull;
    if (!(d0 > 0.0f)) return 0ull;
    return ((unsigned long long)fbits(d0) << 32) | 0xFFFFFFFFull;
}

// Resolve + ComputeInterpolants + pixel shader for one visible pixel
// (Rasterizer.cpp:306-419,460-523; Viewer/Shaders.h:71-130).
__device__ __forceinline__ uint32_t shade_pixel(const FrameDev& F, uint32_t rec, int32_t X0, int32_t Y0, uint32_t px,
                                                uint32_t py)
{
    const uint32_t d = F.rec_meta[rec];
    const DrawDev& D = F.draws[d];
    const float4* S = F.rec_shade + (size_t)rec * 4;
    const float4 s0 = S[0], s1 = S[1], s2 = S[2], s3 = S[3];
    const uint4 q2 = F.rec_raster[(size_t)rec * 4 + 2];
    const uint4 q3 = F.rec_raster[(size_t)rec * 4 + 3];
    const float r0x = ffrom(q2.w), r0y = ffrom(q3.x);
    const float sx = (float)X0 - r0x, sy = (float)Y0 - r0y;
    const float wdx = s0.x, wdy = s0.y;
    const float wc = wdx * sx + wdy * sy + s0.z;
    const float pdx[kMaxSlots] = { s0.w, s1.z, s2.y, s3.x };
    const float pdy[kMaxSlots] = { s1.x, s1.w, s2.z, s3.y };
    const float pf0[kMaxSlots] = { s1.y, s2.x, s2.w, s3.z };
    const float X = (float)px, Y = (float)py;
    const float iq = 1.0f / __builtin_fmaf(X, wdx, __builtin_fmaf(Y, wdy, wc));
    float pc[kMaxSlots], var[kMaxSlots];
    const uint32_t ns = D.nslots;
#pragma unroll
    for (uint32_t k = 0; k < kMaxSlots; ++k) {
        pc[k] = pdx[k] * sx + pdy[k] * sy + pf0[k];
        var[k] = k < ns ? iq * __builtin_fmaf(pdy[k], Y, __builtin_fmaf(pdx[k], X, pc[k])) : 0.0f;
    }
    switch (D.shader) {
    case SRG_SHADER_VISUALIZE_NORMALS:
        return pack_rgba(__builtin_fmaf(var[0], 0.5f, 0.5f), __builtin_fmaf(var[1], 0.5f, 0.5f),
                         __builtin_fmaf(var[2], 0.5f, 0.5f), 1.0f);
    case SRG_SHADER_VISUALIZE_UVS:
        return pack_rgba(var[0], var[1], 0.0f, 0.0f);
    default: break;
    }
    if (D.tex == kNoTex) return 0xFFFFFFFFu;
    const TexDev& T = F.texs[D.tex];
    if (T.empty) return 0xFFFFFFFFu;
    const float X1 = 1.0f + X, Y1 = 1.0f + Y;
    const float i10 = rcpps(__builtin_fmaf(wdx, X1, __builtin_fmaf(wdy, Y, wc)));
    const float i01 = rcpps(__builtin_fmaf(wdx, X, __builtin_fmaf(wdy, Y1, wc)));
    float der[4];
    const bool hi = D.deriv_slot != 0;
#pragma unroll
    for (uint32_t uv = 0; uv < 2; ++uv) {
        const float tdx = hi ? pdx[2 + uv] : pdx[uv], tdy = hi ? pdy[2 + uv] : pdy[uv];
        const float tc = hi ? pc[2 + uv] : pc[uv], tv = hi ? var[2 + uv] : var[uv];
        const float t10 = i10 * __builtin_fmaf(tdx, X1, __builtin_fmaf(tdy, Y, tc));
        const float t01 = i01 * __builtin_fmaf(tdx, X, __builtin_fmaf(tdy, Y1, tc));
        der[2 * uv] = t10 - tv;
        der[2 * uv + 1] = t01 - tv;
    }
    return sample_unlit(T, var[0], var[1], der[0], der[1], der[2], der[3]);
}

__global__ __launch_bounds__(kBackendThreads) void k_backend(FrameDev F)
{
    __shared__ unsigned long long keys[kTilePixels];
    __shared__ uint32_t s_q;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t l = lane & 7u, rr = lane >> 3;
    unsigned long long frags = 0;
    if (F.counters[C_FLAGS] & (FLAG_ENTRY_OVERFLOW | FLAG_REC_OVERFLOW)) return;
    const uint32_t qlen = F.counters[C_QLEN];

    for (;;) {
        if (tid == 0) s_q = atomicAdd(&F.counters[C_QHEAD], 1u);
        __syncthreads();
        const uint32_t qi = s_q;
        __syncthreads();
        if (qi >= qlen) break;
        const uint64_t t_start = (F.flags & FF_TILE_COSTS) ? clock64() : 0;
        const uint32_t gt = F.queue[qi];
        uint32_t vi = 0;
        while (vi + 1 < F.nviews && F.views[vi + 1].tile_base <= gt) ++vi;
        const ViewDev V = F.views[vi];
        const uint32_t lt = gt - V.tile_base;
        const uint32_t tx = lt % V.tiles_x, ty = lt / V.tiles_x;
        const int32_t X0 = (int32_t)(tx * kBin), Y0 = (int32_t)(ty * kBin);
        const size_t tofs = (size_t)lt * kTilePixels;
        const bool clear_d = (V.flags & VF_CLEAR_DEPTH) != 0;
        const bool clear_c = (V.flags & VF_CLEAR_COLOUR) != 0;

        for (uint32_t p = tid; p < kTilePixels; p += kBackendThreads) keys[p] = clear_d ? 0ull : init_key(V.depth[tofs + p]);
        __syncthreads();

        const uint32_t beg = F.tile_start[gt], cnt = F.tile_count[gt];
        for (uint32_t e = wave; e < cnt; e += kBackendWaves) {
            const uint32_t r = F.entries[beg + e];
            const uint4* R = F.rec_raster + (size_t)r * 4;
            const uint4 q0 = R[0], q1 = R[1], q2 = R[2], q3 = R[3];
            const uint32_t order = F.rec_order[r];
            const unsigned long long low = 0xFFFFFFFEull - order;
            // Pre-shift the edge constants and plane constants to the tile
            // origin (Binning.cpp:423-451).
            const uint32_t dxu[3] = { q0.w, q1.x, q1.y };
            const uint32_t dyu[3] = { q1.z, q1.w, q2.x };
            const uint32_t ct[3] = { q0.x + (dxu[0] * (uint32_t)Y0 + dyu[0] * (uint32_t)X0),
                                     q0.y + (dxu[1] * (uint32_t)Y0 + dyu[1] * (uint32_t)X0),
                                     q0.z + (dxu[2] * (uint32_t)Y0 + dyu[2] * (uint32_t)X0) };
            const int32_t xmin = q2.y & 0xFFFF, xmax = q2.y >> 16, ymin = q2.z & 0xFFFF, ymax = q2.z >> 16;
            const uint32_t bminx = (uint32_t)clampi(xmin - X0, 0, 64), bmaxx = (uint32_t)clampi(xmax - X0, 0, 64);
            const uint32_t bminy = (uint32_t)clampi(ymin - Y0, 0, 64), bmaxy = (uint32_t)clampi(ymax - Y0, 0, 64);
            const float r0x = ffrom(q2.w), r0y = ffrom(q3.x);
            const float zdx = ffrom(q3.y), zdy = ffrom(q3.z), zf0 = ffrom(q3.w);
            const float sx = (float)X0 - r0x, sy = (float)Y0 - r0y;
            const float zc0 = zdx * sx + zdy * sy + zf0;
            const float ztop = __builtin_fmaf((float)l, zdx, zc0);
            // Block walk and 64-wide block reject (Rasterizer.cpp:203-261).
            for (uint32_t yb = bminy & ~7u; yb < bmaxy; yb += 8)
                for (uint32_t xb = bminx & ~7u; xb < bmaxx; xb += 8) {
                    const uint32_t Xa = xb, Xb = xb + kBin, Ya = yb, Yb = yb + kBin;
                    bool reject = false, full = true;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t m = ((int32_t)(ct[k] + dyu[k] * Xa + dxu[k] * Ya) > 0 ? 1u : 0u) |
                                           ((int32_t)(ct[k] + dyu[k] * Xa + dxu[k] * Yb) > 0 ? 2u : 0u) |
                                           ((int32_t)(ct[k] + dyu[k] * Xb + dxu[k] * Ya) > 0 ? 4u : 0u) |
                                           ((int32_t)(ct[k] + dyu[k] * Xb + dxu[k] * Yb) > 0 ? 8u : 0u);
                        reject = reject || m == 0;
                        full = full && m == 0xFu;
                    }
                    if (reject) continue;
                    // Per-lane pixel (xb + l, yb + rr) (ComputeBlockMask8x8, :134-192).
                    const uint32_t x = xb + l, y = yb + rr;
                    bool covered = full;
                    if (!full) {
                        const uint32_t e0 = ct[0] + dyu[0] * x + dxu[0] * y;
                        const uint32_t e1 = ct[1] + dyu[1] * x + dxu[1] * y;
                        const uint32_t e2 = ct[2] + dyu[2] * x + dxu[2] * y;
                        covered = (int32_t)(e0 | e1 | e2) >= 0;
                    }
                    float z = __builtin_fmaf((float)yb, zdy, ztop) + (float)xb * zdx;
#pragma unroll
                    for (uint32_t i = 0; i < 7; ++i)
                        if (i < rr) z = z + zdy;
                    if (F.flags & FF_STATS) frags += __popcll(__ballot(covered));
                    if (covered && z > 0.0f)
                        atomicMax(&keys[y * kBin + x], ((unsigned long long)fbits(z) << 32) | low);
                }
        }
        __syncthreads();

        for (uint32_t p = tid; p < kTilePixels; p += kBackendThreads) {
            const unsigned long long k = keys[p];
            const uint32_t lo = (uint32_t)k;
            const bool win = lo != 0xFFFFFFFFu && k != 0ull;
            if (win) {
                const uint32_t order = 0xFFFFFFFEu - lo;
                const uint32_t rec = F.tri_rec[order >> 3] + (order & 7u);
                if (V.flags & VF_HAS_COLOUR) V.colour[tofs + p] = shade_pixel(F, rec, X0, Y0, p & 63u, p >> 6);
                V.depth[tofs + p] = ffrom((uint32_t)(k >> 32));
                if (V.primid) V.primid[tofs + p] = order;
            } else {
                if (clear_c) V.colour[tofs + p] = V.clear_colour;
                if (clear_d) V.depth[tofs + p] = 0.0f;
            }
        }
        if ((F.flags & FF_TILE_COSTS) && tid == 0) F.tile_cost[gt] = clock64() - t_start;
        __syncthreads();
    }
    if (F.flags & FF_STATS) {
        if (lane == 0 && frags) atomicAdd((unsigned long long*)&F.counters[C_FRAGS], frags);
    }
}

// Frame-start fill: counters, per-tile counts and cursors (graph-safe; no memset nodes).
__global__ void k_frame_init(uint32_t* counters, uint32_t* tile_count, uint32_t* tile_cursor, uint32_t ntiles)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C_NCOUNTERS) counters[i] = 0;
    for (uint32_t t = i; t < ntiles; t += gridDim.x * blockDim.x) {
        tile_count[t] = 0;
        tile_cursor[t] = 0;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}

__global__ void k_fill_f32(float* p, float value, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}

// BlitJobFn (Renderer.cpp:319-347): 64x64 tiles -> linear rows.
__global__ void k_detile(const uint32_t* tiles, uint32_t* linear, uint32_t width, uint32_t height, uint32_t tiles_x)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= width || y >= height) return;
    const uint32_t t = (y >> 6) * tiles_x + (x >> 6);
    linear[(size_t)y * width + x] = tiles[(size_t)t * kTilePixels + (y & 63u) * 64u + (x & 63u)];
}

}  // namespace srg

// ==========================================================================
// host side
// ==========================================================================

using namespace srg;

namespace {

#define HIP_TRY(expr)                        \
    do {                                     \
        if ((expr) != hipSuccess) return SRG_ERR_HIP; \
    } while (0)

template <typename T>
struct DevArray {
    T* ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap) return SRG_OK;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(n, 64);
        if (hipMalloc((void**)&ptr, want * sizeof(T)) != hipSuccess) { ptr = nullptr; return SRG_ERR_NOMEM; }
        cap = want;
        return SRG_OK;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct Buffer {
    uint8_t* dev = nullptr;
    size_t bytes = 0;
    bool owned = false;
};

struct Texture {
    uint8_t* dev = nullptr;
    uint64_t bytes = 0;
    uint32_t mip_offsets[14] = {};
    uint32_t wlog2 = 0, hlog2 = 0, nmips = 0;
};

struct PlaneState {
    uint32_t* colour = nullptr;
    float* depth = nullptr;
    uint32_t* primid = nullptr;
    bool pending = false;
    int clear_colour = 0, clear_depth = 0;
    uint32_t clear_value = 0;
};

}  // namespace

struct srg_fb {
    srg_ctx* ctx = nullptr;
    uint32_t width = 0, height = 0, tiles_x = 0, tiles_y = 0, flags = 0;
    PlaneState planes[2];
    uint32_t write_plane = 0;
};

struct srg_ctx {
    srg_config cfg{};
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    std::unordered_map<uint64_t, Buffer> buffers;
    std::unordered_map<uint64_t, Texture> textures;
    uint64_t next_handle = 1;
    bool in_frame = false;
    std::vector<srg_draw> draws;
    std::vector<srg_fb*> fbs;

    // per-frame device tables (one blob)
    DevArray<uint8_t> tables;
    uint8_t* h_tables = nullptr;
    size_t h_tables_cap = 0;
    // arenas
    DevArray<uint4> rec_raster;
    DevArray<float4> rec_shade;
    DevArray<uint32_t> rec_order, rec_meta, tri_rec;
    DevArray<uint32_t> tile_count, tile_start, tile_cursor, queue, entries;
    DevArray<uint64_t> tile_cost;
    DevArray<uint32_t> counters;
    DevArray<uint32_t> blit_staging;
    uint32_t* h_counters = nullptr;
    uint32_t rec_cap = 0, entry_cap = 0;

    hipEvent_t ev[4] = {};
    srg_frame_stats stats{};
};

namespace {

uint64_t new_handle(srg_ctx* ctx) { return ctx->next_handle++; }

int ensure_host_tables(srg_ctx* ctx, size_t n)
{
    if (n <= ctx->h_tables_cap) return SRG_OK;
    if (ctx->h_tables) (void)hipHostFree(ctx->h_tables);
    ctx->h_tables = nullptr;
    ctx->h_tables_cap = 0;
    size_t want = std::max<size_t>(n, 1 << 16);
    if (hipHostMalloc((void**)&ctx->h_tables, want, 0) != hipSuccess) return SRG_ERR_NOMEM;
    ctx->h_tables_cap = want;
    return SRG_OK;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int flush_pending_clear(srg_ctx* ctx, srg_fb* fb, uint32_t plane)
{
    PlaneState& ps = fb->planes[plane];
    if (!ps.pending) return SRG_OK;
    const size_t n = (size_t)fb->tiles_x * fb->tiles_y * kTilePixels;
    const uint32_t b = ps.clear_value & 0xFFu;
    if (ps.clear_colour && ps.colour)
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, ctx->stream, ps.colour, b * 0x01010101u, n);
    if (ps.clear_depth && ps.depth) hipLaunchKernelGGL(k_fill_f32, dim3(1024), dim3(256), 0, ctx->stream, ps.depth, 0.0f, n);
    HIP_TRY(hipGetLastError());
    ps.pending = false;
    return SRG_OK;
}

// Attribute slots the shader reads (Viewer/Shaders.h): the device keeps
// planes only for these; the others never influence the output.
int shader_slots(uint32_t shader, uint32_t uv_offset, uint32_t apt, uint32_t slot_attr[kMaxSlots], uint32_t& nslots,
                 uint32_t& deriv_slot)
{
    for (uint32_t k = 0; k < kMaxSlots; ++k) slot_attr[k] = 0;
    switch (shader) {
    case SRG_SHADER_UNLIT_DIFFUSE:
        if (apt < 8 || uv_offset + 1 >= apt) return SRG_ERR_INVALID;
        slot_attr[0] = 6;
        slot_attr[1] = 7;
        if (uv_offset == 6) {
            nslots = 2;
            deriv_slot = 0;
        } else {
            slot_attr[2] = uv_offset;
            slot_attr[3] = uv_offset + 1;
            nslots = 4;
            deriv_slot = 2;
        }
        return SRG_OK;
    case SRG_SHADER_VISUALIZE_NORMALS:
        if (apt < 6 || uv_offset + 1 >= apt) return SRG_ERR_INVALID;
        slot_attr[0] = 3;
        slot_attr[1] = 4;
        slot_attr[2] = 5;
        nslots = 3;
        deriv_slot = 0;
        return SRG_OK;
    case SRG_SHADER_VISUALIZE_UVS:
        if (apt < 8 || uv_offset + 1 >= apt) return SRG_ERR_INVALID;
        slot_attr[0] = 6;
        slot_attr[1] = 7;
        nslots = 2;
        deriv_slot = 0;
        return SRG_OK;
    default:
        return SRG_ERR_UNKNOWN_SHADER;
    }
}

struct ViewKey {
    srg_fb* fb;
    uint32_t plane;
};

int run_frame(srg_ctx* ctx)
{
    // ---- gather views: planes drawn to, plus planes with a pending clear ----
    std::vector<ViewKey> views;
    auto view_of = [&](srg_fb* fb, uint32_t plane) -> uint32_t {
        for (uint32_t i = 0; i < views.size(); ++i)
            if (views[i].fb == fb && views[i].plane == plane) return i;
        views.push_back({ fb, plane });
        return (uint32_t)views.size() - 1;
    };
    for (auto& d : ctx->draws) view_of(d.fb, d.plane);
    for (srg_fb* fb : ctx->fbs)
        for (uint32_t p = 0; p < 2; ++p)
            if (fb->planes[p].pending) view_of(fb, p);

    if (views.empty()) return SRG_OK;
    std::vector<ViewDev> vdev(views.size());
    uint32_t total_tiles = 0;
    for (size_t i = 0; i < views.size(); ++i) {
        srg_fb* fb = views[i].fb;
        PlaneState& ps = fb->planes[views[i].plane];
        ViewDev& v = vdev[i];
        v.colour = ps.colour;
        v.depth = ps.depth;
        v.primid = (ctx->cfg.flags & SRG_CFG_KEEP_PRIMID) ? ps.primid : nullptr;
        v.width = fb->width;
        v.height = fb->height;
        v.tiles_x = fb->tiles_x;
        v.tiles_y = fb->tiles_y;
        v.tile_base = total_tiles;
        v.flags = (ps.colour ? VF_HAS_COLOUR : 0u);
        if (ps.pending) {
            v.flags |= (ps.clear_colour && ps.colour ? VF_CLEAR_COLOUR : 0u) | (ps.clear_depth ? VF_CLEAR_DEPTH : 0u);
            v.clear_colour = (ps.clear_value & 0xFFu) * 0x01010101u;
        }
        v.pad = 0;
        total_tiles += fb->tiles_x * fb->tiles_y;
    }

    // ---- draws, textures, front-end blocks ----
    std::vector<DrawDev> ddev(ctx->draws.size());
    std::vector<TexDev> tdev;
    std::unordered_map<uint64_t, uint32_t> tex_index;
    std::vector<uint32_t> block_draw, block_tri0;
    uint64_t total_tris = 0;
    for (size_t i = 0; i < ctx->draws.size(); ++i) {
        const srg_draw& s = ctx->draws[i];
        DrawDev& d = ddev[i];
        const Buffer& ib = ctx->buffers[s.index_buffer];
        const Buffer& pb = ctx->buffers[s.position_buffer];
        const Buffer& ab = ctx->buffers[s.attribute_buffer];
        d.idx = ib.dev + s.index_offset;
        d.pos = pb.dev + s.position_offset;
        d.attr = ab.dev + s.attribute_offset;
        d.idx_stride = s.index_stride;
        d.pos_stride = s.position_stride;
        d.attr_stride = s.attribute_stride;
        d.ntris = s.index_num / 3u;
        d.pos_num = s.position_num;
        d.attr_num = s.attribute_num;
        d.tri_base = (uint32_t)total_tris;
        d.view = view_of(s.fb, s.plane);
        d.shader = s.shader;
        int st = shader_slots(s.shader, s.uv_offset, s.attribute_stride / 4u, d.slot_attr, d.nslots, d.deriv_slot);
        if (st) return st;
        d.tex = kNoTex;
        if (s.shader == SRG_SHADER_UNLIT_DIFFUSE && s.texture) {
            auto it = tex_index.find(s.texture);
            if (it == tex_index.end()) {
                const Texture& t = ctx->textures[s.texture];
                TexDev td{};
                td.texels = t.dev;
                std::memcpy(td.mip_offsets, t.mip_offsets, sizeof(td.mip_offsets));
                td.wlog2 = t.wlog2;
                td.hlog2 = t.hlog2;
                td.nmips = t.nmips;
                td.empty = (t.bytes == 0 || t.dev == nullptr) ? 1u : 0u;
                tex_index[s.texture] = (uint32_t)tdev.size();
                d.tex = (uint32_t)tdev.size();
                tdev.push_back(td);
            } else {
                d.tex = it->second;
            }
        }
        std::memcpy(d.mvp, s.mvp, sizeof(d.mvp));
        for (uint32_t t0 = 0; t0 < d.ntris; t0 += kSetupThreads) {
            block_draw.push_back((uint32_t)i);
            block_tri0.push_back(t0);
        }
        total_tris += d.ntris;
    }
    if (total_tris >= (1ull << 29)) return SRG_ERR_INVALID;  // primitive order key is (tri << 3) | fan
    // views may have been added by draws of planes first seen above; refresh
    if (vdev.size() != views.size()) return SRG_ERR_STATE;

    const uint32_t nblocks = (uint32_t)block_draw.size();
    size_t off = 0;
    const size_t o_draw = off; off = align_up(off + sizeof(DrawDev) * ddev.size(), 256);
    const size_t o_view = off; off = align_up(off + sizeof(ViewDev) * vdev.size(), 256);
    const size_t o_tex = off;  off = align_up(off + sizeof(TexDev) * std::max<size_t>(tdev.size(), 1), 256);
    const size_t o_bd = off;   off = align_up(off + sizeof(uint32_t) * std::max<uint32_t>(nblocks, 1), 256);
    const size_t o_bt = off;   off = align_up(off + sizeof(uint32_t) * std::max<uint32_t>(nblocks, 1), 256);
    int st = ensure_host_tables(ctx, off);
    if (st) return st;
    if ((st = ctx->tables.ensure(off))) return st;
    std::memcpy(ctx->h_tables + o_draw, ddev.data(), sizeof(DrawDev) * ddev.size());
    std::memcpy(ctx->h_tables + o_view, vdev.data(), sizeof(ViewDev) * vdev.size());
    if (!tdev.empty()) std::memcpy(ctx->h_tables + o_tex, tdev.data(), sizeof(TexDev) * tdev.size());
    if (nblocks) {
        std::memcpy(ctx->h_tables + o_bd, block_draw.data(), sizeof(uint32_t) * nblocks);
        std::memcpy(ctx->h_tables + o_bt, block_tri0.data(), sizeof(uint32_t) * nblocks);
    }

    // ---- arenas ----
    const uint32_t T = std::max<uint32_t>(total_tiles, 1);
    if ((st = ctx->tile_count.ensure(T)) || (st = ctx->tile_start.ensure(T + 1)) ||
        (st = ctx->tile_cursor.ensure(T)) || (st = ctx->queue.ensure(T)) || (st = ctx->tile_cost.ensure(T)) ||
        (st = ctx->counters.ensure(C_NCOUNTERS)) || (st = ctx->tri_rec.ensure(std::max<uint64_t>(total_tris, 1))))
        return st;
    uint64_t want_rec = ctx->cfg.max_prims ? ctx->cfg.max_prims : total_tris + total_tris / 8 + 1024;
    want_rec = std::max<uint64_t>(want_rec, ctx->rec_cap);
    if (want_rec > 0xFFFFFFF0ull) want_rec = 0xFFFFFFF0ull;
    uint64_t want_entries = ctx->cfg.max_entries ? ctx->cfg.max_entries : 2 * total_tris + 4096;
    want_entries = std::max<uint64_t>(want_entries, ctx->entry_cap);
    if (want_entries > 0xFFFFFFF0ull) want_entries = 0xFFFFFFF0ull;

    for (uint32_t attempt = 0;; ++attempt) {
        if ((st = ctx->rec_raster.ensure(want_rec * 4)) || (st = ctx->rec_shade.ensure(want_rec * 4)) ||
            (st = ctx->rec_order.ensure(want_rec)) || (st = ctx->rec_meta.ensure(want_rec)) ||
            (st = ctx->entries.ensure(want_entries)))
            return st;
        ctx->rec_cap = (uint32_t)want_rec;
        ctx->entry_cap = (uint32_t)want_entries;

        HIP_TRY(hipMemcpyAsync(ctx->tables.ptr, ctx->h_tables, off, hipMemcpyHostToDevice, ctx->stream));

        FrameDev F{};
        F.draws = (const DrawDev*)(ctx->tables.ptr + o_draw);
        F.views = (const ViewDev*)(ctx->tables.ptr + o_view);
        F.texs = (const TexDev*)(ctx->tables.ptr + o_tex);
        F.block_draw = (const uint32_t*)(ctx->tables.ptr + o_bd);
        F.block_tri0 = (const uint32_t*)(ctx->tables.ptr + o_bt);
        F.ndraws = (uint32_t)ddev.size();
        F.nviews = (uint32_t)vdev.size();
        F.total_tiles = total_tiles;
        F.total_tris = (uint32_t)total_tris;
        F.rec_raster = ctx->rec_raster.ptr;
        F.rec_shade = ctx->rec_shade.ptr;
        F.rec_order = ctx->rec_order.ptr;
        F.rec_meta = ctx->rec_meta.ptr;
        F.tri_rec = ctx->tri_rec.ptr;
        F.rec_cap = ctx->rec_cap;
        F.entry_cap = ctx->entry_cap;
        F.tile_count = ctx->tile_count.ptr;
        F.tile_start = ctx->tile_start.ptr;
        F.tile_cursor = ctx->tile_cursor.ptr;
        F.entries = ctx->entries.ptr;
        F.queue = ctx->queue.ptr;
        F.tile_cost = ctx->tile_cost.ptr;
        F.counters = ctx->counters.ptr;
        F.flags = ((ctx->cfg.flags & SRG_CFG_STATS) ? FF_STATS : 0u) |
                  ((ctx->cfg.flags & SRG_CFG_KEEP_PRIMID) ? FF_PRIMID : 0u) |
                  ((ctx->cfg.flags & SRG_CFG_TILE_COSTS) ? FF_TILE_COSTS : 0u);
        F.lds_hist_tiles = kLdsHistTiles;

        const bool timing = (ctx->cfg.flags & SRG_CFG_TIMING) != 0;
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
        hipLaunchKernelGGL(k_frame_init, dim3(std::max<uint32_t>(1, std::min<uint32_t>((T + 255) / 256, 1024))),
                           dim3(256), 0, ctx->stream, ctx->counters.ptr, ctx->tile_count.ptr, ctx->tile_cursor.ptr, T);
        if (ctx->cfg.flags & SRG_CFG_KEEP_PRIMID)
            for (auto& v : vdev)
                if (v.primid)
                    hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, ctx->stream, v.primid, 0xFFFFFFFFu,
                                       (size_t)v.tiles_x * v.tiles_y * kTilePixels);
        if (nblocks) hipLaunchKernelGGL(k_setup, dim3(nblocks), dim3(kSetupThreads), 0, ctx->stream, F);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, ctx->stream, F);
        const uint32_t emit_blocks = (uint32_t)std::min<uint64_t>((want_rec + 255) / 256, 8192);
        hipLaunchKernelGGL(k_emit, dim3(std::max<uint32_t>(emit_blocks, 1)), dim3(256), 0, ctx->stream, F);
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
        const uint32_t wg = ctx->cfg.wg_per_cu ? ctx->cfg.wg_per_cu : 4;
        hipLaunchKernelGGL(k_backend, dim3((uint32_t)ctx->num_cus * wg), dim3(kBackendThreads), 0, ctx->stream, F);
        if (timing) HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(ctx->h_counters, ctx->counters.ptr, sizeof(uint32_t) * C_NCOUNTERS,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));

        const uint32_t flags = ctx->h_counters[C_FLAGS];
        if (flags & FLAG_INDEX_RANGE) return SRG_ERR_INDEX_RANGE;
        if (flags & (FLAG_REC_OVERFLOW | FLAG_ENTRY_OVERFLOW)) {
            if (attempt >= 4) return SRG_ERR_OVERFLOW;
            if (flags & FLAG_REC_OVERFLOW) want_rec = std::min<uint64_t>(want_rec * 2, 0xFFFFFFF0ull);
            if (flags & FLAG_ENTRY_OVERFLOW)
                want_entries = std::min<uint64_t>(std::max<uint64_t>(want_entries * 2, (uint64_t)ctx->h_counters[C_ENTRIES] * 5 / 4),
                                                  0xFFFFFFF0ull);
            ctx->stats.retries++;
            continue;
        }
        ctx->stats.tris_in = total_tris;
        ctx->stats.prims = ctx->h_counters[C_REC];
        ctx->stats.bin_entries = ctx->h_counters[C_ENTRIES];
        ctx->stats.frags_raster = (uint64_t)ctx->h_counters[C_FRAGS] | ((uint64_t)ctx->h_counters[C_FRAGS + 1] << 32);
        ctx->stats.tiles_queued = ctx->h_counters[C_QLEN];
        if (timing) {
            hipEventElapsedTime(&ctx->stats.ms_frontend, ctx->ev[0], ctx->ev[1]);
            hipEventElapsedTime(&ctx->stats.ms_backend, ctx->ev[1], ctx->ev[2]);
            hipEventElapsedTime(&ctx->stats.ms_total, ctx->ev[0], ctx->ev[2]);
        }
        break;
    }
    for (auto& v : views) v.fb->planes[v.plane].pending = false;
    return SRG_OK;
}

bool valid_handle(srg_ctx* ctx, srg_handle h) { return h && ctx->buffers.count(h); }

}  // namespace

extern "C" {

const char* srg_status_string(int s)
{
    switch (s) {
    case SRG_OK: return "ok";
    case SRG_ERR_INVALID: return "invalid argument";
    case SRG_ERR_NOMEM: return "out of memory";
    case SRG_ERR_HIP: return "HIP runtime error";
    case SRG_ERR_INDEX_RANGE: return "index out of range";
    case SRG_ERR_OVERFLOW: return "device arena overflow";
    case SRG_ERR_UNKNOWN_SHADER: return "unknown shader";
    case SRG_ERR_STATE: return "call out of order";
    case SRG_ERR_NO_DEVICE: return "no usable HIP device";
    case SRG_ERR_HANDLE: return "unknown handle";
    default: return "unknown status";
    }
}

int srg_create(const srg_config* cfg, srg_ctx** out)
{
    if (!out) return SRG_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SRG_ERR_NO_DEVICE;
    srg_ctx* ctx = new (std::nothrow) srg_ctx();
    if (!ctx) return SRG_ERR_NOMEM;
    if (cfg) ctx->cfg = *cfg;
    ctx->cfg.struct_size = sizeof(srg_config);
    ctx->device = ctx->cfg.device;
    if (ctx->device < 0 || ctx->device >= ndev) { delete ctx; return SRG_ERR_INVALID; }
    if (hipSetDevice(ctx->device) != hipSuccess) { delete ctx; return SRG_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) { delete ctx; return SRG_ERR_HIP; }
    ctx->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return SRG_ERR_HIP; }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) { srg_destroy(ctx); return SRG_ERR_HIP; }
    if (hipHostMalloc((void**)&ctx->h_counters, sizeof(uint32_t) * C_NCOUNTERS, 0) != hipSuccess) {
        srg_destroy(ctx);
        return SRG_ERR_NOMEM;
    }
    *out = ctx;
    return SRG_OK;
}

void srg_destroy(srg_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->buffers)
        if (kv.second.owned) (void)hipFree(kv.second.dev);
    for (auto& kv : ctx->textures) (void)hipFree(kv.second.dev);
    for (srg_fb* fb : std::vector<srg_fb*>(ctx->fbs)) srg_fb_destroy(ctx, fb);
    ctx->tables.release();
    ctx->rec_raster.release();
    ctx->rec_shade.release();
    ctx->rec_order.release();
    ctx->rec_meta.release();
    ctx->tri_rec.release();
    ctx->tile_count.release();
    ctx->tile_start.release();
    ctx->tile_cursor.release();
    ctx->queue.release();
    ctx->entries.release();
    ctx->tile_cost.release();
    ctx->counters.release();
    ctx->blit_staging.release();
    if (ctx->h_tables) (void)hipHostFree(ctx->h_tables);
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int srg_register_buffer(srg_ctx* ctx, const void* host, size_t bytes, srg_handle* out)
{
    if (!ctx || !out || (!host && bytes)) return SRG_ERR_INVALID;
    Buffer b;
    b.bytes = bytes;
    b.owned = true;
    if (hipMalloc((void**)&b.dev, std::max<size_t>(bytes, 4)) != hipSuccess) return SRG_ERR_NOMEM;
    if (bytes && hipMemcpy(b.dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(b.dev);
        return SRG_ERR_HIP;
    }
    uint64_t h = new_handle(ctx);
    ctx->buffers[h] = b;
    *out = h;
    return SRG_OK;
}

int srg_register_device_buffer(srg_ctx* ctx, void* device_ptr, size_t bytes, srg_handle* out)
{
    if (!ctx || !out || !device_ptr) return SRG_ERR_INVALID;
    Buffer b;
    b.dev = (uint8_t*)device_ptr;
    b.bytes = bytes;
    b.owned = false;
    uint64_t h = new_handle(ctx);
    ctx->buffers[h] = b;
    *out = h;
    return SRG_OK;
}

int srg_update_buffer(srg_ctx* ctx, srg_handle h, size_t offset, const void* host, size_t bytes)
{
    if (!ctx || !host) return SRG_ERR_INVALID;
    auto it = ctx->buffers.find(h);
    if (it == ctx->buffers.end()) return SRG_ERR_HANDLE;
    if (offset > it->second.bytes || bytes > it->second.bytes - offset) return SRG_ERR_INVALID;
    HIP_TRY(hipMemcpy(it->second.dev + offset, host, bytes, hipMemcpyHostToDevice));
    return SRG_OK;
}

int srg_release_buffer(srg_ctx* ctx, srg_handle h)
{
    if (!ctx) return SRG_ERR_INVALID;
    auto it = ctx->buffers.find(h);
    if (it == ctx->buffers.end()) return SRG_ERR_HANDLE;
    (void)hipStreamSynchronize(ctx->stream);
    if (it->second.owned) (void)hipFree(it->second.dev);
    ctx->buffers.erase(it);
    return SRG_OK;
}

int srg_register_texture(srg_ctx* ctx, const void* tiled, uint64_t bytes, const uint32_t mip_offsets[14],
                         uint32_t wlog2, uint32_t hlog2, uint32_t nmips, srg_handle* out)
{
    if (!ctx || !out || (bytes && (!tiled || !mip_offsets))) return SRG_ERR_INVALID;
    Texture t;
    t.bytes = bytes;
    t.wlog2 = wlog2;
    t.hlog2 = hlog2;
    t.nmips = nmips;
    if (bytes) {
        if (nmips == 0 || nmips > 14 || wlog2 >= 14 || hlog2 >= 14) return SRG_ERR_INVALID;
        // Every texel the sampler can address must lie inside the texture.
        for (uint32_t m = 0; m < nmips; ++m) {
            const uint64_t w = std::max<uint64_t>(1, (1ull << wlog2) >> m), h = std::max<uint64_t>(1, (1ull << hlog2) >> m);
            const uint64_t tw = (std::max<uint64_t>(w, 32) + 31) / 32, th = (std::max<uint64_t>(h, 32) + 31) / 32;
            if ((uint64_t)mip_offsets[m] + tw * th * 4096ull > bytes) return SRG_ERR_INVALID;
        }
        std::memcpy(t.mip_offsets, mip_offsets, sizeof(t.mip_offsets));
        if (hipMalloc((void**)&t.dev, bytes) != hipSuccess) return SRG_ERR_NOMEM;
        if (hipMemcpy(t.dev, tiled, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(t.dev);
            return SRG_ERR_HIP;
        }
    }
    uint64_t h = new_handle(ctx);
    ctx->textures[h] = t;
    *out = h;
    return SRG_OK;
}

int srg_release_texture(srg_ctx* ctx, srg_handle h)
{
    if (!ctx) return SRG_ERR_INVALID;
    auto it = ctx->textures.find(h);
    if (it == ctx->textures.end()) return SRG_ERR_HANDLE;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(it->second.dev);
    ctx->textures.erase(it);
    return SRG_OK;
}

int srg_fb_create(srg_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, srg_fb** out)
{
    if (!ctx || !out || !width || !height || width > 16384 || height > 16384 || !(flags & SRG_FB_DEPTH))
        return SRG_ERR_INVALID;
    srg_fb* fb = new (std::nothrow) srg_fb();
    if (!fb) return SRG_ERR_NOMEM;
    fb->ctx = ctx;
    fb->width = width;
    fb->height = height;
    fb->tiles_x = (width + kBin - 1) / kBin;
    fb->tiles_y = (height + kBin - 1) / kBin;
    fb->flags = flags;
    const size_t n = (size_t)fb->tiles_x * fb->tiles_y * kTilePixels;
    for (auto& p : fb->planes) {
        bool okp = true;
        if (flags & SRG_FB_COLOUR) okp = okp && hipMalloc((void**)&p.colour, n * 4) == hipSuccess;
        okp = okp && hipMalloc((void**)&p.depth, n * 4) == hipSuccess;
        if (ctx->cfg.flags & SRG_CFG_KEEP_PRIMID) okp = okp && hipMalloc((void**)&p.primid, n * 4) == hipSuccess;
        if (!okp) {
            ctx->fbs.push_back(fb);
            srg_fb_destroy(ctx, fb);
            return SRG_ERR_NOMEM;
        }
        // FrameBufferPlane::Init leaves the tiles uninitialised; start from a
        // defined state (depth 0 = cleared reverse-Z, colour 0).
        if (p.colour) hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, ctx->stream, p.colour, 0u, n);
        hipLaunchKernelGGL(k_fill_f32, dim3(256), dim3(256), 0, ctx->stream, p.depth, 0.0f, n);
        if (p.primid) hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, ctx->stream, p.primid, 0xFFFFFFFFu, n);
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctx->fbs.push_back(fb);
        srg_fb_destroy(ctx, fb);
        return SRG_ERR_HIP;
    }
    ctx->fbs.push_back(fb);
    *out = fb;
    return SRG_OK;
}

void srg_fb_destroy(srg_ctx* ctx, srg_fb* fb)
{
    if (!ctx || !fb) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : fb->planes) {
        if (p.colour) (void)hipFree(p.colour);
        if (p.depth) (void)hipFree(p.depth);
        if (p.primid) (void)hipFree(p.primid);
    }
    ctx->fbs.erase(std::remove(ctx->fbs.begin(), ctx->fbs.end(), fb), ctx->fbs.end());
    delete fb;
}

uint32_t srg_fb_write_plane(const srg_fb* fb) { return fb ? fb->write_plane : 0u; }
void srg_fb_swap_planes(srg_fb* fb)
{
    if (fb) fb->write_plane ^= 1u;
}

int srg_fb_read_tiles(srg_ctx* ctx, srg_fb* fb, uint32_t plane, uint32_t* colour, float* depth, uint32_t* primid)
{
    if (!ctx || !fb || plane > 1) return SRG_ERR_INVALID;
    int st = flush_pending_clear(ctx, fb, plane);
    if (st) return st;
    const size_t n = (size_t)fb->tiles_x * fb->tiles_y * kTilePixels;
    PlaneState& p = fb->planes[plane];
    if (colour) {
        if (!p.colour) return SRG_ERR_INVALID;
        HIP_TRY(hipMemcpyAsync(colour, p.colour, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (depth) HIP_TRY(hipMemcpyAsync(depth, p.depth, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (primid) {
        if (!p.primid) return SRG_ERR_INVALID;
        HIP_TRY(hipMemcpyAsync(primid, p.primid, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SRG_OK;
}

int srg_fb_write_tiles(srg_ctx* ctx, srg_fb* fb, uint32_t plane, const uint32_t* colour, const float* depth)
{
    if (!ctx || !fb || plane > 1) return SRG_ERR_INVALID;
    const size_t n = (size_t)fb->tiles_x * fb->tiles_y * kTilePixels;
    PlaneState& p = fb->planes[plane];
    if (colour) {
        if (!p.colour) return SRG_ERR_INVALID;
        if (p.pending) p.clear_colour = 0;
        HIP_TRY(hipMemcpyAsync(p.colour, colour, n * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    if (depth) {
        if (p.pending) p.clear_depth = 0;
        HIP_TRY(hipMemcpyAsync(p.depth, depth, n * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    if (p.pending && !p.clear_colour && !p.clear_depth) p.pending = false;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SRG_OK;
}

int srg_begin_frame(srg_ctx* ctx)
{
    if (!ctx) return SRG_ERR_INVALID;
    ctx->draws.clear();
    ctx->in_frame = true;
    return SRG_OK;
}

int srg_clear(srg_ctx* ctx, srg_fb* fb, uint32_t colour, int clear_colour, int clear_depth)
{
    if (!ctx || !fb) return SRG_ERR_INVALID;
    PlaneState& p = fb->planes[fb->write_plane];
    // ClearFrameBuffer acts immediately in the reference, so within a frame
    // it always precedes the rasterisation done by EndFrame; the clear is
    // recorded here and fused into the frame's back end (or flushed by a
    // blit / read of the plane).
    if (!clear_colour && !clear_depth) return SRG_OK;
    if (p.pending) {
        // merge: a later clear overrides the earlier one channel by channel
        if (clear_colour) { p.clear_colour = 1; p.clear_value = colour; }
        if (clear_depth) p.clear_depth = 1;
    } else {
        p.pending = true;
        p.clear_colour = clear_colour ? 1 : 0;
        p.clear_depth = clear_depth ? 1 : 0;
        p.clear_value = colour;
    }
    return SRG_OK;
}

int srg_draw_indexed(srg_ctx* ctx, const srg_draw* d)
{
    if (!ctx || !d) return SRG_ERR_INVALID;
    if (!ctx->in_frame) return SRG_ERR_STATE;
    if (!d->fb || d->plane > 1) return SRG_ERR_INVALID;
    if (std::find(ctx->fbs.begin(), ctx->fbs.end(), d->fb) == ctx->fbs.end()) return SRG_ERR_HANDLE;
    if (!valid_handle(ctx, d->index_buffer) || !valid_handle(ctx, d->position_buffer) ||
        !valid_handle(ctx, d->attribute_buffer))
        return SRG_ERR_HANDLE;
    if (d->index_stride != 1 && d->index_stride != 2 && d->index_stride != 4) return SRG_ERR_INVALID;
    if (d->position_stride < 12 || (d->position_stride & 3) || (d->position_offset & 3)) return SRG_ERR_INVALID;
    if (d->attribute_stride == 0 || d->attribute_stride > 32 || (d->attribute_stride & 3) || (d->attribute_offset & 3))
        return SRG_ERR_INVALID;
    if (d->index_offset % d->index_stride) return SRG_ERR_INVALID;
    uint32_t sa[kMaxSlots], ns, ds;
    int st = shader_slots(d->shader, d->uv_offset, d->attribute_stride / 4u, sa, ns, ds);
    if (st) return st;
    if (d->shader == SRG_SHADER_UNLIT_DIFFUSE && d->texture && !ctx->textures.count(d->texture)) return SRG_ERR_HANDLE;
    const Buffer& ib = ctx->buffers[d->index_buffer];
    const Buffer& pb = ctx->buffers[d->position_buffer];
    const Buffer& ab = ctx->buffers[d->attribute_buffer];
    const uint64_t ntri = d->index_num / 3u;
    if (d->index_offset + ntri * 3u * d->index_stride > ib.bytes) return SRG_ERR_INVALID;
    if (d->position_num && d->position_offset + (uint64_t)(d->position_num - 1) * d->position_stride + 12 > pb.bytes)
        return SRG_ERR_INVALID;
    if (d->attribute_num &&
        d->attribute_offset + (uint64_t)(d->attribute_num - 1) * d->attribute_stride + d->attribute_stride > ab.bytes)
        return SRG_ERR_INVALID;
    ctx->draws.push_back(*d);
    return SRG_OK;
}

int srg_end_frame(srg_ctx* ctx)
{
    if (!ctx) return SRG_ERR_INVALID;
    if (!ctx->in_frame) return SRG_ERR_STATE;
    ctx->in_frame = false;
    ctx->stats = srg_frame_stats{};
    int st = run_frame(ctx);
    ctx->draws.clear();
    return st;
}

int srg_blit(srg_ctx* ctx, srg_fb* fb, void* linear, void (*on_finish)(void*), void* user)
{
    if (!ctx || !fb || !linear) return SRG_ERR_INVALID;
    PlaneState& p = fb->planes[fb->write_plane];
    if (!p.colour) return SRG_ERR_INVALID;
    int st = flush_pending_clear(ctx, fb, fb->write_plane);
    if (st) return st;
    const size_t npx = (size_t)fb->width * fb->height;
    if ((st = ctx->blit_staging.ensure(npx))) return st;
    hipLaunchKernelGGL(k_detile, dim3((fb->width + 255) / 256, fb->height), dim3(256), 0, ctx->stream, p.colour,
                       ctx->blit_staging.ptr, fb->width, fb->height, fb->tiles_x);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(linear, ctx->blit_staging.ptr, npx * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    fb->write_plane ^= 1u;
    if (on_finish) on_finish(user);
    return SRG_OK;
}

int srg_get_frame_stats(srg_ctx* ctx, srg_frame_stats* out)
{
    if (!ctx || !out) return SRG_ERR_INVALID;
    *out = ctx->stats;
    return SRG_OK;
}

// TileTexture (Texture.cpp:73-101): 32x32 tiles, Morton order inside a tile,
// tile rows of AlignUp(width, 32) / 32 tiles.
static void tile_texture(const uint8_t* src, uint8_t* dst, uint32_t w, uint32_t h)
{
    const uint32_t tiles_w = ((w + 31u) & ~31u) >> 5;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            uint32_t m = 0;
            for (uint32_t b = 0; b < 5; ++b) m |= (((x >> b) & 1u) << (2 * b)) | (((y >> b) & 1u) << (2 * b + 1));
            const uint32_t offs = ((y >> 5) * tiles_w + (x >> 5)) * 1024u + m;
            std::memcpy(dst + (size_t)offs * 4, src + ((size_t)y * w + x) * 4, 4);
        }
}

int srg_texture_build_rgba8(const uint8_t* linear, uint32_t width, uint32_t height, int calc_mips, uint8_t* out,
                            uint64_t* out_bytes, uint32_t mip_offsets[14], uint32_t* wlog2, uint32_t* hlog2,
                            uint32_t* nmips)
{
    // CreateFromRGBA8 (Texture.cpp:119-199) requires power-of-two sizes that
    // are multiples of the 32-texel tile and below 2^14.
    auto is_pow2 = [](uint32_t v) { return v && !(v & (v - 1)); };
    if (!out_bytes || !is_pow2(width) || !is_pow2(height) || width < 32 || height < 32 || width >= 16384 ||
        height >= 16384)
        return SRG_ERR_INVALID;
    uint32_t wl = 0, hl = 0;
    while ((1u << wl) < width) ++wl;
    while ((1u << hl) < height) ++hl;
    const uint32_t n = calc_mips ? std::max(wl, hl) + 1 : 1;
    uint32_t offs[14] = {};
    uint64_t total = 0;
    for (uint32_t m = 0; m < n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        offs[m] = (uint32_t)total;
        total += (uint64_t)((mw + 31u) & ~31u) * ((mh + 31u) & ~31u) * 4u;
    }
    *out_bytes = total;
    if (mip_offsets) std::memcpy(mip_offsets, offs, sizeof(offs));
    if (wlog2) *wlog2 = wl;
    if (hlog2) *hlog2 = hl;
    if (nmips) *nmips = n;
    if (!out) return SRG_OK;
    if (!linear) return SRG_ERR_INVALID;
    std::memset(out, 0, total);
    tile_texture(linear, out, width, height);
    // Mip m is a box average of the (2^m x 2^m) base-level footprint, rounded
    // half up, computed from the base level as the reference's chain is.
    std::vector<uint8_t> tmp;
    for (uint32_t m = 1; m < n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        const uint32_t fx = width / mw, fy = height / mh;
        tmp.assign((size_t)mw * mh * 4, 0);
        for (uint32_t y = 0; y < mh; ++y)
            for (uint32_t x = 0; x < mw; ++x)
                for (uint32_t c = 0; c < 4; ++c) {
                    uint64_t s = 0;
                    for (uint32_t yy = 0; yy < fy; ++yy)
                        for (uint32_t xx = 0; xx < fx; ++xx)
                            s += linear[(((size_t)(y * fy + yy)) * width + (x * fx + xx)) * 4 + c];
                    const uint64_t cnt = (uint64_t)fx * fy;
                    tmp[((size_t)y * mw + x) * 4 + c] = (uint8_t)((s + cnt / 2) / cnt);
                }
        tile_texture(tmp.data(), out + offs[m], mw, mh);
    }
    return SRG_OK;
}

}  // extern "C"
