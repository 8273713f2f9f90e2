// sr_frontend.h -- front end: fetch, transform, clip, triangle setup and
// binning (SoftRast/Binning.cpp), as four kernels:
//   k_setup  one lane per triangle of a setup group (draws sharing their
//            geometry), fetched once and set up for each draw of the group;
//            unclipped survivors are written to their own record slot,
//            clipped triangles are queued for k_clip; bin entries are counted
//   k_clip   slow path over the queued triangles: Sutherland-Hodgman clip,
//            fan, setup of each fan triangle, bin counts
//   k_scan   exclusive scan of bin counts + the back end's work queue
//            (three passes; k_scan_single in one workgroup for small frames)
//   k_scatter one lane per (record, bin) pair found by the count walks:
//            writes the record index into its bin list
#pragma once
#include "sr_frame.h"

namespace srg {

struct Prim {
    int32_t c[3], dx[3], dy[3];
    int32_t xmin, xmax, ymin, ymax;
    uint32_t tight;  // bit 0: tight bound holds; bits 1, 2: the bound starts one pixel left of xmin / above ymin
    float r0x, r0y, zdx, zdy, zf0, wdx, wdy, wf0;
    float adx[kWideSlots], ady[kWideSlots], af0[kWideSlots];  // the first NS used by a setup<NS>
};

// kt::Mat4 * kt::Vec4 (Binning.cpp:477-479), order ((c0*x + c1*y) + c2*z) + c3*w,
// w = 1, every operation rounded.
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

// SetupEdge (Binning.cpp:242-259): fill-rule bias, >> 8 truncated to 32 bits.
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

// Which pixels can the reference's coverage test accept? It accepts pixel
// (x, y) iff E_k = ((c64_k + b_k) >> 8) + dy_k*x + dx_k*y >= 0 for the three
// edges (int32, wrapping), b_k in {0, 256} the fill bias. Without wrap this
// is E'_k(P) + b_k >= 0 for the exact edge function E' at P = 256*(x, y),
// and E'_0 + E'_1 + E'_2 = A, the doubled area (A >= 256 after the cull). So
// the barycentric coordinates of an accepted P are >= -256/A, and P lies
// within 2*256*W/A (W = x extent of the vertices) of the vertex x range. With
// A > 2*max(W, H) that is less than one subpixel short of 256: every
// accepted pixel has mnx >> 8 <= x <= ceil(mxx / 256) (same for y). No wrap:
// W, H < 2^18 keeps |E'| < 2^38 over the 64-pixel block extents used, below
// the 2^39 at which E would wrap. The blocks the reference walks may reach
// past the last column or row into the tile's padding, so where the bbox
// maximum was clamped to the framebuffer the bound stays open on that side.
// Returns the record's tight bits: bit 0 valid, bit 1 the x bound starts at
// xmin - 1, bit 2 the y bound at ymin - 1, bit 3 / 4 no x / y upper bound.
__device__ __forceinline__ uint32_t tight_bound(int64_t area2, int32_t mnx, int32_t mxx, int32_t mny, int32_t mxy,
                                                int32_t xmin, int32_t ymin, int32_t xmax, int32_t ymax)
{
    const int64_t w = (int64_t)mxx - mnx, h = (int64_t)mxy - mny;
    if (w >= (1 << 18) || h >= (1 << 18) || area2 <= 2 * (w > h ? w : h)) return 0u;
    const int32_t dxl = xmin - max(mnx >> 8, 0), dyl = ymin - max(mny >> 8, 0);
    if ((uint32_t)dxl > 1u || (uint32_t)dyl > 1u) return 0u;
    const bool xopen = (((int64_t)mxx + 255) >> 8) > xmax, yopen = (((int64_t)mxy + 255) >> 8) > ymax;
    return 1u | ((uint32_t)dxl << 1) | ((uint32_t)dyl << 2) | (xopen ? 8u : 0u) | (yopen ? 16u : 0u);
}

// Viewport, snap, cull, bbox, edges and planes (Binning.cpp:291-350) for the
// kept attribute slots (NS of them at most). Returns false if the triangle is culled.
template <uint32_t NS>
__device__ __forceinline__ bool setup_prim(Prim& s, const float4 v[3], const float a[3][kWideSlots], uint32_t nslots,
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

    const int64_t area2 = (int64_t)(int32_t)((uint32_t)f2x - (uint32_t)f0x) * (int64_t)(int32_t)((uint32_t)f1y - (uint32_t)f0y) -
                          (int64_t)(int32_t)((uint32_t)f2y - (uint32_t)f0y) * (int64_t)(int32_t)((uint32_t)f1x - (uint32_t)f0x);
    if ((area2 >> 8) <= 0) return false;

    const int32_t mnx = min(min(f0x, f1x), f2x), mny = min(min(f0y, f1y), f2y);
    const int32_t mxx = max(max(f0x, f1x), f2x), mxy = max(max(f0y, f1y), f2y);
    s.xmin = (uint16_t)clampi((int32_t)((uint32_t)mnx + 255u) >> 8, 0, (int32_t)width - 1);
    s.ymin = (uint16_t)clampi((int32_t)((uint32_t)mny + 255u) >> 8, 0, (int32_t)height - 1);
    s.xmax = (uint16_t)clampi((int32_t)((uint32_t)mxx + 255u) >> 8, 0, (int32_t)width - 1);
    s.ymax = (uint16_t)clampi((int32_t)((uint32_t)mxy + 255u) >> 8, 0, (int32_t)height - 1);
    s.tight = tight_bound(area2, mnx, mxx, mny, mxy, s.xmin, s.ymin, s.xmax, s.ymax);

    setup_edge(s, 0, f0x, f0y, f1x, f1y);
    setup_edge(s, 1, f1x, f1y, f2x, f2y);
    setup_edge(s, 2, f2x, f2y, f0x, f0y);

    const float d10x = r1x - r0x, d10y = r1y - r0y, d20x = r2x - r0x, d20y = r2y - r0y;
    const float cterm = d10x * d20y - d10y * d20x;
    setup_plane(cterm, d10x, d10y, d20x, d20y, v[1].z * iw1 - v[0].z * iw0, v[2].z * iw2 - v[0].z * iw0, s.zdx, s.zdy);
    setup_plane(cterm, d10x, d10y, d20x, d20y, iw1 - iw0, iw2 - iw0, s.wdx, s.wdy);
#pragma unroll
    for (uint32_t k = 0; k < NS; ++k)
        if (k < nslots)
            setup_plane(cterm, d10x, d10y, d20x, d20y, a[1][k] * iw1 - a[0][k] * iw0, a[2][k] * iw2 - a[0][k] * iw0,
                        s.adx[k], s.ady[k]);
    s.r0x = r0x;
    s.r0y = r0y;
    s.zf0 = v[0].z * iw0;
    s.wf0 = iw0;
#pragma unroll
    for (uint32_t k = 0; k < NS; ++k) s.af0[k] = k < nslots ? a[0][k] * iw0 : 0.0f;
    return true;
}

// True when a set-up primitive provably covers no pixel, so that it needs no
// record, no bin entry and no back-end work. With a tight bound (closed on
// every side) the reference can accept only pixels of that bound lying in the
// blocks it walks: columns from xmin & ~7 up to, not including, xmax rounded
// up to a multiple of 8 (the exclusive block end), rows likewise. It accepts a
// pixel only when the pixel's three edge values E_k = c_k + dy_k x + dx_k y
// (int32, wrapping) are >= 0: all > 0, or some = 0 and the block test passes.
// So a primitive whose candidate pixels (at most four are tested) all have a
// negative edge value draws nothing.
__device__ __forceinline__ bool covers_nothing(const Prim& s)
{
    if (!(s.tight & 1u) || (s.tight & 0x18u)) return false;  // no bound, or open on a side
    const int32_t x0 = max(s.xmin - (int32_t)((s.tight >> 1) & 1u), s.xmin & ~7);
    const int32_t x1 = min(s.xmax, ((s.xmax + 7) & ~7) - 1);
    const int32_t y0 = max(s.ymin - (int32_t)((s.tight >> 2) & 1u), s.ymin & ~7);
    const int32_t y1 = min(s.ymax, ((s.ymax + 7) & ~7) - 1);
    if (x0 > x1 || y0 > y1) return true;
    if ((x1 - x0 + 1) * (y1 - y0 + 1) > 4) return false;
    for (int32_t y = y0; y <= y1; ++y)
        for (int32_t x = x0; x <= x1; ++x) {
            int32_t m = INT32_MAX;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                m = min(m, (int32_t)((uint32_t)s.c[k] + (uint32_t)s.dy[k] * (uint32_t)x + (uint32_t)s.dx[k] * (uint32_t)y));
            if (m >= 0) return false;
        }
    return true;
}

// 64x64 bin corner test (Binning.cpp:376-409): false if for some edge no
// corner of the bin is strictly inside.
__device__ __forceinline__ bool tile_accepts(const int32_t c[3], const int32_t dx[3], const int32_t dy[3], uint32_t X0,
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

// The bins a primitive lands in. The reference appends the triangle to every
// 64x64 bin of its bbox that passes the corner test -- skipped when the bbox
// spans at most two bin rows, because numXbins is computed from Y
// (Binning.cpp:352-370) -- and later walks 8x8 blocks from (bMin & ~7) up to
// (exclusively) bMax inside each bin (Rasterizer.cpp:203-229). Globally those
// blocks start at xmin & ~7 and end before xmax. A work bin of size S keeps
// the primitive iff one of those blocks lies in it and its 64x64 tile passed.
struct BinWalk {
    uint32_t sbx0, sby0, nx, ncand, k, view_bins_x, bin_base;
    bool check;
};

__device__ __forceinline__ BinWalk bin_walk(const ViewDev& V, uint32_t bin_k, uint32_t bin_log2, int32_t xmin,
                                            int32_t xmax, int32_t ymin, int32_t ymax)
{
    BinWalk w;
    const uint32_t xlo = (uint32_t)xmin & ~7u, ylo = (uint32_t)ymin & ~7u;
    w.k = bin_k;
    w.view_bins_x = V.bins_x;
    w.bin_base = V.bin_base;
    w.check = !((((uint32_t)ymax >> 6) - ((uint32_t)ymin >> 6) + 1u) <= 2u);
    if ((uint32_t)xmax <= xlo || (uint32_t)ymax <= ylo) {
        w.sbx0 = w.sby0 = w.nx = w.ncand = 0;
        return w;
    }
    w.sbx0 = xlo >> bin_log2;
    w.sby0 = ylo >> bin_log2;
    w.nx = (((uint32_t)xmax - 1u) >> bin_log2) - w.sbx0 + 1u;
    const uint32_t ny = (((uint32_t)ymax - 1u) >> bin_log2) - w.sby0 + 1u;
    w.ncand = w.nx * ny;
    return w;
}

// Number of 64x64 tiles the reference bins the primitive into (statistics).
__device__ __forceinline__ uint32_t reference_tile_entries(const Prim& s)
{
    const uint32_t bymin = (uint32_t)s.ymin >> 6, bymax = (uint32_t)s.ymax >> 6;
    const uint32_t bxmin = (uint32_t)s.xmin >> 6, bxmax = (uint32_t)s.xmax >> 6;
    const bool check = !(bymax - bymin + 1u <= 2u);
    uint32_t n = 0;
    for (uint32_t by = bymin; by <= bymax; ++by)
        for (uint32_t bx = bxmin; bx <= bxmax; ++bx) n += (!check || tile_accepts(s.c, s.dx, s.dy, bx * kBin, by * kBin)) ? 1u : 0u;
    return n;
}

// Counts a primitive's bin entries, wave-aggregated, and records each
// (record, bin) pair with its slot in the (part, bin) list: the count
// atomic's old value plus the lane's rank among the wave's lanes adding to
// the same counter. Uniform control flow required; `active` lanes carry a
// primitive described by (c, dx, dy, bbox) in view V.
//
// The wave's (primitive, candidate bin) pairs are compacted: an inclusive
// prefix of the candidate counts numbers them, and each pass hands 64 of them
// to the 64 lanes (the owner of pair `it` is the first lane whose prefix
// exceeds it), so a wave takes ceil(sum / 64) passes rather than as many as
// its largest primitive has candidates.
__device__ __forceinline__ void walk_bins(const FrameDev& F, const ViewDev& V, bool active, const int32_t c[3],
                                          const int32_t dx[3], const int32_t dy[3], int32_t xmin, int32_t xmax,
                                          int32_t ymin, int32_t ymax, uint32_t rec, uint32_t* ht, bool clip)
{
    BinWalk w;
    if (active) {
        w = bin_walk(V, F.bin_k, F.bin_log2, xmin, xmax, ymin, ymax);
    } else {
        w.sbx0 = w.sby0 = w.nx = w.ncand = 0;
        w.view_bins_x = w.bin_base = 0;
        w.check = false;
    }
    const uint32_t lane = lane_id();
    const uint32_t incl = wave_incl_sum(w.ncand);
    const uint32_t total = readlane_u32(incl, 63);
    const uint32_t first = incl - w.ncand;
    const uint32_t pk_org = w.sbx0 | (w.sby0 << 16);                  // bin coordinates < 2^16
    const uint32_t pk_nx = max(w.nx, 1u) | ((w.check ? 1u : 0u) << 31);  // nx < 2^31
    const bool any_check = __ballot(w.check && w.ncand != 0) != 0ull;
    // The wave's pairs take one reservation in its region, issued before the
    // passes so that its round trip overlaps their work: candidate `it` goes
    // to slot pbase + it, and one the corner test rejects leaves a hole
    // (record kNoPair). pair_regions is a power of two.
    const uint32_t region =
        clip ? F.pair_regions + blockIdx.x % kClipPairRegions : blockIdx.x & (F.pair_regions - 1u);
    const uint32_t cap = clip ? F.clip_pair_cap : F.pair_region_cap;
    const uint32_t first_p = clip ? F.pair_regions * F.pair_region_cap + (region - F.pair_regions) * F.clip_pair_cap
                                  : region * F.pair_region_cap;
    uint32_t pres = 0;
    if (lane == 0 && total) pres = atomicAdd(&F.pair_count[region * kPairStride], total);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t it = base + lane;
        const bool valid = it < total;
        const uint32_t m = wave_owner(ht + kHashWords, w.ncand, first, base, carry);
        const int src = (int)(valid && m ? min(m - 1u, 63u) : 0u);
        const uint32_t org = __shfl(pk_org, src, 64), pn = __shfl(pk_nx, src, 64);
        const uint32_t i = it - __shfl(first, src, 64);
        const uint32_t bb = __shfl(w.bin_base, src, 64), vbx = __shfl(w.view_bins_x, src, 64);
        const uint32_t orec = __shfl(rec, src, 64);
        const uint32_t nx = pn & 0x7FFFFFFFu;
        // i / nx and i % nx (i < nx * ny <= 2^22, so i / nx < 2^11): the
        // quotient from the fp32 reciprocal is within 2^-11 of i / nx, so its
        // truncation is off by at most one, which one correction step fixes
        uint32_t q = (uint32_t)((float)i * __builtin_amdgcn_rcpf((float)nx));
        int32_t r = (int32_t)(i - __umul24(q, nx));
        if (r < 0) {
            q -= 1u;
            r += (int32_t)nx;
        } else if (r >= (int32_t)nx) {
            q += 1u;
            r -= (int32_t)nx;
        }
        const uint32_t sbx = (org & 0xFFFFu) + (uint32_t)r, sby = (org >> 16) + q;
        bool act = valid;
        if (any_check) {  // the reference's corner test for primitives spanning 3+ bin rows
            int32_t oc[3], odx[3], ody[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                oc[k] = __shfl(c[k], src, 64);
                odx[k] = __shfl(dx[k], src, 64);
                ody[k] = __shfl(dy[k], src, 64);
            }
            if (act && (pn >> 31) &&
                !tile_accepts(oc, odx, ody, (sbx >> F.bin_klog2) * kBin, (sby >> F.bin_klog2) * kBin))
                act = false;
        }
        const uint32_t key = act ? bb + sby * vbx + sbx : 0u;
        const uint32_t pkey = part_of(orec) * F.total_bins + key;  // the (part, bin) counter
        const uint32_t slot = wave_add_slot_lds(ht, F.part_count, pkey, act);
        // the reservation, read here (pin keeps the read, and so its wait,
        // inside the pass instead of right after the atomic)
        const uint32_t pbase = readlane_u32(pin(pres), 0);
        const uint32_t ri = pbase + it;
        if (valid && ri < cap) {
            const uint32_t p = first_p + ri, np = F.pair_total;
            F.pairs[p] = act ? orec : kNoPair;
            if (act) {
                F.pairs[np + p] = pkey;
                F.pairs[2u * np + p] = slot;
            }
        }
        if (base == 0 && lane == 0 && pbase + total > cap) {
            atomicOr(&F.counters[C_FLAGS], FLAG_ENTRY_OVERFLOW);
            atomicMax(&F.counters[clip ? C_CLIP_PAIR_NEED : C_PAIR_NEED], pbase + total);
        }
    }
}

template <uint32_t NS>
__device__ __forceinline__ void write_record(const FrameDev& F, uint32_t r, const Prim& s, uint32_t draw)
{
    uint4* R = F.rec_raster + (size_t)r * 4;
    R[0] = make_uint4((uint32_t)s.c[0], (uint32_t)s.c[1], (uint32_t)s.c[2], (uint32_t)s.dx[0]);
    R[1] = make_uint4((uint32_t)s.dx[1], (uint32_t)s.dx[2], (uint32_t)s.dy[0], (uint32_t)s.dy[1]);
    // bbox fields: bits 0-13 min, 16-29 max; tight bits: x field 14 valid, 15 min - 1,
    // 30 open above; y field 14 min - 1, 30 open above
    R[2] = make_uint4((uint32_t)s.dy[2],
                      (uint32_t)s.xmin | ((s.tight & 3u) << 14) | ((uint32_t)s.xmax << 16) | (((s.tight >> 3) & 1u) << 30),
                      (uint32_t)s.ymin | (((s.tight >> 2) & 1u) << 14) | ((uint32_t)s.ymax << 16) |
                          (((s.tight >> 4) & 1u) << 30),
                      fbits(s.r0x));
    R[3] = make_uint4(fbits(s.r0y), fbits(s.zdx), fbits(s.zdy), fbits(s.zf0));
    // the 1/w plane and the kept attribute planes, packed: 9 floats for a
    // two-plane frame (NS = 2), 15 otherwise
    float* S = F.rec_shade + (size_t)r * (NS <= 2 ? 9u : 15u);
    S[0] = s.wdx; S[1] = s.wdy; S[2] = s.wf0;
#pragma unroll
    for (uint32_t k = 0; k < (NS <= 2 ? 2u : 4u); ++k) {
        S[3 + 3 * k] = s.adx[k];
        S[4 + 3 * k] = s.ady[k];
        S[5 + 3 * k] = s.af0[k];
    }
    if (NS > kMaxSlots) {
        float4* E = F.rec_ext + (size_t)r * 3;
        E[0] = make_float4(s.adx[4], s.ady[4], s.af0[4], s.adx[5]);
        E[1] = make_float4(s.ady[5], s.af0[5], s.adx[6], s.ady[6]);
        E[2] = make_float4(s.af0[6], s.adx[7], s.ady[7], s.af0[7]);
    }
    F.rec_meta[r] = draw;
}

// FetchIndices / FetchPositions / attribute fetch (Binning.cpp:167-221): the
// object-space positions and the kept attribute slots of the three vertices.
// Returns false (and flags the frame) if an index is out of range.
template <uint32_t NS>
__device__ __forceinline__ bool fetch_vertices(const FrameDev& F, const DrawDev& D, uint32_t tri, float p[3][3],
                                               float a[3][kWideSlots])
{
    uint32_t idx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const size_t i = (size_t)tri * 3u + (size_t)k;
        idx[k] = D.idx_stride == 2 ? (uint32_t)global_ptr((const uint16_t*)D.idx)[i]
               : D.idx_stride == 4 ? global_ptr((const uint32_t*)D.idx)[i]
                                   : (uint32_t)global_ptr(D.idx)[i];
    }
    const bool ok = idx[0] < D.pos_num && idx[1] < D.pos_num && idx[2] < D.pos_num && idx[0] < D.attr_num &&
                    idx[1] < D.attr_num && idx[2] < D.attr_num;
    if (!ok) {
        atomicOr(&F.counters[C_FLAGS], FLAG_INDEX_RANGE);
        return false;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const auto q = global_ptr((const float*)(D.pos + (size_t)idx[k] * D.pos_stride));
        p[k][0] = q[0];
        p[k][1] = q[1];
        p[k][2] = q[2];
        const auto at = global_ptr((const float*)(D.attr + (size_t)idx[k] * D.attr_stride));
#pragma unroll
        for (uint32_t s = 0; s < NS; ++s) a[k][s] = s < D.nslots ? at[D.slot_attr[s]] : 0.0f;
    }
    return true;
}

// fetch_vertices + MVP (Binning.cpp:474-479).
template <uint32_t NS>
__device__ __forceinline__ bool fetch_triangle(const FrameDev& F, const DrawDev& D, uint32_t tri, float4 v[3],
                                               float a[3][kWideSlots])
{
    float p[3][3];
    if (!fetch_vertices<NS>(F, D, tri, p, a)) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = transform(D.mvp, p[k][0], p[k][1], p[k][2]);
    return true;
}

constexpr uint32_t kSetupThreads = 256;

// Per-workgroup statistics accumulator: one global atomic per workgroup.
struct BlockStats {
    uint32_t tile_entries, prims;
};

__device__ __forceinline__ void block_stats_add(BlockStats* bs, uint32_t te, uint32_t pr)
{
    te = wave_sum(te);
    pr = wave_sum(pr);
    if (lane_id() == 0) {
        if (te) atomicAdd(&bs->tile_entries, te);
        if (pr) atomicAdd(&bs->prims, pr);
    }
}

__device__ __forceinline__ void block_stats_flush(const FrameDev& F, BlockStats* bs)
{
    if (threadIdx.x == 0) {
        if (bs->tile_entries) atomicAdd(&F.counters[C_TILE_ENTRIES], bs->tile_entries);
        if (bs->prims) atomicAdd(&F.counters[C_PRIMS], bs->prims);
    }
}

// NS: attribute planes kept per record (kMaxSlots, or kWideSlots in a frame
// that has a draw reading eight varyings).
//
// SHARED: the launch covers groups of several draws (a loop over the draws
// per fetched triangle, which needs more registers); otherwise groups of one.
// The launch takes groups [g_lo, g_hi), its workgroup b being the frame's
// setup workgroup b + block_first.
template <uint32_t NS, bool SHARED>
__global__ __launch_bounds__(kSetupThreads) void k_setup(FrameDev F, const SetupGroupDev* __restrict__ groups,
                                                         const uint32_t* __restrict__ members,
                                                         const DrawDev* __restrict__ draws,
                                                         const ViewDev* __restrict__ views, uint32_t g_lo,
                                                         uint32_t g_hi, uint32_t block_first)
{
    // The tables come in as restrict pointers as well: the kernel's stores
    // cannot change them, so the compiler reads each draw's fields (the
    // matrix above all) into scalar registers even between those stores.
    __shared__ BlockStats bs;
    __shared__ uint32_t s_hash[kSetupThreads / 64 * kWaveScratch];
    uint32_t* ht = s_hash + (threadIdx.x >> 6) * kWaveScratch;
    wave_hash_init(ht);
    const bool stats = (F.flags & FF_STATS) != 0;
    if (stats && threadIdx.x == 0) bs = BlockStats{ 0, 0 };
    if (stats) __syncthreads();
    // the setup group of this workgroup: the last with block_base <= b
    const uint32_t b = blockIdx.x + block_first;
    uint32_t gi = g_lo, hi = g_hi;
    while (hi - gi > 1) {
        const uint32_t mid = (gi + hi) >> 1;
        if (groups[mid].block_base <= b) gi = mid;
        else hi = mid;
    }
    const SetupGroupDev G = groups[gi];
    const uint32_t tri = (b - G.block_base) * kSetupThreads + threadIdx.x;
    const bool valid = tri < G.ntris;
    // the triangle's vertices, fetched once for every draw of the group
    float p[3][3];
    float a[3][kWideSlots];
    const bool fetched = valid && fetch_vertices<NS>(F, draws[members[G.member_base]], tri, p, a);
    const uint32_t nm = SHARED ? G.nmembers : 1u;
#pragma unroll 1
    for (uint32_t m = 0; m < nm; ++m) {
        const uint32_t d = members[G.member_base + m];
        const DrawDev& D = draws[d];
        const ViewDev& V = views[D.view];
        const uint32_t g = D.tri_base + tri;
        Prim s;
        bool set_up = false, survive = false, clipped = false;
        if (fetched) {
            float4 v[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = transform(D.mvp, p[k][0], p[k][1], p[k][2]);
            const uint32_t c0 = clip_codes(v[0]), c1 = clip_codes(v[1]), c2 = clip_codes(v[2]);
            const uint32_t mor = c0 | c1 | c2;
            if (mor == 0) set_up = setup_prim<NS>(s, v, a, D.nslots, V.width, V.height);
            else clipped = (c0 & c1 & c2) == 0;
            // a primitive drawing nothing still counts as binned by the reference
            survive = set_up && !covers_nothing(s);
        }
        // a slot without a primitive is never referenced: nothing is written there
        if (survive) write_record<NS>(F, g, s, d);
        const uint32_t q = wave_add_amount(&F.counters[C_CLIPQ], clipped ? 1u : 0u);
        // at most one slot per triangle: never overflows
        if (clipped) F.clip_queue[q] = make_uint2(g, d);

        if (stats) block_stats_add(&bs, set_up ? reference_tile_entries(s) : 0u, set_up ? 1u : 0u);
        walk_bins(F, V, survive, s.c, s.dx, s.dy, s.xmin, s.xmax, s.ymin, s.ymax, g, ht, false);
    }
    if (stats) {
        __syncthreads();
        block_stats_flush(F, &bs);
    }
}

// kt::Dot / kt::Lerp with the operation order the oracle defines.
__device__ __forceinline__ float clip_dot(int plane, float4 v)
{
    const float px = plane == 0 ? 1.0f : (plane == 1 ? -1.0f : 0.0f);
    const float py = plane == 2 ? 1.0f : (plane == 3 ? -1.0f : 0.0f);
    const float pz = plane == 4 ? 1.0f : (plane == 5 ? -1.0f : 0.0f);
    const float pw = 1.0f;
    return ((px * v.x + py * v.y) + pz * v.z) + pw * v.w;
}
__device__ __forceinline__ float lerp_kt(float a, float b, float t) { return a + (b - a) * t; }

// The clipper's two polygon buffers (up to 9 vertices of 4 position floats
// and NS attribute floats) of every lane of a one-wave workgroup, in LDS:
// component-major with the lane innermost, so a wave's accesses to the same
// component are conflict-free.
constexpr uint32_t kClipThreads = 64;
template <uint32_t NS>
struct ClipPoly {
    static constexpr uint32_t kComps = 4 + NS;
    float* lds;
    uint32_t lane;
    __device__ __forceinline__ float& at(uint32_t buf, uint32_t vtx, uint32_t comp) const
    {
        return lds[((buf * 9u + vtx) * kComps + comp) * kClipThreads + lane];
    }
    __device__ __forceinline__ float4 pos(uint32_t buf, uint32_t vtx) const
    {
        return make_float4(at(buf, vtx, 0), at(buf, vtx, 1), at(buf, vtx, 2), at(buf, vtx, 3));
    }
};

// ClipPlane (Binning.cpp:85-165) over the set planes of `mask`, lowest first,
// stopping when the polygon empties (:517-523). Returns the vertex count; the
// result is in buffer `pin`.
template <uint32_t NS>
__device__ __forceinline__ uint32_t clip_polygon(const ClipPoly<NS>& poly, uint32_t mask, uint32_t& pin)
{
    uint32_t n = 3;
    pin = 0;
    do {
        const int plane = __builtin_ctz(mask);
        mask ^= 1u << plane;
        const uint32_t po = pin ^ 1u;
        uint32_t nout = 0;
        uint32_t i0 = n - 1;
        float d0 = clip_dot(plane, poly.pos(pin, i0));
        for (uint32_t i1 = 0; i1 < n; ++i1) {
            const float d1 = clip_dot(plane, poly.pos(pin, i1));
            const bool in0 = d0 >= 0.0f, in1 = d1 >= 0.0f;
            // a convex polygon gains at most one vertex per plane (9 after six
            // planes); the reference asserts that an emission finds room
            // (Binning.cpp:122, :133, :145), so one that would not -- only
            // rounding on a degenerate polygon could ask for it -- is dropped
            if (in0 && nout < 9u) {
                for (uint32_t c = 0; c < ClipPoly<NS>::kComps; ++c) poly.at(po, nout, c) = poly.at(pin, i0, c);
                ++nout;
            }
            if (in0 != in1 && nout < 9u) {
                uint32_t ia, ib;
                float tt;
                if (in1) { tt = d1 / (d1 - d0); ia = i1; ib = i0; }
                else     { tt = d0 / (d0 - d1); ia = i0; ib = i1; }
                for (uint32_t c = 0; c < ClipPoly<NS>::kComps; ++c)
                    poly.at(po, nout, c) = lerp_kt(poly.at(pin, ia, c), poly.at(pin, ib, c), tt);
                ++nout;
            }
            d0 = d1;
            i0 = i1;
        }
        n = nout;
        pin = po;
    } while (mask && n);
    return n;
}

// Slow path: the queued triangles that cross a clip plane, one per lane of
// one-wave workgroups; the fan triangles (0, i-1, i) (Binning.cpp:525-533) get
// records after the submitted triangles' and bin counts like any other.
template <uint32_t NS>
__global__ __launch_bounds__(kClipThreads) void k_clip(FrameDev F)
{
    __shared__ float s_poly[2 * 9 * ClipPoly<NS>::kComps * kClipThreads];
    __shared__ uint32_t s_hash[kClipThreads / 64 * kWaveScratch];
    uint32_t* ht = s_hash + (threadIdx.x >> 6) * kWaveScratch;
    wave_hash_init(ht);
    const uint32_t n = F.counters[C_CLIPQ];
    const uint32_t stride = gridDim.x * blockDim.x;
    const ClipPoly<NS> poly{ s_poly, threadIdx.x };
    for (uint32_t base_i = blockIdx.x * blockDim.x; base_i < n; base_i += stride) {
        const uint32_t qi = base_i + threadIdx.x;
        const bool active = qi < n;
        uint32_t g = 0, d = 0, nfan = 0, pin = 0;
        if (active) {
            const uint2 e = F.clip_queue[qi];
            g = e.x;
            d = e.y;
            const DrawDev& D = F.draws[d];
            float4 v[3];
            float a[3][kWideSlots];
            fetch_triangle<NS>(F, D, g - D.tri_base, v, a);
            const uint32_t mor = clip_codes(v[0]) | clip_codes(v[1]) | clip_codes(v[2]);
            for (uint32_t k = 0; k < 3; ++k) {
                poly.at(0, k, 0) = v[k].x;
                poly.at(0, k, 1) = v[k].y;
                poly.at(0, k, 2) = v[k].z;
                poly.at(0, k, 3) = v[k].w;
                for (uint32_t t = 0; t < NS; ++t) poly.at(0, k, 4 + t) = a[k][t];
            }
            const uint32_t np = clip_polygon(poly, mor, pin);
            nfan = np >= 3 ? np - 2 : 0;
        }
        __syncthreads();  // polygons are read by other lanes below
        const uint32_t base = wave_add_amount(&F.counters[C_CLIPREC], nfan);
        const bool fits = base + nfan <= F.clip_cap;
        if (nfan && !fits) atomicOr(&F.counters[C_FLAGS], FLAG_CLIP_OVERFLOW);
        if (nfan && fits) F.clip_base[g] = base;
        // the wave's fan triangles, one per lane: lane j takes fan item t of
        // the lane whose inclusive fan count first exceeds t
        const uint32_t lane = lane_id();
        const uint32_t incl = wave_incl_sum(fits ? nfan : 0u);
        const uint32_t total = readlane_u32(incl, 63);
        const uint32_t first = incl - (fits ? nfan : 0u);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t t = t0 + lane;
            uint32_t lo = 0, hi = 63;
#pragma unroll
            for (int step = 0; step < 6; ++step) {
                const uint32_t mid = (lo + hi) >> 1;
                if (__shfl(incl, (int)mid, 64) > t) hi = mid;
                else lo = mid + 1;
            }
            const int src = (int)(t < total ? lo : 0u);
            const uint32_t og = __shfl(g, src, 64), od = __shfl(d, src, 64), opin = __shfl(pin, src, 64);
            const uint32_t obase = __shfl(base, src, 64), j = t - __shfl(first, src, 64);
            const DrawDev& D = F.draws[od];
            const ViewDev& V = F.views[D.view];
            Prim s;
            bool ok = false;
            uint32_t r = 0;
            if (t < total) {
                const ClipPoly<NS> owner{ poly.lds, (uint32_t)src };
                const float4 fv[3] = { owner.pos(opin, 0), owner.pos(opin, j + 1), owner.pos(opin, j + 2) };
                float fa[3][kWideSlots];
                for (uint32_t a = 0; a < NS; ++a) {
                    fa[0][a] = owner.at(opin, 0, 4 + a);
                    fa[1][a] = owner.at(opin, j + 1, 4 + a);
                    fa[2][a] = owner.at(opin, j + 2, 4 + a);
                }
                r = F.ntris + obase + j;
                ok = setup_prim<NS>(s, fv, fa, D.nslots, V.width, V.height);
                if (ok) {
                    write_record<NS>(F, r, s, od);
                    F.clip_order[obase + j] = (og << 3) | j;
                }
            }
            if (F.flags & FF_STATS) {
                const uint32_t te = wave_sum(ok ? reference_tile_entries(s) : 0u);
                const uint32_t pr = wave_sum(ok ? 1u : 0u);
                if (lane == 0 && te) atomicAdd(&F.counters[C_TILE_ENTRIES], te);
                if (lane == 0 && pr) atomicAdd(&F.counters[C_PRIMS], pr);
            }
            walk_bins(F, V, ok, s.c, s.dx, s.dy, s.xmin, s.xmax, s.ymin, s.ymax, r, ht, true);
        }
        __syncthreads();  // before the polygons are overwritten
    }
}

// Exclusive scan of bin counts and the back end's queue (every non-empty bin,
// plus every bin of a view with a clear pending), in three coalesced passes
// over kScanThreads-bin blocks: block totals, a scan of the totals, and the
// block-local scan plus the block's offset.
constexpr uint32_t kScanThreads = 1024;

// Exclusive scan of (a, b) over the workgroup; returns the totals in ta, tb.
// All threads of a kScanThreads workgroup must call it.
__device__ __forceinline__ void block_scan2(uint32_t& a, uint32_t& b, uint32_t& ta, uint32_t& tb)
{
    __shared__ uint32_t s_a[kScanThreads / 64], s_b[kScanThreads / 64];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t ia = a, ib = b;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t ya = __shfl_up(ia, o, 64), yb = __shfl_up(ib, o, 64);
        if (lane >= o) {
            ia += ya;
            ib += yb;
        }
    }
    if (lane == 63) {
        s_a[w] = ia;
        s_b[w] = ib;
    }
    __syncthreads();
    uint32_t oa = 0, ob = 0;
    ta = 0;
    tb = 0;
    for (uint32_t i = 0; i < kScanThreads / 64; ++i) {
        const uint32_t va = s_a[i], vb = s_b[i];
        if (i < w) {
            oa += va;
            ob += vb;
        }
        ta += va;
        tb += vb;
    }
    __syncthreads();
    a = oa + ia - a;
    b = ob + ib - b;
}

__device__ __forceinline__ bool view_cleared(const FrameDev& F, uint32_t bin)
{
    uint32_t lo = 0, hi = F.nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (F.views[mid].bin_base <= bin) lo = mid;
        else hi = mid;
    }
    return (F.views[lo].flags & (VF_CLEAR_COLOUR | VF_CLEAR_DEPTH)) != 0;
}

__device__ __forceinline__ void scan_inputs(const FrameDev& F, uint32_t t, uint32_t& c, uint32_t& q)
{
    c = 0;
    q = 0;
    if (t < F.total_bins) {
#pragma unroll
        for (uint32_t k = 0; k < kBinParts; ++k) c += F.part_count[k * F.total_bins + t];
        q = (c || view_cleared(F, t)) ? 1u : 0u;
    }
}

// Pass 1: per block of bins, the entry and queue totals.
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(FrameDev F, uint32_t* blk)
{
    uint32_t c, q, tc, tq;
    scan_inputs(F, blockIdx.x * kScanThreads + threadIdx.x, c, q);
    block_scan2(c, q, tc, tq);
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = tc;
        blk[2 * blockIdx.x + 1] = tq;
    }
}

// The frame totals: entries (the end of the last bin's list) and queue
// length, the entry arena's overflow check and the clip regions' pair count.
__device__ __forceinline__ void scan_totals(const FrameDev& F, uint32_t tc, uint32_t tq)
{
    F.bin_start[F.total_bins] = tc;
    F.counters[C_ENTRIES] = tc;
    F.counters[C_QLEN] = tq;
    if (tc > F.entry_cap) atomicOr(&F.counters[C_FLAGS], FLAG_ENTRY_OVERFLOW);
    uint32_t cp = 0;
    for (uint32_t i = 0; i < kClipPairRegions; ++i) cp += F.pair_count[(F.pair_regions + i) * kPairStride];
    F.counters[C_CLIP_PAIRS] = cp;
}

// Pass 2 (one workgroup): block offsets and the frame totals.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(FrameDev F, uint32_t* blk, uint32_t nblk)
{
    uint32_t tc = 0, tq = 0;
    for (uint32_t base = 0; base < nblk; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        uint32_t c = 0, q = 0, bc, bq;
        if (i < nblk) {
            c = blk[2 * i];
            q = blk[2 * i + 1];
        }
        block_scan2(c, q, bc, bq);
        if (i < nblk) {
            blk[2 * i] = tc + c;
            blk[2 * i + 1] = tq + q;
        }
        tc += bc;
        tq += bq;
    }
    if (threadIdx.x == 0) scan_totals(F, tc, tq);
}

// Pass 3: bin starts and queue slots.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(FrameDev F, const uint32_t* blk)
{
    const uint32_t t = blockIdx.x * kScanThreads + threadIdx.x;
    uint32_t c, q, tc, tq;
    scan_inputs(F, t, c, q);
    const bool enq = q != 0;
    block_scan2(c, q, tc, tq);
    if (t < F.total_bins) {
        F.bin_start[t] = blk[2 * blockIdx.x] + c;
        if (enq) F.queue[blk[2 * blockIdx.x + 1] + q] = t;
        uint32_t off = 0;  // the bin's parts, one after another
#pragma unroll
        for (uint32_t k = 0; k < kBinParts; ++k) {
            F.part_off[k * F.total_bins + t] = off;
            off += F.part_count[k * F.total_bins + t];
        }
        F.bin_count[t] = off;
    }
}

// The three passes in one workgroup, for frames of at most kScanSingleBins
// bins (a view or a few), which saves two launches where a launch costs
// more than the work: blocks of kScanThreads bins one after another with
// running totals, each bin's part counts loaded once and kept in registers
// for its parts' offsets.
constexpr uint32_t kScanSingleBins = 4u * kScanThreads;
__global__ __launch_bounds__(kScanThreads) void k_scan_single(FrameDev F)
{
    const uint32_t nb = F.total_bins;
    uint32_t run_c = 0, run_q = 0;
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        const uint32_t t = base + threadIdx.x;
        uint32_t cnt[kBinParts], c = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBinParts; ++k) {
            cnt[k] = t < nb ? F.part_count[k * nb + t] : 0u;
            c += cnt[k];
        }
        const bool enq = t < nb && (c || view_cleared(F, t));
        uint32_t sc = c, sq = enq ? 1u : 0u, tc, tq;
        block_scan2(sc, sq, tc, tq);
        if (t < nb) {
            F.bin_start[t] = run_c + sc;
            if (enq) F.queue[run_q + sq] = t;
            uint32_t off = 0;  // the bin's parts, one after another
#pragma unroll
            for (uint32_t k = 0; k < kBinParts; ++k) {
                F.part_off[k * nb + t] = off;
                off += cnt[k];
            }
            F.bin_count[t] = off;
        }
        run_c += tc;
        run_q += tq;
    }
    if (threadIdx.x == 0) scan_totals(F, run_c, run_q);
}

// Longest-first ordering of the back end's queue: bins are bucketed by a
// log-scaled cost key (16 buckets per octave) -- the entry count, or the
// learned model's predicted cost -- and written out from the largest bucket
// down. The order inside a bucket is arbitrary; the image does not depend on
// the order bins are processed in. Three passes, so that a queue of any
// length is ordered by many workgroups: per block of kOrderBlock queue
// entries a bucket histogram (k_order_hist), per bucket the offset of every
// block's run (k_order_offsets, buckets descending, blocks ascending inside a
// bucket), and the scatter (k_order_scatter).
constexpr uint32_t kOrderBuckets = 512;
constexpr uint32_t kOrderBlock = 1024;

__device__ __forceinline__ uint32_t cost_bucket(float c)
{
    if (!(c > 1.0f)) return 0u;
    const float b = __log2f(c) * 16.0f;  // scheduling key only; never touches the image
    return b >= (float)(kOrderBuckets - 1) ? kOrderBuckets - 1 : 1u + (uint32_t)b;
}

__global__ __launch_bounds__(kOrderBlock) void k_order_hist(FrameDev F, const uint32_t* in_queue, const float* pred,
                                                           uint16_t* bkt, uint32_t* hist)
{
    __shared__ uint32_t h[kOrderBuckets];
    const uint32_t n = F.counters[C_QLEN], base = blockIdx.x * kOrderBlock;
    if (base >= n) return;  // the whole workgroup
    if (threadIdx.x < kOrderBuckets) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
        const uint32_t gb = in_queue[i];
        const uint32_t b = cost_bucket(pred ? pred[gb] : (float)F.bin_count[gb]);
        bkt[i] = (uint16_t)b;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kOrderBuckets) hist[blockIdx.x * kOrderBuckets + threadIdx.x] = h[threadIdx.x];
}

// One workgroup, a thread per bucket: bucket totals over the blocks, their
// exclusive scan from the largest bucket down, then each block's offset.
__global__ __launch_bounds__(kOrderBuckets) void k_order_offsets(FrameDev F, uint32_t* hist)
{
    __shared__ uint32_t s_w[kOrderBuckets / 64];
    const uint32_t n = F.counters[C_QLEN], nblk = (n + kOrderBlock - 1) / kOrderBlock;
    const uint32_t b = kOrderBuckets - 1u - threadIdx.x;  // thread 0 holds the largest bucket
    uint32_t total = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < nblk; ++k) total += hist[k * kOrderBuckets + b];
    // exclusive scan of the totals over the threads (waves of 64, then the wave sums)
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = total;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t run = incl - total;
    for (uint32_t k = 0; k < w; ++k) run += s_w[k];
#pragma unroll 8
    for (uint32_t k = 0; k < nblk; ++k) {
        const uint32_t c = hist[k * kOrderBuckets + b];
        hist[k * kOrderBuckets + b] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kOrderBlock) void k_order_scatter(FrameDev F, const uint32_t* in_queue,
                                                              const uint16_t* bkt, const uint32_t* hist)
{
    __shared__ uint32_t off[kOrderBuckets];
    const uint32_t n = F.counters[C_QLEN], base = blockIdx.x * kOrderBlock;
    if (base >= n) return;  // the whole workgroup
    if (threadIdx.x < kOrderBuckets) off[threadIdx.x] = hist[blockIdx.x * kOrderBuckets + threadIdx.x];
    __syncthreads();
    const uint32_t i = base + threadIdx.x;
    if (i < n) F.queue[atomicAdd(&off[bkt[i]], 1u)] = in_queue[i];
}

// Bin lists from the pairs of the count walks: entry (bin start + the part's
// offset in the bin + the pair's slot) = record. An overflowed frame (pairs or
// entries past the arena) is rendered again with larger arenas.
//
// grid: (regions, chunks); workgroup (r, c) takes chunks c, c + gridDim.y, ...
// of `chunk` pairs (a multiple of 256) of region r. Workgroups are dispatched
// x fastest, so
// chunk c of every region is written before chunk c + 1 of any: the regions
// fill at the same rate (the set-up's workgroups take them in turn), so the
// pairs in flight were counted at about the same time, and a bin part's
// entries -- whose slots follow the order of counting -- are written close
// together in time, filling cache lines before they leave the cache.
__global__ __launch_bounds__(256) void k_scatter(FrameDev F, uint32_t chunk)
{
    if (F.counters[C_FLAGS] & (FLAG_ENTRY_OVERFLOW | FLAG_CLIP_OVERFLOW)) return;
    const uint32_t region = blockIdx.x;
    const uint32_t n = F.pair_count[region * kPairStride];  // within the region's capacity without overflow
    const uint32_t first_p = region < F.pair_regions
                                 ? region * F.pair_region_cap
                                 : F.pair_regions * F.pair_region_cap + (region - F.pair_regions) * F.clip_pair_cap;
    const uint32_t np = F.pair_total;
    const uint32_t nchunks = n / chunk + (n % chunk != 0u);
    for (uint32_t c = blockIdx.y; c < nchunks; c += gridDim.y) {
        for (uint32_t k = threadIdx.x; k < chunk; k += 256) {
            const uint32_t i = c * chunk + k;  // < n + chunk: no wrap for n < 2^32 - chunk
            if (i >= n) break;
            const uint32_t p = first_p + i;
            const uint32_t rec = F.pairs[p];
            if (rec == kNoPair) continue;  // a candidate bin the corner test rejected
            const uint32_t pkey = F.pairs[np + p], slot = F.pairs[2u * np + p];
            const uint32_t key = pkey - part_of(rec) * F.total_bins;
            F.entries[F.bin_start[key] + F.part_off[pkey] + slot] = rec;
        }
    }
}

}  // namespace srg
