// sr_frame.h -- per-frame device tables shared by the front-end and back-end
// kernels, and the wave-level helpers they use.
#pragma once
#include "sr_device.h"

namespace srg {

// One entry of the frame's draw table: a DrawCall (Renderer.h:119-150)
// resolved to device pointers, plus the attribute slots its shader reads.
struct DrawDev {
    const uint8_t* idx;
    const uint8_t* pos;
    const uint8_t* attr;
    uint32_t idx_stride, pos_stride, attr_stride, ntris;
    uint32_t pos_num, attr_num, tri_base, view;
    uint32_t shader, tex, nslots, deriv_slot;  // deriv_slot: first slot of (u, v) at uv_offset
    uint32_t slot_attr[kWideSlots];
    uint32_t pad_b;
    uint32_t tex_info;    // the texture's max(wlog2, hlog2) | nmips << 8 (0 without one): for scheduling features
    uint32_t pad[2];
    float mvp[16];
};

// SponzaScene::Constants as the SponzaShader reads it (Viewer/SponzaScene.h:15-47):
// one float per m_sunDir / m_ambCol lane-broadcast component, 16 point lights.
constexpr uint32_t kSponzaLights = 16;
struct SponzaLightDev {
    float pos[3], colour[3], intensity, falloff;
};
struct SponzaDev {
    float sun_dir[3], ambient[3];
    SponzaLightDev lights[kSponzaLights];
};

enum : uint32_t {
    VF_CLEAR_COLOUR = 1u,
    VF_CLEAR_DEPTH = 2u,
    VF_HAS_COLOUR = 4u,
};

// A framebuffer plane targeted this frame. Work items ("bins") are
// S x S squares, S = 64 / bin_k, inside the reference's 64x64 tiles.
struct ViewDev {
    uint32_t* colour;
    float* depth;
    uint32_t* primid;
    uint32_t width, height, tiles_x, tiles_y;
    uint32_t bins_x, bins_y, bin_base, flags;
    uint32_t clear_colour, pad0, pad1, pad2;
};

enum : uint32_t {
    C_CLIPQ = 0,       // clipped triangles queued for the slow path
    C_CLIPREC = 1,     // records allocated by the slow path
    C_FLAGS = 2,       // error / overflow bits
    C_QLEN = 3,        // back-end queue length
    C_QHEAD = 4,       // back-end queue head
    C_ENTRIES = 5,     // bin entries
    C_TILE_ENTRIES = 6,// 64x64-tile entries as the reference counts them (stats)
    C_PRIMS = 7,       // primitives that survived the cull (stats)
    C_FRAGS = 8,       // 64-bit rasterised fragment count (words 8, 9)
    C_SCHED_ERR = 10,  // float: summed |log2 predicted - log2 measured| bin cost
    C_SCHED_N = 11,    // bins recorded by the scheduler this frame
    C_PAIR_NEED = 12,  // on a pair region overflow: the largest count a region reached
    C_CLIP_PAIR_NEED = 13, // on an overflow of k_clip's pair region: the count it reached
    C_CLIP_PAIRS = 14, // pairs k_clip found, summed over its regions
    C_NCOUNTERS = 16,
};

enum : uint32_t {
    FLAG_CLIP_OVERFLOW = 1u,
    FLAG_ENTRY_OVERFLOW = 2u,
    FLAG_INDEX_RANGE = 4u,
};

enum : uint32_t { FF_STATS = 1u, FF_PRIMID = 2u, FF_TILE_COSTS = 4u };

constexpr uint32_t kFanUnclipped = 7u;  // fan field of an unclipped triangle's order key
constexpr uint32_t kNoPair = 0xFFFFFFFFu;  // record field of a pair slot left empty

// Primitive records. Index space: [0, ntris) one slot per submitted triangle
// (its unclipped primitive), [ntris, ntris + clip_cap) fan triangles of
// clipped triangles. Order key of a primitive: (global triangle << 3) | fan,
// fan = 7 for an unclipped triangle, 0..6 for fan triangles.
// Draws that share their geometry -- index, position and attribute buffers,
// strides, counts and kept attribute slots, as the views of a batch drawing
// one scene do: k_setup fetches each triangle once for all of them and sets
// it up once per draw. A draw whose geometry no other draw shares is a
// group of one.
struct SetupGroupDev {
    uint32_t block_base;   // first k_setup workgroup of the group (kSetupThreads triangles each)
    uint32_t ntris;
    uint32_t member_base;  // the group's draws: group_members[member_base, member_base + nmembers)
    uint32_t nmembers;
};
constexpr uint32_t kMaxGroupMembers = 16;

struct FrameDev {
    const DrawDev* draws;
    const ViewDev* views;
    const TexDev* texs;
    const SetupGroupDev* groups;
    const uint32_t* group_members;  // draw indices, group by group
    uint32_t ngroups;
    uint32_t ndraws, nviews, total_bins, ntris;
    uint32_t bin_k, bin_size, clip_cap, entry_cap;
    uint32_t bin_log2, bin_klog2;  // log2 of bin_size and of bin_k (both powers of two)
    uint4* rec_raster;     // 4 x uint4 per record: edges, bbox, r0, z plane
    float* rec_shade;      // shade_stride floats per record: 1/w plane, then attribute planes 0-1 (stride 9)
                           // or 0-3 (stride 15), each plane (dx, dy, value at v0)
    uint32_t shade_stride;
    float4* rec_ext;       // 3 x float4 per record: attribute planes 4-7 (wide frames only)
    const SponzaDev* sponza;  // SponzaShader constants (wide frames only)
    const uint32_t* sched_gate;  // learned order: [1] != 0 on frames that predict and record (else null)
    uint32_t* rec_meta;    // draw index (written only for slots that hold a primitive)
    uint2* clip_queue;     // per triangle routed to the clipper: (global triangle id, draw index)
    uint32_t* clip_base;   // per global triangle: first fan record (relative to ntris)
    uint32_t* clip_order;  // per fan record: order key
    uint32_t* bin_count;   // entries per bin (written by the scan)
    uint32_t* bin_start;   // total_bins + 1
    uint32_t* part_count;  // kBinParts x total_bins: entries per (part, bin); the count walks' slot cursors
    uint32_t* part_off;    // kBinParts x total_bins: first entry of a part inside its bin's list
    uint32_t* entries;
    // (record, part-bin key, slot in the part's list) of every entry, as the
    // count walks find them: three arrays of pair_total words, one after
    // another; k_scatter writes entries from them after the scan. Region r <
    // pair_regions (a power of two; k_setup's workgroup index mod
    // pair_regions) fills
    // [r * pair_region_cap, (r + 1) * pair_region_cap) through its own
    // counter pair_count[r * kPairStride]; k_clip's workgroups fill the
    // kClipPairRegions regions after them (workgroup index mod
    // kClipPairRegions), clip_pair_cap words each. k_setup's regions
    // therefore get the same pairs every frame; k_clip takes its triangles in
    // queue order, so after a frame whose clip regions could not each hold
    // all of its clip pairs the host sizes them so (C_CLIP_PAIRS).
    uint32_t* pairs;
    uint32_t* pair_count;
    uint32_t pair_regions, pair_region_cap, clip_pair_cap, pair_total;
    uint32_t* queue;
    uint64_t* bin_cost;    // optional per bin (s_memtime ticks)
    uint32_t* counters;
    uint32_t flags;
};

// Pair regions: one counter per region on its own 64-byte line, so that the
// walks of different workgroups do not contend on one address. A frame uses
// as many regions as k_setup has workgroups, at most kPairRegions, so that
// every region holds a share of the frame's pairs.
constexpr uint32_t kPairRegions = 256, kClipPairRegions = 8, kPairStride = 16;

// A bin's counter and append cursor are split into kBinParts parts, the part
// chosen by the entry's record (64 consecutive records share one), so the
// atomics of a bin that many primitives land in spread over kBinParts
// addresses. The scan sums the parts and lays them out one after another
// inside the bin's entry list.
constexpr uint32_t kBinParts = 16;
__host__ __device__ constexpr uint32_t part_of(uint32_t rec) { return (rec >> 6) & (kBinParts - 1u); }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// v, computed here into a VGPR. An empty asm statement that takes and
// returns the register: the compiler can neither recompute v next to each
// use nor defer the wait for the loads v depends on into the conditional
// blocks that use it.
__device__ __forceinline__ uint32_t pin(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ unsigned long long pin(unsigned long long v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Per-wave LDS hash for aggregating atomics by key, without a serial pass
// per distinct key: kHashSlots keys, then their counts, then their bases.
// Every slot is empty (key kHashEmpty, count 0) between uses. A wave has at
// most 64 distinct keys, so linear probing in 128 slots always ends.
constexpr uint32_t kHashSlots = 128, kHashWords = 3 * kHashSlots, kHashEmpty = 0xFFFFFFFFu;
// A wave's LDS scratch in the front end: the hash, then 64 owner marks.
constexpr uint32_t kWaveScratch = kHashWords + 64;

__device__ __forceinline__ void wave_hash_init(uint32_t* t)
{
    for (uint32_t i = lane_id(); i < kHashSlots; i += 64) {
        t[i] = kHashEmpty;
        t[kHashSlots + i] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Slot of `key` in the wave's table (claimed by compare-and-swap, probing
// onward past other keys) and this lane's rank among the lanes adding to it.
// Keys must differ from kHashEmpty. The probe count is bounded: a lane that
// finds no slot (impossible while every pass empties the slots it used)
// gets kHashSlots and adds on its own.
__device__ __forceinline__ uint32_t wave_hash_insert(uint32_t* t, uint32_t key, uint32_t& rank)
{
    uint32_t h = (key * 0x9E3779B1u) >> 25;
    for (uint32_t probe = 0; probe < kHashSlots; ++probe) {
        const uint32_t prev = atomicCAS(&t[h], kHashEmpty, key);
        if (prev == kHashEmpty || prev == key) {
            rank = atomicAdd(&t[kHashSlots + h], 1u);
            return h;
        }
        h = (h + 1u) & (kHashSlots - 1u);
    }
    rank = 0;
    return kHashSlots;
}

// A wave's LDS operations complete in issue order, so once every lane's
// increment is issued, the rank-0 lane of a key reads the key's full count.
__device__ __forceinline__ uint32_t wave_add_slot_lds(uint32_t* t, uint32_t* counters, uint32_t key, bool active)
{
    uint32_t h = kHashSlots, rank = 0;
    if (active) h = wave_hash_insert(t, key, rank);
    const bool in_table = h < kHashSlots;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t own = 0;
    if (active && !in_table) own = atomicAdd(&counters[key], 1u);
    if (in_table && rank == 0) t[2 * kHashSlots + h] = atomicAdd(&counters[key], t[kHashSlots + h]);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t slot = in_table ? t[2 * kHashSlots + h] + rank : own;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (in_table && rank == 0) {
        t[h] = kHashEmpty;
        t[kHashSlots + h] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return slot;
}

// Inclusive max over the wave's lanes (DPP: row shifts, then row broadcasts).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

// Inclusive sum over the wave's lanes, the same DPP pattern.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Owner of work item `base + lane` when every lane owns `n` consecutive
// items starting at `first` (an exclusive prefix of the n): 1 + the owning
// lane, or 0 past the last item. Lanes whose items start in this pass of 64
// mark their start in the wave's 64-word LDS array; an inclusive max over
// the marks and `carry` (the value of the previous pass's last item, 0 on
// the first pass) names each item's owner. Passes must run in order with
// base = 0, 64, 128, ...; uniform control flow required.
__device__ __forceinline__ uint32_t wave_owner(uint32_t* mark, uint32_t n, uint32_t first, uint32_t base,
                                               uint32_t& carry)
{
    const uint32_t lane = lane_id();
    mark[lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (n && first >= base && first - base < 64u) mark[first - base] = lane + 1u;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t m = max(wave_incl_max(mark[lane]), carry);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    carry = readlane_u32(m, 63);
    return m;
}

// wave_owner that also hands each item its owner's first item: an owner's
// mark is (first << 7) | (lane + 1). Owners' first items strictly increase
// with the lane, so the marks do too and the same max-scan yields both:
// owner = (m & 127) - 1, first = m >> 7 (first < 2^25). Same contract as
// wave_owner.
__device__ __forceinline__ uint32_t wave_owner_first(uint32_t* mark, uint32_t n, uint32_t first, uint32_t base,
                                                     uint32_t& carry)
{
    const uint32_t lane = lane_id();
    mark[lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (n && first >= base && first - base < 64u) mark[first - base] = (first << 7) | (lane + 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t m = max(wave_incl_max(mark[lane]), carry);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    carry = readlane_u32(m, 63);
    return m;
}

// Sum over the wave (result in every lane).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_min_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-aggregated add of per-lane amounts to one counter; returns each lane's
// exclusive offset. Uniform control flow required.
__device__ __forceinline__ uint32_t wave_add_amount(uint32_t* counter, uint32_t amount)
{
    const uint32_t x = wave_incl_sum(amount);
    const uint32_t total = readlane_u32(x, 63);
    uint32_t base = 0;
    if (lane_id() == 0 && total) base = atomicAdd(counter, total);
    return readlane_u32(base, 0) + x - amount;
}

}  // namespace srg
