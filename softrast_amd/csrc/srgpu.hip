// srgpu.hip -- MI355X (gfx950) sort-middle rasteriser behind the C ABI of
// include/srgpu.h.
//
// A frame is six launches on one stream:
//   k_frame_init  zero counters and per-bin arrays (a fill kernel, no memset)
//   k_setup       one lane per submitted triangle (sr_frontend.h)
//   k_clip        slow path for triangles crossing a clip plane
//   k_scan_*      bin offsets and the back end's work queue (three passes)
//   k_scatter     bin lists from the (record, bin) pairs the count walks found
//   k_backend<S>  persistent raster + resolve + shade per S x S bin (sr_backend.h)
//
// Visibility is resolved per pixel as the max of (depth bits << 32 |
// 0xFFFFFFFE - primitive order) in LDS; that is the fragment SoftRast's
// in-order strict reverse-Z test followed by in-order shading leaves in the
// pixel, so bin entries may arrive in any order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "srgpu.h"
#include "sr_sched.h"
#include "sr_texture.h"
#include "sr_obj.h"

namespace srg {

// Frame-start fill: counters, per-(part, bin) counts, per-bin costs
// (graph-safe). With the learned order, gate[1] says whether this frame
// predicts and records (every `every`-th frame, counted in gate[0] on the
// device so that replayed graphs count too).
// Counter words of a frame: the C_NCOUNTERS frame counters (copied back to
// the host), then from word 64 the counters of the pair regions (at most
// kPairRegions for k_setup, kClipPairRegions more for k_clip).
constexpr uint32_t kPairCounterBase = 64,
                   kCounterWords = kPairCounterBase + (kPairRegions + kClipPairRegions) * kPairStride;

// Entry-arena size at which each k_clip pair region (entries / 2 /
// kClipPairRegions pairs) holds `clip_pairs` pairs with a quarter to spare.
static uint64_t clip_pair_want(uint32_t clip_pairs)
{
    return (uint64_t)clip_pairs * 2 * kClipPairRegions * 5 / 4;
}

__global__ void k_frame_init(uint32_t* counters, uint32_t* part_count, uint64_t* bin_cost, uint32_t nbins,
                             uint32_t* gate, uint32_t every)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (gate && i == 0) {
        gate[1] = (every <= 1u || gate[0] % every == 0u) ? 1u : 0u;
        gate[0] = gate[0] + 1u;
    }
    for (uint32_t t = i; t < kCounterWords; t += gridDim.x * blockDim.x) counters[t] = 0;
    for (uint32_t t = i; t < nbins * kBinParts; t += gridDim.x * blockDim.x) {
        part_count[t] = 0;
        if (t < nbins) bin_cost[t] = 0;
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
// With bgra, red and blue are exchanged on the way (the present swizzle of
// Window_Win32::Flip, Viewer/Platform/Window_Win32.cpp:85-112).
__global__ void k_detile(const uint32_t* tiles, uint32_t* linear, uint32_t width, uint32_t height, uint32_t tiles_x,
                         uint32_t bgra)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= width || y >= height) return;
    const uint32_t t = (y >> 6) * tiles_x + (x >> 6);
    uint32_t v = tiles[(size_t)t * kTilePixels + (y & 63u) * 64u + (x & 63u)];
    if (bgra) v = (v & 0xFF00FF00u) | ((v >> 16) & 0xFFu) | ((v & 0xFFu) << 16);
    linear[(size_t)y * width + x] = v;
}

}  // namespace srg

// Scratch memory of srg_end_frame_noalloc: the per-frame arenas of such a
// frame (tables, bin arrays, records, entry lists) are carved from this array
// instead of the context's own device allocations. Each GPU holds one copy,
// from the library's first use on it until the process ends. Host code never
// names it in a kernel argument or copy: it uses the device address that
// hipGetSymbolAddress returned on that GPU (scratch_address below).
static __device__ __align__(256) uint8_t scratch[SRG_NOALLOC_SCRATCH_BYTES];

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

// An asynchronous blit: the host function copies the pinned rows to the
// caller's buffer and calls its callback (BlitJobFn's tail, Renderer.cpp:341-346).
struct BlitJob {
    void* dst = nullptr;
    const void* src = nullptr;
    size_t bytes = 0;
    void (*cb)(void*) = nullptr;
    void* user = nullptr;
};

// Back-end work layout: bin edge in pixels and workgroups per bin.
struct BackendLayout {
    uint32_t bin_size, parts;
};

// Autotuner: every layout of the sweep is timed over a few frames of the same
// frame signature (views and triangle counts), then the fastest is kept until
// the signature changes. Layouts change only the schedule, never a pixel.
struct TuneState {
    bool sweeping = false, locked = false;
    uint64_t sig = 0;
    uint32_t cand = 0, frame = 0;
    float best_ms[8] = {};
    BackendLayout chosen{ 32, 1 };
};

struct PlaneState {
    uint32_t* colour = nullptr;
    float* depth = nullptr;
    uint32_t* primid = nullptr;
    bool pending = false;
    int clear_colour = 0, clear_depth = 0;
    // asynchronous blit of this plane
    uint32_t* pinned = nullptr;      // host rows written by the de-tile kernel
    hipEvent_t read_done = nullptr;  // the plane's tiles have been read
    hipEvent_t blit_done = nullptr;  // the callback has returned
    bool blit_pending = false;
    BlitJob job;
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
    // k_clip workgroups resident on the whole device at once, for two, four
    // and eight kept planes: its grid, which strides over the clip queue, is
    // no larger, so that every workgroup runs in the first round
    uint32_t clip_resident[3] = { 2048, 2048, 2048 };
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // asynchronous blits
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
    DevArray<float> rec_shade;
    DevArray<float4> rec_ext;         // attribute planes 4-7, frames with a SponzaShader draw
    SponzaDev sponza{};               // SponzaShader constants (srg_set_sponza_constants)
    DevArray<uint32_t> rec_meta, clip_base, clip_order;
    DevArray<uint2> clip_queue;
    DevArray<uint32_t> bin_count, bin_start, part_count, part_off, queue, queue2, entries, pairs, scan_blk;
    DevArray<uint16_t> order_bkt;   // per queue entry: its cost bucket
    DevArray<uint32_t> order_hist;  // per (queue block, bucket): count, then offset
    DevArray<uint64_t> bin_cost;
    DevArray<uint32_t> counters;
    DevArray<uint32_t> blit_staging;
    uint32_t* h_counters = nullptr;
    uint32_t clip_cap = 0, entry_cap = 0;
    uint32_t last_bins = 0;
    hipGraphExec_t graph_exec = nullptr;
    // learned tile-cost scheduler
    void* sched_params = nullptr;
    void* sched_adam = nullptr;
    SchedSample* sched_ring = nullptr;
    uint32_t* sched_head = nullptr;
    uint32_t* sched_gate = nullptr;   // [0] frame count, [1] this frame predicts and records
    float sched_last_err = 0.0f;      // error of the last frame that recorded
    bool sched_force_record = true;   // the next frame predicts and records (the model just changed)
    uint64_t sched_layout = 0;        // bin layout (views' bin ranges) of the last learned-order frame
    DevArray<float4> sched_feat;
    DevArray<float> sched_pred;
    bool sched_trained = false;
    bool sched_fallback = false;  // last measured error above kSchedFallbackErr: entry-count order
    uint32_t sched_frames = 0, sched_last_train = 0;
    uint32_t sched_bin_size = 0;  // bin size the recorded samples and the model belong to
    std::unordered_map<uint64_t, TuneState> tunes;  // autotuner, per frame signature
    uint64_t tune_sig = 0;                          // signature of the last frame
    std::vector<uint8_t> graph_sig;

    hipEvent_t ev[5] = {};  // frame start, back-end start, frame end, blit, back-end end
    srg_frame_stats stats{};
    void* scratch_dev = nullptr;  // this device's `scratch` (srg_end_frame_noalloc)
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

// Device address of `scratch` on each GPU, looked up once per device (with
// that device current) by srg_create.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return SRG_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SRG_ERR_HIP;
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return SRG_OK;
}

// Bump allocation of 256-byte-aligned pieces of a fixed block (the scratch).
struct Carver {
    uint8_t* base;
    size_t cap;
    size_t used = 0;
    bool ok = true;
    template <typename T>
    T* take(size_t n)
    {
        const size_t o = align_up(used, 256), b = n * sizeof(T);
        if (o > cap || b > cap - o) {
            ok = false;
            return nullptr;
        }
        used = o + b;
        return (T*)(base + o);
    }
    size_t left() const { return align_up(used, 256) >= cap ? 0 : cap - align_up(used, 256); }
};

// Work on the library stream that touches a plane waits until an
// asynchronous blit of it has read its tiles.
void stream_wait_blit_read(srg_ctx* ctx, PlaneState& p)
{
    if (p.blit_pending) (void)hipStreamWaitEvent(ctx->stream, p.read_done, 0);
}

// Host side: the blit of this plane, callback included, has finished.
int finish_blit(PlaneState& p)
{
    if (!p.blit_pending) return SRG_OK;
    p.blit_pending = false;
    return hipEventSynchronize(p.blit_done) == hipSuccess ? SRG_OK : SRG_ERR_HIP;
}

void blit_host_fn(void* arg)
{
    const BlitJob* j = (const BlitJob*)arg;
    std::memcpy(j->dst, j->src, j->bytes);
    if (j->cb) j->cb(j->user);
}

int flush_pending_clear(srg_ctx* ctx, srg_fb* fb, uint32_t plane)
{
    PlaneState& ps = fb->planes[plane];
    stream_wait_blit_read(ctx, ps);
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
int shader_slots(uint32_t shader, uint32_t uv_offset, uint32_t apt, uint32_t slot_attr[kWideSlots], uint32_t& nslots,
                 uint32_t& deriv_slot)
{
    for (uint32_t k = 0; k < kWideSlots; ++k) slot_attr[k] = 0;
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
    case SRG_SHADER_SPONZA:
        // every varying of an Obj::Vertex (32 bytes): slot k = attribute k
        if (apt != 8 || uv_offset + 1 >= apt) return SRG_ERR_INVALID;
        for (uint32_t k = 0; k < kWideSlots; ++k) slot_attr[k] = k;
        nslots = kWideSlots;
        deriv_slot = uv_offset;
        return SRG_OK;
    default:
        return SRG_ERR_UNKNOWN_SHADER;
    }
}

// Shaders that read their uniform as a texture (UnlitDiffuseShader, SponzaShader).
bool shader_samples(uint32_t shader) { return shader == SRG_SHADER_UNLIT_DIFFUSE || shader == SRG_SHADER_SPONZA; }

struct ViewKey {
    srg_fb* fb;
    uint32_t plane;
};

constexpr uint32_t kSchedRing = 1u << 18;  // recorded (features, cost) samples

int ensure_sched(srg_ctx* ctx)
{
    if (ctx->sched_params) return SRG_OK;
    if (hipMalloc(&ctx->sched_params, sizeof(MlpParams)) != hipSuccess ||
        hipMalloc(&ctx->sched_adam, sizeof(AdamState)) != hipSuccess ||
        hipMalloc((void**)&ctx->sched_ring, sizeof(SchedSample) * kSchedRing) != hipSuccess ||
        hipMalloc((void**)&ctx->sched_head, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&ctx->sched_gate, 2 * sizeof(uint32_t)) != hipSuccess)
        return SRG_ERR_NOMEM;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->sched_gate, 0u, (size_t)2);
    hipLaunchKernelGGL(k_mlp_init, dim3((kMlpParams + 255) / 256), dim3(256), 0, ctx->stream,
                       (MlpParams*)ctx->sched_params, (AdamState*)ctx->sched_adam, 0x5EED5C4Du);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->sched_head, 0u, (size_t)1);
    HIP_TRY(hipGetLastError());
    return SRG_OK;
}

// Mean |log2 predicted - log2 measured| bin cost above which the learned
// order gives way to the entry-count order (a factor of 2 off on average).
constexpr float kSchedFallbackErr = 1.0f;
// Once trained, the model's features, prediction and cost recording run on
// every kSchedEvery-th frame; the frames between order their queue by the
// last prediction (bins keep their numbers across frames of the same views).
constexpr uint32_t kSchedEvery = 16;

// A new bin size: the recorded samples and the model describe other bins, so
// the model starts again from its initial weights and an empty ring.
int reset_sched(srg_ctx* ctx)
{
    hipLaunchKernelGGL(k_mlp_init, dim3((kMlpParams + 255) / 256), dim3(256), 0, ctx->stream,
                       (MlpParams*)ctx->sched_params, (AdamState*)ctx->sched_adam, 0x5EED5C4Du);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->sched_head, 0u, (size_t)1);
    HIP_TRY(hipGetLastError());
    ctx->sched_trained = false;
    ctx->sched_fallback = false;
    ctx->sched_force_record = true;
    ctx->sched_last_err = 0.0f;
    ctx->sched_frames = 0;
    ctx->sched_last_train = 0;
    return SRG_OK;
}

// After a frame of SRG_ORDER_LEARNED: report the model's error and train it
// once enough frames are recorded (and again every retrain interval).
int sched_after_frame(srg_ctx* ctx)
{
    ctx->sched_force_record = false;
    const uint32_t n = ctx->h_counters[C_SCHED_N];
    float err;
    std::memcpy(&err, &ctx->h_counters[C_SCHED_ERR], sizeof(err));
    // a trained model predicts and records every kSchedEvery-th frame; the
    // other frames keep the last measured error and fallback decision
    if (ctx->sched_trained && n) {
        ctx->sched_last_err = err / (float)n;
        ctx->sched_fallback = ctx->sched_last_err > kSchedFallbackErr;
    }
    ctx->stats.sched_err = ctx->sched_trained ? ctx->sched_last_err : 0.0f;
    ctx->sched_frames++;
    const uint32_t warm = ctx->cfg.sched_warmup_frames ? ctx->cfg.sched_warmup_frames : 8u;
    uint32_t steps = 0;
    float lr = 0.0f;
    if (!ctx->sched_trained && ctx->sched_frames >= warm) {
        steps = 600;
        lr = 3e-3f;
    } else if (ctx->sched_trained && ctx->cfg.sched_retrain_interval &&
               ctx->sched_frames - ctx->sched_last_train >= ctx->cfg.sched_retrain_interval) {
        steps = 150;
        lr = 1e-3f;
    }
    if (steps) {
        hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kTrainBatch), 0, ctx->stream, (MlpParams*)ctx->sched_params,
                           (AdamState*)ctx->sched_adam, (const SchedSample*)ctx->sched_ring,
                           (const uint32_t*)ctx->sched_head, kSchedRing, steps, lr, ctx->sched_frames);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->sched_trained = true;
        ctx->sched_force_record = true;
        ctx->sched_last_train = ctx->sched_frames;
        ctx->stats.sched_trainings++;
    }
    ctx->stats.sched_trained = ctx->sched_trained ? 1u : 0u;
    ctx->stats.sched_fallback = ctx->sched_fallback ? 1u : 0u;
    return SRG_OK;
}

// Back-end layouts: work-bin edge S and workgroups per bin (each owning a
// strip of S x S/k pixels). An 8x8 bin is one wave; every other strip holds
// at least one pixel per lane of four waves.
constexpr BackendLayout kLayouts[] = { { 8, 1 }, { 16, 1 }, { 32, 1 }, { 32, 2 }, { 32, 4 }, { 64, 1 }, { 64, 2 }, { 64, 4 } };
constexpr uint32_t kNumLayouts = sizeof(kLayouts) / sizeof(kLayouts[0]);

bool layout_valid(uint32_t bin_size, uint32_t parts)
{
    for (const BackendLayout& l : kLayouts)
        if (l.bin_size == bin_size && l.parts == parts) return true;
    return false;
}

template <uint32_t S, uint32_t H>
void launch_backend_t(srg_ctx* ctx, uint32_t wg, bool stats, const FrameDev& F)
{
    const dim3 grid((uint32_t)ctx->num_cus * wg), block(backend_threads(S, H));
    if (F.rec_ext) {
        if (stats) hipLaunchKernelGGL((k_backend<S, H, true, true>), grid, block, 0, ctx->stream, F);
        else hipLaunchKernelGGL((k_backend<S, H, false, true>), grid, block, 0, ctx->stream, F);
    } else {
        if (stats) hipLaunchKernelGGL((k_backend<S, H, true, false>), grid, block, 0, ctx->stream, F);
        else hipLaunchKernelGGL((k_backend<S, H, false, false>), grid, block, 0, ctx->stream, F);
    }
}

// k_setup over the frame's setup groups: groups [0, nsingle) of one draw
// (workgroups [0, single_blocks)), then the shared ones (the rest).
template <uint32_t NS>
void launch_setup(srg_ctx* ctx, const FrameDev& F, uint32_t nsingle, uint32_t single_blocks, uint32_t nblocks)
{
    if (single_blocks)
        hipLaunchKernelGGL((k_setup<NS, false>), dim3(single_blocks), dim3(kSetupThreads), 0, ctx->stream, F,
                           F.groups, F.group_members, F.draws, F.views, 0u, nsingle, 0u);
    if (nblocks > single_blocks)
        hipLaunchKernelGGL((k_setup<NS, true>), dim3(nblocks - single_blocks), dim3(kSetupThreads), 0, ctx->stream,
                           F, F.groups, F.group_members, F.draws, F.views, nsingle, F.ngroups, single_blocks);
}

int launch_backend(srg_ctx* ctx, BackendLayout L, const FrameDev& F)
{
    // resident workgroups per CU: LDS holds 8 bytes per strip pixel
    const uint32_t wg = ctx->cfg.wg_per_cu ? ctx->cfg.wg_per_cu
                      : L.bin_size * L.bin_size / L.parts >= 4096u ? 4u : (L.bin_size == 8 ? 16u : 8u);
    const bool stats = (F.flags & FF_STATS) != 0;
    switch (L.bin_size * 8u + L.parts) {
    case 8 * 8 + 1: launch_backend_t<8, 8>(ctx, wg, stats, F); break;
    case 16 * 8 + 1: launch_backend_t<16, 16>(ctx, wg, stats, F); break;
    case 32 * 8 + 1: launch_backend_t<32, 32>(ctx, wg, stats, F); break;
    case 32 * 8 + 2: launch_backend_t<32, 16>(ctx, wg, stats, F); break;
    case 32 * 8 + 4: launch_backend_t<32, 8>(ctx, wg, stats, F); break;
    case 64 * 8 + 1: launch_backend_t<64, 64>(ctx, wg, stats, F); break;
    case 64 * 8 + 2: launch_backend_t<64, 32>(ctx, wg, stats, F); break;
    case 64 * 8 + 4: launch_backend_t<64, 16>(ctx, wg, stats, F); break;
    default: return SRG_ERR_INVALID;
    }
    return SRG_OK;
}

BackendLayout config_layout(const srg_config& cfg)
{
    return { cfg.bin_size ? cfg.bin_size : 32u, cfg.wg_per_tile ? cfg.wg_per_tile : 1u };
}

uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h * 0xFF51AFD7ED558CCDull;
}

// The autotuner's layout for a frame of signature sig: the next layout of
// the sweep, or the chosen one.
BackendLayout tune_layout(srg_ctx* ctx, uint64_t sig)
{
    auto it = ctx->tunes.find(sig);
    if (it == ctx->tunes.end()) {
        if (ctx->tunes.size() >= 64) ctx->tunes.clear();
        TuneState t;
        t.sweeping = true;
        t.sig = sig;
        it = ctx->tunes.emplace(sig, t).first;
    }
    return it->second.sweeping ? kLayouts[it->second.cand] : it->second.chosen;
}

// After a frame of the sweep: keep the layout's best device time; after the
// last layout, choose the fastest.
void tune_after_frame(srg_ctx* ctx, uint64_t sig, float ms)
{
    TuneState& t = ctx->tunes[sig];
    if (!t.sweeping) return;
    float& b = t.best_ms[t.cand];
    if (b == 0.0f || ms < b) b = ms;
    const uint32_t frames = ctx->cfg.autotune_frames ? ctx->cfg.autotune_frames : 3u;
    if (++t.frame < frames) return;
    t.frame = 0;
    if (++t.cand < kNumLayouts) return;
    uint32_t best = 0;
    for (uint32_t i = 1; i < kNumLayouts; ++i)
        if (t.best_ms[i] < t.best_ms[best]) best = i;
    t.chosen = kLayouts[best];
    t.sweeping = false;
    t.locked = true;
}

// Device memory one frame works in: the context's arenas (grown as frames
// need them), or pieces of the scratch array for srg_end_frame_noalloc.
struct FrameArenas {
    uint8_t* tables = nullptr;
    uint4* rec_raster = nullptr;
    float* rec_shade = nullptr;
    float4* rec_ext = nullptr;
    uint32_t *rec_meta = nullptr, *clip_base = nullptr, *clip_order = nullptr;
    uint2* clip_queue = nullptr;
    uint32_t *bin_count = nullptr, *bin_start = nullptr, *part_count = nullptr, *pairs = nullptr;
    uint32_t *part_off = nullptr, *queue = nullptr, *queue2 = nullptr, *entries = nullptr, *scan_blk = nullptr;
    uint16_t* order_bkt = nullptr;
    uint32_t* order_hist = nullptr;
    uint64_t* bin_cost = nullptr;
    uint32_t* counters = nullptr;
};

// One EndFrame. With noalloc, nothing is allocated: the arenas are carved
// from the scratch array (entry lists take what the rest leaves), the queue
// is in natural or entry-count order (the learned order's buffers are not
// used), the autotuner's locked choice or the configured layout is used
// without sweeping, and the launches run without a graph.
int run_frame(srg_ctx* ctx, bool noalloc)
{
    // ---- views: planes drawn to, plus planes with a pending clear ----
    std::vector<ViewKey> views;
    uint32_t last_view = ~0u;  // consecutive draws mostly share a plane
    auto view_of = [&](srg_fb* fb, uint32_t plane) -> uint32_t {
        if (last_view < views.size() && views[last_view].fb == fb && views[last_view].plane == plane) return last_view;
        for (uint32_t i = 0; i < views.size(); ++i)
            if (views[i].fb == fb && views[i].plane == plane) return last_view = i;
        views.push_back({ fb, plane });
        return last_view = (uint32_t)views.size() - 1;
    };
    std::vector<uint32_t> draw_view;
    draw_view.reserve(ctx->draws.size());
    for (auto& d : ctx->draws) draw_view.push_back(view_of(d.fb, d.plane));
    for (srg_fb* fb : ctx->fbs)
        for (uint32_t p = 0; p < 2; ++p)
            if (fb->planes[p].pending) view_of(fb, p);
    if (views.empty()) return SRG_OK;

    // back-end layout: the configured one, or the autotuner's for this
    // frame's signature (view sizes, per-draw triangle counts)
    const bool autotune = (ctx->cfg.flags & SRG_CFG_AUTOTUNE) != 0;
    uint64_t sig = 0;
    BackendLayout L = config_layout(ctx->cfg);
    if (autotune) {
        sig = mix64(sig, views.size());
        for (const ViewKey& v : views) sig = mix64(mix64(sig, v.fb->width), v.fb->height);
        for (size_t i = 0; i < ctx->draws.size(); ++i) sig = mix64(mix64(sig, ctx->draws[i].index_num), draw_view[i]);
        if (noalloc) {
            const auto it = ctx->tunes.find(sig);
            if (it != ctx->tunes.end() && it->second.locked) L = it->second.chosen;
        } else {
            L = tune_layout(ctx, sig);
            ctx->tune_sig = sig;
        }
    }
    const bool sweeping = autotune && !noalloc && ctx->tunes[sig].sweeping;
    const uint32_t K = kBin / L.bin_size;
    std::vector<ViewDev> vdev(views.size());
    uint64_t total_bins = 0;
    for (size_t i = 0; i < views.size(); ++i) {
        srg_fb* fb = views[i].fb;
        PlaneState& ps = fb->planes[views[i].plane];
        ViewDev& v = vdev[i];
        std::memset(&v, 0, sizeof(v));
        v.colour = ps.colour;
        v.depth = ps.depth;
        v.primid = (ctx->cfg.flags & SRG_CFG_KEEP_PRIMID) ? ps.primid : nullptr;
        v.width = fb->width;
        v.height = fb->height;
        v.tiles_x = fb->tiles_x;
        v.tiles_y = fb->tiles_y;
        v.bins_x = fb->tiles_x * K;
        v.bins_y = fb->tiles_y * K;
        v.bin_base = (uint32_t)total_bins;
        v.flags = ps.colour ? VF_HAS_COLOUR : 0u;
        if (ps.pending) {
            v.flags |= (ps.clear_colour && ps.colour ? VF_CLEAR_COLOUR : 0u) | (ps.clear_depth ? VF_CLEAR_DEPTH : 0u);
            v.clear_colour = (ps.clear_value & 0xFFu) * 0x01010101u;
        }
        total_bins += (uint64_t)v.bins_x * v.bins_y;
    }
    if (total_bins * kBinParts >= 0xFFFFFFF0ull) return SRG_ERR_INVALID;  // (part, bin) indices are 32-bit

    // ---- draws, textures, front-end blocks ----
    std::vector<DrawDev> ddev(ctx->draws.size());
    std::vector<TexDev> tdev;
    std::unordered_map<uint64_t, uint32_t> tex_index;
    uint64_t total_tris = 0;
    uint32_t nblocks = 0;
    bool wide = false;  // a draw keeps more than kMaxSlots attribute planes
    uint32_t max_slots = 0;
    for (size_t i = 0; i < ctx->draws.size(); ++i) {
        const srg_draw& s = ctx->draws[i];
        DrawDev& d = ddev[i];
        std::memset(&d, 0, sizeof(d));
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
        d.view = draw_view[i];
        d.shader = s.shader;
        int st = shader_slots(s.shader, s.uv_offset, s.attribute_stride / 4u, d.slot_attr, d.nslots, d.deriv_slot);
        if (st) return st;
        wide = wide || d.nslots > kMaxSlots;
        max_slots = std::max(max_slots, d.nslots);
        d.tex = kNoTex;
        if (shader_samples(s.shader) && s.texture) {
            auto it = tex_index.find(s.texture);
            if (it == tex_index.end()) {
                const Texture& t = ctx->textures[s.texture];
                TexDev td;
                std::memset(&td, 0, sizeof(td));
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
            const TexDev& td = tdev[d.tex];
            d.tex_info = td.empty ? 0u : (std::max(td.wlog2, td.hlog2) | (td.nmips << 8));
        }
        std::memcpy(d.mvp, s.mvp, sizeof(d.mvp));
        total_tris += d.ntris;
    }
    // ---- setup groups: draws that share their geometry, at most
    // kMaxGroupMembers each, in draw order ----
    // groups of one first (k_setup<NS, false>), then the shared ones
    // (k_setup<NS, true>), each launch over its own range of workgroups
    std::vector<SetupGroupDev> gdev;
    std::vector<uint32_t> gmem;
    uint32_t nsingle = 0, single_blocks = 0;
    {
        // the geometry a draw fetches: buffers, strides, counts, kept slots
        struct GeomKey {
            const uint8_t* ptr[3];
            uint32_t f[8];  // seven fields and a zero word
            uint32_t slot_attr[kWideSlots];
            bool operator==(const GeomKey& o) const { return std::memcmp(this, &o, sizeof(GeomKey)) == 0; }
        };
        struct GeomHash {
            size_t operator()(const GeomKey& k) const
            {
                uint64_t h = 0;
                const uint64_t* w = (const uint64_t*)&k;
                for (size_t i = 0; i < sizeof(GeomKey) / 8; ++i) h = mix64(h, w[i]);
                return (size_t)h;
            }
        };
        static_assert(sizeof(GeomKey) % 8 == 0 && sizeof(DrawDev::slot_attr) == sizeof(GeomKey::slot_attr),
                      "geometry key: whole words, every kept slot");
        std::vector<std::vector<uint32_t>> lists;
        std::unordered_map<GeomKey, uint32_t, GeomHash> open_group;
        open_group.reserve(ddev.size());
        for (size_t i = 0; i < ddev.size(); ++i) {
            const DrawDev& d = ddev[i];
            if (d.ntris == 0) continue;
            GeomKey key;
            std::memset(&key, 0, sizeof(key));
            key.ptr[0] = d.idx; key.ptr[1] = d.pos; key.ptr[2] = d.attr;
            const uint32_t fields[] = { d.idx_stride, d.pos_stride, d.attr_stride, d.ntris, d.pos_num, d.attr_num,
                                        d.nslots };
            std::memcpy(key.f, fields, sizeof(fields));
            std::memcpy(key.slot_attr, d.slot_attr, sizeof(key.slot_attr));
            auto it = open_group.find(key);
            if (it == open_group.end() || lists[it->second].size() >= kMaxGroupMembers) {
                open_group[key] = (uint32_t)lists.size();
                lists.emplace_back();
                it = open_group.find(key);
            }
            lists[it->second].push_back((uint32_t)i);
        }
        for (int pass = 0; pass < 2; ++pass)
            for (const auto& l : lists) {
                if ((l.size() > 1) != (pass == 1)) continue;
                const uint32_t nt = ddev[l[0]].ntris;
                gdev.push_back({ nblocks, nt, (uint32_t)gmem.size(), (uint32_t)l.size() });
                gmem.insert(gmem.end(), l.begin(), l.end());
                nblocks += (nt + kSetupThreads - 1) / kSetupThreads;
                if (pass == 0) {
                    nsingle = (uint32_t)gdev.size();
                    single_blocks = nblocks;
                }
            }
    }
    if (total_tris >= (1ull << 29)) return SRG_ERR_INVALID;  // order key is (triangle << 3) | fan
    // setup records keep two attribute planes (9 floats of shade record) when
    // no draw of the frame reads more, four (15 floats) otherwise; planes 4-7
    // of a frame with a SponzaShader draw go to rec_ext
    const bool narrow = !wide && max_slots <= 2;
    const uint32_t shade_stride = narrow ? 9u : 15u;
    const uint32_t NT = (uint32_t)total_tris;

    size_t off = 0;
    const size_t o_draw = off; off = align_up(off + sizeof(DrawDev) * std::max<size_t>(ddev.size(), 1), 256);
    const size_t o_view = off; off = align_up(off + sizeof(ViewDev) * vdev.size(), 256);
    const size_t o_tex = off;  off = align_up(off + sizeof(TexDev) * std::max<size_t>(tdev.size(), 1), 256);
    const size_t o_sponza = off; off = align_up(off + (wide ? sizeof(SponzaDev) : 0), 256);
    const size_t o_group = off; off = align_up(off + sizeof(SetupGroupDev) * std::max<size_t>(gdev.size(), 1), 256);
    const size_t o_member = off; off = align_up(off + sizeof(uint32_t) * std::max<size_t>(gmem.size(), 1), 256);
    int st = SRG_OK;
    FrameArenas A;
    Carver C{ (uint8_t*)ctx->scratch_dev, noalloc ? (size_t)SRG_NOALLOC_SCRATCH_BYTES : 0u };
    if (noalloc) {
        if (!ctx->scratch_dev) return SRG_ERR_HIP;
        if (off > ctx->h_tables_cap) return SRG_ERR_OVERFLOW;  // pinned staging sized at srg_create
        A.tables = C.take<uint8_t>(off);
    } else {
        if ((st = ensure_host_tables(ctx, off)) || (st = ctx->tables.ensure(off))) return st;
        A.tables = ctx->tables.ptr;
    }
    if (!ddev.empty()) std::memcpy(ctx->h_tables + o_draw, ddev.data(), sizeof(DrawDev) * ddev.size());
    std::memcpy(ctx->h_tables + o_view, vdev.data(), sizeof(ViewDev) * vdev.size());
    if (!tdev.empty()) std::memcpy(ctx->h_tables + o_tex, tdev.data(), sizeof(TexDev) * tdev.size());
    if (wide) std::memcpy(ctx->h_tables + o_sponza, &ctx->sponza, sizeof(SponzaDev));
    if (!gdev.empty()) std::memcpy(ctx->h_tables + o_group, gdev.data(), sizeof(SetupGroupDev) * gdev.size());
    if (!gmem.empty()) std::memcpy(ctx->h_tables + o_member, gmem.data(), sizeof(uint32_t) * gmem.size());

    // ---- arenas ----
    const uint32_t TB = (uint32_t)std::max<uint64_t>(total_bins, 1);
    const uint32_t NQ = std::max<uint32_t>(NT, 1);
    const size_t n_hist = (size_t)((TB + kOrderBlock - 1) / kOrderBlock) * kOrderBuckets;
    const size_t n_scan = 2 * ((TB + kScanThreads - 1) / kScanThreads);
    if (noalloc) {
        A.counters = C.take<uint32_t>(kCounterWords);
        A.bin_count = C.take<uint32_t>(TB);
        A.bin_start = C.take<uint32_t>(TB + 1);
        A.part_count = C.take<uint32_t>((size_t)kBinParts * TB);
        A.part_off = C.take<uint32_t>((size_t)kBinParts * TB);
        A.queue = C.take<uint32_t>(TB);
        A.queue2 = C.take<uint32_t>(TB);
        A.order_bkt = C.take<uint16_t>(TB);
        A.order_hist = C.take<uint32_t>(n_hist);
        A.scan_blk = C.take<uint32_t>(n_scan);
        A.bin_cost = C.take<uint64_t>(TB);
        A.clip_queue = C.take<uint2>(NQ);
        A.clip_base = C.take<uint32_t>(NQ);
        if (!C.ok) return SRG_ERR_OVERFLOW;
    } else {
        if ((st = ctx->bin_count.ensure(TB)) || (st = ctx->bin_start.ensure(TB + 1)) ||
            (st = ctx->part_count.ensure((size_t)kBinParts * TB)) ||
            (st = ctx->part_off.ensure((size_t)kBinParts * TB)) || (st = ctx->queue.ensure(TB)) ||
            (st = ctx->queue2.ensure(TB)) || (st = ctx->order_bkt.ensure(TB)) || (st = ctx->order_hist.ensure(n_hist)) ||
            (st = ctx->scan_blk.ensure(n_scan)) || (st = ctx->bin_cost.ensure(TB)) ||
            (st = ctx->counters.ensure(kCounterWords)) || (st = ctx->clip_queue.ensure(NQ)) ||
            (st = ctx->clip_base.ensure(NQ)))
            return st;
        A.counters = ctx->counters.ptr;
        A.bin_count = ctx->bin_count.ptr;
        A.bin_start = ctx->bin_start.ptr;
        A.part_count = ctx->part_count.ptr;
        A.part_off = ctx->part_off.ptr;
        A.queue = ctx->queue.ptr;
        A.queue2 = ctx->queue2.ptr;
        A.order_bkt = ctx->order_bkt.ptr;
        A.order_hist = ctx->order_hist.ptr;
        A.scan_blk = ctx->scan_blk.ptr;
        A.bin_cost = ctx->bin_cost.ptr;
        A.clip_queue = ctx->clip_queue.ptr;
        A.clip_base = ctx->clip_base.ptr;
    }
    uint64_t want_clip = ctx->cfg.max_prims ? ctx->cfg.max_prims : (uint64_t)NT / 8 + 4096;
    uint64_t want_entries = ctx->cfg.max_entries ? ctx->cfg.max_entries : 2ull * NT + 4096;
    if (!noalloc) {
        want_clip = std::max<uint64_t>(want_clip, ctx->clip_cap);
        want_entries = std::max<uint64_t>(want_entries, ctx->entry_cap);
    }

    // planes an asynchronous blit is still reading are written only after it
    for (const ViewKey& v : views) stream_wait_blit_read(ctx, v.fb->planes[v.plane]);

    for (uint32_t attempt = 0;; ++attempt) {
        if (want_clip > 0xFFFFFFF0ull - NT || want_entries > 0xFFFFFFF0ull) return SRG_ERR_OVERFLOW;
        const uint64_t nrec = (uint64_t)NT + want_clip;
        if (noalloc) {
            // records first; the entry lists take what the scratch has left
            Carver R = C;
            A.rec_raster = R.take<uint4>(nrec * 4);
            A.rec_shade = R.take<float>(nrec * shade_stride);
            A.rec_meta = R.take<uint32_t>(nrec);
            A.clip_order = R.take<uint32_t>(want_clip);
            A.rec_ext = wide ? R.take<float4>(nrec * 3) : nullptr;
            // each entry takes 4 bytes of list and 18 of (record, key, slot)
            // pairs (setup regions plus k_clip's half-size region), 22 in
            // all; 64 entries' worth is left for the alignment of the two
            const size_t left = R.left();
            if (!R.ok || left < (4096 + 64) * 22) return SRG_ERR_OVERFLOW;
            want_entries = std::min<uint64_t>(left / 22 - 64, 0xFFFFFFF0ull);
            A.entries = R.take<uint32_t>(want_entries);
            A.pairs = R.take<uint32_t>(3 * (want_entries + want_entries / 2));
            if (!R.ok) return SRG_ERR_OVERFLOW;
        } else {
            if ((st = ctx->rec_raster.ensure(nrec * 4)) || (st = ctx->rec_shade.ensure(nrec * shade_stride)) ||
                (st = ctx->rec_meta.ensure(nrec)) || (st = ctx->clip_order.ensure(want_clip)) ||
                (st = ctx->entries.ensure(want_entries)) || (st = ctx->pairs.ensure(3 * (want_entries + want_entries / 2))) ||
                (wide && (st = ctx->rec_ext.ensure(nrec * 3))))
                return st;
            ctx->clip_cap = (uint32_t)want_clip;
            ctx->entry_cap = (uint32_t)want_entries;
            A.rec_raster = ctx->rec_raster.ptr;
            A.rec_shade = ctx->rec_shade.ptr;
            A.rec_meta = ctx->rec_meta.ptr;
            A.clip_order = ctx->clip_order.ptr;
            A.entries = ctx->entries.ptr;
            A.pairs = ctx->pairs.ptr;
            A.rec_ext = wide ? ctx->rec_ext.ptr : nullptr;
        }

        FrameDev F;
        std::memset(&F, 0, sizeof(F));
        F.draws = (const DrawDev*)(A.tables + o_draw);
        F.views = (const ViewDev*)(A.tables + o_view);
        F.texs = (const TexDev*)(A.tables + o_tex);
        F.groups = (const SetupGroupDev*)(A.tables + o_group);
        F.group_members = (const uint32_t*)(A.tables + o_member);
        F.ngroups = (uint32_t)gdev.size();
        F.ndraws = (uint32_t)ddev.size();
        F.nviews = (uint32_t)vdev.size();
        F.total_bins = (uint32_t)total_bins;
        F.ntris = NT;
        F.bin_k = K;
        F.bin_size = kBin / K;
        F.bin_log2 = (uint32_t)__builtin_ctz(F.bin_size);
        F.bin_klog2 = (uint32_t)__builtin_ctz(K);
        F.clip_cap = (uint32_t)want_clip;
        F.entry_cap = (uint32_t)want_entries;
        F.rec_raster = A.rec_raster;
        F.rec_shade = A.rec_shade;
        F.shade_stride = shade_stride;
        F.rec_ext = A.rec_ext;
        F.sponza = wide ? (const SponzaDev*)(A.tables + o_sponza) : nullptr;
        F.rec_meta = A.rec_meta;
        F.clip_queue = A.clip_queue;
        F.clip_base = A.clip_base;
        F.clip_order = A.clip_order;
        F.bin_count = A.bin_count;
        F.bin_start = A.bin_start;
        F.part_count = A.part_count;
        F.part_off = A.part_off;
        F.entries = A.entries;
        F.pairs = A.pairs;
        F.pair_count = A.counters + kPairCounterBase;
        // a power of two, so that k_setup finds its region with a mask
        F.pair_regions = 1u << (31 - __builtin_clz(std::max<uint32_t>(1, std::min<uint32_t>(nblocks, kPairRegions))));
        F.pair_region_cap = (uint32_t)(want_entries / F.pair_regions);
        F.clip_pair_cap = (uint32_t)(want_entries / 2 / kClipPairRegions);
        F.pair_total = F.pair_regions * F.pair_region_cap + kClipPairRegions * F.clip_pair_cap;
        F.queue = A.queue;
        F.bin_cost = A.bin_cost;
        F.counters = A.counters;
        F.flags = ((ctx->cfg.flags & SRG_CFG_STATS) ? FF_STATS : 0u) |
                  ((ctx->cfg.flags & SRG_CFG_KEEP_PRIMID) ? FF_PRIMID : 0u) |
                  ((ctx->cfg.flags & SRG_CFG_TILE_COSTS) ? FF_TILE_COSTS : 0u);

        // k_scatter: a workgroup per (region, chunk of a set-up region's
        // capacity); a pair per thread while the chunks number at most 512,
        // up to eight above (fewer, longer-lived workgroups)
        const uint32_t emit_chunk =
            256u * std::max<uint32_t>(1, std::min<uint32_t>(8, (F.pair_region_cap + 256u * 512u - 1) / (256u * 512u)));
        const uint32_t emit_blocks =
            std::max<uint32_t>(1, std::min<uint32_t>((F.pair_region_cap + emit_chunk - 1) / emit_chunk, 4096));
        const uint32_t clip_blocks = std::max<uint32_t>(
            1, std::min<uint32_t>((NT + kClipThreads - 1) / kClipThreads, ctx->clip_resident[wide ? 2 : (narrow ? 0 : 1)]));
        const uint32_t init_blocks = std::max<uint32_t>(1, std::min<uint32_t>((kBinParts * TB + 255) / 256, 1024));
        const uint32_t scan_blocks = (TB + kScanThreads - 1) / kScanThreads;
        const bool ordered = ctx->cfg.tile_order != SRG_ORDER_NATURAL;
        // while the autotuner sweeps, the learned order falls back to entry
        // counts and records nothing (its samples belong to one bin size)
        const bool learned = ctx->cfg.tile_order == SRG_ORDER_LEARNED && !sweeping && !noalloc;
        if (learned) {
            if ((st = ensure_sched(ctx)) || (st = ctx->sched_feat.ensure(TB)) || (st = ctx->sched_pred.ensure(TB)))
                return st;
            if (ctx->sched_bin_size != L.bin_size) {
                if (ctx->sched_bin_size && (st = reset_sched(ctx))) return st;
                ctx->sched_bin_size = L.bin_size;
            }
            F.flags |= FF_TILE_COSTS;
        }
        // the model orders the queue once trained, unless its last measured
        // error was too large: then the entry-count order is the fallback
        const bool use_model = ctx->sched_trained && !ctx->sched_fallback;
        // predict and record every frame until trained, then every
        // kSchedEvery-th frame (the count restarts, outside any captured
        // graph, when the model changes, so that frame records); a caller
        // that asked for tile costs gets them every frame
        const uint32_t sched_every = (!ctx->sched_trained || (ctx->cfg.flags & SRG_CFG_TILE_COSTS)) ? 1u : kSchedEvery;
        F.sched_gate = learned ? ctx->sched_gate : nullptr;
        if (learned) {
            // predictions are per bin: a frame whose bins are laid out
            // differently from the last one's (other views, sizes or bin
            // size) predicts afresh, and so does a frame rendered again after
            // an overflow (its first attempt advanced the frame count)
            uint64_t lay = mix64(vdev.size(), K);
            for (const ViewDev& v : vdev) lay = mix64(mix64(mix64(lay, v.bin_base), v.bins_x), v.bins_y);
            if (lay != ctx->sched_layout || attempt > 0) ctx->sched_force_record = true;
            ctx->sched_layout = lay;
        }
        if (learned && ctx->sched_force_record) {
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->sched_gate, 0u, (size_t)1);
            HIP_TRY(hipGetLastError());
        }
        // The whole frame on the stream: tables, six kernels, counters back.
        auto enqueue = [&](bool split_events) -> int {
            HIP_TRY(hipMemcpyAsync(A.tables, ctx->h_tables, off, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(k_frame_init, dim3(init_blocks), dim3(256), 0, ctx->stream, A.counters,
                               A.part_count, A.bin_cost, TB,
                               learned ? ctx->sched_gate : nullptr, sched_every);
            if (ctx->cfg.flags & SRG_CFG_KEEP_PRIMID)
                for (auto& v : vdev)
                    if (v.primid)
                        hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, ctx->stream, v.primid, 0xFFFFFFFFu,
                                           (size_t)v.tiles_x * v.tiles_y * kTilePixels);
            if (wide) {
                launch_setup<kWideSlots>(ctx, F, nsingle, single_blocks, nblocks);
                if (NT) hipLaunchKernelGGL(k_clip<kWideSlots>, dim3(clip_blocks), dim3(kClipThreads), 0, ctx->stream, F);
            } else if (narrow) {
                launch_setup<2>(ctx, F, nsingle, single_blocks, nblocks);
                if (NT) hipLaunchKernelGGL(k_clip<2>, dim3(clip_blocks), dim3(kClipThreads), 0, ctx->stream, F);
            } else {
                launch_setup<kMaxSlots>(ctx, F, nsingle, single_blocks, nblocks);
                if (NT) hipLaunchKernelGGL(k_clip<kMaxSlots>, dim3(clip_blocks), dim3(kClipThreads), 0, ctx->stream, F);
            }
            FrameDev Fs = F;  // with ordering, the natural queue goes to the spare buffer first
            if (ordered) Fs.queue = A.queue2;
            if (TB <= kScanSingleBins) {
                hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanThreads), 0, ctx->stream, Fs);
            } else {
                hipLaunchKernelGGL(k_scan_blocks, dim3(scan_blocks), dim3(kScanThreads), 0, ctx->stream, Fs,
                                   A.scan_blk);
                hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, ctx->stream, Fs, A.scan_blk,
                                   scan_blocks);
                hipLaunchKernelGGL(k_scan_apply, dim3(scan_blocks), dim3(kScanThreads), 0, ctx->stream, Fs,
                                   (const uint32_t*)A.scan_blk);
            }
            hipLaunchKernelGGL(k_scatter, dim3(F.pair_regions + kClipPairRegions, emit_blocks), dim3(256), 0,
                               ctx->stream, F, emit_chunk);
            if (learned) {
                hipLaunchKernelGGL(k_bin_predict, dim3(std::min<uint32_t>((TB + 3) / 4, 16384u)),
                                   dim3(256), 0, ctx->stream, F, ctx->sched_feat.ptr,
                                   (const MlpParams*)(ctx->sched_trained ? ctx->sched_params : nullptr),
                                   ctx->sched_pred.ptr);
            }
            if (ordered) {
                const float* pred = (const float*)(learned && use_model ? ctx->sched_pred.ptr : nullptr);
                const uint32_t oblk = (TB + kOrderBlock - 1) / kOrderBlock;
                hipLaunchKernelGGL(k_order_hist, dim3(oblk), dim3(kOrderBlock), 0, ctx->stream, F,
                                   (const uint32_t*)A.queue2, pred, A.order_bkt, A.order_hist);
                hipLaunchKernelGGL(k_order_offsets, dim3(1), dim3(kOrderBuckets), 0, ctx->stream, F, A.order_hist);
                hipLaunchKernelGGL(k_order_scatter, dim3(oblk), dim3(kOrderBlock), 0, ctx->stream, F,
                                   (const uint32_t*)A.queue2, (const uint16_t*)A.order_bkt,
                                   (const uint32_t*)A.order_hist);
            }
            if (split_events) HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
            if (const int lst = launch_backend(ctx, L, F)) return lst;
            if (split_events) HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
            if (learned)
                hipLaunchKernelGGL(k_mlp_record, dim3((TB + 255) / 256), dim3(256), 0, ctx->stream, F,
                                   (const float4*)ctx->sched_feat.ptr,
                                   (const float*)(ctx->sched_trained ? ctx->sched_pred.ptr : nullptr), ctx->sched_ring,
                                   kSchedRing, ctx->sched_head);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(ctx->h_counters, A.counters, sizeof(uint32_t) * C_NCOUNTERS,
                                   hipMemcpyDeviceToHost, ctx->stream));
            return SRG_OK;
        };

        const bool timing = (ctx->cfg.flags & SRG_CFG_TIMING) != 0 || sweeping;
        bool split = false;
        if ((ctx->cfg.flags & SRG_CFG_GRAPH) && !noalloc) {
            // Replay the captured frame when nothing it depends on changed:
            // the table bytes, every pointer and capacity in F, and the
            // launch shapes.
            std::vector<uint8_t> sig(ctx->h_tables, ctx->h_tables + off);
            const uint64_t extra[] = { (uint64_t)(uintptr_t)ctx->h_tables, (uint64_t)(uintptr_t)A.tables,
                                       (uint64_t)(uintptr_t)ctx->h_counters, nblocks, emit_blocks, clip_blocks,
                                       init_blocks, TB, K, L.parts, ctx->cfg.flags, ctx->cfg.wg_per_cu,
                                       (uint64_t)ctx->sched_trained | ((uint64_t)ctx->sched_fallback << 1) |
                                           ((uint64_t)sched_every << 2), ctx->cfg.tile_order,
                                       (uint64_t)(uintptr_t)ctx->sched_feat.ptr, (uint64_t)(uintptr_t)ctx->sched_pred.ptr };
            sig.insert(sig.end(), (const uint8_t*)&F, (const uint8_t*)&F + sizeof(F));
            sig.insert(sig.end(), (const uint8_t*)extra, (const uint8_t*)extra + sizeof(extra));
            if (!ctx->graph_exec || sig != ctx->graph_sig) {
                if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
                ctx->graph_exec = nullptr;
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                const int est = enqueue(false);
                const hipError_t cerr = hipStreamEndCapture(ctx->stream, &graph);
                if (est) return est;
                if (cerr != hipSuccess) return SRG_ERR_HIP;
                const hipError_t ierr = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ierr != hipSuccess) {
                    ctx->graph_exec = nullptr;
                    return SRG_ERR_HIP;
                }
                ctx->graph_sig.swap(sig);
                ctx->stats.graph_captures++;
            }
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
            HIP_TRY(hipGraphLaunch(ctx->graph_exec, ctx->stream));
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
        } else {
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
            if ((st = enqueue(timing))) return st;
            if (timing) HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
            split = timing;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));

        const uint32_t flags = ctx->h_counters[C_FLAGS];
        if (flags & FLAG_INDEX_RANGE) return SRG_ERR_INDEX_RANGE;
        if (flags & (FLAG_CLIP_OVERFLOW | FLAG_ENTRY_OVERFLOW)) {
            if (attempt >= 4) return SRG_ERR_OVERFLOW;
            // the scratch cannot grow: only more fan records at the entries' expense
            if (noalloc && (flags & FLAG_ENTRY_OVERFLOW)) return SRG_ERR_OVERFLOW;
            if (flags & FLAG_CLIP_OVERFLOW)
                want_clip = std::max<uint64_t>(want_clip * 2, (uint64_t)ctx->h_counters[C_CLIPREC] + 64);
            if (flags & FLAG_ENTRY_OVERFLOW)
                want_entries = std::max({ want_entries * 2, (uint64_t)ctx->h_counters[C_ENTRIES] * 5 / 4,
                                          (uint64_t)ctx->h_counters[C_PAIR_NEED] * F.pair_regions * 5 / 4,
                                          (uint64_t)ctx->h_counters[C_CLIP_PAIR_NEED] * 2 * kClipPairRegions * 5 / 4,
                                          clip_pair_want(ctx->h_counters[C_CLIP_PAIRS]) });
            ctx->stats.retries++;
            continue;
        }
        // The clip fans' pairs reach their regions in the order the clip
        // queue was filled, which changes from frame to frame: when a region
        // could not hold all of them, the next frame sizes every region so
        // (the total is the same for the same triangles).
        if (!noalloc && clip_pair_want(ctx->h_counters[C_CLIP_PAIRS]) > want_entries)
            ctx->entry_cap = (uint32_t)std::min<uint64_t>(clip_pair_want(ctx->h_counters[C_CLIP_PAIRS]), 0xFFFFFFF0ull);
        ctx->stats.tris_in = total_tris;
        ctx->stats.prims = ctx->h_counters[C_PRIMS];
        ctx->stats.bin_entries = ctx->h_counters[C_TILE_ENTRIES];
        ctx->stats.work_entries = ctx->h_counters[C_ENTRIES];
        ctx->stats.tris_clipped = ctx->h_counters[C_CLIPQ];
        ctx->stats.frags_raster = (uint64_t)ctx->h_counters[C_FRAGS] | ((uint64_t)ctx->h_counters[C_FRAGS + 1] << 32);
        ctx->stats.tiles_queued = ctx->h_counters[C_QLEN];
        if (learned && (st = sched_after_frame(ctx))) return st;
        if (timing) {
            if (split) {
                (void)hipEventElapsedTime(&ctx->stats.ms_frontend, ctx->ev[0], ctx->ev[1]);
                (void)hipEventElapsedTime(&ctx->stats.ms_backend, ctx->ev[1], ctx->ev[4]);
            }
            (void)hipEventElapsedTime(&ctx->stats.ms_total, ctx->ev[0], ctx->ev[2]);
        }
        if (sweeping) tune_after_frame(ctx, sig, ctx->stats.ms_total);
        ctx->stats.bin_size = L.bin_size;
        ctx->stats.wg_per_tile = L.parts;
        break;
    }
    for (auto& v : views) v.fb->planes[v.plane].pending = false;
    ctx->last_bins = (uint32_t)total_bins;
    return SRG_OK;
}

bool valid_handle(srg_ctx* ctx, srg_handle h) { return h && ctx->buffers.count(h); }

bool fb_known(const srg_ctx* ctx, const srg_fb* fb)
{
    return fb && std::find(ctx->fbs.begin(), ctx->fbs.end(), fb) != ctx->fbs.end();
}

// The checks srg_draw_indexed makes. srg_end_frame makes them again for every
// recorded draw: a buffer, texture or framebuffer may have been released
// between the two calls, and the frame must then fail with a status instead of
// reading a released handle.
int validate_draw(srg_ctx* ctx, const srg_draw* d)
{
    if (!d->fb || (d->plane > 1 && d->plane != SRG_PLANE_WRITE)) return SRG_ERR_INVALID;
    if (!fb_known(ctx, d->fb)) return SRG_ERR_HANDLE;
    if (!valid_handle(ctx, d->index_buffer) || !valid_handle(ctx, d->position_buffer) ||
        !valid_handle(ctx, d->attribute_buffer))
        return SRG_ERR_HANDLE;
    if (d->index_stride != 1 && d->index_stride != 2 && d->index_stride != 4) return SRG_ERR_INVALID;
    if (d->position_stride < 12 || (d->position_stride & 3) || (d->position_offset & 3)) return SRG_ERR_INVALID;
    if (d->attribute_stride == 0 || d->attribute_stride > 32 || (d->attribute_stride & 3) || (d->attribute_offset & 3))
        return SRG_ERR_INVALID;
    if (d->index_offset % d->index_stride) return SRG_ERR_INVALID;
    uint32_t sa[kWideSlots], ns, ds;
    int st = shader_slots(d->shader, d->uv_offset, d->attribute_stride / 4u, sa, ns, ds);
    if (st) return st;
    if (shader_samples(d->shader) && d->texture && !ctx->textures.count(d->texture)) return SRG_ERR_HANDLE;
    const Buffer& ib = ctx->buffers.at(d->index_buffer);
    const Buffer& pb = ctx->buffers.at(d->position_buffer);
    const Buffer& ab = ctx->buffers.at(d->attribute_buffer);
    const uint64_t ntri = d->index_num / 3u;
    if (d->index_offset + ntri * 3u * d->index_stride > ib.bytes) return SRG_ERR_INVALID;
    if (d->position_num && d->position_offset + (uint64_t)(d->position_num - 1) * d->position_stride + 12 > pb.bytes)
        return SRG_ERR_INVALID;
    if (d->attribute_num &&
        d->attribute_offset + (uint64_t)(d->attribute_num - 1) * d->attribute_stride + d->attribute_stride > ab.bytes)
        return SRG_ERR_INVALID;
    return SRG_OK;
}

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
    if (cfg && (cfg->struct_size != sizeof(srg_config) ||
                !layout_valid(cfg->bin_size ? cfg->bin_size : 32u, cfg->wg_per_tile ? cfg->wg_per_tile : 1u) ||
                cfg->tile_order > SRG_ORDER_LEARNED))
        return SRG_ERR_INVALID;
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
    {
        const void* kc[3] = { reinterpret_cast<const void*>(&k_clip<2>), reinterpret_cast<const void*>(&k_clip<kMaxSlots>),
                              reinterpret_cast<const void*>(&k_clip<kWideSlots>) };
        for (int i = 0; i < 3; ++i) {
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kc[i], (int)kClipThreads, 0) == hipSuccess &&
                per_cu > 0)
                ctx->clip_resident[i] = (uint32_t)per_cu * (uint32_t)ctx->num_cus;
        }
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return SRG_ERR_HIP; }
    if (hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess) {
        srg_destroy(ctx);
        return SRG_ERR_HIP;
    }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) { srg_destroy(ctx); return SRG_ERR_HIP; }
    if (hipHostMalloc((void**)&ctx->h_counters, sizeof(uint32_t) * C_NCOUNTERS, 0) != hipSuccess ||
        ensure_host_tables(ctx, SRG_NOALLOC_TABLE_BYTES) != SRG_OK) {
        srg_destroy(ctx);
        return SRG_ERR_NOMEM;
    }
    // srg_end_frame_noalloc's scratch on this device (looked up once per device)
    if (const int sst = scratch_address(ctx->device, &ctx->scratch_dev)) {
        srg_destroy(ctx);
        return sst;
    }
    *out = ctx;
    return SRG_OK;
}

void srg_destroy(srg_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->sched_params) (void)hipFree(ctx->sched_params);
    if (ctx->sched_adam) (void)hipFree(ctx->sched_adam);
    if (ctx->sched_ring) (void)hipFree(ctx->sched_ring);
    if (ctx->sched_head) (void)hipFree(ctx->sched_head);
    if (ctx->sched_gate) (void)hipFree(ctx->sched_gate);
    ctx->sched_feat.release();
    ctx->sched_pred.release();
    for (auto& kv : ctx->buffers)
        if (kv.second.owned) (void)hipFree(kv.second.dev);
    for (auto& kv : ctx->textures) (void)hipFree(kv.second.dev);
    for (srg_fb* fb : std::vector<srg_fb*>(ctx->fbs)) srg_fb_destroy(ctx, fb);
    ctx->tables.release();
    ctx->rec_raster.release();
    ctx->rec_shade.release();
    ctx->rec_ext.release();
    ctx->rec_meta.release();
    ctx->clip_queue.release();
    ctx->clip_base.release();
    ctx->clip_order.release();
    ctx->bin_count.release();
    ctx->bin_start.release();
    ctx->part_count.release();
    ctx->pairs.release();
    ctx->part_off.release();
    ctx->queue.release();
    ctx->queue2.release();
    ctx->order_bkt.release();
    ctx->order_hist.release();
    ctx->scan_blk.release();
    ctx->entries.release();
    ctx->bin_cost.release();
    ctx->counters.release();
    ctx->blit_staging.release();
    if (ctx->h_tables) (void)hipHostFree(ctx->h_tables);
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
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
    if (!ctx || !fb_known(ctx, fb)) return;
    for (srg_draw& d : ctx->draws)
        if (d.fb == fb) d.fb = nullptr;  // srg_end_frame rejects the frame
    for (auto& p : fb->planes) (void)finish_blit(p);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    for (auto& p : fb->planes) {
        if (p.colour) (void)hipFree(p.colour);
        if (p.depth) (void)hipFree(p.depth);
        if (p.primid) (void)hipFree(p.primid);
        if (p.pinned) (void)hipHostFree(p.pinned);
        if (p.read_done) (void)hipEventDestroy(p.read_done);
        if (p.blit_done) (void)hipEventDestroy(p.blit_done);
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
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
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
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
    const size_t n = (size_t)fb->tiles_x * fb->tiles_y * kTilePixels;
    PlaneState& p = fb->planes[plane];
    stream_wait_blit_read(ctx, p);
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
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
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

int srg_clear_n(srg_ctx* ctx, srg_fb* const* fbs, uint32_t n, uint32_t colour, int clear_colour, int clear_depth)
{
    if (!ctx || (n && !fbs)) return SRG_ERR_INVALID;
    // all or nothing: no plane is cleared when one handle is bad
    for (uint32_t i = 0; i < n; ++i) {
        if (!fbs[i]) return SRG_ERR_INVALID;
        if (!fb_known(ctx, fbs[i])) return SRG_ERR_HANDLE;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (const int st = srg_clear(ctx, fbs[i], colour, clear_colour, clear_depth)) return st;
    return SRG_OK;
}

int srg_draw_indexed(srg_ctx* ctx, const srg_draw* d)
{
    if (!ctx || !d) return SRG_ERR_INVALID;
    if (!ctx->in_frame) return SRG_ERR_STATE;
    if (const int st = validate_draw(ctx, d)) return st;
    ctx->draws.push_back(*d);
    if (d->plane == SRG_PLANE_WRITE) ctx->draws.back().plane = d->fb->write_plane;
    return SRG_OK;
}

int srg_draw_indexed_n(srg_ctx* ctx, const srg_draw* draws, uint32_t n)
{
    if (!ctx || (!draws && n)) return SRG_ERR_INVALID;
    for (uint32_t i = 0; i < n; ++i)
        if (const int st = srg_draw_indexed(ctx, draws + i)) return st;
    return SRG_OK;
}

static int end_frame(srg_ctx* ctx, bool noalloc)
{
    if (!ctx) return SRG_ERR_INVALID;
    if (!ctx->in_frame) return SRG_ERR_STATE;
    ctx->in_frame = false;
    ctx->stats = srg_frame_stats{};
    for (const srg_draw& d : ctx->draws) {
        // a released framebuffer leaves its draws with fb == nullptr
        const int vst = d.fb ? validate_draw(ctx, &d) : SRG_ERR_HANDLE;
        if (vst) {
            ctx->draws.clear();
            return vst;
        }
    }
    int st = run_frame(ctx, noalloc);
    ctx->draws.clear();
    return st;
}

int srg_end_frame(srg_ctx* ctx) { return end_frame(ctx, false); }

int srg_end_frame_noalloc(srg_ctx* ctx) { return end_frame(ctx, true); }

int srg_blit(srg_ctx* ctx, srg_fb* fb, void* linear, void (*on_finish)(void*), void* user)
{
    if (!ctx || !fb || !linear) return SRG_ERR_INVALID;
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
    PlaneState& p = fb->planes[fb->write_plane];
    if (!p.colour) return SRG_ERR_INVALID;
    int st = flush_pending_clear(ctx, fb, fb->write_plane);
    if (st) return st;
    const size_t npx = (size_t)fb->width * fb->height;
    if ((st = ctx->blit_staging.ensure(npx))) return st;
    hipLaunchKernelGGL(k_detile, dim3((fb->width + 255) / 256, fb->height), dim3(256), 0, ctx->stream, p.colour,
                       ctx->blit_staging.ptr, fb->width, fb->height, fb->tiles_x, 0u);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(linear, ctx->blit_staging.ptr, npx * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    fb->write_plane ^= 1u;
    if (on_finish) on_finish(user);
    return SRG_OK;
}

int srg_blit_async(srg_ctx* ctx, srg_fb* fb, void* linear, uint32_t flags, void (*on_finish)(void*), void* user)
{
    if (!ctx || !fb || !linear || (flags & ~SRG_BLIT_BGRA)) return SRG_ERR_INVALID;
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
    if (ctx->in_frame) return SRG_ERR_STATE;
    const uint32_t wp = fb->write_plane;
    PlaneState& p = fb->planes[wp];
    if (!p.colour) return SRG_ERR_INVALID;
    // the reference waits for the previous blit, of the other plane (Renderer.cpp:353)
    int st = finish_blit(fb->planes[wp ^ 1u]);
    if (!st) st = finish_blit(p);
    if (!st) st = flush_pending_clear(ctx, fb, wp);
    if (st) return st;
    const size_t npx = (size_t)fb->width * fb->height;
    if (!p.pinned && hipHostMalloc((void**)&p.pinned, npx * 4, 0) != hipSuccess) return SRG_ERR_NOMEM;
    if (!p.read_done) HIP_TRY(hipEventCreateWithFlags(&p.read_done, hipEventDisableTiming));
    if (!p.blit_done) HIP_TRY(hipEventCreateWithFlags(&p.blit_done, hipEventDisableTiming));
    uint32_t* dst = nullptr;
    HIP_TRY(hipHostGetDevicePointer((void**)&dst, p.pinned, 0));
    // the copy stream starts once the frame that wrote the plane is done
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
    HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->ev[3], 0));
    hipLaunchKernelGGL(k_detile, dim3((fb->width + 255) / 256, fb->height), dim3(256), 0, ctx->copy_stream, p.colour,
                       dst, fb->width, fb->height, fb->tiles_x, (flags & SRG_BLIT_BGRA) ? 1u : 0u);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(p.read_done, ctx->copy_stream));
    p.job = BlitJob{ linear, p.pinned, npx * 4, on_finish, user };
    HIP_TRY(hipLaunchHostFunc(ctx->copy_stream, blit_host_fn, &p.job));
    HIP_TRY(hipEventRecord(p.blit_done, ctx->copy_stream));
    p.blit_pending = true;
    fb->write_plane ^= 1u;
    return SRG_OK;
}

int srg_blit_wait(srg_ctx* ctx, srg_fb* fb)
{
    if (!ctx || !fb) return SRG_ERR_INVALID;
    if (!fb_known(ctx, fb)) return SRG_ERR_HANDLE;
    int st = finish_blit(fb->planes[0]);
    const int st1 = finish_blit(fb->planes[1]);
    return st ? st : st1;
}

int srg_get_frame_stats(srg_ctx* ctx, srg_frame_stats* out)
{
    if (!ctx || !out) return SRG_ERR_INVALID;
    *out = ctx->stats;
    return SRG_OK;
}

int srg_set_sponza_constants(srg_ctx* ctx, const srg_sponza_constants* k)
{
    static_assert(sizeof(srg_sponza_constants) == sizeof(SponzaDev), "SponzaScene::Constants layout");
    if (!ctx || !k) return SRG_ERR_INVALID;
    std::memcpy(&ctx->sponza, k, sizeof(SponzaDev));
    return SRG_OK;
}

int srg_get_tuning(srg_ctx* ctx, srg_tuning* out)
{
    if (!ctx || !out) return SRG_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->nlayouts = kNumLayouts;
    static_assert(kNumLayouts == SRG_TUNE_LAYOUTS, "tuning table size");
    const auto it = ctx->tunes.find(ctx->tune_sig);
    const TuneState* t = ((ctx->cfg.flags & SRG_CFG_AUTOTUNE) && it != ctx->tunes.end()) ? &it->second : nullptr;
    for (uint32_t i = 0; i < kNumLayouts; ++i) {
        out->layouts[i].bin_size = kLayouts[i].bin_size;
        out->layouts[i].wg_per_tile = kLayouts[i].parts;
        out->layouts[i].ms = t ? t->best_ms[i] : 0.0f;
    }
    if (t) {
        out->state = t->sweeping ? 1u : 2u;
        if (t->locked) {
            out->bin_size = t->chosen.bin_size;
            out->wg_per_tile = t->chosen.parts;
        }
    }
    return SRG_OK;
}

int srg_sched_get_model(srg_ctx* ctx, float* params, uint32_t count)
{
    if (!ctx || !params || count != kMlpParams) return SRG_ERR_INVALID;
    int st = ensure_sched(ctx);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(params, ctx->sched_params, sizeof(MlpParams), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SRG_OK;
}

int srg_sched_set_model(srg_ctx* ctx, const float* params, uint32_t count)
{
    if (!ctx || !params || count != kMlpParams) return SRG_ERR_INVALID;
    int st = ensure_sched(ctx);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(ctx->sched_params, params, sizeof(MlpParams), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->sched_trained = true;
    ctx->sched_fallback = false;
    ctx->sched_force_record = true;
    return SRG_OK;
}

int srg_sched_eval(srg_ctx* ctx, const float* features, uint32_t n, float* log2_cost)
{
    if (!ctx || (n && (!features || !log2_cost))) return SRG_ERR_INVALID;
    int st = ensure_sched(ctx);
    if (st || !n) return st;
    DevArray<float4> x;
    DevArray<float> out;
    if ((st = x.ensure(n)) || (st = out.ensure(n))) return st;
    int rc = SRG_OK;
    if (hipMemcpyAsync(x.ptr, features, sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = SRG_ERR_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_eval, dim3(std::min<uint32_t>((n + 3) / 4, 4096u)), dim3(256), 0, ctx->stream,
                           (const MlpParams*)ctx->sched_params, (const float4*)x.ptr, n, out.ptr);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(log2_cost, out.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = SRG_ERR_HIP;
    }
    (void)hipStreamSynchronize(ctx->stream);
    x.release();
    out.release();
    return rc;
}

int srg_sched_train(srg_ctx* ctx, const float* samples, uint32_t n, uint32_t steps, float lr, uint32_t seed)
{
    if (!ctx || !samples || n == 0 || n > kSchedRing) return SRG_ERR_INVALID;
    int st = ensure_sched(ctx);
    if (st) return st;
    static_assert(sizeof(SchedSample) == 5 * sizeof(float), "sample rows of 5 floats");
    static_assert(sizeof(AdamState) % sizeof(uint32_t) == 0, "Adam state in words");
    HIP_TRY(hipMemcpyAsync(ctx->sched_ring, samples, sizeof(SchedSample) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->sched_head, &n, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(16), dim3(256), 0, ctx->stream, (uint32_t*)ctx->sched_adam, 0u,
                       sizeof(AdamState) / sizeof(uint32_t));
    hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kTrainBatch), 0, ctx->stream, (MlpParams*)ctx->sched_params,
                       (AdamState*)ctx->sched_adam, (const SchedSample*)ctx->sched_ring,
                       (const uint32_t*)ctx->sched_head, kSchedRing, steps, lr, seed);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->sched_trained = true;
    ctx->sched_fallback = false;
    ctx->sched_force_record = true;
    return SRG_OK;
}

int srg_get_bin_stats(srg_ctx* ctx, uint32_t* counts, uint64_t* costs, uint32_t capacity, uint32_t* nbins)
{
    if (!ctx || !nbins) return SRG_ERR_INVALID;
    *nbins = ctx->last_bins;
    const uint32_t n = std::min(capacity, ctx->last_bins);
    if (counts && n) HIP_TRY(hipMemcpyAsync(counts, ctx->bin_count.ptr, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (costs && n) HIP_TRY(hipMemcpyAsync(costs, ctx->bin_cost.ptr, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
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

// ---- mip filter: stbir_resize_uint8 of stb_image_resize v0.96 ----
// Texture.cpp:196 builds each mip with stbir_resize_uint8(base, w, h, 0, out,
// mw, mh, 0, 4): Mitchell filter in both axes (scale <= 1), clamped edges,
// linear colour space, no alpha weighting (stb_image_resize.h:2464-2471).
// The library scatters every input sample into the outputs it reaches; here
// each output gathers the same weighted terms in the same order, so every
// float operation and rounding matches.

// stbir__filter_mitchell (stb_image_resize.h:824-836), float throughout.
static float mitchell(float x)
{
    x = std::fabs(x);
    if (x < 1.0f) return (16.0f + (x * x) * (21.0f * x - 36.0f)) / 18.0f;
    if (x < 2.0f) return (32.0f + x * (-60.0f + x * (36.0f - 7.0f * x))) / 18.0f;
    return 0.0f;
}

// Per output index o: taps [first[o], first[o+1]) of (clamped input index,
// weight), in ascending input order.
struct AxisTaps {
    std::vector<uint32_t> first;
    std::vector<uint32_t> src;
    std::vector<float> weight;
};

// Contributor ranges and weights (stb_image_resize.h:1023-1035, 1087-1115),
// per-output normalisation (:1117-1160), zero/edge trimming (:1162-1191).
static bool mitchell_axis(uint32_t in, uint32_t out, AxisTaps& t)
{
    const float scale = ((float)out / (float)in) / 1.0f;
    const float radius = 2.0f / scale;
    const int margin = (int)std::ceil((double)(4.0f / scale)) / 2;
    const int count = (int)in + 2 * margin;
    std::vector<int> n0(count), n1(count);
    std::vector<float> w((size_t)count * 4 + 4, 0.0f);
    for (int n = 0; n < count; ++n) {
        const float c = (float)(n - margin) + 0.5f;
        const float oc = c * scale - 0.0f;
        const int lo = (int)std::floor((double)((c - radius) * scale - 0.0f) + 0.5);
        const int hi = (int)std::floor((double)((c + radius) * scale - 0.0f) - 0.5);
        if (hi - lo > 3) return false;  // four taps per input (exact for power-of-two sizes)
        n0[n] = lo;
        n1[n] = hi;
        for (int i = 0; i <= hi - lo; ++i) w[(size_t)n * 4 + i] = mitchell(((float)(lo + i) + 0.5f) - oc) * scale;
        for (int i = hi - lo; i >= 0 && w[(size_t)n * 4 + i] == 0.0f; --i) n1[n] = lo + i - 1;
    }
    for (int o = 0; o < (int)out; ++o) {
        float sum = 0.0f;
        for (int n = 0; n < count && o >= n0[n]; ++n)
            if (o <= n1[n]) sum += w[(size_t)n * 4 + (o - n0[n])];
        const float inv = 1.0f / sum;
        for (int n = 0; n < count && o >= n0[n]; ++n)
            if (o <= n1[n]) w[(size_t)n * 4 + (o - n0[n])] *= inv;
    }
    std::vector<std::vector<std::pair<uint32_t, float>>> gather(out);
    for (int n = 0; n < count; ++n) {
        float* wn = &w[(size_t)n * 4];
        int skip = 0;
        while (skip < 4 && wn[skip] == 0.0f) ++skip;
        int a = n0[n] + skip;
        if (a < 0) {
            skip -= a;
            a = 0;
        }
        const int b = std::min(n1[n], (int)out - 1);
        const uint32_t srcn = (uint32_t)std::min(std::max(n - margin, 0), (int)in - 1);
        for (int k = a; k <= b; ++k) {
            const int i = k - a;
            // a weight the shift does not reach keeps its unshifted value
            const float wk = (i + skip < 4) ? wn[i + skip] : wn[i];
            gather[k].emplace_back(srcn, wk);
        }
    }
    t.first.assign(out + 1, 0);
    t.src.clear();
    t.weight.clear();
    for (uint32_t o = 0; o < out; ++o) {
        t.first[o] = (uint32_t)t.src.size();
        for (auto& p : gather[o]) {
            t.src.push_back(p.first);
            t.weight.push_back(p.second);
        }
    }
    t.first[out] = (uint32_t)t.src.size();
    return true;
}

// Runs f(i) for i in [0, n) on up to 16 host threads.
extern "C++" template <class F>
static void parallel_rows(uint32_t n, F f)
{
    const uint32_t nt = std::max(1u, std::min({ n, 16u, std::max(1u, std::thread::hardware_concurrency()) }));
    std::atomic<uint32_t> next{ 0 };
    auto run = [&]() {
        for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i);
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(run);
    run();
    for (auto& th : pool) th.join();
}

static bool mitchell_resize_rgba8(const uint8_t* in, uint32_t w, uint32_t h, uint8_t* out, uint32_t ow, uint32_t oh)
{
    AxisTaps tx, ty;
    if (!mitchell_axis(w, ow, tx) || !mitchell_axis(h, oh, ty)) return false;
    float dec[256];
    for (int i = 0; i < 256; ++i) dec[i] = (float)i / 255.0f;
    // horizontal pass for every input row the vertical taps name
    std::vector<float> hrows((size_t)h * ow * 4);
    parallel_rows(h, [&](uint32_t y) {
        const uint8_t* src = in + (size_t)y * w * 4;
        float* dst = &hrows[(size_t)y * ow * 4];
        for (uint32_t x = 0; x < ow; ++x) {
            float a[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
            for (uint32_t t = tx.first[x]; t < tx.first[x + 1]; ++t) {
                const uint8_t* p = src + (size_t)tx.src[t] * 4;
                const float k = tx.weight[t];
                for (int c = 0; c < 4; ++c) a[c] += dec[p[c]] * k;
            }
            for (int c = 0; c < 4; ++c) dst[x * 4 + c] = a[c];
        }
    });
    // vertical pass and encode (stb_image_resize.h:1736-1757)
    parallel_rows(oh, [&](uint32_t y) {
        for (uint32_t x = 0; x < ow; ++x)
            for (int c = 0; c < 4; ++c) {
                float a = 0.0f;
                for (uint32_t t = ty.first[y]; t < ty.first[y + 1]; ++t)
                    a += hrows[((size_t)ty.src[t] * ow + x) * 4 + c] * ty.weight[t];
                a = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
                out[((size_t)y * ow + x) * 4 + c] = (uint8_t)(int)((double)(a * 255.0f) + 0.5);
            }
    });
    return true;
}

// Layout of CreateFromRGBA8 (Texture.cpp:119-177): power-of-two sizes, each
// a multiple of the 32-texel tile and below 2^14; a full chain when mips are
// made; every level padded to whole 32x32 tiles.
struct TexLayout {
    uint32_t wl = 0, hl = 0, n = 0;
    uint32_t offs[14] = {};
    uint64_t total = 0;
};

static int texture_layout(uint32_t width, uint32_t height, int calc_mips, TexLayout& L)
{
    auto is_pow2 = [](uint32_t v) { return v && !(v & (v - 1)); };
    if (!is_pow2(width) || !is_pow2(height) || width < 32 || height < 32 || width >= 16384 || height >= 16384)
        return SRG_ERR_INVALID;
    while ((1u << L.wl) < width) ++L.wl;
    while ((1u << L.hl) < height) ++L.hl;
    L.n = calc_mips ? std::max(L.wl, L.hl) + 1 : 1;
    L.total = 0;
    for (uint32_t m = 0; m < L.n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        L.offs[m] = (uint32_t)L.total;
        L.total += (uint64_t)((mw + 31u) & ~31u) * ((mh + 31u) & ~31u) * 4u;
    }
    return SRG_OK;
}

int srg_texture_build_rgba8(const uint8_t* linear, uint32_t width, uint32_t height, int calc_mips, uint8_t* out,
                            uint64_t* out_bytes, uint32_t mip_offsets[14], uint32_t* wlog2, uint32_t* hlog2,
                            uint32_t* nmips)
{
    TexLayout L;
    if (!out_bytes) return SRG_ERR_INVALID;
    int st = texture_layout(width, height, calc_mips, L);
    if (st) return st;
    const uint32_t n = L.n;
    const uint32_t* offs = L.offs;
    const uint64_t total = L.total;
    *out_bytes = total;
    if (mip_offsets) std::memcpy(mip_offsets, offs, sizeof(L.offs));
    if (wlog2) *wlog2 = L.wl;
    if (hlog2) *hlog2 = L.hl;
    if (nmips) *nmips = n;
    if (!out) return SRG_OK;
    if (!linear) return SRG_ERR_INVALID;
    std::memset(out, 0, total);
    tile_texture(linear, out, width, height);
    // Every mip is resized from the base level (Texture.cpp:188-198).
    std::vector<uint8_t> tmp;
    for (uint32_t m = 1; m < n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        tmp.assign((size_t)mw * mh * 4, 0);
        if (!mitchell_resize_rgba8(linear, width, height, tmp.data(), mw, mh)) return SRG_ERR_INVALID;
        tile_texture(tmp.data(), out + offs[m], mw, mh);
    }
    return SRG_OK;
}

int srg_texture_create_rgba8(srg_ctx* ctx, const uint8_t* linear, uint32_t width, uint32_t height, int calc_mips,
                             srg_handle* out)
{
    TexLayout L;
    if (!ctx || !out || !linear) return SRG_ERR_INVALID;
    int st = texture_layout(width, height, calc_mips, L);
    if (st) return st;
    // per-level tap lists, built on the host before anything is allocated
    std::vector<AxisTaps> tx(L.n), ty(L.n);
    size_t ntaps = 0;
    for (uint32_t m = 1; m < L.n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        if (!mitchell_axis(width, mw, tx[m]) || !mitchell_axis(height, mh, ty[m])) return SRG_ERR_INVALID;
        // words: first (n + 1), src and weight per axis
        ntaps += tx[m].first.size() + 2 * tx[m].src.size() + ty[m].first.size() + 2 * ty[m].src.size();
    }
    Texture t;
    t.bytes = L.total;
    t.wlog2 = L.wl;
    t.hlog2 = L.hl;
    t.nmips = L.n;
    std::memcpy(t.mip_offsets, L.offs, sizeof(L.offs));
    uint32_t* base = nullptr;
    float4* rows = nullptr;
    uint8_t* taps = nullptr;
    std::vector<uint8_t> htaps;
    const size_t row_bytes = L.n > 1 ? sizeof(float4) * (size_t)height * std::max(1u, width >> 1) : 0;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(ctx->stream);
        if (base) (void)hipFree(base);
        if (rows) (void)hipFree(rows);
        if (taps) (void)hipFree(taps);
        if (code && t.dev) (void)hipFree(t.dev);
        return code;
    };
    if (hipMalloc((void**)&t.dev, L.total) != hipSuccess || hipMalloc((void**)&base, (size_t)width * height * 4) != hipSuccess ||
        (row_bytes && hipMalloc((void**)&rows, row_bytes) != hipSuccess) ||
        (ntaps && hipMalloc((void**)&taps, ntaps * 4) != hipSuccess))
        return cleanup(SRG_ERR_NOMEM);
    if (hipMemcpyAsync(base, linear, (size_t)width * height * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return cleanup(SRG_ERR_HIP);
    // all tap lists in one upload: per level first_x, src_x, w_x, first_y, src_y, w_y
    htaps.resize(ntaps * 4);
    std::vector<MipTaps> dx(L.n), dy(L.n);
    size_t at = 0;
    auto put = [&](const void* p, size_t words) {
        const size_t o = at;
        if (at + words <= ntaps) std::memcpy(htaps.data() + o * 4, p, words * 4);
        at += words;
        return o;
    };
    for (uint32_t m = 1; m < L.n; ++m) {
        const size_t fx = put(tx[m].first.data(), tx[m].first.size()), sx = put(tx[m].src.data(), tx[m].src.size());
        const size_t wx = put(tx[m].weight.data(), tx[m].weight.size());
        const size_t fy = put(ty[m].first.data(), ty[m].first.size()), sy = put(ty[m].src.data(), ty[m].src.size());
        const size_t wy = put(ty[m].weight.data(), ty[m].weight.size());
        const uint32_t* T = (const uint32_t*)taps;
        dx[m] = MipTaps{ T + fx, T + sx, (const float*)(T + wx) };
        dy[m] = MipTaps{ T + fy, T + sy, (const float*)(T + wy) };
    }
    if (at != ntaps) return cleanup(SRG_ERR_INVALID);  // tap lists must fill the upload exactly
    if (ntaps && hipMemcpyAsync(taps, htaps.data(), at * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return cleanup(SRG_ERR_HIP);
    for (uint32_t m = 0; m < L.n; ++m) {
        const uint32_t mw = std::max(1u, width >> m), mh = std::max(1u, height >> m);
        const uint32_t tiles_w = (mw + 31u) >> 5, tiles_h = (mh + 31u) >> 5;
        const uint32_t ntex = tiles_w * tiles_h * 1024u;
        uint32_t* dst = (uint32_t*)(t.dev + L.offs[m]);
        if (m == 0) {
            hipLaunchKernelGGL(k_tex_tile, dim3((ntex + 255) / 256), dim3(256), 0, ctx->stream, base, width, height, dst,
                               tiles_w, ntex);
        } else {
            hipLaunchKernelGGL(k_mip_rows, dim3((height * mw + 255) / 256), dim3(256), 0, ctx->stream, base, width, height,
                               mw, dx[m], rows);
            hipLaunchKernelGGL(k_mip_cols, dim3((ntex + 255) / 256), dim3(256), 0, ctx->stream, rows, mw, mh, dy[m], dst,
                               tiles_w, ntex);
        }
        if (hipGetLastError() != hipSuccess) return cleanup(SRG_ERR_HIP);
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return cleanup(SRG_ERR_HIP);
    cleanup(SRG_OK);
    const uint64_t h = new_handle(ctx);
    ctx->textures[h] = t;
    *out = h;
    return SRG_OK;
}

int srg_texture_read(srg_ctx* ctx, srg_handle h, uint8_t* out, uint64_t capacity, uint64_t* bytes,
                     uint32_t mip_offsets[14], uint32_t* wlog2, uint32_t* hlog2, uint32_t* nmips)
{
    if (!ctx) return SRG_ERR_INVALID;
    auto it = ctx->textures.find(h);
    if (it == ctx->textures.end()) return SRG_ERR_HANDLE;
    const Texture& t = it->second;
    if (bytes) *bytes = t.bytes;
    if (mip_offsets) std::memcpy(mip_offsets, t.mip_offsets, sizeof(t.mip_offsets));
    if (wlog2) *wlog2 = t.wlog2;
    if (hlog2) *hlog2 = t.hlog2;
    if (nmips) *nmips = t.nmips;
    if (!out) return SRG_OK;
    if (capacity < t.bytes) return SRG_ERR_INVALID;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t.bytes) HIP_TRY(hipMemcpy(out, t.dev, t.bytes, hipMemcpyDeviceToHost));
    return SRG_OK;
}

}  // extern "C"

struct srg_obj_model {
    srg_obj::Model model;
};

namespace {

std::string dir_of(const char* path)
{
    const char* slash = std::strrchr(path, '/');
    return slash ? std::string(path, (size_t)(slash - path) + 1) : std::string();
}

int obj_parse(const char* text, size_t bytes, std::string dir, uint32_t flags, srg_obj_model** out)
{
    srg_obj_model* m = new (std::nothrow) srg_obj_model();
    if (!m) return SRG_ERR_NOMEM;
    srg_obj::Parser parser(m->model, flags, std::move(dir));
    if (!srg_obj::parse_text(parser, text, bytes)) {
        delete m;
        return SRG_ERR_INVALID;
    }
    *out = m;
    return SRG_OK;
}

}  // namespace

extern "C" {

int srg_obj_parse(const char* text, size_t bytes, const char* dir, uint32_t flags, srg_obj_model** out)
{
    if (!out || (!text && bytes)) return SRG_ERR_INVALID;
    *out = nullptr;
    std::string d = dir ? std::string(dir) : std::string();
    if (!d.empty() && d.back() != '/') d += '/';
    return obj_parse(text, bytes, std::move(d), flags, out);
}

int srg_obj_load(const char* path, uint32_t flags, srg_obj_model** out)
{
    if (!out || !path) return SRG_ERR_INVALID;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return SRG_ERR_INVALID;
    std::vector<char> text;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.insert(text.end(), buf, buf + n);
    std::fclose(f);
    return obj_parse(text.data(), text.size(), dir_of(path), flags, out);
}

uint32_t srg_obj_mesh_count(const srg_obj_model* m) { return m ? (uint32_t)m->model.meshes.size() : 0u; }

int srg_obj_get_mesh(const srg_obj_model* m, uint32_t index, srg_obj_mesh* out)
{
    if (!m || !out || index >= m->model.meshes.size()) return SRG_ERR_INVALID;
    const srg_obj::Mesh& s = m->model.meshes[index];
    out->index_stride = s.index_stride;
    out->num_indices = s.num_indices;
    out->indices = s.index_data.data();
    out->num_vertices = (uint32_t)s.vertices.size();
    out->vertices = s.vertices.empty() ? nullptr : s.vertices[0].pos;
    out->material = s.material;
    return SRG_OK;
}

uint32_t srg_obj_material_count(const srg_obj_model* m) { return m ? (uint32_t)m->model.materials.size() : 0u; }

const char* srg_obj_material_name(const srg_obj_model* m, uint32_t index)
{
    return (m && index < m->model.materials.size()) ? m->model.materials[index].name.c_str() : nullptr;
}

const char* srg_obj_material_diffuse(const srg_obj_model* m, uint32_t index)
{
    return (m && index < m->model.materials.size()) ? m->model.materials[index].diffuse.c_str() : nullptr;
}

void srg_obj_free(srg_obj_model* m) { delete m; }

}  // extern "C"
