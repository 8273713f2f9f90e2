// sr_sched.h -- learned work-bin cost model for the back end's queue order.
//
// The back end is persistent and takes bins from a queue; starting the most
// expensive bins first shortens the frame's tail. Every kernel here changes
// only the order bins are processed in, never the arithmetic of a pixel.
//
//   k_bin_predict   one wave per work bin: the features, from its entry list:
//                   x0 = log2(1 + entries) (the bin count),
//                   x1 = log2(1 + summed clipped triangle area / bin area),
//                        a triangle's clipped area taken as the smaller of
//                        its area and its bounding box inside the bin,
//                   x2 = log2(1 + depth-range overlap): the entries' depth
//                        extents over the bin summed, over the span of all of
//                        them (about 1 when the ranges tile the span, up to
//                        the entry count when they all coincide),
//                   x3 = mean mip level of the textured entries (from the
//                        uv planes at the first vertex, 0 without texture);
//                   the sums are estimated from at most 16 entries spread
//                   evenly over the list, scaled by entries / sampled. Then,
//                   once trained, the 4 -> 32 -> 32 -> 1 ReLU MLP (fp32) ->
//                   predicted log2 cost, one hidden unit per lane
//   k_mlp_record    after the back end: (features, log2 measured s_memtime
//                   cost) samples into a ring buffer, and the model's error
//   k_mlp_train     one workgroup, mini-batch Adam on the recorded samples
#pragma once
#include "sr_backend.h"

namespace srg {

constexpr uint32_t kFeat = 4, kHid = 32;
constexpr uint32_t kMlpParams = kHid * kFeat + kHid + kHid * kHid + kHid + kHid + 1;  // 1249

struct MlpParams {
    float w1[kHid][kFeat];
    float b1[kHid];
    float w2[kHid][kHid];
    float b2[kHid];
    float w3[kHid];
    float b3;
};
static_assert(sizeof(MlpParams) == kMlpParams * sizeof(float), "packed parameters");

struct SchedSample {
    float x[kFeat];
    float y;  // log2 measured cost
};

// The MLP as k_bin_predict holds it in LDS: layer 2 transposed, so that lane
// j reading weight (j, i) for every i walks consecutive banks.
struct MlpLds {
    float w1[kHid][kFeat];
    float b1[kHid];
    float w2t[kHid][kHid];  // w2t[i][j] = w2[j][i]
    float b2[kHid];
    float w3[kHid];
    float b3;
};

// The model's output, the predicted log2 cost, for features x (one wave, the
// result in every lane): hidden unit j = lane & 31 (the upper half repeats
// it); layer 2 reads unit i of layer 1 from lane i.
__device__ __forceinline__ float mlp_infer_wave(const MlpLds& P, const float x[kFeat], uint32_t lane)
{
    const uint32_t j = lane & 31u;
    float a = P.b1[j];
#pragma unroll
    for (uint32_t k = 0; k < kFeat; ++k) a += P.w1[j][k] * x[k];
    const float h1 = a > 0.0f ? a : 0.0f;
    float b = P.b2[j];
#pragma unroll
    for (uint32_t i = 0; i < kHid; ++i) b += P.w2t[i][j] * __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, h1), (int)i));
    const float h2 = b > 0.0f ? b : 0.0f;
    float o = lane < kHid ? P.w3[j] * h2 : 0.0f;
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) o += __shfl_xor(o, m, 64);
    return P.b3 + o;
}

// The parameters into LDS (layer 2 transposed); the whole workgroup takes
// part, a barrier follows.
__device__ __forceinline__ void load_mlp_lds(MlpLds& P, const MlpParams* params)
{
    for (uint32_t i = threadIdx.x; i < kHid * kFeat; i += blockDim.x) (&P.w1[0][0])[i] = (&params->w1[0][0])[i];
    for (uint32_t i = threadIdx.x; i < kHid * kHid; i += blockDim.x) P.w2t[i % kHid][i / kHid] = (&params->w2[0][0])[i];
    if (threadIdx.x < kHid) {
        P.b1[threadIdx.x] = params->b1[threadIdx.x];
        P.b2[threadIdx.x] = params->b2[threadIdx.x];
        P.w3[threadIdx.x] = params->w3[threadIdx.x];
    }
    if (threadIdx.x == 0) P.b3 = params->b3;
    __syncthreads();
}

// Features of bin gb and, once trained, its predicted cost (one wave).
__device__ __forceinline__ void bin_predict_one(const FrameDev& F, const MlpLds& P, bool trained, uint32_t gb,
                                                uint32_t lane, float4* feat, float* pred)
{
    const ViewDev V = F.views[view_of_bin(F, gb)];
    const uint32_t S = F.bin_size, K = F.bin_k;
    const uint32_t lb = gb - V.bin_base, sbx = lb % V.bins_x, sby = lb / V.bins_x;
    const int32_t X0 = (int32_t)((sbx / K) * kBin), Y0 = (int32_t)((sby / K) * kBin);
    const uint32_t sx0 = (sbx % K) * S, sy0 = (sby % K) * S;
    const uint32_t beg = F.bin_start[gb], cnt = F.bin_count[gb];
    const uint32_t ns = cnt < 16u ? cnt : 16u;  // sampled entries: lane j takes entry j * cnt / ns
    float area = 0.0f, zlo = 3.0e38f, zhi = -3.0e38f, zext = 0.0f, mip = 0.0f, textured = 0.0f;
    if (lane < ns) {
        const uint32_t e = (uint32_t)(((uint64_t)lane * cnt) / ns);
        const uint32_t rec = F.entries[beg + e];
        const uint4 q0 = F.rec_raster[(size_t)rec * 4], q2 = F.rec_raster[(size_t)rec * 4 + 2];
        const uint4 q3 = F.rec_raster[(size_t)rec * 4 + 3];
        const int32_t xmin = q2.y & 0x3FFF, xmax = (q2.y >> 16) & 0x3FFF, ymin = q2.z & 0x3FFF, ymax = (q2.z >> 16) & 0x3FFF;
        // bounding box inside the bin (pixels, ends exclusive)
        const int32_t bx0 = max(xmin - X0, (int32_t)sx0), bx1 = min(xmax - X0 + 1, (int32_t)(sx0 + S));
        const int32_t by0 = max(ymin - Y0, (int32_t)sy0), by1 = min(ymax - Y0 + 1, (int32_t)(sy0 + S));
        const float box = (float)max(0, bx1 - bx0) * (float)max(0, by1 - by0);
        // the three edge constants sum to the doubled area in subpixel x pixel units
        const float tri = (float)max(0, (int32_t)(q0.x + q0.y + q0.z)) * (1.0f / 512.0f);
        area = fminf(tri, box);
        // depth over the box corners: z = zdx (x - r0x) + zdy (y - r0y) + z0
        const float zdx = ffrom(q3.y), zdy = ffrom(q3.z);
        const float ux = (float)(X0 + bx0) - ffrom(q2.w), uy = (float)(Y0 + by0) - ffrom(q3.x);
        const float za = zdx * ux + zdy * uy + ffrom(q3.w);
        const float wx = zdx * (float)max(0, bx1 - bx0), wy = zdy * (float)max(0, by1 - by0);
        const float z0 = fminf(za, za + wx) + fminf(0.0f, wy), z1 = fmaxf(za, za + wx) + fmaxf(0.0f, wy);
        if (z0 == z0 && z1 == z1 && z1 - z0 < 1.0e30f) {
            zlo = z0;
            zhi = z1;
            zext = z1 - z0;
        }
        // mip level from the uv planes at the first vertex (slots 0 and 1)
        const float* Sp = F.rec_shade + (size_t)rec * F.shade_stride;
        const float4 s0 = make_float4(Sp[0], Sp[1], Sp[2], Sp[3]), s1 = make_float4(Sp[4], Sp[5], Sp[6], Sp[7]);
        const DrawDev& D = F.draws[F.rec_meta[rec]];
        const uint32_t ti = D.tex_info;
        if (ti && D.deriv_slot == 0) {
            const float g = fmaxf(fmaxf(fabsf(s0.w), fabsf(s1.x)), fmaxf(fabsf(s1.z), fabsf(s1.w))) /
                            fmaxf(fabsf(s0.z), 1.0e-30f) * (float)(1u << (ti & 31u));
            const float m = g > 1.0f ? __log2f(g) : 0.0f;
            mip = m == m ? fminf(m, (float)((ti >> 8) ? (ti >> 8) - 1u : 0u)) : 0.0f;
            textured = 1.0f;
        }
    }
    area = wave_sum_f(area);
    zext = wave_sum_f(zext);
    zlo = wave_min_f(zlo);
    zhi = wave_max_f(zhi);
    mip = wave_sum_f(mip);
    textured = wave_sum_f(textured);
    const float scale = ns ? (float)cnt / (float)ns : 0.0f;
    const float overlap = zhi > zlo ? zext / (zhi - zlo) : (float)ns;  // one shared depth: all coincide
    const float x[kFeat] = { __log2f(1.0f + (float)cnt), __log2f(1.0f + scale * area / (float)(S * S)),
                             __log2f(1.0f + scale * overlap), textured > 0.0f ? mip / textured : 0.0f };
    if (lane == 0) feat[gb] = make_float4(x[0], x[1], x[2], x[3]);
    if (!trained) return;
    const float o = mlp_infer_wave(P, x, lane);
    if (lane == 0) pred[gb] = cnt ? exp2f(fminf(o, 120.0f)) : 0.0f;
}

__device__ __forceinline__ float mlp_forward(const MlpParams& P, const float x[kFeat], float h1[kHid], float h2[kHid])
{
#pragma unroll
    for (uint32_t i = 0; i < kHid; ++i) {
        float a = P.b1[i];
#pragma unroll
        for (uint32_t j = 0; j < kFeat; ++j) a += P.w1[i][j] * x[j];
        h1[i] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll
    for (uint32_t i = 0; i < kHid; ++i) {
        float a = P.b2[i];
#pragma unroll
        for (uint32_t j = 0; j < kHid; ++j) a += P.w2[i][j] * h1[j];
        h2[i] = a > 0.0f ? a : 0.0f;
    }
    float o = P.b3;
#pragma unroll
    for (uint32_t i = 0; i < kHid; ++i) o += P.w3[i] * h2[i];
    return o;
}

// Grid-stride over bins, one wave per bin at a time, so each workgroup loads
// the parameters once.
__global__ __launch_bounds__(256) void k_bin_predict(FrameDev F, float4* feat, const MlpParams* params, float* pred)
{
    __shared__ MlpLds P;
    // An overflowed frame has bin starts past the entry arena and no entry
    // lists (k_emit skipped it); it is rendered again with larger arenas.
    if (F.counters[C_FLAGS] & (FLAG_ENTRY_OVERFLOW | FLAG_CLIP_OVERFLOW)) return;
    if (F.sched_gate && !F.sched_gate[1]) return;  // a frame between recordings keeps the last prediction
    if (params) load_mlp_lds(P, params);
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t gb = blockIdx.x * 4u + (threadIdx.x >> 6); gb < F.total_bins; gb += gridDim.x * 4u)
        bin_predict_one(F, P, params != nullptr, gb, lane, feat, pred);
}

// The model's output for n given feature rows (srg_sched_eval): the inference
// k_bin_predict runs, one wave per row.
__global__ __launch_bounds__(256) void k_mlp_eval(const MlpParams* params, const float4* x, uint32_t n, float* out)
{
    __shared__ MlpLds P;
    load_mlp_lds(P, params);
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6); r < n; r += gridDim.x * 4u) {
        const float4 f = x[r];
        const float xx[kFeat] = { f.x, f.y, f.z, f.w };
        const float o = mlp_infer_wave(P, xx, lane);
        if (lane == 0) out[r] = o;
    }
}

// After the back end: append (features, log2 cost) of every non-empty bin to
// the ring buffer and accumulate |predicted - measured| in log2 units.
__global__ __launch_bounds__(256) void k_mlp_record(FrameDev F, const float4* feat, const float* pred,
                                                    SchedSample* ring, uint32_t ring_cap, uint32_t* ring_head)
{
    if (F.sched_gate && !F.sched_gate[1]) return;  // not a recording frame
    const uint32_t gb = blockIdx.x * 256u + threadIdx.x;
    // a frame that overflowed an arena is rendered again; its costs are not recorded
    const bool clean = (F.counters[C_FLAGS] & (FLAG_CLIP_OVERFLOW | FLAG_ENTRY_OVERFLOW)) == 0;
    const bool ok = clean && gb < F.total_bins && F.bin_count[gb] > 0 && F.bin_cost[gb] > 0;
    const uint32_t slot = wave_add_amount(ring_head, ok ? 1u : 0u);
    const uint32_t nok = wave_sum(ok ? 1u : 0u);
    const float y = ok ? __log2f((float)F.bin_cost[gb]) : 0.0f;
    float err = (ok && pred) ? fabsf(__log2f(fmaxf(pred[gb], 1.0f)) - y) : 0.0f;
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) err += __shfl_xor(err, m, 64);
    if (lane_id() == 0 && nok) {  // one atomic per wave, not per bin
        atomicAdd(&F.counters[C_SCHED_N], nok);
        if (pred) atomicAdd((float*)&F.counters[C_SCHED_ERR], err);
    }
    if (!ok) return;
    const float4 f = feat[gb];
    SchedSample s;
    s.x[0] = f.x; s.x[1] = f.y; s.x[2] = f.z; s.x[3] = f.w;
    s.y = y;
    ring[slot % ring_cap] = s;
}

struct AdamState {
    float m[kMlpParams];
    float v[kMlpParams];
    uint32_t t;
};

constexpr uint32_t kTrainBatch = 256;

// Mini-batch Adam on the squared log2-cost error, one workgroup of 256
// threads, one sample per thread per step. Activations and output deltas of
// the batch live in LDS; each thread then owns a slice of the parameters and
// sums its gradient over the batch.
__global__ __launch_bounds__(kTrainBatch) void k_mlp_train(MlpParams* params, AdamState* adam, const SchedSample* ring,
                                                           const uint32_t* ring_head, uint32_t ring_cap, uint32_t steps,
                                                           float lr, uint32_t seed)
{
    __shared__ MlpParams P;
    __shared__ float sx[kTrainBatch][kFeat];
    __shared__ float sh1[kTrainBatch][kHid];
    __shared__ float sh2[kTrainBatch][kHid];
    __shared__ float sd2[kTrainBatch][kHid];   // dL/d(pre-activation of layer 2)
    __shared__ float sd1[kTrainBatch][kHid];   // dL/d(pre-activation of layer 1)
    __shared__ float sdo[kTrainBatch];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kMlpParams; i += kTrainBatch) ((float*)&P)[i] = ((const float*)params)[i];
    const uint32_t n = min(*ring_head, ring_cap);
    __syncthreads();
    if (n == 0) return;
    uint32_t rng = seed * 747796405u + t * 2891336453u + 1u;
    const uint32_t t0 = adam->t;
    const float b1c = 0.9f, b2c = 0.999f, eps = 1e-8f;
    for (uint32_t step = 0; step < steps; ++step) {
        rng = rng * 1664525u + 1013904223u;
        const SchedSample s = ring[(rng >> 8) % n];
        float x[kFeat] = { s.x[0], s.x[1], s.x[2], s.x[3] };
        float h1[kHid], h2[kHid];
        const float o = mlp_forward(P, x, h1, h2);
        const float dout = 2.0f * (o - s.y) / (float)kTrainBatch;
        for (uint32_t j = 0; j < kFeat; ++j) sx[t][j] = x[j];
        for (uint32_t i = 0; i < kHid; ++i) {
            sh1[t][i] = h1[i];
            sh2[t][i] = h2[i];
            sd2[t][i] = h2[i] > 0.0f ? dout * P.w3[i] : 0.0f;
        }
        sdo[t] = dout;
        for (uint32_t j = 0; j < kHid; ++j) {
            float a = 0.0f;
            for (uint32_t i = 0; i < kHid; ++i) a += P.w2[i][j] * sd2[t][i];
            sd1[t][j] = h1[j] > 0.0f ? a : 0.0f;
        }
        __syncthreads();
        // gradients: thread t owns parameters t, t + 256, ...
        const float tt = (float)(t0 + step + 1u);
        for (uint32_t p = t; p < kMlpParams; p += kTrainBatch) {
            float g = 0.0f;
            if (p < kHid * kFeat) {
                const uint32_t i = p / kFeat, j = p % kFeat;
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sd1[b][i] * sx[b][j];
            } else if (p < kHid * kFeat + kHid) {
                const uint32_t i = p - kHid * kFeat;
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sd1[b][i];
            } else if (p < kHid * kFeat + kHid + kHid * kHid) {
                const uint32_t q = p - (kHid * kFeat + kHid), i = q / kHid, j = q % kHid;
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sd2[b][i] * sh1[b][j];
            } else if (p < kHid * kFeat + 2 * kHid + kHid * kHid) {
                const uint32_t i = p - (kHid * kFeat + kHid + kHid * kHid);
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sd2[b][i];
            } else if (p < kMlpParams - 1) {
                const uint32_t i = p - (kHid * kFeat + 2 * kHid + kHid * kHid);
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sdo[b] * sh2[b][i];
            } else {
                for (uint32_t b = 0; b < kTrainBatch; ++b) g += sdo[b];
            }
            const float m = b1c * adam->m[p] + (1.0f - b1c) * g;
            const float v = b2c * adam->v[p] + (1.0f - b2c) * g * g;
            adam->m[p] = m;
            adam->v[p] = v;
            const float mh = m / (1.0f - powf(b1c, tt)), vh = v / (1.0f - powf(b2c, tt));
            ((float*)&P)[p] -= lr * mh / (sqrtf(vh) + eps);
        }
        __syncthreads();
    }
    for (uint32_t i = t; i < kMlpParams; i += kTrainBatch) ((float*)params)[i] = ((const float*)&P)[i];
    if (t == 0) adam->t = t0 + steps;
}

// Deterministic small initialisation (He-style scale) and cleared Adam state.
__global__ void k_mlp_init(MlpParams* params, AdamState* adam, uint32_t seed)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    uint32_t h = (i + 1u) * 2654435761u ^ seed;
    h ^= h >> 16;
    h *= 2246822519u;
    h ^= h >> 13;
    const float u = (float)(h & 0xFFFFFFu) / 16777216.0f - 0.5f;
    float scale = 0.0f;
    if (i < kHid * kFeat) scale = 2.0f * sqrtf(2.0f / kFeat);
    else if (i >= kHid * kFeat + kHid && i < kHid * kFeat + kHid + kHid * kHid) scale = 2.0f * sqrtf(2.0f / kHid);
    else if (i >= kHid * kFeat + 2 * kHid && i < kMlpParams - 1) scale = 2.0f * sqrtf(1.0f / kHid);
    ((float*)params)[i] = (i == kMlpParams - 1) ? 16.0f : u * scale;  // bias: about 2^16 ticks
    adam->m[i] = 0.0f;
    adam->v[i] = 0.0f;
    if (i == 0) adam->t = 0;
}

}  // namespace srg
