/*
 * srgpu.h -- C ABI of the MI355X (gfx950) sort-middle rasteriser.
 *
 * This is the drop-in boundary for SoftRast's frame / draw-call API
 * (SoftRast/Renderer.h:110-177). Every entry point takes plain pointers and
 * sizes and returns an int status (0 = SRG_OK, negative = error); nothing
 * asserts. The C++ facade in include/sr/Renderer.h re-creates the
 * reference's RenderContext / FrameBuffer / DrawCall types on top of it.
 *
 * Threading: a context is driven from one host thread, as the reference's
 * RenderContext is (TaskSystem.cpp:25). Contexts are independent.
 */
#ifndef SRGPU_H
#define SRGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRG_API_VERSION 1u

typedef struct srg_ctx srg_ctx;
typedef struct srg_fb srg_fb;
typedef uint64_t srg_handle; /* 0 is never a valid handle */

enum srg_status {
    SRG_OK = 0,
    SRG_ERR_INVALID = -1,        /* bad argument or draw-call state */
    SRG_ERR_NOMEM = -2,          /* host or device allocation failed */
    SRG_ERR_HIP = -3,            /* a HIP runtime call failed */
    SRG_ERR_INDEX_RANGE = -4,    /* an index addressed past the vertex buffer */
    SRG_ERR_OVERFLOW = -5,       /* a device arena stayed too small after growing */
    SRG_ERR_UNKNOWN_SHADER = -6, /* shader id has no device implementation */
    SRG_ERR_STATE = -7,          /* call out of order (e.g. draw outside a frame) */
    SRG_ERR_NO_DEVICE = -8,      /* no HIP device, or the kernels are not loadable */
    SRG_ERR_HANDLE = -9          /* unknown or released handle */
};

/* Built-in pixel shaders: device implementations of Viewer/Shaders.h. */
enum srg_shader {
    SRG_SHADER_UNLIT_DIFFUSE = 0,    /* Shaders.h:71-104: texture sample, white without texture */
    SRG_SHADER_VISUALIZE_NORMALS = 1,/* Shaders.h:106-121 */
    SRG_SHADER_VISUALIZE_UVS = 2,    /* Shaders.h:123-130 */
    SRG_SHADER_SPONZA = 3            /* SponzaScene.cpp:13-104: sun + 16 point lights + ambient,
                                        times the texture sample; white without texture. Reads all
                                        eight varyings: attribute_stride must be 32 (Obj::Vertex) */
};

/* SponzaScene::Constants (Viewer/SponzaScene.h:15-47), the global the
 * SponzaShader reads (SponzaScene.cpp:11): each m_sunDir / m_ambCol lane
 * broadcast as one float, 16 PointLights (falloff is carried, not read). */
#define SRG_SPONZA_LIGHTS 16
typedef struct srg_sponza_light {
    float pos[3], colour[3], intensity, falloff;
} srg_sponza_light;
typedef struct srg_sponza_constants {
    float sun_dir[3], ambient[3];
    srg_sponza_light lights[SRG_SPONZA_LIGHTS];
} srg_sponza_constants;

/* Back-end tile-queue ordering. Every policy yields the same image. */
enum srg_tile_order {
    SRG_ORDER_NATURAL = 0,  /* row-major, view by view (Renderer.cpp:264-306) */
    SRG_ORDER_TRICOUNT = 1, /* longest-first by binned-triangle count */
    SRG_ORDER_LEARNED = 2   /* longest-first by the learned tile-cost model */
};

enum srg_config_flags {
    SRG_CFG_KEEP_PRIMID = 1u << 0, /* keep a per-pixel primitive-order plane (tests) */
    SRG_CFG_STATS = 1u << 1,       /* count rasterised fragments and binned entries */
    SRG_CFG_TIMING = 1u << 2,      /* time each kernel with HIP events */
    SRG_CFG_TILE_COSTS = 1u << 3,  /* record per-tile s_memtime cost in the back end */
    SRG_CFG_GRAPH = 1u << 4,       /* capture the frame's launches in a hipGraph and replay it while
                                      the draw list, views and arenas are unchanged */
    SRG_CFG_AUTOTUNE = 1u << 5     /* pick bin_size x wg_per_tile per frame signature (view sizes and
                                      draw triangle counts): each layout of srg_tuning is timed over
                                      autotune_frames frames, then the fastest is kept. Every layout
                                      gives the same image. */
};

typedef struct srg_config {
    uint32_t struct_size;   /* sizeof(srg_config) */
    int32_t device;         /* HIP device ordinal */
    uint32_t flags;         /* srg_config_flags */
    uint32_t tile_order;    /* srg_tile_order */
    uint32_t max_prims;     /* initial primitive-record capacity, 0 = sized per frame */
    uint32_t max_entries;   /* initial tile-entry capacity, 0 = sized per frame */
    uint32_t wg_per_cu;     /* persistent back-end workgroups per CU, 0 = default */
    uint32_t bin_size;      /* back-end work bin edge: 64, 32, 16 or 8 pixels, 0 = 32. The
                               reference's 64x64 bin semantics are kept for every size. */
    uint32_t sched_warmup_frames;    /* SRG_ORDER_LEARNED: frames recorded before the first
                                        training (0 = 8); entry-count order meanwhile */
    uint32_t sched_retrain_interval; /* SRG_ORDER_LEARNED: retrain every N frames (0 = never) */
    uint32_t wg_per_tile;   /* workgroups sharing one work bin, each owning a strip of
                               bin_size / wg_per_tile rows: 1, 2 or 4 (0 = 1). Valid layouts
                               (bin_size, wg_per_tile): (8,1) (16,1) (32,1|2|4) (64,1|2|4) */
    uint32_t autotune_frames;        /* SRG_CFG_AUTOTUNE: frames timed per layout (0 = 3) */
    uint32_t reserved[4];
} srg_config;

/* Mirrors DrawCall (Renderer.h:119-150). Buffers are registered handles plus
 * a byte offset; strides and counts have the reference's meaning:
 * index stride 1/2/4, position = 3 floats at pos_offset + i*position_stride,
 * attributes = attribute_stride/4 (<= 8) floats per vertex, uv_offset the
 * float index of (u, v) inside the attributes (SetAttributeBuffer, Renderer.cpp:119-126).
 * mvp is column-major: clip = mvp * (x, y, z, 1). */
typedef struct srg_draw {
    uint32_t shader;              /* srg_shader */
    srg_handle texture;           /* UnlitDiffuse uniforms; 0 = none (writes 0xFFFFFFFF) */
    srg_handle index_buffer;   uint64_t index_offset;     uint32_t index_stride;     uint32_t index_num;
    srg_handle position_buffer; uint64_t position_offset; uint32_t position_stride;  uint32_t position_num;
    srg_handle attribute_buffer; uint64_t attribute_offset; uint32_t attribute_stride; uint32_t attribute_num;
    uint32_t uv_offset;
    srg_fb* fb;                   /* target framebuffer */
    uint32_t plane;               /* its plane (srg_fb_write_plane at SetFrameBuffer time), or
                                     SRG_PLANE_WRITE: the write plane when the draw is submitted */
    float mvp[16];
} srg_draw;

typedef struct srg_frame_stats {
    uint64_t tris_in;         /* triangles submitted */
    uint64_t tris_clipped;    /* triangles routed through the clipper */
    uint64_t prims;           /* primitives surviving clip and cull (SRG_CFG_STATS) */
    uint64_t bin_entries;     /* triangle-64x64-tile entries as the reference bins them (SRG_CFG_STATS) */
    uint64_t work_entries;    /* entries in the back end's work bins */
    uint64_t frags_raster;    /* covered pixel samples before the depth test (SRG_CFG_STATS) */
    uint32_t tiles_queued;    /* tiles processed by the back end */
    uint32_t retries;         /* frame re-runs after arena growth */
    uint32_t graph_captures;  /* SRG_CFG_GRAPH: graphs (re)captured by this frame */
    float ms_frontend;        /* SRG_CFG_TIMING without SRG_CFG_GRAPH: setup + clip + scan + emit */
    float ms_backend;         /* SRG_CFG_TIMING without SRG_CFG_GRAPH: the raster + shade kernel alone */
    float ms_total;           /* SRG_CFG_TIMING: whole frame on the device */
    float sched_err;          /* SRG_ORDER_LEARNED: mean |log2 predicted - log2 measured| bin cost */
    uint32_t sched_trained;   /* SRG_ORDER_LEARNED: 1 once the model orders the queue */
    uint32_t sched_trainings; /* SRG_ORDER_LEARNED: training runs done by this frame */
    uint32_t bin_size;        /* back-end layout of this frame: work-bin edge ... */
    uint32_t wg_per_tile;     /* ... and workgroups per bin */
    uint32_t sched_fallback;  /* SRG_ORDER_LEARNED: 1 while the model's last measured error
                                 (sched_err) exceeds 1.0 and the entry-count order is used */
} srg_frame_stats;

/* SRG_CFG_AUTOTUNE state for the signature of the last frame. */
#define SRG_TUNE_LAYOUTS 8
typedef struct srg_tuning {
    uint32_t state;           /* 0 = autotuning off or no frame yet, 1 = sweeping, 2 = layout chosen */
    uint32_t bin_size;        /* layout chosen (state 2) */
    uint32_t wg_per_tile;
    uint32_t nlayouts;        /* SRG_TUNE_LAYOUTS */
    struct {
        uint32_t bin_size, wg_per_tile;
        float ms;             /* best whole-frame device time measured, 0 = not yet */
    } layouts[SRG_TUNE_LAYOUTS];
} srg_tuning;

/* ---- context: RenderContext ctor / Shutdown (Renderer.cpp:138-159) ---- */
int srg_create(const srg_config* cfg, srg_ctx** out);
void srg_destroy(srg_ctx* ctx);
const char* srg_status_string(int status);

/* ---- resources ---- */
/* Copies `bytes` of host memory to a new device buffer. Registered data is
 * treated as immutable; re-register (or srg_update_buffer) after changes. */
int srg_register_buffer(srg_ctx* ctx, const void* host, size_t bytes, srg_handle* out);
/* Wraps existing device memory (not copied, not owned). */
int srg_register_device_buffer(srg_ctx* ctx, void* device_ptr, size_t bytes, srg_handle* out);
int srg_update_buffer(srg_ctx* ctx, srg_handle h, size_t offset, const void* host, size_t bytes);
int srg_release_buffer(srg_ctx* ctx, srg_handle h);
/* Tex::TextureData (Texture.h:21-41): tiled + Morton-swizzled RGBA8 mip chain
 * as produced by srg_texture_build_rgba8 / TextureData::CreateFromRGBA8. */
int srg_register_texture(srg_ctx* ctx, const void* tiled_texels, uint64_t bytes, const uint32_t mip_offsets[14],
                         uint32_t width_log2, uint32_t height_log2, uint32_t num_mips, srg_handle* out);
int srg_release_texture(srg_ctx* ctx, srg_handle h);

/* ---- framebuffers: FrameBuffer (Renderer.h:75-108, Renderer.cpp:23-57) ---- */
enum srg_fb_flags { SRG_FB_COLOUR = 1u, SRG_FB_DEPTH = 2u };
int srg_fb_create(srg_ctx* ctx, uint32_t width, uint32_t height, uint32_t flags, srg_fb** out);
/* Destroying a framebuffer that draws of the open frame target makes
 * srg_end_frame fail with SRG_ERR_HANDLE. Every entry point taking an srg_fb
 * returns SRG_ERR_HANDLE for one that is not (or no longer) ctx's; a destroyed
 * framebuffer is not dereferenced. */
void srg_fb_destroy(srg_ctx* ctx, srg_fb* fb);
uint32_t srg_fb_write_plane(const srg_fb* fb);
void srg_fb_swap_planes(srg_fb* fb);
/* Copies a plane's tiled contents (tiles_x*tiles_y*4096 elements each) to host
 * memory; any pointer may be NULL. primid needs SRG_CFG_KEEP_PRIMID. */
int srg_fb_read_tiles(srg_ctx* ctx, srg_fb* fb, uint32_t plane, uint32_t* colour, float* depth, uint32_t* primid);
/* Uploads tiled contents into a plane (e.g. a non-cleared depth buffer). */
int srg_fb_write_tiles(srg_ctx* ctx, srg_fb* fb, uint32_t plane, const uint32_t* colour, const float* depth);

/* ---- frame: BeginFrame / ClearFrameBuffer / DrawIndexed / EndFrame / Blit ---- */
int srg_begin_frame(srg_ctx* ctx);
int srg_clear(srg_ctx* ctx, srg_fb* fb, uint32_t colour, int clear_colour, int clear_depth);
/* srg_clear for n framebuffers in one call (a batch of views). Every handle
 * is checked first: with one unknown handle nothing is cleared and the call
 * returns SRG_ERR_HANDLE. No counterpart in the reference: n ClearFrameBuffer
 * calls (Renderer.cpp:168-194). */
int srg_clear_n(srg_ctx* ctx, srg_fb* const* fbs, uint32_t n, uint32_t colour, int clear_colour, int clear_depth);
int srg_draw_indexed(srg_ctx* ctx, const srg_draw* draw);
/* n draws in one call, each as srg_draw_indexed; returns the first failing
 * status (the draws before it are recorded). No counterpart in the reference,
 * whose DrawIndexed takes one DrawCall (Renderer.cpp:161-166): a binding
 * that pays per foreign call submits a prepared draw list with it, and
 * SRG_PLANE_WRITE keeps such a list valid across SwapPlanes. */
#define SRG_PLANE_WRITE 0xFFFFFFFFu
int srg_draw_indexed_n(srg_ctx* ctx, const srg_draw* draws, uint32_t n);
/* Synchronous like the reference: on return every tile is rastered and shaded.
 * Every recorded draw is checked again first (buffers, texture and framebuffer
 * still registered, sizes still in range): a handle released since its
 * srg_draw_indexed fails the frame with SRG_ERR_HANDLE, drawing nothing. */
int srg_end_frame(srg_ctx* ctx);
/* srg_end_frame without allocating: no hipMalloc, hipHostMalloc, hipFree or
 * graph instantiation, for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid allocation). No counterpart
 * in the reference, whose EndFrame (Renderer.cpp:209-317) bump-allocates from
 * per-thread arenas. The frame's working memory (tables, bin arrays, setup
 * records, entry lists) is carved from a static device array of
 * SRG_NOALLOC_SCRATCH_BYTES held by each GPU the library runs on until the
 * process ends, and the tables are staged through SRG_NOALLOC_TABLE_BYTES of
 * pinned host memory allocated by srg_create. Same image as srg_end_frame.
 * Differences: the queue is in natural or entry-count order (a context
 * configured for the learned order uses entry counts; its model is neither
 * used nor trained), the autotuner's choice for the frame's signature is used
 * only once it is locked (no sweep; otherwise the configured layout), and no
 * graph is captured. A frame whose arenas do not fit returns SRG_ERR_OVERFLOW
 * and draws nothing (its planes keep their pending clears); about 115-135
 * bytes per triangle plus 16 per triangle-bin entry, i.e. roughly 350k
 * triangles per frame. Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (two threads, two contexts) race and must be serialised by the
 * caller. Each call is synchronous and initialises every part of the array it
 * reads. */
#define SRG_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define SRG_NOALLOC_TABLE_BYTES (256u << 10)
int srg_end_frame_noalloc(srg_ctx* ctx);
/* Copies the write plane's colour to linear RGBA8 rows (width*height*4 bytes),
 * swaps the planes, then calls on_finish(user) (Renderer.cpp:350-372). */
int srg_blit(srg_ctx* ctx, srg_fb* fb, void* linear_rgba8, void (*on_finish)(void*), void* user);
/* Blit (Renderer.cpp:350-372) with the reference's asynchrony: returns at
 * once; on a second stream the write plane is de-tiled into pinned memory
 * (overlapping the next frame), copied to linear_rgba8 (width*height*4
 * bytes) and on_finish(user) is called from a runtime thread, which must not
 * call into HIP. The planes are swapped at once. Like the reference, a blit
 * first waits for the previous blit, of the other plane, to finish; later
 * frames wait on the device until the blitted plane has been read.
 * SRG_BLIT_BGRA exchanges red and blue on the way (the present swizzle of
 * Window_Win32::Flip, Viewer/Platform/Window_Win32.cpp:85-112). */
#define SRG_BLIT_BGRA 1u
int srg_blit_async(srg_ctx* ctx, srg_fb* fb, void* linear_rgba8, uint32_t flags, void (*on_finish)(void*), void* user);
/* Waits until every asynchronous blit of fb, callback included, is done. */
int srg_blit_wait(srg_ctx* ctx, srg_fb* fb);
int srg_get_frame_stats(srg_ctx* ctx, srg_frame_stats* out);
/* Per work-bin data of the last frame, bins numbered view by view in the
 * order views were first drawn to, row-major inside a view: entry counts and,
 * with SRG_CFG_TILE_COSTS, the back end's measured cost (s_memtime ticks).
 * Either array may be NULL; *nbins receives the bin count. No counterpart in
 * the reference: its per-tile scope timing lives in microprofile
 * (Rasterizer.cpp:527). */
int srg_get_bin_stats(srg_ctx* ctx, uint32_t* counts, uint64_t* costs, uint32_t capacity, uint32_t* nbins);
/* The learned tile-cost model's parameters (4-32-32-1 fp32 MLP, SRG_SCHED_PARAMS
 * floats): saving and restoring them is the scheduler's checkpoint. Setting
 * them marks the model trained. No counterpart in the reference. */
#define SRG_SCHED_PARAMS 1249u
int srg_sched_get_model(srg_ctx* ctx, float* params, uint32_t count);
int srg_sched_set_model(srg_ctx* ctx, const float* params, uint32_t count);
/* Test access to the model's device kernels (no counterpart in the
 * reference). srg_sched_eval: the predicted log2 cost of n rows of 4
 * features (the inference the learned order runs per bin), with the current
 * parameters. srg_sched_train: replaces the recorded samples with n rows of
 * (4 features, log2 cost), clears the Adam state and runs `steps` steps of
 * the device's mini-batch Adam (batch 256, each thread drawing its sample
 * with the generator seeded by `seed`) at learning rate lr; the parameters
 * are then read with srg_sched_get_model. n <= 262144. */
int srg_sched_eval(srg_ctx* ctx, const float* features, uint32_t n, float* log2_cost);
int srg_sched_train(srg_ctx* ctx, const float* samples, uint32_t n, uint32_t steps, float lr, uint32_t seed);
/* The autotuner's table for the last frame's signature (SRG_CFG_AUTOTUNE).
 * No counterpart in the reference, whose bin size is a compile-time constant
 * (Config.h:18-21) and whose tiles each take one task (Renderer.cpp:264-306). */
int srg_get_tuning(srg_ctx* ctx, srg_tuning* out);
/* Sets the SponzaShader constants for frames ended after the call (the
 * reference writes the g_constants global in SponzaScene::Init / Update,
 * SponzaScene.cpp:127-186, before issuing its draws). Zero until set. */
int srg_set_sponza_constants(srg_ctx* ctx, const srg_sponza_constants* k);

/* ---- scene ingest: Obj::Model::Load (Viewer/Obj.cpp:374-567) ----
 * Host code (no device needed). Reads a Wavefront OBJ as the reference's
 * viewer does: v / vt / vn / f (triangles and quads, 1-based or negative
 * indices), one vertex per distinct (position, uv, normal) triple of a mesh,
 * "g" starts a new mesh, usemtl / mtllib / newmtl / map_Kd name materials.
 * Meshes come out as Obj::Mesh (Obj.h:24-42): Obj::Vertex {pos3, norm3, uv2}
 * and u16 indices, u32 above 65,535 vertices. The flags are Obj::LoadFlags
 * (Obj.h:54-60). The reference's "<path>.bin" cache (kt serializer) is not
 * written or read. Parse errors and unreadable files return SRG_ERR_INVALID. */
enum srg_obj_flags {
    SRG_OBJ_FLIP_WINDING = 1u, /* swap the 2nd and 3rd index of every triangle */
    SRG_OBJ_FLIP_UVS = 4u      /* v -> 1 - v */
};
typedef struct srg_obj_model srg_obj_model;
typedef struct srg_obj_mesh {
    uint32_t index_stride;     /* 2 (IndexType::u16) or 4 (IndexType::u32) */
    uint32_t num_indices;
    const void* indices;       /* owned by the model */
    uint32_t num_vertices;
    const float* vertices;     /* 8 floats per vertex, owned by the model */
    uint32_t material;         /* index into the model's materials (0 without any) */
} srg_obj_mesh;
int srg_obj_load(const char* path, uint32_t flags, srg_obj_model** out);
/* The same from text in memory; mtllib names are read relative to dir
 * (NULL: the working directory). */
int srg_obj_parse(const char* text, size_t bytes, const char* dir, uint32_t flags, srg_obj_model** out);
uint32_t srg_obj_mesh_count(const srg_obj_model* model);
int srg_obj_get_mesh(const srg_obj_model* model, uint32_t index, srg_obj_mesh* out);
uint32_t srg_obj_material_count(const srg_obj_model* model);
const char* srg_obj_material_name(const srg_obj_model* model, uint32_t index);
/* The map_Kd path of a material, joined to the OBJ's directory ("" if none). */
const char* srg_obj_material_diffuse(const srg_obj_model* model, uint32_t index);
void srg_obj_free(srg_obj_model* model);

/* ---- host helper: TextureData::CreateFromRGBA8 (Texture.cpp:119-199) ----
 * Tiles a power-of-two RGBA8 image (>= 32 in both dimensions) into 32x32
 * Morton-swizzled tiles and, if calc_mips, appends the mip chain. Call with
 * out_tiled == NULL to get the byte size in *out_bytes. */
int srg_texture_build_rgba8(const uint8_t* linear_rgba8, uint32_t width, uint32_t height, int calc_mips,
                            uint8_t* out_tiled, uint64_t* out_bytes, uint32_t mip_offsets[14],
                            uint32_t* width_log2, uint32_t* height_log2, uint32_t* num_mips);

/* ---- TextureData::CreateFromRGBA8 on the device (Texture.cpp:119-199) ----
 * Uploads a linear power-of-two RGBA8 image (>= 32 in both dimensions),
 * builds the tiled level 0 and, if calc_mips, every mip on the GPU, and
 * registers the result as a texture. The bytes equal what
 * srg_texture_build_rgba8 makes (each mip is stbir_resize_uint8 of the base
 * level, Texture.cpp:196). */
int srg_texture_create_rgba8(srg_ctx* ctx, const uint8_t* linear_rgba8, uint32_t width, uint32_t height,
                             int calc_mips, srg_handle* out);
/* A registered texture's tiled bytes and layout (TextureData's m_texels,
 * m_mipOffsets, m_widthLog2, m_heightLog2, m_numMips; Texture.h:21-41).
 * With out == NULL only the layout is returned. */
int srg_texture_read(srg_ctx* ctx, srg_handle h, uint8_t* out, uint64_t capacity, uint64_t* bytes,
                     uint32_t mip_offsets[14], uint32_t* width_log2, uint32_t* height_log2, uint32_t* num_mips);

#ifdef __cplusplus
}
#endif
#endif
