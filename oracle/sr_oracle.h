/*
 * sr_oracle.h -- CPU oracle for SoftRast's sort-middle rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of the reference
 * renderer (SoftRast/Binning.cpp, Rasterizer.cpp, Texture.cpp, SIMDUtil.h,
 * Renderer.cpp, Viewer/Shaders.h) used to check the HIP implementation and,
 * timed on host cores, as the CPU baseline. Nothing in the product library
 * (softrast_amd/) links, loads or calls it.
 *
 * It keeps the reference's own structure: per-thread bins of 32-triangle
 * chunks, a per-tile fragment list in submission order, interpolation and
 * shading of every depth-passing fragment in that order, so that the last
 * write wins. The GPU library resolves visibility differently (a per-pixel
 * lexicographic max of (depth, primitive order)); agreement between the two
 * is what the parity tests establish.
 *
 * Float semantics: every float operation is rounded to fp32 (built with
 * -ffp-contract=off); fmaf() appears exactly where the reference uses an
 * explicit FMA intrinsic. Parity status of each function is listed in the
 * header comment of sr_oracle.c.
 */
#ifndef SR_ORACLE_H
#define SR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRO_BIN_SIZE 64
#define SRO_MAX_VARYINGS 8
#define SRO_MAX_MIPS 14

enum { SRO_SHADER_UNLIT_DIFFUSE = 0, SRO_SHADER_VISUALIZE_NORMALS = 1, SRO_SHADER_VISUALIZE_UVS = 2,
       SRO_SHADER_SPONZA = 3 };

/* SponzaScene::Constants (Viewer/SponzaScene.h:36-47), the global the
 * SponzaShader reads: each m_sunDir / m_ambCol component is one float
 * broadcast over the lanes; 16 PointLights (m_falloff is not read). */
#define SRO_SPONZA_LIGHTS 16
typedef struct sro_sponza_light {
    float pos[3], colour[3], intensity, falloff;
} sro_sponza_light;
typedef struct sro_sponza_constants {
    float sun_dir[3], ambient[3];
    sro_sponza_light lights[SRO_SPONZA_LIGHTS];
} sro_sponza_constants;

/* Tex::TextureData (SoftRast/Texture.h:21-41): 32x32-tiled, Morton-swizzled
 * RGBA8 mip chain. texels == NULL or size == 0 means "empty texture". */
typedef struct sro_texture {
    const uint8_t* texels;
    uint64_t size;
    uint32_t mip_offsets[SRO_MAX_MIPS];
    uint32_t wlog2, hlog2, nmips;
} sro_texture;

/* DrawCall (SoftRast/Renderer.h:119-150), with the pixel shader named by id
 * and its uniforms (the texture) given directly. */
typedef struct sro_draw {
    uint32_t shader;
    const sro_texture* tex;
    const void* idx;  uint32_t idx_num;  uint32_t idx_stride;  /* stride 1, 2 or 4 bytes */
    const void* pos;  uint32_t pos_num;  uint32_t pos_stride;
    const void* attr; uint32_t attr_num; uint32_t attr_stride; /* stride <= 32 bytes */
    uint32_t uv_offset;
    uint32_t plane;    /* index of the target plane in the plane array */
    float mvp[16];     /* column-major, clip = M * (x, y, z, 1) */
} sro_draw;

/* FrameBufferPlane (SoftRast/Renderer.h:58-73): 64x64 tiles, RGBA8 colour and
 * fp32 depth, row-major inside a tile. All arrays are caller-owned and hold
 * tiles_x * tiles_y * 4096 elements. primid (may be NULL) receives, for the
 * pixels written during sro_render, the primitive order key of the final
 * fragment, ((first_tri_of_draw + tri) << 3) | fan, where fan is the clip-fan
 * index (0..6) of a clipped triangle and 7 for an unclipped one; untouched
 * pixels are set to 0xFFFFFFFF. */
typedef struct sro_plane {
    uint32_t width, height, tiles_x, tiles_y;
    uint32_t* colour;
    float* depth;
    uint32_t* primid;
} sro_plane;

typedef struct sro_stats {
    uint64_t tris_in;        /* triangles submitted */
    uint64_t tris_clipped;   /* triangles that went through the clipper */
    uint64_t prims_binned;   /* (sub-)triangles that survived cull and were binned */
    uint64_t bin_entries;    /* triangle-tile entries */
    uint64_t frags_raster;   /* pixel samples passing coverage (before depth) */
    uint64_t frags_passed;   /* fragments passing the depth test (shaded) */
    uint64_t blocks_tested;  /* 8x8 blocks that pass the block reject test (Rasterizer.cpp:230-261) and reach
                                ComputeBlockMask8x8 or its depth-only variant (:263-272); no pixel depends on
                                which blocks these are, so this count is what observes the reject test */
} sro_stats;

typedef struct sro_ctx sro_ctx;

/* RenderContext (SoftRast/Renderer.cpp:138-150): nthreads = total threads
 * including the calling thread (1 = SR_DEBUG_SINGLE_THREADED). */
sro_ctx* sro_create(uint32_t nthreads);
void sro_destroy(sro_ctx* ctx);

void sro_plane_dims(uint32_t width, uint32_t height, uint32_t* tiles_x, uint32_t* tiles_y);

/* RenderContext::ClearFrameBuffer (SoftRast/Renderer.cpp:168-194). */
void sro_clear(sro_plane* plane, uint32_t colour, int clear_colour, int clear_depth);

/* BeginFrame + DrawIndexed x n + EndFrame (SoftRast/Renderer.cpp:161-317).
 * tile_entries (may be NULL) receives, per plane in order and per tile
 * row-major, the number of triangle entries binned to that tile.
 * Returns 0, or a negative value on invalid input. */
int sro_render(sro_ctx* ctx, const sro_draw* draws, uint32_t ndraws, sro_plane* planes,
               uint32_t nplanes, sro_stats* stats, uint32_t* tile_entries);

/* The SponzaShader's global constants (g_constants, Viewer/SponzaScene.cpp:11)
 * for later renders of this context. */
void sro_set_sponza_constants(sro_ctx* ctx, const sro_sponza_constants* k);

/* BlitJobFn (SoftRast/Renderer.cpp:319-347): tiled colour -> linear rows. */
void sro_blit(const sro_plane* plane, uint32_t* linear);

/* ---- single-function entry points, for known-answer tests ---- */
float sro_rcpps(float x);
float sro_rsqrtps(float x);
uint32_t sro_pack_rgba(float r, float g, float b, float a);
uint32_t sro_morton5(uint32_t x, uint32_t y);
int32_t sro_calc_mip(const sro_texture* tex, float dudx, float dudy, float dvdx, float dvdy);
uint32_t sro_sample_wrap(const sro_texture* tex, float u, float v, float dudx, float dudy,
                         float dvdx, float dvdy, float rgba_out[4]);
void sro_transform(const float mvp[16], const float p[3], float out[4]);
uint32_t sro_clip_codes(const float v[4]);
/* Clips one triangle (3 x vec4 + 3 x 8 attributes); returns the polygon vertex
 * count (0 when fully clipped) and writes up to 9 vertices. */
uint32_t sro_clip_triangle(const float verts[12], const float attribs[24], uint32_t mask,
                           float out_verts[36], float out_attribs[72]);
/* Triangle setup of BinTransformedAndClippedTri for a W x H target.
 * Returns 0 if culled, else 1 and fills:
 *   ints[0..2]=c, [3..5]=dx, [6..8]=dy, [9..12]=xmin,xmax,ymin,ymax, [13..18]=fixed-point verts
 *   floats[0..1]=r0, [2..4]=z plane dx,dy,f0, [5..7]=1/w plane dx,dy,f0,
 *   [8 + 3k .. 10 + 3k]=attribute k plane dx,dy,f0 for k < nattr. */
int sro_setup_triangle(const float v[12], const float attribs[24], uint32_t nattr, uint32_t width,
                       uint32_t height, int32_t ints[19], float floats[32]);

/* One mip level as TextureData::CreateFromRGBA8 makes it (Texture.cpp:188-198):
 * stbir_resize_uint8(base, w, h, 0, out, ow, oh, 0, 4) of the vendored
 * stb_image_resize v0.96, i.e. Mitchell filter in both axes, clamped edges,
 * linear colour space, no alpha handling. Downsampling only (ow <= w,
 * oh <= h); returns 0, or -1 for other sizes. */
int sro_mip_resize_rgba8(const uint8_t* in, uint32_t w, uint32_t h, uint8_t* out, uint32_t ow, uint32_t oh);

#ifdef __cplusplus
}
#endif
#endif
