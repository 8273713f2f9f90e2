// sr/Renderer.h -- source-level replacement for SoftRast's Renderer.h
// (SoftRast/Renderer.h:20-177), Texture.h (:21-56) and the shader contract
// of Viewer/Shaders.h, implemented over the C ABI in srgpu.h.
//
// Field and method names follow the reference so that code like
// Viewer/Scene.cpp:32-65 compiles unchanged apart from the include and the
// matrix type. Differences a caller can observe:
//   * RenderContext methods that can fail return an int status (0 = ok)
//     instead of asserting. A FrameBuffer lives in one context's device
//     memory: FrameBuffer(w, h) binds to the most recently created live
//     RenderContext, or, when none exists yet, to the first context that
//     clears, draws to or blits it; FrameBuffer(ctx, w, h) names the context.
//   * A PixelShaderFn pointer selects a device shader: the three shaders of
//     Viewer/Shaders.h are provided here; any other pointer makes DrawIndexed
//     return SRG_ERR_UNKNOWN_SHADER (there is no CPU fallback).
//   * Buffers referenced by a DrawCall are uploaded on first use and cached
//     by (pointer, size); call RenderContext::InvalidateBuffer after changing
//     their contents.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "../srgpu.h"

namespace sr
{

namespace Config
{
constexpr uint32_t c_binWidth = 64;   // Config.h:18-21
constexpr uint32_t c_binHeight = 64;
constexpr uint32_t c_maxVaryings = 8; // Config.h:27
constexpr uint32_t c_maxTexDimLog2 = 14;
}

// Column-major 4x4 matrix, clip = M * (x, y, z, 1) (kt::Mat4 layout).
struct Mat4
{
    float m[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    static Mat4 Identity() { return Mat4{}; }
};

// Interpolants (Renderer.h:39-56): the pixel-shader input. Device shaders
// receive the same quantities per pixel; the host struct exists so that
// shader function signatures match the reference's PixelShaderFn.
struct Interpolants
{
    float* m_derivs[4];
    float* m_varyings[Config::c_maxVaryings];
};

using PixelShaderFn = void(void const* _uniforms, Interpolants const& _interpolants, uint32_t o_texels[8],
                           uint32_t _execMask);

namespace Tex
{

// Tex::TextureData (Texture.h:21-41).
struct TextureData
{
    std::vector<uint8_t> m_texels;
    uint32_t m_mipOffsets[Config::c_maxTexDimLog2] = {};
    uint32_t m_widthLog2 = 0;
    uint32_t m_heightLog2 = 0;
    uint32_t m_numMips = 0;
    uint32_t m_bytesPerPixel = 0;

    // Texture.cpp:119-199. Returns 0 or an srg_status.
    int CreateFromRGBA8(uint8_t const* _texels, uint32_t _width, uint32_t _height, bool _calcMips = false)
    {
        uint64_t bytes = 0;
        int st = srg_texture_build_rgba8(_texels, _width, _height, _calcMips ? 1 : 0, nullptr, &bytes, m_mipOffsets,
                                         &m_widthLog2, &m_heightLog2, &m_numMips);
        if (st) return st;
        m_texels.assign(bytes, 0);
        st = srg_texture_build_rgba8(_texels, _width, _height, _calcMips ? 1 : 0, m_texels.data(), &bytes,
                                     m_mipOffsets, &m_widthLog2, &m_heightLog2, &m_numMips);
        m_bytesPerPixel = 4;
        return st;
    }

    void Clear() { m_texels.clear(); }
};

}  // namespace Tex

namespace shader
{
// Viewer/Shaders.h:71-130. The host bodies are never executed: the address
// selects the device implementation.
inline void UnlitDiffuseShader(void const*, Interpolants const&, uint32_t o_texels[8], uint32_t) { memset(o_texels, 0, 32); }
inline void VisualizeNormalsShader(void const*, Interpolants const&, uint32_t o_texels[8], uint32_t) { memset(o_texels, 0, 32); }
inline void VisualizeUVsShader(void const*, Interpolants const&, uint32_t o_texels[8], uint32_t) { memset(o_texels, 0, 32); }
// SponzaShader (Viewer/SponzaScene.cpp:13-104; a file-static function there):
// its constants come from RenderContext::SetSponzaConstants instead of the
// g_constants global.
inline void SponzaShader(void const*, Interpolants const&, uint32_t o_texels[8], uint32_t) { memset(o_texels, 0, 32); }
}  // namespace shader

class RenderContext;

// FrameBufferPlane / FrameBuffer (Renderer.h:58-108): double-buffered
// planes of 64x64 tiles held in device memory.
struct FrameBuffer;

struct FrameBufferPlane
{
    srg_fb* m_fb = nullptr;
    uint32_t m_plane = 0;
    uint32_t m_width = 0, m_height = 0, m_tilesX = 0, m_tilesY = 0;
    FrameBuffer* m_owner = nullptr;  // binds the framebuffer when a draw reaches it unbound
};

struct FrameBuffer
{
    FrameBuffer(uint32_t _width, uint32_t _height, bool _colour = true, bool _depth = true);
    FrameBuffer(RenderContext& _ctx, uint32_t _width, uint32_t _height, bool _colour = true, bool _depth = true);
    ~FrameBuffer();
    FrameBuffer(FrameBuffer const&) = delete;
    FrameBuffer& operator=(FrameBuffer const&) = delete;

    FrameBufferPlane* WritePlane()
    {
        m_bufferedPlanes[m_writePlane].m_plane = m_writePlane;
        return &m_bufferedPlanes[m_writePlane];
    }
    FrameBufferPlane const* ReadPlane() const { return &m_bufferedPlanes[m_writePlane ^ 1]; }
    void SwapPlanes()
    {
        srg_fb_swap_planes(m_fb);
        m_writePlane ^= 1;
    }

    FrameBufferPlane m_bufferedPlanes[2];
    uint32_t m_writePlane = 0;
    srg_fb* m_fb = nullptr;
    RenderContext* m_ctx = nullptr;
    int m_status = 0;
    uint32_t m_width = 0, m_height = 0, m_flags = 0;

private:
    void Init(uint32_t _width, uint32_t _height, bool _colour, bool _depth);
    int BindTo(RenderContext& _ctx);
    friend class RenderContext;
};

// TaskSystem's per-thread scratch allocator (TaskSystem.h:13, kt::LinearAllocator):
// a host bump allocator per thread for host code written against
// RenderContext::ThreadAllocator. The device path does not use it: its
// scratch is the library's device arenas.
class ThreadScratchAllocator
{
public:
    void* Alloc(size_t _bytes, size_t _align = 16)
    {
        size_t at = (m_used + _align - 1) / _align * _align;
        if (m_blocks.empty() || at + _bytes > m_cap) {
            const size_t cap = _bytes + _align > (size_t(1) << 20) ? _bytes + _align : (size_t(1) << 20);
            m_blocks.emplace_back(new uint8_t[cap]);
            m_cap = cap;
            m_used = 0;
            const uintptr_t base = (uintptr_t)m_blocks.back().get();
            at = (size_t)(((base + _align - 1) / _align * _align) - base);
        }
        m_used = at + _bytes;
        return m_blocks.back().get() + at;
    }
    // Frees everything allocated since the last Reset (the frame's scratch).
    void Reset()
    {
        if (m_blocks.size() > 1) m_blocks.erase(m_blocks.begin(), m_blocks.end() - 1);
        m_used = 0;
    }

private:
    std::vector<std::unique_ptr<uint8_t[]>> m_blocks;
    size_t m_cap = 0, m_used = 0;
};

struct GenericDrawBuffer
{
    void const* m_ptr = nullptr;
    uint32_t m_num = 0;
    uint32_t m_stride = 0;
};

// DrawCall (Renderer.h:119-150).
struct DrawCall
{
    static const uint32_t UV_OFFSET_INVALID = 0xFFFFFFFF;

    DrawCall& SetPixelShader(PixelShaderFn* _fn, void const* _uniforms)
    {
        m_pixelShader = _fn;
        m_pixelUniforms = _uniforms;
        return *this;
    }
    DrawCall& SetIndexBuffer(void const* _buffer, uint32_t const _stride, uint32_t const _num)
    {
        m_indexBuffer = { _buffer, _num, _stride };
        return *this;
    }
    DrawCall& SetPositionBuffer(void const* _buffer, uint32_t const _stride, uint32_t const _num)
    {
        m_positionBuffer = { _buffer, _num, _stride };
        return *this;
    }
    DrawCall& SetAttributeBuffer(void const* _buffer, uint32_t const _stride, uint32_t const _num,
                                 uint32_t const _uvOffset = 0)
    {
        m_uvOffset = _uvOffset;
        m_attributeBuffer = { _buffer, _num, _stride };
        return *this;
    }
    DrawCall& SetFrameBuffer(FrameBuffer* _buffer)
    {
        m_frameBuffer = _buffer->WritePlane();
        return *this;
    }
    DrawCall& SetMVP(Mat4 const& _mvp)
    {
        m_mvp = _mvp;
        return *this;
    }

    PixelShaderFn* m_pixelShader = nullptr;
    void const* m_pixelUniforms = nullptr;
    GenericDrawBuffer m_indexBuffer;
    GenericDrawBuffer m_positionBuffer;
    GenericDrawBuffer m_attributeBuffer;
    uint32_t m_uvOffset = 0;
    FrameBufferPlane const* m_frameBuffer = nullptr;
    Mat4 m_mvp = Mat4::Identity();
    uint32_t m_drawCallIdx = 0;
    uint32_t m_colourWrite : 1;
    uint32_t m_depthWrite : 1;
    uint32_t m_depthRead : 1;

    DrawCall() : m_colourWrite(1), m_depthWrite(1), m_depthRead(1) {}
};

// RenderContext (Renderer.h:153-177).
class RenderContext
{
public:
    explicit RenderContext(srg_config const* _cfg = nullptr)
    {
        srg_config cfg{};
        cfg.struct_size = sizeof(cfg);
        if (_cfg) cfg = *_cfg;
        m_status = srg_create(&cfg, &m_ctx);
        if (!m_status) m_prevCurrent = std::exchange(CurrentSlot(), this);
    }
    ~RenderContext()
    {
        Shutdown();
        if (CurrentSlot() == this) CurrentSlot() = m_prevCurrent;
    }
    RenderContext(RenderContext const&) = delete;
    RenderContext& operator=(RenderContext const&) = delete;

    int Status() const { return m_status; }
    srg_ctx* Handle() const { return m_ctx; }

    // The most recently created RenderContext still alive (FrameBuffer(w, h)
    // binds to it), or nullptr.
    static RenderContext* Current() { return CurrentSlot(); }

    // RenderContext::ThreadAllocator (Renderer.h:165): the calling thread's
    // host scratch allocator.
    ThreadScratchAllocator& ThreadAllocator()
    {
        thread_local ThreadScratchAllocator alloc;
        return alloc;
    }

    void Shutdown()
    {
        // srg_destroy frees the framebuffers' planes: detach them first
        for (FrameBuffer* fb : m_fbs) {
            fb->m_fb = nullptr;
            fb->m_ctx = nullptr;
            fb->m_bufferedPlanes[0].m_fb = fb->m_bufferedPlanes[1].m_fb = nullptr;
        }
        m_fbs.clear();
        if (m_ctx) srg_destroy(m_ctx);
        m_ctx = nullptr;
        m_buffers.clear();
        m_textures.clear();
    }

    int BeginFrame()
    {
        m_numDraws = 0;
        return srg_begin_frame(m_ctx);
    }

    // SponzaScene::Constants for frames ended after the call (the reference's
    // g_constants global, written by SponzaScene::Init / Update).
    int SetSponzaConstants(srg_sponza_constants const& _k) { return srg_set_sponza_constants(m_ctx, &_k); }

    int ClearFrameBuffer(FrameBuffer& _buffer, uint32_t _color = 0x00000000, bool _clearColour = true,
                         bool _clearDepth = true)
    {
        if (int st = Bind(_buffer)) return st;
        return srg_clear(m_ctx, _buffer.m_fb, _color, _clearColour ? 1 : 0, _clearDepth ? 1 : 0);
    }

    int DrawIndexed(DrawCall const& _call)
    {
        if (!_call.m_indexBuffer.m_ptr || !_call.m_frameBuffer) return SRG_ERR_INVALID;
        FrameBuffer* owner = _call.m_frameBuffer->m_owner;
        if (owner && !owner->m_fb) {
            if (int st = Bind(*owner)) return st;
        }
        srg_draw d{};
        if (_call.m_pixelShader == &shader::UnlitDiffuseShader) d.shader = SRG_SHADER_UNLIT_DIFFUSE;
        else if (_call.m_pixelShader == &shader::VisualizeNormalsShader) d.shader = SRG_SHADER_VISUALIZE_NORMALS;
        else if (_call.m_pixelShader == &shader::VisualizeUVsShader) d.shader = SRG_SHADER_VISUALIZE_UVS;
        else if (_call.m_pixelShader == &shader::SponzaShader) d.shader = SRG_SHADER_SPONZA;
        else return SRG_ERR_UNKNOWN_SHADER;
        int st = 0;
        if ((d.shader == SRG_SHADER_UNLIT_DIFFUSE || d.shader == SRG_SHADER_SPONZA) && _call.m_pixelUniforms) {
            Tex::TextureData const* tex = static_cast<Tex::TextureData const*>(_call.m_pixelUniforms);
            if (!tex->m_texels.empty() && (st = Texture(tex, &d.texture))) return st;
        }
        GenericDrawBuffer const& ib = _call.m_indexBuffer;
        GenericDrawBuffer const& pb = _call.m_positionBuffer;
        GenericDrawBuffer const& ab = _call.m_attributeBuffer;
        if ((st = Buffer(ib.m_ptr, (size_t)ib.m_num * ib.m_stride, &d.index_buffer))) return st;
        if ((st = Buffer(pb.m_ptr, (size_t)pb.m_num * pb.m_stride, &d.position_buffer))) return st;
        if ((st = Buffer(ab.m_ptr, (size_t)ab.m_num * ab.m_stride, &d.attribute_buffer))) return st;
        d.index_stride = ib.m_stride;
        d.index_num = ib.m_num;
        d.position_stride = pb.m_stride;
        d.position_num = pb.m_num;
        d.attribute_stride = ab.m_stride;
        d.attribute_num = ab.m_num;
        d.uv_offset = _call.m_uvOffset;
        d.fb = owner ? owner->m_fb : _call.m_frameBuffer->m_fb;
        d.plane = _call.m_frameBuffer->m_plane;
        memcpy(d.mvp, _call.m_mvp.m, sizeof(d.mvp));
        ++m_numDraws;
        return srg_draw_indexed(m_ctx, &d);
    }

    int EndFrame() { return srg_end_frame(m_ctx); }

    int Blit(FrameBuffer& _fb, uint8_t* _linearPixels, void (*_onFinishBlit)(void*) = nullptr,
             void* _onFinishUser = nullptr)
    {
        if (int bst = Bind(_fb)) return bst;
        int st = srg_blit(m_ctx, _fb.m_fb, _linearPixels, _onFinishBlit, _onFinishUser);
        if (!st) _fb.m_writePlane ^= 1;
        return st;
    }

    void InvalidateBuffer(void const* _ptr)
    {
        for (auto it = m_buffers.begin(); it != m_buffers.end();) {
            if (it->first.first == _ptr) {
                srg_release_buffer(m_ctx, it->second);
                it = m_buffers.erase(it);
            } else {
                ++it;
            }
        }
    }

private:
    static RenderContext*& CurrentSlot()
    {
        static RenderContext* current = nullptr;
        return current;
    }

    // A framebuffer made without a context is created in this one on first
    // use; one bound to another context is an error.
    int Bind(FrameBuffer& _fb)
    {
        if (_fb.m_fb) return _fb.m_ctx == this ? SRG_OK : SRG_ERR_HANDLE;
        return _fb.BindTo(*this);
    }

    int Buffer(void const* _ptr, size_t _bytes, srg_handle* _out)
    {
        auto key = std::make_pair(_ptr, _bytes);
        auto it = m_buffers.find(key);
        if (it != m_buffers.end()) {
            *_out = it->second;
            return SRG_OK;
        }
        int st = srg_register_buffer(m_ctx, _ptr, _bytes, _out);
        if (!st) m_buffers[key] = *_out;
        return st;
    }

    int Texture(Tex::TextureData const* _tex, srg_handle* _out)
    {
        auto it = m_textures.find(_tex);
        if (it != m_textures.end()) {
            *_out = it->second;
            return SRG_OK;
        }
        int st = srg_register_texture(m_ctx, _tex->m_texels.data(), _tex->m_texels.size(), _tex->m_mipOffsets,
                                      _tex->m_widthLog2, _tex->m_heightLog2, _tex->m_numMips, _out);
        if (!st) m_textures[_tex] = *_out;
        return st;
    }

    srg_ctx* m_ctx = nullptr;
    int m_status = 0;
    uint32_t m_numDraws = 0;
    std::map<std::pair<void const*, size_t>, srg_handle> m_buffers;
    std::map<Tex::TextureData const*, srg_handle> m_textures;
    RenderContext* m_prevCurrent = nullptr;
    std::vector<FrameBuffer*> m_fbs;  // framebuffers bound to this context

    friend struct FrameBuffer;
};

inline void FrameBuffer::Init(uint32_t _width, uint32_t _height, bool _colour, bool _depth)
{
    m_width = _width;
    m_height = _height;
    m_flags = (_colour ? SRG_FB_COLOUR : 0u) | (_depth ? SRG_FB_DEPTH : 0u);
    for (uint32_t i = 0; i < 2; ++i) {
        m_bufferedPlanes[i].m_plane = i;
        m_bufferedPlanes[i].m_width = _width;
        m_bufferedPlanes[i].m_height = _height;
        m_bufferedPlanes[i].m_tilesX = (_width + 63) / 64;
        m_bufferedPlanes[i].m_tilesY = (_height + 63) / 64;
        m_bufferedPlanes[i].m_owner = this;
    }
}

inline int FrameBuffer::BindTo(RenderContext& _ctx)
{
    m_status = srg_fb_create(_ctx.m_ctx, m_width, m_height, m_flags, &m_fb);
    if (m_status) {
        m_fb = nullptr;
        return m_status;
    }
    m_ctx = &_ctx;
    _ctx.m_fbs.push_back(this);
    for (uint32_t i = 0; i < 2; ++i) m_bufferedPlanes[i].m_fb = m_fb;
    if (m_writePlane) srg_fb_swap_planes(m_fb);  // swapped before it was bound
    return SRG_OK;
}

inline FrameBuffer::FrameBuffer(uint32_t _width, uint32_t _height, bool _colour, bool _depth)
{
    Init(_width, _height, _colour, _depth);
    if (RenderContext* ctx = RenderContext::Current()) BindTo(*ctx);
}

inline FrameBuffer::FrameBuffer(RenderContext& _ctx, uint32_t _width, uint32_t _height, bool _colour, bool _depth)
{
    Init(_width, _height, _colour, _depth);
    BindTo(_ctx);
}

inline FrameBuffer::~FrameBuffer()
{
    if (!m_ctx) return;  // never bound, or its context has shut down
    if (m_ctx->m_ctx) srg_fb_destroy(m_ctx->m_ctx, m_fb);
    auto& v = m_ctx->m_fbs;
    v.erase(std::remove(v.begin(), v.end(), this), v.end());
}

}  // namespace sr
