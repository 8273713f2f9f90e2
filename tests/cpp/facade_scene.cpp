// A host program written against the reference's frame API the way
// Viewer/Scene.cpp:32-65 and Viewer/Main.cpp:41-86 use it, compiled against
// include/sr/Renderer.h. Draws a small quad grid textured with a checker
// (UnlitDiffuseShader, then SponzaShader with constants set the way
// SponzaScene::Init / Update set them) and prints the number of non-clear
// pixels after the last blit. With a path argument it also writes its inputs
// and both blitted frames there, as records of (u32 byte count, bytes):
// width and height, vertices, u16 indices, the checker's RGBA8 texels, the
// MVP, the Sponza constants, frame 0's rows, frame 1's rows.
// Exit codes: 0 = rendered (or no GPU: prints "no-device"), 1 = failure.
#include <cstdio>
#include <vector>

#include "sr/Renderer.h"

struct Vertex { float pos[3]; float norm[3]; float uv[2]; };

static void put(std::FILE* f, const void* p, size_t n)
{
    const uint32_t k = (uint32_t)n;
    std::fwrite(&k, 4, 1, f);
    std::fwrite(p, 1, n, f);
}

int main(int argc, char** argv)
{
    // made before any context exists: bound by the first context that uses
    // it, and destroyed after that context shut down
    sr::FrameBuffer early(100, 70);
    if (early.m_ctx) return 1;
    sr::RenderContext renderCtx;
    if (renderCtx.Status() == SRG_ERR_NO_DEVICE) {
        std::printf("no-device\n");
        return 0;
    }
    if (renderCtx.Status()) return 1;
    const uint32_t W = 320, H = 200;
    // as Viewer/Main.cpp:41-42: the framebuffer is made after the context, without naming it
    sr::FrameBuffer framebuffer(W, H);
    if (framebuffer.m_status || framebuffer.m_ctx != &renderCtx) return 1;
    // per-thread host scratch, as the reference's tasks take it
    sr::ThreadScratchAllocator& scratch = renderCtx.ThreadAllocator();
    float* tmp = (float*)scratch.Alloc(4096 * sizeof(float), 64);
    if (!tmp || ((uintptr_t)tmp & 63u)) return 1;
    tmp[4095] = 1.0f;
    scratch.Reset();
    if (renderCtx.BeginFrame() || renderCtx.ClearFrameBuffer(early, 0x5A) || renderCtx.EndFrame()) return 1;
    if (early.m_ctx != &renderCtx || !early.m_fb) return 1;
    {
        std::vector<uint32_t> rows(100 * 70);
        if (renderCtx.Blit(early, (uint8_t*)rows.data())) return 1;
        for (uint32_t p : rows)
            if (p != 0x5A5A5A5Au) return 1;  // ClearFrameBuffer's low-byte memset
    }

    std::vector<uint8_t> checker(64 * 64 * 4);
    for (uint32_t y = 0; y < 64; ++y)
        for (uint32_t x = 0; x < 64; ++x)
            for (uint32_t c = 0; c < 4; ++c) checker[(y * 64 + x) * 4 + c] = ((x / 8 + y / 8) & 1) ? 230 : 20;
    sr::Tex::TextureData tex;
    if (tex.CreateFromRGBA8(checker.data(), 64, 64, true)) return 1;

    // 4x4 quads facing the camera at z = 4 (clip = (x, y, z * 0.25, z))
    std::vector<Vertex> verts;
    std::vector<uint16_t> indices;
    for (int qy = 0; qy < 4; ++qy)
        for (int qx = 0; qx < 4; ++qx) {
            uint16_t base = (uint16_t)verts.size();
            for (int k = 0; k < 4; ++k) {
                float x = -3.0f + 1.5f * qx + ((k & 1) ? 1.2f : 0.0f);
                float y = -3.0f + 1.5f * qy + ((k & 2) ? 1.2f : 0.0f);
                verts.push_back(Vertex{ { x, y, 4.0f }, { 0, 0, -1 }, { (k & 1) ? 1.0f : 0.0f, (k & 2) ? 1.0f : 0.0f } });
            }
            uint16_t q[6] = { (uint16_t)base, (uint16_t)(base + 1), (uint16_t)(base + 2),
                              (uint16_t)(base + 1), (uint16_t)(base + 3), (uint16_t)(base + 2) };
            indices.insert(indices.end(), q, q + 6);
        }
    sr::Mat4 mvp;  // x, y scaled into the frustum; w = z
    float m[16] = { 0.5f, 0, 0, 0, 0, 0.8f, 0, 0, 0, 0, 0.25f, 1, 0, 0, 0, 0 };
    for (int i = 0; i < 16; ++i) mvp.m[i] = m[i];

    srg_sponza_constants sponza{};
    for (int c = 0; c < 3; ++c) {
        sponza.sun_dir[c] = 0.49f;
        sponza.ambient[c] = 0.1f;
    }
    for (int i = 0; i < SRG_SPONZA_LIGHTS; ++i) {
        srg_sponza_light& l = sponza.lights[i];
        l.pos[0] = -3.0f + 0.4f * (float)i;
        l.pos[1] = 1.0f;
        l.pos[2] = 2.0f;
        l.colour[0] = 1.0f;
        l.colour[1] = 0.5f + 0.03f * (float)i;
        l.colour[2] = 0.25f;
        l.intensity = 0.5f;
        l.falloff = 1000.0f;
    }
    if (int st = renderCtx.SetSponzaConstants(sponza)) { std::printf("SetSponzaConstants: %s\n", srg_status_string(st)); return 1; }

    std::vector<uint32_t> pixels(W * H), frames[2];
    for (int frame = 0; frame < 2; ++frame) {
        renderCtx.BeginFrame();
        renderCtx.ClearFrameBuffer(framebuffer, 0);
        sr::DrawCall call;
        call.SetFrameBuffer(&framebuffer);
        call.m_mvp = mvp;
        call.SetAttributeBuffer(verts.data(), sizeof(Vertex), (uint32_t)verts.size(), 6);
        call.m_positionBuffer.m_ptr = verts.data();
        call.m_positionBuffer.m_stride = sizeof(Vertex);
        call.m_positionBuffer.m_num = (uint32_t)verts.size();
        call.m_indexBuffer.m_ptr = indices.data();
        call.m_indexBuffer.m_num = (uint32_t)indices.size();
        call.m_indexBuffer.m_stride = sizeof(uint16_t);
        call.m_pixelUniforms = &tex;
        call.m_pixelShader = frame == 0 ? sr::shader::UnlitDiffuseShader : sr::shader::SponzaShader;
        if (int st = renderCtx.DrawIndexed(call)) { std::printf("DrawIndexed: %s\n", srg_status_string(st)); return 1; }
        if (int st = renderCtx.EndFrame()) { std::printf("EndFrame: %s\n", srg_status_string(st)); return 1; }
        if (int st = renderCtx.Blit(framebuffer, (uint8_t*)pixels.data())) { std::printf("Blit: %s\n", srg_status_string(st)); return 1; }
        frames[frame] = pixels;
    }
    if (argc > 1) {
        std::FILE* f = std::fopen(argv[1], "wb");
        if (!f) return 1;
        const uint32_t wh[2] = { W, H };
        put(f, wh, sizeof(wh));
        put(f, verts.data(), verts.size() * sizeof(Vertex));
        put(f, indices.data(), indices.size() * sizeof(uint16_t));
        put(f, checker.data(), checker.size());
        put(f, m, sizeof(m));
        put(f, &sponza, sizeof(sponza));
        put(f, frames[0].data(), frames[0].size() * 4);
        put(f, frames[1].data(), frames[1].size() * 4);
        std::fclose(f);
    }
    uint32_t drawn = 0;
    for (uint32_t p : pixels) drawn += p != 0;
    std::printf("drawn=%u\n", drawn);
    return drawn > 1000 ? 0 : 1;
}
