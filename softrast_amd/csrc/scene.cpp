// scene.cpp -- seeded synthetic scenes and textures for benchmarks and tests.
//
// Host C++ (no GPU code). Builds the inputs the rasteriser is measured on:
// meshes in SoftRast's Obj::Vertex layout {pos3, norm3, uv2} (32 bytes,
// Viewer/Obj.h:16-21), index buffers, column-major view-projection matrices
// (reverse-Z PerspectiveLH_ZO as Viewer/Scene.cpp:20-22 and Camera.cpp:170
// use it) and RGBA8 value-noise textures.
//
// Random numbers: SplitMix64 expands a 64-bit seed, PCG32 (XSH-RR) draws.
// Generation arithmetic is double precision and deterministic on x86-64.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

namespace {

uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t mix2(uint64_t a, uint64_t b)
{
    uint64_t s = a ^ (b * 0xD1B54A32D192ED03ull);
    return splitmix64(s);
}

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    explicit Pcg32(uint64_t seed, uint64_t stream = 0)
    {
        uint64_t s = seed;
        uint64_t a = splitmix64(s), b = splitmix64(s) ^ stream;
        inc = (b << 1) | 1u;
        state = 0;
        next();
        state += a;
        next();
    }
    uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    double uniform() { return (double)(next() >> 8) * (1.0 / 16777216.0); }   // [0,1)
    double normal()
    {
        double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct V3 { double x, y, z; };
V3 sub(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
V3 cross(V3 a, V3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 norm(V3 a) { double l = std::sqrt(dot(a, a)); return { a.x / l, a.y / l, a.z / l }; }

struct Vertex { float pos[3], nrm[3], uv[2]; };
static_assert(sizeof(Vertex) == 32, "Obj::Vertex layout");

// Appends one triangle so that it is front-facing for a viewer outside the
// surface whose outward normal is n: SoftRast keeps triangles whose
// screen-space (y down) signed area is positive (Binning.cpp:305-311), which
// for a proper rotation means cross(v1-v0, v2-v0) . n_out < 0 in world space.
template <typename Idx>
void push_tri(std::vector<Idx>& idx, const std::vector<Vertex>& vtx, uint32_t a, uint32_t b, uint32_t c, V3 n_out)
{
    V3 p0 = { vtx[a].pos[0], vtx[a].pos[1], vtx[a].pos[2] };
    V3 p1 = { vtx[b].pos[0], vtx[b].pos[1], vtx[b].pos[2] };
    V3 p2 = { vtx[c].pos[0], vtx[c].pos[1], vtx[c].pos[2] };
    if (dot(cross(sub(p1, p0), sub(p2, p0)), n_out) < 0) { idx.push_back((Idx)a); idx.push_back((Idx)b); idx.push_back((Idx)c); }
    else { idx.push_back((Idx)a); idx.push_back((Idx)c); idx.push_back((Idx)b); }
}

}  // namespace

extern "C" {

// Column-major view-projection: LookAt (left-handed, +y up) then
// PerspectiveLH_ZO(fov_y, aspect, near, far). Passing near > far gives the
// reverse-Z projection SoftRast uses (near = 10000, far = 0.1).
void srs_camera_mvp(const double eye[3], const double target[3], double fov_y_rad, double aspect,
                    double znear, double zfar, float out[16])
{
    V3 e = { eye[0], eye[1], eye[2] }, t = { target[0], target[1], target[2] };
    V3 f = norm(sub(t, e));
    V3 r = norm(cross({ 0, 1, 0 }, f));
    V3 u = cross(f, r);
    double view[16] = { r.x, u.x, f.x, 0, r.y, u.y, f.y, 0, r.z, u.z, f.z, 0, -dot(r, e), -dot(u, e), -dot(f, e), 1 };
    double h = 1.0 / std::tan(fov_y_rad * 0.5), w = h / aspect, q = zfar / (zfar - znear);
    double proj[16] = { w, 0, 0, 0, 0, h, 0, 0, 0, 0, q, 1, 0, 0, -q * znear, 0 };
    for (int c = 0; c < 4; ++c)
        for (int rr = 0; rr < 4; ++rr) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += proj[k * 4 + rr] * view[c * 4 + k];
            out[c * 4 + rr] = (float)s;
        }
}

// RGBA8 lattice value noise, linear row-major: per RGB channel 4 octaves of
// fBm, base period 16 texels (halving per octave), smoothstep interpolation,
// lattice wrapped so the texture tiles; alpha = 255.
void srs_value_noise_rgba8(uint64_t seed, uint32_t width, uint32_t height, uint8_t* out)
{
    const int octaves = 4;
    for (int c = 0; c < 3; ++c) {
        std::vector<float> acc((size_t)width * height, 0.0f);
        double amp = 1.0, total = 0.0;
        for (int o = 0; o < octaves; ++o) {
            uint32_t period = 16u >> o;
            uint32_t gw = std::max(1u, width / period), gh = std::max(1u, height / period);
            std::vector<float> lat((size_t)gw * gh);
            Pcg32 rng(mix2(seed, (uint64_t)(c * 16 + o)), (uint64_t)(c * 16 + o));
            for (auto& v : lat) v = (float)rng.uniform();
            auto rows = [&](uint32_t ya, uint32_t yz) {
                for (uint32_t y = ya; y < yz; ++y) {
                    double fy = (double)y / period;
                    uint32_t y0 = (uint32_t)fy % gh, y1 = (y0 + 1) % gh;
                    double ty = fy - std::floor(fy);
                    ty = ty * ty * (3.0 - 2.0 * ty);
                    for (uint32_t x = 0; x < width; ++x) {
                        double fx = (double)x / period;
                        uint32_t x0 = (uint32_t)fx % gw, x1 = (x0 + 1) % gw;
                        double tx = fx - std::floor(fx);
                        tx = tx * tx * (3.0 - 2.0 * tx);
                        double a = lat[(size_t)y0 * gw + x0], b = lat[(size_t)y0 * gw + x1];
                        double cc = lat[(size_t)y1 * gw + x0], d = lat[(size_t)y1 * gw + x1];
                        double v = (a + (b - a) * tx) + ((cc + (d - cc) * tx) - (a + (b - a) * tx)) * ty;
                        acc[(size_t)y * width + x] += (float)(amp * v);
                    }
                }
            };
            // rows are independent: split them over up to 16 threads (same result)
            const uint32_t nt = std::max(1u, std::min({ 16u, std::thread::hardware_concurrency(), height / 64u }));
            std::vector<std::thread> pool;
            for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(rows, height * t / nt, height * (t + 1) / nt);
            rows(0, height / nt);
            for (auto& th : pool) th.join();
            total += amp;
            amp *= 0.5;
        }
        for (size_t i = 0; i < acc.size(); ++i) {
            double v = acc[i] / total;
            int b = (int)std::floor(v * 255.0 + 0.5);
            out[i * 4 + c] = (uint8_t)std::min(255, std::max(0, b));
        }
    }
    for (size_t i = 0; i < (size_t)width * height; ++i) out[i * 4 + 3] = 255;
}

// Unit-cube grid: gx * gz cubes on a 2-unit cell pitch centred on the origin,
// each with a seeded yaw in [0, 2pi) and height scale U[0.5, 2], base on y = 0.
// 24 vertices (per-face normals, UVs 0..1) and 12 triangles per cube, u16
// indices. Returns the vertex count; call with null buffers to size.
uint32_t srs_cube_grid(uint64_t seed, uint32_t gx, uint32_t gz, Vertex* vout, uint16_t* iout)
{
    uint32_t nv = gx * gz * 24u;
    if (!vout) return nv;
    std::vector<Vertex> vtx;
    std::vector<uint16_t> idx;
    vtx.reserve(nv);
    Pcg32 rng(seed, 1);
    static const int faces[6][3] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    for (uint32_t iz = 0; iz < gz; ++iz)
        for (uint32_t ix = 0; ix < gx; ++ix) {
            double yaw = rng.uniform() * 6.283185307179586;
            double hs = 0.5 + 1.5 * rng.uniform();
            double cx = ((double)ix - (gx - 1) * 0.5) * 2.0, cz = ((double)iz - (gz - 1) * 0.5) * 2.0;
            double cy = std::cos(yaw), sy = std::sin(yaw);
            for (int f = 0; f < 6; ++f) {
                V3 n = { (double)faces[f][0], (double)faces[f][1], (double)faces[f][2] };
                // two axes spanning the face
                V3 a = std::fabs(n.y) > 0.5 ? V3{ 1, 0, 0 } : V3{ 0, 1, 0 };
                V3 b = cross(n, a);
                uint32_t base = (uint32_t)vtx.size();
                static const double corner[4][2] = { { -1, -1 }, { 1, -1 }, { 1, 1 }, { -1, 1 } };
                for (int k = 0; k < 4; ++k) {
                    V3 p = { 0.5 * (n.x + corner[k][0] * a.x + corner[k][1] * b.x),
                             0.5 * (n.y + corner[k][0] * a.y + corner[k][1] * b.y),
                             0.5 * (n.z + corner[k][0] * a.z + corner[k][1] * b.z) };
                    p.y = (p.y + 0.5) * hs;
                    V3 q = { cy * p.x + sy * p.z + cx, p.y, -sy * p.x + cy * p.z + cz };
                    V3 nn = { cy * n.x + sy * n.z, n.y, -sy * n.x + cy * n.z };
                    Vertex v;
                    v.pos[0] = (float)q.x; v.pos[1] = (float)q.y; v.pos[2] = (float)q.z;
                    v.nrm[0] = (float)nn.x; v.nrm[1] = (float)nn.y; v.nrm[2] = (float)nn.z;
                    v.uv[0] = (float)(0.5 * (corner[k][0] + 1.0)); v.uv[1] = (float)(0.5 * (corner[k][1] + 1.0));
                    vtx.push_back(v);
                }
                V3 nn = { cy * n.x + sy * n.z, n.y, -sy * n.x + cy * n.z };
                push_tri(idx, vtx, base, base + 1, base + 2, nn);
                push_tri(idx, vtx, base, base + 2, base + 3, nn);
            }
        }
    std::memcpy(vout, vtx.data(), sizeof(Vertex) * vtx.size());
    std::memcpy(iout, idx.data(), sizeof(uint16_t) * idx.size());
    return nv;
}

// UV spheres (slices x stacks, (slices+1)(stacks+1) vertices, 2*slices*(stacks-1)
// triangles, poles as single triangles) with log-normal radii exp(N(0, sigma)),
// centres uniform in the view frustum of a camera at the origin looking down
// +z (fov_y, aspect) between depths zmin and zmax. The radius scale is chosen
// so that the summed projected sphere area is `depth_complexity` times the
// screen area of a width x height target. UVs: u = 4*slice/slices + U[0,1),
// v = stack/stacks + U[0,1) per sphere. Indices u32, per-sphere vertex blocks
// consecutive. Returns the vertex count; call with null buffers to size.
uint32_t srs_spheres(uint64_t seed, uint32_t nspheres, uint32_t slices, uint32_t stacks, double sigma,
                     double fov_y_rad, double aspect, double zmin, double zmax, uint32_t width, uint32_t height,
                     double depth_complexity, Vertex* vout, uint32_t* iout, double* radius_scale_out)
{
    uint32_t vps = (slices + 1) * (stacks + 1);
    if (!vout) return nspheres * vps;
    Pcg32 rng(seed, 2);
    std::vector<V3> centre(nspheres);
    std::vector<double> rad(nspheres), uo(nspheres), vo(nspheres);
    double tan_y = std::tan(fov_y_rad * 0.5), tan_x = tan_y * aspect;
    double focal = 0.5 * height / tan_y, sum = 0.0;
    for (uint32_t s = 0; s < nspheres; ++s) {
        // uniform in the frustum volume: z with density ~ z^2
        double u = rng.uniform();
        double z = std::cbrt(zmin * zmin * zmin + u * (zmax * zmax * zmax - zmin * zmin * zmin));
        double x = (2.0 * rng.uniform() - 1.0) * tan_x * z;
        double y = (2.0 * rng.uniform() - 1.0) * tan_y * z;
        centre[s] = { x, y, z };
        rad[s] = std::exp(sigma * rng.normal());
        uo[s] = rng.uniform();
        vo[s] = rng.uniform();
        double pr = rad[s] * focal / z;
        sum += 3.141592653589793 * pr * pr;
    }
    double scale = std::sqrt(depth_complexity * (double)width * height / sum);
    if (radius_scale_out) *radius_scale_out = scale;
    std::vector<Vertex> vtx;
    vtx.reserve((size_t)nspheres * vps);
    std::vector<uint32_t> idx;
    idx.reserve((size_t)nspheres * slices * (stacks - 1) * 6);
    for (uint32_t s = 0; s < nspheres; ++s) {
        double r = rad[s] * scale;
        uint32_t base = (uint32_t)vtx.size();
        for (uint32_t i = 0; i <= stacks; ++i) {
            double th = 3.141592653589793 * i / stacks;
            for (uint32_t j = 0; j <= slices; ++j) {
                double ph = 6.283185307179586 * j / slices;
                V3 n = { std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph) };
                Vertex v;
                v.pos[0] = (float)(centre[s].x + r * n.x);
                v.pos[1] = (float)(centre[s].y + r * n.y);
                v.pos[2] = (float)(centre[s].z + r * n.z);
                v.nrm[0] = (float)n.x; v.nrm[1] = (float)n.y; v.nrm[2] = (float)n.z;
                v.uv[0] = (float)(4.0 * j / slices + uo[s]);
                v.uv[1] = (float)((double)i / stacks + vo[s]);
                vtx.push_back(v);
            }
        }
        for (uint32_t i = 0; i < stacks; ++i)
            for (uint32_t j = 0; j < slices; ++j) {
                uint32_t a = base + i * (slices + 1) + j, b = a + 1, c = a + (slices + 1), d = c + 1;
                double thm = 3.141592653589793 * (i + 0.5) / stacks, phm = 6.283185307179586 * (j + 0.5) / slices;
                V3 n = { std::sin(thm) * std::cos(phm), std::cos(thm), std::sin(thm) * std::sin(phm) };
                if (i != 0) push_tri(idx, vtx, a, b, d, n);
                if (i != stacks - 1) push_tri(idx, vtx, a, d, c, n);
            }
    }
    std::memcpy(vout, vtx.data(), sizeof(Vertex) * vtx.size());
    std::memcpy(iout, idx.data(), sizeof(uint32_t) * idx.size());
    return (uint32_t)vtx.size();
}

// Cubes (24 vertices with per-face normals, 12 triangles) with log-normal
// edge lengths exp(N(0, sigma)), centres uniform in the view frustum of a
// camera at the origin looking down +z (as srs_spheres), random yaw and
// pitch. The size scale is chosen so that the summed mean projected area
// (1.5 s^2 for edge s, a convex body's surface area / 4) is
// `depth_complexity` times the screen area. UVs 0..1 per face plus a per-cube
// offset in [0, 4). u32 indices. Returns the vertex count; call with null
// buffers to size.
uint32_t srs_cubes(uint64_t seed, uint32_t ncubes, double sigma, double fov_y_rad, double aspect, double zmin,
                   double zmax, uint32_t width, uint32_t height, double depth_complexity, Vertex* vout,
                   uint32_t* iout)
{
    if (!vout) return ncubes * 24u;
    Pcg32 rng(seed, 4);
    std::vector<V3> centre(ncubes);
    std::vector<double> edge(ncubes), yaw(ncubes), pitch(ncubes), uo(ncubes), vo(ncubes);
    const double tan_y = std::tan(fov_y_rad * 0.5), tan_x = tan_y * aspect;
    const double focal = 0.5 * height / tan_y;
    double sum = 0.0;
    for (uint32_t c = 0; c < ncubes; ++c) {
        const double u = rng.uniform();
        const double z = std::cbrt(zmin * zmin * zmin + u * (zmax * zmax * zmax - zmin * zmin * zmin));
        centre[c] = { (2.0 * rng.uniform() - 1.0) * tan_x * z, (2.0 * rng.uniform() - 1.0) * tan_y * z, z };
        edge[c] = std::exp(sigma * rng.normal());
        yaw[c] = 6.283185307179586 * rng.uniform();
        pitch[c] = 3.141592653589793 * (rng.uniform() - 0.5);
        uo[c] = 4.0 * rng.uniform();
        vo[c] = 4.0 * rng.uniform();
        const double ps = edge[c] * focal / z;
        sum += 1.5 * ps * ps;
    }
    const double scale = ncubes ? std::sqrt(depth_complexity * (double)width * height / sum) : 1.0;
    std::vector<Vertex> vtx;
    vtx.reserve((size_t)ncubes * 24);
    std::vector<uint32_t> idx;
    idx.reserve((size_t)ncubes * 36);
    static const int faces[6][3] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    static const double corner[4][2] = { { -1, -1 }, { 1, -1 }, { 1, 1 }, { -1, 1 } };
    for (uint32_t c = 0; c < ncubes; ++c) {
        const double h = 0.5 * edge[c] * scale;
        const double cyw = std::cos(yaw[c]), syw = std::sin(yaw[c]), cp = std::cos(pitch[c]), sp = std::sin(pitch[c]);
        // rotation: pitch about x, then yaw about y
        auto rot = [&](V3 p) {
            const V3 q = { p.x, cp * p.y - sp * p.z, sp * p.y + cp * p.z };
            return V3{ cyw * q.x + syw * q.z, q.y, -syw * q.x + cyw * q.z };
        };
        for (int f = 0; f < 6; ++f) {
            const V3 n = { (double)faces[f][0], (double)faces[f][1], (double)faces[f][2] };
            const V3 a = std::fabs(n.y) > 0.5 ? V3{ 1, 0, 0 } : V3{ 0, 1, 0 };
            const V3 b = cross(n, a);
            const uint32_t base = (uint32_t)vtx.size();
            const V3 nr = rot(n);
            for (int k = 0; k < 4; ++k) {
                const V3 p = rot({ h * (n.x + corner[k][0] * a.x + corner[k][1] * b.x),
                                   h * (n.y + corner[k][0] * a.y + corner[k][1] * b.y),
                                   h * (n.z + corner[k][0] * a.z + corner[k][1] * b.z) });
                Vertex v;
                v.pos[0] = (float)(centre[c].x + p.x);
                v.pos[1] = (float)(centre[c].y + p.y);
                v.pos[2] = (float)(centre[c].z + p.z);
                v.nrm[0] = (float)nr.x; v.nrm[1] = (float)nr.y; v.nrm[2] = (float)nr.z;
                v.uv[0] = (float)(uo[c] + 0.5 * (corner[k][0] + 1.0));
                v.uv[1] = (float)(vo[c] + 0.5 * (corner[k][1] + 1.0));
                vtx.push_back(v);
            }
            push_tri(idx, vtx, base, base + 1, base + 2, nr);
            push_tri(idx, vtx, base, base + 2, base + 3, nr);
        }
    }
    std::memcpy(vout, vtx.data(), sizeof(Vertex) * vtx.size());
    std::memcpy(iout, idx.data(), sizeof(uint32_t) * idx.size());
    return (uint32_t)vtx.size();
}

// Independent triangles for a width x height target, given directly in clip
// space (draw them with the identity MVP): a fraction `clustered` of the
// triangles around 3 Gaussian cluster centres (sigma = sigma_frac * width,
// centres uniform in the middle half of the screen), the rest uniform over
// the screen. Screen area log-uniform in [area_min, area_max] px^2, random
// orientation and aspect in [1, 3]; depth z/w uniform in (0.05, 0.95) so the
// reverse-Z test orders them. Normals random unit, UVs uniform in [0, 4).
// Triangle t uses vertices 3t..3t+2; 6 * n * 4 bytes of u32 indices are not
// needed, so only vertices are written (3 * n).
void srs_clustered_tris(uint64_t seed, uint32_t n, uint32_t width, uint32_t height, double clustered,
                        double sigma_frac, double area_min, double area_max, Vertex* vout)
{
    Pcg32 rng(seed, 3);
    double cx[3], cy[3];
    for (int k = 0; k < 3; ++k) {
        cx[k] = (0.25 + 0.5 * rng.uniform()) * width;
        cy[k] = (0.25 + 0.5 * rng.uniform()) * height;
    }
    const double sig = sigma_frac * width, la = std::log(area_min), lb = std::log(area_max);
    for (uint32_t t = 0; t < n; ++t) {
        double px, py;
        if (rng.uniform() < clustered) {
            const int k = (int)(rng.next() % 3u);
            px = cx[k] + sig * rng.normal();
            py = cy[k] + sig * rng.normal();
        } else {
            px = rng.uniform() * width;
            py = rng.uniform() * height;
        }
        const double area = std::exp(la + (lb - la) * rng.uniform());
        const double aspect = 1.0 + 2.0 * rng.uniform(), rot = 6.283185307179586 * rng.uniform();
        // triangle with base b and height h = aspect * b / 2: area = b * h / 2
        const double b = std::sqrt(4.0 * area / aspect), h = aspect * b * 0.5;
        const double lx[3] = { -0.5 * b, 0.5 * b, (rng.uniform() - 0.5) * b };
        const double ly[3] = { 0.0, 0.0, h };
        const double z = 0.05 + 0.9 * rng.uniform();
        const double c = std::cos(rot), sn = std::sin(rot);
        for (int k = 0; k < 3; ++k) {
            // counter-clockwise in pixel space (y down) is front-facing here
            const int kk = k == 0 ? 0 : 3 - k;
            const double sx = px + c * lx[kk] - sn * ly[kk], sy = py + sn * lx[kk] + c * ly[kk];
            Vertex v;
            v.pos[0] = (float)(2.0 * sx / width - 1.0);
            v.pos[1] = (float)(1.0 - 2.0 * sy / height);
            v.pos[2] = (float)z;
            V3 nn = norm({ rng.normal(), rng.normal(), rng.normal() });
            v.nrm[0] = (float)nn.x; v.nrm[1] = (float)nn.y; v.nrm[2] = (float)nn.z;
            v.uv[0] = (float)(4.0 * rng.uniform());
            v.uv[1] = (float)(4.0 * rng.uniform());
            vout[(size_t)t * 3 + k] = v;
        }
    }
}

}  // extern "C"
