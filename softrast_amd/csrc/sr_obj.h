// sr_obj.h -- scene ingest: Obj::Model::Load (Viewer/Obj.cpp:374-567) on the
// host, producing meshes in the layout the rasteriser draws (Obj::Vertex
// {pos3, norm3, uv2}, u16 or u32 indices).
//
// Parse rules, as the reference applies them:
//   * each line is stripped of leading blanks and trailing blanks / CR / LF
//     (StripWhiteSpaceAndNewLine, Obj.cpp:136-155);
//   * "v x y z", "vt u v" (v -> 1 - v with FLIP_UVS), "vn x y z" are read with
//     %f and must yield 3 / 2 / 3 numbers, else the load fails (:423-477);
//   * "f": up to four corners "p", "p/t", "p//n" or "p/t/n"; an index is
//     optional blanks, an optional '-', decimal digits; a position index of 0
//     (or no digits) ends the face (:164-234). Four corners are split into
//     (0, 1, 2) and (0, 2, 3) (:236-252). Index k > 0 means element k - 1,
//     k < 0 the k-th from the end of the elements read so far, 0 element 0
//     (FixupFaceIdx, :157-162);
//   * each distinct (position, uv, normal) index triple of the current mesh
//     becomes one vertex, in order of first use (:255-313); a normal or uv
//     list that is empty gives zeros, an index past its list fails the load;
//   * "g" closes the current mesh if it has vertices (:479-486); positions,
//     uvs and normals stay shared across meshes; "usemtl name" selects the
//     material of that name for meshes closed later (:487-502); "mtllib file"
//     reads "newmtl" / "map_Kd" lines from file, relative to the .obj's
//     directory (:318-359, :504-529);
//   * a mesh has u32 indices when it has more than 65,535 vertices, else u16
//     (FinalizeMesh, :90-123); FLIP_WINDING swaps the 2nd and 3rd index of
//     every triangle of every mesh (:62-70, :538-551).
// The reference also writes and reads a binary cache ("<path>.bin") through
// its kt serializer; that format belongs to the absent kt library and is not
// reproduced. Texture files named by map_Kd are reported, not decoded.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace srg_obj {

struct Vertex {
    float pos[3], norm[3], uv[2];
};
static_assert(sizeof(Vertex) == 32, "Obj::Vertex layout");

struct Mesh {
    std::vector<uint8_t> index_data;
    uint32_t index_stride = 2;
    uint32_t num_indices = 0;
    std::vector<Vertex> vertices;
    uint32_t material = 0;
};

struct Material {
    std::string name, diffuse;
};

struct Model {
    std::vector<Mesh> meshes;
    std::vector<Material> materials;
};

enum : uint32_t { kFlipWinding = 1u, kFlipUvs = 4u };

inline char* strip(char* s)
{
    while (*s == ' ' || *s == '\t') ++s;
    size_t n = std::strlen(s);
    while (n > 0 && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\r' || s[n - 1] == '\n')) s[--n] = '\0';
    return s;
}

inline uint32_t fix_index(int32_t k, uint32_t count)
{
    if (k == 0) return 0;
    return k < 0 ? (uint32_t)((int32_t)count + k) : (uint32_t)(k - 1);
}

struct Triple {
    uint32_t p, t, n;
    bool operator==(const Triple& o) const { return p == o.p && t == o.t && n == o.n; }
};
struct TripleHash {
    size_t operator()(const Triple& k) const
    {
        uint64_t h = k.p * 0x9E3779B97F4A7C15ull;
        h ^= (k.t + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
        h ^= (k.n + 0x85EBCA77C2B2AE63ull) * 0x165667B19E3779F9ull;
        return (size_t)(h ^ (h >> 29));
    }
};

class Parser {
public:
    Parser(Model& m, uint32_t flags, std::string root) : model_(m), flags_(flags), root_(std::move(root)) {}

    // One stripped line; false on a malformed line.
    bool line(char* s)
    {
        switch (s[0]) {
        case 'v':
            if (s[1] == ' ') {
                float v[3];
                if (std::sscanf(s + 2, "%f %f %f", v, v + 1, v + 2) != 3) return false;
                pos_.insert(pos_.end(), v, v + 3);
            } else if (s[1] == 't') {
                float v[2];
                if (std::sscanf(s + 3, "%f %f", v, v + 1) != 2) return false;
                if (flags_ & kFlipUvs) v[1] = 1.0f - v[1];
                uv_.insert(uv_.end(), v, v + 2);
            } else if (s[1] == 'n') {
                float v[3];
                if (std::sscanf(s + 3, "%f %f %f", v, v + 1, v + 2) != 3) return false;
                nrm_.insert(nrm_.end(), v, v + 3);
            }
            return true;
        case 'f': return face(s + 2);
        case 'g':
            if (!verts_.empty()) close_mesh();
            return true;
        case 'u':
            if (std::strncmp(s, "usemtl", 6) == 0) {
                const char* name = strip(s + 6);
                for (uint32_t i = 0; i < model_.materials.size(); ++i)
                    if (model_.materials[i].name == name) {
                        material_ = i;
                        break;
                    }
            }
            return true;
        case 'm':
            if (std::strncmp(s, "mtllib", 6) == 0) {
                const char* name = s + 6;
                while (*name == ' ' || *name == '\t') ++name;
                if (*name) read_mtl(root_ + name);
            }
            return true;
        default: return true;
        }
    }

    void finish()
    {
        if (!verts_.empty()) close_mesh();
        if (flags_ & kFlipWinding)
            for (Mesh& m : model_.meshes) {
                for (uint32_t i = 0; i + 2 < m.num_indices; i += 3) swap_indices(m, i + 1, i + 2);
            }
    }

private:
    static void swap_indices(Mesh& m, uint32_t a, uint32_t b)
    {
        const uint32_t s = m.index_stride;
        uint8_t t[4];
        std::memcpy(t, &m.index_data[a * s], s);
        std::memcpy(&m.index_data[a * s], &m.index_data[b * s], s);
        std::memcpy(&m.index_data[b * s], t, s);
    }

    static int32_t read_int(const char*& p)
    {
        while (*p == ' ' || *p == '\t') ++p;
        int32_t sign = 1;
        if (*p == '-') {
            sign = -1;
            ++p;
        }
        int32_t v = 0;
        while (*p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
        return v * sign;
    }

    // One corner of an "f" line as written: position, uv and normal indices
    // (0 where the corner names none).
    struct Corner {
        int32_t p = 0, t = 0, n = 0;
    };

    // Reads "p", "p/t", "p//n" or "p/t/n" at s. A position index of 0 (or no
    // digits) is no corner: the face ends before it.
    static bool read_corner(const char*& s, Corner& c)
    {
        c = Corner{};
        c.p = read_int(s);
        if (c.p == 0) return false;
        if (*s != '/') return true;
        ++s;
        if (*s != '/') {  // "p/t" or "p/t/n"
            c.t = read_int(s);
            if (*s != '/') return true;
        }
        ++s;  // "p//n" or the second '/' of "p/t/n"
        c.n = read_int(s);
        return true;
    }

    // Appends the mesh vertex of corner c to the index list: the existing
    // vertex of the same (position, uv, normal) triple, else a new one. A new
    // vertex whose index lies past its list fails the load; an empty uv or
    // normal list gives zeros.
    bool emit_corner(const Corner& c)
    {
        const uint32_t np = (uint32_t)(pos_.size() / 3), nt = (uint32_t)(uv_.size() / 2), nn = (uint32_t)(nrm_.size() / 3);
        const Triple k{ fix_index(c.p, np), fix_index(c.t, nt), fix_index(c.n, nn) };
        const auto found = map_.find(k);
        if (found != map_.end()) {
            idx_.push_back(found->second);
            return true;
        }
        if (k.p >= np || (nt && k.t >= nt) || (nn && k.n >= nn)) return false;
        Vertex v{};
        std::memcpy(v.pos, &pos_[3 * (size_t)k.p], 12);
        if (nn) std::memcpy(v.norm, &nrm_[3 * (size_t)k.n], 12);
        if (nt) std::memcpy(v.uv, &uv_[2 * (size_t)k.t], 8);
        const uint32_t id = (uint32_t)verts_.size();
        map_.emplace(k, id);
        idx_.push_back(id);
        verts_.push_back(v);
        return true;
    }

    // "f": at most four corners are read. Three are one triangle; four are the
    // triangles (0 1 2) and (0 2 3); one or two corners are appended as they
    // are, so their indices stay in the mesh's index list.
    bool face(const char* s)
    {
        Corner corner[4];
        int n = 0;
        while (*s && n < 4 && read_corner(s, corner[n])) ++n;
        static const int kQuadFan[6] = { 0, 1, 2, 0, 2, 3 };
        if (n == 4) {
            for (int i : kQuadFan)
                if (!emit_corner(corner[i])) return false;
            return true;
        }
        for (int i = 0; i < n; ++i)
            if (!emit_corner(corner[i])) return false;
        return true;
    }

    void close_mesh()
    {
        Mesh m;
        m.material = material_;
        m.index_stride = verts_.size() > 0xFFFFu ? 4u : 2u;
        m.num_indices = (uint32_t)idx_.size();
        m.index_data.resize((size_t)m.num_indices * m.index_stride);
        for (uint32_t i = 0; i < m.num_indices; ++i) {
            if (m.index_stride == 4) std::memcpy(&m.index_data[4 * (size_t)i], &idx_[i], 4);
            else {
                const uint16_t v = (uint16_t)idx_[i];
                std::memcpy(&m.index_data[2 * (size_t)i], &v, 2);
            }
        }
        m.vertices.swap(verts_);
        model_.meshes.push_back(std::move(m));
        verts_.clear();
        idx_.clear();
        map_.clear();
    }

    void read_mtl(const std::string& path)
    {
        FILE* f = std::fopen(path.c_str(), "r");
        if (!f) return;
        char buf[2048];
        Material* cur = nullptr;
        while (std::fgets(buf, sizeof(buf), f)) {
            char* s = strip(buf);
            if (std::strncmp(s, "newmtl", 6) == 0) {
                model_.materials.push_back(Material{ strip(s + 6), std::string() });
                cur = &model_.materials.back();
            } else if (std::strncmp(s, "map_Kd", 6) == 0 && cur) {
                cur->diffuse = root_ + strip(s + 6);
            }
        }
        std::fclose(f);
    }

    Model& model_;
    uint32_t flags_;
    std::string root_;
    std::vector<float> pos_, uv_, nrm_;
    std::vector<Vertex> verts_;
    std::vector<uint32_t> idx_;
    std::unordered_map<Triple, uint32_t, TripleHash> map_;
    uint32_t material_ = 0;
};

// Feeds `text` to the parser line by line as fgets(buf, 2048, file) splits
// it (a line longer than 2,047 bytes arrives in pieces), then finishes.
inline bool parse_text(Parser& parser, const char* text, size_t bytes)
{
    char buf[2048];
    size_t i = 0;
    while (i < bytes) {
        size_t n = 0;
        while (i < bytes && n < sizeof(buf) - 1) {
            const char c = text[i++];
            buf[n++] = c;
            if (c == '\n') break;
        }
        buf[n] = '\0';
        if (!parser.line(strip(buf))) return false;
    }
    parser.finish();
    return true;
}

}  // namespace srg_obj
