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

    bool face(const char* p)
    {
        int32_t ip[6] = {}, it[6] = {}, in[6] = {};
        int nv = 0;
        while (*p && nv < 4) {
            ip[nv] = read_int(p);
            if (ip[nv] == 0) break;
            const int c = nv++;
            if (*p != '/') continue;
            ++p;
            if (*p == '/') {
                ++p;
                in[c] = read_int(p);
                continue;
            }
            it[c] = read_int(p);
            if (*p != '/') continue;
            ++p;
            in[c] = read_int(p);
        }
        if (nv == 4) {  // corners 0 1 2 3 -> triangles (0 1 2) (0 2 3)
            ip[5] = ip[3]; it[5] = it[3]; in[5] = in[3];
            ip[3] = ip[0]; it[3] = it[0]; in[3] = in[0];
            ip[4] = ip[2]; it[4] = it[2]; in[4] = in[2];
            nv = 6;
        }
        const uint32_t np = (uint32_t)(pos_.size() / 3), nt = (uint32_t)(uv_.size() / 2), nn = (uint32_t)(nrm_.size() / 3);
        for (int c = 0; c < nv; ++c) {
            const Triple k{ fix_index(ip[c], np), fix_index(it[c], nt), fix_index(in[c], nn) };
            const auto found = map_.find(k);
            if (found != map_.end()) {
                idx_.push_back(found->second);
                continue;
            }
            const uint32_t id = (uint32_t)verts_.size();
            idx_.push_back(id);
            map_.emplace(k, id);
            if (k.p >= np) return false;
            Vertex v;
            std::memcpy(v.pos, &pos_[3 * (size_t)k.p], 12);
            if (nn) {
                if (k.n >= nn) return false;
                std::memcpy(v.norm, &nrm_[3 * (size_t)k.n], 12);
            } else {
                std::memset(v.norm, 0, 12);
            }
            if (nt) {
                if (k.t >= nt) return false;
                std::memcpy(v.uv, &uv_[2 * (size_t)k.t], 8);
            } else {
                std::memset(v.uv, 0, 8);
            }
            verts_.push_back(v);
        }
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
