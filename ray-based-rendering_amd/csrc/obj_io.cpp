// Wavefront OBJ meshes for the device tables: the triangles ObjImporter::parse builds
// (obj_importer.cpp:59-150), as plain vertex triples.  Host code.
//
// Records and rules kept from the reference:
//   "v x y z" (a record that does not read three floats adds no vertex), "vt", "vn",
//   "f c0 c1 c2 ..." with corners v, v/vt, v//vn or v/vt/vn; indices are 1-based, negative
//   ones count back from the last record of their kind; a vertex index outside the list fails
//   the whole parse, texture and normal indices outside theirs are ignored; polygons are
//   fanned around their first corner (c0 ci ci+1); a line whose first token starts with '#'
//   is skipped, and so are unknown records.  Numbers are read with the C++ stream and atoi
//   conversions the reference uses, so the vertex values are the same floats.
// Texture coordinates and vertex normals are parsed and counted; the device path shades
// triangles with their geometric normal.
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "rbr.h"

namespace {

// 1-based or negative (relative) index -> 0-based, or -1 when out of range or 0.
int toZeroBased(int idx, size_t count) {
    if (idx > 0) return idx <= static_cast<int>(count) ? idx - 1 : -1;
    if (idx < 0) {
        const int r = static_cast<int>(count) + idx;
        return r >= 0 ? r : -1;
    }
    return -1;
}

// "v", "v/vt", "v//vn", "v/vt/vn" -> the three integers (0 where absent or empty).
void splitCorner(const std::string& tok, int& v, int& vt, int& vn) {
    vt = vn = 0;
    const size_t a = tok.find('/');
    v = std::atoi(tok.substr(0, a).c_str());
    if (a == std::string::npos) return;
    const size_t b = tok.find('/', a + 1);
    const std::string t = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
    if (!t.empty()) vt = std::atoi(t.c_str());
    if (b != std::string::npos) {
        const std::string n = tok.substr(b + 1);
        if (!n.empty()) vn = std::atoi(n.c_str());
    }
}

struct ObjResult {
    std::vector<float> tris;
    size_t normals = 0, uvs = 0;
    bool ok = true;
};

ObjResult parseObj(const std::string& text) {
    ObjResult r;
    std::vector<float> pos;   // 3 per vertex
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            float x, y, z;
            if (ls >> x >> y >> z) pos.insert(pos.end(), {x, y, z});
        } else if (tag == "vn") {
            float x, y, z;
            if (ls >> x >> y >> z) ++r.normals;
        } else if (tag == "vt") {
            float u;
            if (ls >> u) ++r.uvs;
        } else if (tag == "f") {
            std::vector<int> corner;   // vertex index per corner
            std::string tok;
            while (ls >> tok) {
                int v, vt, vn;
                splitCorner(tok, v, vt, vn);
                const int p = toZeroBased(v, pos.size() / 3);
                if (p < 0) {
                    r.ok = false;
                    r.tris.clear();
                    return r;
                }
                corner.push_back(p);
            }
            for (size_t i = 1; i + 1 < corner.size(); ++i)
                for (int c : {corner[0], corner[i], corner[i + 1]})
                    r.tris.insert(r.tris.end(), {pos[3 * size_t(c)], pos[3 * size_t(c) + 1], pos[3 * size_t(c) + 2]});
        }
    }
    if (r.tris.empty()) r.ok = false;   // an empty mesh is a failed import, as in the reference
    return r;
}

}  // namespace

extern "C" {

int rbr_obj_parse(const char* text, size_t length, float* tris, size_t capacity, size_t* count, size_t* normals,
                  size_t* uvs) {
    if (!text || !count) return RBR_ERR_INVALID;
    const ObjResult r = parseObj(std::string(text, length));
    *count = r.tris.size() / 9;
    if (normals) *normals = r.normals;
    if (uvs) *uvs = r.uvs;
    if (!r.ok) return RBR_ERR_INVALID;
    if (tris) {
        if (capacity < r.tris.size() / 9) return RBR_ERR_INVALID;
        for (size_t i = 0; i < r.tris.size(); ++i) tris[i] = r.tris[i];
    }
    return RBR_OK;
}

}  // extern "C"
