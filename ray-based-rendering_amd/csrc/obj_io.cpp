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
// Vertex normals follow the reference's rule too: a corner takes the normal its "vn" index
// names (none when the index is absent or out of range), and the mesh keeps normals only when
// some corner has one and every corner's normal is nonzero (obj_importer.cpp:135-143);
// rbr_obj_parse_normals returns them per triangle corner, as written in the file (the scene
// normalizes them, as Triangle's constructor does).  Texture coordinates are parsed and
// counted.
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
    std::vector<float> triNormals;   // 9 per triangle, when the mesh keeps vertex normals
    size_t normals = 0, uvs = 0;
    bool ok = true;
};

ObjResult parseObj(const std::string& text) {
    ObjResult r;
    std::vector<float> pos;    // 3 per vertex
    std::vector<float> nrm;    // 3 per "vn" record
    std::vector<float> cornerN;   // normal of every face corner, 3 per corner (0 when none)
    bool anyNormals = false;
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
            if (ls >> x >> y >> z) {
                nrm.insert(nrm.end(), {x, y, z});
                ++r.normals;
            }
        } else if (tag == "vt") {
            float u;
            if (ls >> u) ++r.uvs;
        } else if (tag == "f") {
            std::vector<int> corner;   // vertex index per corner
            std::vector<int> cn;       // normal index per corner (-1: none)
            std::string tok;
            while (ls >> tok) {
                int v, vt, vn;
                splitCorner(tok, v, vt, vn);
                const int p = toZeroBased(v, pos.size() / 3);
                if (p < 0) {
                    r.ok = false;
                    r.tris.clear();
                    r.triNormals.clear();
                    return r;
                }
                const int ni = toZeroBased(vn, nrm.size() / 3);
                corner.push_back(p);
                cn.push_back(ni);
                if (ni >= 0) anyNormals = true;
                for (int k = 0; k < 3; ++k) cornerN.push_back(ni >= 0 ? nrm[3 * size_t(ni) + k] : 0.0f);
            }
            for (size_t i = 1; i + 1 < corner.size(); ++i)
                for (size_t c : {size_t(0), i, i + 1}) {
                    const size_t pc = size_t(corner[c]);
                    r.tris.insert(r.tris.end(), {pos[3 * pc], pos[3 * pc + 1], pos[3 * pc + 2]});
                    for (int k = 0; k < 3; ++k)
                        r.triNormals.push_back(cn[c] >= 0 ? nrm[3 * size_t(cn[c]) + k] : 0.0f);
                }
        }
    }
    if (r.tris.empty()) r.ok = false;   // an empty mesh is a failed import, as in the reference
    bool complete = anyNormals;
    for (size_t k = 0; complete && k < cornerN.size(); k += 3) {
        const float x = cornerN[k], y = cornerN[k + 1], z = cornerN[k + 2];
        if (x * x + y * y + z * z == 0.0f) complete = false;   // Vec3::lengthSquared() == 0
    }
    if (!complete) r.triNormals.clear();
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

int rbr_obj_parse_normals(const char* text, size_t length, float* normals, size_t capacity, size_t* count) {
    if (!text || !count) return RBR_ERR_INVALID;
    const ObjResult r = parseObj(std::string(text, length));
    *count = r.triNormals.size() / 9;
    if (!r.ok) return RBR_ERR_INVALID;
    if (normals) {
        if (capacity < *count) return RBR_ERR_INVALID;
        for (size_t i = 0; i < r.triNormals.size(); ++i) normals[i] = r.triNormals[i];
    }
    return RBR_OK;
}

}  // extern "C"
