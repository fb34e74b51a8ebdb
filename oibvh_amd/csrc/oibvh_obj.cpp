// oibvh_obj.cpp -- a minimal Wavefront OBJ reader: the mesh ingest the
// reference gets from assimp's Model (model.cpp:40-186, Model::loadModel with
// aiProcess_Triangulate), reduced to what the collision path reads: vertex
// positions and triangle indices.
//
// Accepted:
//   v x y z [w]        position (w ignored)
//   f c1 c2 c3 ...     polygon; each corner is i, i/t, i//n or i/t/n with i
//                      1-based, or negative (relative to the positions read
//                      so far); polygons of k > 3 corners become the fan
//                      (c1, cj, cj+1), j = 2..k-1
//   # ...              comment; every other statement (vt, vn, o, g, s,
//                      usemtl, mtllib, l, p, ...) is skipped
// All groups and objects form one mesh, and every position is kept in file
// order (assimp splits meshes by material and keeps only referenced corners,
// so its vertex numbering differs; positions and triangles are the same for a
// single-object file whose vertices are all referenced).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oibvh_internal.h"

namespace {

struct ObjData {
    std::vector<float> xyz;
    std::vector<uint32_t> idx;
    std::string err;
};

bool parse_obj(const char* path, ObjData& o)
{
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        o.err = std::string("cannot open ") + path + ": " + std::strerror(errno);
        return false;
    }
    std::string line;
    std::vector<int64_t> poly;
    char buf[4096];
    uint64_t lineno = 0;
    bool ok = true;
    auto fail = [&](const char* what) {
        char m[256];
        std::snprintf(m, sizeof m, "%s:%llu: %s", path, (unsigned long long)lineno, what);
        o.err = m;
        ok = false;
    };
    while (ok) {
        line.clear();
        bool got = false;
        while (std::fgets(buf, sizeof buf, fp)) {
            got = true;
            line += buf;
            if (!line.empty() && line.back() == '\n') break;
        }
        if (!got) break;
        lineno++;
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') s++;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            char* e = nullptr;
            const char* q = s + 2;
            float p[3];
            for (int k = 0; k < 3; k++) {
                p[k] = std::strtof(q, &e);
                if (e == q) {
                    fail("a position needs three numbers");
                    break;
                }
                q = e;
            }
            if (!ok) break;
            o.xyz.insert(o.xyz.end(), p, p + 3);
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            poly.clear();
            const char* q = s + 2;
            for (;;) {
                while (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n') q++;
                if (!*q) break;
                char* e = nullptr;
                const long long v = std::strtoll(q, &e, 10);
                if (e == q || v == 0) {
                    fail("bad face corner");
                    break;
                }
                const int64_t nv = (int64_t)(o.xyz.size() / 3);
                const int64_t i = v > 0 ? v - 1 : nv + v;
                if (i < 0 || i >= nv) {
                    fail("face corner refers to a position not yet read");
                    break;
                }
                poly.push_back(i);
                q = e;
                while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n') q++;  // skip /t/n
            }
            if (!ok) break;
            if (poly.size() < 3) {
                fail("a face needs three corners");
                break;
            }
            for (size_t j = 1; j + 1 < poly.size(); j++) {
                o.idx.push_back((uint32_t)poly[0]);
                o.idx.push_back((uint32_t)poly[j]);
                o.idx.push_back((uint32_t)poly[j + 1]);
            }
        }
        if (o.xyz.size() / 3 > 0xffffffffull || o.idx.size() / 3 > 0xffffffffull) fail("mesh too large");
    }
    std::fclose(fp);
    return ok;
}

}  // namespace

extern "C" oibvh_status oibvh_obj_read(oibvh_ctx* ctx, const char* path, float* xyz, uint64_t xyz_cap,
                                       uint32_t* idx, uint64_t idx_cap, uint32_t* V, uint32_t* F)
{
    if (!path || !V || !F) return oi_set_err(ctx, OIBVH_ERR_INVALID, "oibvh_obj_read: null argument");
    ObjData o;
    if (!parse_obj(path, o)) return oi_set_err(ctx, OIBVH_ERR_INVALID, "%s", o.err.c_str());
    *V = (uint32_t)(o.xyz.size() / 3);
    *F = (uint32_t)(o.idx.size() / 3);
    if (!xyz || !idx || xyz_cap < o.xyz.size() || idx_cap < o.idx.size())
        return oi_set_err(ctx, OIBVH_ERR_CAPACITY, "need %u positions and %u faces", *V, *F);
    if (!o.xyz.empty()) std::memcpy(xyz, o.xyz.data(), o.xyz.size() * sizeof(float));
    if (!o.idx.empty()) std::memcpy(idx, o.idx.data(), o.idx.size() * sizeof(uint32_t));
    return OIBVH_OK;
}
