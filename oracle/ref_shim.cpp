// ref_shim.cpp -- C entry points over the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY. Compiled together with the reference's
// src/utils/utils.cpp (triangleIntersect / project6, utils.cpp:71-169) and its
// headers include/utils/utils.h (aabb_box_t, utils.h:15-50) and the vendored
// glm 0.9.9.8, straight from the reference tree, into oracle/_ref/. Tests use
// it to check the oracle's restatement of those functions against the
// reference itself. Nothing here is part of the product.
#include <cstdint>

#include <glm/glm.hpp>

#include "utils/utils.h"

static glm::vec3 v3(const float* p) { return glm::vec3(p[0], p[1], p[2]); }

extern "C" {

// triangleIntersect(P1..P3, Q1..Q3), utils.cpp:97-169
int ref_tri_intersect(const float* P, const float* Q)
{
    return triangleIntersect(v3(P), v3(P + 3), v3(P + 6), v3(Q), v3(Q + 3), v3(Q + 6)) ? 1 : 0;
}

// batched form: n pairs, P and Q are n*9 floats
void ref_tri_intersect_many(const float* P, const float* Q, uint64_t n, uint8_t* out)
{
    for (uint64_t i = 0; i < n; i++) out[i] = (uint8_t)ref_tri_intersect(P + 9 * i, Q + 9 * i);
}

// SimpleBVH leaf box: default aabb_box_t merged with each vertex's point box
// (simpleBVH.cpp:136-139)
void ref_leaf_box(const float* v0, const float* v1, const float* v2, float* out)
{
    aabb_box_t b;
    b.merge(aabb_box_t{v3(v0), v3(v0)});
    b.merge(aabb_box_t{v3(v1), v3(v1)});
    b.merge(aabb_box_t{v3(v2), v3(v2)});
    for (int k = 0; k < 3; k++) {
        out[k] = b.m_minimum[k];
        out[3 + k] = b.m_maximum[k];
    }
}

// internal node: init(left) then merge(right) (simpleBVH.cpp:107-113)
void ref_merge(const float* l, const float* r, float* out)
{
    aabb_box_t b;
    b.init(aabb_box_t{v3(l), v3(l + 3)});
    b.merge(aabb_box_t{v3(r), v3(r + 3)});
    for (int k = 0; k < 3; k++) {
        out[k] = b.m_minimum[k];
        out[3 + k] = b.m_maximum[k];
    }
}

// aabb_box_t::overlap, utils.h:39-44
int ref_overlap(const float* a, const float* b)
{
    aabb_box_t x{v3(a), v3(a + 3)};
    aabb_box_t y{v3(b), v3(b + 3)};
    return x.overlap(y) ? 1 : 0;
}

// the Mesh::setupAABB loop (mesh.cpp:91-98) with the reference's glm min/max
void ref_mesh_box(const float* xyz, uint32_t V, float* out)
{
    aabb_box_t b;
    for (uint32_t i = 0; i < V; i++) {
        glm::vec3 p = v3(xyz + 3 * i);
        b.m_maximum = glm::max(p, b.m_maximum);
        b.m_minimum = glm::min(p, b.m_minimum);
    }
    for (int k = 0; k < 3; k++) {
        out[k] = b.m_minimum[k];
        out[3 + k] = b.m_maximum[k];
    }
}

// glm::mat4 * glm::vec4(p, 1) as transform_vec4_kernel applies it
// (transform.cu:35-41); m is 16 floats, column-major (glm::make_mat4 order)
void ref_xform(const float* m, const float* xyz, uint32_t n, float* out)
{
    glm::mat4 M;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) M[c][r] = m[4 * c + r];
    for (uint32_t i = 0; i < n; i++) {
        const glm::vec4 p = M * glm::vec4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0f);
        out[3 * i] = p.x;
        out[3 * i + 1] = p.y;
        out[3 * i + 2] = p.z;
    }
}

}  // extern "C"
