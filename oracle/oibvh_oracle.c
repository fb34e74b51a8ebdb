/*
 * oibvh_oracle.c -- CPU restatement of the reference oibvh collision path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the parity checker for the HIP
 * library in oibvh_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. Nothing in the product path links
 * or calls it.
 *
 * Every function follows a reference file:line (paths relative to the
 * hhhcbw/oibvh source tree). Floating-point arithmetic is written out
 * operation by operation and the file is compiled with -ffp-contract=off and
 * without -ffast-math, so each float expression rounds exactly as the
 * reference's glm expressions do.
 *
 * Semantics pinned here (single-threaded, like src/cpu):
 *   - glm::min(a,b) = (b < a) ? b : a and glm::max(a,b) = (a < b) ? b : a
 *     (glm 0.9.9.8 func_common.inl:17-30), thrust::min/max the same contract.
 *   - Morton30: oibvh.cu:6-22 and :42-70; a NaN quantisation input maps to 0
 *     (the reference's (unsigned)NaN cast is undefined; the HIP kernels apply
 *     the same policy).
 *   - stable sort of the 30-bit keys: oibvhTree.cu:287-297.
 *   - tree shape / AABBs: SimpleBVH::recursiveBuild simpleBVH.cpp:122-160,
 *     recursiveRefit :96-115, serialised breadth-first like SimpleBVH::log
 *     :56-79 (the breadth-first order is the oibvh "real" node order).
 *   - traversal: SimpleCollide::detect simpleCollide.cpp:47-161.
 *   - tri-tri: triangleIntersect / project6 utils.cpp:71-169 (17-axis SAT).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* index math: utils.cuh:66-86, oibvh.cuh:56-182                            */
/* ------------------------------------------------------------------------ */

/* next_power_of_two, utils.cuh:66-76 (returns 0 for 0, like the reference) */
EXPORT uint32_t orc_next_pow2(uint32_t x)
{
    x--;
    x |= x >> 1;
    x |= x >> 2;
    x |= x >> 4;
    x |= x >> 8;
    x |= x >> 16;
    x++;
    return x;
}

/* ilog2, utils.cuh:83-86 (x > 0) */
EXPORT uint32_t orc_ilog2(uint32_t x) { return 31u - (uint32_t)__builtin_clz(x); }

/* oibvh_get_size, oibvh.cuh:56-59 */
EXPORT uint32_t orc_tree_size(uint32_t F)
{
    return 2u * F + (uint32_t)__builtin_popcount(orc_next_pow2(F) - F) - 1u;
}

/* oibvh_level_real_node_count, oibvh.cuh:81-85 */
EXPORT uint32_t orc_level_real_count(uint32_t l, uint32_t L, uint32_t vl)
{
    return (1u << l) - (vl >> (L - l));
}

/* oibvh_implicit_to_real, oibvh.cuh:108-113 (implicit index -> real index),
 * computed here by counting the virtual nodes of the levels above. */
EXPORT uint32_t orc_implicit_to_real(uint32_t i, uint32_t L, uint32_t vl)
{
    uint32_t l = orc_ilog2(i + 1u), virt = 0;
    for (uint32_t k = 0; k < l; k++) virt += vl >> (L - k);
    return i - virt;
}

/* oibvh_real_to_implicit, oibvh.cuh:122-136, by scanning the level sizes. */
EXPORT uint32_t orc_real_to_implicit(uint32_t r, uint32_t L, uint32_t vl)
{
    uint32_t start = 0;
    for (uint32_t l = 0; l <= L; l++) {
        uint32_t cnt = orc_level_real_count(l, L, vl);
        if (r < start + cnt) return (1u << l) - 1u + (r - start);
        start += cnt;
    }
    return 0xFFFFFFFFu;
}

/* ------------------------------------------------------------------------ */
/* glm / thrust compare-select                                              */
/* ------------------------------------------------------------------------ */
static inline float gmin(float a, float b) { return (b < a) ? b : a; }
static inline float gmax(float a, float b) { return (a < b) ? b : a; }

/* aabb_box_t default, utils.h:17-18 */
static inline void box_default(float* b)
{
    b[0] = b[1] = b[2] = FLT_MAX;
    b[3] = b[4] = b[5] = -FLT_MAX;
}
/* aabb_box_t::merge, utils.h:33-37 */
static inline void box_merge(float* b, const float* o)
{
    for (int k = 0; k < 3; k++) {
        b[k] = gmin(b[k], o[k]);
        b[3 + k] = gmax(b[3 + k], o[3 + k]);
    }
}
/* aabb_box_t::overlap, utils.h:39-44 (closed intervals) */
static inline int box_overlap(const float* a, const float* b)
{
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) &&
           (a[2] <= b[5] && a[5] >= b[2]);
}
EXPORT int orc_box_overlap(const float* a, const float* b) { return box_overlap(a, b); }

/* ------------------------------------------------------------------------ */
/* mesh AABB: Mesh::setupAABB, mesh.cpp:91-98                                */
/* m_max = glm::max(v, m_max); m_min = glm::min(v, m_min), default box.      */
/* ------------------------------------------------------------------------ */
EXPORT void orc_mesh_box(const float* xyz, uint32_t V, float* out)
{
    float b[6];
    box_default(b);
    for (uint32_t v = 0; v < V; v++)
        for (int k = 0; k < 3; k++) {
            float x = xyz[3 * v + k];
            b[3 + k] = gmax(x, b[3 + k]);
            b[k] = gmin(x, b[k]);
        }
    memcpy(out, b, sizeof b);
}

/* ------------------------------------------------------------------------ */
/* Mesh::transform (mesh.cpp:189-214; transform_vec4_kernel, transform.cu:35-41): */
/* p' = M * vec4(p, 1) in glm's mat4 * vec4 order (type_mat4x4.inl:561-573):   */
/* (m[0] x + m[1] y) + (m[2] z + m[3] w), m[c] the c-th column, w = 1.          */
/* m is column-major (glm::mat4 memory layout).                                 */
/* ------------------------------------------------------------------------ */
EXPORT void orc_xform(const float* m, const float* xyz, uint32_t n, float* out)
{
    for (uint32_t i = 0; i < n; i++) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2], w = 1.0f;
        for (int r = 0; r < 3; r++) {
            const float mul0 = m[r] * x, mul1 = m[4 + r] * y;
            const float add0 = mul0 + mul1;
            const float mul2 = m[8 + r] * z, mul3 = m[12 + r] * w;
            const float add1 = mul2 + mul3;
            out[3 * i + r] = add0 + add1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Morton30: strech_by_3 / morton3D oibvh.cu:6-22, leaf kernel oibvh.cu:42-70 */
/* ------------------------------------------------------------------------ */
static inline uint32_t spread10(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
}

/* thrust::min(thrust::max(p * 1024.f, 0.f), 1023.f) then (unsigned) */
EXPORT uint32_t orc_quantise10(float p)
{
    float q = p * 1024.0f;
    q = (q < 0.0f) ? 0.0f : q;       /* thrust::max(q, 0): q < 0 ? 0 : q   */
    q = (1023.0f < q) ? 1023.0f : q; /* thrust::min(q, 1023): 1023 < q ... */
    if (q != q) return 0u;           /* NaN policy: 0                     */
    return (uint32_t)q;
}

EXPORT uint32_t orc_morton30(float x, float y, float z)
{
    return (spread10(orc_quantise10(x)) << 2) | (spread10(orc_quantise10(y)) << 1) |
           spread10(orc_quantise10(z));
}

/* leaf AABB = min(min(v0, v1), v2) (oibvh.cu:54-59) */
static inline void leaf_box(const uint32_t* f, const float* xyz, float* b)
{
    const float* v0 = xyz + 3u * f[0];
    const float* v1 = xyz + 3u * f[1];
    const float* v2 = xyz + 3u * f[2];
    for (int k = 0; k < 3; k++) {
        b[k] = gmin(gmin(v0[k], v1[k]), v2[k]);
        b[3 + k] = gmax(gmax(v0[k], v1[k]), v2[k]);
    }
}

/* calculate_aabb_and_morton_kernel, oibvh.cu:42-70: per face (original
 * order) the leaf AABB and the Morton key of its centroid in `box`. */
EXPORT void orc_leaf_keys(const uint32_t* faces, const float* xyz, uint32_t F, const float* box,
                          uint32_t* keys, float* leaves /* F*6 or NULL */)
{
    for (uint32_t i = 0; i < F; i++) {
        float b[6], p[3];
        leaf_box(faces + 3u * i, xyz, b);
        if (leaves) memcpy(leaves + 6u * i, b, sizeof b);
        for (int k = 0; k < 3; k++) {
            float c = (b[k] + b[3 + k]) * 0.5f;
            float off = c - box[k];
            float len = box[3 + k] - box[k];
            p[k] = off / len;
        }
        keys[i] = orc_morton30(p[0], p[1], p[2]);
    }
}

/* ------------------------------------------------------------------------ */
/* Morton63: no reference counterpart. The 64-bit form of morton3D with 21    */
/* bits per axis, quantised like quantise10:                                  */
/* min(max(p * 2^21, 0), 2^21 - 1) then (unsigned), NaN -> 0; code =         */
/* x << 2 | y << 1 | z over the bit-spread axes. Changes the leaf order (and  */
/* so the tree), never the pair set.                                          */
/* ------------------------------------------------------------------------ */
static inline uint64_t spread21(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

EXPORT uint32_t orc_quantise21(float p)
{
    float q = p * 2097152.0f;
    q = (q < 0.0f) ? 0.0f : q;
    q = (2097151.0f < q) ? 2097151.0f : q;
    if (q != q) return 0u;
    return (uint32_t)q;
}

EXPORT uint64_t orc_morton63(float x, float y, float z)
{
    return (spread21(orc_quantise21(x)) << 2) | (spread21(orc_quantise21(y)) << 1) | spread21(orc_quantise21(z));
}

/* as orc_leaf_keys, 63-bit keys */
EXPORT void orc_leaf_keys63(const uint32_t* faces, const float* xyz, uint32_t F, const float* box, uint64_t* keys)
{
    for (uint32_t i = 0; i < F; i++) {
        float b[6], p[3];
        leaf_box(faces + 3u * i, xyz, b);
        for (int k = 0; k < 3; k++) {
            float c = (b[k] + b[3 + k]) * 0.5f;
            float off = c - box[k];
            float len = box[3 + k] - box[k];
            p[k] = off / len;
        }
        keys[i] = orc_morton63(p[0], p[1], p[2]);
    }
}

/* stable order of 64-bit keys (the same comparison sort over (key, index)) */
static const uint64_t* g_sort_keys64;
static int cmp_key64_idx(const void* a, const void* b)
{
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    uint64_t ka = g_sort_keys64[ia], kb = g_sort_keys64[ib];
    if (ka != kb) return ka < kb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}
EXPORT void orc_stable_order64(const uint64_t* keys, uint32_t F, uint32_t* perm)
{
    for (uint32_t i = 0; i < F; i++) perm[i] = i;
    g_sort_keys64 = keys;
    qsort(perm, F, sizeof(uint32_t), cmp_key64_idx);
    g_sort_keys64 = NULL;
}

/* stable sort of keys -> permutation (sorted position -> original face),
 * thrust::stable_sort_by_key semantics (oibvhTree.cu:287-297): a
 * comparison sort over (key, original index), which is the stable order. */
static const uint32_t* g_sort_keys;
static int cmp_key_idx(const void* a, const void* b)
{
    uint32_t ia = *(const uint32_t*)a, ib = *(const uint32_t*)b;
    uint32_t ka = g_sort_keys[ia], kb = g_sort_keys[ib];
    if (ka != kb) return ka < kb ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}
EXPORT void orc_stable_order(const uint32_t* keys, uint32_t F, uint32_t* perm)
{
    for (uint32_t i = 0; i < F; i++) perm[i] = i;
    g_sort_keys = keys;
    qsort(perm, F, sizeof(uint32_t), cmp_key_idx);
    g_sort_keys = NULL;
}

/* ------------------------------------------------------------------------ */
/* SimpleBVH restatement: simpleBVH.cpp:96-160                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    float b[6];
    int32_t left, right; /* -1 = none */
    int32_t tri;         /* -1 = internal; else position in the tree's face order */
} onode;

typedef struct orc_tree {
    onode* nodes;
    uint32_t n, cap;
    int32_t root;
    uint32_t depth; /* m_depth = log2(next_power2(F)) */
    uint32_t F, V;
    uint32_t* faces;    /* F*3, in tree order (face i of the tree) */
    uint32_t* orig;     /* F: tree order -> caller's original face id */
    float* xyz;         /* V*3 current positions */
    int32_t* bfs;       /* node ids in breadth-first (real) order */
} orc_tree;

static int32_t new_node(orc_tree* t)
{
    if (t->n == t->cap) {
        t->cap = t->cap ? t->cap * 2 : 1024;
        t->nodes = (onode*)realloc(t->nodes, sizeof(onode) * t->cap);
    }
    onode* n = &t->nodes[t->n];
    box_default(n->b);
    n->left = n->right = -1;
    n->tri = -1;
    return (int32_t)t->n++;
}

/* recursiveBuild, simpleBVH.cpp:122-160 */
static int32_t rec_build(orc_tree* t, int32_t lo, int32_t hi, uint32_t depth)
{
    int32_t id = new_node(t);
    if (lo == hi) {
        if (depth < t->depth) {
            int32_t c = rec_build(t, lo, hi, depth + 1);
            t->nodes[id].left = c;
            box_merge(t->nodes[id].b, t->nodes[c].b);
            return id;
        }
        const uint32_t* f = t->faces + 3u * (uint32_t)lo;
        for (int v = 0; v < 3; v++) {
            float vb[6];
            memcpy(vb, t->xyz + 3u * f[v], 3 * sizeof(float));
            memcpy(vb + 3, t->xyz + 3u * f[v], 3 * sizeof(float));
            box_merge(t->nodes[id].b, vb);
        }
        t->nodes[id].tri = lo;
        return id;
    }
    int32_t count = hi - lo + 1;
    int32_t mid = lo + (int32_t)orc_next_pow2((uint32_t)count) / 2 - 1;
    if ((uint32_t)count <= (1u << (t->depth - depth - 1))) {
        int32_t c = rec_build(t, lo, hi, depth + 1);
        t->nodes[id].left = c;
        box_merge(t->nodes[id].b, t->nodes[c].b);
        return id;
    }
    int32_t l = rec_build(t, lo, mid, depth + 1);
    int32_t r = rec_build(t, mid + 1, hi, depth + 1);
    t->nodes[id].left = l;
    t->nodes[id].right = r;
    box_merge(t->nodes[id].b, t->nodes[l].b);
    box_merge(t->nodes[id].b, t->nodes[r].b);
    return id;
}

/* recursiveRefit, simpleBVH.cpp:96-115: leaf init(v0,v0) merge v1, v2;
 * internal init(left) merge(right). */
static void rec_refit(orc_tree* t, int32_t id)
{
    onode* n = &t->nodes[id];
    if (n->tri != -1) {
        const uint32_t* f = t->faces + 3u * (uint32_t)n->tri;
        const float* v0 = t->xyz + 3u * f[0];
        memcpy(n->b, v0, 3 * sizeof(float));
        memcpy(n->b + 3, v0, 3 * sizeof(float));
        for (int v = 1; v < 3; v++) {
            float vb[6];
            memcpy(vb, t->xyz + 3u * f[v], 3 * sizeof(float));
            memcpy(vb + 3, t->xyz + 3u * f[v], 3 * sizeof(float));
            box_merge(n->b, vb);
        }
        return;
    }
    rec_refit(t, n->left);
    n = &t->nodes[id];
    memcpy(n->b, t->nodes[n->left].b, sizeof n->b);
    if (n->right != -1) {
        rec_refit(t, n->right);
        n = &t->nodes[id];
        box_merge(n->b, t->nodes[n->right].b);
    }
}

/* breadth-first order, SimpleBVH::log simpleBVH.cpp:56-79 */
static void make_bfs(orc_tree* t)
{
    t->bfs = (int32_t*)malloc(sizeof(int32_t) * t->n);
    uint32_t head = 0, tail = 0;
    t->bfs[tail++] = t->root;
    while (head < tail) {
        const onode* n = &t->nodes[t->bfs[head++]];
        if (n->left != -1) t->bfs[tail++] = n->left;
        if (n->right != -1) t->bfs[tail++] = n->right;
    }
}

/* Build a SimpleBVH-shaped tree over the faces taken in `order` (order[i] =
 * original face id of tree leaf i; NULL = original mesh order, which is what
 * src/cpu does; the Morton-stable order reproduces the oibvh leaf order). */
EXPORT orc_tree* orc_tree_build(const uint32_t* faces, uint32_t F, const float* xyz, uint32_t V,
                                const uint32_t* order)
{
    if (F == 0) return NULL;
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->F = F;
    t->V = V;
    t->faces = (uint32_t*)malloc(sizeof(uint32_t) * 3u * F);
    t->orig = (uint32_t*)malloc(sizeof(uint32_t) * F);
    for (uint32_t i = 0; i < F; i++) {
        uint32_t o = order ? order[i] : i;
        t->orig[i] = o;
        memcpy(t->faces + 3u * i, faces + 3u * o, 3 * sizeof(uint32_t));
    }
    t->xyz = (float*)malloc(sizeof(float) * 3u * V);
    memcpy(t->xyz, xyz, sizeof(float) * 3u * V);
    t->depth = orc_ilog2(orc_next_pow2(F));
    t->root = rec_build(t, 0, (int32_t)F - 1, 0);
    make_bfs(t);
    return t;
}

EXPORT void orc_tree_free(orc_tree* t)
{
    if (!t) return;
    free(t->nodes);
    free(t->faces);
    free(t->orig);
    free(t->xyz);
    free(t->bfs);
    free(t);
}

EXPORT uint32_t orc_tree_node_count(const orc_tree* t) { return t->n; }
EXPORT uint32_t orc_tree_depth(const orc_tree* t) { return t->depth; }

/* AABBs in breadth-first (real) order, N*6 floats */
EXPORT void orc_tree_nodes(const orc_tree* t, float* out)
{
    for (uint32_t i = 0; i < t->n; i++) memcpy(out + 6u * i, t->nodes[t->bfs[i]].b, 6 * sizeof(float));
}

/* per real node: tree-order face index for leaves, -1 for internal nodes,
 * plus the number of children (the topology) */
EXPORT void orc_tree_topology(const orc_tree* t, int32_t* tri, int32_t* nchild)
{
    for (uint32_t i = 0; i < t->n; i++) {
        const onode* n = &t->nodes[t->bfs[i]];
        tri[i] = n->tri;
        nchild[i] = (n->left != -1) + (n->right != -1);
    }
}

/* SimpleBVH::refit with new positions (simpleBVH.cpp:81-94) */
EXPORT void orc_tree_refit(orc_tree* t, const float* xyz)
{
    memcpy(t->xyz, xyz, sizeof(float) * 3u * t->V);
    rec_refit(t, t->root);
}

/* ------------------------------------------------------------------------ */
/* SAT: project6 + triangleIntersect, utils.cpp:71-169 (float), and an      */
/* identical expression evaluated in double from widened float inputs.      */
/* ------------------------------------------------------------------------ */
#define DEFINE_SAT(T, NAME, FMAX, FMIN)                                                          \
    typedef struct { T x, y, z; } v3_##T;                                                      \
    static inline v3_##T sub_##T(v3_##T a, v3_##T b)                                           \
    {                                                                                           \
        v3_##T r = {a.x - b.x, a.y - b.y, a.z - b.z};                                           \
        return r;                                                                               \
    }                                                                                           \
    /* glm::cross, func_geometric.inl:68-79 */                                                  \
    static inline v3_##T cross_##T(v3_##T a, v3_##T b)                                         \
    {                                                                                           \
        v3_##T r = {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};       \
        return r;                                                                               \
    }                                                                                           \
    /* glm::dot, func_geometric.inl:48-55: (x + y) + z */                                       \
    static inline T dot_##T(v3_##T a, v3_##T b)                                                 \
    {                                                                                           \
        T x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;                                          \
        return (x + y) + z;                                                                     \
    }                                                                                           \
    static inline int project6_##T(v3_##T ax, v3_##T p1, v3_##T p2, v3_##T p3, v3_##T q1,      \
                                   v3_##T q2, v3_##T q3)                                        \
    {                                                                                           \
        T P1 = dot_##T(ax, p1), P2 = dot_##T(ax, p2), P3 = dot_##T(ax, p3);                     \
        T Q1 = dot_##T(ax, q1), Q2 = dot_##T(ax, q2), Q3 = dot_##T(ax, q3);                     \
        T mx1 = FMAX(FMAX(P1, P2), P3), mn1 = FMIN(FMIN(P1, P2), P3);                           \
        T mx2 = FMAX(FMAX(Q1, Q2), Q3), mn2 = FMIN(FMIN(Q1, Q2), Q3);                           \
        if (mn1 > mx2) return 0;                                                                \
        if (mn2 > mx1) return 0;                                                                \
        return 1;                                                                               \
    }                                                                                           \
    static int NAME(v3_##T P1, v3_##T P2, v3_##T P3, v3_##T Q1, v3_##T Q2, v3_##T Q3)          \
    {                                                                                           \
        v3_##T p1 = {0, 0, 0};                                                                  \
        v3_##T p2 = sub_##T(P2, P1), p3 = sub_##T(P3, P1);                                      \
        v3_##T q1 = sub_##T(Q1, P1), q2 = sub_##T(Q2, P1), q3 = sub_##T(Q3, P1);                \
        v3_##T e1 = sub_##T(p2, p1), e2 = sub_##T(p3, p2), e3 = sub_##T(p1, p3);                \
        v3_##T f1 = sub_##T(q2, q1), f2 = sub_##T(q3, q2), f3 = sub_##T(q1, q3);                \
        v3_##T n1 = cross_##T(e1, e2), m1 = cross_##T(f1, f2);                                  \
        v3_##T g1 = cross_##T(e1, n1), g2 = cross_##T(e2, n1), g3 = cross_##T(e3, n1);          \
        v3_##T h1 = cross_##T(f1, m1), h2 = cross_##T(f2, m1), h3 = cross_##T(f3, m1);          \
        v3_##T ax[17] = {n1,                                                                    \
                         m1,                                                                    \
                         cross_##T(e1, f1),                                                     \
                         cross_##T(e1, f2),                                                     \
                         cross_##T(e1, f3),                                                     \
                         cross_##T(e2, f1),                                                     \
                         cross_##T(e2, f2),                                                     \
                         cross_##T(e2, f3),                                                     \
                         cross_##T(e3, f1),                                                     \
                         cross_##T(e3, f2),                                                     \
                         cross_##T(e3, f3),                                                     \
                         g1,                                                                    \
                         g2,                                                                    \
                         g3,                                                                    \
                         h1,                                                                    \
                         h2,                                                                    \
                         h3};                                                                   \
        for (int k = 0; k < 17; k++)                                                            \
            if (!project6_##T(ax[k], p1, p2, p3, q1, q2, q3)) return 0;                         \
        return 1;                                                                               \
    }

DEFINE_SAT(float, sat_f32, fmaxf, fminf)
DEFINE_SAT(double, sat_f64, fmax, fmin)

/* P = 9 floats (P1,P2,P3), Q = 9 floats (Q1,Q2,Q3) */
EXPORT int orc_tri_intersect(const float* P, const float* Q)
{
    v3_float a = {P[0], P[1], P[2]}, b = {P[3], P[4], P[5]}, c = {P[6], P[7], P[8]};
    v3_float d = {Q[0], Q[1], Q[2]}, e = {Q[3], Q[4], Q[5]}, f = {Q[6], Q[7], Q[8]};
    return sat_f32(a, b, c, d, e, f);
}

EXPORT int orc_tri_intersect_f64(const float* P, const float* Q)
{
    v3_double a = {P[0], P[1], P[2]}, b = {P[3], P[4], P[5]}, c = {P[6], P[7], P[8]};
    v3_double d = {Q[0], Q[1], Q[2]}, e = {Q[3], Q[4], Q[5]}, f = {Q[6], Q[7], Q[8]};
    return sat_f64(a, b, c, d, e, f);
}

/* ------------------------------------------------------------------------ */
/* SimpleCollide::detect, simpleCollide.cpp:47-161                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t* v; /* 4 per pair: meshA, meshB, origTriA, origTriB */
    uint64_t n, cap;
} pairbuf;

static void pb_push(pairbuf* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    if (p->n == p->cap) {
        p->cap = p->cap ? p->cap * 2 : 4096;
        p->v = (uint32_t*)realloc(p->v, sizeof(uint32_t) * 4u * p->cap);
    }
    uint32_t* q = p->v + 4u * p->n++;
    q[0] = a;
    q[1] = b;
    q[2] = c;
    q[3] = d;
}

typedef struct {
    int32_t na, nb;
    uint32_t ma, mb;
} sentry;

typedef struct {
    sentry* v;
    uint64_t n, cap;
} sstack;

static void ss_push(sstack* s, int32_t na, uint32_t ma, int32_t nb, uint32_t mb)
{
    if (s->n == s->cap) {
        s->cap = s->cap ? s->cap * 2 : 4096;
        s->v = (sentry*)realloc(s->v, sizeof(sentry) * s->cap);
    }
    sentry e = {na, nb, ma, mb};
    s->v[s->n++] = e;
}

static int tri_share_vertex(const uint32_t* a, const uint32_t* b)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (a[i] == b[j]) return 1;
    return 0;
}

static inline void tri_coords(const orc_tree* t, int32_t tri, v3_float* v)
{
    const uint32_t* f = t->faces + 3u * (uint32_t)tri;
    for (int k = 0; k < 3; k++) {
        const float* p = t->xyz + 3u * f[k];
        v[k].x = p[0];
        v[k].y = p[1];
        v[k].z = p[2];
    }
}

static inline void tri_coords_d(const orc_tree* t, int32_t tri, v3_double* v)
{
    const uint32_t* f = t->faces + 3u * (uint32_t)tri;
    for (int k = 0; k < 3; k++) {
        const float* p = t->xyz + 3u * f[k];
        v[k].x = p[0];
        v[k].y = p[1];
        v[k].z = p[2];
    }
}

/* flags for orc_detect */
#define ORC_SELF 1u   /* also test every tree against itself (no counterpart in
                         the reference; pairs sharing a vertex are skipped, as in
                         gProximity's disabled skipAdjacent kernel,
                         cuda_intersect_tritri.h:459-465) */
#define ORC_FP64 2u   /* narrow phase in double (FP64 oracle) */
#define ORC_BROAD 4u  /* emit every overlapping leaf pair (no narrow phase) */

/* Narrow phase for one leaf pair. P comes from tree `ta` (the lower mesh
 * id, simpleCollide.cpp:78-86). For a pair inside one mesh (self collision,
 * no reference counterpart) P is the face with the smaller original id, so
 * the result does not depend on the tree's leaf order. Emits the pair in
 * original ids. */
static void narrow_pair(pairbuf* pb, const orc_tree* ta, uint32_t ma, int32_t triA, const orc_tree* tb,
                        uint32_t mb, int32_t triB, uint32_t flags)
{
    uint32_t oa = ta->orig[triA], ob = tb->orig[triB];
    if (ma == mb && oa > ob) {
        int32_t t = triA;
        triA = triB;
        triB = t;
        uint32_t u = oa;
        oa = ob;
        ob = u;
    }
    int hit;
    if (flags & ORC_BROAD)
        hit = 1;
    else if (flags & ORC_FP64) {
        v3_double P[3], Q[3];
        tri_coords_d(ta, triA, P);
        tri_coords_d(tb, triB, Q);
        hit = sat_f64(P[0], P[1], P[2], Q[0], Q[1], Q[2]);
    } else {
        v3_float P[3], Q[3];
        tri_coords(ta, triA, P);
        tri_coords(tb, triB, Q);
        hit = sat_f32(P[0], P[1], P[2], Q[0], Q[1], Q[2]);
    }
    if (hit) pb_push(pb, ma, mb, oa, ob);
}

/* Run the reference traversal over n trees; returns the number of pairs and a
 * malloc'd array of 4*n uint32 (free with orc_free). Pairs are (i, j, a, b)
 * with i < j (or i == j and a < b for self pairs) in original face ids. The
 * order is the traversal's order; callers sort. stats[0] = BVTT nodes popped,
 * stats[1] = leaf pairs that reached the narrow phase. */
EXPORT uint64_t orc_detect(orc_tree** trees, uint32_t n, uint32_t flags, uint32_t** out,
                           uint64_t* stats)
{
    pairbuf pb = {0, 0, 0};
    sstack st = {0, 0, 0};
    uint64_t popped = 0, leafpairs = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (flags & ORC_SELF) ss_push(&st, trees[i]->root, i, trees[i]->root, i);
        for (uint32_t j = i + 1; j < n; j++) ss_push(&st, trees[i]->root, i, trees[j]->root, j);
    }
    while (st.n) {
        sentry e = st.v[--st.n];
        popped++;
        const orc_tree* ta = trees[e.ma];
        const orc_tree* tb = trees[e.mb];
        const onode* A = &ta->nodes[e.na];
        const onode* B = &tb->nodes[e.nb];
        const int self = (e.ma == e.mb && e.na == e.nb);
        if (!self && !box_overlap(A->b, B->b)) continue;
        if (A->tri != -1 && B->tri != -1) {
            if (self) continue; /* a triangle against itself */
            if (e.ma == e.mb && tri_share_vertex(ta->faces + 3u * (uint32_t)A->tri,
                                                 tb->faces + 3u * (uint32_t)B->tri))
                continue;
            leafpairs++;
            narrow_pair(&pb, ta, e.ma, A->tri, tb, e.mb, B->tri, flags);
            continue;
        }
        int32_t ca[2], cb[2];
        int nca = 0, ncb = 0;
        if (A->tri != -1)
            ca[nca++] = e.na;
        else {
            ca[nca++] = A->left;
            if (A->right != -1) ca[nca++] = A->right;
        }
        if (B->tri != -1)
            cb[ncb++] = e.nb;
        else {
            cb[ncb++] = B->left;
            if (B->right != -1) cb[ncb++] = B->right;
        }
        if (self) {
            /* a subtree against itself: unordered child pairs (x, y), x <= y */
            for (int x = 0; x < nca; x++)
                for (int y = x; y < nca; y++) ss_push(&st, ca[x], e.ma, ca[y], e.mb);
        } else {
            for (int x = 0; x < nca; x++)
                for (int y = 0; y < ncb; y++) ss_push(&st, ca[x], e.ma, cb[y], e.mb);
        }
    }
    free(st.v);
    if (stats) {
        stats[0] = popped;
        stats[1] = leafpairs;
    }
    *out = pb.v;
    return pb.n;
}

/* Brute force over every face pair (tests only: closes the broad phase). */
EXPORT uint64_t orc_brute_force(orc_tree** trees, uint32_t n, uint32_t flags, uint32_t** out)
{
    pairbuf pb = {0, 0, 0};
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = i; j < n; j++) {
            if (i == j && !(flags & ORC_SELF)) continue;
            const orc_tree* ta = trees[i];
            const orc_tree* tb = trees[j];
            for (uint32_t a = 0; a < ta->F; a++) {
                float ba[6];
                leaf_box(ta->faces + 3u * a, ta->xyz, ba);
                for (uint32_t b = (i == j ? a + 1 : 0); b < tb->F; b++) {
                    float bb[6];
                    leaf_box(tb->faces + 3u * b, tb->xyz, bb);
                    if (!box_overlap(ba, bb)) continue;
                    if (i == j && tri_share_vertex(ta->faces + 3u * a, tb->faces + 3u * b)) continue;
                    narrow_pair(&pb, ta, i, (int32_t)a, tb, j, (int32_t)b, flags);
                }
            }
        }
    *out = pb.v;
    return pb.n;
}

EXPORT void orc_free(void* p) { free(p); }
