// oibvh_collide.hip -- gfx950 broad + narrow phase for Scene::detectCollision.
//
// The reference relaunches traversal_kernel once per BVTT level with a host
// round trip for the front size (scene.cu:272-311, collide.cu:76-246) and then
// runs triangle_intersect_kernel over the candidate list (collide.cu:248-300).
// Here the steps are launched back to back with no host round trip:
//   k_seed        zeroes the level counts and detect counters;
//   k_bvtt_step   one BVTT step, grid-stride over the current front (its
//                 length is read from device memory): both sides of an
//                 overlapping pair descend up to E levels (a leaf stands for
//                 itself, collide.cu:165-176) and every pair of the two
//                 descendant runs is tested for AABB overlap (collide.cu:
//                 12-17); overlapping leaf pairs go to the candidate list,
//                 other overlapping pairs to the next front. Outputs are
//                 compacted with a block prefix sum and one device-scope
//                 fetch_add per block;
//   k_narrow      17-axis SAT (utils.cpp:71-169 / cuda_intersect_tritri.h:
//                 350-434) over the candidate list, hits appended as
//                 (meshA, meshB, origTriA, origTriB) in original face ids.
// Self mode (no reference counterpart) also walks each mesh against itself:
// a node paired with itself expands to the unordered pairs of its children;
// leaf pairs sharing a vertex index are dropped before the SAT.
//
// BVTT entries are two 32-bit implicit indices in the scene's concatenated
// implicit space (mesh m spans [ibase_m, ibase_m + 2^(L_m+1) - 1)).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "oibvh_common.h"

#define TB 256
#define MESH_LDS 16
#define INLINE_MESHES 4

typedef unsigned long long u64;

// 0, computed from v by an instruction the compiler cannot see through: an
// address offset by it is formed only once v is available
__device__ __forceinline__ uint32_t opaque_zero(float v)
{
    uint32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z) : "v"(v));
    return z;
}

__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* s_wsum, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = oi_wave_scan_incl(v);
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t w0 = s_wsum[0], w1 = s_wsum[1], w2 = s_wsum[2], w3 = s_wsum[3];
    uint32_t before = (wave > 0 ? w0 : 0u) + (wave > 1 ? w1 : 0u) + (wave > 2 ? w2 : 0u);
    total = w0 + w1 + w2 + w3;
    __syncthreads();
    return before + x - v;
}

__device__ __forceinline__ uint32_t add_u32(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void or_u32(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// SAT, FP32 (utils.cpp:71-169; glm cross/dot op order, no contraction)
// ---------------------------------------------------------------------------
struct v3f {
    float x, y, z;
};
struct v3d {
    double x, y, z;
};

template <class V>
__device__ __forceinline__ V vsub(V a, V b)
{
    return V{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class V>
__device__ __forceinline__ V vcross(V a, V b)
{
    return V{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T, class V>
__device__ __forceinline__ T vdot(V a, V b)
{
    const T x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return (x + y) + z;
}

// the 17 axes in the reference's test order (utils.cpp:146-168)
enum { AX_N1, AX_M1, AX_E11, AX_E12, AX_E13, AX_E21, AX_E22, AX_E23, AX_E31, AX_E32, AX_E33,
       AX_G1, AX_G2, AX_G3, AX_H1, AX_H2, AX_H3 };

template <class T, class V>
struct SatFrame {
    V p2, p3, q1, q2, q3;
    V e1, e2, e3, f1, f2, f3;
    V n1, m1;
    __device__ __forceinline__ void init(const V* P, const V* Q)
    {
        const V p1 = V{0, 0, 0};
        p2 = vsub(P[1], P[0]);
        p3 = vsub(P[2], P[0]);
        q1 = vsub(Q[0], P[0]);
        q2 = vsub(Q[1], P[0]);
        q3 = vsub(Q[2], P[0]);
        e1 = vsub(p2, p1);
        e2 = vsub(p3, p2);
        e3 = vsub(p1, p3);
        f1 = vsub(q2, q1);
        f2 = vsub(q3, q2);
        f3 = vsub(q1, q3);
        n1 = vcross(e1, e2);
        m1 = vcross(f1, f2);
    }
    __device__ __forceinline__ V axis(int k) const
    {
        switch (k) {
        case AX_N1: return n1;
        case AX_M1: return m1;
        case AX_E11: return vcross(e1, f1);
        case AX_E12: return vcross(e1, f2);
        case AX_E13: return vcross(e1, f3);
        case AX_E21: return vcross(e2, f1);
        case AX_E22: return vcross(e2, f2);
        case AX_E23: return vcross(e2, f3);
        case AX_E31: return vcross(e3, f1);
        case AX_E32: return vcross(e3, f2);
        case AX_E33: return vcross(e3, f3);
        case AX_G1: return vcross(e1, n1);
        case AX_G2: return vcross(e2, n1);
        case AX_G3: return vcross(e3, n1);
        case AX_H1: return vcross(f1, m1);
        case AX_H2: return vcross(f2, m1);
        default: return vcross(f3, m1);
        }
    }
};

__device__ __forceinline__ float fmx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float fmn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double fmx(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ double fmn(double a, double b) { return fmin(a, b); }

// project6 (utils.cpp:71-95): 1 if the axis does NOT separate
template <class T, class V>
__device__ __forceinline__ bool project6(const SatFrame<T, V>& s, V ax)
{
    const V p1 = V{0, 0, 0};
    const T P1 = vdot<T>(ax, p1), P2 = vdot<T>(ax, s.p2), P3 = vdot<T>(ax, s.p3);
    const T Q1 = vdot<T>(ax, s.q1), Q2 = vdot<T>(ax, s.q2), Q3 = vdot<T>(ax, s.q3);
    const T mx1 = fmx(fmx(P1, P2), P3), mn1 = fmn(fmn(P1, P2), P3);
    const T mx2 = fmx(fmx(Q1, Q2), Q3), mn2 = fmn(fmn(Q1, Q2), Q3);
    if (mn1 > mx2) return false;
    if (mn2 > mx1) return false;
    return true;
}

__device__ bool sat_f32(const v3f* P, const v3f* Q)
{
    SatFrame<float, v3f> s;
    s.init(P, Q);
    for (int k = 0; k < 17; k++)
        if (!project6<float, v3f>(s, s.axis(k))) return false;
    return true;
}

// ---------------------------------------------------------------------------
// FP32 filter with FP64 re-check (precision form).
// Every FP32 SAT quantity z has a magnitude S(z) >= |z| built by the same
// expression over |operands| (sums of magnitudes for +/-, products for *),
// and |fl(z) - z| <= c * 2^-24 * S(z) (+ O(2^-48)) at rounding depth c, c
// counted as roundings along the chain (a product adds its operands' depths
// plus one, a sum takes the larger plus one): the differences p = P - P1
// have c = 1, edges 2, n1 = e x e 6, g = e x n1 10, the products of a g/h
// projection 12 and its two additions 14, the deepest SAT quantity.
// Here the magnitudes are bounded per component through three scalars per
// triangle pair: M = (Mx, My, Mz), the largest |x|, |y| and |z| components of
// the five differences p2, p3, q1, q2, q3. Every magnitude is a sum of
// products of leaf magnitudes with non-negative coefficients, so it grows
// with them: edges have magnitudes <= 2M (componentwise), n1 <= 4C, m1 and
// e x f <= 8C with C = (My Mz, Mz Mx, Mx My), g <= 8D and h <= 16D with
// D = M x C (magnitude cross product) = (Mx (My^2 + Mz^2), My (Mz^2 + Mx^2),
// Mz (Mx^2 + My^2)); a projection pair S_P + S_Q <= 2 (axis bound . M):
// 24 Mx My Mz (n1), 48 Mx My Mz (m1, e x f), 32 W (g), 64 W (h), with
// W = Mx^2 My^2 + My^2 Mz^2 + Mz^2 Mx^2. A flat axis (a plane perpendicular
// to it) keeps the n1 bound near zero. A comparison mn1 > mx2 is decided in
// FP32 only when |mn1 - mx2| exceeds T = 32 * 2^-24 * (that bound) + 2^-100,
// which covers the FP32 error of both sides (c <= 14), the FP64 oracle's own
// error, the rounding of the subtraction, of the bound and of T; the FP32
// decision then equals the exact one and the FP64 one. A pair with some
// undecided axis and no decided separating axis is evaluated again, every
// axis in double exactly as the FP64 oracle evaluates it (k_narrow's second
// pass). Non-finite or NaN values never decide.
// ---------------------------------------------------------------------------
enum { SAT_SEPARATED = 0, SAT_HIT = 1, SAT_UNDECIDED = 2 };

__device__ __forceinline__ v3f vabs(v3f a) { return v3f{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ v3f vmaxc(v3f a, v3f b) { return v3f{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

__device__ int sat_filter(const v3f* P, const v3f* Q)
{
    SatFrame<float, v3f> s;
    s.init(P, Q);
    const v3f M = vmaxc(vmaxc(vmaxc(vabs(s.p2), vabs(s.p3)), vmaxc(vabs(s.q1), vabs(s.q2))), vabs(s.q3));
    const float xyz = M.x * M.y * M.z;
    const float xx = M.x * M.x, yy = M.y * M.y, zz = M.z * M.z;
    const float W = (xx * yy + yy * zz) + zz * xx;
    constexpr float U = 32.0f * 5.9604645e-8f, TINY = 7.888609e-31f;
    const float t_n1 = U * 24.0f * xyz + TINY, t_ef = U * 48.0f * xyz + TINY;
    const float t_g = U * 32.0f * W + TINY, t_h = U * 64.0f * W + TINY;
    bool undecided = false;
    for (int k = 0; k < 17; k++) {
        const v3f ax = s.axis(k);
        const v3f p1 = v3f{0, 0, 0};
        const float P1 = vdot<float>(ax, p1), P2 = vdot<float>(ax, s.p2), P3 = vdot<float>(ax, s.p3);
        const float Q1 = vdot<float>(ax, s.q1), Q2 = vdot<float>(ax, s.q2), Q3 = vdot<float>(ax, s.q3);
        const float mx1 = fmaxf(fmaxf(P1, P2), P3), mn1 = fminf(fminf(P1, P2), P3);
        const float mx2 = fmaxf(fmaxf(Q1, Q2), Q3), mn2 = fminf(fminf(Q1, Q2), Q3);
        const float T = k == AX_N1 ? t_n1 : k <= AX_E33 ? t_ef : k <= AX_G3 ? t_g : t_h;
        const bool finite = (T < __builtin_inff());
        // separation on this axis: mn1 > mx2 or mn2 > mx1
        if (finite && (mn1 - mx2 > T || mn2 - mx1 > T)) return SAT_SEPARATED;
        if (!(finite && mx2 - mn1 > T && mx1 - mn2 > T)) undecided = true;
    }
    return undecided ? SAT_UNDECIDED : SAT_HIT;
}

// the FP64 oracle's SAT over the widened FP32 inputs (every axis until one
// separates); counts the axes evaluated
__device__ bool sat_f64_counted(const v3f* P, const v3f* Q, uint32_t& n_fp64)
{
    v3d Pd[3], Qd[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Pd[i] = v3d{(double)P[i].x, (double)P[i].y, (double)P[i].z};
        Qd[i] = v3d{(double)Q[i].x, (double)Q[i].y, (double)Q[i].z};
    }
    SatFrame<double, v3d> sd;
    sd.init(Pd, Qd);
    for (int k = 0; k < 17; k++) {
        n_fp64++;
        if (!project6<double, v3d>(sd, sd.axis(k))) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// scene tables
// ---------------------------------------------------------------------------
struct MeshTab {
    const oi_mesh_desc* meshes;       // device table (scenes of more than INLINE_MESHES meshes)
    uint32_t n_meshes;
    oi_mesh_desc inl[INLINE_MESHES];  // small scenes: the table travels in the kernel arguments
};

struct MeshRef {
    const float* nodes;
    const uint32_t* sfaces;
    const float* pos;
    const uint32_t* perm;
    const float* xf;
    uint32_t L, vl, ibase;
};

// copy the first MESH_LDS mesh descriptors to LDS
__device__ __forceinline__ void load_mesh_tab(const MeshTab& T, oi_mesh_desc* s_desc, uint32_t* s_ibase)
{
    const uint32_t tid = threadIdx.x;
    if (tid < MESH_LDS && tid < T.n_meshes) {
        if (T.n_meshes <= INLINE_MESHES) {
            // constant indices keep the argument block out of scratch memory
            if (tid == 0) s_desc[0] = T.inl[0];
            if (tid == 1) s_desc[1] = T.inl[1];
            if (tid == 2) s_desc[2] = T.inl[2];
            if (tid == 3) s_desc[3] = T.inl[3];
        } else {
            s_desc[tid] = T.meshes[tid];
        }
    }
    __syncthreads();
    if (tid < MESH_LDS && tid < T.n_meshes) s_ibase[tid] = s_desc[tid].ibase;
    __syncthreads();
}

__device__ __forceinline__ uint32_t find_mesh(const MeshTab& T, const uint32_t* s_ibase, uint32_t g)
{
    uint32_t lo = 0, hi = T.n_meshes - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t b = (mid < MESH_LDS) ? s_ibase[mid] : T.meshes[mid].ibase;
        if (b <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ MeshRef mesh_ref(const MeshTab& T, const oi_mesh_desc* s_desc, uint32_t m)
{
    const oi_mesh_desc& d = (m < MESH_LDS) ? s_desc[m] : T.meshes[m];
    return MeshRef{d.nodes, d.sfaces, d.pos, d.perm, d.xf, d.L, d.vl, d.ibase};
}

// A batch's mesh table: every mesh's implicit range starts at a multiple of
// 2^shift and is padded to one, so the mesh of global implicit index g is
// chunk_mesh[g >> shift] (one lookup instead of a search over the meshes).
struct ChunkTab {
    const oi_mesh_desc* meshes;
    const uint32_t* chunk_mesh;
    uint32_t shift;
};

// uniform access to both table kinds: per-workgroup set-up, then lookups
__device__ __forceinline__ void tab_init(const MeshTab& T, oi_mesh_desc* s_desc, uint32_t* s_ibase)
{
    load_mesh_tab(T, s_desc, s_ibase);
}
__device__ __forceinline__ void tab_init(const ChunkTab&, oi_mesh_desc*, uint32_t*) {}

__device__ __forceinline__ MeshRef tab_ref(const MeshTab& T, const oi_mesh_desc* s_desc, const uint32_t* s_ibase,
                                           uint32_t g, uint32_t& m)
{
    m = find_mesh(T, s_ibase, g);
    return mesh_ref(T, s_desc, m);
}
__device__ __forceinline__ MeshRef tab_ref(const ChunkTab& T, const oi_mesh_desc*, const uint32_t*, uint32_t g,
                                           uint32_t& m)
{
    m = T.chunk_mesh[g >> T.shift];
    const oi_mesh_desc& d = T.meshes[m];
    return MeshRef{d.nodes, d.sfaces, d.pos, d.perm, d.xf, d.L, d.vl, d.ibase};
}

// 16 zeros, read in place of a model matrix by meshes that have none
__device__ const float k_no_xf[16] = {};

// a vertex through the mesh's model matrix, if any
__device__ __forceinline__ v3f vertex(const float* pos, const float* xf, uint32_t v)
{
    float p[3];
    oi_ld3(pos, v, p);
    if (!xf) return v3f{p[0], p[1], p[2]};
    float m[16];
#pragma unroll
    for (int k = 0; k < 16; k++) m[k] = oi_g(xf)[k];
    float o[3];
    oi_xform(m, p[0], p[1], p[2], o);
    return v3f{o[0], o[1], o[2]};
}

__device__ __forceinline__ bool overlap6(const float* a, const float* b)
{
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
}

__device__ __forceinline__ void load_box(const float* nodes, uint32_t r, float* o)
{
    const OI_GLOBAL float* p = oi_g(nodes) + 6ull * r;
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = p[k];
}

// Expansion of one side: the descendants `d` levels below implicit node i
// (d = min(expand, L - level)); start = 2^d (i + 1) - 1, count clipped at
// the level's right-most real node (collide.cu:165-176).
struct Side {
    uint32_t start, count;
};

__device__ __forceinline__ Side expand_side(uint32_t i, uint32_t lvl, uint32_t L, uint32_t vl, uint32_t e)
{
    if (lvl == L) return Side{i, 1u};
    const uint32_t d = (L - lvl < e) ? (L - lvl) : e;
    const uint32_t nl = lvl + d;
    const uint32_t st = ((i + 1u) << d) - 1u;
    const uint32_t last = oi_lvl_last_implicit(nl, L, vl);
    const uint32_t c = last - st + 1u;
    return Side{st, c < (1u << d) ? c : (1u << d)};
}

// ---------------------------------------------------------------------------
// k_bvtt_step: one step of the BVTT walk, testing on expansion
//
// A front entry is a node pair whose boxes are known to overlap (or, in self
// mode, a node paired with itself). A step takes both sides up to E levels
// down (a leaf stands for itself; collide.cu:165-176) and tests every pair of
// the two descendant runs, which are contiguous in the real node order, so
// each side is one coalesced run of at most 2^E boxes. Overlapping pairs go
// to the next front, or to the candidate list when both sides are leaves.
// Only overlapping pairs are ever written, and parents contain their children
// exactly (compare-select merges), so the candidate set equals that of a walk
// which tests every level (simpleCollide.cpp:88-143) for finite coordinates.
// The first step also tests the entry pairs themselves (the initial front is
// every pair of entry-level nodes, scene.cu:192-223).
// ---------------------------------------------------------------------------
template <class Tab>
struct StepArgs {
    Tab tab;
    const uint2* front_in;
    const uint32_t* count_in;   // length of front_in; null: n_in
    uint32_t n_in;
    uint32_t cap_in;            // entries front_in can hold
    uint2* front_out;
    uint32_t* count_out;        // zeroed by the seed
    uint32_t cap_out;
    uint2* cands;
    uint32_t* cand_count;
    uint32_t cand_cap;
    uint32_t* error;
    unsigned long long* tests;  // box pair tests (statistics)
    unsigned long long* adj;    // self-mode leaf pairs dropped as adjacent
    uint4* hits;                // fused narrow phase: hits and their count
    uint32_t* n_hits;
    uint32_t hits_cap;
    uint2* defer;               // fused FP32 filter: undecided leaf pairs for the FP64 pass
    uint32_t* defer_count;
    uint32_t defer_cap;
    uint32_t implicit_end;
};

// A run of n >= 1 real nodes from real index r0. Slots k >= n repeat the
// last node (their results are never used), so every load is unconditional
// and the run's loads are issued back to back.
template <int E>
__device__ __forceinline__ void load_run(const float* nodes, uint32_t r0, uint32_t n, float (&b)[1 << E][6])
{
#pragma unroll
    for (int k = 0; k < (1 << E); k++) load_box(nodes, r0 + min((uint32_t)k, n - 1u), b[k]);
}

// faces of leaf run [leaf0, leaf0 + n), n >= 1 (leaf positions); slots
// k >= n repeat the last face as load_run does
template <int E>
__device__ __forceinline__ void load_faces(const uint32_t* sfaces, uint32_t leaf0, uint32_t n,
                                           uint32_t (&f)[1 << E][3])
{
#pragma unroll
    for (int k = 0; k < (1 << E); k++) oi_ld3(sfaces, leaf0 + min((uint32_t)k, n - 1u), f[k]);
}

__device__ __forceinline__ bool share_vertex(const uint32_t* a, const uint32_t* b)
{
    bool s = false;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) s |= (a[i] == b[j]);
    return s;
}

template <bool FP64, class Tab>
__device__ __forceinline__ void narrow_one(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                           uint2 c, bool& hit, uint4& out, uint32_t& n_fp64, uint32_t& n_adj);
template <class Tab>
__device__ __forceinline__ void narrow_one_filter(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                                  uint2 c, bool& hit, bool& defer, uint4& out, uint32_t& n_adj);

// position of the k-th set bit of m (k < popcount(m))
__device__ __forceinline__ uint32_t select_bit(u64 m, uint32_t k)
{
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << w) - 1ull));
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += w;
        }
    }
    return pos;
}

// bits x * K + y of the K x K child pairs (K = 2^E) with y > x, and y == x
template <int E>
struct PairMasks {
    static constexpr int K = 1 << E;
    static constexpr u64 make(bool diagonal)
    {
        u64 m = 0;
        for (int x = 0; x < K; x++)
            for (int y = 0; y < K; y++)
                if (diagonal ? (y == x) : (y > x)) m |= 1ull << (x * K + y);
        return m;
    }
    static constexpr u64 upper = make(false);
    static constexpr u64 diag = make(true);
};

// One BVTT entry (a pair of implicit indices): with FIRST the pair itself is
// tested first (an initial-front entry); then both sides descend up to E
// levels and every pair of the two descendant runs is tested. Bit x * K + y
// of `mask` keeps pair (sa.start + x, sb.start + y); `leafout` says the kept
// pairs are leaf pairs. In self mode leaf pairs of one mesh that share a
// vertex are dropped here.
template <int E, bool FIRST, class Tab>
__device__ __forceinline__ void entry_expand(const Tab& T, const oi_mesh_desc* s_desc, const uint32_t* s_ibase,
                                             uint2 e, u64& mask, bool& leafout, Side& sa, Side& sb,
                                             uint32_t& baseA, uint32_t& baseB, uint32_t& n_tests, uint32_t& n_adj)
{
    constexpr int K = 1 << E;
    uint32_t ma, mb;
    const MeshRef ra = tab_ref(T, s_desc, s_ibase, e.x, ma);
    const MeshRef rb = tab_ref(T, s_desc, s_ibase, e.y, mb);
    baseA = ra.ibase;
    baseB = rb.ibase;
    const uint32_t ia = e.x - ra.ibase, ib = e.y - rb.ibase;
    const uint32_t la = oi_ilog2(ia + 1u), lb = oi_ilog2(ib + 1u);
    const bool selfnode = (e.x == e.y);
    const bool leafleaf = (la == ra.L && lb == rb.L);
    bool go = !(selfnode && leafleaf);  // a triangle against itself
    if (FIRST && go && !selfnode && !leafleaf) {
        float ba[6], bb[6];
        load_box(ra.nodes, ia - oi_virt_above(la, ra.L, ra.vl), ba);
        load_box(rb.nodes, ib - oi_virt_above(lb, rb.L, rb.vl), bb);
        go = overlap6(ba, bb);
        n_tests++;
    }
    if (!go) return;
    sa = expand_side(ia, la, ra.L, ra.vl, E);
    sb = selfnode ? sa : expand_side(ib, lb, rb.L, rb.vl, E);
    const uint32_t ca = oi_ilog2(sa.start + 1u), cb = oi_ilog2(sb.start + 1u);
    leafout = (ca == ra.L && cb == rb.L);
    // both runs unconditionally (a node paired with itself reads its run twice)
    float bxa[K][6], bxb[K][6];
    load_run<E>(ra.nodes, sa.start - oi_virt_above(ca, ra.L, ra.vl), sa.count, bxa);
    load_run<E>(rb.nodes, sb.start - oi_virt_above(cb, rb.L, rb.vl), sb.count, bxb);
    // every slot pair is tested with no branch (slots past a run's end repeat
    // its last node), then masked to the pairs that exist: x < sa.count,
    // y < sb.count, and for a node paired with itself the unordered child
    // pairs, y > x tested and y == x kept unless they are leaves
    u64 hit = 0;
#pragma unroll
    for (int x = 0; x < K; x++)
#pragma unroll
        for (int y = 0; y < K; y++)
            if (overlap6(bxa[x], bxb[y])) hit |= 1ull << (x * K + y);
    u64 rows = 0;
#pragma unroll
    for (int x = 0; x < K; x++)
        if ((uint32_t)x < sa.count) rows |= 1ull << (x * K);
    const u64 exist = rows * ((1ull << sb.count) - 1ull);  // sb.count <= K: no carry between rows
    const u64 tested = selfnode ? (exist & PairMasks<E>::upper) : exist;
    n_tests += (uint32_t)__popcll(tested);
    mask = hit & tested;
    if (selfnode && !leafout) mask |= exist & PairMasks<E>::diag;
    // inside one mesh (self mode) leaf pairs sharing a vertex are not
    // collisions: drop them before they become candidates
    if (leafout && ma == mb && mask) {
        uint32_t fa[K][3], fb[K][3];
        const uint32_t leaf0 = (1u << ra.L) - 1u;
        load_faces<E>(ra.sfaces, sa.start - leaf0, sa.count, fa);
        load_faces<E>(ra.sfaces, sb.start - leaf0, sb.count, fb);
#pragma unroll
        for (int x = 0; x < K; x++)
#pragma unroll
            for (int y = 0; y < K; y++) {
                const u64 bit = 1ull << (x * K + y);
                if ((mask & bit) && share_vertex(fa[x], fb[y])) {
                    mask &= ~bit;
                    n_adj++;
                }
            }
    }
}

// LDS the workgroup needs to test its leaf pairs cooperatively
struct FusedLds {
    u64 cmask[TB];
    uint32_t coff[TB], ga[TB], gb[TB];
    uint32_t hbase, dbase;
};

// The workgroup's C leaf pairs (thread t owns the set bits of cmask[t],
// numbered from coff[t]) through the FP32 SAT (FILTER: the FP32 filter,
// undecided pairs to the defer list for the FP64 pass): pair c is found by a
// binary search over the offsets and a select in the owner's mask; hits (and
// deferred pairs) are appended with one fetch_add each per round.
template <int E, bool FILTER, class Tab>
__device__ __forceinline__ void fused_narrow(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                             FusedLds& F, uint32_t C, uint32_t* s_wsum, uint4* hits,
                                             uint32_t* n_hits, uint32_t hits_cap, uint2* defer,
                                             uint32_t* defer_count, uint32_t defer_cap, uint32_t* error,
                                             uint32_t& n_fp64, uint32_t& n_adj)
{
    constexpr int K = 1 << E;
    const uint32_t tid = threadIdx.x;
    for (uint32_t c0 = 0; c0 < C; c0 += TB) {
        const uint32_t c = c0 + tid;
        bool hit = false, dfr = false;
        uint2 pr = make_uint2(0, 0);
        uint4 h = make_uint4(0, 0, 0, 0);
        if (c < C) {
            uint32_t lo = 0, hi = TB - 1;  // last thread whose offset is <= c
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (F.coff[mid] <= c)
                    lo = mid;
                else
                    hi = mid - 1u;
            }
            const uint32_t bit = select_bit(F.cmask[lo], c - F.coff[lo]);
            pr = make_uint2(F.ga[lo] + (bit >> E), F.gb[lo] + (bit & (K - 1)));
            if (FILTER)
                narrow_one_filter<Tab>(T, s_ibase, s_desc, pr, hit, dfr, h, n_adj);
            else
                narrow_one<false, Tab>(T, s_ibase, s_desc, pr, hit, h, n_fp64, n_adj);
        }
        // one scan for both counts: hits in the high half, deferred pairs in
        // the low half (at most 256 each)
        uint32_t HD;
        const uint32_t o = block_scan((hit ? 65536u : 0u) + (dfr ? 1u : 0u), s_wsum, HD);
        const uint32_t H = HD >> 16, D = HD & 0xffffu;
        if (tid == 0) {
            F.hbase = H ? add_u32(n_hits, H) : 0u;
            if (FILTER) F.dbase = D ? add_u32(defer_count, D) : 0u;
        }
        __syncthreads();
        if (hit) {
            const uint32_t pos = F.hbase + (o >> 16);
            if (pos < hits_cap)
                hits[pos] = h;
            else
                or_u32(error, OI_ERR_HITS_OVERFLOW);
        }
        if (FILTER && dfr) {
            const uint32_t pos = F.dbase + (o & 0xffffu);
            if (pos < defer_cap)
                defer[pos] = pr;
            else
                or_u32(error, OI_ERR_CAND_OVERFLOW);
        }
    }
}

// FUSE 1: overlapping leaf pairs are tested with the FP32 SAT by the
// workgroup that found them (fused_narrow) instead of going to the candidate
// list; FUSE 2: with the FP32 filter, the pairs it leaves undecided going to
// A.defer for the FP64 pass.
// Occupancy: the step waits on its child-run loads, so resident waves matter
// more than loads in flight per wave. E = 3 holds two runs of 8 boxes and
// would take 153-161 VGPRs (3 waves per SIMD); capped at 128 (4 waves) it
// keeps the runs in registers (the fused forms spill 2-5 dwords). E = 2 fits
// 80 VGPRs (6 waves) unspilled.
template <int E, bool FIRST, int FUSE, class Tab>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(E == 3 ? 4 : E == 2 ? 6 : 1, 8)))
k_bvtt_step(StepArgs<Tab> A)
{
    constexpr bool FUSED = FUSE != 0;
    constexpr int K = 1 << E;
    __shared__ oi_mesh_desc s_desc[MESH_LDS];
    __shared__ uint32_t s_ibase[MESH_LDS];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_obase, s_cbase;
    __shared__ std::conditional_t<FUSED, FusedLds, char> s_f[1];
    // after an overflow the count keeps growing past the buffer (writes were
    // dropped and the detect will be re-run): read only what exists
    uint32_t n_in = A.count_in ? *A.count_in : A.n_in;
    if (n_in > A.cap_in) n_in = A.cap_in;
    if (blockIdx.x * TB >= n_in) return;
    tab_init(A.tab, s_desc, s_ibase);
    const uint32_t tid = threadIdx.x;
    uint32_t n_tests = 0, n_adj = 0, n_cands = 0, n_fp64 = 0;
    for (uint32_t base = blockIdx.x * TB; base < n_in; base += gridDim.x * TB) {
        const uint32_t i = base + tid;
        bool have = i < n_in;
        const uint2 e = have ? A.front_in[i] : make_uint2(0, 0);
        if (have && (e.x >= A.implicit_end || e.y >= A.implicit_end)) {
            or_u32(A.error, OI_ERR_CORRUPT);  // never dereference a bad entry
            have = false;
        }
        u64 mask = 0;          // bit (x << E) + y: pair (start_a + x, start_b + y) kept
        bool leafout = false;  // kept pairs are leaf pairs (candidates)
        Side sa = {0, 0}, sb = {0, 0};
        uint32_t baseA = 0, baseB = 0;
        if (have) entry_expand<E, FIRST, Tab>(A.tab, s_desc, s_ibase, e, mask, leafout, sa, sb, baseA, baseB,
                                              n_tests, n_adj);
        const uint32_t nout = (uint32_t)__popcll(mask);
        // one scan for both counts: front pairs in the high half, leaf pairs
        // in the low half (at most 256 x 64 each, so no carry crosses)
        uint32_t TC;
        const uint32_t oc = block_scan(leafout ? nout : (nout << 16), s_wsum, TC);
        const uint32_t T = TC >> 16, C = TC & 0xffffu, off = oc >> 16, coff = oc & 0xffffu;
        if (tid == 0) {
            s_obase = T ? add_u32(A.count_out, T) : 0u;
            if (!FUSED) s_cbase = C ? add_u32(A.cand_count, C) : 0u;
        }
        if constexpr (FUSED) {
            s_f[0].cmask[tid] = leafout ? mask : 0ull;
            s_f[0].coff[tid] = coff;
            s_f[0].ga[tid] = baseA + sa.start;
            s_f[0].gb[tid] = baseB + sb.start;
        }
        __syncthreads();
        if (nout && !(FUSED && leafout)) {
            uint2* dst = leafout ? A.cands : A.front_out;
            const uint32_t cap = leafout ? A.cand_cap : A.cap_out;
            uint32_t o = leafout ? s_cbase + coff : s_obase + off;
            if (o + nout > cap) or_u32(A.error, leafout ? OI_ERR_CAND_OVERFLOW : OI_ERR_QUEUE_OVERFLOW);
            u64 m = mask;
            while (m) {
                const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                m &= m - 1ull;
                if (o < cap) dst[o] = make_uint2(baseA + sa.start + (k >> E), baseB + sb.start + (k & (K - 1)));
                o++;
            }
        }
        if constexpr (FUSED) {
            if (C) {
                if (tid == 0) n_cands += C;
                fused_narrow<E, FUSE == 2, Tab>(A.tab, s_ibase, s_desc, s_f[0], C, s_wsum, A.hits, A.n_hits,
                                                A.hits_cap, A.defer, A.defer_count, A.defer_cap, A.error, n_fp64,
                                                n_adj);
            }
        }
        __syncthreads();
    }
    uint32_t nt, na;
    block_scan(n_tests, s_wsum, nt);
    block_scan(n_adj, s_wsum, na);
    if (tid == 0 && nt) __hip_atomic_fetch_add(A.tests, (u64)nt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0 && na) __hip_atomic_fetch_add(A.adj, (u64)na, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (FUSED && tid == 0 && n_cands) add_u32(A.cand_count, n_cands);
}

// ---------------------------------------------------------------------------
// k_bvtt_local: J steps of the walk inside workgroups. Workgroup b takes
// entries [c0, c0 + chunk) for c0 = b * chunk, (b + grid) * chunk, ... of
// its input front and walks them J levels of E (J x E tree levels) breadth
// first with its front in LDS, testing leaf pairs as it meets them (fused
// SAT, or the candidate list); the pairs kept by the J-th level go to the
// global front, where the steps take over. Two uses:
//   head: the first J steps over the initial front (test_entries: the entry
//         pairs are tested first, as the first step does), whose fronts are
//         small, in one launch instead of J;
//   tail: the last steps over a front the steps wrote (its length read from
//         device memory), down to the leaves: nothing reaches the global
//         front, and the front is read and written once instead of once per
//         step.
// A workgroup whose front outgrows its LDS queue sets `overflow_bit` and the
// detect is run again without that walk.
// ---------------------------------------------------------------------------
#define LQ 2048

template <class Tab>
struct LocalArgs {
    Tab tab;
    const uint2* front_in;
    const uint32_t* count_in;  // length of front_in; null: n_in
    uint32_t n_in;
    uint32_t cap_in;           // entries front_in can hold
    uint32_t chunk;
    uint32_t levels;           // J >= 1
    uint32_t test_entries;     // test the input pairs themselves (initial front)
    uint32_t overflow_bit;     // OI_ERR_* bit set when a workgroup's LDS front overflows
    uint2* front_out;    // pairs kept by level J
    uint32_t* count_out;
    uint32_t cap_out;
    uint2* cands;
    uint32_t* cand_count;
    uint32_t cand_cap;
    uint32_t* error;
    unsigned long long* tests;
    unsigned long long* adj;
    uint4* hits;
    uint32_t* n_hits;
    uint32_t hits_cap;
    uint32_t implicit_end;
};

template <int E, bool FUSED, class Tab>
__global__ void __launch_bounds__(TB) k_bvtt_local(LocalArgs<Tab> A)
{
    constexpr int K = 1 << E;
    __shared__ oi_mesh_desc s_desc[MESH_LDS];
    __shared__ uint32_t s_ibase[MESH_LDS];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_qn, s_obase, s_cbase;
    __shared__ uint2 s_q[2][LQ];
    __shared__ std::conditional_t<FUSED, FusedLds, char> s_f[1];
    // after an overflow the count keeps growing past the buffer (writes were
    // dropped and the detect will be re-run): read only what exists
    uint32_t n_in = A.count_in ? *A.count_in : A.n_in;
    if (n_in > A.cap_in) n_in = A.cap_in;
    if (blockIdx.x * A.chunk >= n_in) return;
    tab_init(A.tab, s_desc, s_ibase);
    const uint32_t tid = threadIdx.x;
    uint32_t n_tests = 0, n_adj = 0, n_cands = 0, n_fp64 = 0;
    bool overflow = false;
    for (uint32_t c0 = blockIdx.x * A.chunk; c0 < n_in; c0 += gridDim.x * A.chunk) {
        uint32_t n_cur = min(A.chunk, n_in - c0);
        uint32_t cur = 0;
        for (uint32_t lvl = 0; n_cur; lvl++) {
            const bool first = lvl == 0;
            const bool test = first && A.test_entries;
            const bool last = lvl + 1 >= A.levels;  // kept pairs go to the global front
            if (tid == 0) s_qn = 0;
            __syncthreads();
            for (uint32_t base = 0; base < n_cur; base += TB) {
                const uint32_t i = base + tid;
                bool have = i < n_cur;
                uint2 e = make_uint2(0, 0);
                if (have) e = first ? A.front_in[c0 + i] : s_q[cur][i];
                if (have && (e.x >= A.implicit_end || e.y >= A.implicit_end)) {
                    or_u32(A.error, OI_ERR_CORRUPT);
                    have = false;
                }
                u64 mask = 0;
                bool leafout = false;
                Side sa = {0, 0}, sb = {0, 0};
                uint32_t baseA = 0, baseB = 0;
                if (have) {
                    if (test)
                        entry_expand<E, true, Tab>(A.tab, s_desc, s_ibase, e, mask, leafout, sa, sb, baseA, baseB,
                                                   n_tests, n_adj);
                    else
                        entry_expand<E, false, Tab>(A.tab, s_desc, s_ibase, e, mask, leafout, sa, sb, baseA, baseB,
                                                    n_tests, n_adj);
                }
                const uint32_t nout = (uint32_t)__popcll(mask);
                // one scan for both counts (front pairs high, leaf pairs low)
                uint32_t TC;
                const uint32_t oc = block_scan(leafout ? nout : (nout << 16), s_wsum, TC);
                const uint32_t T = TC >> 16, C = TC & 0xffffu, off = oc >> 16, coff = oc & 0xffffu;
                if (tid == 0) {
                    if (last) {
                        s_obase = T ? add_u32(A.count_out, T) : 0u;
                    } else {
                        s_obase = s_qn;
                        s_qn += T;
                    }
                    if (!FUSED) s_cbase = C ? add_u32(A.cand_count, C) : 0u;
                }
                if constexpr (FUSED) {
                    s_f[0].cmask[tid] = leafout ? mask : 0ull;
                    s_f[0].coff[tid] = coff;
                    s_f[0].ga[tid] = baseA + sa.start;
                    s_f[0].gb[tid] = baseB + sb.start;
                }
                __syncthreads();
                if (nout && !leafout) {
                    uint32_t o = s_obase + off;
                    if (last && o + nout > A.cap_out) or_u32(A.error, OI_ERR_QUEUE_OVERFLOW);
                    u64 m = mask;
                    while (m) {
                        const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                        m &= m - 1ull;
                        const uint2 pr = make_uint2(baseA + sa.start + (k >> E), baseB + sb.start + (k & (K - 1)));
                        if (last) {
                            if (o < A.cap_out) A.front_out[o] = pr;
                        } else if (o < LQ) {
                            s_q[cur ^ 1u][o] = pr;
                        }
                        o++;
                    }
                }
                if (!FUSED && nout && leafout) {
                    uint32_t o = s_cbase + coff;
                    if (o + nout > A.cand_cap) or_u32(A.error, OI_ERR_CAND_OVERFLOW);
                    u64 m = mask;
                    while (m) {
                        const uint32_t k = (uint32_t)__ffsll((long long)m) - 1u;
                        m &= m - 1ull;
                        if (o < A.cand_cap)
                            A.cands[o] = make_uint2(baseA + sa.start + (k >> E), baseB + sb.start + (k & (K - 1)));
                        o++;
                    }
                }
                if constexpr (FUSED) {
                    if (C) {
                        if (tid == 0) n_cands += C;
                        fused_narrow<E, false, Tab>(A.tab, s_ibase, s_desc, s_f[0], C, s_wsum, A.hits, A.n_hits,
                                                    A.hits_cap, nullptr, nullptr, 0, A.error, n_fp64, n_adj);
                    }
                }
                __syncthreads();
            }
            if (last) break;
            const uint32_t q = s_qn;
            if (q > LQ) overflow = true;
            n_cur = q < LQ ? q : LQ;
            cur ^= 1u;
            __syncthreads();
        }
    }
    if (tid == 0 && overflow) or_u32(A.error, A.overflow_bit);
    uint32_t nt, na;
    block_scan(n_tests, s_wsum, nt);
    block_scan(n_adj, s_wsum, na);
    if (tid == 0 && nt) __hip_atomic_fetch_add(A.tests, (u64)nt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0 && na) __hip_atomic_fetch_add(A.adj, (u64)na, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (FUSED && tid == 0 && n_cands) add_u32(A.cand_count, n_cands);
}

// ---------------------------------------------------------------------------
// k_narrow: SAT over the candidate list
// ---------------------------------------------------------------------------
template <class Tab>
struct NarrowArgs {
    Tab tab;
    const uint2* cands;
    const uint32_t* cand_count;
    uint32_t cand_cap;
    uint4* hits;
    uint32_t* n_hits;
    uint32_t hits_cap;
    uint32_t* error;
    unsigned long long* fp64;   // SAT axis comparisons evaluated in double
    unsigned long long* adj;    // self-mode leaf pairs dropped as adjacent
    uint2* defer;               // NARROW_FILTER: undecided candidates for NARROW_FP64
    uint32_t* defer_count;
    uint32_t defer_cap;
};

// NARROW_FP32: the reference's FP32 SAT. NARROW_FILTER: the FP32 filter;
// undecided candidates go to the defer list. NARROW_FP64: the FP64 oracle's
// SAT over the defer list (the precision form's second pass).
enum { NARROW_FP32 = 0, NARROW_FILTER = 1, NARROW_FP64 = 2 };

// The two triangles of candidate c as the SAT takes them: sorted faces and
// original ids, then the vertices through the model matrices; inside one
// mesh P is the face with the smaller original id, across meshes the face of
// the lower mesh. False (nothing to test) for self-mode pairs that share a
// vertex, counted in n_adj when count_adj.
template <class Tab>
__device__ __forceinline__ bool narrow_gather(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                              uint2 c, bool count_adj, uint32_t& n_adj, v3f* P, v3f* Q, uint4& id)
{
    uint32_t ma, mb;
    const MeshRef ra = tab_ref(T, s_desc, s_ibase, c.x, ma);
    const MeshRef rb = tab_ref(T, s_desc, s_ibase, c.y, mb);
    const uint32_t sa = c.x - ra.ibase - ((1u << ra.L) - 1u);
    const uint32_t sb = c.y - rb.ibase - ((1u << rb.L) - 1u);
    // the original ids are read with the faces, and the vertices before the
    // ids arrive: two dependent rounds of loads per candidate
    uint32_t fa[3], fb[3];
    oi_ld3(ra.sfaces, sa, fa);
    oi_ld3(rb.sfaces, sb, fb);
    uint32_t oa = oi_g(ra.perm)[sa], ob = oi_g(rb.perm)[sb];
    if (ma == mb) {
        bool adj = false;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) adj |= (fa[i] == fb[j]);
        if (adj) {
            if (count_adj) n_adj++;
            return false;
        }
    }
    // all six vertices, then both model matrices, are loaded before any is
    // used (one round of loads, not one per vertex)
    float pa[3][3], pb[3][3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        oi_ld3(ra.pos, fa[k], pa[k]);
        oi_ld3(rb.pos, fb[k], pb[k]);
    }
    // (a mesh without a matrix reads a zero block in its place, unused)
    const float* sxa = ra.xf ? ra.xf : k_no_xf;
    const float* sxb = rb.xf ? rb.xf : k_no_xf;
    float xa[16], xb[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        xa[k] = oi_g(sxa)[k];
        xb[k] = oi_g(sxb)[k];
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float o[3];
        if (ra.xf) {
            oi_xform(xa, pa[k][0], pa[k][1], pa[k][2], o);
            P[k] = v3f{o[0], o[1], o[2]};
        } else {
            P[k] = v3f{pa[k][0], pa[k][1], pa[k][2]};
        }
        if (rb.xf) {
            oi_xform(xb, pb[k][0], pb[k][1], pb[k][2], o);
            Q[k] = v3f{o[0], o[1], o[2]};
        } else {
            Q[k] = v3f{pb[k][0], pb[k][1], pb[k][2]};
        }
    }
    if (ma == mb && oa > ob) {
        // both faces share the mesh's positions and model matrix, so the
        // loaded vertices are swapped
        const uint32_t t = oa;
        oa = ob;
        ob = t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const v3f u = P[k];
            P[k] = Q[k];
            Q[k] = u;
        }
    }
    id = make_uint4(ma, mb, oa, ob);
    return true;
}

// FP32 narrow phase of one candidate (the fused BVTT steps use it)
template <bool FP64, class Tab>
__device__ __forceinline__ void narrow_one(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                           uint2 c, bool& hit, uint4& out, uint32_t& n_fp64, uint32_t& n_adj)
{
    static_assert(!FP64, "the FP64 fallback runs in k_narrow's two passes");
    hit = false;
    v3f P[3], Q[3];
    uint4 id;
    if (!narrow_gather(T, s_ibase, s_desc, c, true, n_adj, P, Q, id)) return;
    if (sat_f32(P, Q)) {
        hit = true;
        out = id;
    }
}

// the FP32 filter on one candidate (the fused steps of an FP64-fallback
// query): a hit, a pair left to the FP64 pass (defer), or neither
template <class Tab>
__device__ __forceinline__ void narrow_one_filter(const Tab& T, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                                  uint2 c, bool& hit, bool& defer, uint4& out, uint32_t& n_adj)
{
    hit = false;
    defer = false;
    v3f P[3], Q[3];
    uint4 id;
    if (!narrow_gather(T, s_ibase, s_desc, c, true, n_adj, P, Q, id)) return;
    const int r = sat_filter(P, Q);
    hit = r == SAT_HIT;
    defer = r == SAT_UNDECIDED;
    out = id;
}

template <int MODE, class Tab>
__global__ void __launch_bounds__(TB) k_narrow(NarrowArgs<Tab> A)
{
    __shared__ oi_mesh_desc s_desc[MESH_LDS];
    __shared__ uint32_t s_ibase[MESH_LDS];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_hbase, s_dbase;
    uint32_t n = *A.cand_count;
    if (n > A.cand_cap) n = A.cand_cap;
    if (blockIdx.x * TB >= n) return;
    tab_init(A.tab, s_desc, s_ibase);
    const uint32_t tid = threadIdx.x;
    uint32_t n_fp64 = 0, n_adj = 0;
    const uint32_t stride = gridDim.x * TB;
    uint2 c_next = (blockIdx.x * TB + tid < n) ? A.cands[blockIdx.x * TB + tid] : make_uint2(0, 0);
    for (uint32_t base = blockIdx.x * TB; base < n; base += stride) {
        const uint32_t i = base + tid;
        bool hit = false, defer = false;
        uint4 out = make_uint4(0, 0, 0, 0);
        const uint2 c = c_next;
        if (i < n) {
            v3f P[3], Q[3];
            // deferred candidates were checked for adjacency by the first pass
            const bool live = narrow_gather(A.tab, s_ibase, s_desc, c, MODE != NARROW_FP64, n_adj, P, Q, out);
            // the next candidate is loaded once this one's vertices have
            // arrived: loads complete in issue order, so issued earlier it
            // would hold up the wait for them
            if (i + stride < n) {
                const float dep = live ? P[2].z + Q[2].z : 0.0f;
                c_next = A.cands[i + stride + opaque_zero(dep)];
            }
            if (live) {
                if (MODE == NARROW_FP32) {
                    hit = sat_f32(P, Q);
                } else if (MODE == NARROW_FILTER) {
                    const int r = sat_filter(P, Q);
                    hit = r == SAT_HIT;
                    defer = r == SAT_UNDECIDED;
                } else {
                    hit = sat_f64_counted(P, Q, n_fp64);
                }
            }
            if (defer) out = make_uint4(c.x, c.y, 0, 0);
        }
        // one scan for both counts: hits in the high half, deferred pairs in
        // the low half (at most 256 each, so no carry crosses)
        uint32_t HD;
        const uint32_t o = block_scan((hit ? 65536u : 0u) + (defer ? 1u : 0u), s_wsum, HD);
        const uint32_t H = HD >> 16, D = HD & 0xffffu;
        if (tid == 0) {
            s_hbase = H ? add_u32(A.n_hits, H) : 0u;
            if (MODE == NARROW_FILTER) s_dbase = D ? add_u32(A.defer_count, D) : 0u;
        }
        __syncthreads();
        if (hit) {
            const uint32_t pos = s_hbase + (o >> 16);
            if (pos < A.hits_cap)
                A.hits[pos] = out;
            else
                or_u32(A.error, OI_ERR_HITS_OVERFLOW);
        }
        if (MODE == NARROW_FILTER && defer) {
            const uint32_t pos = s_dbase + (o & 0xffffu);
            if (pos < A.defer_cap)
                A.defer[pos] = make_uint2(out.x, out.y);
            else
                or_u32(A.error, OI_ERR_QUEUE_OVERFLOW);
        }
        __syncthreads();
    }
    uint32_t f, a;
    block_scan(n_fp64, s_wsum, f);
    block_scan(n_adj, s_wsum, a);
    if (tid == 0) {
        if (f) __hip_atomic_fetch_add(A.fp64, (u64)f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (a) __hip_atomic_fetch_add(A.adj, (u64)a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// seed: zeroed level counts and result counters (counts[0] = initial front)
__global__ void k_seed(uint32_t n0, uint32_t* counts, uint32_t n_counts, oi_detect_ctl* ctl, uint32_t reset_sticky)
{
    const uint32_t i = threadIdx.x;
    if (i < n_counts) counts[i] = (i == 0) ? n0 : 0u;
    if (i == 0) {
        // errors of the previous detect move to `sticky` unless the caller
        // checked it synchronously
        ctl->sticky = reset_sticky ? 0u : (ctl->sticky | ctl->error);
        ctl->error = 0;
        ctl->n_hits = 0;
        ctl->cands = 0;
        ctl->fp64 = 0;
        ctl->adj = 0;
        ctl->tests = 0;
        ctl->deferred = 0;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// a launch whose completion also records `done` (when not null) as part of
// the dispatch, which saves a separate hipEventRecord call on the host
template <class K, class... Args>
static void launch_done(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t done, Args... args)
{
    if (done)
        hipExtLaunchKernelGGL(kernel, g, b, 0, st, nullptr, done, 0, args...);
    else
        hipLaunchKernelGGL(kernel, g, b, 0, st, args...);
}

static MeshTab make_tab(const oi_mesh_desc* d_meshes, const oi_mesh_desc* h_meshes, uint32_t n)
{
    MeshTab t{};
    t.meshes = d_meshes;
    t.n_meshes = n;
    for (uint32_t m = 0; m < n && m < INLINE_MESHES; m++) t.inl[m] = h_meshes[m];
    return t;
}

template <class Tab>
static hipError_t launch_step(hipStream_t st, uint32_t grid, const Tab& tab, int first, const uint2* front_in,
                              const uint32_t* count_in, uint32_t n_in, uint32_t cap_in, uint2* front_out,
                              uint32_t* count_out, uint32_t cap_out, uint2* cands, uint32_t* cand_count,
                              uint32_t cand_cap, uint4* hits, uint32_t hits_cap, oi_detect_ctl* ctl, uint32_t expand,
                              uint32_t implicit_end, int fused, hipEvent_t done)
{
    if (fused < 0 || fused > 2) return hipErrorInvalidValue;
    StepArgs<Tab> a{};
    a.tab = tab;
    a.front_in = front_in;
    a.count_in = count_in;
    a.n_in = n_in;
    a.cap_in = cap_in;
    a.front_out = front_out;
    a.count_out = count_out;
    a.cap_out = cap_out;
    a.cands = cands;
    a.cand_count = cand_count;
    a.cand_cap = cand_cap;
    a.error = &ctl->error;
    a.tests = &ctl->tests;
    a.adj = &ctl->adj;
    a.hits = hits;
    a.n_hits = &ctl->n_hits;
    a.hits_cap = hits_cap;
    // the filter's deferred pairs go to the candidate buffer, which fused
    // steps leave unused
    a.defer = cands;
    a.defer_count = &ctl->deferred;
    a.defer_cap = cand_cap;
    a.implicit_end = implicit_end;
    const dim3 g(grid), b(TB);
#define OI_STEP(EE)                                                                    \
    if (first && fused == 2)                                                           \
        launch_done((k_bvtt_step<EE, true, 2, Tab>), g, b, st, done, a);               \
    else if (first && fused)                                                           \
        launch_done((k_bvtt_step<EE, true, 1, Tab>), g, b, st, done, a);               \
    else if (first)                                                                    \
        launch_done((k_bvtt_step<EE, true, 0, Tab>), g, b, st, done, a);               \
    else if (fused == 2)                                                               \
        launch_done((k_bvtt_step<EE, false, 2, Tab>), g, b, st, done, a);              \
    else if (fused)                                                                    \
        launch_done((k_bvtt_step<EE, false, 1, Tab>), g, b, st, done, a);              \
    else                                                                               \
        launch_done((k_bvtt_step<EE, false, 0, Tab>), g, b, st, done, a);
    switch (expand) {
    case 1: OI_STEP(1) break;
    case 2: OI_STEP(2) break;
    case 3: OI_STEP(3) break;
    default: return hipErrorInvalidValue;
    }
#undef OI_STEP
    return hipGetLastError();
}

template <class Tab>
static hipError_t launch_local(hipStream_t st, const Tab& tab, const oi_local_args& w, uint2* cands,
                               uint32_t* cand_count, uint32_t cand_cap, uint4* hits, uint32_t hits_cap,
                               oi_detect_ctl* ctl, uint32_t expand, uint32_t implicit_end, int fused)
{
    // without a device count the grid is one workgroup per chunk
    const uint64_t chunks = w.count_in ? ((uint64_t)w.cap_in + w.chunk - 1) / w.chunk
                                       : ((uint64_t)w.n_in + w.chunk - 1) / w.chunk;
    const uint64_t grid = (w.count_in && w.grid && w.grid < chunks) ? w.grid : chunks;
    if (grid == 0) return w.done ? hipEventRecord(w.done, st) : hipSuccess;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    LocalArgs<Tab> a{};
    a.tab = tab;
    a.front_in = w.front_in;
    a.count_in = w.count_in;
    a.n_in = w.n_in;
    a.cap_in = w.count_in ? w.cap_in : w.n_in;
    a.chunk = w.chunk;
    a.levels = w.levels;
    a.test_entries = w.test_entries;
    a.overflow_bit = w.overflow_bit;
    a.front_out = w.front_out;
    a.count_out = w.count_out;
    a.cap_out = w.cap_out;
    a.cands = cands;
    a.cand_count = cand_count;
    a.cand_cap = cand_cap;
    a.error = &ctl->error;
    a.tests = &ctl->tests;
    a.adj = &ctl->adj;
    a.hits = hits;
    a.n_hits = &ctl->n_hits;
    a.hits_cap = hits_cap;
    a.implicit_end = implicit_end;
    const dim3 g((uint32_t)grid), b(TB);
#define OI_LOCAL(EE)                                                           \
    if (fused)                                                                 \
        launch_done((k_bvtt_local<EE, true, Tab>), g, b, st, w.done, a);       \
    else                                                                       \
        launch_done((k_bvtt_local<EE, false, Tab>), g, b, st, w.done, a);
    switch (expand) {
    case 1: OI_LOCAL(1) break;
    case 2: OI_LOCAL(2) break;
    case 3: OI_LOCAL(3) break;
    default: return hipErrorInvalidValue;
    }
#undef OI_LOCAL
    return hipGetLastError();
}

template <class Tab>
static hipError_t launch_narrow(hipStream_t st, uint32_t grid, const Tab& tab, const uint2* cands,
                                const uint32_t* cand_count, uint32_t cand_cap, uint4* hits, uint32_t* n_hits,
                                uint32_t hits_cap, oi_detect_ctl* ctl, int mode, uint2* defer, uint32_t defer_cap,
                                hipEvent_t done)
{
    NarrowArgs<Tab> a{};
    a.tab = tab;
    a.cands = cands;
    a.cand_count = cand_count;
    a.cand_cap = cand_cap;
    a.hits = hits;
    a.n_hits = n_hits;
    a.hits_cap = hits_cap;
    a.error = &ctl->error;
    a.fp64 = &ctl->fp64;
    a.adj = &ctl->adj;
    a.defer = defer;
    a.defer_count = &ctl->deferred;
    a.defer_cap = defer_cap;
    switch (mode) {
    case NARROW_FP32: launch_done((k_narrow<NARROW_FP32, Tab>), dim3(grid), dim3(TB), st, done, a); break;
    case NARROW_FILTER: launch_done((k_narrow<NARROW_FILTER, Tab>), dim3(grid), dim3(TB), st, done, a); break;
    case NARROW_FP64: launch_done((k_narrow<NARROW_FP64, Tab>), dim3(grid), dim3(TB), st, done, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" {

hipError_t oi_launch_seed(hipStream_t st, uint32_t n0, uint32_t* counts, uint32_t n_counts, oi_detect_ctl* ctl,
                          uint32_t reset_sticky)
{
    if (n_counts > 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(256), 0, st, n0, counts, n_counts, ctl, reset_sticky);
    return hipGetLastError();
}

// one BVTT step; front_in's length is *count_in, or n_in when count_in is null
// fused 1: overlapping leaf pairs are tested with the FP32 SAT inside the step
// (hits appended directly) instead of going to the candidate list; fused 2:
// with the FP32 filter, undecided pairs to the candidate buffer (count in
// ctl->deferred) for the FP64 pass
hipError_t oi_launch_bvtt_step(hipStream_t st, uint32_t grid, const oi_tab_args* t, int first,
                               const uint2* front_in, const uint32_t* count_in, uint32_t n_in, uint32_t cap_in,
                               uint2* front_out, uint32_t* count_out, uint32_t cap_out, uint2* cands,
                               uint32_t* cand_count, uint32_t cand_cap, uint4* hits, uint32_t hits_cap,
                               oi_detect_ctl* ctl, uint32_t expand, uint32_t implicit_end, int fused, hipEvent_t done)
{
    if (t->chunk_mesh)
        return launch_step(st, grid, ChunkTab{t->d_meshes, t->chunk_mesh, t->shift}, first, front_in, count_in, n_in,
                           cap_in, front_out, count_out, cap_out, cands, cand_count, cand_cap, hits, hits_cap, ctl,
                           expand, implicit_end, fused, done);
    return launch_step(st, grid, make_tab(t->d_meshes, t->h_meshes, t->n_meshes), first, front_in, count_in, n_in,
                       cap_in, front_out, count_out, cap_out, cands, cand_count, cand_cap, hits, hits_cap, ctl,
                       expand, implicit_end, fused, done);
}

// the first `levels` steps in one launch, one workgroup per `chunk` initial
// entries; the pairs kept by the last of them go to front_out / count_out
hipError_t oi_launch_bvtt_local(hipStream_t st, const oi_tab_args* t, const oi_local_args* w, uint2* cands,
                                uint32_t* cand_count, uint32_t cand_cap, uint4* hits, uint32_t hits_cap,
                                oi_detect_ctl* ctl, uint32_t expand, uint32_t implicit_end, int fused)
{
    if (w->chunk == 0 || w->levels == 0 || w->chunk > LQ) return hipErrorInvalidValue;
    if (t->chunk_mesh)
        return launch_local(st, ChunkTab{t->d_meshes, t->chunk_mesh, t->shift}, *w, cands, cand_count, cand_cap,
                            hits, hits_cap, ctl, expand, implicit_end, fused);
    return launch_local(st, make_tab(t->d_meshes, t->h_meshes, t->n_meshes), *w, cands, cand_count, cand_cap, hits,
                        hits_cap, ctl, expand, implicit_end, fused);
}

// Scene::convertToVertexArray (scene.cu:69-93): six vertices per hit, the
// three of triangle A then the three of triangle B, each through its mesh's
// model matrix as the detect saw it. One thread per output vertex.
__global__ void __launch_bounds__(256)
k_pair_vertices(const oi_build_desc* __restrict__ g, const uint4* __restrict__ pairs, uint32_t n,
                float* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= 6ull * n) return;
    const uint32_t p = (uint32_t)(i / 6u), r = (uint32_t)(i % 6u);
    const uint4 h = pairs[p];
    const bool b = r >= 3;
    const oi_build_desc& d = g[b ? h.y : h.x];
    const uint32_t tri = b ? h.w : h.z;
    const uint32_t v = oi_g(d.faces)[3ull * tri + (r % 3u)];
    const v3f q = vertex(d.pos, d.xf, v);
    OI_GLOBAL float* o = oi_gw(out) + 3ull * i;
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
}

// ---------------------------------------------------------------------------
// The hit list in get_pairs order (meshA, meshB, triA, triB, lexicographic),
// sorted on the device: k_hits_digits splits the hits into four word arrays
// and builds the 256-bin histogram of every byte the sort visits (the bytes
// that can be non-zero, from the mesh count and the largest face count);
// LSD radix passes (k_sort_pass, keys looked up by hit index) then order the
// hit indices by each byte in turn, least significant first, and
// k_hits_gather writes the hits in that order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* __restrict__ p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) p[i] = 0u;
}

__global__ void __launch_bounds__(256)
k_hits_digits(const uint4* __restrict__ hits, uint32_t n, uint32_t* __restrict__ words, uint32_t* __restrict__ hist,
              oi_hit_digits dg)
{
    __shared__ uint32_t s_h[OI_HIT_DIGITS][256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t p = 0; p < dg.nd; p++) s_h[p][tid] = 0u;
    __syncthreads();
    OI_GLOBAL uint32_t* w = oi_gw(words);
    for (uint32_t i = blockIdx.x * 256u + tid; i < n; i += gridDim.x * 256u) {
        const uint4 h = hits[i];
        w[i] = h.x;
        w[(uint64_t)n + i] = h.y;
        w[2ull * n + i] = h.z;
        w[3ull * n + i] = h.w;
        for (uint32_t p = 0; p < dg.nd; p++) {
            const uint32_t k = dg.word[p] == 0 ? h.x : dg.word[p] == 1 ? h.y : dg.word[p] == 2 ? h.z : h.w;
            atomicAdd(&s_h[p][(k >> dg.shift[p]) & 0xffu], 1u);
        }
    }
    __syncthreads();
    for (uint32_t p = 0; p < dg.nd; p++) {
        const uint32_t c = s_h[p][tid];
        if (c) atomicAdd(&hist[256u * p + tid], c);
    }
}

__global__ void __launch_bounds__(256)
k_hits_gather(const uint4* __restrict__ hits, const uint32_t* __restrict__ order, uint32_t n, uint4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = hits[order[i]];
}

// A hit list of at most HS_MAX pairs sorted by one workgroup of HS_WAVES
// waves: the hit indices stay in LDS and every byte the sort visits is one
// pass inside the kernel (one launch instead of the zeroing, histogram, pass
// and gather launches). Pass layout as k_sort_pass: wave w owns positions
// [512 w, 512 w + 512), round j lane l holds position 512 w + 64 j + l; each
// thread ranks its keys with 8 ballots per round (stable), the per-wave digit
// counts give each key its slot, and the indices move to the other LDS
// buffer. The key byte of hit v is read from hits[v]. The sorted hits are
// then gathered to out.
#define HS_WAVES 16
#define HS_THREADS (64 * HS_WAVES)
#define HS_ROUNDS 8
#define HS_MAX (HS_THREADS * HS_ROUNDS)

__global__ void __launch_bounds__(HS_THREADS)
k_hits_sort_small(const uint4* __restrict__ hits, uint32_t n, oi_hit_digits dg, uint4* __restrict__ out)
{
    __shared__ uint32_t s_idx[2][HS_MAX];
    __shared__ uint32_t s_whist[HS_WAVES][256];
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_wsum[HS_WAVES];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t i = tid; i < n; i += HS_THREADS) s_idx[0][i] = i;
    const unsigned long long lt = (1ull << lane) - 1ull;
    uint32_t cur = 0;
    for (uint32_t p = 0; p < dg.nd; p++) {
        const uint32_t word = dg.word[p], shift = dg.shift[p];
        for (uint32_t i = tid; i < HS_WAVES * 256u; i += HS_THREADS) (&s_whist[0][0])[i] = 0u;
        __syncthreads();
        uint32_t v[HS_ROUNDS], d[HS_ROUNDS], rk[HS_ROUNDS];
#pragma unroll
        for (int j = 0; j < HS_ROUNDS; j++) {
            const uint32_t pos = 512u * wave + 64u * j + lane;
            v[j] = pos < n ? s_idx[cur][pos] : 0u;
        }
#pragma unroll
        for (int j = 0; j < HS_ROUNDS; j++) {
            const uint32_t pos = 512u * wave + 64u * j + lane;
            const OI_GLOBAL uint32_t* h = (const OI_GLOBAL uint32_t*)oi_g(hits) + 4ull * v[j] + word;
            d[j] = pos < n ? (*h >> shift) & 0xffu : 0u;
        }
#pragma unroll
        for (int j = 0; j < HS_ROUNDS; j++) {
            const uint32_t pos = 512u * wave + 64u * j + lane;
            const bool ok = pos < n;
            unsigned long long peers = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const bool bit = (d[j] >> b) & 1u;
                const unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const uint32_t prev = s_whist[wave][d[j]];
            rk[j] = prev + (uint32_t)__popcll(peers & lt);
            if (ok && (peers & lt) == 0ull) s_whist[wave][d[j]] = prev + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        // threads tid < 256 own digit tid: per-wave offsets, then the digit
        // starts (a scan over the first four waves)
        uint32_t tot = 0;
        if (tid < 256u) {
#pragma unroll
            for (int w = 0; w < HS_WAVES; w++) {
                const uint32_t c = s_whist[w][tid];
                s_whist[w][tid] = tot;
                tot += c;
            }
        }
        const uint32_t x = oi_wave_scan_incl(tot);
        if (lane == 63u) s_wsum[wave] = x;
        __syncthreads();
        if (tid < 256u) {
            uint32_t before = 0;
#pragma unroll
            for (int w = 0; w < 4; w++) before += (w < (int)wave) ? s_wsum[w] : 0u;
            s_start[tid] = before + x - tot;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < HS_ROUNDS; j++) {
            const uint32_t pos = 512u * wave + 64u * j + lane;
            if (pos < n) s_idx[cur ^ 1u][s_start[d[j]] + s_whist[wave][d[j]] + rk[j]] = v[j];
        }
        cur ^= 1u;
        __syncthreads();
    }
    for (uint32_t i = tid; i < n; i += HS_THREADS) out[i] = hits[s_idx[cur][i]];
}

hipError_t oi_launch_hits_sort_small(hipStream_t st, const uint4* hits, uint32_t n, const oi_hit_digits* dg,
                                     uint4* out)
{
    if (n == 0) return hipSuccess;
    if (n > HS_MAX || dg->nd > OI_HIT_DIGITS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_hits_sort_small, dim3(1), dim3(HS_THREADS), 0, st, hits, n, *dg, out);
    return hipGetLastError();
}

uint32_t oi_hits_sort_small_max() { return HS_MAX; }

hipError_t oi_launch_zero_u32(hipStream_t st, uint32_t* p, uint32_t n)
{
    if (n == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 1024u);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t oi_launch_hits_digits(hipStream_t st, const uint4* hits, uint32_t n, uint32_t* words, uint32_t* hist,
                                 const oi_hit_digits* dg)
{
    if (n == 0) return hipSuccess;
    if (dg->nd > OI_HIT_DIGITS) return hipErrorInvalidValue;
    const uint32_t blocks = std::min<uint32_t>((n + 255u) / 256u, 2048u);
    hipLaunchKernelGGL(k_hits_digits, dim3(blocks), dim3(256), 0, st, hits, n, words, hist, *dg);
    return hipGetLastError();
}

hipError_t oi_launch_hits_gather(hipStream_t st, const uint4* hits, const uint32_t* order, uint32_t n, uint4* out)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_hits_gather, dim3((n + 255u) / 256u), dim3(256), 0, st, hits, order, n, out);
    return hipGetLastError();
}

hipError_t oi_launch_pair_vertices(hipStream_t st, const oi_build_desc* g, const uint4* pairs, uint32_t n, float* out)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (6ull * n + 255u) / 256u;
    hipLaunchKernelGGL(k_pair_vertices, dim3((uint32_t)blocks), dim3(256), 0, st, g, pairs, n, out);
    return hipGetLastError();
}

// mode: 0 the FP32 SAT over the candidates; 1 the FP32 filter, undecided
// candidates appended to `defer` (capacity defer_cap, count ctl->deferred);
// 2 the FP64 SAT over those (cands = defer, cand_count = &ctl->deferred)
hipError_t oi_launch_narrow(hipStream_t st, uint32_t grid, const oi_tab_args* t, const uint2* cands,
                            const uint32_t* cand_count, uint32_t cand_cap, uint4* hits, uint32_t* n_hits,
                            uint32_t hits_cap, oi_detect_ctl* ctl, int mode, uint2* defer, uint32_t defer_cap,
                            hipEvent_t done)
{
    if (t->chunk_mesh)
        return launch_narrow(st, grid, ChunkTab{t->d_meshes, t->chunk_mesh, t->shift}, cands, cand_count, cand_cap,
                             hits, n_hits, hits_cap, ctl, mode, defer, defer_cap, done);
    return launch_narrow(st, grid, make_tab(t->d_meshes, t->h_meshes, t->n_meshes), cands, cand_count, cand_cap,
                         hits, n_hits, hits_cap, ctl, mode, defer, defer_cap, done);
}

}  // extern "C"
