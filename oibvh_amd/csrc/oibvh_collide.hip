// oibvh_collide.hip -- gfx950 broad + narrow phase for Scene::detectCollision.
//
// The reference relaunches traversal_kernel once per BVTT level with a host
// round trip for the front size (scene.cu:272-311, collide.cu:76-246) and then
// runs triangle_intersect_kernel over the candidate list (collide.cu:248-300).
// Here one persistent launch does both:
//   * each workgroup keeps a BVTT stack in LDS and pops 256 entries per step
//     (one per thread): AABB overlap test (collide.cu:12-17), leaf x leaf ->
//     candidate, otherwise both sides are expanded `expand` levels (a leaf
//     stands for itself, collide.cu:165-176) and the children are pushed back
//     with a block prefix sum;
//   * a global ring queue takes spills and donations to idle workgroups
//     (work stealing); a `pending` counter (ring entries + workgroups holding
//     entries, updated only when work moves through the ring or a workgroup
//     changes between busy and idle) lets idle workgroups stop;
//   * candidates are compacted into an LDS buffer and tested 256 at a time
//     with the 17-axis SAT (utils.cpp:71-169 / cuda_intersect_tritri.h:350-434),
//     hits appended with one atomic per batch as (meshA, meshB, origTriA,
//     origTriB) -- original face ids, mapped through the build permutation.
// Self mode (no reference counterpart) also walks each mesh against itself:
// a node paired with itself expands to the unordered pairs of its children and
// leaf pairs sharing a vertex index are dropped before the narrow phase.
//
// BVTT entries are two 32-bit implicit indices in the scene's concatenated
// implicit space (mesh m spans [ibase_m, ibase_m + 2^(L_m+1) - 1)).
#include <hip/hip_runtime.h>

#include "oibvh_common.h"

#define TB 256
#define STACK_CAP 3072
#define CAND_CAP 512
#define MESH_LDS 16

typedef unsigned long long u64;

__device__ __forceinline__ u64 ld64(const u64* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st64(u64* p, u64 v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ int ld32i(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ uint32_t ld32(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Reserve `need` ring slots. A detect never wraps the ring: every index it
// reserves lies in [base, base + cap), so each slot it writes was emptied by an
// earlier detect (or the reset) and pushes need not wait. Fails when this
// detect would exceed the capacity; the host then grows the ring and re-runs.
__device__ bool ring_reserve(oi_trav_ctl* ctl, u64 base, uint32_t cap, uint32_t need, u64& out)
{
    for (;;) {
        const u64 t = ld64(&ctl->tail);
        if (t + need - base > cap) return false;
        u64 expected = t;
        if (__hip_atomic_compare_exchange_strong(&ctl->tail, &expected, t + need, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            out = t;
            return true;
        }
    }
}

// Ring slots. Ring index i maps to slot i & (cap - 1) and to lap
// (i >> log2 cap) & 15. A slot moves through EMPTY(lap) -> FULL(lap) ->
// EMPTY(lap + 1): the pusher of index i waits for EMPTY(lap(i)) and the popper
// of index i waits for FULL(lap(i)), so neither can act on a value of another
// lap when the ring wraps. Entries are two 30-bit implicit indices (the host
// keeps a scene's implicit space below 2^30 - 1); the lap tag lives in bits
// 31:30 and 63:62, and EMPTY is the all-ones index pair.
#define RING_IDX 0x3fffffffu
__device__ __forceinline__ u64 ring_pack(uint32_t a, uint32_t b, uint32_t lap)
{
    return (u64)(a | ((lap & 3u) << 30)) | ((u64)(b | ((lap >> 2) << 30)) << 32);
}
__device__ __forceinline__ u64 ring_empty(uint32_t lap) { return ring_pack(RING_IDX, RING_IDX, lap & 15u); }
__device__ __forceinline__ uint32_t ring_lap_of(u64 v) { return ((uint32_t)v >> 30) | (((uint32_t)(v >> 32) >> 30) << 2); }

// Push: the slot of a reserved index is free (see ring_reserve), so a plain
// atomic store publishes the entry.
__device__ __forceinline__ void slot_put(u64* ring, uint32_t shift, u64 idx, uint2 e)
{
    const uint32_t mask = (1u << shift) - 1u;
    const uint32_t lap = (uint32_t)(idx >> shift) & 15u;
    st64(ring + (idx & mask), ring_pack(e.x, e.y, lap));
}

// Pop: a reserved index is written promptly by its pusher, so the wait ends;
// the wall-clock limit only guards against a fault. On expiry it flags
// OI_ERR_TIMEOUT and reports failure, and the caller drops the entry.
__device__ __forceinline__ bool slot_take(u64* ring, uint32_t shift, u64 idx, uint2& e, oi_trav_ctl* ctl, u64 t0,
                                          u64 limit)
{
    const uint32_t mask = (1u << shift) - 1u;
    const uint32_t lap = (uint32_t)(idx >> shift) & 15u;
    u64* slot = ring + (idx & mask);
    u64 v;
    for (;;) {
        v = ld64(slot);
        if (ring_lap_of(v) == lap && ((uint32_t)v & RING_IDX) != RING_IDX) break;
        __builtin_amdgcn_s_sleep(1);
        if (wall_clock64() - t0 > limit) {
            __hip_atomic_fetch_or(&ctl->error, OI_ERR_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    e = make_uint2((uint32_t)v & RING_IDX, (uint32_t)(v >> 32) & RING_IDX);
    st64(slot, ring_empty(lap + 1u));
    return true;
}

__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* s_wsum, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t w0 = s_wsum[0], w1 = s_wsum[1], w2 = s_wsum[2], w3 = s_wsum[3];
    uint32_t before = (wave > 0 ? w0 : 0u) + (wave > 1 ? w1 : 0u) + (wave > 2 ? w2 : 0u);
    total = w0 + w1 + w2 + w3;
    __syncthreads();
    return before + x - v;
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
// Every FP32 quantity z carries a magnitude S(z) >= |z|, |exact z| built by
// the same expression over |operands| (sums of magnitudes for +/-, products
// for *). A value at rounding depth c has |z - exact z| <= c * 2^-24 * S(z)
// (+ O(2^-48)); the deepest SAT projection (a g/h axis) has c = 14. A
// comparison mn1 > mx2 is decided in FP32 only when |mn1 - mx2| exceeds
// T = 32 * 2^-24 * (S_P + S_Q) + 2^-100, which covers the FP32 error of both
// sides, the FP64 error of the oracle's own evaluation and the rounding of T;
// then the FP32 decision equals the exact one and the FP64 one. Otherwise the
// axis is evaluated again in double exactly as the FP64 oracle evaluates it.
// Non-finite magnitudes never pass the test, so they also go to FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ v3f vabs(v3f a) { return v3f{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ v3f vaddm(v3f a, v3f b) { return v3f{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3f vcrossm(v3f a, v3f b)
{
    return v3f{a.y * b.z + b.y * a.z, a.z * b.x + b.z * a.x, a.x * b.y + b.x * a.y};
}
__device__ __forceinline__ float vdotm(v3f a, v3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

struct SatMag {
    v3f p2, p3, q1, q2, q3, e1, e2, e3, f1, f2, f3, n1, m1;
    __device__ __forceinline__ void init(const v3f* P, const v3f* Q)
    {
        const v3f a0 = vabs(P[0]);
        p2 = vaddm(vabs(P[1]), a0);
        p3 = vaddm(vabs(P[2]), a0);
        q1 = vaddm(vabs(Q[0]), a0);
        q2 = vaddm(vabs(Q[1]), a0);
        q3 = vaddm(vabs(Q[2]), a0);
        e1 = p2;
        e2 = vaddm(p3, p2);
        e3 = p3;
        f1 = vaddm(q2, q1);
        f2 = vaddm(q3, q2);
        f3 = vaddm(q1, q3);
        n1 = vcrossm(e1, e2);
        m1 = vcrossm(f1, f2);
    }
    __device__ __forceinline__ v3f axis(int k) const
    {
        switch (k) {
        case AX_N1: return n1;
        case AX_M1: return m1;
        case AX_E11: return vcrossm(e1, f1);
        case AX_E12: return vcrossm(e1, f2);
        case AX_E13: return vcrossm(e1, f3);
        case AX_E21: return vcrossm(e2, f1);
        case AX_E22: return vcrossm(e2, f2);
        case AX_E23: return vcrossm(e2, f3);
        case AX_E31: return vcrossm(e3, f1);
        case AX_E32: return vcrossm(e3, f2);
        case AX_E33: return vcrossm(e3, f3);
        case AX_G1: return vcrossm(e1, n1);
        case AX_G2: return vcrossm(e2, n1);
        case AX_G3: return vcrossm(e3, n1);
        case AX_H1: return vcrossm(f1, m1);
        case AX_H2: return vcrossm(f2, m1);
        default: return vcrossm(f3, m1);
        }
    }
};

// returns 1 = not separated, 0 = separated; counts FP64 re-evaluations
__device__ bool sat_filtered(const v3f* P, const v3f* Q, uint32_t& n_fp64)
{
    SatFrame<float, v3f> s;
    s.init(P, Q);
    SatMag m;
    m.init(P, Q);
    bool have_d = false;
    SatFrame<double, v3d> sd;
    for (int k = 0; k < 17; k++) {
        const v3f ax = s.axis(k);
        const v3f am = m.axis(k);
        const v3f p1 = v3f{0, 0, 0};
        const float P1 = vdot<float>(ax, p1), P2 = vdot<float>(ax, s.p2), P3 = vdot<float>(ax, s.p3);
        const float Q1 = vdot<float>(ax, s.q1), Q2 = vdot<float>(ax, s.q2), Q3 = vdot<float>(ax, s.q3);
        const float mx1 = fmaxf(fmaxf(P1, P2), P3), mn1 = fminf(fminf(P1, P2), P3);
        const float mx2 = fmaxf(fmaxf(Q1, Q2), Q3), mn2 = fminf(fminf(Q1, Q2), Q3);
        const float SP = fmaxf(vdotm(am, m.p2), vdotm(am, m.p3));
        const float SQ = fmaxf(fmaxf(vdotm(am, m.q1), vdotm(am, m.q2)), vdotm(am, m.q3));
        const float T = 32.0f * 5.9604645e-8f * (SP + SQ) + 7.888609e-31f;
        const bool finite = (T < __builtin_inff());
        // separation on this axis: mn1 > mx2 or mn2 > mx1
        const bool sep1_sure = finite && (mn1 - mx2 > T);
        const bool nsep1_sure = finite && (mx2 - mn1 > T);
        const bool sep2_sure = finite && (mn2 - mx1 > T);
        const bool nsep2_sure = finite && (mx1 - mn2 > T);
        if (sep1_sure || sep2_sure) return false;
        if (nsep1_sure && nsep2_sure) continue;
        // uncertain: evaluate this axis exactly like the FP64 oracle
        if (!have_d) {
            v3d Pd[3], Qd[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                Pd[i] = v3d{(double)P[i].x, (double)P[i].y, (double)P[i].z};
                Qd[i] = v3d{(double)Q[i].x, (double)Q[i].y, (double)Q[i].z};
            }
            sd.init(Pd, Qd);
            have_d = true;
        }
        n_fp64++;
        if (!project6<double, v3d>(sd, sd.axis(k))) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
#define INLINE_MESHES 4
struct TravArgs {
    const oi_mesh_desc* meshes;  // mesh table in device memory (scenes of more than INLINE_MESHES meshes)
    uint32_t n_meshes;
    oi_mesh_desc inl[INLINE_MESHES];  // small scenes: the table travels in the kernel arguments
    u64* ring;
    uint32_t ring_shift;  // log2 of the ring capacity
    oi_trav_ctl* ctl;
    uint4* hits;
    uint32_t hits_cap;
    uint32_t expand;
    uint32_t flags;
    u64 timeout_ticks;
    uint32_t implicit_end;  // end of the scene's implicit index space
    unsigned long long* trace;  // OI_TRACE_WORDS per workgroup, or nullptr
};

struct MeshRef {
    const float* nodes;
    const uint32_t* sfaces;
    const float* pos;
    const uint32_t* perm;
    uint32_t L, vl, ibase;
};

__device__ __forceinline__ uint32_t find_mesh(const oi_mesh_desc* md, const uint32_t* s_ibase, uint32_t n,
                                              uint32_t g)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t b = (n <= MESH_LDS) ? s_ibase[mid] : md[mid].ibase;
        if (b <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// mesh m's descriptor: LDS copy for the first MESH_LDS meshes, else device memory
__device__ __forceinline__ MeshRef mesh_ref(const oi_mesh_desc* md, const oi_mesh_desc* s_desc, uint32_t m)
{
    const oi_mesh_desc& d = (m < MESH_LDS) ? s_desc[m] : md[m];
    return MeshRef{d.nodes, d.sfaces, d.pos, d.perm, d.L, d.vl, d.ibase};
}

__device__ __forceinline__ bool overlap6(const float* a, const float* b)
{
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
}

__device__ __forceinline__ void load_box(const float* nodes, uint32_t r, float* o)
{
    const float2* p = reinterpret_cast<const float2*>(nodes + 6u * r);
    const float2 a = p[0], b = p[1], c = p[2];
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
    o[4] = c.x;
    o[5] = c.y;
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

// k-th unordered pair (x <= y) of n items, row-major: (0,0),(0,1)..(0,n-1),(1,1)..
__device__ __forceinline__ void tri_pair(uint32_t k, uint32_t n, uint32_t& x, uint32_t& y)
{
    x = 0;
    uint32_t row = n;
    while (k >= row) {
        k -= row;
        x++;
        row--;
    }
    y = x + k;
}

template <bool FP64>
__device__ __forceinline__ void narrow_one(const TravArgs& A, const uint32_t* s_ibase, const oi_mesh_desc* s_desc,
                                           uint2 c, bool valid,
                                           bool& hit, uint4& out, uint32_t& n_fp64, uint32_t& n_adj)
{
    hit = false;
    if (!valid) return;
    const uint32_t ma = find_mesh(A.meshes, s_ibase, A.n_meshes, c.x);
    const uint32_t mb = find_mesh(A.meshes, s_ibase, A.n_meshes, c.y);
    const MeshRef ra = mesh_ref(A.meshes, s_desc, ma), rb = mesh_ref(A.meshes, s_desc, mb);
    uint32_t sa = c.x - ra.ibase - ((1u << ra.L) - 1u);
    uint32_t sb = c.y - rb.ibase - ((1u << rb.L) - 1u);
    uint32_t fa[3], fb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        fa[k] = ra.sfaces[3u * sa + k];
        fb[k] = rb.sfaces[3u * sb + k];
    }
    if (ma == mb) {
        bool adj = false;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) adj |= (fa[i] == fb[j]);
        if (adj) {
            n_adj++;
            return;
        }
    }
    uint32_t oa = ra.perm[sa], ob = rb.perm[sb];
    const float* pa = ra.pos;
    const float* pb = rb.pos;
    if (ma == mb && oa > ob) {
        // inside one mesh P is the face with the smaller original id
        uint32_t t = oa;
        oa = ob;
        ob = t;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            uint32_t u = fa[k];
            fa[k] = fb[k];
            fb[k] = u;
        }
    }
    v3f P[3], Q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        P[k] = v3f{pa[3u * fa[k]], pa[3u * fa[k] + 1], pa[3u * fa[k] + 2]};
        Q[k] = v3f{pb[3u * fb[k]], pb[3u * fb[k] + 1], pb[3u * fb[k] + 2]};
    }
    const bool h = FP64 ? sat_filtered(P, Q, n_fp64) : sat_f32(P, Q);
    if (h) {
        hit = true;
        out = make_uint4(ma, mb, oa, ob);
    }
}

template <bool FP64>
__global__ void __launch_bounds__(TB) k_traverse(TravArgs A)
{
    __shared__ uint2 s_stack[STACK_CAP];
    __shared__ uint2 s_cand[CAND_CAP];
    __shared__ uint32_t s_ibase[MESH_LDS];
    __shared__ oi_mesh_desc s_desc[MESH_LDS];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_n, s_cn, s_pop_n, s_flag, s_idle_hint;
    __shared__ u64 s_pop_h, s_push_base;

    const uint32_t tid = threadIdx.x;
    oi_trav_ctl* ctl = A.ctl;
    if (tid < MESH_LDS && tid < A.n_meshes) {
        if (A.n_meshes <= INLINE_MESHES) {
            // constant indices keep the argument block out of scratch memory
            if (tid == 0) s_desc[0] = A.inl[0];
            if (tid == 1) s_desc[1] = A.inl[1];
            if (tid == 2) s_desc[2] = A.inl[2];
            if (tid == 3) s_desc[3] = A.inl[3];
        } else {
            s_desc[tid] = A.meshes[tid];
        }
    }
    __syncthreads();
    if (tid < MESH_LDS && tid < A.n_meshes) s_ibase[tid] = s_desc[tid].ibase;
    if (tid == 0) {
        s_n = 0;
        s_cn = 0;
        s_idle_hint = 1;
    }
    __syncthreads();

    u64 n_tests = 0, n_cands = 0;
    uint32_t n_fp64 = 0, n_adj = 0;
    bool idle = true;  // every workgroup starts without work (the seed counts them all idle)
    const u64 t0 = wall_clock64();
    // trace counters (thread 0; only written out when A.trace is set)
    u64 tr[OI_TRACE_WORDS];
#pragma unroll
    for (int k = 0; k < OI_TRACE_WORDS; k++) tr[k] = 0;
    const bool tracing = A.trace != nullptr;
    u64 tmark = t0;
#define TRACE_T(slot)                                       \
    if (tracing && tid == 0) {                              \
        const u64 now_ = wall_clock64();                    \
        tr[slot] += now_ - tmark;                           \
        tmark = now_;                                       \
    }
#define TRACE_N(slot, v) \
    if (tracing && tid == 0) tr[slot] += (v);
    const uint32_t cap = 1u << A.ring_shift;
    const u64 ring_base = ld64(&ctl->base);

    const uint32_t fan = 1u << (2u * A.expand);  // worst-case children per entry
    uint32_t backoff = 1;

    // Move the bottom D entries of the LDS stack to the ring (the oldest
    // entries: the largest subtrees, the most useful work for another
    // workgroup) and shift the rest down. Called by every thread with
    // uniform D; s_push_base holds the reserved ring index.
    auto move_bottom = [&](uint32_t D) {
        const uint32_t m = s_n;
        for (uint32_t i = tid; i < D; i += TB)
            slot_put(A.ring, A.ring_shift, s_push_base + i, s_stack[i]);
        __syncthreads();
        // chunk r's destination lies below every later chunk's source
        for (uint32_t r = 0; D + r * TB < m; r++) {
            const uint32_t i = D + r * TB + tid;
            uint2 x = make_uint2(0, 0);
            if (i < m) x = s_stack[i];
            __syncthreads();
            if (i < m) s_stack[i - D] = x;
            __syncthreads();
        }
        if (tid == 0) s_n = m - D;
        __syncthreads();
    };

    for (;;) {
        // ---- refill from the ring when the local stack is short. `pending`
        // counts ring entries plus busy workgroups: a workgroup that turns
        // busy adds itself before it takes entries out of the ring, so the
        // count never drops to zero while work exists (a stale low read only
        // lets an idle workgroup leave early, which loses no work: no
        // workgroup leaves while it holds entries or sees a non-empty ring).
        if (tid == 0) {
            s_pop_n = 0;
            s_flag = 0;
            if (s_n < TB) {
                uint32_t want = TB - s_n;
                const u64 h = ld64(&ctl->head);
                const u64 t = ld64(&ctl->tail);
                if (h < t) {
                    const u64 avail = t - h;
                    uint32_t take = (avail < want) ? (uint32_t)avail : want;
                    // an idle workgroup takes its share of a short ring
                    if (s_n == 0) {
                        const uint32_t share = (uint32_t)(avail / (u64)(s_idle_hint > 0 ? s_idle_hint : 1u));
                        const uint32_t lim = share < 32u ? 32u : share;
                        if (take > lim) take = lim;
                    }
                    if (idle) __hip_atomic_fetch_add(&ctl->pending, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    u64 expected = h;
                    if (__hip_atomic_compare_exchange_strong(&ctl->head, &expected, h + take, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        s_pop_h = h;
                        s_pop_n = take;
                        __hip_atomic_fetch_sub(&ctl->pending, (int)take, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    } else if (idle) {
                        __hip_atomic_fetch_sub(&ctl->pending, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t pn = s_pop_n;
        if (pn) {
            TRACE_N(OI_TR_POPS, 1);
            TRACE_N(OI_TR_POPPED, pn);
            const u64 h = s_pop_h;
            for (uint32_t i = tid; i < pn; i += TB) {
                uint2 v;
                if (!slot_take(A.ring, A.ring_shift, h + i, v, ctl, t0, A.timeout_ticks))
                    v = make_uint2(OI_BAD_ENTRY, OI_BAD_ENTRY);
                s_stack[s_n + i] = v;
            }
            __syncthreads();
            if (tid == 0) s_n += pn;
            __syncthreads();
        }
        TRACE_T(idle ? OI_TR_T_IDLE : OI_TR_T_REFILL);
        const uint32_t n = s_n;
        if (n == 0) {
            // ---- idle: finished when nothing is pending anywhere
            if (tid == 0) {
                if (!idle) {
                    // busy -> idle
                    __hip_atomic_fetch_sub(&ctl->pending, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(&ctl->idle, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                const int pend = ld32i(&ctl->pending);
                const uint32_t err = ld32(&ctl->error);
                s_idle_hint = ld32(&ctl->idle);
                uint32_t f = 0;
                if (pend <= 0 || err) f = 1;
                else if (wall_clock64() - t0 > A.timeout_ticks) {
                    __hip_atomic_fetch_or(&ctl->error, OI_ERR_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    f = 1;
                }
                s_flag = f;
            }
            idle = true;
            __syncthreads();
            TRACE_N(OI_TR_IDLE_LOOPS, 1);
            if (s_flag) break;
            // exponential back-off keeps idle workgroups off the shared counters
            for (uint32_t b = 0; b < backoff; b++) __builtin_amdgcn_s_sleep(32);
            if (backoff < 4) backoff <<= 1;
            continue;
        }
        backoff = 1;
        if (idle) {
            // idle -> busy (its +1 on `pending` preceded the pop)
            if (tid == 0) __hip_atomic_fetch_sub(&ctl->idle, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tracing && tid == 0 && tr[OI_TR_T_FIRST] == 0) tr[OI_TR_T_FIRST] = wall_clock64() - t0;
            idle = false;
        }
        if (tracing && tid == 0 && n > tr[OI_TR_MAX_STACK]) tr[OI_TR_MAX_STACK] = n;

        // ---- a stack too full to expand even one entry gives half to the ring
        if (STACK_CAP - n < fan - 1u) {
            if (tid == 0) {
                u64 b = 0;
                __hip_atomic_fetch_add(&ctl->pending, (int)(n / 2), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_flag = ring_reserve(ctl, ring_base, cap, n / 2, b) ? 0u : 1u;
                if (s_flag)
                    __hip_atomic_fetch_or(&ctl->error, OI_ERR_QUEUE_OVERFLOW, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                s_push_base = b;
            }
            __syncthreads();
            if (s_flag) break;
            move_bottom(n / 2);
            TRACE_N(OI_TR_MOVES, 1);
            TRACE_N(OI_TR_MOVED, n / 2);
            TRACE_T(OI_TR_T_MOVE);
            continue;
        }

        // idle count of the other workgroups, loaded now, used after the step
        uint32_t idle_now = 0;
        if (tid == 0) idle_now = ld32(&ctl->idle);

        // ---- take as many entries from the top as the stack can absorb the
        // children of (take * fan <= free slots + take), at most one per thread
        const uint32_t fit = (STACK_CAP - n) / (fan - 1u);
        uint32_t take = (n < TB) ? n : TB;
        if (take > fit) take = fit;
        bool have = tid < take;
        uint2 e = have ? s_stack[n - take + tid] : make_uint2(0, 0);
        __syncthreads();
        const uint32_t nb = n - take;

        // ---- test
        uint32_t nchild = 0;
        bool cand = false;
        Side sa = {0, 0}, sb = {0, 0};
        bool selfnode = false;
        uint32_t baseA = 0, baseB = 0;
        if (have && (e.x >= A.implicit_end || e.y >= A.implicit_end)) {
            // an entry dropped after a failed hand-over: never dereference it
            __hip_atomic_fetch_or(&ctl->error, OI_ERR_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            have = false;
        }
        if (have) {
            const uint32_t ma = find_mesh(A.meshes, s_ibase, A.n_meshes, e.x);
            const uint32_t mb = find_mesh(A.meshes, s_ibase, A.n_meshes, e.y);
            const MeshRef ra = mesh_ref(A.meshes, s_desc, ma), rb = mesh_ref(A.meshes, s_desc, mb);
            baseA = ra.ibase;
            baseB = rb.ibase;
            const uint32_t ia = e.x - ra.ibase, ib = e.y - rb.ibase;
            const uint32_t la = oi_ilog2(ia + 1u), lb = oi_ilog2(ib + 1u);
            selfnode = (e.x == e.y);
            bool ov = true;
            if (!selfnode) {
                float ba[6], bb[6];
                load_box(ra.nodes, ia - oi_virt_above(la, ra.L, ra.vl), ba);
                load_box(rb.nodes, ib - oi_virt_above(lb, rb.L, rb.vl), bb);
                ov = overlap6(ba, bb);
            }
            if (ov) {
                const bool leafA = (la == ra.L), leafB = (lb == rb.L);
                if (leafA && leafB) {
                    cand = !selfnode;
                } else {
                    sa = expand_side(ia, la, ra.L, ra.vl, A.expand);
                    sb = expand_side(ib, lb, rb.L, rb.vl, A.expand);
                    nchild = selfnode ? sa.count * (sa.count + 1u) / 2u : sa.count * sb.count;
                }
            }
        }
        n_tests += have ? 1u : 0u;

        // ---- children onto the stack (block prefix sum; they always fit)
        uint32_t T;
        const uint32_t off = block_scan(nchild, s_wsum, T);
        for (uint32_t c = 0; c < nchild; c++) {
            uint32_t x, y;
            if (selfnode) {
                tri_pair(c, sa.count, x, y);
                x = sa.start + x;
                y = sa.start + y;
            } else {
                x = sa.start + c / sb.count;
                y = sb.start + c % sb.count;
            }
            s_stack[nb + off + c] = make_uint2(baseA + x, baseB + y);
        }
        // ---- candidates into the LDS buffer
        uint32_t CT;
        const uint32_t coff = block_scan(cand ? 1u : 0u, s_wsum, CT);
        if (cand) s_cand[s_cn + coff] = e;
        n_cands += cand ? 1u : 0u;
        __syncthreads();
        if (tid == 0) {
            s_n = nb + T;
            s_cn += CT;
        }
        __syncthreads();
        TRACE_N(OI_TR_STEPS, 1);
        TRACE_N(OI_TR_TESTS, take);
        TRACE_T(OI_TR_T_STEP);

        // ---- narrow phase on a full batch of candidates
        if (s_cn >= TB) {
            const uint32_t cn = s_cn;
            uint2 c = s_cand[cn - TB + tid];
            bool hit;
            uint4 out;
            narrow_one<FP64>(A, s_ibase, s_desc, c, true, hit, out, n_fp64, n_adj);
            uint32_t HT;
            const uint32_t hoff = block_scan(hit ? 1u : 0u, s_wsum, HT);
            if (tid == 0) {
                s_cn = cn - TB;
                s_pop_h = HT ? __hip_atomic_fetch_add(&ctl->n_hits, HT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : 0;
            }
            __syncthreads();
            if (hit) {
                const uint32_t pos = (uint32_t)s_pop_h + hoff;
                if (pos < A.hits_cap)
                    A.hits[pos] = out;
                else
                    __hip_atomic_fetch_or(&ctl->error, OI_ERR_HITS_OVERFLOW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            TRACE_N(OI_TR_NARROW, 1);
            TRACE_T(OI_TR_T_NARROW);
        }

        // ---- while other workgroups are idle, everything beyond the next
        // batch goes to the ring (oldest entries: the largest subtrees)
        if (tid == 0) {
            s_flag = 0;
            s_idle_hint = idle_now;
            if (idle_now > 0u && s_n > TB) {
                const uint32_t D = s_n - TB;
                u64 b = 0;
                __hip_atomic_fetch_add(&ctl->pending, (int)D, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (ring_reserve(ctl, ring_base, cap, D, b)) {
                    s_push_base = b;
                    s_flag = D;
                } else {
                    __hip_atomic_fetch_sub(&ctl->pending, (int)D, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        __syncthreads();
        if (s_flag) {
            TRACE_N(OI_TR_MOVES, 1);
            TRACE_N(OI_TR_MOVED, s_flag);
            move_bottom(s_flag);
            TRACE_T(OI_TR_T_MOVE);
        }
    }

    // ---- drain the remaining candidates (fewer than one batch)
    __syncthreads();
    const uint32_t cn = s_cn;
    if (cn) {
        bool hit = false;
        uint4 out;
        uint2 c = (tid < cn) ? s_cand[tid] : make_uint2(0, 0);
        narrow_one<FP64>(A, s_ibase, s_desc, c, tid < cn, hit, out, n_fp64, n_adj);
        uint32_t HT;
        const uint32_t hoff = block_scan(hit ? 1u : 0u, s_wsum, HT);
        if (tid == 0)
            s_pop_h = HT ? __hip_atomic_fetch_add(&ctl->n_hits, HT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        __syncthreads();
        if (hit) {
            const uint32_t pos = (uint32_t)s_pop_h + hoff;
            if (pos < A.hits_cap)
                A.hits[pos] = out;
            else
                __hip_atomic_fetch_or(&ctl->error, OI_ERR_HITS_OVERFLOW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    if (tracing && tid == 0) {
        tr[OI_TR_T_EXIT] = wall_clock64() - t0;
        for (int k = 0; k < OI_TRACE_WORDS; k++) A.trace[(u64)blockIdx.x * OI_TRACE_WORDS + k] = tr[k];
    }
#undef TRACE_T
#undef TRACE_N
    // ---- statistics (one atomic per workgroup per counter)
    uint32_t t_;
    const uint32_t f64tot = block_scan(n_fp64, s_wsum, t_);
    (void)f64tot;
    if (t_) {
        if (tid == 0) __hip_atomic_fetch_add(&ctl->fp64, (u64)t_, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint32_t a_;
    block_scan(n_adj, s_wsum, a_);
    if (tid == 0) {
        if (a_) __hip_atomic_fetch_add(&ctl->adj, (u64)a_, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    uint32_t te, ce_;
    block_scan((uint32_t)n_tests, s_wsum, te);
    block_scan((uint32_t)n_cands, s_wsum, ce_);
    if (tid == 0) {
        __hip_atomic_fetch_add(&ctl->tests, (u64)te, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&ctl->cands, (u64)ce_, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&ctl->exited, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// seed: place the initial front at the ring's current head (after a clean
// traversal head == tail and every slot waits for its next lap) and reset the
// counters. Only block 0 / thread 0 writes the counters; every thread reads
// `head`, which no one writes during this kernel.
__global__ void k_seed(u64* ring, uint32_t shift, const uint2* front, uint32_t n0, oi_trav_ctl* ctl,
                       uint32_t reset_sticky, uint32_t n_wg)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 base = ctl->head;
    if (i < n0) {
        const u64 idx = base + i;
        ring[idx & ((1u << shift) - 1u)] = ring_pack(front[i].x, front[i].y, (uint32_t)(idx >> shift) & 15u);
    }
    if (i == 0) {
        // errors of the previous traversal move to `sticky` unless the caller
        // checked it synchronously
        ctl->sticky = reset_sticky ? 0u : (ctl->sticky | ctl->error);
        ctl->base = base;
        ctl->tail = base + n0;
        ctl->pending = (int)n0;
        ctl->n_hits = 0;
        ctl->error = 0;
        ctl->idle = n_wg;
        ctl->tests = 0;
        ctl->cands = 0;
        ctl->fp64 = 0;
        ctl->adj = 0;
        ctl->exited = 0;
    }
}

// ring reset (new ring, or after an aborted traversal): every slot EMPTY(0),
// head = tail = 0
__global__ void k_ring_reset(u64* ring, uint32_t n, oi_trav_ctl* ctl)
{
    const u64 e0 = ring_empty(0);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ring[i] = e0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->head = 0;
        ctl->tail = 0;
        ctl->base = 0;
    }
}

__global__ void k_fill64(u64* p, uint32_t n, u64 v)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

extern "C" {

hipError_t oi_launch_fill64(hipStream_t st, unsigned long long* p, uint32_t n, unsigned long long v)
{
    uint32_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_fill64, dim3(g), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t oi_launch_ring_reset(hipStream_t st, unsigned long long* ring, uint32_t cap, oi_trav_ctl* ctl)
{
    uint32_t g = (cap + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_ring_reset, dim3(g), dim3(256), 0, st, ring, cap, ctl);
    return hipGetLastError();
}

hipError_t oi_launch_seed(hipStream_t st, unsigned long long* ring, uint32_t cap, const uint2* front, uint32_t n0,
                          oi_trav_ctl* ctl, uint32_t reset_sticky, uint32_t n_wg)
{
    uint32_t g = (n0 + 255) / 256;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_seed, dim3(g), dim3(256), 0, st, ring, oi_ilog2(cap), front, n0, ctl, reset_sticky, n_wg);
    return hipGetLastError();
}

// resident workgroups per CU of the plain [0] and FP64-filter [1] traversal
hipError_t oi_traverse_occupancy(int* blocks_per_cu)
{
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu[0], k_traverse<false>, TB, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu[1], k_traverse<true>, TB, 0);
    return e;
}

hipError_t oi_launch_traverse(hipStream_t st, uint32_t grid, const oi_mesh_desc* meshes, const oi_mesh_desc* host_meshes,
                              uint32_t n_meshes,
                              unsigned long long* ring, uint32_t ring_cap, oi_trav_ctl* ctl,
                              uint4* hits, uint32_t hits_cap, uint32_t expand, uint32_t flags,
                              unsigned long long timeout_ticks, uint32_t implicit_end,
                              unsigned long long* trace)
{
    TravArgs a{};
    a.meshes = meshes;
    a.n_meshes = n_meshes;
    for (uint32_t m = 0; m < n_meshes && m < INLINE_MESHES; m++) a.inl[m] = host_meshes[m];
    a.ring = ring;
    a.ring_shift = oi_ilog2(ring_cap);
    a.ctl = ctl;
    a.hits = hits;
    a.hits_cap = hits_cap;
    a.expand = expand;
    a.flags = flags;
    a.timeout_ticks = timeout_ticks;
    a.implicit_end = implicit_end;
    a.trace = trace;
    if (flags & OI_DETECT_FP64_FALLBACK)
        hipLaunchKernelGGL(k_traverse<true>, dim3(grid), dim3(TB), 0, st, a);
    else
        hipLaunchKernelGGL(k_traverse<false>, dim3(grid), dim3(TB), 0, st, a);
    return hipGetLastError();
}

}  // extern "C"
