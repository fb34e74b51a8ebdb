// oibvh_build.hip -- gfx950 kernels for oibvh tree build and refit.
//
// Build (OibvhTree::build, oibvhTree.cu:237-388):
//   k_mesh_box     per-block vertex AABBs (the Morton quantisation box,
//                  Mesh::setupAABB mesh.cpp:91-98); also clears the sort's
//                  scratch counters
//   k_morton       folds the block boxes, then leaf AABB centroid -> 30-bit
//                  Morton key per face
//                  (calculate_aabb_and_morton_kernel, oibvh.cu:42-70) and the
//                  three 10-bit digit histograms of the keys (8-bit digits
//                  for 63-bit keys and batches)
//   k_sort_pass x3 onesweep LSD radix pass with decoupled look-back; stable,
//                  so the permutation equals thrust::stable_sort_by_key's
//                  (oibvhTree.cu:287-297)
//   k_reduce       leaves in sorted order + bottom-up AABB reduce, one subtree
//                  per workgroup; k_reduce_apex climbs the rest in one workgroup
//                  (oibvh_tree_construction_kernel2, oibvh.cu:156-219)
// Refit (OibvhTree::refit, oibvhTree.cu:193-235) is k_reduce alone over the
// already sorted faces (calculate_aabb_kernel + oibvh_tree_construction_kernel,
// oibvh.cu:24-40, 80-154).
//
// Every kernel is table-driven: a BuildTab names the meshes (oi_build_desc)
// and maps each workgroup to a (mesh, chunk) work item, so one launch per
// stage serves one tree or a batch of hundreds of trees. Vertex positions go
// through the mesh's model matrix when it has one (Mesh::transform,
// mesh.cpp:189-214), evaluated on the fly wherever a position is read.
//
// Compiled with -ffp-contract=off. min/max are compare-selects with glm's tie
// rules: min(a,b) = b < a ? b : a, max(a,b) = a < b ? b : a.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "oibvh_common.h"

#define OI_BLOCK 256
#define OI_MORTON_CHUNK 2048
#define OI_HIST_WORDS 3072  // k_morton's digit histograms: 8 x 256 or 3 x 1024 bins

__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ uint32_t ld_acq_relaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// block-wide exclusive scan of one uint32 per thread (256 threads, wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wsum, uint32_t& total)
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

// ---------------------------------------------------------------------------
// mesh tables
// ---------------------------------------------------------------------------
// Meshes come from a device table with a work list, or (d = null) travel in
// the kernel arguments: up to OI_INLINE_MESHES meshes, mesh m owning the
// workgroups [boff[m], boff[m + 1]).
#define OI_INLINE_MESHES 4
struct BuildTab {
    const oi_build_desc* d;  // device table, or null
    const uint2* work;       // workgroup -> (mesh, chunk) for the device table
    uint32_t n_inl;
    uint32_t boff[OI_INLINE_MESHES + 1];
    oi_build_desc inl[OI_INLINE_MESHES];
    __device__ __forceinline__ uint2 item() const { return item_at(blockIdx.x, gridDim.x); }
    // work item of virtual workgroup v of n (a launch of fewer workgroups
    // than items loops v over b, b + gridDim.x, ...: with gridDim.x a
    // multiple of 8, v stays on workgroup b's XCD)
    __device__ __forceinline__ uint2 item_at(uint32_t v, uint32_t n) const
    {
        const uint32_t w = oi_xcd_remap(v, n);
        if (work) return work[w];
        uint32_t m = 0;
#pragma unroll
        for (uint32_t k = 1; k < OI_INLINE_MESHES; k++)
            if (k < n_inl && w >= boff[k]) m = k;
        return make_uint2(m, w - boff[m]);
    }
    __device__ __forceinline__ oi_build_desc mesh(uint32_t m) const
    {
        if (d) return d[m];
        // constant indices keep the argument block out of scratch memory
        switch (m) {
        case 0: return inl[0];
        case 1: return inl[1];
        case 2: return inl[2];
        default: return inl[3];
        }
    }
};

// a mesh's model matrix, held in registers
struct Xf {
    float m[16];
    bool on;
    __device__ __forceinline__ void load(const float* xf)
    {
        on = xf != nullptr;
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = on ? oi_g(xf)[k] : 0.0f;
    }
    // a loaded position through the matrix (or as is)
    __device__ __forceinline__ void apply(const float* p, float* o) const
    {
        if (on) {
            oi_xform(m, p[0], p[1], p[2], o);
        } else {
            o[0] = p[0];
            o[1] = p[1];
            o[2] = p[2];
        }
    }
    __device__ __forceinline__ void vertex(const float* pos, uint32_t v, float* o) const
    {
        float p[3];
        oi_ld3(pos, v, p);
        apply(p, o);
    }
};

// ---------------------------------------------------------------------------
// k_mesh_box: vertex AABB with the sequential fold's exact result.
// The fold acc = glm::min(v, acc) keeps the later value on ties, so partial
// results combine as lo(L, R) = L < R ? L : R and hi(L, R) = R < L ? L : R in
// index order; +-inf are the identities, the default box (+-FLT_MAX) is folded
// in first, as aabb_box_t's initial value (utils.h:17-18).
// Work item (m, c): vertices [c * vchunk, (c + 1) * vchunk) of mesh m, partial
// box c of the mesh (at partials + 6 (pbase + c), or with partials null in the
// mesh's own box buffer at box + 6 + 6 c).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lo_c(float l, float r) { return (l < r) ? l : r; }
__device__ __forceinline__ float hi_c(float l, float r) { return (r < l) ? l : r; }

__global__ void __launch_bounds__(OI_BLOCK)
k_mesh_box(BuildTab T, float* __restrict__ partials, uint32_t* __restrict__ clear_u32, uint32_t clear_n)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = blockIdx.x * OI_BLOCK + tid; i < clear_n; i += gridDim.x * OI_BLOCK) clear_u32[i] = 0u;

    __shared__ float s_v[6][OI_BLOCK];
    const uint2 it = T.item();
    const oi_build_desc D = T.mesh(it.x);
    Xf X;
    X.load(D.xf);
    const float INF = __builtin_inff();
    float mn[3] = {INF, INF, INF}, mx[3] = {-INF, -INF, -INF};
    const uint32_t chunk = D.vchunk;
    const uint32_t b0 = it.y * chunk;
    const uint32_t b1 = min(D.V, b0 + chunk);
    const uint32_t sub = (chunk + OI_BLOCK - 1) / OI_BLOCK;
    const uint32_t s0 = b0 + tid * sub;
    const uint32_t s1 = min(b1, s0 + sub);
    // the thread's run in index order, eight vertices loaded per round trip
    // (positions past the run load vertex s0 and are not folded)
    for (uint32_t v0 = s0; v0 < s1; v0 += 8u) {
        float p[8][3];
#pragma unroll
        for (int j = 0; j < 8; j++) X.vertex(D.pos, v0 + j < s1 ? v0 + j : s0, p[j]);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (v0 + j >= s1) continue;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                mn[k] = lo_c(mn[k], p[j][k]);
                mx[k] = hi_c(mx[k], p[j][k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s_v[k][tid] = mn[k];
        s_v[3 + k][tid] = mx[k];
    }
    __syncthreads();
    for (uint32_t s = 1; s < OI_BLOCK; s <<= 1) {
        if ((tid & (2u * s - 1u)) == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                s_v[k][tid] = lo_c(s_v[k][tid], s_v[k][tid + s]);
                s_v[3 + k][tid] = hi_c(s_v[3 + k][tid], s_v[3 + k][tid + s]);
            }
        }
        __syncthreads();
    }
    float* P = partials ? partials + 6u * D.pbase : D.box + 6;
    if (tid < 6) oi_gw(P)[6u * it.y + tid] = s_v[tid][0];
}

// Fold the per-block partial boxes (at most OI_BLOCK of them) in block order and
// then with the default box (+-FLT_MAX) in front; every thread gets the result.
__device__ void fold_partials(const float* __restrict__ partials, uint32_t nparts, float (*s_v)[OI_BLOCK],
                              float* bmn, float* bmx)
{
    const uint32_t tid = threadIdx.x;
    const float INF = __builtin_inff();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        s_v[k][tid] = (tid < nparts) ? partials[6u * tid + k] : INF;
        s_v[3 + k][tid] = (tid < nparts) ? partials[6u * tid + 3 + k] : -INF;
    }
    __syncthreads();
    for (uint32_t s = 1; s < OI_BLOCK; s <<= 1) {
        if ((tid & (2u * s - 1u)) == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                s_v[k][tid] = lo_c(s_v[k][tid], s_v[k][tid + s]);
                s_v[3 + k][tid] = hi_c(s_v[3 + k][tid], s_v[3 + k][tid + s]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        bmn[k] = lo_c(__FLT_MAX__, s_v[k][0]);
        bmx[k] = hi_c(-__FLT_MAX__, s_v[3 + k][0]);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// k_morton: Morton30 of each face's leaf-box centroid (oibvh.cu:6-22, 42-70)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x30000ffu;
    x = (x | (x << 8)) & 0x300f00fu;
    x = (x | (x << 4)) & 0x30c30c3u;
    x = (x | (x << 2)) & 0x9249249u;
    return x;
}

// thrust::min(thrust::max(p * 1024, 0), 1023) -> unsigned; NaN -> 0
__device__ __forceinline__ uint32_t quantise10(float p)
{
    float q = p * 1024.0f;
    q = (q < 0.0f) ? 0.0f : q;
    q = (1023.0f < q) ? 1023.0f : q;
    return (q != q) ? 0u : (uint32_t)q;
}

// Work item (m, c): faces [c * OI_MORTON_CHUNK, (c + 1) * OI_MORTON_CHUNK) of
// mesh m. Keys go to keys[fbase + face]; the 8-bit digit histograms of the
// keys accumulate in ghist + m * hstride (hstride 0: one set for all keys of
// the launch); fmesh (if not null) receives the mesh id of each face
// (fmesh[fbase + face] = m). The mesh-box partials are read from partials +
// 6 pbase, or with partials null from the mesh's own box buffer (box + 6).
// 63-bit form (no reference counterpart): 21 bits per axis, quantised as
// min(max(p * 2^21, 0), 2^21 - 1), NaN -> 0
__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t quantise21(float p)
{
    float q = p * 2097152.0f;
    q = (q < 0.0f) ? 0.0f : q;
    q = (2097151.0f < q) ? 2097151.0f : q;
    return (q != q) ? 0u : (uint32_t)q;
}

// h[bin] += 1 for every active lane, one LDS atomic per run of consecutive
// active lanes with the same bin: consecutive faces lie close together, so
// the high digits of neighbouring lanes mostly agree, and atomics of a wave
// to one address would serialise.
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t bin, bool active)
{
    const uint32_t lane = threadIdx.x & 63u;
    // the bin of the lane below (lane 0 reads its own; it is always a head)
    const uint32_t below = (uint32_t)__shfl_up((int)bin, 1u, 64);
    const unsigned long long act = __ballot(active);
    const bool below_active = lane > 0 && ((act >> (lane - 1u)) & 1ull);
    const bool head = active && (!below_active || below != bin);
    const unsigned long long heads = __ballot(head);
    if (head) {
        // the run ends at the next head or the next inactive lane
        const unsigned long long above = (lane == 63u) ? 0ull : ~0ull << (lane + 1u);
        const unsigned long long stop = (heads | ~act) & above;
        const uint32_t end = stop ? (uint32_t)__builtin_ctzll(stop) : 64u;
        atomicAdd(&h[bin], end - lane);
    }
}

// keys_hi == null: 30-bit keys and the histograms of 4 8-bit digits (hbits 8)
// or of 3 10-bit digits (hbits 10); else 63-bit keys split into keys (low
// word) and keys_hi (high word) and 8 8-bit digit histograms. Digit d's
// histogram is ghist[d << hbits .. (d + 1) << hbits).
__global__ void __launch_bounds__(OI_BLOCK)
k_morton(BuildTab T, const float* __restrict__ partials, uint32_t* __restrict__ keys, uint32_t* __restrict__ keys_hi,
         uint32_t* __restrict__ ghist, uint32_t hstride, uint32_t* __restrict__ fmesh, uint32_t hbits)
{
    __shared__ uint32_t s_h[OI_HIST_WORDS];
    __shared__ float s_v[6][OI_BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint2 it = T.item();
    const oi_build_desc D = T.mesh(it.x);
    Xf X;
    X.load(D.xf);
    const bool h10 = hbits == 10u;
    const uint32_t nd = keys_hi ? 8u : h10 ? 3u : 4u;
    for (uint32_t i = tid; i < OI_HIST_WORDS; i += OI_BLOCK) s_h[i] = 0u;
    float bmn[3], bmx[3];
    fold_partials(partials ? partials + 6u * D.pbase : D.box + 6, D.nparts, s_v, bmn, bmx);
    ghist += it.x * hstride;
    if (it.y == 0 && tid < 3) {
        oi_gw(D.box)[tid] = bmn[tid];
        oi_gw(D.box)[3 + tid] = bmx[tid];
    }
    const uint32_t f0 = it.y * OI_MORTON_CHUNK;
    const uint32_t f1 = min(D.F, f0 + OI_MORTON_CHUNK);  // f0 < f1: a work item holds a face
    // a thread's faces i = f0 + tid + 256 j: every face load is issued, then
    // every vertex load, then the keys are formed (two rounds of memory
    // latency per thread instead of two per face); faces past f1 load face f0
    // and are dropped
    constexpr int NJ = OI_MORTON_CHUNK / OI_BLOCK;
    uint32_t fv[NJ][3];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const uint32_t i = f0 + tid + (uint32_t)j * OI_BLOCK;
        oi_ld3(D.faces, i < f1 ? i : f0, fv[j]);
    }
    float p[NJ][9];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        oi_ld3(D.pos, fv[j][0], &p[j][0]);
        oi_ld3(D.pos, fv[j][1], &p[j][3]);
        oi_ld3(D.pos, fv[j][2], &p[j][6]);
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const uint32_t i = f0 + tid + (uint32_t)j * OI_BLOCK;
        const bool live = i < f1;
        float a[3], b[3], c[3];
        X.apply(&p[j][0], a);
        X.apply(&p[j][3], b);
        X.apply(&p[j][6], c);
        float u[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float lo = gmin(gmin(a[k], b[k]), c[k]);
            const float hi = gmax(gmax(a[k], b[k]), c[k]);
            const float cen = (lo + hi) * 0.5f;
            const float off = cen - bmn[k];
            const float len = bmx[k] - bmn[k];
            u[k] = off / len;
        }
        uint32_t key, khi = 0;
        if (keys_hi) {
            const uint64_t k64 = (spread21(quantise21(u[0])) << 2) | (spread21(quantise21(u[1])) << 1) |
                                 spread21(quantise21(u[2]));
            key = (uint32_t)k64;
            khi = (uint32_t)(k64 >> 32);
            if (live) oi_gw(keys_hi)[D.fbase + i] = khi;
        } else {
            key = (spread10(quantise10(u[0])) << 2) | (spread10(quantise10(u[1])) << 1) | spread10(quantise10(u[2]));
        }
        if (live) {
            oi_gw(keys)[D.fbase + i] = key;
            if (fmesh) oi_gw(fmesh)[D.fbase + i] = it.x;
        }
        // every lane of the wave takes part (ballots), dead lanes add nothing
        if (h10) {
#pragma unroll
            for (int d = 0; d < 3; d++) hist_add(s_h + 1024 * d, (key >> (10 * d)) & 0x3ffu, live);
        } else {
#pragma unroll
            for (int d = 0; d < 4; d++) hist_add(s_h + 256 * d, (key >> (8 * d)) & 0xffu, live);
        }
        if (keys_hi) {
#pragma unroll
            for (int d = 0; d < 4; d++) hist_add(s_h + 256 * (4 + d), (khi >> (8 * d)) & 0xffu, live);
        }
    }
    __syncthreads();
    const uint32_t words = nd << hbits;
    for (uint32_t i = tid; i < words; i += OI_BLOCK) {
        const uint32_t c = s_h[i];
        if (c) atomicAdd(&ghist[i], c);
    }
}

// ---------------------------------------------------------------------------
// k_sort_pass<DB>: one stable LSD pass over a DB-bit digit (8 or 10),
// onesweep style, over up to
// OI_SORT_SEGS independent segments of the concatenated key arrays (one per
// tree when several trees are built in one launch).
// A tile is 16 waves x 8 rounds x 64 lanes = 8192 keys of one segment in
// input order (wave w owns keys [512w, 512w + 512) of the tile, round j lane l
// holds key 512w + 64j + l): fewer, larger tiles than one wave per SIMD would
// give shorten the look-back chains. Keys are ranked per wave with 8 ballots
// per round (DB ballots; stable: earlier rounds and lower lanes come first), staged
// through LDS in tile-local sorted order and written out in runs. Segment s
// owns the workgroups and tile ids [tile0[s], tile0[s + 1]); a workgroup takes
// its tile from its segment's atomic counter (tile_counter[s]), so every tile
// a look-back waits on is already running, and the look-back ends at the
// segment's own first tile.
// Look-back status word: bits 31:30 = 1 aggregate / 2 inclusive prefix,
// bits 29:0 = count. A digit that no key of the segment has (its histogram
// entry is 0) is neither published nor looked back on. Values default to the
// key's index in the concatenated arrays. With key_of_val the key of a value v
// is key_of_val[v] (a batch's mesh-id passes: a stable partition of the
// Morton-sorted faces of all meshes back into their meshes).
// ---------------------------------------------------------------------------
#define RS_WAVES 16
#define RS_THREADS (64 * RS_WAVES)
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define LB_AGG (1u << 30)
#define LB_PRE (2u << 30)
#define LB_CNT 0x3fffffffu
#define LB_WIN 8  // look-back statuses loaded per round trip
#define OI_SORT_SEGS 4

struct SortSegs {
    uint32_t ns;                        // segments, 1..OI_SORT_SEGS
    uint32_t hstride;                   // words between the segments' digit histograms
    uint32_t tile0[OI_SORT_SEGS + 1];   // segment s owns tiles [tile0[s], tile0[s + 1])
    uint32_t base[OI_SORT_SEGS];        // its first key in the concatenated arrays
    uint32_t n[OI_SORT_SEGS];           // its keys
};

// block-wide exclusive scan over NW waves (one uint32 per thread)
template <int NW>
__device__ __forceinline__ uint32_t block_excl_scan_w(uint32_t v, uint32_t* s_wsum, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = oi_wave_scan_incl(v);
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        const uint32_t t = s_wsum[w];
        before += (w < (int)wave) ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();
    return before + x - v;
}

template <int DB>
__global__ void __launch_bounds__(RS_THREADS)
k_sort_pass(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
            uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, SortSegs S, uint32_t shift,
            const uint32_t* __restrict__ ghist, uint32_t* lookback, uint32_t* lookback_next,
            uint32_t* tile_counter, const uint32_t* __restrict__ key_of_val)
{
    constexpr uint32_t NB = 1u << DB;  // bins; at most RS_THREADS
    static_assert(NB <= RS_THREADS, "one digit per thread");
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_whist[RS_WAVES][NB];
    __shared__ uint32_t s_lstart[NB];
    __shared__ uint32_t s_gbase[NB];
    __shared__ uint32_t s_wsum[RS_WAVES];
    __shared__ uint32_t s_keys[RS_TILE];
    __shared__ uint32_t s_vals[RS_TILE];

    // threads tid < NB also own digit tid
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool dig = tid < NB;
    // the workgroup's segment (constant indices keep the argument block out
    // of scratch memory)
    uint32_t seg = 0, seg0 = S.tile0[0], sbase = S.base[0], n = S.n[0];
#pragma unroll
    for (int k = 1; k < OI_SORT_SEGS; k++) {
        if (k < (int)S.ns && blockIdx.x >= S.tile0[k]) {
            seg = k;
            seg0 = S.tile0[k];
            sbase = S.base[k];
            n = S.n[k];
        }
    }
    if (tid == 0)
        s_tile = __hip_atomic_fetch_add(tile_counter + seg, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the histogram entry is loaded while the tile id is in flight
    const uint32_t hcnt = dig ? ghist[seg * S.hstride + tid] : 0u;
    for (uint32_t i = tid; i < RS_WAVES * NB; i += RS_THREADS) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t gt = seg0 + tile;  // tile id over all segments
    const uint32_t base = tile * RS_TILE;
    if (lookback_next && dig) lookback_next[gt * NB + tid] = 0u;

    // exclusive prefix of the segment's digit histogram
    uint32_t gtot;
    const uint32_t gexcl = block_excl_scan_w<RS_WAVES>(hcnt, s_wsum, gtot);

    uint32_t k[RS_ITEMS], v[RS_ITEMS], rk[RS_ITEMS];
    const uint32_t wbase = base + wave * (64u * RS_ITEMS);
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        const uint32_t idx = wbase + 64u * j + lane;
        const bool ok = idx < n;
        const uint32_t gi = sbase + idx;
        v[j] = ok ? (vals_in ? vals_in[gi] : gi) : 0u;
        k[j] = ok ? (key_of_val ? key_of_val[v[j]] : keys_in[gi]) : 0u;
    }
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        const uint32_t idx = wbase + 64u * j + lane;
        const bool ok = idx < n;
        const uint32_t d = (k[j] >> shift) & (NB - 1u);
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t prev = s_whist[wave][d];
        const uint32_t cnt = (uint32_t)__popcll(peers);
        rk[j] = prev + (uint32_t)__popcll(peers & lt);
        const bool leader = ok && ((peers & lt) == 0ull);
        if (leader) s_whist[wave][d] = prev + cnt;
    }
    __syncthreads();
    // per-wave exclusive offsets and tile totals (thread tid owns digit tid)
    uint32_t tot = 0;
    if (dig) {
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            const uint32_t c = s_whist[w][tid];
            s_whist[w][tid] = tot;
            tot += c;
        }
    }
    // publish this tile's aggregate (or inclusive prefix for the segment's
    // first tile)
    uint32_t* my = lookback + gt * NB + tid;
    const bool live = dig && hcnt != 0u;
    if (live) st_relaxed(my, (tile == 0 ? LB_PRE : LB_AGG) | tot);
    uint32_t ltot;
    const uint32_t lstart = block_excl_scan_w<RS_WAVES>(tot, s_wsum, ltot);
    if (dig) s_lstart[tid] = lstart;
    // decoupled look-back over the segment's preceding tiles for digit tid:
    // the statuses of LB_WIN tiles i, i - 1, ... are loaded in one round trip
    // (the loads cross XCDs) and consumed in order, summing aggregates until
    // an inclusive prefix; a tile not yet published ends the window, which is
    // loaded again from there. The segment's first tile publishes its prefix
    // at once, so every walk ends there at the latest (statuses loaded beyond
    // it are not consumed).
    uint32_t excl = 0;
    if (tile > 0 && live) {
        int32_t i = (int32_t)gt - 1;
        bool done = false;
        while (!done) {
            uint32_t st[LB_WIN];
#pragma unroll
            for (int w = 0; w < LB_WIN; w++) {
                const int32_t t = i - w;
                st[w] = t >= 0 ? ld_acq_relaxed(lookback + (uint32_t)t * NB + tid) : 0u;
            }
            int32_t used = LB_WIN;
#pragma unroll
            for (int w = 0; w < LB_WIN; w++) {
                if (used != LB_WIN || done) continue;
                const uint32_t fl = st[w] & ~LB_CNT;
                if (fl == 0u) {
                    used = w;
                } else {
                    excl += st[w] & LB_CNT;
                    if (fl == LB_PRE) done = true;
                }
            }
            if (done) break;
            i -= used;
            if (used != LB_WIN) __builtin_amdgcn_s_sleep(1);
        }
        st_relaxed(my, LB_PRE | (excl + tot));
    }
    if (dig) s_gbase[tid] = sbase + gexcl + excl;
    __syncthreads();
    // stage in tile-local sorted order
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        const uint32_t idx = wbase + 64u * j + lane;
        if (idx < n) {
            const uint32_t d = (k[j] >> shift) & (NB - 1u);
            const uint32_t lp = s_lstart[d] + s_whist[wave][d] + rk[j];
            s_keys[lp] = k[j];
            s_vals[lp] = v[j];
        }
    }
    __syncthreads();
    const uint32_t cnt_tile = (n - base < RS_TILE) ? (n - base) : RS_TILE;
    for (uint32_t i = tid; i < cnt_tile; i += RS_THREADS) {
        const uint32_t key = s_keys[i];
        const uint32_t d = (key >> shift) & (NB - 1u);
        const uint32_t g = s_gbase[d] + (i - s_lstart[d]);
        if (keys_out) keys_out[g] = key;
        vals_out[g] = s_vals[i];
    }
}

// ---------------------------------------------------------------------------
// k_reduce: leaves + bottom-up reduce.
// Stage 1: work item (m, w) owns the subtree of mesh m of height
// h = min(L, 8 + log2 M) rooted at level L - h, position w (its 2^h leaves).
// Each thread computes M consecutive leaves and folds them log2(M) levels in
// registers; the 256 results climb the remaining levels through LDS. Every
// level of the subtree is a contiguous run of real nodes: it is staged in LDS
// and written out as whole cache lines from consecutive lanes (one thread
// writing its own 24-byte nodes would leave lines partially written). Every
// real node is written once, at its real index.
// Stage 2 (k_reduce_apex, a second launch): one workgroup per mesh reads the
// stage-1 roots and climbs to the root in chunks of up to 2^11 nodes.
// A parent is real iff its left child is; with no real right child the
// parent equals the left child (oibvh.cu:139-146).
// ---------------------------------------------------------------------------
struct Box6 {
    float v[6];
};

__device__ __forceinline__ void box_merge_into(Box6& a, const Box6& r)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a.v[k] = gmin(a.v[k], r.v[k]);
        a.v[3 + k] = gmax(a.v[3 + k], r.v[3 + k]);
    }
}

__device__ __forceinline__ Box6 load_node(const float* nodes, uint32_t r)
{
    const OI_GLOBAL float* p = oi_g(nodes) + 6ull * r;
    Box6 o;
#pragma unroll
    for (int k = 0; k < 6; k++) o.v[k] = p[k];
    return o;
}

struct ReduceArgs {
    BuildTab T;
    const uint32_t* vals;  // build: Morton-sorted face ids (fbase + face); null: refit over sfaces
    uint32_t write_perm;   // build: also write perm (faces local to the mesh)
    uint32_t nwork;        // k_reduce_w: work items (0: one per workgroup)
};

struct LeafLoad {
    const oi_build_desc& D;
    const Xf& X;
    const uint32_t* vals;
    uint32_t write_perm;
    __device__ __forceinline__ Box6 operator()(uint32_t p) const
    {
        uint32_t f[3];
        if (vals) {
            const uint32_t fi = oi_g(vals)[D.fbase + p] - D.fbase;
            oi_ld3(D.faces, fi, f);
            OI_GLOBAL uint32_t* q = oi_gw(D.sfaces) + 3ull * p;
            q[0] = f[0];
            q[1] = f[1];
            q[2] = f[2];
            if (write_perm) oi_gw(D.perm)[p] = fi;
        } else {
            oi_ld3(D.sfaces, p, f);
        }
        float a[3], b[3], c[3];
        X.vertex(D.pos, f[0], a);
        X.vertex(D.pos, f[1], b);
        X.vertex(D.pos, f[2], c);
        Box6 r;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            r.v[k] = gmin(gmin(a[k], b[k]), c[k]);
            r.v[3 + k] = gmax(gmax(a[k], b[k]), c[k]);
        }
        return r;
    }
};

struct NodeLoad {
    const float* nodes;
    uint32_t lvl_start;
    __device__ __forceinline__ Box6 operator()(uint32_t p) const { return load_node(nodes, lvl_start + p); }
};

// LDS staging of one level's nodes in node layout (6 floats per node)
__device__ __forceinline__ void put_box(float* stage, uint32_t i, const Box6& b)
{
    float2* d = reinterpret_cast<float2*>(stage + 6u * i);
    d[0] = make_float2(b.v[0], b.v[1]);
    d[1] = make_float2(b.v[2], b.v[3]);
    d[2] = make_float2(b.v[4], b.v[5]);
}
__device__ __forceinline__ Box6 get_box(const float* stage, uint32_t i)
{
    const float2* d = reinterpret_cast<const float2*>(stage + 6u * i);
    const float2 a = d[0], b = d[1], c = d[2];
    Box6 o;
    o.v[0] = a.x;
    o.v[1] = a.y;
    o.v[2] = b.x;
    o.v[3] = b.y;
    o.v[4] = c.x;
    o.v[5] = c.y;
    return o;
}

// Write `count` staged nodes to real index `first` onwards as one contiguous
// run of float2 (whole cache lines from consecutive lanes); the block's
// threads share the copy.
__device__ __forceinline__ void copy_out(float* nodes, uint32_t first, uint32_t count, const float* stage)
{
    OI_GLOBAL float* dst = oi_gw(nodes) + 6ull * first;
    const float2* src = reinterpret_cast<const float2*>(stage);
    for (uint32_t i = threadIdx.x; i < 3u * count; i += OI_BLOCK) {
        const float2 v = src[i];
        dst[2u * i] = v.x;
        dst[2u * i + 1u] = v.y;
    }
}

// LDS for one subtree reduce with M leaves per thread: two staging buffers
// used in turn, 256 M and max(128 M, 256) nodes
template <int M>
struct StageSize {
    static constexpr uint32_t n0 = 256u * M;
    static constexpr uint32_t n1 = (128u * M > 256u) ? 128u * M : 256u;
    static constexpr uint32_t floats = 6u * (n0 + n1);
};

// Reduce the subtree whose bottom level is lc, h levels up, root position w
// at level lc - h. Requires 2^h <= 256 * M and stage >= StageSize<M>::floats.
// Each level is staged in LDS and written out contiguously; consecutive
// levels use the two buffers in turn, so one barrier per level orders both
// the next level's reads and the previous buffer's reuse.
// Returns with every node of the subtree written; ends with a barrier.
template <int M, class Load>
__device__ void reduce_subtree(const Load& load, float* nodes, uint32_t L, uint32_t vl, uint32_t lc, uint32_t h,
                               uint32_t w, float* stage)
{
    constexpr uint32_t LGM = (M == 1) ? 0 : (M == 2) ? 1 : (M == 4) ? 2 : 3;
    float* const s0 = stage;
    float* const s1 = stage + 6u * StageSize<M>::n0;
    const uint32_t tid = threadIdx.x;
    const bool act = (tid * M) < (1u << h);
    const uint32_t c0 = (w << h) + tid * M;
    uint32_t ncnt = oi_lvl_cnt(lc, L, vl);
    Box6 b[M];
#pragma unroll
    for (int j = 0; j < M; j++)
        if (act && c0 + j < ncnt) {
            b[j] = load(c0 + j);
            put_box(s0, tid * M + j, b[j]);
        }
    __syncthreads();
    copy_out(nodes, oi_lvl_start(lc, L, vl) + (w << h), min(1u << h, ncnt - (w << h)), s0);
    uint32_t cur = 0;  // buffer holding the last level
    const uint32_t hl = (h < LGM) ? h : LGM;
    uint32_t l = lc;
#pragma unroll
    for (uint32_t s = 1; s <= LGM; s++) {
        if (s <= hl) {
            const uint32_t cc = ncnt;  // real count of the child level
            l -= 1;
            ncnt = oi_lvl_cnt(l, L, vl);
            cur ^= 1u;
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)(M >> s); j++) {
                const uint32_t q = (c0 >> s) + j;
                if (act && 2u * q < cc) {
                    Box6 x = b[2 * j];
                    if (2u * q + 1u < cc) box_merge_into(x, b[2 * j + 1]);
                    b[j] = x;
                    put_box(cur ? s1 : s0, tid * (uint32_t)(M >> s) + j, x);
                }
            }
            __syncthreads();
            const uint32_t f = w << (h - s);
            copy_out(nodes, oi_lvl_start(l, L, vl) + f, min(1u << (h - s), ncnt - f), cur ? s1 : s0);
        }
    }
    const uint32_t hs = h - hl;
    // LDS climb: slot t of the current buffer holds node (w << hs) + t of level l
    for (uint32_t s = 1; s <= hs; s++) {
        const uint32_t cc = ncnt;
        l -= 1;
        ncnt = oi_lvl_cnt(l, L, vl);
        const uint32_t width = 1u << (hs - s);
        const uint32_t f = w << (hs - s);
        const uint32_t q = f + tid;
        const float* src = cur ? s1 : s0;
        float* dst = cur ? s0 : s1;
        if (tid < width && 2u * q < cc) {
            Box6 x = get_box(src, 2u * tid);
            if (2u * q + 1u < cc) box_merge_into(x, get_box(src, 2u * tid + 1u));
            put_box(dst, tid, x);
        }
        cur ^= 1u;
        __syncthreads();
        copy_out(nodes, oi_lvl_start(l, L, vl) + f, min(width, ncnt - f), dst);
    }
    __syncthreads();
}

template <int M>
__global__ void __launch_bounds__(OI_BLOCK) k_reduce(ReduceArgs a)
{
    __shared__ float stage[StageSize<M>::floats];
    constexpr uint32_t LGM = (M == 1) ? 0 : (M == 2) ? 1 : (M == 4) ? 2 : 3;
    const uint2 it = a.T.item();
    const oi_build_desc D = a.T.mesh(it.x);
    Xf X;
    X.load(D.xf);
    const uint32_t L = D.L, vl = D.vl;
    const uint32_t h = (L < 8 + LGM) ? L : 8 + LGM;
    LeafLoad ll{D, X, a.vals, a.write_perm};
    reduce_subtree<M>(ll, D.nodes, L, vl, L, h, it.y, stage);
}

// ---------------------------------------------------------------------------
// k_reduce_w: stage 1 with one subtree per wave (the alternative to
// k_reduce's one subtree per workgroup). Wave ws owns the 2^hw leaves
// [ws 2^hw, (ws + 1) 2^hw) of the leaf level, hw = min(L, 8); lane l holds
// leaves 64k + l (k < 4), so every face, vertex and node access of one
// instruction covers consecutive leaves. Each level is staged in the wave's
// own LDS and written out as a contiguous run; the wave climbs hw levels
// with no workgroup barrier. The workgroup's waves then meet once, and
// thread 0 merges their roots hb = min(L - hw, 2) levels further, so stage 1
// ends at level L - hw - hb as k_reduce with M = 4 does (apex lgm = 2).
// ---------------------------------------------------------------------------
#define RW_LG 8
#define RW_W (1u << RW_LG)

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Write `count` staged nodes (at most CMAX) to real index `first` onwards as
// one contiguous run of float2 per lane; all LDS reads are issued before the
// first store.
template <uint32_t CMAX>
__device__ __forceinline__ void wave_copy_out(float* nodes, uint32_t first, uint32_t count, const float* stage,
                                              uint32_t lane)
{
    constexpr uint32_t IT = (3u * CMAX + 63u) / 64u;
    OI_GLOBAL float* dst = oi_gw(nodes) + 6ull * first;
    const float2* src = reinterpret_cast<const float2*>(stage);
    const uint32_t n = 3u * count;
    float2 v[IT];
#pragma unroll
    for (uint32_t it = 0; it < IT; it++) {
        const uint32_t i = lane + 64u * it;
        if (i < n) v[it] = src[i];
    }
#pragma unroll
    for (uint32_t it = 0; it < IT; it++) {
        const uint32_t i = lane + 64u * it;
        if (i < n) {
            dst[2u * i] = v[it].x;
            dst[2u * i + 1u] = v[it].y;
        }
    }
}

// One level of a wave climb in place: slots 2t, 2t + 1 of stage (the child
// level, first node 2f, cc real) -> slot t (first node f). Every read of the
// level precedes every write.
template <uint32_t WMAX>
__device__ __forceinline__ void wave_level(float* stage, uint32_t width, uint32_t f, uint32_t cc, uint32_t lane)
{
    constexpr uint32_t IT = (WMAX + 63u) / 64u;
    Box6 x[IT], y[IT];
    bool has[IT], two[IT];
#pragma unroll
    for (uint32_t it = 0; it < IT; it++) {
        const uint32_t t = lane + 64u * it;
        const uint32_t q = f + t;
        has[it] = t < width && 2u * q < cc;
        two[it] = has[it] && 2u * q + 1u < cc;
        if (has[it]) x[it] = get_box(stage, 2u * t);
        if (two[it]) y[it] = get_box(stage, 2u * t + 1u);
    }
    wave_sync();
#pragma unroll
    for (uint32_t it = 0; it < IT; it++) {
        if (two[it]) box_merge_into(x[it], y[it]);
        if (has[it]) put_box(stage, lane + 64u * it, x[it]);
    }
    wave_sync();
}

// the refit's sorted faces of the leaves a wave owns
__device__ __forceinline__ void wave_sfaces(const oi_build_desc& D, uint32_t f0, uint32_t hw, uint32_t nleaf,
                                            uint32_t lane, uint32_t (&f)[RW_W / 64][3])
{
#pragma unroll
    for (int k = 0; k < RW_W / 64; k++) {
        const uint32_t off = 64u * k + lane;
        const uint32_t p = f0 + off;
        const bool ok = off < (1u << hw) && p < nleaf;
        oi_ld3(D.sfaces, ok ? p : f0, f[k]);  // in range; only ok results are kept
    }
}

// the leaf boxes of a wave's faces f (vertex gathers, model matrix if XF),
// staged at their leaf slots
template <bool XF>
__device__ __forceinline__ void wave_boxes(const oi_build_desc& D, uint32_t f0, uint32_t hw, uint32_t nleaf,
                                           uint32_t lane, float* stage, const uint32_t (&f)[RW_W / 64][3])
{
    constexpr int K = RW_W / 64;
    float m[16];
    if constexpr (XF) {
#pragma unroll
        for (int k = 0; k < 16; k++) m[k] = oi_g(D.xf)[k];
    }
    float v[K][3][3];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if constexpr (XF) {
                float q[3];
                oi_ld3(D.pos, f[k][j], q);
                oi_xform(m, q[0], q[1], q[2], v[k][j]);
            } else {
                oi_ld3(D.pos, f[k][j], v[k][j]);
            }
        }
#pragma unroll
    for (int k = 0; k < K; k++) {
        Box6 r;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            r.v[c] = gmin(gmin(v[k][0][c], v[k][1][c]), v[k][2][c]);
            r.v[3 + c] = gmax(gmax(v[k][0][c], v[k][1][c]), v[k][2][c]);
        }
        const uint32_t off = 64u * k + lane;
        if (off < (1u << hw) && f0 + off < nleaf) put_box(stage, off, r);
    }
}

// The leaf level of one wave subtree into stage (node layout): every load is
// issued before any box is formed. BUILD reads the sorted face ids and writes
// sfaces and perm; XF applies the model matrix. Both are compile-time, so the
// loads form one straight run.
template <bool BUILD, bool XF>
__device__ __forceinline__ void wave_leaves(const ReduceArgs& A, const oi_build_desc& D, uint32_t f0, uint32_t hw,
                                            uint32_t nleaf, uint32_t lane, float* stage)
{
    constexpr int K = RW_W / 64;
    uint32_t f[K][3], fid[K];
    if constexpr (BUILD) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t off = 64u * k + lane;
            const uint32_t p = f0 + off;
            const uint32_t pc = (off < (1u << hw) && p < nleaf) ? p : f0;  // only in-range results are kept
            fid[k] = oi_g(A.vals)[D.fbase + pc] - D.fbase;
            oi_ld3(D.faces, fid[k], f[k]);
        }
    } else {
        wave_sfaces(D, f0, hw, nleaf, lane, f);
    }
    wave_boxes<XF>(D, f0, hw, nleaf, lane, stage, f);
    if constexpr (BUILD) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t off = 64u * k + lane;
            const uint32_t p = f0 + off;
            if (off < (1u << hw) && p < nleaf) {
                OI_GLOBAL uint32_t* q = oi_gw(D.sfaces) + 3ull * p;
                q[0] = f[k][0];
                q[1] = f[k][1];
                q[2] = f[k][2];
                if (A.write_perm) oi_gw(D.perm)[p] = fid[k];
            }
        }
    }
}

// climb levels S..8 of a full-height (hw = 8) wave subtree; level s has
// 256 >> s nodes
template <uint32_t S>
__device__ __forceinline__ void wave_climb8(float* stage, float* nodes, uint32_t L, uint32_t vl, uint32_t f0,
                                            uint32_t cc, uint32_t lane)
{
    if constexpr (S <= RW_LG) {
        constexpr uint32_t WIDTH = RW_W >> S;
        const uint32_t l = L - S;
        const uint32_t ncnt = oi_lvl_cnt(l, L, vl);
        const uint32_t f = f0 >> S;
        wave_level<WIDTH>(stage, WIDTH, f, cc, lane);
        wave_copy_out<WIDTH>(nodes, oi_lvl_start(l, L, vl) + f, min(WIDTH, ncnt - f), stage, lane);
        wave_climb8<S + 1>(stage, nodes, L, vl, f0, ncnt, lane);
    }
}

// a wave's part of one work item once its leaf boxes are staged: write the
// leaves, climb hw levels (wave 0 .. 2^hb - 1 of the workgroup)
__device__ __forceinline__ void wave_climb_all(const oi_build_desc& D, uint32_t f0, uint32_t hw, uint32_t nleaf,
                                               uint32_t lane, float* W)
{
    const uint32_t L = D.L, vl = D.vl;
    wave_sync();
    wave_copy_out<RW_W>(D.nodes, oi_lvl_start(L, L, vl) + f0, min(1u << hw, nleaf - f0), W, lane);
    if (hw == RW_LG) {
        wave_climb8<1>(W, D.nodes, L, vl, f0, nleaf, lane);
    } else {
        // short trees: the same climb with run-time widths
        uint32_t ncnt = nleaf;
        for (uint32_t s = 1; s <= hw; s++) {
            const uint32_t l = L - s;
            const uint32_t cc = ncnt;
            ncnt = oi_lvl_cnt(l, L, vl);
            const uint32_t width = 1u << (hw - s);
            const uint32_t f = f0 >> s;
            wave_level<RW_W / 2>(W, width, f, cc, lane);
            wave_copy_out<RW_W / 2>(D.nodes, oi_lvl_start(l, L, vl) + f, min(width, ncnt - f), W, lane);
        }
    }
}

// thread 0 merges the wave roots of work item `it` (slot 0 of each wave's
// stage) hb levels further; called by every thread after a barrier
__device__ __forceinline__ void item_top(const oi_build_desc& D, uint2 it, uint32_t hw, uint32_t hb,
                                         float (*S)[6 * RW_W])
{
    if (threadIdx.x != 0) return;
    const uint32_t L = D.L, vl = D.vl;
    const uint32_t lw = L - hw;
    uint32_t cc = oi_lvl_cnt(lw, L, vl);
    Box6 m[2];
    const uint32_t n1 = 1u << (hb - 1u);  // nodes of the first merged level
#pragma unroll
    for (uint32_t t = 0; t < 2u; t++) {
        const uint32_t q = (it.y << (hb - 1u)) + t;
        if (t < n1 && 2u * q < cc) {
            Box6 x = get_box(S[2u * t], 0);
            if (2u * q + 1u < cc) box_merge_into(x, get_box(S[2u * t + 1u], 0));
            m[t] = x;
            OI_GLOBAL float* o = oi_gw(D.nodes) + 6ull * (oi_lvl_start(lw - 1u, L, vl) + q);
#pragma unroll
            for (int c = 0; c < 6; c++) o[c] = x.v[c];
        }
    }
    if (hb == 2) {
        cc = oi_lvl_cnt(lw - 1u, L, vl);
        const uint32_t q = it.y;
        if (2u * q < cc) {
            Box6 x = m[0];
            if (2u * q + 1u < cc) box_merge_into(x, m[1]);
            OI_GLOBAL float* o = oi_gw(D.nodes) + 6ull * (oi_lvl_start(lw - 2u, L, vl) + q);
#pragma unroll
            for (int c = 0; c < 6; c++) o[c] = x.v[c];
        }
    }
}

// what one wave of a workgroup does for work item `it` (1024 leaves: a
// subtree per wave, then up to two levels above them)
struct WaveItem {
    oi_build_desc D;
    uint32_t hw, hb, f0, nleaf;
    bool on;  // this wave owns leaves of the item
};
__device__ __forceinline__ WaveItem wave_item(const ReduceArgs& A, uint2 it, uint32_t wave)
{
    WaveItem w;
    w.D = A.T.mesh(it.x);
    const uint32_t L = w.D.L;
    w.hw = (L < RW_LG) ? L : RW_LG;
    w.hb = (L - w.hw < 2u) ? L - w.hw : 2u;
    w.f0 = ((it.y << w.hb) + wave) << w.hw;
    w.nleaf = oi_lvl_cnt(L, L, w.D.vl);
    w.on = wave < (1u << w.hb) && w.f0 < w.nleaf;
    return w;
}

// one work item; S is the workgroup's staging, free again on return
__device__ __forceinline__ void reduce_w_item(const ReduceArgs& A, uint2 it, float (*S)[6 * RW_W])
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const WaveItem w = wave_item(A, it, wave);
    const oi_build_desc& D = w.D;
    float* W = S[wave];
    if (w.on) {
        if (A.vals) {
            if (D.xf)
                wave_leaves<true, true>(A, D, w.f0, w.hw, w.nleaf, lane, W);
            else
                wave_leaves<true, false>(A, D, w.f0, w.hw, w.nleaf, lane, W);
        } else {
            if (D.xf)
                wave_leaves<false, true>(A, D, w.f0, w.hw, w.nleaf, lane, W);
            else
                wave_leaves<false, false>(A, D, w.f0, w.hw, w.nleaf, lane, W);
        }
        wave_climb_all(D, w.f0, w.hw, w.nleaf, lane, W);
    }
    if (w.hb == 0) return;
    __syncthreads();
    item_top(D, it, w.hw, w.hb, S);
}

__global__ void __launch_bounds__(OI_BLOCK) k_reduce_w(ReduceArgs A)
{
    __shared__ float S[OI_BLOCK / 64][6 * RW_W];
    reduce_w_item(A, A.T.item(), S);
}

// k_reduce_w for a refit with fewer workgroups than work items (A.nwork):
// each takes items b, b + gridDim.x, ...; the barrier keeps a workgroup's
// next item off S until thread 0 has read the wave roots of this one (hb is
// uniform over the workgroup). A kernel of its own: the loop costs
// registers. (Loading the next item's faces before the current item's
// vertices was measured: the refit was no faster beside a detect, and the
// detect slower.)
__global__ void __launch_bounds__(OI_BLOCK) k_reduce_wl(ReduceArgs A)
{
    __shared__ float S[OI_BLOCK / 64][6 * RW_W];
    for (uint32_t v = blockIdx.x; v < A.nwork; v += gridDim.x) {
        reduce_w_item(A, A.T.item_at(v, A.nwork), S);
        if (v + gridDim.x < A.nwork) __syncthreads();
    }
}

// Stage 2, one workgroup per mesh: the levels above the stage-1 subtree roots
// (level lc = L - h), in chunks of up to 2^11 nodes per climb, with as many
// nodes per thread as keep all 256 threads busy. A separate launch, so the
// roots written by every stage-1 workgroup are visible without device-scope
// fences.
template <int M>
__device__ __forceinline__ void apex_chunks(const oi_build_desc& D, uint32_t lc, uint32_t h2, float* stage)
{
    const uint32_t nchunk = oi_lvl_cnt(lc - h2, D.L, D.vl);
    NodeLoad nl{D.nodes, oi_lvl_start(lc, D.L, D.vl)};
    for (uint32_t c = 0; c < nchunk; c++) reduce_subtree<M>(nl, D.nodes, D.L, D.vl, lc, h2, c, stage);
}

// zeroed level counts and result counters for the detect that follows
// (the work of the collide side's k_seed, one workgroup)
__device__ __forceinline__ void seed_counters(const oi_seed_args& s)
{
    const uint32_t i = threadIdx.x;
    if (i < s.n_counts) s.counts[i] = (i == 0) ? s.n0 : 0u;
    if (i == 0) {
        s.ctl->sticky = s.reset_sticky ? 0u : (s.ctl->sticky | s.ctl->error);
        s.ctl->error = 0;
        s.ctl->n_hits = 0;
        s.ctl->cands = 0;
        s.ctl->fp64 = 0;
        s.ctl->adj = 0;
        s.ctl->tests = 0;
        s.ctl->deferred = 0;
    }
}

// with seed.counts set, the launch has one more workgroup, which resets the
// counters of the detect that follows instead of climbing a mesh. Work items
// are numbered through the XCD remap (BuildTab::item), so the extra workgroup
// is the one whose remapped index is the last: the others then cover exactly
// the items 0 .. n - 1 of the work list.
__global__ void __launch_bounds__(OI_BLOCK) k_reduce_apex(ReduceArgs a, uint32_t lgm, oi_seed_args seed)
{
    __shared__ float stage[StageSize<8>::floats];
    if (seed.counts && oi_xcd_remap(blockIdx.x, gridDim.x) == gridDim.x - 1) {
        seed_counters(seed);
        return;
    }
    const uint2 it = a.T.item();
    const oi_build_desc D = a.T.mesh(it.x);
    const uint32_t L = D.L;
    const uint32_t h = (L < 8 + lgm) ? L : 8 + lgm;
    uint32_t lc = L - h;
    while (lc > 0) {
        const uint32_t h2 = (lc < 11u) ? lc : 11u;
        if (h2 <= 8)
            apex_chunks<1>(D, lc, h2, stage);
        else if (h2 == 9)
            apex_chunks<2>(D, lc, h2, stage);
        else if (h2 == 10)
            apex_chunks<4>(D, lc, h2, stage);
        else
            apex_chunks<8>(D, lc, h2, stage);
        lc -= h2;
        __threadfence_block();
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host-side launchers (called from oibvh_host.cpp)
// ---------------------------------------------------------------------------
// a tree's model matrix, written in stream order (the host copy may go away)
struct Mat16 {
    float m[16];
};
__global__ void k_set_xf(float* dst, Mat16 m)
{
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 16; k++) dst[k] = m.m[k];
    }
}

// point mesh m of a batch at other vertex memory, in stream order (so that a
// captured frame replays against it without capturing again)
__global__ void k_set_mesh_pos(oi_build_desc* bd, oi_mesh_desc* md, uint32_t m, const float* pos)
{
    if (threadIdx.x == 0) {
        bd[m].pos = pos;
        md[m].pos = pos;
    }
}

static BuildTab make_btab(const oi_build_desc* d, const uint2* work, const oi_build_desc* one)
{
    BuildTab t{};
    t.d = d;
    t.work = work;
    t.n_inl = 1;
    if (one) t.inl[0] = *one;
    return t;
}

// n (<= OI_INLINE_MESHES) meshes in the kernel arguments; mesh m owns
// workgroups [boff[m], boff[m + 1])
static BuildTab make_btab_inline(const oi_build_desc* meshes, uint32_t n, const uint32_t* boff)
{
    BuildTab t{};
    t.n_inl = n;
    for (uint32_t m = 0; m < n && m < OI_INLINE_MESHES; m++) {
        t.inl[m] = meshes[m];
        t.boff[m] = boff[m];
    }
    t.boff[n < OI_INLINE_MESHES ? n : OI_INLINE_MESHES] = boff[n];
    return t;
}

// a launch that records t0 as it starts and t1 as it completes (when set)
template <class K, class... Args>
static void launch_timed(K kernel, dim3 g, dim3 b, hipStream_t st, hipEvent_t t0, hipEvent_t t1, Args... args)
{
    if (t1)
        hipExtLaunchKernelGGL(kernel, g, b, 0, st, t0, t1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, g, b, 0, st, args...);
}

extern "C" {

// Every launcher takes the mesh table as (d, work, one): a device descriptor
// table and work list, or d = work = null and `one` for a single mesh whose
// workgroup b takes chunk b.
hipError_t oi_launch_mesh_box(hipStream_t st, const oi_build_desc* d, const uint2* work, const oi_build_desc* one,
                              uint32_t nblocks, float* partials, uint32_t* clear_u32, uint32_t clear_n)
{
    hipLaunchKernelGGL(k_mesh_box, dim3(nblocks), dim3(OI_BLOCK), 0, st, make_btab(d, work, one), partials,
                       clear_u32, clear_n);
    return hipGetLastError();
}

uint32_t oi_morton_chunk() { return OI_MORTON_CHUNK; }

hipError_t oi_launch_set_xf(hipStream_t st, float* dst, const float* mat4)
{
    Mat16 m;
    for (int k = 0; k < 16; k++) m.m[k] = mat4[k];
    hipLaunchKernelGGL(k_set_xf, dim3(1), dim3(64), 0, st, dst, m);
    return hipGetLastError();
}

// keys_hi: null for 30-bit keys, else the high words of 63-bit keys
hipError_t oi_launch_morton(hipStream_t st, const oi_build_desc* d, const uint2* work, const oi_build_desc* one,
                            uint32_t nblocks, const float* partials, uint32_t* keys, uint32_t* keys_hi,
                            uint32_t* ghist, uint32_t* fmesh, uint32_t hbits)
{
    if ((hbits != 8 && hbits != 10) || (keys_hi && hbits != 8)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_morton, dim3(nblocks), dim3(OI_BLOCK), 0, st, make_btab(d, work, one), partials, keys,
                       keys_hi, ghist, 0u, fmesh, hbits);
    return hipGetLastError();
}

// n (<= OI_INLINE_MESHES) meshes in the kernel arguments, mesh m owning the
// workgroups [boff[m], boff[m + 1]): the mesh boxes (partials in each mesh's
// own box buffer), and the Morton keys with mesh m's digit histograms at
// ghist + m * hstride
hipError_t oi_launch_mesh_box_inline(hipStream_t st, const oi_build_desc* meshes, uint32_t n, const uint32_t* boff,
                                     uint32_t* clear_u32, uint32_t clear_n)
{
    if (n == 0 || n > OI_INLINE_MESHES) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mesh_box, dim3(boff[n]), dim3(OI_BLOCK), 0, st, make_btab_inline(meshes, n, boff), nullptr,
                       clear_u32, clear_n);
    return hipGetLastError();
}

hipError_t oi_launch_morton_inline(hipStream_t st, const oi_build_desc* meshes, uint32_t n, const uint32_t* boff,
                                   uint32_t* keys, uint32_t* keys_hi, uint32_t* ghist, uint32_t hstride,
                                   uint32_t hbits)
{
    if (n == 0 || n > OI_INLINE_MESHES) return hipErrorInvalidValue;
    if ((hbits != 8 && hbits != 10) || (keys_hi && hbits != 8)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_morton, dim3(boff[n]), dim3(OI_BLOCK), 0, st, make_btab_inline(meshes, n, boff), nullptr,
                       keys, keys_hi, ghist, hstride, nullptr, hbits);
    return hipGetLastError();
}

hipError_t oi_launch_set_mesh_pos(hipStream_t st, oi_build_desc* bd, oi_mesh_desc* md, uint32_t m, const float* pos)
{
    hipLaunchKernelGGL(k_set_mesh_pos, dim3(1), dim3(64), 0, st, bd, md, m, pos);
    return hipGetLastError();
}

uint32_t oi_sort_tile() { return RS_TILE; }

hipError_t oi_launch_sort_pass_segs(hipStream_t st, const uint32_t* keys_in, const uint32_t* vals_in,
                                    uint32_t* keys_out, uint32_t* vals_out, uint32_t ns, const uint32_t* base,
                                    const uint32_t* n, uint32_t shift, const uint32_t* ghist, uint32_t hstride,
                                    uint32_t* lookback, uint32_t* lookback_next, uint32_t* tile_counter,
                                    const uint32_t* key_of_val, uint32_t bits);

hipError_t oi_launch_sort_pass(hipStream_t st, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                               uint32_t* vals_out, uint32_t n, uint32_t shift, const uint32_t* ghist,
                               uint32_t* lookback, uint32_t* lookback_next, uint32_t* tile_counter,
                               const uint32_t* key_of_val, uint32_t bits)
{
    const uint32_t base = 0;
    return oi_launch_sort_pass_segs(st, keys_in, vals_in, keys_out, vals_out, 1, &base, &n, shift, ghist, 0,
                                    lookback, lookback_next, tile_counter, key_of_val, bits);
}

uint32_t oi_sort_segs_max() { return OI_SORT_SEGS; }

// ns segments: segment s is keys [base[s], base[s] + n[s]) of the arrays, its
// digit histogram at ghist + s * hstride, its tile counter tile_counter[s];
// look-back statuses are indexed by the tile id over all segments (sum of
// ceil(n[s] / tile) tiles)
hipError_t oi_launch_sort_pass_segs(hipStream_t st, const uint32_t* keys_in, const uint32_t* vals_in,
                                    uint32_t* keys_out, uint32_t* vals_out, uint32_t ns, const uint32_t* base,
                                    const uint32_t* n, uint32_t shift, const uint32_t* ghist, uint32_t hstride,
                                    uint32_t* lookback, uint32_t* lookback_next, uint32_t* tile_counter,
                                    const uint32_t* key_of_val, uint32_t bits)
{
    if (ns == 0 || ns > OI_SORT_SEGS || (bits != 8 && bits != 10)) return hipErrorInvalidValue;
    SortSegs S{};
    S.ns = ns;
    S.hstride = hstride;
    uint32_t tiles = 0;
    for (uint32_t k = 0; k < ns; k++) {
        if (n[k] == 0) return hipErrorInvalidValue;
        S.tile0[k] = tiles;
        S.base[k] = base[k];
        S.n[k] = n[k];
        tiles += (n[k] + RS_TILE - 1) / RS_TILE;
    }
    S.tile0[ns] = tiles;
    if (bits == 10)
        hipLaunchKernelGGL(k_sort_pass<10>, dim3(tiles), dim3(RS_THREADS), 0, st, keys_in, vals_in, keys_out,
                           vals_out, S, shift, ghist, lookback, lookback_next, tile_counter, key_of_val);
    else
        hipLaunchKernelGGL(k_sort_pass<8>, dim3(tiles), dim3(RS_THREADS), 0, st, keys_in, vals_in, keys_out,
                           vals_out, S, shift, ghist, lookback, lookback_next, tile_counter, key_of_val);
    return hipGetLastError();
}

// subtree height of stage 1 for M leaves per thread
// M = 16 selects k_reduce_w, whose stage 1 climbs as high as k_reduce<4>
static uint32_t lgm_of(int M) { return (M == 1) ? 0 : (M == 2) ? 1 : (M == 4 || M == 16) ? 2 : 3; }

uint32_t oi_reduce_height(int M, uint32_t L)
{
    const uint32_t lgm = lgm_of(M);
    return (L < 8 + lgm) ? L : 8 + lgm;
}

// stage 1 (M = leaves per thread: 1, 2, 4 or 8; 16: wave subtrees); vals:
// build, null: refit. With t0 and t1 (both or neither) the launch records
// them as it starts and completes. max_blocks > 0 (refits with M = 16 only):
// at most that many workgroups (rounded up to a multiple of 8), each looping
// over work items, as in oi_launch_reduce_inline
hipError_t oi_launch_reduce(hipStream_t st, int M, const oi_build_desc* d, const uint2* work,
                            const oi_build_desc* one, uint32_t nblocks, const uint32_t* vals, uint32_t write_perm,
                            hipEvent_t t0, hipEvent_t t1, uint32_t max_blocks)
{
    if (!t0 != !t1) return hipErrorInvalidValue;
    ReduceArgs a{make_btab(d, work, one), vals, write_perm};
    dim3 g(nblocks);
    const dim3 b(OI_BLOCK);
    if (M == 16 && !vals && max_blocks && max_blocks < nblocks) {
        const uint32_t cap = (max_blocks + 7u) & ~7u;
        if (cap < nblocks) {
            a.nwork = nblocks;
            g = dim3(cap);
        }
    }
    switch (M) {
    case 1: launch_timed(k_reduce<1>, g, b, st, t0, t1, a); break;
    case 2: launch_timed(k_reduce<2>, g, b, st, t0, t1, a); break;
    case 4: launch_timed(k_reduce<4>, g, b, st, t0, t1, a); break;
    case 16:
        if (a.nwork)
            launch_timed(k_reduce_wl, g, b, st, t0, t1, a);
        else
            launch_timed(k_reduce_w, g, b, st, t0, t1, a);
        break;
    default: launch_timed(k_reduce<8>, g, b, st, t0, t1, a); break;
    }
    return hipGetLastError();
}

uint32_t oi_inline_meshes() { return OI_INLINE_MESHES; }

// stage 1 over up to OI_INLINE_MESHES meshes passed by value; boff[n + 1]
// gives each mesh's first workgroup
// with t0 and t1 (both or neither) the launch itself records them as it
// starts and completes: their elapsed time is the kernel's own duration.
// max_blocks > 0 (refits with k_reduce_w only): at most that many workgroups (rounded up
// to a multiple of 8), each looping over work items, which leaves the rest
// of the GPU to work running beside the launch
hipError_t oi_launch_reduce_inline(hipStream_t st, int M, const oi_build_desc* meshes, uint32_t n,
                                   const uint32_t* boff, const uint32_t* vals, uint32_t write_perm, hipEvent_t t0,
                                   hipEvent_t t1, uint32_t max_blocks)
{
    if (n == 0 || n > OI_INLINE_MESHES) return hipErrorInvalidValue;
    if (!t0 != !t1) return hipErrorInvalidValue;
    ReduceArgs a{make_btab_inline(meshes, n, boff), vals, write_perm};
    dim3 g(boff[n]);
    const dim3 b(OI_BLOCK);
    if (M == 16 && !vals && max_blocks && max_blocks < boff[n]) {
        const uint32_t cap = (max_blocks + 7u) & ~7u;
        if (cap < boff[n]) {
            a.nwork = boff[n];
            g = dim3(cap);
        }
    }
    switch (M) {
    case 1: launch_timed(k_reduce<1>, g, b, st, t0, t1, a); break;
    case 2: launch_timed(k_reduce<2>, g, b, st, t0, t1, a); break;
    case 4: launch_timed(k_reduce<4>, g, b, st, t0, t1, a); break;
    case 16:
        if (a.nwork)
            launch_timed(k_reduce_wl, g, b, st, t0, t1, a);
        else
            launch_timed(k_reduce_w, g, b, st, t0, t1, a);
        break;
    default: launch_timed(k_reduce<8>, g, b, st, t0, t1, a); break;
    }
    return hipGetLastError();
}

// stage 2 for n inline meshes, one workgroup each; with seed (may be null)
// one more workgroup resets the following detect's counters
hipError_t oi_launch_reduce_apex_inline(hipStream_t st, int M, const oi_build_desc* meshes, uint32_t n,
                                        const oi_seed_args* seed, hipEvent_t done)
{
    if (n == 0 || n > OI_INLINE_MESHES) return hipErrorInvalidValue;
    if (seed && seed->n_counts > OI_BLOCK) return hipErrorInvalidValue;
    uint32_t boff[OI_INLINE_MESHES + 1];
    for (uint32_t m = 0; m <= n; m++) boff[m] = m;
    ReduceArgs a{make_btab_inline(meshes, n, boff), nullptr, 0};
    const uint32_t lgm = lgm_of(M);
    const oi_seed_args sd = seed ? *seed : oi_seed_args{};
    // with done the launch records it on completion (no separate event call)
    if (done)
        hipExtLaunchKernelGGL(k_reduce_apex, dim3(n + (seed ? 1u : 0u)), dim3(OI_BLOCK), 0, st, nullptr, done, 0, a,
                              lgm, sd);
    else
        hipLaunchKernelGGL(k_reduce_apex, dim3(n + (seed ? 1u : 0u)), dim3(OI_BLOCK), 0, st, a, lgm, sd);
    return hipGetLastError();
}

// stage 2, one workgroup per work item (mesh, 0) whose L exceeds the stage-1 height
hipError_t oi_launch_reduce_apex(hipStream_t st, int M, const oi_build_desc* d, const uint2* work,
                                 const oi_build_desc* one, uint32_t nblocks, const oi_seed_args* seed)
{
    if (seed && seed->n_counts > OI_BLOCK) return hipErrorInvalidValue;
    ReduceArgs a{make_btab(d, work, one), nullptr, 0};
    const uint32_t lgm = lgm_of(M);
    const oi_seed_args sd = seed ? *seed : oi_seed_args{};
    hipLaunchKernelGGL(k_reduce_apex, dim3(nblocks + (seed ? 1u : 0u)), dim3(OI_BLOCK), 0, st, a, lgm, sd);
    return hipGetLastError();
}

}  // extern "C"
