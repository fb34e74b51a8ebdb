// oibvh_common.h -- device descriptors and device helpers shared by host code
// and kernels; the layout math lives in oibvh_layout.h.
#pragma once
#include "oibvh_layout.h"

#if defined(__HIPCC__) || defined(__HIP__)
// Pointers read from descriptor tables are generic to the compiler, which
// then emits flat memory instructions; these views make them global so that
// loads and stores are global_* and adjacent words combine into one
// multi-dword access (indices are 64-bit so adjacency is provable).
#define OI_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const OI_GLOBAL T* oi_g(const T* p)
{
    return (const OI_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ OI_GLOBAL T* oi_gw(T* p)
{
    return (OI_GLOBAL T*)p;
}
// three consecutive 32-bit words at base + 3 i (a vertex or a face)
__device__ __forceinline__ void oi_ld3(const float* base, uint32_t i, float* o)
{
    const OI_GLOBAL float* q = oi_g(base) + 3ull * i;
    o[0] = q[0];
    o[1] = q[1];
    o[2] = q[2];
}
__device__ __forceinline__ void oi_ld3(const uint32_t* base, uint32_t i, uint32_t* o)
{
    const OI_GLOBAL uint32_t* q = oi_g(base) + 3ull * i;
    o[0] = q[0];
    o[1] = q[1];
    o[2] = q[2];
}
// inclusive prefix sum over the 64 lanes of a wave (every lane active) with
// DPP moves instead of LDS permutes: a Kogge-Stone scan inside each row of
// 16 lanes (row_shr 1, 2, 4, 8; a lane shifted in from outside its row
// reads 0), then rows 1 and 3 add the last lane of the row before them
// (row_bcast:15), then rows 2 and 3 add lane 31 (row_bcast:31)
__device__ __forceinline__ uint32_t oi_wave_scan_incl(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    return x;
}
#endif

// ---------------------------------------------------------------------------
// device descriptors
// ---------------------------------------------------------------------------

// p' = M * vec4(p, 1) with glm's evaluation order for mat4 * vec4
// (type_mat4x4.inl:561-573): (m[0] x + m[1] y) + (m[2] z + m[3] * 1), m[c] the
// c-th column, stored column-major like glm::mat4 (xf[4c + r]). This is the
// vertex transform of Mesh::transform (mesh.cpp:189-214, transform.cu:35-41).
OI_HD void oi_xform(const float* xf, float x, float y, float z, float* o)
{
    for (int r = 0; r < 3; r++) {
        const float a = xf[r] * x;
        const float b = xf[4 + r] * y;
        const float c = xf[8 + r] * z;
        const float d = xf[12 + r] * 1.0f;
        o[r] = (a + b) + (c + d);
    }
}

// One mesh as the build and refit kernels see it.
struct oi_build_desc {
    const float* pos;        // V*3 vertex positions
    const float* xf;         // model matrix applied to pos (oi_xform), or null
    const uint32_t* faces;   // F*3 original faces (vertex ids local to the mesh)
    uint32_t* sfaces;        // F*3 faces in leaf order (written by the build, read by the refit)
    uint32_t* perm;          // F: leaf -> original face id
    float* nodes;            // N*6 floats, real order
    float* box;              // 6 floats: the Morton quantisation box (written by the build)
    uint32_t F, V, L, vl;
    uint32_t fbase;          // offset in the concatenated per-face sort arrays
    uint32_t pbase;          // first mesh-box partial of this mesh
    uint32_t nparts;         // mesh-box partials (1..256)
    uint32_t vchunk;         // vertices per mesh-box partial
};

// One mesh as the traversal sees it. `ibase` is the mesh's offset in the
// scene's concatenated implicit-index space (each mesh spans 2^(L+1) - 1).
struct oi_mesh_desc {
    const float* nodes;      // N*6 floats, real order: min.xyz, max.xyz
    const uint32_t* sfaces;  // F*3, faces in leaf (Morton) order
    const float* pos;        // V*3 vertex positions
    const uint32_t* perm;    // F: leaf position -> original face id
    const float* xf;         // model matrix applied to pos, or null
    uint32_t F, L, vl, N;
    uint32_t ibase;
    uint32_t pad0;
};

// A mesh table is either a scene's (d_meshes / h_meshes, n_meshes) or, when
// chunk_mesh is not null, a batch's (d_meshes indexed through chunk_mesh).
struct oi_tab_args {
    const oi_mesh_desc* d_meshes;
    const oi_mesh_desc* h_meshes;
    uint32_t n_meshes;
    const uint32_t* chunk_mesh;
    uint32_t shift;
};

// Detect counters, reset by the seed kernel before each detect.
struct oi_detect_ctl {
    uint32_t n_hits;           // hits found (may exceed capacity)
    uint32_t cands;            // candidate leaf pairs (may exceed capacity)
    uint32_t error;            // OI_ERR_* bits of this detect
    uint32_t sticky;           // OI_ERR_* bits of earlier asynchronous detects not checked since
    unsigned long long fp64;   // SAT axis comparisons re-evaluated in double
    unsigned long long adj;    // leaf pairs dropped as adjacent (self mode)
    unsigned long long tests;  // BVTT box pair tests
    uint32_t deferred;         // FP64 fallback: candidates the FP32 filter left undecided
    uint32_t pad;
};

// a detect's counter reset (zeroed level counts and result counters), done
// by the last workgroup of a refit's apex launch when a detect follows it in
// the same frame call, instead of a launch of its own
struct oi_seed_args {
    uint32_t* counts;  // null: no reset in this launch
    oi_detect_ctl* ctl;
    uint32_t n0, n_counts, reset_sticky, pad;
};

#define OI_ERR_QUEUE_OVERFLOW 1u  // a BVTT front exceeded its buffer
#define OI_ERR_TIMEOUT 2u
#define OI_ERR_CORRUPT 4u
#define OI_ERR_HITS_OVERFLOW 8u
#define OI_ERR_CAND_OVERFLOW 16u
#define OI_ERR_LOCAL_OVERFLOW 32u  // a workgroup's front outgrew its LDS queue (head local walk)
#define OI_ERR_TAIL_OVERFLOW 64u   // the same in the tail local walk

// one launch of the workgroup-local walk (k_bvtt_local)
struct oi_local_args {
    const uint2* front_in;
    const uint32_t* count_in;  // device length of front_in (null: n_in, one workgroup per chunk)
    uint32_t n_in;
    uint32_t cap_in;           // entries front_in holds (with count_in)
    uint32_t chunk;            // entries per workgroup and pass, at most 2048
    uint32_t levels;           // steps walked; the last one's non-leaf pairs go to front_out
    uint32_t grid;             // with count_in: at most this many workgroups (0: one per chunk of cap_in)
    uint32_t test_entries;     // test the input pairs themselves (the initial front)
    uint32_t overflow_bit;     // error bit set on an LDS front overflow
    uint2* front_out;
    uint32_t* count_out;
    uint32_t cap_out;
    hipEvent_t done;           // recorded when the launch completes (null: none)
};

// the bytes a device sort of the hit list visits, least significant first:
// digit p is byte shift[p] / 8 of word word[p] (0 meshA, 1 meshB, 2 triA,
// 3 triB)
#define OI_HIT_DIGITS 16
struct oi_hit_digits {
    uint32_t nd;
    uint8_t word[OI_HIT_DIGITS];
    uint8_t shift[OI_HIT_DIGITS];
};

// a scene's concatenated implicit index space must stay below this bound
#define OI_IMPLICIT_LIMIT 0xfffffffeu

// detect flags (mirror include/oibvh.h)
#define OI_DETECT_SELF 1u
#define OI_DETECT_FP64_FALLBACK 2u
