// oibvh_noalloc.hip -- oibvh_build_noalloc: the tree build over caller-owned
// device buffers with its scratch in a fixed device array, for callers that
// cannot allocate (HIP graph capture, serving loops, harnesses that forbid
// hipMalloc). No reference counterpart; the launches are oibvh_tree_build's
// (oi_enqueue_build), so the outputs are that entry point's bit for bit.
//
// The array `scratch` below exists once per GPU the library runs on (each
// device holds its own copy, at its own address, until the process ends).
// In host code its name is a host-side stand-in: its device address is taken
// with hipGetSymbolAddress once per device, with that device current, when a
// context is created there, and only that address is handed to kernels.
// Every call on one GPU shares that GPU's array, which keeps the last call's
// contents; everything the build needs zeroed (histograms, tile counters,
// look-back flags) is cleared by the build's own first kernel. Calls that
// overlap on one GPU (different streams or contexts of that device) race on
// the array; calls in one stream are ordered.
#include <mutex>

#include "oibvh_internal.h"

// layout of the scratch (32-bit words), each part 256-byte aligned:
// keys | keys_alt | vals_tmp (F each) | sort head + look-back (OI_SORT_HEAD +
// 2 OI_SORT_LB words per sort tile) | Morton box + mesh-box partials (6 + 6 * 256)
#define OI_NA_ALIGN 64u
#define OI_NA_ROUND(x) (((x) + OI_NA_ALIGN - 1u) & ~(OI_NA_ALIGN - 1u))
// the sort's tile is 8192 keys; the run-time check below refuses a build
// whose parts do not fit (should that tile ever shrink)
#define OI_NA_TILE_MIN 8192u
#define OI_NA_WORDS                                                                                  \
    (3u * OI_NA_ROUND(OIBVH_NOALLOC_MAX_FACES) +                                                     \
     OI_NA_ROUND(OI_SORT_HEAD + 2u * OI_SORT_LB * ((OIBVH_NOALLOC_MAX_FACES + OI_NA_TILE_MIN - 1u) / OI_NA_TILE_MIN)) + \
     OI_NA_ROUND(6u + 6u * 256u))

static __device__ __align__(256) uint32_t scratch[OI_NA_WORDS];

#define OI_NA_MAX_DEVICES 64
static std::mutex g_mu;
static void* g_ptr[OI_NA_MAX_DEVICES];      // device address of `scratch`, per device
static hipError_t g_err[OI_NA_MAX_DEVICES]; // hipGetSymbolAddress's result, per device
static bool g_done[OI_NA_MAX_DEVICES];

// once per device, with that device current (called by oibvh_ctx_create)
hipError_t oi_noalloc_attach(int device)
{
    if (device < 0 || device >= OI_NA_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        void* p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = g_err[device] == hipSuccess ? p : nullptr;
        g_done[device] = true;
    }
    return g_err[device];
}

static uint32_t* scratch_of(int device)
{
    if (device < 0 || device >= OI_NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_done[device] ? (uint32_t*)g_ptr[device] : nullptr;
}

extern "C" oibvh_status oibvh_build_noalloc(oibvh_ctx* ctx, const float* xyz, uint32_t V, const uint32_t* idx,
                                           uint32_t F, float* nodes, uint32_t* perm, uint32_t* sorted_faces)
{
    if (!ctx) return OIBVH_ERR_INVALID;
    if (!xyz || !idx || !nodes || !perm || !sorted_faces || V == 0 || F == 0)
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "oibvh_build_noalloc: invalid argument");
    if (F > OIBVH_NOALLOC_MAX_FACES)
        return oi_set_err(ctx, OIBVH_ERR_CAPACITY, "oibvh_build_noalloc: %u faces exceed the maximum of %u", F,
                          OIBVH_NOALLOC_MAX_FACES);
    const uint32_t tiles = (F + oi_sort_tile() - 1) / oi_sort_tile();
    const uint64_t need = 3ull * OI_NA_ROUND(F) + OI_NA_ROUND(OI_SORT_HEAD + 2ull * OI_SORT_LB * tiles) +
                          OI_NA_ROUND(6u + 6u * 256u);
    if (need > OI_NA_WORDS)
        return oi_set_err(ctx, OIBVH_ERR_CAPACITY, "oibvh_build_noalloc: %llu scratch words needed, %u held",
                          (unsigned long long)need, (unsigned)OI_NA_WORDS);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint32_t* base = scratch_of(ctx->device);
    if (!base) {
        HIPCHK(ctx, oi_noalloc_attach(ctx->device));
        base = scratch_of(ctx->device);
        if (!base) return oi_set_err(ctx, OIBVH_ERR_HIP, "oibvh_build_noalloc: no scratch address on this device");
    }
    HIPCHK(ctx, oi_ctx_join(ctx));
    uint32_t* w = base;
    oi_build_scratch s{};
    s.keys = w;
    w += OI_NA_ROUND(F);
    s.keys_alt = w;
    w += OI_NA_ROUND(F);
    s.vals_tmp = w;
    w += OI_NA_ROUND(F);
    s.sort = w;
    w += OI_NA_ROUND(OI_SORT_HEAD + 2u * OI_SORT_LB * tiles);
    s.keys_hi = nullptr;  // 30-bit keys (the reference's morton3D)
    oi_build_desc d{};
    d.pos = xyz;
    d.xf = nullptr;
    d.faces = idx;
    d.sfaces = sorted_faces;
    d.perm = perm;
    d.nodes = nodes;
    d.box = reinterpret_cast<float*>(w);
    d.F = F;
    d.V = V;
    d.L = oi_ilog2(oi_np2(F));
    d.vl = oi_np2(F) - F;
    d.fbase = 0;
    d.pbase = 0;
    d.nparts = oi_box_parts(V);
    d.vchunk = (V + d.nparts - 1) / d.nparts;
    return oi_enqueue_build(ctx, d, s, oi_build_lpt(ctx, F));
}
