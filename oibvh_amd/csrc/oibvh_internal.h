// oibvh_internal.h -- host-side internals shared by the C ABI implementation
// files: the context, error handling, device allocation helpers, the kernel
// launchers and the BVTT detector used by scenes and batches.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oibvh.h"
#include "oibvh_common.h"

extern "C" {
hipError_t oi_launch_mesh_box(hipStream_t, const oi_build_desc*, const uint2*, const oi_build_desc*, uint32_t, float*,
                              uint32_t*, uint32_t);
uint32_t oi_morton_chunk();
hipError_t oi_launch_mesh_box_inline(hipStream_t, const oi_build_desc*, uint32_t, const uint32_t*, uint32_t*,
                                     uint32_t);
// hbits: digit width of the histograms the Morton launch accumulates (8, or
// 10 for 30-bit keys)
hipError_t oi_launch_morton_inline(hipStream_t, const oi_build_desc*, uint32_t, const uint32_t*, uint32_t*,
                                   uint32_t*, uint32_t*, uint32_t, uint32_t hbits = 8);
hipError_t oi_launch_morton(hipStream_t, const oi_build_desc*, const uint2*, const oi_build_desc*, uint32_t,
                            const float*, uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t hbits = 8);

// sort scratch: digit histograms (8 x 256 or 3 x 1024 bins) | per-pass tile
// counters | two look-back buffers of OI_SORT_LB words per tile (256 per
// tile suffice for 8-bit digits)
#define OI_SORT_HIST 3072u
#define OI_SORT_CTRS 16u
#define OI_SORT_HEAD (OI_SORT_HIST + OI_SORT_CTRS)
#define OI_SORT_LB 1024u
// 30-bit Morton keys of a tree or scene build: three passes of 10-bit digits
#define OI_M30_PASSES 3
#define OI_M30_BITS 10u
uint32_t oi_sort_tile();
hipError_t oi_launch_sort_pass(hipStream_t, const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, uint32_t,
                               uint32_t, const uint32_t*, uint32_t*, uint32_t*, uint32_t*, const uint32_t*,
                               uint32_t bits = 8);
uint32_t oi_sort_segs_max();
// bits: digit width, 8 or 10 (look-back buffers of 2^bits words per tile)
hipError_t oi_launch_sort_pass_segs(hipStream_t, const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, uint32_t,
                                    const uint32_t*, const uint32_t*, uint32_t, const uint32_t*, uint32_t, uint32_t*,
                                    uint32_t*, uint32_t*, const uint32_t*, uint32_t bits = 8);
uint32_t oi_reduce_height(int, uint32_t);
hipError_t oi_launch_reduce(hipStream_t, int, const oi_build_desc*, const uint2*, const oi_build_desc*, uint32_t,
                            const uint32_t*, uint32_t, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr,
                            uint32_t max_blocks = 0);
hipError_t oi_launch_reduce_apex(hipStream_t, int, const oi_build_desc*, const uint2*, const oi_build_desc*,
                                 uint32_t, const oi_seed_args*);
hipError_t oi_launch_set_xf(hipStream_t, float*, const float*);
uint32_t oi_inline_meshes();
hipError_t oi_launch_reduce_inline(hipStream_t, int, const oi_build_desc*, uint32_t, const uint32_t*, const uint32_t*,
                                   uint32_t, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr,
                                   uint32_t max_blocks = 0);
hipError_t oi_launch_reduce_apex_inline(hipStream_t, int, const oi_build_desc*, uint32_t, const oi_seed_args*,
                                        hipEvent_t);
hipError_t oi_launch_set_mesh_pos(hipStream_t, oi_build_desc*, oi_mesh_desc*, uint32_t, const float*);
hipError_t oi_launch_seed(hipStream_t, uint32_t, uint32_t*, uint32_t, oi_detect_ctl*, uint32_t);
hipError_t oi_launch_bvtt_step(hipStream_t, uint32_t, const oi_tab_args*, int, const uint2*, const uint32_t*,
                               uint32_t, uint32_t, uint2*, uint32_t*, uint32_t, uint2*, uint32_t*, uint32_t, uint4*,
                               uint32_t, oi_detect_ctl*, uint32_t, uint32_t, int, hipEvent_t);
hipError_t oi_launch_narrow(hipStream_t, uint32_t, const oi_tab_args*, const uint2*, const uint32_t*, uint32_t,
                            uint4*, uint32_t*, uint32_t, oi_detect_ctl*, int, uint2*, uint32_t, hipEvent_t);
hipError_t oi_launch_bvtt_local(hipStream_t, const oi_tab_args*, const oi_local_args*, uint2*, uint32_t*, uint32_t,
                                uint4*, uint32_t, oi_detect_ctl*, uint32_t, uint32_t, int);
hipError_t oi_launch_pair_vertices(hipStream_t, const oi_build_desc*, const uint4*, uint32_t, float*);
hipError_t oi_launch_zero_u32(hipStream_t, uint32_t*, uint32_t);
hipError_t oi_launch_hits_digits(hipStream_t, const uint4*, uint32_t, uint32_t*, uint32_t*, const oi_hit_digits*);
hipError_t oi_launch_hits_gather(hipStream_t, const uint4*, const uint32_t*, uint32_t, uint4*);
hipError_t oi_launch_hits_sort_small(hipStream_t, const uint4*, uint32_t, const oi_hit_digits*, uint4*);
uint32_t oi_hits_sort_small_max();
}

// side streams of overlapped frames' detects, and the ring of their events
#ifndef OI_SIDE_STREAMS
#define OI_SIDE_STREAMS 2u
#endif
#ifndef OI_SIDE_RING
#define OI_SIDE_RING 4u
#endif
static_assert(OI_SIDE_RING % OI_SIDE_STREAMS == 0,
              "a ring event is recorded again only on the side stream of the detect it marked");
// node buffers per tree and counter sets per detector of overlapped frames:
// frame f's refit rewrites the buffer and resets the set that frame
// f - OI_FRAME_DEPTH's detect used, so it waits for that detect only and may
// run beside the detects of the frames in between. Measured on the deforming
// UV pair (2 x 1M triangles, the 200-frame sequence with rebuilds), with two
// side streams: 3 gave 63.3 us per frame against 66.5 us with 2, and 4
// (the host up to four frames ahead, which absorbs the stalls of the
// rebuild frames) 56.7-58.5 against 59.2-62.9 us with 3, in three A/B runs;
// 5 and 6 gave 57.7-58.7 us against 57.3-57.7 us with 4 (no gain).
// With one side stream 3 had been the slower (75.9-80.5 us against 73.9 us):
// the refit ran ahead into detects that could not start before the one
// ahead. Three side streams measured 61.5-61.6 us (depth 3 or 4)
#ifndef OI_FRAME_DEPTH
#define OI_FRAME_DEPTH 4u
#endif
static_assert(OI_SIDE_RING >= OI_FRAME_DEPTH, "the ring must still hold the event a refit waits for");

// ---------------------------------------------------------------------------
struct oibvh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int cus = 0;
    uint64_t clock_khz = 100000;
    uint32_t leaves_per_thread = 0;
    uint32_t trav_blocks = 0;
    uint32_t trav_expand = 0;
    uint32_t trav_schedule = 0;
    uint32_t trav_local = 0;         // 0: BVTT steps only; n: the first steps as a local walk, n pairs per workgroup
    uint32_t trav_local_levels = 0;  // steps the local walk replaces (0: all)
    uint32_t trav_tail = 0;          // 0: no tail walk; n: the last steps as a local walk, n pairs per workgroup
    uint32_t trav_tail_steps = 0;    // steps the tail walk replaces (0: all after the head walk)
    uint32_t refit_blocks = 0;       // at most this many workgroups in a scene refit's k_reduce_w (0: one per item)
    uint32_t narrow_separate = 0;    // FP32 queries test leaf pairs in one k_narrow launch
    uint64_t tuning_gen = 0;         // counts oibvh_ctx_set_tuning calls
    uint32_t queue_cap = 1u << 22;
    uint32_t hits_cap = 1u << 20;
    uint32_t cand_cap = 1u << 20;
    // overlapped frames (OIBVH_FRAME_OVERLAP) run their detects on the side
    // streams, created at default priority on first use: detect q (side_seq
    // counts them) runs on side[q % OI_SIDE_STREAMS] and ends with
    // ev_ring[q % OI_SIDE_RING], so that consecutive detects may run at once
    hipStream_t side[OI_SIDE_STREAMS] = {};
    hipEvent_t ev_main = nullptr;  // hand-over from the context's stream (a frame's refit done)
    hipEvent_t ev_ring[OI_SIDE_RING] = {};
    uint64_t side_seq = 0;
    bool side_pending = false;     // side work the context's stream has not waited for
    int live = 0;                 // trees, scenes and batches of this context not yet destroyed
    bool destroy_pending = false; // oibvh_ctx_destroy called while objects were live
};


// device address of oibvh_build_noalloc's scratch array on `device`, taken
// once per device with that device current (oibvh_noalloc.hip)
hipError_t oi_noalloc_attach(int device);

// drop one reference held by a tree, scene or batch (frees a context whose
// destroy was deferred)
void oi_ctx_release(oibvh_ctx* ctx);
// order the context's stream after everything queued on its side stream
// (every entry point but an overlapped frame calls it before queueing work
// or reading results)
hipError_t oi_ctx_join(oibvh_ctx* ctx);
// wait until both streams are idle (before memory they may use is freed or
// rewritten from the host)
hipError_t oi_ctx_quiesce(oibvh_ctx* ctx);
// create the side streams and their events of overlapped frames (first use)
oibvh_status oi_ctx_side_init(oibvh_ctx* c);
// the side stream of detect q
inline hipStream_t oi_side_stream(const oibvh_ctx* c, uint64_t q) { return c->side[q % OI_SIDE_STREAMS]; }
// order the context's stream after side detects: tickets[i] is 1 + a detect
// (0: none). Per side stream only the latest ticket is waited for (a stream
// runs its detects in order); once OI_SIDE_RING detects later its ring event
// has been recorded again, on the same stream, and the wait is for that
// later point
hipError_t oi_ctx_wait_tickets(oibvh_ctx* c, const uint64_t* tickets, uint32_t n);
// the same wait on the host: returns once those detects have ended, so that
// the context's stream needs no wait of its own
hipError_t oi_ctx_host_wait_tickets(oibvh_ctx* c, const uint64_t* tickets, uint32_t n);

oibvh_status oi_set_err(oibvh_ctx* ctx, oibvh_status st, const char* fmt, ...);

// device scratch of one build: sort keys (two buffers, plus the high words
// of 63-bit keys or null), the values' ping-pong buffer and the sort's
// histograms, tile counters and look-back tiles (OI_SORT_HEAD + 512 words
// per sort tile)
struct oi_build_scratch {
    uint32_t* keys;
    uint32_t* keys_alt;
    uint32_t* keys_hi;
    uint32_t* vals_tmp;
    uint32_t* sort;
};
// the build's launches over mesh d (d.box holds 6 + 6 * oi_box_parts(V)
// floats) with M leaves per thread in the reduce; no allocation
oibvh_status oi_enqueue_build(oibvh_ctx* c, const oi_build_desc& d, const oi_build_scratch& s, int M,
                              hipStream_t stream = nullptr);
// mesh-box partials (workgroups of k_mesh_box) for V vertices, 1..256
uint32_t oi_box_parts(uint32_t V);
// reduce variant by launch size, and the one a build of F faces uses
int oi_default_lpt(uint64_t F, int cus);
int oi_build_lpt(const oibvh_ctx* c, uint32_t F);

#define HIPCHK(ctx, expr)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return oi_set_err((ctx), OIBVH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <class T>
static inline hipError_t dalloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <class T>
static inline void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

static inline hipMemcpyKind kind_of(uint32_t ptr_kind)
{
    return ptr_kind == OIBVH_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
}

// workgroups per BVTT / narrow-phase launch
static inline uint32_t oi_grid_for(const oibvh_ctx* c)
{
    return c->trav_blocks ? c->trav_blocks : (uint32_t)std::max(1, c->cus) * 4u;
}

// ---------------------------------------------------------------------------
// BVTT detector: the device buffers and launch sequence of one collision
// query (Scene::detectCollisionOnGPU, scene.cu:225-446), shared by scenes and
// batches. The owner provides the mesh table, the initial front and the
// depth; enqueue only launches kernels (it can be captured in a graph),
// prepare sizes the buffers beforehand, check waits, grows what overflowed
// and runs the query again.
// ---------------------------------------------------------------------------
struct oi_detector {
    oibvh_ctx* ctx = nullptr;
    uint2* d_init = nullptr;  // initial BVTT front
    uint32_t init_cap = 0;
    uint32_t n_front = 0;
    uint2* front[2] = {nullptr, nullptr};  // ping-pong BVTT fronts
    uint32_t front_cap = 0;
    // OI_FRAME_DEPTH sets of level counts and counters: overlapped frames
    // take them in turn, so a frame's refit can reset one while the detects
    // of the frames before it still use the others (set k: counts_base +
    // 256 k, ctl_base + k)
    uint32_t* counts_base = nullptr;
    uint32_t counts_cap = 0;
    oi_detect_ctl* ctl_base = nullptr;
    uint32_t set = 0;
    uint64_t set_seq[OI_FRAME_DEPTH] = {};       // 1 + the side-stream detect that last used set k (0: none)
    bool set_unchecked[OI_FRAME_DEPTH] = {};     // set k (not the current one) holds an unchecked result
    uint32_t* counts = nullptr;  // front length per step (+1), of the current set
    uint32_t n_levels = 0;
    uint2* cands = nullptr;
    uint32_t cand_cap = 0;
    uint4* hits = nullptr;
    uint32_t hits_cap = 0;
    // front, cands and hits above are the current set's: every counter set
    // has working buffers of its own (the detects of consecutive overlapped
    // frames may run at once, on different side streams), kept here while
    // another set is current (oi_det_use_set)
    uint2* set_front[OI_FRAME_DEPTH][2] = {};
    uint32_t set_front_cap[OI_FRAME_DEPTH] = {};
    uint2* set_cands[OI_FRAME_DEPTH] = {};
    uint32_t set_cand_cap[OI_FRAME_DEPTH] = {};
    uint4* set_hits[OI_FRAME_DEPTH] = {};
    uint32_t set_hits_cap[OI_FRAME_DEPTH] = {};
    // sizes every set's buffers grow to when it is next prepared (raised by
    // overflows, so that a set does not overflow where another already did)
    uint32_t want_front = 0, want_cands = 0, want_hits = 0;
    oi_detect_ctl* ctl = nullptr;    // counters of the current set
    oi_detect_ctl* h_ctl = nullptr;  // pinned, OI_FRAME_DEPTH (the current set's, then the others read)
    std::vector<uint32_t> h_counts;
    // the query (kept for re-runs)
    oi_tab_args tab{};
    uint32_t depth = 0, expand = 1, flags = 0, implicit_end = 0;
    uint32_t schedule = 0;      // levels per step of the first steps (2 bits each), then `expand`
    uint8_t step_e[256] = {};   // levels descended by each step of the current query
    uint32_t local_chunk = 0;   // this query: initial pairs per workgroup of the local walk (0: steps)
    uint32_t local_levels = 0;  // this query: steps done by the local walk
    uint64_t local_off = 0;     // 1 + the tuning generation in which the head walk overflowed (0: none)
    uint32_t tail_chunk = 0;    // this query: front pairs per workgroup of the tail walk (0: none)
    uint32_t tail_from = 0;     // this query: first step done by the tail walk
    uint64_t tail_off = 0;      // the same for the tail walk
    bool fused = false;  // narrow phase inside the BVTT steps (FP32 queries)
    hipStream_t stream = nullptr;  // stream of the next enqueue (null: the context's)
    bool unchecked = false;
    bool have_result = false;
    uint32_t n_pairs = 0;
    bool sorted_ok = false;       // d_sorted holds the checked hits in get_pairs order
    uint4* d_sorted = nullptr;
    uint32_t d_sorted_cap = 0;
    uint32_t* sort_words = nullptr;  // device sort of the hits: the four words of every hit
    uint32_t* sort_vals = nullptr;   // two buffers of hit indices
    uint32_t* sort_scratch = nullptr;  // histograms | tile counters | two look-back buffers
    uint32_t sort_cap = 0;           // hits the sort buffers hold
    float* d_stage = nullptr;     // vertex export staging for host output
    uint64_t stage_cap = 0;
    oibvh_detect_stats stats{};
    uint64_t generation = 0;  // bumped whenever a device buffer is replaced
};

oibvh_status oi_det_init(oi_detector* d, oibvh_ctx* c);
void oi_det_free(oi_detector* d);
// upload the initial front (the vector must stay unchanged until the stream
// has consumed it)
oibvh_status oi_det_set_front(oi_detector* d, const std::vector<uint2>& front);
// size every buffer for the current query (allocates; not capturable)
oibvh_status oi_det_prepare(oi_detector* d);
// launch seed, steps and narrow phase (no allocation, no synchronisation);
// seeded: the counters were reset by a launch already queued (see
// oi_det_seed_args), so no seed launch
// done (may be null) is recorded as the query's last launch completes
oibvh_status oi_det_enqueue(oi_detector* d, bool reset_sticky, bool seeded = false, hipEvent_t done = nullptr);
// the counter reset of the prepared query, for a launch queued before it
oi_seed_args oi_det_seed_args(const oi_detector* d, bool reset_sticky);
// switch to the other counter set (an overlapped frame, before prepare)
void oi_det_flip_set(oi_detector* d);
// use counter set k (a batch's overlapped frames tie it to their slot)
void oi_det_use_set(oi_detector* d, uint32_t k);
// wait; on overflow grow the buffers and run the query again
oibvh_status oi_det_check(oi_detector* d);
// sorted hits (4 words each)
oibvh_status oi_det_get_pairs(oi_detector* d, uint32_t* out, uint64_t cap, uint32_t* n);
// the same on the device (valid until the next detect)
oibvh_status oi_det_sorted_device(oi_detector* d, const uint4** dptr, uint32_t* n);
oibvh_status oi_det_level_counts(oi_detector* d, uint32_t* out, uint64_t cap, uint32_t* n);
// six transformed vertices (18 floats) per hit in get_pairs order; geom[m] =
// positions, faces and model matrix of mesh m (device table)
oibvh_status oi_det_pair_vertices(oi_detector* d, const oi_build_desc* geom, float* out, uint64_t cap,
                                  uint32_t ptr_kind, uint32_t* n);
