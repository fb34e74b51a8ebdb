// oibvh_host.cpp -- C ABI implementation (include/oibvh.h): contexts, trees,
// scenes. Host orchestration of the reference's OibvhTree (oibvhTree.cu) and
// Scene (scene.cu): device buffers are allocated once per object and sized
// from the inputs (the reference uses fixed 10M-element buffers,
// oibvhTree.cu:11-13, scene.cu:15-21), every launch goes to the context's
// stream and only result accessors synchronise.
#include <new>

#include "oibvh_internal.h"

static void ctx_free(oibvh_ctx* ctx)
{
    (void)hipSetDevice(ctx->device);
    (void)oi_ctx_quiesce(ctx);
    if (ctx->side[0]) {
        (void)hipEventDestroy(ctx->ev_main);
        for (hipEvent_t e : ctx->ev_ring) (void)hipEventDestroy(e);
        for (hipStream_t x : ctx->side) (void)hipStreamDestroy(x);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// the latest of `tickets` (1 + a detect, 0: none) on each side stream
static void latest_per_stream(const uint64_t* tickets, uint32_t n, uint64_t* last)
{
    for (uint32_t k = 0; k < OI_SIDE_STREAMS; k++) last[k] = 0;
    for (uint32_t i = 0; i < n; i++)
        if (tickets[i]) {
            uint64_t& l = last[(tickets[i] - 1) % OI_SIDE_STREAMS];
            l = std::max(l, tickets[i]);
        }
}

hipError_t oi_ctx_wait_tickets(oibvh_ctx* c, const uint64_t* tickets, uint32_t n)
{
    uint64_t last[OI_SIDE_STREAMS];
    latest_per_stream(tickets, n, last);
    for (uint64_t t : last)
        if (t) {
            hipError_t e = hipStreamWaitEvent(c->stream, c->ev_ring[(t - 1) % OI_SIDE_RING], 0);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t oi_ctx_host_wait_tickets(oibvh_ctx* c, const uint64_t* tickets, uint32_t n)
{
    uint64_t last[OI_SIDE_STREAMS];
    latest_per_stream(tickets, n, last);
    for (uint64_t t : last)
        if (t) {
            hipError_t e = hipEventSynchronize(c->ev_ring[(t - 1) % OI_SIDE_RING]);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t oi_ctx_join(oibvh_ctx* ctx)
{
    if (!ctx->side_pending) return hipSuccess;
    ctx->side_pending = false;
    // the last detect of each side stream
    uint64_t t[OI_SIDE_STREAMS];
    for (uint32_t k = 0; k < OI_SIDE_STREAMS; k++) t[k] = ctx->side_seq > k ? ctx->side_seq - k : 0;
    return oi_ctx_wait_tickets(ctx, t, OI_SIDE_STREAMS);
}

hipError_t oi_ctx_quiesce(oibvh_ctx* ctx)
{
    hipError_t e = oi_ctx_join(ctx);
    for (hipStream_t x : ctx->side)
        if (e == hipSuccess && x) e = hipStreamSynchronize(x);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e;
}

// the side streams of overlapped frames and their events (first use)
oibvh_status oi_ctx_side_init(oibvh_ctx* c)
{
    if (c->side[0]) return OIBVH_OK;
    hipStream_t s[OI_SIDE_STREAMS] = {};
    hipError_t e = hipSuccess;
    for (uint32_t k = 0; k < OI_SIDE_STREAMS && e == hipSuccess; k++)
        e = hipStreamCreateWithFlags(&s[k], hipStreamNonBlocking);
    // recorded by kernel launches (hipExtLaunchKernel's stop event), so
    // created as ordinary timing events
    if (e == hipSuccess) e = hipEventCreate(&c->ev_main);
    for (uint32_t k = 0; k < OI_SIDE_RING && e == hipSuccess; k++) e = hipEventCreate(&c->ev_ring[k]);
    if (e != hipSuccess) {
        if (c->ev_main) (void)hipEventDestroy(c->ev_main);
        c->ev_main = nullptr;
        for (hipEvent_t& x : c->ev_ring) {
            if (x) (void)hipEventDestroy(x);
            x = nullptr;
        }
        for (hipStream_t x : s)
            if (x) (void)hipStreamDestroy(x);
        return oi_set_err(c, OIBVH_ERR_HIP, "side streams: %s", hipGetErrorString(e));
    }
    std::copy(s, s + OI_SIDE_STREAMS, c->side);
    return OIBVH_OK;
}

// a context outlives its trees, scenes and batches: destroying it while they
// are live only marks it, and the last object to go frees it
void oi_ctx_release(oibvh_ctx* ctx)
{
    if (--ctx->live == 0 && ctx->destroy_pending) ctx_free(ctx);
}

oibvh_status oi_set_err(oibvh_ctx* ctx, oibvh_status st, const char* fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return st;
}

// ---------------------------------------------------------------------------
struct oibvh_tree {
    oibvh_ctx* ctx = nullptr;
    uint32_t F = 0, V = 0, N = 0, L = 0, vl = 0;
    float* pos_own = nullptr;   // owned positions
    const float* pos = nullptr; // current positions (owned or borrowed)
    uint32_t* faces = nullptr;  // original order
    uint32_t* sfaces = nullptr; // leaf order
    uint32_t* perm = nullptr;   // leaf -> original
    float* nodes = nullptr;     // nbuf[cur]
    // node buffers: overlapped frames refit into them in turn, each into
    // the one the detects of the frames before it are not reading (the
    // others are allocated at the first such frames); read_seq[k] is 1 + the
    // side-stream detect that last read nbuf[k] (0: none)
    float* nbuf[OI_FRAME_DEPTH] = {};
    uint32_t cur = 0;
    uint64_t read_seq[OI_FRAME_DEPTH] = {};
    // build scratch (allocated at the first build)
    uint32_t* keys = nullptr;
    uint32_t* keys_hi = nullptr; // high words of 63-bit keys
    uint32_t* keys_alt = nullptr;
    uint32_t* vals_tmp = nullptr;
    uint32_t* sort_scratch = nullptr;
    uint32_t tiles = 0;
    uint32_t morton_bits = 30;
    float* box = nullptr;        // 6 floats + per-block partial boxes
    float* xf = nullptr;         // model matrix (16 floats, device)
    const float* xf_cur = nullptr; // the matrix later launches read: xf, or a scene's table / caller memory
    bool has_xf = false;
    bool built = false;
};

struct oibvh_scene {
    oibvh_ctx* ctx = nullptr;
    std::vector<oibvh_tree*> trees;
    // mesh tables of scenes of more than four trees, one per detect counter
    // set: an overlapped frame writes its own while the previous frame's
    // detect may still read the other
    oi_mesh_desc* d_desc[OI_FRAME_DEPTH] = {};
    uint32_t desc_cap[OI_FRAME_DEPTH] = {};
    std::vector<oi_mesh_desc> h_desc;  // host copies stay alive until the next detect
    oi_build_desc* d_geom = nullptr;   // per-tree geometry for the vertex export
    uint32_t geom_cap = 0;
    float* d_xf = nullptr;             // model matrices copied by oibvh_scene_set_transforms (16 per tree)
    uint32_t xf_cap = 0;
    // optional HIP-event timing of the refit's stage-1 launches and of the
    // detects (pairs of events, read and reset by oibvh_scene_get_timing)
    uint32_t timing = 0;              // 0: off; n: every n-th refit and detect
    uint64_t seq_refit = 0, seq_detect = 0;
    std::vector<hipEvent_t> ev_refit, ev_detect;
    uint32_t n_refit = 0, n_detect = 0;
    std::vector<uint2> h_front;
    uint32_t front_key[3] = {~0u, ~0u, ~0u}; // entry, flags&SELF, #trees
    oi_detector det;
    uint32_t last_entry = 0, last_expand = 1, last_flags = 0;
    // scratch of oibvh_scene_build, grown on demand: per-face keys, alternate
    // keys, values, sorted values and 63-bit high words over the concatenated
    // faces of up to four trees, and the sort's digit histograms (one set per
    // tree), pass counters and two look-back buffers
    uint32_t *b_keys = nullptr, *b_keys_alt = nullptr, *b_vals_tmp = nullptr, *b_vals = nullptr;
    uint32_t *b_keys_hi = nullptr, *b_sort = nullptr;
    uint64_t b_cap = 0, b_hi_cap = 0, b_sort_cap = 0;
};

// ---------------------------------------------------------------- context
extern "C" oibvh_status oibvh_ctx_create(int device, void* stream, oibvh_ctx** out)
{
    if (!out) return OIBVH_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return OIBVH_ERR_NO_DEVICE;
    oibvh_ctx* c = new (std::nothrow) oibvh_ctx;
    if (!c) return OIBVH_ERR_INVALID;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return OIBVH_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return OIBVH_ERR_NO_DEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        delete c;
        return OIBVH_ERR_NO_DEVICE;
    }
    c->cus = prop.multiProcessorCount;
    // the no-allocation build's scratch address on this device (a failure
    // is reported by that entry point, which tries again)
    (void)oi_noalloc_attach(device);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->clock_khz = (uint64_t)khz;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return OIBVH_ERR_HIP;
        }
        c->own_stream = true;
    }
    *out = c;
    return OIBVH_OK;
}

extern "C" void oibvh_ctx_destroy(oibvh_ctx* ctx)
{
    if (!ctx) return;
    if (ctx->live > 0) {
        ctx->destroy_pending = true;
        return;
    }
    ctx_free(ctx);
}

extern "C" const char* oibvh_last_error(const oibvh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" oibvh_status oibvh_ctx_synchronize(oibvh_ctx* ctx)
{
    if (!ctx) return OIBVH_ERR_INVALID;
    HIPCHK(ctx, oi_ctx_quiesce(ctx));
    return OIBVH_OK;
}

extern "C" void* oibvh_ctx_stream(oibvh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" oibvh_status oibvh_ctx_set_tuning(oibvh_ctx* ctx, const oibvh_tuning* t)
{
    if (!ctx || !t) return OIBVH_ERR_INVALID;
    const uint32_t lpt = t->reduce_leaves_per_thread;
    if (lpt != 0 && lpt != 1 && lpt != 2 && lpt != 4 && lpt != 8 && lpt != 16)
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "reduce_leaves_per_thread must be 0, 1, 2, 4, 8 or 16");
    if (t->front_capacity && (t->front_capacity & (t->front_capacity - 1)))
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "front_capacity must be a power of two");
    if (t->traverse_expand > 3) return oi_set_err(ctx, OIBVH_ERR_INVALID, "traverse_expand must be 0..3");
    if (t->traverse_local > 2048 || t->traverse_tail > 2048)
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "traverse_local and traverse_tail must be 0..2048");
    if (t->refit_blocks > (1u << 20)) return oi_set_err(ctx, OIBVH_ERR_INVALID, "refit_blocks must be 0..2^20");
    ctx->leaves_per_thread = lpt;
    ctx->trav_blocks = t->traverse_blocks;
    ctx->trav_expand = t->traverse_expand;
    ctx->queue_cap = t->front_capacity ? t->front_capacity : (1u << 22);
    ctx->cand_cap = t->candidate_capacity ? t->candidate_capacity : (1u << 20);
    ctx->hits_cap = t->hits_capacity ? t->hits_capacity : (1u << 20);
    ctx->trav_schedule = t->traverse_schedule;
    ctx->trav_local = t->traverse_local;
    ctx->trav_local_levels = t->traverse_local_levels;
    ctx->trav_tail = t->traverse_tail;
    ctx->trav_tail_steps = t->traverse_tail_steps;
    ctx->narrow_separate = t->narrow_separate;
    ctx->refit_blocks = t->refit_blocks;
    ctx->tuning_gen++;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_ctx_get_tuning(const oibvh_ctx* ctx, oibvh_tuning* t)
{
    if (!ctx || !t) return OIBVH_ERR_INVALID;
    t->reduce_leaves_per_thread = ctx->leaves_per_thread;
    t->traverse_blocks = ctx->trav_blocks;
    t->traverse_expand = ctx->trav_expand;
    t->front_capacity = ctx->queue_cap;
    t->candidate_capacity = ctx->cand_cap;
    t->traverse_schedule = ctx->trav_schedule;
    t->traverse_local = ctx->trav_local;
    t->traverse_local_levels = ctx->trav_local_levels;
    t->traverse_tail = ctx->trav_tail;
    t->traverse_tail_steps = ctx->trav_tail_steps;
    t->narrow_separate = ctx->narrow_separate;
    t->hits_capacity = ctx->hits_cap;
    t->refit_blocks = ctx->refit_blocks;
    return OIBVH_OK;
}

// ---------------------------------------------------------------- trees
static void tree_free(oibvh_tree* t)
{
    dfree(t->pos_own);
    dfree(t->faces);
    dfree(t->sfaces);
    dfree(t->perm);
    for (float*& b : t->nbuf) dfree(b);
    t->nodes = nullptr;
    dfree(t->keys);
    dfree(t->keys_hi);
    dfree(t->keys_alt);
    dfree(t->vals_tmp);
    dfree(t->sort_scratch);
    dfree(t->box);
    dfree(t->xf);
}

static oibvh_status tree_alloc_common(oibvh_tree* t)
{
    oibvh_ctx* c = t->ctx;
    t->L = oi_ilog2(oi_np2(t->F));
    t->vl = oi_np2(t->F) - t->F;
    t->N = oi_tree_size(t->F);
    HIPCHK(c, dalloc(&t->pos_own, 3ull * t->V));
    HIPCHK(c, dalloc(&t->faces, 3ull * t->F));
    HIPCHK(c, dalloc(&t->sfaces, 3ull * t->F));
    HIPCHK(c, dalloc(&t->perm, t->F));
    HIPCHK(c, dalloc(&t->nbuf[0], 6ull * t->N));
    t->nodes = t->nbuf[0];
    HIPCHK(c, dalloc(&t->box, 6ull + 6ull * 256));
    HIPCHK(c, dalloc(&t->xf, 16));
    t->pos = t->pos_own;
    return OIBVH_OK;
}

uint32_t oi_box_parts(uint32_t V) { return std::max(1u, std::min(256u, (V + 2047u) / 2048u)); }

// the tree as a one-mesh build table
static oi_build_desc tree_desc(const oibvh_tree* t)
{
    oi_build_desc d{};
    d.pos = t->pos;
    d.xf = t->has_xf ? t->xf_cur : nullptr;
    d.faces = t->faces;
    d.sfaces = t->sfaces;
    d.perm = t->perm;
    d.nodes = t->nodes;
    d.box = t->box;
    d.F = t->F;
    d.V = t->V;
    d.L = t->L;
    d.vl = t->vl;
    d.fbase = 0;
    d.pbase = 0;
    d.nparts = oi_box_parts(t->V);
    d.vchunk = (t->V + d.nparts - 1) / d.nparts;
    return d;
}

static oibvh_status set_positions(oibvh_tree* t, const float* xyz, uint32_t ptr_kind)
{
    oibvh_ctx* c = t->ctx;
    if (ptr_kind == OIBVH_PTR_DEVICE_BORROW) {
        t->pos = xyz;
        return OIBVH_OK;
    }
    HIPCHK(c, hipMemcpyAsync(t->pos_own, xyz, 12ull * t->V, kind_of(ptr_kind), c->stream));
    t->pos = t->pos_own;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_create(oibvh_ctx* ctx, const float* xyz, uint32_t V, const uint32_t* idx,
                                         uint32_t F, uint32_t ptr_kind, oibvh_tree** out)
{
    if (!out) return OIBVH_ERR_INVALID;
    *out = nullptr;
    if (!ctx || !xyz || !idx || V == 0 || F == 0 || ptr_kind > OIBVH_PTR_DEVICE_BORROW)
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "oibvh_tree_create: invalid argument");
    if (F > (1u << 24))
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "oibvh_tree_create: at most 2^24 faces per tree");
    if (ptr_kind == OIBVH_PTR_HOST) {
        for (uint64_t i = 0; i < 3ull * F; i++)
            if (idx[i] >= V) return oi_set_err(ctx, OIBVH_ERR_INVALID, "face index %u out of range", idx[i]);
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    oibvh_tree* t = new (std::nothrow) oibvh_tree;
    if (!t) return OIBVH_ERR_INVALID;
    t->ctx = ctx;
    t->F = F;
    t->V = V;
    oibvh_status st = tree_alloc_common(t);
    if (st == OIBVH_OK) {
        hipError_t e = hipMemcpyAsync(t->faces, idx, 12ull * F,
                                      ptr_kind == OIBVH_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                                      ctx->stream);
        if (e != hipSuccess) st = oi_set_err(ctx, OIBVH_ERR_HIP, "face upload: %s", hipGetErrorString(e));
    }
    if (st == OIBVH_OK) st = set_positions(t, xyz, ptr_kind);
    if (st == OIBVH_OK && ptr_kind == OIBVH_PTR_HOST) {
        // host inputs may be freed by the caller right after return
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = oi_set_err(ctx, OIBVH_ERR_HIP, "upload: %s", hipGetErrorString(e));
    }
    if (st != OIBVH_OK) {
        tree_free(t);
        delete t;
        return st;
    }
    ctx->live++;
    *out = t;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_create_like(oibvh_tree* other, const float* xyz, uint32_t V, uint32_t ptr_kind,
                                              oibvh_tree** out)
{
    if (!out) return OIBVH_ERR_INVALID;
    *out = nullptr;
    if (!other) return OIBVH_ERR_INVALID;
    oibvh_ctx* ctx = other->ctx;
    if (!xyz || V != other->V || ptr_kind > OIBVH_PTR_DEVICE_BORROW)
        return oi_set_err(ctx, OIBVH_ERR_INVALID, "oibvh_tree_create_like: invalid argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    oibvh_tree* t = new (std::nothrow) oibvh_tree;
    if (!t) return OIBVH_ERR_INVALID;
    t->ctx = ctx;
    t->F = other->F;
    t->V = V;
    oibvh_status st = tree_alloc_common(t);
    hipStream_t s = ctx->stream;
    if (st == OIBVH_OK) {
        hipError_t e = hipMemcpyAsync(t->faces, other->faces, 12ull * t->F, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(t->sfaces, other->sfaces, 12ull * t->F, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(t->perm, other->perm, 4ull * t->F, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(t->nodes, other->nodes, 24ull * t->N, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) st = oi_set_err(ctx, OIBVH_ERR_HIP, "create_like copy: %s", hipGetErrorString(e));
    }
    if (st == OIBVH_OK) st = set_positions(t, xyz, ptr_kind);
    if (st == OIBVH_OK && ptr_kind == OIBVH_PTR_HOST) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = oi_set_err(ctx, OIBVH_ERR_HIP, "upload: %s", hipGetErrorString(e));
    }
    if (st != OIBVH_OK) {
        tree_free(t);
        delete t;
        return st;
    }
    t->built = other->built;
    t->morton_bits = other->morton_bits;
    ctx->live++;
    *out = t;
    return OIBVH_OK;
}

extern "C" void oibvh_tree_destroy(oibvh_tree* t)
{
    if (!t) return;
    oibvh_ctx* c = t->ctx;
    (void)hipSetDevice(c->device);
    (void)oi_ctx_quiesce(c);
    tree_free(t);
    delete t;
    oi_ctx_release(c);
}

// reduce variant for F faces per launch: the wave-subtree kernel (16) once
// there are about as many of its 1024-leaf workgroups as CUs, else M leaves
// per thread with >= ~2 workgroups per CU
int oi_default_lpt(uint64_t F, int cus)
{
    if (F >= (1ull << 18)) return 16;
    const uint64_t target = 2ull * (uint64_t)(cus > 0 ? cus : 256) * 256ull;
    int m = 1;
    while (m < 8 && F / (2ull * m) >= target) m *= 2;
    return m;
}

static int pick_lpt(const oibvh_tree* t)
{
    uint32_t lpt = t->ctx->leaves_per_thread;
    if (lpt) return (int)lpt;
    return oi_default_lpt(t->F, t->ctx->cus);
}

static oibvh_status launch_reduce(oibvh_tree* t, bool build)
{
    oibvh_ctx* c = t->ctx;
    const int M = pick_lpt(t);
    const uint32_t h = oi_reduce_height(M, t->L);
    const oi_build_desc d = tree_desc(t);
    HIPCHK(c, oi_launch_reduce(c->stream, M, nullptr, nullptr, &d, oi_lvl_cnt(t->L - h, t->L, t->vl),
                               build ? t->perm : nullptr, 0));
    if (t->L > h) HIPCHK(c, oi_launch_reduce_apex(c->stream, M, nullptr, nullptr, &d, 1, nullptr));
    return OIBVH_OK;
}

// The launches of one build (OibvhTree::build, oibvhTree.cu:237-388) of the
// mesh in d (d.box: the Morton box followed by the mesh-box partials) with
// scratch s: mesh box, Morton keys, stable LSD radix sort (values end in
// d.perm), then the reduce in build mode (writes d.sfaces and d.nodes). No
// allocation and no synchronisation, so oibvh_build_noalloc can share it.
oibvh_status oi_enqueue_build(oibvh_ctx* c, const oi_build_desc& d, const oi_build_scratch& s, int M,
                              hipStream_t stream)
{
    const uint32_t F = d.F;
    const uint32_t tiles = (F + oi_sort_tile() - 1) / oi_sort_tile();
    const bool m63 = s.keys_hi != nullptr;
    uint32_t* ghist = s.sort;
    uint32_t* tctr = s.sort + OI_SORT_HIST;
    uint32_t* lb[2] = {s.sort + OI_SORT_HEAD, s.sort + OI_SORT_HEAD + (uint64_t)OI_SORT_LB * tiles};
    hipStream_t st = stream ? stream : c->stream;
    // the mesh-box launch also clears the histograms, tile counters and the
    // first look-back buffer (a kernel, not a memset: capturable)
    HIPCHK(c, oi_launch_mesh_box(st, nullptr, nullptr, &d, d.nparts, d.box + 6, s.sort,
                                 OI_SORT_HEAD + OI_SORT_LB * tiles));
    const uint32_t mblocks = (F + oi_morton_chunk() - 1) / oi_morton_chunk();
    // 30-bit keys: three passes of 10-bit digits; 63-bit keys: four 8-bit
    // passes over the low word (keys move with the values), then four over
    // the high word looked up by face id. Values end in perm.
    const uint32_t bits = m63 ? 8u : OI_M30_BITS;
    HIPCHK(c, oi_launch_morton(st, nullptr, nullptr, &d, mblocks, d.box + 6, s.keys, s.keys_hi, ghist, nullptr,
                               bits));
    const int np = m63 ? 8 : OI_M30_PASSES, nlo = m63 ? 4 : OI_M30_PASSES;
    for (int p = 0; p < np; p++) {
        const bool lo = p < nlo;
        const uint32_t* kin = lo ? ((p & 1) ? s.keys_alt : s.keys) : nullptr;
        uint32_t* kout = (lo && p + 1 < nlo) ? ((p & 1) ? s.keys : s.keys_alt) : nullptr;
        // the passes alternate between perm and vals_tmp, the last writing perm
        uint32_t* vout = ((np - 1 - p) & 1) ? s.vals_tmp : d.perm;
        const uint32_t* vin = p == 0 ? nullptr : (((np - p) & 1) ? s.vals_tmp : d.perm);
        const uint32_t shift = bits * (uint32_t)(p % nlo);
        HIPCHK(c, oi_launch_sort_pass(st, kin, vin, kout, vout, F, shift, ghist + (p << bits), lb[p & 1],
                                      p + 1 < np ? lb[(p + 1) & 1] : nullptr, tctr + p, lo ? nullptr : s.keys_hi,
                                      bits));
    }
    const uint32_t h = oi_reduce_height(M, d.L);
    HIPCHK(c, oi_launch_reduce(st, M, nullptr, nullptr, &d, oi_lvl_cnt(d.L - h, d.L, d.vl), d.perm, 0));
    if (d.L > h) HIPCHK(c, oi_launch_reduce_apex(st, M, nullptr, nullptr, &d, 1, nullptr));
    return OIBVH_OK;
}

int oi_build_lpt(const oibvh_ctx* c, uint32_t F)
{
    return c->leaves_per_thread ? (int)c->leaves_per_thread : oi_default_lpt(F, c->cus);
}

// the tree's build scratch (allocated at its first build)
static oibvh_status build_scratch(oibvh_tree* t, oi_build_scratch* out)
{
    oibvh_ctx* c = t->ctx;
    const uint32_t F = t->F;
    const uint32_t tiles = (F + oi_sort_tile() - 1) / oi_sort_tile();
    const bool m63 = t->morton_bits == 63;
    if (!t->keys) {
        HIPCHK(c, dalloc(&t->keys, F));
        HIPCHK(c, dalloc(&t->keys_alt, F));
        HIPCHK(c, dalloc(&t->vals_tmp, F));
        t->tiles = tiles;
        HIPCHK(c, dalloc(&t->sort_scratch, OI_SORT_HEAD + 2ull * OI_SORT_LB * tiles));
    }
    if (m63 && !t->keys_hi) HIPCHK(c, dalloc(&t->keys_hi, F));
    *out = oi_build_scratch{t->keys, t->keys_alt, m63 ? t->keys_hi : nullptr, t->vals_tmp, t->sort_scratch};
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_build(oibvh_tree* t)
{
    if (!t) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = t->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    oi_build_scratch s;
    oibvh_status st = build_scratch(t, &s);
    if (st == OIBVH_OK) st = oi_enqueue_build(c, tree_desc(t), s, pick_lpt(t));
    if (st != OIBVH_OK) return st;
    t->built = true;
    return OIBVH_OK;
}

// key width of later builds: 30 (oibvh.cu:6-22) or 63 (21 bits per axis, no
// reference counterpart)
extern "C" oibvh_status oibvh_tree_set_morton_bits(oibvh_tree* t, uint32_t bits)
{
    if (!t) return OIBVH_ERR_INVALID;
    if (bits != 30 && bits != 63) return oi_set_err(t->ctx, OIBVH_ERR_INVALID, "morton bits must be 30 or 63");
    t->morton_bits = bits;
    return OIBVH_OK;
}

// Mesh::transform (mesh.cpp:189-214): every later build, refit and detect
// sees positions M * (x, y, z, 1); null restores the identity
extern "C" oibvh_status oibvh_tree_set_transform(oibvh_tree* t, const float* mat4)
{
    if (!t) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = t->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    if (!mat4) {
        t->has_xf = false;
        return OIBVH_OK;
    }
    HIPCHK(c, oi_launch_set_xf(c->stream, t->xf, mat4));
    t->xf_cur = t->xf;
    t->has_xf = true;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_refit(oibvh_tree* t, const float* xyz, uint32_t ptr_kind)
{
    if (!t) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = t->ctx;
    if (ptr_kind > OIBVH_PTR_DEVICE_BORROW) return oi_set_err(c, OIBVH_ERR_INVALID, "bad ptr_kind");
    if (!t->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "refit before build");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    if (xyz) {
        oibvh_status st = set_positions(t, xyz, ptr_kind);
        if (st != OIBVH_OK) return st;
    }
    oibvh_status st = launch_reduce(t, false);
    if (st == OIBVH_OK && xyz && ptr_kind == OIBVH_PTR_HOST) HIPCHK(c, hipStreamSynchronize(c->stream));
    return st;
}

extern "C" oibvh_status oibvh_tree_set_positions(oibvh_tree* t, const float* xyz, uint32_t ptr_kind)
{
    if (!t || !xyz) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = t->ctx;
    if (ptr_kind > OIBVH_PTR_DEVICE_BORROW) return oi_set_err(c, OIBVH_ERR_INVALID, "bad ptr_kind");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    oibvh_status st = set_positions(t, xyz, ptr_kind);
    if (st == OIBVH_OK && ptr_kind == OIBVH_PTR_HOST) HIPCHK(c, hipStreamSynchronize(c->stream));
    return st;
}

extern "C" oibvh_status oibvh_tree_get_depth(const oibvh_tree* t, uint32_t* d)
{
    if (!t || !d) return OIBVH_ERR_INVALID;
    *d = oi_ilog2(t->N);  // getDepth: ilog2(m_aabbTree.size()), oibvhTree.cu:46-49
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_get_prim_count(const oibvh_tree* t, uint32_t* n)
{
    if (!t || !n) return OIBVH_ERR_INVALID;
    *n = t->F;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_get_node_count(const oibvh_tree* t, uint32_t* n)
{
    if (!t || !n) return OIBVH_ERR_INVALID;
    *n = t->N;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_get_vertex_count(const oibvh_tree* t, uint32_t* n)
{
    if (!t || !n) return OIBVH_ERR_INVALID;
    *n = t->V;
    return OIBVH_OK;
}

static oibvh_status readback(oibvh_tree* t, void* dst, const void* src, size_t bytes)
{
    oibvh_ctx* c = t->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_tree_get_nodes(oibvh_tree* t, float* out, uint64_t cap, uint32_t* n)
{
    if (!t) return OIBVH_ERR_INVALID;
    if (n) *n = t->N;
    if (!t->built) return oi_set_err(t->ctx, OIBVH_ERR_NOT_BUILT, "tree not built");
    if (!out || cap < t->N) return oi_set_err(t->ctx, OIBVH_ERR_CAPACITY, "need %u nodes", t->N);
    return readback(t, out, t->nodes, 24ull * t->N);
}

extern "C" oibvh_status oibvh_tree_get_sorted_faces(oibvh_tree* t, uint32_t* out, uint64_t cap)
{
    if (!t) return OIBVH_ERR_INVALID;
    if (!t->built) return oi_set_err(t->ctx, OIBVH_ERR_NOT_BUILT, "tree not built");
    if (!out || cap < t->F) return oi_set_err(t->ctx, OIBVH_ERR_CAPACITY, "need %u faces", t->F);
    return readback(t, out, t->sfaces, 12ull * t->F);
}

extern "C" oibvh_status oibvh_tree_get_perm(oibvh_tree* t, uint32_t* out, uint64_t cap)
{
    if (!t) return OIBVH_ERR_INVALID;
    if (!t->built) return oi_set_err(t->ctx, OIBVH_ERR_NOT_BUILT, "tree not built");
    if (!out || cap < t->F) return oi_set_err(t->ctx, OIBVH_ERR_CAPACITY, "need %u faces", t->F);
    return readback(t, out, t->perm, 4ull * t->F);
}

extern "C" oibvh_status oibvh_tree_get_positions(oibvh_tree* t, float* out, uint64_t cap)
{
    if (!t) return OIBVH_ERR_INVALID;
    if (!out || cap < 3ull * t->V) return oi_set_err(t->ctx, OIBVH_ERR_CAPACITY, "need %u floats", 3 * t->V);
    return readback(t, out, t->pos, 12ull * t->V);
}

extern "C" oibvh_status oibvh_tree_positions_device(oibvh_tree* t, float** dptr)
{
    if (!t || !dptr) return OIBVH_ERR_INVALID;
    *dptr = const_cast<float*>(t->pos);
    return OIBVH_OK;
}

// ---------------------------------------------------------------- scenes
extern "C" oibvh_status oibvh_scene_create(oibvh_ctx* ctx, oibvh_scene** out)
{
    if (!out) return OIBVH_ERR_INVALID;
    *out = nullptr;
    if (!ctx) return OIBVH_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    oibvh_scene* s = new (std::nothrow) oibvh_scene;
    if (!s) return OIBVH_ERR_INVALID;
    s->ctx = ctx;
    oibvh_status r = oi_det_init(&s->det, ctx);
    if (r != OIBVH_OK) {
        delete s;
        return r;
    }
    ctx->live++;
    *out = s;
    return OIBVH_OK;
}

extern "C" void oibvh_scene_destroy(oibvh_scene* s)
{
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)oi_ctx_quiesce(s->ctx);
    for (uint32_t k = 0; k < OI_FRAME_DEPTH; k++) dfree(s->d_desc[k]);
    dfree(s->d_geom);
    dfree(s->d_xf);
    dfree(s->b_keys);
    dfree(s->b_keys_alt);
    dfree(s->b_vals_tmp);
    dfree(s->b_vals);
    dfree(s->b_keys_hi);
    dfree(s->b_sort);
    for (hipEvent_t e : s->ev_refit) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->ev_detect) (void)hipEventDestroy(e);
    oi_det_free(&s->det);
    oibvh_ctx* c = s->ctx;
    delete s;
    oi_ctx_release(c);
}

extern "C" oibvh_status oibvh_scene_add_tree(oibvh_scene* s, oibvh_tree* t)
{
    if (!s || !t) return OIBVH_ERR_INVALID;
    if (t->ctx != s->ctx) return oi_set_err(s->ctx, OIBVH_ERR_INVALID, "tree from another context");
    if (!t->built) return oi_set_err(s->ctx, OIBVH_ERR_NOT_BUILT, "addOibvhTree: tree not built");
    s->trees.push_back(t);
    s->det.have_result = false;
    return OIBVH_OK;
}

// initial BVTT front, Scene::expandBvttNodes (scene.cu:192-223): real nodes at
// level min(L, entry) of mesh i x those of mesh j, i < j; in self mode also the
// unordered pairs (x <= y) of each mesh's entry-level nodes.
static void make_front(oibvh_scene* s, uint32_t entry, bool self)
{
    s->h_front.clear();
    const uint32_t n = (uint32_t)s->trees.size();
    std::vector<uint32_t> ibase(n);
    uint64_t acc = 0;
    for (uint32_t m = 0; m < n; m++) {
        ibase[m] = (uint32_t)acc;
        acc += (2ull << s->trees[m]->L) - 1ull;
    }
    auto lvl_range = [&](uint32_t m, uint32_t& a, uint32_t& b) {
        const oibvh_tree* t = s->trees[m];
        const uint32_t l = std::min(t->L, entry);
        a = (1u << l) - 1u;
        b = oi_lvl_last_implicit(l, t->L, t->vl);
    };
    for (uint32_t i = 0; i < n; i++) {
        uint32_t a0, a1;
        lvl_range(i, a0, a1);
        if (self)
            for (uint32_t x = a0; x <= a1; x++)
                for (uint32_t y = x; y <= a1; y++) s->h_front.push_back(make_uint2(ibase[i] + x, ibase[i] + y));
        for (uint32_t j = i + 1; j < n; j++) {
            uint32_t b0, b1;
            lvl_range(j, b0, b1);
            for (uint32_t x = a0; x <= a1; x++)
                for (uint32_t y = b0; y <= b1; y++) s->h_front.push_back(make_uint2(ibase[i] + x, ibase[j] + y));
        }
    }
}

// size the detect buffers and set up the query of s->last_* (allocates)
static oibvh_status prepare_scene_detect(oibvh_scene* s)
{
    oibvh_ctx* c = s->ctx;
    oi_detector& d = s->det;
    const uint32_t n = (uint32_t)s->trees.size();
    // mesh table (positions may be borrowed and change between frames)
    std::vector<oi_mesh_desc>& desc = s->h_desc;
    desc.resize(n);
    uint64_t acc = 0;
    uint32_t depth = 0;
    for (uint32_t m = 0; m < n; m++) {
        const oibvh_tree* t = s->trees[m];
        oi_mesh_desc& md = desc[m];
        std::memset(&md, 0, sizeof md);
        md.nodes = t->nodes;
        md.sfaces = t->sfaces;
        md.pos = t->pos;
        md.perm = t->perm;
        md.xf = t->has_xf ? t->xf_cur : nullptr;
        md.F = t->F;
        md.L = t->L;
        md.vl = t->vl;
        md.N = t->N;
        md.ibase = (uint32_t)acc;
        acc += (2ull << t->L) - 1ull;
        depth = std::max(depth, t->L - std::min(t->L, s->last_entry));
    }
    if (acc > OI_IMPLICIT_LIMIT)
        return oi_set_err(c, OIBVH_ERR_INVALID, "scene too large: %llu implicit BVTT indices (limit %u)",
                          (unsigned long long)acc, OI_IMPLICIT_LIMIT);
    // small scenes pass the table in the kernel arguments; larger ones use the
    // table of the detect's counter set, which no queued detect reads any
    // more (an overlapped frame has made the stream wait for the last one
    // that used this set)
    const uint32_t ks = d.set;
    if (n > 4) {
        if (s->desc_cap[ks] < n) {
            HIPCHK(c, oi_ctx_quiesce(c));
            dfree(s->d_desc[ks]);
            s->desc_cap[ks] = 0;
            HIPCHK(c, dalloc(&s->d_desc[ks], n));
            s->desc_cap[ks] = n;
        }
        HIPCHK(c, hipMemcpyAsync(s->d_desc[ks], desc.data(), n * sizeof(oi_mesh_desc), hipMemcpyHostToDevice,
                                 c->stream));
    }
    const bool self = (s->last_flags & OIBVH_DETECT_SELF) != 0;
    const uint32_t key[3] = {s->last_entry, self ? 1u : 0u, n};
    if (std::memcmp(key, s->front_key, sizeof key) != 0) {
        make_front(s, s->last_entry, self);
        oibvh_status r = oi_det_set_front(&d, s->h_front);
        if (r != OIBVH_OK) return r;
        std::memcpy(s->front_key, key, sizeof key);
    }
    d.tab = oi_tab_args{n > 4 ? s->d_desc[ks] : nullptr, desc.data(), n, nullptr, 0};
    d.depth = depth;
    d.expand = c->trav_expand ? c->trav_expand : std::min(3u, std::max(1u, s->last_expand));
    d.schedule = c->trav_schedule;
    d.flags = s->last_flags;
    d.implicit_end = (uint32_t)acc;
    return oi_det_prepare(&d);
}

static oibvh_status enqueue_detect(oibvh_scene* s, bool reset_sticky)
{
    oibvh_status r = prepare_scene_detect(s);
    if (r != OIBVH_OK) return r;
    return oi_det_enqueue(&s->det, reset_sticky);
}

// the next event pair of a timing list (created on first use)
static oibvh_status timing_pair(oibvh_ctx* c, std::vector<hipEvent_t>& ev, uint32_t& n, hipEvent_t* a, hipEvent_t* b)
{
    if (2ull * n + 2 > ev.size()) {
        for (int k = 0; k < 2; k++) {
            hipEvent_t e = nullptr;
            HIPCHK(c, hipEventCreate(&e));
            ev.push_back(e);
        }
    }
    *a = ev[2 * n];
    *b = ev[2 * n + 1];
    n++;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_detect(oibvh_scene* s, uint32_t entry, uint32_t expand, uint32_t flags)
{
    if (!s) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    for (oibvh_tree* t : s->trees)
        if (!t->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "detect: tree not built");
    const bool async = (flags & OIBVH_DETECT_ASYNC) != 0;
    oi_detector& d = s->det;
    if (d.unchecked && !async) {
        oibvh_status r = oi_det_check(&d);
        if (r != OIBVH_OK) return r;
    }
    s->last_entry = entry;
    s->last_expand = expand;
    s->last_flags = flags;
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    if (s->trees.empty() || (s->trees.size() == 1 && !(flags & OIBVH_DETECT_SELF))) {
        d.have_result = true;
        d.unchecked = false;
        d.sorted_ok = false;
        d.n_levels = 0;
        return OIBVH_OK;
    }
    // an asynchronous detect keeps the error history of unchecked earlier
    // ones (reported as async_errors); a synchronous one starts afresh
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->timing && (s->seq_detect++ % s->timing) == 0) {
        oibvh_status r = timing_pair(c, s->ev_detect, s->n_detect, &e0, &e1);
        if (r != OIBVH_OK) return r;
        HIPCHK(c, hipEventRecord(e0, c->stream));
    }
    oibvh_status r = enqueue_detect(s, !async);
    if (r != OIBVH_OK) return r;
    if (e1) HIPCHK(c, hipEventRecord(e1, c->stream));
    if (async) return OIBVH_OK;
    return oi_det_check(&d);
}

// Mesh::transform (mesh.cpp:189-214) for every tree of the scene in one call:
// host matrices are staged per tree in stream order, device ones copied into
// the scene's table with one copy, borrowed ones read in place; each tree then
// reads its matrix from there
extern "C" oibvh_status oibvh_scene_set_transforms(oibvh_scene* s, const float* mats, uint32_t ptr_kind)
{
    if (!s || !mats || ptr_kind > OIBVH_PTR_DEVICE_BORROW) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    const uint32_t n = (uint32_t)s->trees.size();
    if (n == 0) return OIBVH_OK;
    HIPCHK(c, hipSetDevice(c->device));
    if (ptr_kind == OIBVH_PTR_DEVICE_BORROW) {
        // no device work: launches queued later read the new pointers, earlier
        // ones keep the ones they were given
        for (uint32_t i = 0; i < n; i++) {
            s->trees[i]->xf_cur = mats + 16ull * i;
            s->trees[i]->has_xf = true;
        }
        return OIBVH_OK;
    }
    HIPCHK(c, oi_ctx_join(c));
    if (ptr_kind == OIBVH_PTR_HOST) {
        for (uint32_t i = 0; i < n; i++) {
            oibvh_tree* t = s->trees[i];
            HIPCHK(c, oi_launch_set_xf(c->stream, t->xf, mats + 16ull * i));
            t->xf_cur = t->xf;
            t->has_xf = true;
        }
        return OIBVH_OK;
    }
    if (s->xf_cap < n) {
        // the table may still be read by queued work
        HIPCHK(c, oi_ctx_quiesce(c));
        dfree(s->d_xf);
        s->xf_cap = 0;
        HIPCHK(c, dalloc(&s->d_xf, 16ull * n));
        s->xf_cap = n;
    }
    HIPCHK(c, hipMemcpyAsync(s->d_xf, mats, 64ull * n, hipMemcpyDeviceToDevice, c->stream));
    for (uint32_t i = 0; i < n; i++) {
        s->trees[i]->xf_cur = s->d_xf + 16ull * i;
        s->trees[i]->has_xf = true;
    }
    return OIBVH_OK;
}

// Every tree of the scene refit at once: up to four trees per launch pair
// (their descriptors travel in the kernel arguments), so the reduce fills the
// GPU once instead of once per tree. Same results as oibvh_tree_refit on each.
// the scene refit; with seed, the last apex launch (if there is one) also
// resets the counters of the detect that follows, and *seeded says so
static oibvh_status scene_refit(oibvh_scene* s, const oi_seed_args* seed, bool* seeded, hipEvent_t done = nullptr)
{
    oibvh_ctx* c = s->ctx;
    if (seeded) *seeded = false;
    HIPCHK(c, hipSetDevice(c->device));
    for (oibvh_tree* t : s->trees)
        if (!t->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "scene refit: tree not built");
    const uint32_t per = oi_inline_meshes();
    for (size_t i0 = 0; i0 < s->trees.size(); i0 += per) {
        const uint32_t n = (uint32_t)std::min<size_t>(per, s->trees.size() - i0);
        uint64_t F = 0;
        for (uint32_t k = 0; k < n; k++) F += s->trees[i0 + k]->F;
        int M = (int)c->leaves_per_thread;
        if (!M) M = oi_default_lpt(F, c->cus);
        oi_build_desc d[8], apex[8];
        uint32_t boff[9], na = 0;
        boff[0] = 0;
        for (uint32_t k = 0; k < n; k++) {
            const oibvh_tree* t = s->trees[i0 + k];
            d[k] = tree_desc(t);
            const uint32_t h = oi_reduce_height(M, t->L);
            boff[k + 1] = boff[k] + oi_lvl_cnt(t->L - h, t->L, t->vl);
            if (t->L > h) apex[na++] = d[k];
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (s->timing && i0 == 0 && (s->seq_refit++ % s->timing) == 0) {
            oibvh_status r = timing_pair(c, s->ev_refit, s->n_refit, &e0, &e1);
            if (r != OIBVH_OK) return r;
        }
        // the timing pair is recorded by the stage-1 launch itself
        HIPCHK(c, oi_launch_reduce_inline(c->stream, M, d, n, boff, nullptr, 0, e0, e1, c->refit_blocks));
        const bool last = i0 + per >= s->trees.size();
        const oi_seed_args* sd = (last && seed && na) ? seed : nullptr;
        // `done` (may be null) is recorded as the refit's last launch completes
        if (na) HIPCHK(c, oi_launch_reduce_apex_inline(c->stream, M, apex, na, sd, last ? done : nullptr));
        if (last && done && !na) HIPCHK(c, hipEventRecord(done, c->stream));
        if (sd && seeded) *seeded = true;
    }
    return OIBVH_OK;
}

// trees per launch of the scene build; tile counters of its sort (one per
// pass and tree)
#define OI_SCENE_BUILD_GROUP 4
#define OI_SCENE_CTRS (8u * OI_SCENE_BUILD_GROUP)

// The scene build's scratch for a group of n trees with F faces and `tiles`
// sort tiles in all (grown on demand, after the context's streams drain).
static oibvh_status scene_build_scratch(oibvh_scene* s, uint64_t F, uint32_t n, uint32_t tiles, bool m63)
{
    oibvh_ctx* c = s->ctx;
    const uint64_t words = (uint64_t)n * OI_SORT_HIST + OI_SCENE_CTRS + 2ull * OI_SORT_LB * tiles;
    const bool grow = F > s->b_cap || words > s->b_sort_cap || (m63 && F > s->b_hi_cap);
    if (!grow) return OIBVH_OK;
    HIPCHK(c, oi_ctx_quiesce(c));
    if (F > s->b_cap) {
        dfree(s->b_keys);
        dfree(s->b_keys_alt);
        dfree(s->b_vals_tmp);
        dfree(s->b_vals);
        s->b_cap = 0;
        HIPCHK(c, dalloc(&s->b_keys, F));
        HIPCHK(c, dalloc(&s->b_keys_alt, F));
        HIPCHK(c, dalloc(&s->b_vals_tmp, F));
        HIPCHK(c, dalloc(&s->b_vals, F));
        s->b_cap = F;
    }
    if (m63 && F > s->b_hi_cap) {
        dfree(s->b_keys_hi);
        s->b_hi_cap = 0;
        HIPCHK(c, dalloc(&s->b_keys_hi, F));
        s->b_hi_cap = F;
    }
    if (words > s->b_sort_cap) {
        dfree(s->b_sort);
        s->b_sort_cap = 0;
        HIPCHK(c, dalloc(&s->b_sort, words));
        s->b_sort_cap = words;
    }
    return OIBVH_OK;
}

// One launch group of the scene build: n <= 4 consecutive trees of one
// Morton width, their build descriptors over the concatenated faces and the
// launch geometry.
struct SceneBuildGroup {
    uint32_t n = 0, tiles = 0;
    bool m63 = false;
    int M = 0;
    uint64_t F = 0;
    oi_build_desc d[OI_SCENE_BUILD_GROUP];
    uint32_t box_off[OI_SCENE_BUILD_GROUP + 1] = {0}, mort_off[OI_SCENE_BUILD_GROUP + 1] = {0};
    uint32_t red_off[OI_SCENE_BUILD_GROUP + 1] = {0};
    uint32_t base[OI_SCENE_BUILD_GROUP] = {0}, nf[OI_SCENE_BUILD_GROUP] = {0};
};

static void plan_group(const oibvh_ctx* c, oibvh_tree* const* trees, uint32_t n, SceneBuildGroup& g)
{
    const uint32_t tile = oi_sort_tile(), chunk = oi_morton_chunk();
    g.n = n;
    g.m63 = trees[0]->morton_bits == 63;
    for (uint32_t k = 0; k < n; k++) {
        oi_build_desc& d = g.d[k];
        d = tree_desc(trees[k]);
        d.fbase = (uint32_t)g.F;
        g.base[k] = (uint32_t)g.F;
        g.nf[k] = d.F;
        g.box_off[k + 1] = g.box_off[k] + d.nparts;
        g.mort_off[k + 1] = g.mort_off[k] + (d.F + chunk - 1) / chunk;
        g.tiles += (d.F + tile - 1) / tile;
        g.F += d.F;
    }
    g.M = (int)c->leaves_per_thread;
    if (!g.M) g.M = oi_default_lpt(g.F, c->cus);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t h = oi_reduce_height(g.M, g.d[k].L);
        g.red_off[k + 1] = g.red_off[k] + oi_lvl_cnt(g.d[k].L - h, g.d[k].L, g.d[k].vl);
    }
}

// OibvhTree::build (oibvhTree.cu:237-388) of one group, every stage one
// launch for all its trees: mesh boxes, Morton keys over the concatenated
// faces (each tree's own digit histograms), the LSD passes sorting each
// tree's keys as a segment of their own, and the reduce, which also writes
// each tree's permutation from the sorted values. Enqueues only (capturable).
static oibvh_status enqueue_group(oibvh_scene* s, const SceneBuildGroup& g)
{
    oibvh_ctx* c = s->ctx;
    const uint32_t n = g.n, tiles = g.tiles;
    hipStream_t st = c->stream;
    uint32_t* ghist = s->b_sort;
    uint32_t* ctrs = ghist + (uint64_t)n * OI_SORT_HIST;
    uint32_t* lb[2] = {ctrs + OI_SCENE_CTRS, ctrs + OI_SCENE_CTRS + (uint64_t)OI_SORT_LB * tiles};
    // the mesh-box launch also clears the histograms, counters and the first
    // look-back buffer (a kernel, not a memset: capturable)
    HIPCHK(c, oi_launch_mesh_box_inline(st, g.d, n, g.box_off, ghist,
                                        n * OI_SORT_HIST + OI_SCENE_CTRS + OI_SORT_LB * tiles));
    // 30-bit keys: three passes of 10-bit digits; 63-bit keys: four 8-bit
    // passes over the low words, then four over the high words looked up by
    // value. The sorted values (face ids over the concatenated faces) end in
    // b_vals.
    const uint32_t bits = g.m63 ? 8u : OI_M30_BITS;
    HIPCHK(c, oi_launch_morton_inline(st, g.d, n, g.mort_off, s->b_keys, g.m63 ? s->b_keys_hi : nullptr, ghist,
                                      OI_SORT_HIST, bits));
    const int np = g.m63 ? 8 : OI_M30_PASSES, nlo = g.m63 ? 4 : OI_M30_PASSES;
    for (int p = 0; p < np; p++) {
        const bool lo = p < nlo;
        const uint32_t* kin = lo ? ((p & 1) ? s->b_keys_alt : s->b_keys) : nullptr;
        uint32_t* kout = (lo && p + 1 < nlo) ? ((p & 1) ? s->b_keys : s->b_keys_alt) : nullptr;
        // the passes alternate between b_vals and b_vals_tmp, the last writing b_vals
        uint32_t* vout = ((np - 1 - p) & 1) ? s->b_vals_tmp : s->b_vals;
        const uint32_t* vin = p == 0 ? nullptr : (((np - p) & 1) ? s->b_vals_tmp : s->b_vals);
        HIPCHK(c, oi_launch_sort_pass_segs(st, kin, vin, kout, vout, n, g.base, g.nf, bits * (uint32_t)(p % nlo),
                                           ghist + (p << bits), OI_SORT_HIST, lb[p & 1],
                                           p + 1 < np ? lb[(p + 1) & 1] : nullptr, ctrs + p * OI_SCENE_BUILD_GROUP,
                                           lo ? nullptr : s->b_keys_hi, bits));
    }
    uint32_t na = 0;
    oi_build_desc apex[OI_SCENE_BUILD_GROUP];
    for (uint32_t k = 0; k < n; k++)
        if (g.d[k].L > oi_reduce_height(g.M, g.d[k].L)) apex[na++] = g.d[k];
    HIPCHK(c, oi_launch_reduce_inline(st, g.M, g.d, n, g.red_off, s->b_vals, 1, nullptr, nullptr, 0));
    if (na) HIPCHK(c, oi_launch_reduce_apex_inline(st, g.M, apex, na, nullptr, nullptr));
    return OIBVH_OK;
}

// OibvhTree::build on every tree of the scene (oibvhTree.cu:237-388). A single
// tree's build is a chain of small launches that leave most of the GPU idle,
// so up to four consecutive trees of the same Morton width share each launch
// of the chain. Plain launches: replaying them from a captured hipGraph was
// measured slower (164 against 157 us for two 1M-face trees, 94 against 90 us
// for two 20K-face trees), the graph's launch costing more than the host
// enqueues, which overlap the GPU's work. Results equal oibvh_tree_build on
// each tree.
extern "C" oibvh_status oibvh_scene_build(oibvh_scene* s)
{
    if (!s) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    const uint32_t per = std::min<uint32_t>(OI_SCENE_BUILD_GROUP, std::min(oi_inline_meshes(), oi_sort_segs_max()));
    std::vector<SceneBuildGroup> groups;
    uint64_t Fmax = 0;
    uint32_t nmax = 0, tmax = 0;
    bool m63 = false;
    for (size_t i = 0; i < s->trees.size();) {
        uint32_t n = 1;
        while (n < per && i + n < s->trees.size() && s->trees[i + n]->morton_bits == s->trees[i]->morton_bits) n++;
        groups.emplace_back();
        plan_group(c, &s->trees[i], n, groups.back());
        const SceneBuildGroup& g = groups.back();
        if (g.F > 0xffffffffull) return oi_set_err(c, OIBVH_ERR_CAPACITY, "scene build: too many faces in one group");
        Fmax = std::max(Fmax, g.F);
        nmax = std::max(nmax, g.n);
        tmax = std::max(tmax, g.tiles);
        m63 = m63 || g.m63;
        i += n;
    }
    if (groups.empty()) return OIBVH_OK;
    oibvh_status r = scene_build_scratch(s, Fmax, nmax, tmax, m63);
    if (r != OIBVH_OK) return r;
    for (const SceneBuildGroup& g : groups)
        if ((r = enqueue_group(s, g)) != OIBVH_OK) return r;
    for (oibvh_tree* t : s->trees) t->built = true;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_refit(oibvh_scene* s)
{
    if (!s) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    return scene_refit(s, nullptr, nullptr);
}

// An overlapped frame (OIBVH_FRAME_OVERLAP): every tree is refit on the
// context's stream into its next node buffer, once the last detect that read
// that buffer has finished (a wait on the host), and the detect runs on a
// side stream after the refit. The refit of the next frame so runs while
// this frame's detect does, and consecutive detects, on alternate side
// streams with their own counter sets, may run at once; any other call first
// orders the context's stream after them (oi_ctx_join).
static oibvh_status overlapped_frame(oibvh_scene* s, uint32_t entry, uint32_t expand, uint32_t flags)
{
    oibvh_ctx* c = s->ctx;
    oi_detector& d = s->det;
    const bool async = (flags & OIBVH_DETECT_ASYNC) != 0;
    oibvh_status r = oi_ctx_side_init(c);
    if (r != OIBVH_OK) return r;
    if (d.unchecked && !async && (r = oi_det_check(&d)) != OIBVH_OK) return r;
    // wait for the side detects that read a buffer about to be rewritten or
    // used the counter set (and its working buffers) about to be reset
    bool set_unchecked[OI_FRAME_DEPTH];
    std::copy(d.set_unchecked, d.set_unchecked + OI_FRAME_DEPTH, set_unchecked);
    const uint32_t prev_set = d.set;
    oi_det_flip_set(&d);
    uint64_t need[2] = {d.set_seq[d.set], 0};
    for (oibvh_tree* t : s->trees) {
        const uint32_t k = (t->cur + 1u) % OI_FRAME_DEPTH;
        if (!t->nbuf[k]) {
            HIPCHK(c, oi_ctx_quiesce(c));  // the allocation is the first of this tree's buffer k
            HIPCHK(c, dalloc(&t->nbuf[k], 6ull * t->N));
        }
        need[1] = std::max(need[1], t->read_seq[k]);
        t->cur = k;
        t->nodes = t->nbuf[k];
    }
    // on the host: a stream wait's barrier packet held the context's stream
    // about 10 us per frame (2 x 1M faces), and that detect, OI_FRAME_DEPTH
    // frames back, has usually ended; the host stays at most that far ahead
    HIPCHK(c, oi_ctx_host_wait_tickets(c, need, 2));
    s->last_entry = entry;
    s->last_expand = expand;
    s->last_flags = flags;
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    if ((r = prepare_scene_detect(s)) != OIBVH_OK) {
        // nothing was queued: the trees keep the buffers they were refit into
        for (oibvh_tree* t : s->trees) {
            t->cur = (t->cur + OI_FRAME_DEPTH - 1u) % OI_FRAME_DEPTH;
            t->nodes = t->nbuf[t->cur];
        }
        oi_det_use_set(&d, prev_set);
        std::copy(set_unchecked, set_unchecked + OI_FRAME_DEPTH, d.set_unchecked);
        return r;
    }
    // the refit's apex launch also resets this frame's counter set
    const oi_seed_args seed = oi_det_seed_args(&d, !async);
    bool seeded = false;
    if ((r = scene_refit(s, &seed, &seeded, c->ev_main)) != OIBVH_OK) return r;
    const uint64_t q = c->side_seq;
    hipStream_t side = oi_side_stream(c, q);
    HIPCHK(c, hipStreamWaitEvent(side, c->ev_main, 0));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->timing && (s->seq_detect++ % s->timing) == 0) {
        if ((r = timing_pair(c, s->ev_detect, s->n_detect, &e0, &e1)) != OIBVH_OK) return r;
        HIPCHK(c, hipEventRecord(e0, side));
    }
    // the detect's last launch records this detect's ring event
    d.stream = side;
    r = oi_det_enqueue(&d, !async, seeded, c->ev_ring[q % OI_SIDE_RING]);
    d.stream = nullptr;
    if (r != OIBVH_OK) return r;
    c->side_seq++;
    if (e1) HIPCHK(c, hipEventRecord(e1, side));
    for (oibvh_tree* t : s->trees) t->read_seq[t->cur] = q + 1;
    d.set_seq[d.set] = q + 1;
    c->side_pending = true;
    if (async) return OIBVH_OK;
    return oi_det_check(&d);
}

// The reference's per-frame sequence (main.cpp:284 after the trees' refit):
// point trees at new vertex memory, refit every tree, detect. One call, so a
// frame costs the host one entry into the library.
extern "C" oibvh_status oibvh_scene_frame(oibvh_scene* s, const float* const* positions, uint32_t entry,
                                          uint32_t expand, uint32_t flags)
{
    if (!s) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    if (positions)
        for (size_t i = 0; i < s->trees.size(); i++)
            if (positions[i]) s->trees[i]->pos = positions[i];
    const bool async = (flags & OIBVH_DETECT_ASYNC) != 0;
    oi_detector& d = s->det;
    const bool trivial = s->trees.empty() || (s->trees.size() == 1 && !(flags & OIBVH_DETECT_SELF));
    for (oibvh_tree* t : s->trees)
        if (!t->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "frame: tree not built");
    if ((flags & OIBVH_FRAME_OVERLAP) && !trivial) return overlapped_frame(s, entry, expand, flags);
    HIPCHK(c, oi_ctx_join(c));
    if (trivial) {
        oibvh_status r = scene_refit(s, nullptr, nullptr);
        if (r != OIBVH_OK) return r;
        return oibvh_scene_detect(s, entry, expand, flags);
    }
    // as oibvh_scene_detect, with the detect's counter reset folded into the
    // refit's apex launch: check an earlier synchronous result first, size
    // the query, refit, then the detect's own launches
    if (d.unchecked && !async) {
        oibvh_status r = oi_det_check(&d);
        if (r != OIBVH_OK) return r;
    }
    s->last_entry = entry;
    s->last_expand = expand;
    s->last_flags = flags;
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    oibvh_status r = prepare_scene_detect(s);
    if (r != OIBVH_OK) return r;
    const oi_seed_args seed = oi_det_seed_args(&d, !async);
    bool seeded = false;
    if ((r = scene_refit(s, &seed, &seeded)) != OIBVH_OK) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->timing && (s->seq_detect++ % s->timing) == 0) {
        if ((r = timing_pair(c, s->ev_detect, s->n_detect, &e0, &e1)) != OIBVH_OK) return r;
        HIPCHK(c, hipEventRecord(e0, c->stream));
    }
    if ((r = oi_det_enqueue(&d, !async, seeded)) != OIBVH_OK) return r;
    if (e1) HIPCHK(c, hipEventRecord(e1, c->stream));
    if (async) return OIBVH_OK;
    return oi_det_check(&d);
}

extern "C" oibvh_status oibvh_scene_set_timing(oibvh_scene* s, uint32_t every)
{
    if (!s) return OIBVH_ERR_INVALID;
    s->timing = every;
    s->seq_refit = s->seq_detect = 0;
    s->n_refit = s->n_detect = 0;
    return OIBVH_OK;
}

static oibvh_status sum_pairs(oibvh_ctx* c, const std::vector<hipEvent_t>& ev, uint32_t n, double* ms)
{
    *ms = 0.0;
    for (uint32_t k = 0; k < n; k++) {
        HIPCHK(c, hipEventSynchronize(ev[2 * k + 1]));
        float v = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&v, ev[2 * k], ev[2 * k + 1]));
        *ms += v;
    }
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_get_timing(oibvh_scene* s, oibvh_scene_timing* out)
{
    if (!s || !out) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    std::memset(out, 0, sizeof *out);
    oibvh_status r = sum_pairs(c, s->ev_refit, s->n_refit, &out->refit_ms);
    if (r == OIBVH_OK) r = sum_pairs(c, s->ev_detect, s->n_detect, &out->detect_ms);
    if (r != OIBVH_OK) return r;
    out->refits = s->n_refit;
    out->detects = s->n_detect;
    s->n_refit = s->n_detect = 0;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_get_pair_count(oibvh_scene* s, uint32_t* count)
{
    if (!s || !count) return OIBVH_ERR_INVALID;
    oibvh_status r = oi_det_check(&s->det);
    if (r != OIBVH_OK) return r;
    *count = s->det.n_pairs;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_get_stats(oibvh_scene* s, oibvh_detect_stats* st)
{
    if (!s || !st) return OIBVH_ERR_INVALID;
    oibvh_status r = oi_det_check(&s->det);
    if (r != OIBVH_OK) return r;
    *st = s->det.stats;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_hits_device(oibvh_scene* s, void** dptr, uint32_t* n)
{
    if (!s || !dptr || !n) return OIBVH_ERR_INVALID;
    oibvh_status r = oi_det_check(&s->det);
    if (r != OIBVH_OK) return r;
    *dptr = s->det.hits;
    *n = s->det.n_pairs;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_scene_pairs_device(oibvh_scene* s, const uint32_t** dptr, uint32_t* n)
{
    if (!s || !dptr || !n) return OIBVH_ERR_INVALID;
    HIPCHK(s->ctx, hipSetDevice(s->ctx->device));
    const uint4* p = nullptr;
    oibvh_status r = oi_det_sorted_device(&s->det, &p, n);
    *dptr = reinterpret_cast<const uint32_t*>(p);
    return r;
}

extern "C" oibvh_status oibvh_scene_get_pairs(oibvh_scene* s, uint32_t* out, uint64_t cap, uint32_t* n)
{
    if (!s) return OIBVH_ERR_INVALID;
    return oi_det_get_pairs(&s->det, out, cap, n);
}

// Scene::convertToVertexArray (scene.cu:69-93), gathered on the device
extern "C" oibvh_status oibvh_scene_get_pair_vertices(oibvh_scene* s, float* out, uint64_t cap, uint32_t ptr_kind,
                                                      uint32_t* n)
{
    if (!s) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = s->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, oi_ctx_join(c));
    const uint32_t nt = (uint32_t)s->trees.size();
    if (s->det.have_result && s->det.n_pairs && nt) {
        std::vector<oi_build_desc> g(nt);
        for (uint32_t m = 0; m < nt; m++) g[m] = tree_desc(s->trees[m]);
        if (s->geom_cap < nt) {
            dfree(s->d_geom);
            s->geom_cap = 0;
            HIPCHK(c, dalloc(&s->d_geom, nt));
            s->geom_cap = nt;
        }
        // synchronous: the table is read by the export launched below
        HIPCHK(c, hipMemcpyAsync(s->d_geom, g.data(), nt * sizeof(oi_build_desc), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return oi_det_pair_vertices(&s->det, s->d_geom, out, cap, ptr_kind, n);
}

extern "C" oibvh_status oibvh_scene_get_level_counts(oibvh_scene* s, uint32_t* out, uint64_t cap, uint32_t* n)
{
    if (!s || !n) return OIBVH_ERR_INVALID;
    return oi_det_level_counts(&s->det, out, cap, n);
}
