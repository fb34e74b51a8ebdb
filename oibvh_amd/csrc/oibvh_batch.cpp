// oibvh_batch.cpp -- many independent scenes built, refit and queried
// together (no reference counterpart; the reference handles one scene of a
// few meshes, scene.cu:225-446).
//
// Every mesh of every scene lives in shared concatenated buffers, and every
// stage is one launch over all meshes: the build and refit kernels walk a
// (mesh, chunk) work table, the Morton sort sorts all faces at once (four
// 8-bit Morton passes, then 8-bit passes over the mesh id that partition the
// faces back into their meshes, stably), and the BVTT steps and the narrow
// phase consume one front that holds the pairs of all scenes. Meshes of a
// scene are tested against each other (and, in self mode, each against
// itself); meshes of different scenes never meet: the initial front only
// pairs meshes of one scene. A frame (refit + detect) can be replayed from a
// captured hipGraph.
//
// Overlapped frames (OIBVH_FRAME_OVERLAP) keep two slots of per-frame state:
// node buffer, model-matrix table and descriptor tables. Frame f refits into
// one slot on the context's stream while frame f - 1's detect reads the other
// on a side stream (the context's two alternate by frame); staging calls
// after an overlapped frame write the slot the next one will refit, once the
// detect that last read it has finished.
#include <new>

#include "oibvh_internal.h"

#define OI_CHUNK_SHIFT 12u  // granularity of a batch's implicit index space

struct oibvh_batch {
    oibvh_ctx* ctx = nullptr;
    // meshes as added (host staging until the first build)
    struct Mesh {
        uint32_t scene, V, F;
        uint32_t src;  // mesh whose vertices and faces this one shares (itself if owned)
        uint32_t L, vl, N;
        uint64_t vbase, fbase, nbase, ibase;  // offsets in the concatenated buffers
        uint32_t pbase, nparts, vchunk;
    };
    std::vector<Mesh> meshes;
    std::vector<std::vector<float>> h_xyz;      // per owned mesh (freed after upload)
    std::vector<std::vector<uint32_t>> h_idx;
    bool committed = false, built = false;
    uint64_t Vt = 0, Ft = 0, Nt = 0, It = 0, Pt = 0;
    uint32_t n_scenes = 0;
    // device
    float* pos = nullptr;
    uint32_t* faces = nullptr;
    uint32_t* sfaces = nullptr;
    uint32_t* perm = nullptr;
    float* boxes = nullptr;
    float* partials = nullptr;
    bool has_xf = false;
    // per-frame state: slot[cur] is the one the last build or refit wrote
    // (and the one every non-overlapped call uses); the other exists once
    // an overlapped frame has run
    struct Slot {
        float* nodes = nullptr;
        float* xf = nullptr;                  // model matrices (16 floats per mesh), if has_xf
        oi_build_desc* d_bdesc = nullptr;     // build/refit table
        oi_mesh_desc* d_mdesc = nullptr;      // traversal table
        std::vector<oi_build_desc> h_bdesc;   // host copies (positions track k_set_mesh_pos updates)
        std::vector<oi_mesh_desc> h_mdesc;
        uint64_t read_seq = 0;                // 1 + the side-stream detect that last read this slot (0: none)
    };
    Slot slot[2];
    uint32_t cur = 0;
    bool ov_active = false;     // the last frame was overlapped: staging goes to the other slot
    bool staged_other = false;  // the other slot has been opened for the next overlapped frame
    bool borrowed = false;  // some mesh reads caller-owned vertex memory
    uint32_t *keys = nullptr, *keys_alt = nullptr, *vals = nullptr, *vals_tmp = nullptr, *fmesh = nullptr;
    uint32_t* keys_hi = nullptr;       // high words of 63-bit keys
    uint32_t morton_bits = 30;
    uint32_t* sort_scratch = nullptr;  // histograms | tile counters | lookback0 | lookback1
    uint32_t* mesh_hist = nullptr;     // digit histograms of the mesh ids, 256 per pass
    uint32_t tiles = 0, mesh_passes = 0;
    uint2* w_box = nullptr;
    uint32_t n_box = 0;
    uint2* w_morton = nullptr;
    uint32_t n_morton = 0;
    uint2* w_reduce = nullptr;
    uint32_t n_reduce = 0;
    uint2* w_apex = nullptr;
    uint32_t n_apex = 0;
    int reduce_M = 0;  // M the reduce tables were made for
    // traversal
    uint32_t* chunk_mesh = nullptr;
    std::vector<uint2> h_front;
    uint32_t front_key[2] = {~0u, ~0u};  // entry, self
    oi_detector det;
    // captured graphs: the whole frame, and an overlapped frame's refit and
    // detect halves, per slot
    struct Graph {
        hipGraphExec_t exec = nullptr;
        uint64_t key[8] = {};
    };
    Graph g_frame[2], g_refit[2], g_detect[2];
    uint32_t desc_gen = 0;
    // HIP-event timing of the refit's stage-1 launch and of detects
    uint32_t timing = 0;
    uint64_t seq_refit = 0, seq_detect = 0;
    std::vector<hipEvent_t> ev_refit, ev_detect;
    uint32_t n_refit = 0, n_detect = 0;
};

static inline oibvh_batch::Slot& cur_slot(oibvh_batch* b) { return b->slot[b->cur]; }

static oibvh_status batch_err(oibvh_batch* b, oibvh_status st, const char* what)
{
    return oi_set_err(b->ctx, st, "%s", what);
}

extern "C" oibvh_status oibvh_batch_create(oibvh_ctx* ctx, oibvh_batch** out)
{
    if (!out) return OIBVH_ERR_INVALID;
    *out = nullptr;
    if (!ctx) return OIBVH_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    oibvh_batch* b = new (std::nothrow) oibvh_batch;
    if (!b) return OIBVH_ERR_INVALID;
    b->ctx = ctx;
    oibvh_status r = oi_det_init(&b->det, ctx);
    if (r != OIBVH_OK) {
        delete b;
        return r;
    }
    ctx->live++;
    *out = b;
    return OIBVH_OK;
}

static void batch_free_device(oibvh_batch* b)
{
    dfree(b->pos);
    dfree(b->faces);
    dfree(b->sfaces);
    dfree(b->perm);
    for (oibvh_batch::Slot& sl : b->slot) {
        dfree(sl.nodes);
        dfree(sl.xf);
        dfree(sl.d_bdesc);
        dfree(sl.d_mdesc);
    }
    dfree(b->boxes);
    dfree(b->partials);
    dfree(b->keys);
    dfree(b->keys_alt);
    dfree(b->keys_hi);
    dfree(b->vals);
    dfree(b->vals_tmp);
    dfree(b->fmesh);
    dfree(b->sort_scratch);
    dfree(b->mesh_hist);
    dfree(b->w_box);
    dfree(b->w_morton);
    dfree(b->w_reduce);
    dfree(b->w_apex);
    dfree(b->chunk_mesh);
}

extern "C" void oibvh_batch_destroy(oibvh_batch* b)
{
    if (!b) return;
    oibvh_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)oi_ctx_quiesce(c);
    for (oibvh_batch::Graph* gs : {b->g_frame, b->g_refit, b->g_detect})
        for (int k = 0; k < 2; k++)
            if (gs[k].exec) (void)hipGraphExecDestroy(gs[k].exec);
    for (hipEvent_t e : b->ev_refit) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ev_detect) (void)hipEventDestroy(e);
    batch_free_device(b);
    oi_det_free(&b->det);
    delete b;
    oi_ctx_release(c);
}

extern "C" oibvh_status oibvh_batch_add_mesh(oibvh_batch* b, uint32_t scene, const float* xyz, uint32_t V,
                                             const uint32_t* idx, uint32_t F, uint32_t* index)
{
    if (!b) return OIBVH_ERR_INVALID;
    if (b->committed) return batch_err(b, OIBVH_ERR_INVALID, "oibvh_batch_add_mesh: batch already built");
    if (!xyz || !idx || V == 0 || F == 0) return batch_err(b, OIBVH_ERR_INVALID, "oibvh_batch_add_mesh: invalid argument");
    if (F > (1u << 24)) return batch_err(b, OIBVH_ERR_INVALID, "oibvh_batch_add_mesh: at most 2^24 faces per mesh");
    for (uint64_t i = 0; i < 3ull * F; i++)
        if (idx[i] >= V) return oi_set_err(b->ctx, OIBVH_ERR_INVALID, "face index %u out of range", idx[i]);
    oibvh_batch::Mesh m{};
    m.scene = scene;
    m.V = V;
    m.F = F;
    m.src = (uint32_t)b->meshes.size();
    b->meshes.push_back(m);
    b->h_xyz.emplace_back(xyz, xyz + 3ull * V);
    b->h_idx.emplace_back(idx, idx + 3ull * F);
    if (index) *index = m.src;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_add_instance(oibvh_batch* b, uint32_t scene, uint32_t source, uint32_t* index)
{
    if (!b) return OIBVH_ERR_INVALID;
    if (b->committed) return batch_err(b, OIBVH_ERR_INVALID, "oibvh_batch_add_instance: batch already built");
    if (source >= b->meshes.size()) return batch_err(b, OIBVH_ERR_INVALID, "oibvh_batch_add_instance: bad source");
    oibvh_batch::Mesh m = b->meshes[source];
    m.scene = scene;
    m.src = b->meshes[source].src;
    b->meshes.push_back(m);
    b->h_xyz.emplace_back();
    b->h_idx.emplace_back();
    if (index) *index = (uint32_t)b->meshes.size() - 1;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_mesh_count(const oibvh_batch* b, uint32_t* n)
{
    if (!b || !n) return OIBVH_ERR_INVALID;
    *n = (uint32_t)b->meshes.size();
    return OIBVH_OK;
}

static int batch_lpt(const oibvh_batch* b)
{
    if (b->ctx->leaves_per_thread) return (int)b->ctx->leaves_per_thread;
    // enough faces for about one wave-subtree workgroup per CU: k_reduce_w
    if (b->Ft >= (1ull << 18)) return 16;
    // aim for >= ~2 workgroups per CU in the reduce
    const uint64_t target = 2ull * (uint64_t)(b->ctx->cus > 0 ? b->ctx->cus : 256) * 256ull;
    uint32_t m = 1;
    while (m < 8 && b->Ft / (2ull * m) >= target) m *= 2;
    return (int)m;
}

template <class T>
static oibvh_status upload(oibvh_ctx* c, T*& dst, const std::vector<T>& src)
{
    HIPCHK(c, dalloc(&dst, src.size()));
    if (!src.empty()) HIPCHK(c, hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return OIBVH_OK;
}

// the descriptor tables of slot k (positions: the batch's own buffers)
static void fill_descs(oibvh_batch* b, uint32_t k)
{
    oibvh_batch::Slot& sl = b->slot[k];
    const uint32_t n = (uint32_t)b->meshes.size();
    sl.h_bdesc.assign(n, oi_build_desc{});
    sl.h_mdesc.assign(n, oi_mesh_desc{});
    for (uint32_t i = 0; i < n; i++) {
        const oibvh_batch::Mesh& m = b->meshes[i];
        const oibvh_batch::Mesh& s = b->meshes[m.src];
        oi_build_desc& d = sl.h_bdesc[i];
        d.pos = b->pos + 3ull * s.vbase;
        d.xf = b->has_xf ? sl.xf + 16ull * i : nullptr;
        d.faces = b->faces + 3ull * s.fbase;
        d.sfaces = b->sfaces + 3ull * m.fbase;
        d.perm = b->perm + m.fbase;
        d.nodes = sl.nodes + 6ull * m.nbase;
        d.box = b->boxes + 6ull * i;
        d.F = m.F;
        d.V = m.V;
        d.L = m.L;
        d.vl = m.vl;
        d.fbase = (uint32_t)m.fbase;
        d.pbase = m.pbase;
        d.nparts = m.nparts;
        d.vchunk = m.vchunk;
        oi_mesh_desc& t = sl.h_mdesc[i];
        t.nodes = d.nodes;
        t.sfaces = d.sfaces;
        t.pos = d.pos;
        t.perm = d.perm;
        t.xf = d.xf;
        t.F = m.F;
        t.L = m.L;
        t.vl = m.vl;
        t.N = m.N;
        t.ibase = (uint32_t)m.ibase;
    }
}

// refill and upload the tables of every allocated slot, keeping each slot's
// vertex pointers (borrowed ones included)
static oibvh_status upload_descs(oibvh_batch* b)
{
    oibvh_ctx* c = b->ctx;
    for (uint32_t k = 0; k < 2; k++) {
        oibvh_batch::Slot& sl = b->slot[k];
        if (!sl.d_bdesc) continue;
        std::vector<const float*> keep;
        for (const oi_build_desc& d : sl.h_bdesc) keep.push_back(d.pos);
        fill_descs(b, k);
        for (size_t i = 0; i < keep.size(); i++) {
            sl.h_bdesc[i].pos = keep[i];
            sl.h_mdesc[i].pos = keep[i];
        }
        HIPCHK(c, hipMemcpy(sl.d_bdesc, sl.h_bdesc.data(), sl.h_bdesc.size() * sizeof(oi_build_desc),
                            hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(sl.d_mdesc, sl.h_mdesc.data(), sl.h_mdesc.size() * sizeof(oi_mesh_desc),
                            hipMemcpyHostToDevice));
    }
    b->desc_gen++;
    return OIBVH_OK;
}

// reduce work tables for M leaves per thread
static oibvh_status make_reduce_tables(oibvh_batch* b, int M)
{
    oibvh_ctx* c = b->ctx;
    std::vector<uint2> wr, wa;
    for (uint32_t i = 0; i < b->meshes.size(); i++) {
        const oibvh_batch::Mesh& m = b->meshes[i];
        const uint32_t h = oi_reduce_height(M, m.L);
        const uint32_t nb = oi_lvl_cnt(m.L - h, m.L, m.vl);
        for (uint32_t w = 0; w < nb; w++) wr.push_back(make_uint2(i, w));
        if (m.L > h) wa.push_back(make_uint2(i, 0));
    }
    dfree(b->w_reduce);
    dfree(b->w_apex);
    oibvh_status r;
    if ((r = upload(c, b->w_reduce, wr)) != OIBVH_OK) return r;
    if ((r = upload(c, b->w_apex, wa)) != OIBVH_OK) return r;
    b->n_reduce = (uint32_t)wr.size();
    b->n_apex = (uint32_t)wa.size();
    b->reduce_M = M;
    return OIBVH_OK;
}

static oibvh_status commit(oibvh_batch* b)
{
    if (b->committed) return OIBVH_OK;
    oibvh_ctx* c = b->ctx;
    const uint32_t n = (uint32_t)b->meshes.size();
    if (n == 0) return batch_err(b, OIBVH_ERR_INVALID, "batch has no meshes");
    // layout
    uint64_t V = 0, F = 0, N = 0, I = 0, P = 0;
    uint32_t scenes = 0;
    for (uint32_t i = 0; i < n; i++) {
        oibvh_batch::Mesh& m = b->meshes[i];
        m.L = oi_ilog2(oi_np2(m.F));
        m.vl = oi_np2(m.F) - m.F;
        m.N = oi_tree_size(m.F);
        if (m.src == i) {
            m.vbase = V;
            V += m.V;
        } else {
            m.vbase = b->meshes[m.src].vbase;
        }
        m.fbase = F;
        F += m.F;
        m.nbase = N;
        N += m.N;
        m.ibase = I;
        const uint64_t span = (2ull << m.L) - 1ull;
        I += (span + (1ull << OI_CHUNK_SHIFT) - 1) >> OI_CHUNK_SHIFT << OI_CHUNK_SHIFT;
        m.nparts = std::max(1u, std::min(256u, (m.V + 4095u) / 4096u));
        m.vchunk = (m.V + m.nparts - 1) / m.nparts;
        m.pbase = (uint32_t)P;
        P += m.nparts;
        scenes = std::max(scenes, m.scene + 1);
    }
    if (F >= (1ull << 30)) return batch_err(b, OIBVH_ERR_INVALID, "batch exceeds 2^30 faces");
    if (I > OI_IMPLICIT_LIMIT)
        return oi_set_err(c, OIBVH_ERR_INVALID, "batch too large: %llu implicit BVTT indices (limit %u)",
                          (unsigned long long)I, OI_IMPLICIT_LIMIT);
    b->Vt = V;
    b->Ft = F;
    b->Nt = N;
    b->It = I;
    b->Pt = P;
    b->n_scenes = scenes;
    HIPCHK(c, hipSetDevice(c->device));
    // geometry
    HIPCHK(c, dalloc(&b->pos, 3 * V));
    HIPCHK(c, dalloc(&b->faces, 3 * F));
    for (uint32_t i = 0; i < n; i++) {
        const oibvh_batch::Mesh& m = b->meshes[i];
        if (m.src != i) continue;
        HIPCHK(c, hipMemcpy(b->pos + 3 * m.vbase, b->h_xyz[i].data(), 12ull * m.V, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(b->faces + 3 * m.fbase, b->h_idx[i].data(), 12ull * m.F, hipMemcpyHostToDevice));
        std::vector<float>().swap(b->h_xyz[i]);
        std::vector<uint32_t>().swap(b->h_idx[i]);
    }
    HIPCHK(c, dalloc(&b->sfaces, 3 * F));
    HIPCHK(c, dalloc(&b->perm, F));
    HIPCHK(c, dalloc(&b->slot[0].nodes, 6 * N));
    HIPCHK(c, dalloc(&b->boxes, 6ull * n));
    HIPCHK(c, dalloc(&b->partials, 6 * P));
    // sort
    HIPCHK(c, dalloc(&b->keys, F));
    HIPCHK(c, dalloc(&b->keys_alt, F));
    HIPCHK(c, dalloc(&b->vals, F));
    HIPCHK(c, dalloc(&b->vals_tmp, F));
    b->mesh_passes = n <= 1 ? 0 : n <= 256 ? 1 : n <= 65536 ? 2 : 3;
    if (b->mesh_passes) HIPCHK(c, dalloc(&b->fmesh, F));
    b->tiles = (uint32_t)((F + oi_sort_tile() - 1) / oi_sort_tile());
    HIPCHK(c, dalloc(&b->sort_scratch, OI_SORT_HEAD + 512ull * b->tiles));
    {
        std::vector<uint32_t> mh(768, 0u);
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t d = 0; d < 3; d++) mh[256 * d + ((i >> (8 * d)) & 0xffu)] += b->meshes[i].F;
        oibvh_status r = upload(c, b->mesh_hist, mh);
        if (r != OIBVH_OK) return r;
    }
    // work tables
    std::vector<uint2> wb, wm;
    for (uint32_t i = 0; i < n; i++) {
        const oibvh_batch::Mesh& m = b->meshes[i];
        for (uint32_t k = 0; k < m.nparts; k++) wb.push_back(make_uint2(i, k));
        const uint32_t mc = (m.F + oi_morton_chunk() - 1) / oi_morton_chunk();
        for (uint32_t k = 0; k < mc; k++) wm.push_back(make_uint2(i, k));
    }
    oibvh_status r;
    if ((r = upload(c, b->w_box, wb)) != OIBVH_OK) return r;
    if ((r = upload(c, b->w_morton, wm)) != OIBVH_OK) return r;
    b->n_box = (uint32_t)wb.size();
    b->n_morton = (uint32_t)wm.size();
    // traversal: mesh of every 2^OI_CHUNK_SHIFT implicit indices
    {
        std::vector<uint32_t> cm(std::max<uint64_t>(1, I >> OI_CHUNK_SHIFT), 0u);
        for (uint32_t i = 0; i < n; i++) {
            const oibvh_batch::Mesh& m = b->meshes[i];
            const uint64_t span = (2ull << m.L) - 1ull;
            for (uint64_t k = m.ibase >> OI_CHUNK_SHIFT; k < (m.ibase + span + (1ull << OI_CHUNK_SHIFT) - 1) >> OI_CHUNK_SHIFT; k++)
                cm[k] = i;
        }
        if ((r = upload(c, b->chunk_mesh, cm)) != OIBVH_OK) return r;
    }
    HIPCHK(c, dalloc(&b->slot[0].d_bdesc, n));
    HIPCHK(c, dalloc(&b->slot[0].d_mdesc, n));
    fill_descs(b, 0);
    if ((r = upload_descs(b)) != OIBVH_OK) return r;
    b->committed = true;
    return OIBVH_OK;
}

// ---------------------------------------------------------------- slots
// Make slot k's buffers and tables exist (the second slot is allocated by
// the first overlapped frame).
static oibvh_status ensure_slot(oibvh_batch* b, uint32_t k)
{
    oibvh_ctx* c = b->ctx;
    oibvh_batch::Slot& sl = b->slot[k];
    const uint32_t n = (uint32_t)b->meshes.size();
    if (sl.d_bdesc && (!b->has_xf || sl.xf)) return OIBVH_OK;
    // the tables are uploaded again below with synchronous copies, which are
    // not ordered with the library's streams: let queued work finish first
    HIPCHK(c, oi_ctx_quiesce(c));
    if (!sl.nodes) HIPCHK(c, dalloc(&sl.nodes, 6 * b->Nt));
    if (b->has_xf && !sl.xf) HIPCHK(c, dalloc(&sl.xf, 16ull * n));
    if (!sl.d_bdesc) {
        HIPCHK(c, dalloc(&sl.d_bdesc, n));
        HIPCHK(c, dalloc(&sl.d_mdesc, n));
        // a new slot starts with the current slot's vertex pointers
        fill_descs(b, k);
        const oibvh_batch::Slot& cs = b->slot[b->cur];
        for (uint32_t i = 0; i < n && i < cs.h_bdesc.size(); i++) {
            sl.h_bdesc[i].pos = cs.h_bdesc[i].pos;
            sl.h_mdesc[i].pos = cs.h_mdesc[i].pos;
        }
    }
    // upload_descs refills every allocated slot (xf pointers may have appeared)
    return upload_descs(b);
}

// main stream waits for the side-stream detect that last read slot k
static oibvh_status wait_slot_readers(oibvh_batch* b, uint32_t k)
{
    oibvh_ctx* c = b->ctx;
    if (c->side[0]) HIPCHK(c, oi_ctx_wait_tickets(c, &b->slot[k].read_seq, 1));
    return OIBVH_OK;
}

// copy slot `from`'s per-frame inputs (matrices, vertex pointers) to slot
// `to`, in stream order
static oibvh_status copy_slot_inputs(oibvh_batch* b, uint32_t from, uint32_t to)
{
    oibvh_ctx* c = b->ctx;
    oibvh_batch::Slot& f = b->slot[from];
    oibvh_batch::Slot& t = b->slot[to];
    if (b->has_xf)
        HIPCHK(c, hipMemcpyAsync(t.xf, f.xf, 64ull * b->meshes.size(), hipMemcpyDeviceToDevice, c->stream));
    for (uint32_t i = 0; i < b->meshes.size(); i++)
        if (t.h_bdesc[i].pos != f.h_bdesc[i].pos) {
            HIPCHK(c, oi_launch_set_mesh_pos(c->stream, t.d_bdesc, t.d_mdesc, i, f.h_bdesc[i].pos));
            t.h_bdesc[i].pos = f.h_bdesc[i].pos;
            t.h_mdesc[i].pos = f.h_bdesc[i].pos;
        }
    return OIBVH_OK;
}

// The slot staging calls write. After an overlapped frame it is the other
// slot, opened once per frame: wait for its last reader, then carry the
// current slot's inputs over, so that meshes not staged again keep theirs.
static oibvh_status stage_slot(oibvh_batch* b, uint32_t* k)
{
    if (!b->ov_active) {
        *k = b->cur;
        return OIBVH_OK;
    }
    const uint32_t o = b->cur ^ 1u;
    if (!b->staged_other) {
        oibvh_status r = ensure_slot(b, o);
        if (r == OIBVH_OK) r = wait_slot_readers(b, o);
        if (r == OIBVH_OK) r = copy_slot_inputs(b, b->cur, o);
        if (r != OIBVH_OK) return r;
        b->staged_other = true;
    }
    *k = o;
    return OIBVH_OK;
}

// Before a call that is not an overlapped frame: order the context's stream
// after the queued side-stream detects and leave overlapped mode. A call
// that computes (build, refit, detect, frame) first takes over inputs staged
// for the next overlapped frame into the current slot.
static oibvh_status settle(oibvh_batch* b, bool compute)
{
    oibvh_ctx* c = b->ctx;
    if (b->ov_active) {
        HIPCHK(c, oi_ctx_join(c));
        b->ov_active = false;
    }
    if (compute && b->staged_other) {
        oibvh_status r = copy_slot_inputs(b, b->cur ^ 1u, b->cur);
        if (r != OIBVH_OK) return r;
        b->staged_other = false;
    }
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_set_transforms(oibvh_batch* b, const float* mats, uint32_t ptr_kind)
{
    if (!b || !mats || ptr_kind > OIBVH_PTR_DEVICE_BORROW) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = commit(b);
    if (r != OIBVH_OK) return r;
    const size_t bytes = 64ull * b->meshes.size();
    if (!b->has_xf) {
        // the first matrices: every allocated slot gets a table
        HIPCHK(c, oi_ctx_quiesce(c));
        b->has_xf = true;
        for (uint32_t k = 0; k < 2; k++) {
            oibvh_batch::Slot& sl = b->slot[k];
            if (!sl.d_bdesc) continue;
            HIPCHK(c, dalloc(&sl.xf, 16ull * b->meshes.size()));
        }
        if ((r = upload_descs(b)) != OIBVH_OK) return r;
    }
    uint32_t k = 0;
    if ((r = stage_slot(b, &k)) != OIBVH_OK) return r;
    float* dst = b->slot[k].xf;
    if (ptr_kind == OIBVH_PTR_HOST) {
        // host memory: copied before return, after the work queued before
        // (the stream has waited for the slot's last reader)
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(dst, mats, bytes, hipMemcpyHostToDevice));
    } else {
        HIPCHK(c, hipMemcpyAsync(dst, mats, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_positions_device(oibvh_batch* b, uint32_t mesh, float** dptr)
{
    if (!b || !dptr || mesh >= b->meshes.size()) return OIBVH_ERR_INVALID;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    oibvh_status r = commit(b);
    if (r != OIBVH_OK) return r;
    *dptr = b->pos + 3ull * b->meshes[b->meshes[mesh].src].vbase;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_set_positions(oibvh_batch* b, uint32_t mesh, const float* xyz, uint32_t ptr_kind)
{
    if (!b || !xyz || mesh >= b->meshes.size() || ptr_kind > OIBVH_PTR_DEVICE_BORROW) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = commit(b);
    if (r != OIBVH_OK) return r;
    const oibvh_batch::Mesh& m = b->meshes[mesh];
    if (ptr_kind == OIBVH_PTR_DEVICE_BORROW) {
        // the kernels read the descriptor tables at run time: update the
        // staged slot's in stream order (the host copies follow)
        uint32_t k = 0;
        if ((r = stage_slot(b, &k)) != OIBVH_OK) return r;
        oibvh_batch::Slot& sl = b->slot[k];
        HIPCHK(c, oi_launch_set_mesh_pos(c->stream, sl.d_bdesc, sl.d_mdesc, mesh, xyz));
        sl.h_bdesc[mesh].pos = xyz;
        sl.h_mdesc[mesh].pos = xyz;
        b->borrowed = true;
        return OIBVH_OK;
    }
    // a copy into the batch's own vertex buffer, which both slots read: the
    // queued overlapped detects finish first
    if ((r = settle(b, true)) != OIBVH_OK) return r;
    oibvh_batch::Slot& sl = cur_slot(b);
    float* own = b->pos + 3ull * b->meshes[m.src].vbase;
    if (sl.h_bdesc[mesh].pos != own) {
        HIPCHK(c, oi_launch_set_mesh_pos(c->stream, sl.d_bdesc, sl.d_mdesc, mesh, own));
        sl.h_bdesc[mesh].pos = own;
        sl.h_mdesc[mesh].pos = own;
    }
    if (ptr_kind == OIBVH_PTR_HOST) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(own, xyz, 12ull * m.V, hipMemcpyHostToDevice));
    } else {
        HIPCHK(c, hipMemcpyAsync(own, xyz, 12ull * m.V, hipMemcpyDeviceToDevice, c->stream));
    }
    return OIBVH_OK;
}

// the next event pair of a timing list (created on first use)
static oibvh_status timing_pair(oibvh_ctx* c, std::vector<hipEvent_t>& ev, uint32_t& n, hipEvent_t* e0, hipEvent_t* e1)
{
    if (2ull * n + 2 > ev.size()) {
        for (int k = 0; k < 2; k++) {
            hipEvent_t e = nullptr;
            HIPCHK(c, hipEventCreate(&e));
            ev.push_back(e);
        }
    }
    *e0 = ev[2 * n];
    *e1 = ev[2 * n + 1];
    n++;
    return OIBVH_OK;
}

// refit launches over the current slot; with seed, the apex launch (if any)
// also resets the counters of the detect that follows and *seeded says so;
// t0/t1 (both or neither) are recorded by the stage-1 launch itself;
// max_blocks caps the stage-1 launch (k_reduce_wl)
static oibvh_status enqueue_refit(oibvh_batch* b, const oi_seed_args* seed = nullptr, bool* seeded = nullptr,
                                  hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr, uint32_t max_blocks = 0)
{
    oibvh_ctx* c = b->ctx;
    const int M = b->reduce_M;
    const oibvh_batch::Slot& sl = cur_slot(b);
    if (seeded) *seeded = false;
    HIPCHK(c, oi_launch_reduce(c->stream, M, sl.d_bdesc, b->w_reduce, nullptr, b->n_reduce, nullptr, 0, t0, t1,
                               max_blocks));
    if (b->n_apex) {
        HIPCHK(c, oi_launch_reduce_apex(c->stream, M, sl.d_bdesc, b->w_apex, nullptr, b->n_apex, seed));
        if (seed && seeded) *seeded = true;
    }
    return OIBVH_OK;
}

// event pair for this refit when timing samples it
static oibvh_status refit_timing(oibvh_batch* b, hipEvent_t* t0, hipEvent_t* t1)
{
    *t0 = *t1 = nullptr;
    if (b->timing && (b->seq_refit++ % b->timing) == 0) return timing_pair(b->ctx, b->ev_refit, b->n_refit, t0, t1);
    return OIBVH_OK;
}

static oibvh_status detect_timing(oibvh_batch* b, hipEvent_t* t0, hipEvent_t* t1)
{
    *t0 = *t1 = nullptr;
    if (b->timing && (b->seq_detect++ % b->timing) == 0)
        return timing_pair(b->ctx, b->ev_detect, b->n_detect, t0, t1);
    return OIBVH_OK;
}

static oibvh_status ensure_reduce_tables(oibvh_batch* b)
{
    const int M = batch_lpt(b);
    if (M != b->reduce_M) {
        HIPCHK(b->ctx, oi_ctx_quiesce(b->ctx));
        return make_reduce_tables(b, M);
    }
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_build(oibvh_batch* b)
{
    if (!b) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = commit(b);
    if (r != OIBVH_OK) return r;
    if ((r = settle(b, true)) != OIBVH_OK) return r;
    if ((r = ensure_reduce_tables(b)) != OIBVH_OK) return r;
    hipStream_t s = c->stream;
    const oibvh_batch::Slot& sl = cur_slot(b);
    const uint32_t F = (uint32_t)b->Ft;
    const bool m63 = b->morton_bits == 63;
    if (m63 && !b->keys_hi) HIPCHK(c, dalloc(&b->keys_hi, F));
    uint32_t* ghist = b->sort_scratch;
    uint32_t* tctr = b->sort_scratch + OI_SORT_HIST;
    uint32_t* lb[2] = {b->sort_scratch + OI_SORT_HEAD, b->sort_scratch + OI_SORT_HEAD + 256ull * b->tiles};
    HIPCHK(c, oi_launch_mesh_box(s, sl.d_bdesc, b->w_box, nullptr, b->n_box, b->partials, b->sort_scratch,
                                 OI_SORT_HEAD + 256u * b->tiles));
    HIPCHK(c, oi_launch_morton(s, sl.d_bdesc, b->w_morton, nullptr, b->n_morton, b->partials, b->keys,
                               m63 ? b->keys_hi : nullptr, ghist, b->fmesh));
    // Morton passes (four over the low word with the keys moving along; for
    // 63-bit keys four more over the high word, looked up by face id), then
    // the mesh-id passes; lookback buffers alternate
    const uint32_t mpasses = m63 ? 8 : 4;
    const uint32_t passes = mpasses + b->mesh_passes;
    uint32_t* kin = b->keys;
    uint32_t* kout = b->keys_alt;
    const uint32_t* vin = nullptr;
    uint32_t* vout = b->vals_tmp;
    for (uint32_t p = 0; p < passes; p++) {
        const bool lo = p < 4;
        const bool morton = p < mpasses;
        const bool last = (p + 1 == passes);
        const uint32_t* hist = morton ? ghist + 256 * p : b->mesh_hist + 256 * (p - mpasses);
        const uint32_t shift = morton ? 8 * (p & 3) : 8 * (p - mpasses);
        uint32_t* ko = (lo && p < 3) ? kout : nullptr;
        const uint32_t* kov = lo ? nullptr : (morton ? b->keys_hi : b->fmesh);
        HIPCHK(c, oi_launch_sort_pass(s, lo ? kin : nullptr, vin, ko, vout, F, shift, hist, lb[p & 1],
                                      last ? nullptr : lb[(p + 1) & 1], tctr + p, kov));
        std::swap(kin, kout);
        vin = vout;
        vout = (vout == b->vals_tmp) ? b->vals : b->vals_tmp;
    }
    HIPCHK(c, oi_launch_reduce(s, b->reduce_M, sl.d_bdesc, b->w_reduce, nullptr, b->n_reduce, vin, 1));
    if (b->n_apex)
        HIPCHK(c, oi_launch_reduce_apex(s, b->reduce_M, sl.d_bdesc, b->w_apex, nullptr, b->n_apex, nullptr));
    b->built = true;
    return OIBVH_OK;
}

// key width of later builds: 30 (oibvh.cu:6-22) or 63 (21 bits per axis)
extern "C" oibvh_status oibvh_batch_set_morton_bits(oibvh_batch* b, uint32_t bits)
{
    if (!b) return OIBVH_ERR_INVALID;
    if (bits != 30 && bits != 63) return oi_set_err(b->ctx, OIBVH_ERR_INVALID, "morton bits must be 30 or 63");
    b->morton_bits = bits;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_refit(oibvh_batch* b)
{
    if (!b) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    if (!b->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "batch refit before build");
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = settle(b, true);
    if (r == OIBVH_OK) r = ensure_reduce_tables(b);
    if (r != OIBVH_OK) return r;
    hipEvent_t t0, t1;
    if ((r = refit_timing(b, &t0, &t1)) != OIBVH_OK) return r;
    return enqueue_refit(b, nullptr, nullptr, t0, t1);
}

// initial front: for every scene, the entry-level node pairs of every mesh
// pair i < j (Scene::expandBvttNodes, scene.cu:192-223), and in self mode
// the unordered entry-level pairs of each mesh with itself
static void make_front(oibvh_batch* b, uint32_t entry, bool self)
{
    b->h_front.clear();
    const uint32_t n = (uint32_t)b->meshes.size();
    std::vector<std::vector<uint32_t>> by_scene(b->n_scenes);
    for (uint32_t i = 0; i < n; i++) by_scene[b->meshes[i].scene].push_back(i);
    auto lvl_range = [&](uint32_t m, uint32_t& a, uint32_t& z) {
        const oibvh_batch::Mesh& t = b->meshes[m];
        const uint32_t l = std::min(t.L, entry);
        a = (1u << l) - 1u;
        z = oi_lvl_last_implicit(l, t.L, t.vl);
    };
    for (const std::vector<uint32_t>& ms : by_scene) {
        for (size_t p = 0; p < ms.size(); p++) {
            const uint32_t i = ms[p];
            const uint32_t bi = (uint32_t)b->meshes[i].ibase;
            uint32_t a0, a1;
            lvl_range(i, a0, a1);
            if (self)
                for (uint32_t x = a0; x <= a1; x++)
                    for (uint32_t y = x; y <= a1; y++) b->h_front.push_back(make_uint2(bi + x, bi + y));
            for (size_t q = p + 1; q < ms.size(); q++) {
                const uint32_t j = ms[q];
                const uint32_t bj = (uint32_t)b->meshes[j].ibase;
                uint32_t b0, b1;
                lvl_range(j, b0, b1);
                for (uint32_t x = a0; x <= a1; x++)
                    for (uint32_t y = b0; y <= b1; y++) b->h_front.push_back(make_uint2(bi + x, bj + y));
            }
        }
    }
}

// set the detector up for (entry, expand, flags) over the current slot:
// front, table, depth, buffers
static oibvh_status prepare_detect(oibvh_batch* b, uint32_t entry, uint32_t expand, uint32_t flags)
{
    oibvh_ctx* c = b->ctx;
    oi_detector& d = b->det;
    const bool self = (flags & OIBVH_DETECT_SELF) != 0;
    const uint32_t key[2] = {entry, self ? 1u : 0u};
    if (std::memcmp(key, b->front_key, sizeof key) != 0) {
        HIPCHK(c, oi_ctx_quiesce(c));  // the previous front may still be in flight
        make_front(b, entry, self);
        oibvh_status r = oi_det_set_front(&d, b->h_front);
        if (r != OIBVH_OK) return r;
        std::memcpy(b->front_key, key, sizeof key);
    }
    uint32_t depth = 0;
    for (const oibvh_batch::Mesh& m : b->meshes) depth = std::max(depth, m.L - std::min(m.L, entry));
    const oibvh_batch::Slot& sl = cur_slot(b);
    d.tab = oi_tab_args{sl.d_mdesc, sl.h_mdesc.data(), (uint32_t)b->meshes.size(), b->chunk_mesh, OI_CHUNK_SHIFT};
    d.depth = depth;
    d.expand = c->trav_expand ? c->trav_expand : std::min(3u, std::max(1u, expand));
    d.schedule = c->trav_schedule;
    d.flags = flags;
    d.implicit_end = (uint32_t)b->It;
    return oi_det_prepare(&d);
}

extern "C" oibvh_status oibvh_batch_detect(oibvh_batch* b, uint32_t entry, uint32_t expand, uint32_t flags)
{
    if (!b) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    if (!b->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "batch detect before build");
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = settle(b, true);
    if (r != OIBVH_OK) return r;
    oi_detector& d = b->det;
    const bool async = (flags & OIBVH_DETECT_ASYNC) != 0;
    if (d.unchecked && !async) {
        r = oi_det_check(&d);
        if (r != OIBVH_OK) return r;
    }
    oi_det_use_set(&d, b->cur);
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    if ((r = prepare_detect(b, entry, expand, flags)) != OIBVH_OK) return r;
    hipEvent_t e0, e1;
    if ((r = detect_timing(b, &e0, &e1)) != OIBVH_OK) return r;
    if (e0) HIPCHK(c, hipEventRecord(e0, c->stream));
    if ((r = oi_det_enqueue(&d, !async)) != OIBVH_OK) return r;
    if (e1) HIPCHK(c, hipEventRecord(e1, c->stream));
    if (async) return OIBVH_OK;
    return oi_det_check(&d);
}

// capture the work `enqueue` queues on the context's stream into g, unless
// g already holds a graph made under the same key
template <class F>
static oibvh_status captured(oibvh_batch* b, oibvh_batch::Graph& g, const uint64_t (&key)[8], F enqueue)
{
    oibvh_ctx* c = b->ctx;
    if (g.exec && std::memcmp(key, g.key, sizeof key) == 0) return OIBVH_OK;
    if (g.exec) {
        HIPCHK(c, oi_ctx_quiesce(c));
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    oibvh_status rr = enqueue();
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rr != OIBVH_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rr;
    }
    if (e != hipSuccess) return oi_set_err(c, OIBVH_ERR_HIP, "frame capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g.exec = nullptr;
        return oi_set_err(c, OIBVH_ERR_HIP, "frame instantiate: %s", hipGetErrorString(e));
    }
    std::memcpy(g.key, key, sizeof key);
    return OIBVH_OK;
}

// what a captured detect depends on beyond the buffers (slot, query, tuning)
static void detect_key(const oibvh_batch* b, uint32_t flags, uint64_t* k)
{
    const oibvh_ctx* c = b->ctx;
    const oi_detector& d = b->det;
    k[0] = b->front_key[0] | ((uint64_t)d.local_chunk << 32) | ((uint64_t)d.local_levels << 48);
    k[1] = d.expand | ((uint64_t)d.schedule << 32);
    k[2] = (flags & ~(OIBVH_DETECT_ASYNC | OIBVH_FRAME_OVERLAP)) | ((uint64_t)c->narrow_separate << 32);
    k[3] = (uint64_t)b->reduce_M | ((uint64_t)d.tail_chunk << 16) | ((uint64_t)d.tail_from << 40);
    k[4] = d.generation;
    k[5] = (uint64_t)b->desc_gen ^ ((uint64_t)oi_grid_for(c) << 32);
    k[6] = (uint64_t)b->cur | ((uint64_t)d.n_levels << 8) | ((uint64_t)d.n_front << 32);
    k[7] = (uint64_t)c->refit_blocks;
}

// An overlapped frame: refit into the slot the staging calls since the last
// frame wrote, on the context's stream, then this frame's detect on the side
// stream, while the previous frame's detect (reading the other slot) may
// still run there. Each half is a captured graph per slot.
static oibvh_status overlapped_frame(oibvh_batch* b, uint32_t entry, uint32_t expand, uint32_t flags)
{
    oibvh_ctx* c = b->ctx;
    oi_detector& d = b->det;
    const bool async = (flags & OIBVH_DETECT_ASYNC) != 0;
    oibvh_status r = oi_ctx_side_init(c);
    if (r != OIBVH_OK) return r;
    if (d.unchecked && !async && (r = oi_det_check(&d)) != OIBVH_OK) return r;
    if (b->ov_active) {
        // the staged slot becomes current (opened now if nothing was staged)
        uint32_t k = 0;
        if ((r = stage_slot(b, &k)) != OIBVH_OK) return r;
        b->cur = k;
        b->staged_other = false;
    } else {
        // first overlapped frame: staging went to the current slot, whose
        // readers the last settle already waited for
        if ((r = wait_slot_readers(b, b->cur)) != OIBVH_OK) return r;
    }
    // the detector's counter set follows the slot, so the refit's apex
    // launch resets the set the previous frame's detect is not using
    oi_det_use_set(&d, b->cur);
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    if ((r = prepare_detect(b, entry, expand, flags & ~(OIBVH_DETECT_ASYNC | OIBVH_FRAME_OVERLAP))) != OIBVH_OK)
        return r;
    const oi_seed_args seed = oi_det_seed_args(&d, false);
    bool seeded = b->n_apex != 0;
    uint64_t key[8];
    detect_key(b, flags, key);
    // refit half (eager when timing samples it)
    hipEvent_t t0, t1;
    if ((r = refit_timing(b, &t0, &t1)) != OIBVH_OK) return r;
    if (t0) {
        r = enqueue_refit(b, &seed, &seeded, t0, t1, c->refit_blocks);
    } else {
        oibvh_batch::Graph& gr = b->g_refit[b->cur];
        r = captured(b, gr, key, [&]() { return enqueue_refit(b, &seed, &seeded, nullptr, nullptr, c->refit_blocks); });
        if (r == OIBVH_OK) HIPCHK(c, hipGraphLaunch(gr.exec, c->stream));
    }
    if (r != OIBVH_OK) return r;
    HIPCHK(c, hipEventRecord(c->ev_main, c->stream));
    // detect half, captured on the context's stream, run on this detect's side stream
    oibvh_batch::Graph& gd = b->g_detect[b->cur];
    key[7] |= seeded ? (1ull << 40) : 0ull;
    if ((r = captured(b, gd, key, [&]() { return oi_det_enqueue(&d, false, seeded); })) != OIBVH_OK) return r;
    const uint64_t q = c->side_seq;
    hipStream_t side = oi_side_stream(c, q);
    HIPCHK(c, hipStreamWaitEvent(side, c->ev_main, 0));
    hipEvent_t e0, e1;
    if ((r = detect_timing(b, &e0, &e1)) != OIBVH_OK) return r;
    if (e0) HIPCHK(c, hipEventRecord(e0, side));
    HIPCHK(c, hipGraphLaunch(gd.exec, side));
    if (e1) HIPCHK(c, hipEventRecord(e1, side));
    c->side_seq++;
    HIPCHK(c, hipEventRecord(c->ev_ring[q % OI_SIDE_RING], side));
    b->slot[b->cur].read_seq = q + 1;
    c->side_pending = true;
    b->ov_active = true;
    d.unchecked = true;
    d.have_result = true;
    d.sorted_ok = false;
    if (async) return OIBVH_OK;
    return oi_det_check(&d);
}

// refit + detect replayed from a captured graph; the graph is captured again
// whenever the query, the reduce shape or a buffer changes
extern "C" oibvh_status oibvh_batch_frame(oibvh_batch* b, uint32_t entry, uint32_t expand, uint32_t flags)
{
    if (!b) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    if (!b->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "batch frame before build");
    HIPCHK(c, hipSetDevice(c->device));
    oibvh_status r = ensure_reduce_tables(b);
    if (r != OIBVH_OK) return r;
    if (flags & OIBVH_FRAME_OVERLAP) {
        if ((r = ensure_slot(b, b->cur ^ 1u)) != OIBVH_OK) return r;
        return overlapped_frame(b, entry, expand, flags);
    }
    if ((r = settle(b, true)) != OIBVH_OK) return r;
    oi_detector& d = b->det;
    if (d.unchecked && !(flags & OIBVH_DETECT_ASYNC) && (r = oi_det_check(&d)) != OIBVH_OK) return r;
    oi_det_use_set(&d, b->cur);
    d.n_pairs = 0;
    d.stats = oibvh_detect_stats{};
    if ((r = prepare_detect(b, entry, expand, flags & ~OIBVH_DETECT_ASYNC)) != OIBVH_OK) return r;
    // the detect's counter reset rides on the refit's apex launch
    const oi_seed_args seed = oi_det_seed_args(&d, false);
    bool seeded = false;
    hipEvent_t t0, t1, e0, e1;
    if ((r = refit_timing(b, &t0, &t1)) != OIBVH_OK) return r;
    if ((r = detect_timing(b, &e0, &e1)) != OIBVH_OK) return r;
    if (t0 || e0) {
        // a sampled frame runs eagerly, with its event pairs
        if ((r = enqueue_refit(b, &seed, &seeded, t0, t1)) != OIBVH_OK) return r;
        if (e0) HIPCHK(c, hipEventRecord(e0, c->stream));
        if ((r = oi_det_enqueue(&d, false, seeded)) != OIBVH_OK) return r;
        if (e1) HIPCHK(c, hipEventRecord(e1, c->stream));
    } else {
        uint64_t key[8];
        detect_key(b, flags, key);
        key[7] = 0;
        oibvh_batch::Graph& g = b->g_frame[b->cur];
        r = captured(b, g, key, [&]() {
            oibvh_status rr = enqueue_refit(b, &seed, &seeded);
            if (rr == OIBVH_OK) rr = oi_det_enqueue(&d, false, seeded);
            return rr;
        });
        if (r != OIBVH_OK) return r;
        HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
    }
    d.unchecked = true;
    d.have_result = true;
    d.sorted_ok = false;
    if (flags & OIBVH_DETECT_ASYNC) return OIBVH_OK;
    return oi_det_check(&d);
}

// HIP-event timing of the refit's stage-1 launch and of detects (every n-th
// of each; 0 stops), as oibvh_scene_set_timing
extern "C" oibvh_status oibvh_batch_set_timing(oibvh_batch* b, uint32_t every)
{
    if (!b) return OIBVH_ERR_INVALID;
    b->timing = every;
    b->seq_refit = b->seq_detect = 0;
    b->n_refit = b->n_detect = 0;
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

extern "C" oibvh_status oibvh_batch_get_timing(oibvh_batch* b, oibvh_scene_timing* out)
{
    if (!b || !out) return OIBVH_ERR_INVALID;
    oibvh_ctx* c = b->ctx;
    HIPCHK(c, hipSetDevice(c->device));
    std::memset(out, 0, sizeof *out);
    oibvh_status r = sum_pairs(c, b->ev_refit, b->n_refit, &out->refit_ms);
    if (r == OIBVH_OK) r = sum_pairs(c, b->ev_detect, b->n_detect, &out->detect_ms);
    if (r != OIBVH_OK) return r;
    out->refits = b->n_refit;
    out->detects = b->n_detect;
    b->n_refit = b->n_detect = 0;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_get_pair_count(oibvh_batch* b, uint32_t* count)
{
    if (!b || !count) return OIBVH_ERR_INVALID;
    oibvh_status r = oi_det_check(&b->det);
    if (r != OIBVH_OK) return r;
    *count = b->det.n_pairs;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_get_pairs(oibvh_batch* b, uint32_t* out, uint64_t cap, uint32_t* n)
{
    if (!b) return OIBVH_ERR_INVALID;
    return oi_det_get_pairs(&b->det, out, cap, n);
}

extern "C" oibvh_status oibvh_batch_pairs_device(oibvh_batch* b, const uint32_t** dptr, uint32_t* n)
{
    if (!b || !dptr || !n) return OIBVH_ERR_INVALID;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    const uint4* p = nullptr;
    oibvh_status r = oi_det_sorted_device(&b->det, &p, n);
    *dptr = reinterpret_cast<const uint32_t*>(p);
    return r;
}

// Scene::convertToVertexArray (scene.cu:69-93) over the batch's meshes
extern "C" oibvh_status oibvh_batch_get_pair_vertices(oibvh_batch* b, float* out, uint64_t cap, uint32_t ptr_kind,
                                                      uint32_t* n)
{
    if (!b) return OIBVH_ERR_INVALID;
    HIPCHK(b->ctx, hipSetDevice(b->ctx->device));
    return oi_det_pair_vertices(&b->det, cur_slot(b).d_bdesc, out, cap, ptr_kind, n);
}

extern "C" oibvh_status oibvh_batch_get_stats(oibvh_batch* b, oibvh_detect_stats* st)
{
    if (!b || !st) return OIBVH_ERR_INVALID;
    oibvh_status r = oi_det_check(&b->det);
    if (r != OIBVH_OK) return r;
    *st = b->det.stats;
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_get_level_counts(oibvh_batch* b, uint32_t* out, uint64_t cap, uint32_t* n)
{
    if (!b || !n) return OIBVH_ERR_INVALID;
    return oi_det_level_counts(&b->det, out, cap, n);
}

static oibvh_status mesh_readback(oibvh_batch* b, uint32_t mesh, void* dst, const void* src, size_t bytes)
{
    oibvh_ctx* c = b->ctx;
    if (mesh >= b->meshes.size()) return batch_err(b, OIBVH_ERR_INVALID, "bad mesh index");
    if (!b->built) return oi_set_err(c, OIBVH_ERR_NOT_BUILT, "batch not built");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OIBVH_OK;
}

extern "C" oibvh_status oibvh_batch_get_nodes(oibvh_batch* b, uint32_t mesh, float* out, uint64_t cap, uint32_t* n)
{
    if (!b || mesh >= b->meshes.size()) return OIBVH_ERR_INVALID;
    const oibvh_batch::Mesh& m = b->meshes[mesh];
    if (n) *n = oi_tree_size(m.F);
    if (!out || cap < oi_tree_size(m.F)) return oi_set_err(b->ctx, OIBVH_ERR_CAPACITY, "need %u nodes", oi_tree_size(m.F));
    return mesh_readback(b, mesh, out, cur_slot(b).nodes + 6ull * m.nbase, 24ull * m.N);
}

extern "C" oibvh_status oibvh_batch_get_perm(oibvh_batch* b, uint32_t mesh, uint32_t* out, uint64_t cap)
{
    if (!b || mesh >= b->meshes.size()) return OIBVH_ERR_INVALID;
    const oibvh_batch::Mesh& m = b->meshes[mesh];
    if (!out || cap < m.F) return oi_set_err(b->ctx, OIBVH_ERR_CAPACITY, "need %u faces", m.F);
    return mesh_readback(b, mesh, out, b->perm + m.fbase, 4ull * m.F);
}
