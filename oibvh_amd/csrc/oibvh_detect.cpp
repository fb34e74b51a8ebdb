// oibvh_detect.cpp -- the BVTT detector shared by scenes and batches: buffer
// management, the launch sequence of one query and the overflow re-run loop
// (Scene::detectCollisionOnGPU, scene.cu:225-446, which instead copies the
// front size to the host after every level and uses fixed 10M-entry buffers).
#include <cstddef>

#include "oibvh_internal.h"

// before a buffer an overlapped frame's detect may still use is freed or
// rewritten
static oibvh_status idle(oibvh_ctx* c)
{
    if (c->side[0]) HIPCHK(c, oi_ctx_quiesce(c));
    return OIBVH_OK;
}

template <class T>
static oibvh_status grow(oibvh_ctx* c, T*& p, uint32_t& cap, uint64_t need, uint64_t& gen)
{
    if (cap >= need) return OIBVH_OK;
    if (need > 0x7fffffffull) return oi_set_err(c, OIBVH_ERR_OVERFLOW, "detect buffer would exceed 2^31 entries");
    oibvh_status r = idle(c);
    if (r != OIBVH_OK) return r;
    dfree(p);
    cap = 0;
    gen++;
    HIPCHK(c, dalloc(&p, need));
    cap = (uint32_t)need;
    return OIBVH_OK;
}

oibvh_status oi_det_init(oi_detector* d, oibvh_ctx* c)
{
    d->ctx = c;
    hipError_t e = dalloc(&d->ctl_base, OI_FRAME_DEPTH);
    if (e == hipSuccess) e = hipMemsetAsync(d->ctl_base, 0, OI_FRAME_DEPTH * sizeof(oi_detect_ctl), c->stream);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&d->h_ctl, OI_FRAME_DEPTH * sizeof(oi_detect_ctl), hipHostMallocDefault);
    if (e != hipSuccess) {
        oi_det_free(d);
        return oi_set_err(c, OIBVH_ERR_HIP, "detector alloc: %s", hipGetErrorString(e));
    }
    std::memset(d->h_ctl, 0, OI_FRAME_DEPTH * sizeof(oi_detect_ctl));
    d->ctl = d->ctl_base;
    return OIBVH_OK;
}

// the current set's working buffers to / from their per-set slots
static void save_view(oi_detector* d)
{
    const uint32_t k = d->set;
    d->set_front[k][0] = d->front[0];
    d->set_front[k][1] = d->front[1];
    d->set_front_cap[k] = d->front_cap;
    d->set_cands[k] = d->cands;
    d->set_cand_cap[k] = d->cand_cap;
    d->set_hits[k] = d->hits;
    d->set_hits_cap[k] = d->hits_cap;
}

static void load_view(oi_detector* d)
{
    const uint32_t k = d->set;
    d->front[0] = d->set_front[k][0];
    d->front[1] = d->set_front[k][1];
    d->front_cap = d->set_front_cap[k];
    d->cands = d->set_cands[k];
    d->cand_cap = d->set_cand_cap[k];
    d->hits = d->set_hits[k];
    d->hits_cap = d->set_hits_cap[k];
}

void oi_det_use_set(oi_detector* d, uint32_t k)
{
    if (k == d->set) return;
    save_view(d);
    // the set left behind keeps an unchecked result until a check reads it;
    // the set taken holds none any more: an asynchronous seed folds its
    // errors into its sticky bits, and a synchronous frame checks first
    if (d->unchecked) d->set_unchecked[d->set] = true;
    d->set = k % OI_FRAME_DEPTH;
    load_view(d);
    d->set_unchecked[d->set] = false;
    d->ctl = d->ctl_base + d->set;
    d->counts = d->counts_base ? d->counts_base + 256u * d->set : nullptr;
}

void oi_det_flip_set(oi_detector* d) { oi_det_use_set(d, (d->set + 1u) % OI_FRAME_DEPTH); }

void oi_det_free(oi_detector* d)
{
    dfree(d->d_init);
    save_view(d);
    for (uint32_t k = 0; k < OI_FRAME_DEPTH; k++) {
        dfree(d->set_front[k][0]);
        dfree(d->set_front[k][1]);
        dfree(d->set_cands[k]);
        dfree(d->set_hits[k]);
        d->set_front_cap[k] = d->set_cand_cap[k] = d->set_hits_cap[k] = 0;
    }
    load_view(d);
    dfree(d->counts_base);
    d->counts = nullptr;
    dfree(d->d_sorted);
    dfree(d->sort_words);
    dfree(d->sort_vals);
    dfree(d->sort_scratch);
    dfree(d->d_stage);
    dfree(d->ctl_base);
    d->ctl = nullptr;
    if (d->h_ctl) (void)hipHostFree(d->h_ctl);
    d->h_ctl = nullptr;
}

oibvh_status oi_det_set_front(oi_detector* d, const std::vector<uint2>& front)
{
    oibvh_ctx* c = d->ctx;
    oibvh_status r = idle(c);
    if (r != OIBVH_OK) return r;
    d->n_front = (uint32_t)front.size();
    r = grow(c, d->d_init, d->init_cap, std::max(1u, d->n_front), d->generation);
    if (r != OIBVH_OK) return r;
    if (d->n_front)
        HIPCHK(c, hipMemcpyAsync(d->d_init, front.data(), d->n_front * sizeof(uint2), hipMemcpyHostToDevice,
                                 c->stream));
    return OIBVH_OK;
}

oibvh_status oi_det_prepare(oi_detector* d)
{
    oibvh_ctx* c = d->ctx;
    // every BVTT step takes each non-leaf side up to `expand` levels down and
    // tests the pairs it reaches: the walk ends at the leaves after
    // ceil(depth / expand) steps (one step tests entry-level leaf pairs)
    // the schedule's leading steps, then `expand` per step, until the deepest
    // tree's leaves are reached (at least one step)
    uint32_t k = 0, done = 0, sch = d->schedule;
    bool lead = true;
    while (k == 0 || done < d->depth) {
        uint32_t e = lead ? (sch & 3u) : 0u;
        if (e == 0) {
            lead = false;
            e = d->expand;
        }
        sch >>= 2;
        if (k + 1 >= 256) return oi_set_err(c, OIBVH_ERR_INVALID, "detect: more than 255 BVTT steps");
        d->step_e[k++] = (uint8_t)e;
        done += e;
    }
    d->n_levels = k;
    // the first steps as one local walk when asked for, unless it overflowed
    // recently; its steps all descend `expand` levels
    const uint32_t mode = c->trav_local;
    d->local_chunk = 0;
    d->local_levels = 0;
    // a walk that overflowed stays off until the tuning is set again: an
    // asynchronous detect cannot re-run, so one that overflowed again
    // would leave its frame incomplete
    if (mode && d->local_off != c->tuning_gen + 1 && d->n_front) {
        d->local_chunk = mode;
        k = 0;
        done = 0;
        while (k == 0 || done < d->depth) {
            d->step_e[k++] = (uint8_t)d->expand;
            done += d->expand;
        }
        d->n_levels = k;
        d->local_levels = c->trav_local_levels ? std::min(c->trav_local_levels, k) : k;
    }
    // the last steps as a tail walk when asked for (unless it overflowed
    // recently): it starts after the head walk and after every step whose
    // depth the schedule set, since it descends `expand` levels per step
    d->tail_chunk = 0;
    d->tail_from = 0;
    if (c->trav_tail && d->tail_off != c->tuning_gen + 1 && d->n_front) {
        const uint32_t S = c->trav_tail_steps;
        uint32_t T = (S && S < d->n_levels) ? d->n_levels - S : 0u;
        T = std::max(T, d->local_chunk ? d->local_levels : 0u);
        for (uint32_t j = T; j < d->n_levels; j++)
            if (d->step_e[j] != d->expand) T = j + 1;
        if (T < d->n_levels) {
            d->tail_chunk = c->trav_tail;
            d->tail_from = T;
        }
    }
    const uint32_t fcap = std::max(std::max(c->queue_cap, d->want_front), oi_np2(std::max(1u, d->n_front)));
    if (d->front_cap < fcap) {
        oibvh_status ri = idle(c);
        if (ri != OIBVH_OK) return ri;
        dfree(d->front[0]);
        dfree(d->front[1]);
        d->front_cap = 0;
        d->generation++;
        HIPCHK(c, dalloc(&d->front[0], fcap));
        HIPCHK(c, dalloc(&d->front[1], fcap));
        d->front_cap = fcap;
    }
    oibvh_status r;
    if ((r = grow(c, d->counts_base, d->counts_cap, 256u * OI_FRAME_DEPTH, d->generation)) != OIBVH_OK) return r;
    d->counts = d->counts_base + 256u * d->set;
    if ((r = grow(c, d->cands, d->cand_cap, std::max(c->cand_cap, d->want_cands), d->generation)) != OIBVH_OK)
        return r;
    if ((r = grow(c, d->hits, d->hits_cap, std::max(c->hits_cap, d->want_hits), d->generation)) != OIBVH_OK)
        return r;
    return OIBVH_OK;
}

oi_seed_args oi_det_seed_args(const oi_detector* d, bool reset_sticky)
{
    oi_seed_args s{};
    s.counts = d->counts;
    s.ctl = d->ctl;
    s.n0 = d->n_front;
    s.n_counts = d->n_levels + 1;
    s.reset_sticky = reset_sticky ? 1u : 0u;
    return s;
}

oibvh_status oi_det_enqueue(oi_detector* d, bool reset_sticky, bool seeded, hipEvent_t done)
{
    oibvh_ctx* c = d->ctx;
    hipStream_t st = d->stream ? d->stream : c->stream;
    if (!seeded)
        HIPCHK(c, oi_launch_seed(st, d->n_front, d->counts, d->n_levels + 1, d->ctl, reset_sticky ? 1u : 0u));
    const uint32_t grid = oi_grid_for(c);
    const bool fp64 = (d->flags & OIBVH_DETECT_FP64_FALLBACK) != 0;
    const uint32_t J = d->local_chunk ? d->local_levels : 0u;
    const uint32_t T = d->tail_chunk ? d->tail_from : d->n_levels;
    // Leaf pairs are tested inside the BVTT launches that find them (unless
    // narrow_separate): FP32 queries with the SAT; FP64-fallback queries with
    // the FP32 filter in the steps (not in local walks), the pairs it leaves
    // undecided collected in the candidate buffer for one FP64 pass
    d->fused = !c->narrow_separate && (!fp64 || (J == 0 && T >= d->n_levels));
    const int fuse = d->fused ? (fp64 ? 2 : 1) : 0;
    // `done` rides on the last launch: k_narrow, else the tail walk, else
    // the last step, else the head walk
    const bool narrow = !d->fused || fp64, tail = T < d->n_levels, steps = T > J;
    hipEvent_t done_head = (!narrow && !tail && !steps) ? done : nullptr;
    hipEvent_t done_tail = (!narrow && tail) ? done : nullptr;
    if (J) {
        // steps 0 .. J-1 in one launch; its front goes where step J - 1's would
        oi_local_args w{};
        w.front_in = d->d_init;
        w.n_in = d->n_front;
        w.chunk = d->local_chunk;
        w.levels = J;
        w.test_entries = 1;
        w.overflow_bit = OI_ERR_LOCAL_OVERFLOW;
        w.front_out = d->front[(J - 1) & 1];
        w.count_out = d->counts + J;
        w.cap_out = d->front_cap;
        w.done = done_head;
        HIPCHK(c, oi_launch_bvtt_local(st, &d->tab, &w, d->cands, &d->ctl->cands, d->cand_cap, d->hits, d->hits_cap,
                                       d->ctl, d->expand, d->implicit_end, d->fused ? 1 : 0));
        done_head = nullptr;
    }
    for (uint32_t k = J; k < T; k++) {
        // step 0 reads the initial front in place; then the fronts alternate
        const uint2* in = k ? d->front[(k - 1) & 1] : d->d_init;
        HIPCHK(c, oi_launch_bvtt_step(st, grid, &d->tab, k == 0, in, k ? d->counts + k : nullptr, d->n_front,
                                      k ? d->front_cap : d->n_front, d->front[k & 1], d->counts + k + 1,
                                      d->front_cap, d->cands, &d->ctl->cands, d->cand_cap, d->hits, d->hits_cap,
                                      d->ctl, d->step_e[k], d->implicit_end, fuse,
                                      (!narrow && !tail && k + 1 == T) ? done : nullptr));
    }
    if (tail) {
        // steps T .. n_levels - 1 in one launch over the front step T - 1
        // left: every pair ends at the leaves, so the last level writes no
        // front (its count, checked below, stays 0); front_out is the buffer
        // not being read
        oi_local_args w{};
        w.front_in = T ? d->front[(T - 1) & 1] : d->d_init;
        w.count_in = T ? d->counts + T : nullptr;
        w.n_in = d->n_front;
        w.cap_in = T ? d->front_cap : d->n_front;
        w.chunk = d->tail_chunk;
        w.levels = d->n_levels - T;
        w.grid = 4u * (uint32_t)std::max(1, c->cus);
        w.test_entries = T == 0 ? 1u : 0u;
        w.overflow_bit = OI_ERR_TAIL_OVERFLOW;
        w.front_out = d->front[T & 1];
        w.count_out = d->counts + d->n_levels;
        w.cap_out = d->front_cap;
        w.done = done_tail;
        HIPCHK(c, oi_launch_bvtt_local(st, &d->tab, &w, d->cands, &d->ctl->cands, d->cand_cap, d->hits, d->hits_cap,
                                       d->ctl, d->expand, d->implicit_end, d->fused ? 1 : 0));
    }
    if (narrow && !fp64)
        HIPCHK(c, oi_launch_narrow(st, grid, &d->tab, d->cands, &d->ctl->cands, d->cand_cap, d->hits,
                                   &d->ctl->n_hits, d->hits_cap, d->ctl, 0, nullptr, 0, done));
    if (fp64 && d->fused) {
        // the FP64 SAT over the pairs the steps' filter left undecided
        HIPCHK(c, oi_launch_narrow(st, grid, &d->tab, d->cands, &d->ctl->deferred, d->cand_cap, d->hits,
                                   &d->ctl->n_hits, d->hits_cap, d->ctl, 2, nullptr, 0, done));
    } else if (narrow && fp64) {
        // the FP32 filter over the candidates, then the FP64 SAT over the
        // pairs it left undecided; the BVTT fronts are idle by now, and the
        // first holds that list (an overflow grows them like a front's)
        HIPCHK(c, oi_launch_narrow(st, grid, &d->tab, d->cands, &d->ctl->cands, d->cand_cap, d->hits,
                                   &d->ctl->n_hits, d->hits_cap, d->ctl, 1, d->front[0], d->front_cap, nullptr));
        HIPCHK(c, oi_launch_narrow(st, grid, &d->tab, d->front[0], &d->ctl->deferred, d->front_cap, d->hits,
                                   &d->ctl->n_hits, d->hits_cap, d->ctl, 2, nullptr, 0, done));
    }
    else if (done_head)  // a query with no BVTT launch
        HIPCHK(c, hipEventRecord(done, st));
    d->unchecked = true;
    d->have_result = true;
    d->sorted_ok = false;
    return OIBVH_OK;
}

oibvh_status oi_det_check(oi_detector* d)
{
    oibvh_ctx* c = d->ctx;
    if (!d->unchecked) return OIBVH_OK;
    HIPCHK(c, oi_ctx_join(c));  // an overlapped frame's detect ran on a side stream
    uint32_t retries = 0, earlier = 0;
    for (;;) {
        d->h_counts.resize(d->n_levels + 1);
        HIPCHK(c, hipMemcpyAsync(d->h_ctl, d->ctl, sizeof(oi_detect_ctl), hipMemcpyDeviceToHost, c->stream));
        // the other sets still holding unchecked overlapped frames' results
        uint32_t others[OI_FRAME_DEPTH], n_others = 0;
        for (uint32_t s = 0; s < OI_FRAME_DEPTH; s++)
            if (s != d->set && d->set_unchecked[s]) others[n_others++] = s;
        for (uint32_t j = 0; j < n_others; j++)
            HIPCHK(c, hipMemcpyAsync(d->h_ctl + 1 + j, d->ctl_base + others[j], sizeof(oi_detect_ctl),
                                     hipMemcpyDeviceToHost, c->stream));
        // what is read here is reported once: clear the error and sticky
        // bits of every set read, so later asynchronous seeds do not fold
        // them in again (error and sticky are adjacent words)
        static_assert(offsetof(oi_detect_ctl, sticky) == offsetof(oi_detect_ctl, error) + 4, "ctl layout");
        HIPCHK(c, hipMemsetAsync(&d->ctl->error, 0, 8, c->stream));
        for (uint32_t j = 0; j < n_others; j++)
            HIPCHK(c, hipMemsetAsync(&d->ctl_base[others[j]].error, 0, 8, c->stream));
        HIPCHK(c, hipMemcpyAsync(d->h_counts.data(), d->counts, (d->n_levels + 1) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const oi_detect_ctl& k = d->h_ctl[0];
        earlier |= k.sticky;  // errors of unchecked asynchronous detects before this one
        for (uint32_t j = 0; j < n_others; j++) {
            // earlier overlapped frames' detects and the unchecked ones before them
            earlier |= d->h_ctl[1 + j].sticky | d->h_ctl[1 + j].error;
            d->set_unchecked[others[j]] = false;
        }
        if (d->h_counts[d->n_levels] != 0) {
            // the last step reaches the leaves on every side: it writes no front
            d->unchecked = false;
            d->have_result = false;
            return oi_set_err(c, OIBVH_ERR_HIP, "detect: %u BVTT pairs left after the last step",
                              d->h_counts[d->n_levels]);
        }
        if (k.error & OI_ERR_CORRUPT) {
            d->unchecked = false;
            d->have_result = false;
            return oi_set_err(c, OIBVH_ERR_HIP, "detect: corrupt BVTT entry (error bits 0x%x)", k.error);
        }
        bool rerun = false;
        oibvh_status r = OIBVH_OK;
        if (k.error & OI_ERR_LOCAL_OVERFLOW) {
            // a workgroup's front outgrew LDS: steps for the next queries
            d->local_off = c->tuning_gen + 1;
            rerun = true;
        }
        if (k.error & OI_ERR_TAIL_OVERFLOW) {
            d->tail_off = c->tuning_gen + 1;
            rerun = true;
        }
        uint32_t maxfront = 0;
        for (uint32_t v : d->h_counts) maxfront = std::max(maxfront, v);
        if (k.error & OI_ERR_QUEUE_OVERFLOW) {
            uint64_t need = 1;
            while (need < std::max<uint64_t>(maxfront, 4ull * d->front_cap)) need <<= 1;
            if (need > (1u << 30)) {
                d->unchecked = false;
                d->have_result = false;
                return oi_set_err(c, OIBVH_ERR_OVERFLOW, "BVTT front exceeds %u entries", 1u << 30);
            }
            dfree(d->front[0]);
            dfree(d->front[1]);
            d->front_cap = 0;
            d->generation++;
            HIPCHK(c, dalloc(&d->front[0], need));
            HIPCHK(c, dalloc(&d->front[1], need));
            d->front_cap = (uint32_t)need;
            d->want_front = d->front_cap;
            rerun = true;
        }
        if ((k.error & OI_ERR_CAND_OVERFLOW) || (!d->fused && k.cands > d->cand_cap)) {
            r = grow(c, d->cands, d->cand_cap, std::max<uint64_t>(k.cands, 2ull * d->cand_cap), d->generation);
            if (r != OIBVH_OK) return r;
            d->want_cands = d->cand_cap;
            rerun = true;
        }
        if ((k.error & OI_ERR_HITS_OVERFLOW) || k.n_hits > d->hits_cap) {
            const uint64_t need = std::max<uint64_t>((uint64_t)k.n_hits + (k.n_hits >> 3), 2ull * d->hits_cap);
            if ((r = grow(c, d->hits, d->hits_cap, need, d->generation)) != OIBVH_OK) return r;
            d->want_hits = d->hits_cap;
            rerun = true;
        }
        if (!rerun) {
            d->n_pairs = k.n_hits;
            d->stats.bvtt_tests = k.tests;
            d->stats.candidates = k.cands;
            d->stats.fp64_rechecks = k.fp64;
            d->stats.fp64_candidates = k.deferred;
            d->stats.adjacent = k.adj;
            d->stats.pairs = k.n_hits;
            d->stats.retries = retries;
            d->stats.async_errors = earlier;
            d->unchecked = false;
            d->sorted_ok = false;  // a new result: sorted again on demand
            return OIBVH_OK;
        }
        retries++;
        if (retries > 16) return oi_set_err(c, OIBVH_ERR_OVERFLOW, "detect: too many retries");
        if ((r = oi_det_prepare(d)) != OIBVH_OK) return r;
        if ((r = oi_det_enqueue(d, true)) != OIBVH_OK) return r;
    }
}

static uint32_t bits_for(uint32_t max_value)
{
    return max_value ? 32u - (uint32_t)__builtin_clz(max_value) : 0u;
}

// The checked hits sorted on the device into d_sorted, in get_pairs order
// (meshA, meshB, triA, triB): LSD radix passes over the bytes that can be
// non-zero (mesh ids below the table's mesh count, face ids below its
// largest face count), sorting hit indices with keys looked up by index,
// then one gather. Once per query.
static oibvh_status sort_hits(oi_detector* d)
{
    oibvh_ctx* c = d->ctx;
    const uint32_t n = d->n_pairs;
    if (d->sorted_ok || n == 0) return OIBVH_OK;
    if (n >= (1u << 30)) return oi_set_err(c, OIBVH_ERR_OVERFLOW, "pair sort: more than 2^30 pairs");
    uint32_t max_f = 1;
    for (uint32_t m = 0; m < d->tab.n_meshes; m++) max_f = std::max(max_f, d->tab.h_meshes[m].F);
    const uint32_t mbits = bits_for(d->tab.n_meshes ? d->tab.n_meshes - 1 : 0), tbits = bits_for(max_f - 1);
    oi_hit_digits dg{};
    for (uint32_t w : {3u, 2u, 1u, 0u}) {
        const uint32_t nb = ((w >= 2 ? tbits : mbits) + 7u) / 8u;
        for (uint32_t k = 0; k < nb; k++) {
            dg.word[dg.nd] = (uint8_t)w;
            dg.shift[dg.nd] = (uint8_t)(8u * k);
            dg.nd++;
        }
    }
    const uint32_t tiles = (n + oi_sort_tile() - 1) / oi_sort_tile();
    uint64_t gen = 0;
    oibvh_status r;
    if (n <= oi_hits_sort_small_max()) {
        // one workgroup, every pass in LDS, one launch
        if ((r = grow(c, d->d_sorted, d->d_sorted_cap, n, gen)) != OIBVH_OK) return r;
        HIPCHK(c, oi_launch_hits_sort_small(c->stream, d->hits, n, &dg, d->d_sorted));
        d->sorted_ok = true;
        return OIBVH_OK;
    }
    if (d->sort_cap < n) {
        // sized with headroom, so that the next queries rarely grow them
        const uint64_t cap = std::min<uint64_t>((1ull << 30) - 1, (uint64_t)n + (n >> 2));
        const uint64_t ctiles = (cap + oi_sort_tile() - 1) / oi_sort_tile();
        dfree(d->sort_words);
        dfree(d->sort_vals);
        dfree(d->sort_scratch);
        d->sort_cap = 0;
        HIPCHK(c, dalloc(&d->sort_words, 4 * cap));
        HIPCHK(c, dalloc(&d->sort_vals, 2 * cap));
        HIPCHK(c, dalloc(&d->sort_scratch, 256ull * OI_HIT_DIGITS + OI_HIT_DIGITS + 512ull * ctiles));
        d->sort_cap = (uint32_t)cap;
    }
    if ((r = grow(c, d->d_sorted, d->d_sorted_cap, n, gen)) != OIBVH_OK) return r;
    uint32_t* hist = d->sort_scratch;
    uint32_t* tctr = hist + 256u * OI_HIT_DIGITS;
    uint32_t* lb[2] = {tctr + OI_HIT_DIGITS, tctr + OI_HIT_DIGITS + 256ull * tiles};
    uint32_t* vals[2] = {d->sort_vals, d->sort_vals + d->sort_cap};
    hipStream_t st = c->stream;
    // histograms, tile counters and the first look-back buffer start at zero
    HIPCHK(c, oi_launch_zero_u32(st, hist, 256u * OI_HIT_DIGITS + OI_HIT_DIGITS + 256u * tiles));
    HIPCHK(c, oi_launch_hits_digits(st, d->hits, n, d->sort_words, hist, &dg));
    const uint32_t* order = nullptr;  // identity
    for (uint32_t p = 0; p < dg.nd; p++) {
        HIPCHK(c, oi_launch_sort_pass(st, nullptr, order, nullptr, vals[p & 1], n, dg.shift[p], hist + 256u * p,
                                      lb[p & 1], p + 1 < dg.nd ? lb[(p + 1) & 1] : nullptr, tctr + p,
                                      d->sort_words + (uint64_t)n * dg.word[p]));
        order = vals[p & 1];
    }
    if (order) {
        HIPCHK(c, oi_launch_hits_gather(st, d->hits, order, n, d->d_sorted));
    } else {
        // one mesh of one face: at most one distinct pair
        HIPCHK(c, hipMemcpyAsync(d->d_sorted, d->hits, n * sizeof(uint4), hipMemcpyDeviceToDevice, st));
    }
    d->sorted_ok = true;
    return OIBVH_OK;
}

oibvh_status oi_det_get_pairs(oi_detector* d, uint32_t* out, uint64_t cap, uint32_t* n)
{
    oibvh_ctx* c = d->ctx;
    oibvh_status r = oi_det_check(d);
    if (r != OIBVH_OK) return r;
    if (n) *n = d->n_pairs;
    if (d->n_pairs == 0) return OIBVH_OK;
    if (!out || cap < d->n_pairs) return oi_set_err(c, OIBVH_ERR_CAPACITY, "need %u pairs", d->n_pairs);
    if ((r = sort_hits(d)) != OIBVH_OK) return r;
    HIPCHK(c, hipMemcpyAsync(out, d->d_sorted, d->n_pairs * sizeof(uint4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OIBVH_OK;
}

oibvh_status oi_det_sorted_device(oi_detector* d, const uint4** dptr, uint32_t* n)
{
    oibvh_status r = oi_det_check(d);
    if (r != OIBVH_OK) return r;
    if ((r = sort_hits(d)) != OIBVH_OK) return r;
    // the sort runs on the context's stream: complete it before the pointer
    // is handed to a consumer on any stream or the host
    HIPCHK(d->ctx, hipStreamSynchronize(d->ctx->stream));
    *dptr = d->d_sorted;
    *n = d->n_pairs;
    return OIBVH_OK;
}

oibvh_status oi_det_pair_vertices(oi_detector* d, const oi_build_desc* geom, float* out, uint64_t cap,
                                  uint32_t ptr_kind, uint32_t* n)
{
    oibvh_ctx* c = d->ctx;
    if (ptr_kind != OIBVH_PTR_HOST && ptr_kind != OIBVH_PTR_DEVICE)
        return oi_set_err(c, OIBVH_ERR_INVALID, "pair vertices: ptr_kind must be HOST or DEVICE");
    oibvh_status r = oi_det_check(d);
    if (r != OIBVH_OK) return r;
    if (n) *n = d->n_pairs;
    if (d->n_pairs == 0) return OIBVH_OK;
    const uint64_t need = 18ull * d->n_pairs;
    if (!out || cap < need) return oi_set_err(c, OIBVH_ERR_CAPACITY, "need %llu floats", (unsigned long long)need);
    if ((r = sort_hits(d)) != OIBVH_OK) return r;
    float* dst = out;
    if (ptr_kind == OIBVH_PTR_HOST) {
        if (d->stage_cap < need) {
            dfree(d->d_stage);
            d->stage_cap = 0;
            HIPCHK(c, dalloc(&d->d_stage, need));
            d->stage_cap = need;
        }
        dst = d->d_stage;
    }
    HIPCHK(c, oi_launch_pair_vertices(c->stream, geom, d->d_sorted, d->n_pairs, dst));
    if (ptr_kind == OIBVH_PTR_HOST) {
        HIPCHK(c, hipMemcpyAsync(out, d->d_stage, need * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    // host output must land; device output is complete when this returns
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OIBVH_OK;
}

oibvh_status oi_det_level_counts(oi_detector* d, uint32_t* out, uint64_t cap, uint32_t* n)
{
    oibvh_status r = oi_det_check(d);
    if (r != OIBVH_OK) return r;
    *n = d->have_result ? d->n_levels : 0;
    if (*n == 0) return OIBVH_OK;
    if (!out || cap < *n) return oi_set_err(d->ctx, OIBVH_ERR_CAPACITY, "need %u entries", *n);
    std::memcpy(out, d->h_counts.data(), *n * sizeof(uint32_t));
    return OIBVH_OK;
}
