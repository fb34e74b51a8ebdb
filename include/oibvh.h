/*
 * oibvh.h -- C ABI of the MI355X (gfx950) oibvh collision library
 * (liboibvh_hip.so).
 *
 * Drop-in boundary for the reference's OibvhTree (include/cuda/oibvhTree.cuh:41-174,
 * src/cuda/oibvhTree.cu) and Scene (include/cuda/scene.cuh:26-182, src/cuda/scene.cu)
 * classes. The reference exposes plain C++ classes with void methods; here every
 * entry point returns an oibvh_status and nothing throws across the ABI. Handles
 * are opaque; the library copies caller data into device memory it owns unless a
 * *_BORROW flag says otherwise. All device work of a context runs on the
 * context's stream, in call order.
 */
#ifndef OIBVH_H
#define OIBVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define OIBVH_API __declspec(dllexport)
#else
#define OIBVH_API __attribute__((visibility("default")))
#endif

typedef enum {
    OIBVH_OK = 0,
    OIBVH_ERR_INVALID = 1,    /* bad argument (null handle, zero faces, index out of range, ...) */
    OIBVH_ERR_NOT_BUILT = 2,  /* tree used before build (the reference asserts, scene.cu:97) */
    OIBVH_ERR_HIP = 3,        /* a HIP runtime call failed; see oibvh_last_error */
    OIBVH_ERR_CAPACITY = 4,   /* caller buffer too small; the required size is reported */
    OIBVH_ERR_OVERFLOW = 5,   /* a detect buffer would exceed its maximum size */
    OIBVH_ERR_TIMEOUT = 6,    /* traversal exceeded its time limit */
    OIBVH_ERR_NO_DEVICE = 7   /* no usable gfx950 device */
} oibvh_status;

typedef struct oibvh_ctx oibvh_ctx;
typedef struct oibvh_tree oibvh_tree;
typedef struct oibvh_scene oibvh_scene;
typedef struct oibvh_batch oibvh_batch;

/* pointer kinds for positions / indices */
#define OIBVH_PTR_HOST 0u          /* host memory, copied */
#define OIBVH_PTR_DEVICE 1u        /* device memory, copied device-to-device */
#define OIBVH_PTR_DEVICE_BORROW 2u /* device memory used in place until replaced; caller keeps it alive */

/* detect flags */
#define OIBVH_DETECT_SELF 1u          /* also collide every mesh with itself (pairs sharing a vertex skipped) */
#define OIBVH_DETECT_FP64_FALLBACK 2u /* FP32 SAT with a forward error filter; pairs it leaves undecided are
                                         tested again by the FP64 SAT */
#define OIBVH_DETECT_ASYNC 4u         /* enqueue only: no host wait. Results are checked at the next
                                         get/stats call or synchronous detect; a further ASYNC detect
                                         replaces an unchecked one, whose errors go to async_errors */
#define OIBVH_FRAME_OVERLAP 8u        /* oibvh_scene_frame and oibvh_batch_frame: this frame's refit may
                                         run while the previous frame's detect still runs (see there) */

/* ---------------------------------------------------------------- context */

/* Context on HIP device `device` with stream `stream` (a hipStream_t; NULL =
 * the library creates one at default priority). No reference counterpart: the
 * reference uses cudaSetDevice and the default stream (scene.cu:229). */
OIBVH_API oibvh_status oibvh_ctx_create(int device, void* stream, oibvh_ctx** out);
/* Trees and scenes keep their context alive: destroying a context that still
 * has live objects defers the release until the last of them is destroyed. */
OIBVH_API void oibvh_ctx_destroy(oibvh_ctx* ctx);
/* text of the last error on this context ("" if none) */
OIBVH_API const char* oibvh_last_error(const oibvh_ctx* ctx);
/* wait for all work queued on the context's stream */
OIBVH_API oibvh_status oibvh_ctx_synchronize(oibvh_ctx* ctx);
/* the context's hipStream_t */
OIBVH_API void* oibvh_ctx_stream(oibvh_ctx* ctx);

/* ---------------------------------------------------------------- trees */

/* OibvhTree(std::shared_ptr<Mesh>) (oibvhTree.cuh:50, oibvhTree.cu:9-16 with
 * setup() :157-191): a tree over V vertices (xyz, 3 floats each) and F faces
 * (idx, 3 vertex indices each). Nothing is built yet. */
OIBVH_API oibvh_status oibvh_tree_create(oibvh_ctx* ctx, const float* xyz, uint32_t V, const uint32_t* idx,
                                         uint32_t F, uint32_t ptr_kind, oibvh_tree** out);

/* OibvhTree(std::shared_ptr<OibvhTree> other, std::shared_ptr<Mesh>)
 * (oibvhTree.cuh:57, oibvhTree.cu:17-33): a tree over a mesh with other's
 * topology that reuses other's sorted faces and build state; new vertex
 * positions xyz (V must equal other's). Call refit before use. */
OIBVH_API oibvh_status oibvh_tree_create_like(oibvh_tree* other, const float* xyz, uint32_t V, uint32_t ptr_kind,
                                              oibvh_tree** out);

OIBVH_API void oibvh_tree_destroy(oibvh_tree* tree);

/* OibvhTree::build() (oibvhTree.cu:237-388): Morton30 codes of the leaf-box
 * centroids in the vertex AABB, stable radix sort, bottom-up AABB reduce. Uses
 * the tree's current positions. */
OIBVH_API oibvh_status oibvh_tree_build(oibvh_tree* tree);

/* OibvhTree::refit() (oibvhTree.cu:193-235): recompute every AABB over the
 * fixed sorted face order. xyz == NULL refits the current positions, else the
 * positions are replaced first (ptr_kind as in oibvh_tree_create). */
OIBVH_API oibvh_status oibvh_tree_refit(oibvh_tree* tree, const float* xyz, uint32_t ptr_kind);

/* Replace the vertex positions without refitting: the counterpart of the
 * reference's Mesh::transform / rotate / translate (mesh.cpp:170-214), whose
 * new m_vertices the next refit() uploads (oibvhTree.cu:196-199). The next
 * build or refit uses them. */
OIBVH_API oibvh_status oibvh_tree_set_positions(oibvh_tree* tree, const float* xyz, uint32_t ptr_kind);

/* Mesh::transform(glm::mat4) (mesh.cpp:189-214, transform.cu:35-41) applied on
 * the fly: every later build, refit and detect of this tree sees each vertex
 * as mat4 * (x, y, z, 1), evaluated in glm's order for mat4 * vec4
 * ((m0 x + m1 y) + (m2 z + m3), columns m0..m3, no fused multiply-add).
 * mat4: 16 floats, column-major like glm::mat4 (host memory, read before the
 * call returns); NULL restores the identity. The stored positions are not
 * changed. Refit (or build) afterwards. */
OIBVH_API oibvh_status oibvh_tree_set_transform(oibvh_tree* tree, const float* mat4);

/* Morton key width of later builds: 30 (the reference's morton3D, 10 bits per
 * axis, oibvh.cu:6-22; the default) or 63 (21 bits per axis: the same
 * centroid, box and quantisation with 2^21 cells, NaN -> 0). No reference
 * counterpart: 63-bit keys change the leaf order (and so the tree) of meshes
 * whose centroids share a 30-bit cell; the pair set of a detect does not
 * depend on it. */
OIBVH_API oibvh_status oibvh_tree_set_morton_bits(oibvh_tree* tree, uint32_t bits);

/* OibvhTree::getDepth() (oibvhTree.cuh:87, oibvhTree.cu:46-49): ilog2(node count) */
OIBVH_API oibvh_status oibvh_tree_get_depth(const oibvh_tree* tree, uint32_t* depth);
/* OibvhTree::getPrimCount() (oibvhTree.cuh:93, oibvhTree.cu:51-54) */
OIBVH_API oibvh_status oibvh_tree_get_prim_count(const oibvh_tree* tree, uint32_t* count);
/* number of tree nodes, oibvh_get_size (oibvh.cuh:56-59) */
OIBVH_API oibvh_status oibvh_tree_get_node_count(const oibvh_tree* tree, uint32_t* count);
OIBVH_API oibvh_status oibvh_tree_get_vertex_count(const oibvh_tree* tree, uint32_t* count);

/* The tree's AABBs, 6 floats per node (min xyz, max xyz) in real (level-major)
 * order -- the content of OibvhTree::m_aabbTree (oibvhTree.cuh:127) after
 * build/refit. cap = capacity of out in nodes; *n receives the node count.
 * Synchronises the stream. */
OIBVH_API oibvh_status oibvh_tree_get_nodes(oibvh_tree* tree, float* out, uint64_t cap, uint32_t* n);

/* Faces in leaf order, 3 indices each -- OibvhTree::m_faces after build
 * (oibvhTree.cu:377). cap in faces. Synchronises. */
OIBVH_API oibvh_status oibvh_tree_get_sorted_faces(oibvh_tree* tree, uint32_t* out, uint64_t cap);

/* Leaf position -> original face id (no reference accessor: the reference drops
 * the permutation). cap in faces. Synchronises. */
OIBVH_API oibvh_status oibvh_tree_get_perm(oibvh_tree* tree, uint32_t* out, uint64_t cap);

/* Current vertex positions (V*3 floats). Synchronises. */
OIBVH_API oibvh_status oibvh_tree_get_positions(oibvh_tree* tree, float* out, uint64_t cap);

/* Device pointer of the tree's current positions (V*3 floats), for callers
 * that update vertices in place on the context's stream. */
OIBVH_API oibvh_status oibvh_tree_positions_device(oibvh_tree* tree, float** dptr);

/* ---------------------------------------------------------------- scenes */

/* Scene() (scene.cuh:32, scene.cu:12-34) */
OIBVH_API oibvh_status oibvh_scene_create(oibvh_ctx* ctx, oibvh_scene** out);
OIBVH_API void oibvh_scene_destroy(oibvh_scene* scene);

/* Scene::addOibvhTree (scene.cuh:44, scene.cu:95-129): the tree must be built
 * (OIBVH_ERR_NOT_BUILT otherwise). Mesh ids in results are the insertion order. */
OIBVH_API oibvh_status oibvh_scene_add_tree(oibvh_scene* scene, oibvh_tree* tree);

/* Scene::detectCollision(DeviceType, entryLevel, expandLevels) (scene.cuh:53-55,
 * scene.cu:157-185, 225-446): broad phase over every mesh pair i < j (and, with
 * OIBVH_DETECT_SELF, every mesh with itself) and narrow phase. entryLevel sets
 * the initial BVTT front (scene.cu:192-223); expandLevels the levels each side
 * descends per step (values above 3 act as 3). Neither changes the result.
 * Synchronous unless OIBVH_DETECT_ASYNC; retries internally with larger
 * buffers when the hit list or work queue overflows. */
OIBVH_API oibvh_status oibvh_scene_detect(oibvh_scene* scene, uint32_t entry_level, uint32_t expand_levels,
                                          uint32_t flags);

/* Scene::getIntTriPairCount (scene.cuh:61, scene.cu:50-53) */
OIBVH_API oibvh_status oibvh_scene_get_pair_count(oibvh_scene* scene, uint32_t* count);

/* Colliding triangle pairs, 4 uint32 each: meshA, meshB (meshA < meshB, or equal
 * in self mode), original face id in A, original face id in B; sorted
 * lexicographically. The reference keeps them in Scene::m_intTriPairs
 * (scene.cuh:133) with sorted-order face ids and has no getter. cap in pairs;
 * *n receives the pair count (OIBVH_ERR_CAPACITY if it exceeds cap). */
OIBVH_API oibvh_status oibvh_scene_get_pairs(oibvh_scene* scene, uint32_t* out, uint64_t cap, uint32_t* n);

/* Scene::convertToVertexArray (scene.cu:69-93; Scene::draw uploads the
 * result to a GL buffer, scene.cu:55-66): the triangles of every colliding
 * pair as 18 floats, the three xyz vertices of triangle A then the three of
 * triangle B, in oibvh_scene_get_pairs order. Vertices are the trees' current
 * positions through their model matrices (as a detect sees them), gathered on
 * the device. out: host memory (ptr_kind OIBVH_PTR_HOST) or device memory
 * (OIBVH_PTR_DEVICE, e.g. a mapped vertex buffer); cap in floats; *n receives
 * the pair count (OIBVH_ERR_CAPACITY if 18 * n exceeds cap). Synchronises. */
OIBVH_API oibvh_status oibvh_scene_get_pair_vertices(oibvh_scene* scene, float* out, uint64_t cap,
                                                     uint32_t ptr_kind, uint32_t* n);

/* Device pointer of the unsorted hit list (uint4 per pair, as above) and its
 * length, valid until the next detect. Synchronises. */
OIBVH_API oibvh_status oibvh_scene_hits_device(oibvh_scene* scene, void** dptr, uint32_t* n);

/* The colliding pairs in oibvh_scene_get_pairs order, sorted on the device:
 * device pointer (4 uint32 per pair) and count, valid until the next detect.
 * The reference downloads its pair list (scene.cu:414-416); here it can stay
 * in HBM for a device-side consumer. Synchronises (the sort is complete on
 * return). */
OIBVH_API oibvh_status oibvh_scene_pairs_device(oibvh_scene* scene, const uint32_t** dptr, uint32_t* n);

typedef struct {
    uint64_t bvtt_tests;    /* BVTT node pairs tested */
    uint64_t candidates;    /* leaf pairs sent to the narrow phase */
    uint64_t fp64_rechecks; /* SAT axis comparisons evaluated in FP64 (OIBVH_DETECT_FP64_FALLBACK) */
    uint64_t adjacent;      /* self-mode leaf pairs skipped for sharing a vertex */
    uint32_t pairs;         /* colliding pairs */
    uint32_t retries;       /* detect re-runs after buffer growth */
    uint32_t async_errors;  /* error bits of earlier OIBVH_DETECT_ASYNC detects whose results were
                               never checked (their pair lists were incomplete); 0 when all were clean */
    uint32_t fp64_candidates; /* candidate pairs the FP32 filter left to the FP64 SAT */
} oibvh_detect_stats;

/* counters of the last detect (no reference counterpart; the reference prints
 * counts, scene.cu:318-321, 417-421). Synchronises. */
OIBVH_API oibvh_status oibvh_scene_get_stats(oibvh_scene* scene, oibvh_detect_stats* stats);

/* Tuning knobs (no reference counterpart; none changes results). Zero fields
 * select the defaults. */
typedef struct {
    uint32_t reduce_leaves_per_thread; /* 1, 2, 4 or 8: the build/refit reduce gives each workgroup a
                                          subtree of 256 x this many leaves; 16: one subtree of 256
                                          leaves per wave (4 per lane, interleaved), four per
                                          workgroup (0 = by mesh size) */
    uint32_t traverse_blocks;          /* workgroups per BVTT-level and narrow-phase launch (0 = 4 per CU) */
    uint32_t traverse_expand;          /* levels each side descends per BVTT step, 1..3 (0 = the
                                          detect call's expandLevels, clamped to 1..3) */
    uint32_t front_capacity;           /* initial BVTT front entries per buffer, a power of two (0 = 2^22) */
    uint32_t hits_capacity;            /* initial hit-list entries (0 = 2^20) */
    uint32_t candidate_capacity;       /* initial candidate leaf-pair entries (0 = 2^20) */
    uint32_t traverse_schedule;        /* levels descended by the first BVTT steps, 2 bits per step
                                          (step k: bits 2k+1..2k, values 1..3); the first zero field
                                          ends the list and later steps use the uniform setting
                                          (0 = uniform) */
    uint32_t traverse_local;           /* 0: BVTT steps only (one launch per step, each spread over the
                                          whole GPU); n >= 1: the first steps as one launch in which
                                          workgroups of n initial pairs walk breadth first with their
                                          fronts in LDS (1..2048). When a workgroup's front outgrows
                                          LDS the detect is re-run with steps and the walk stays off
                                          until the tuning is set again; an OIBVH_DETECT_ASYNC detect
                                          whose walk overflowed is incomplete (async_errors) */
    uint32_t traverse_local_levels;    /* steps the local walk replaces (0 = all of them) */
    uint32_t narrow_separate;          /* FP32 queries: 0 tests leaf pairs inside the BVTT launches that
                                          find them; 1 collects them and tests them in one k_narrow
                                          launch over the whole GPU (the FP64 fallback always does) */
    uint32_t traverse_tail;            /* 0: none; n >= 1: the last steps as one launch in which
                                          workgroups take n pairs at a time of the front the steps
                                          left and walk them to the leaves with their fronts in LDS
                                          (1..2048); an overflow is handled as for traverse_local */
    uint32_t traverse_tail_steps;      /* steps the tail walk replaces, counted from the last (0 = all
                                          after the head local walk, or all) */
    uint32_t refit_blocks;             /* scene refits whose stage 1 is the wave kernel
                                          (reduce_leaves_per_thread 16, or 0 on launches of 2^18
                                          faces or more): at most this many workgroups in that
                                          launch (rounded up to a multiple of 8, up to 2^20), each
                                          looping over subtrees of 1024 leaves, so that an
                                          overlapped frame's refit leaves room on the GPU for the
                                          detect beside it (0 = one workgroup per 1024 leaves) */
} oibvh_tuning;

/* The detect buffers grow on overflow (the detect is re-run). */
OIBVH_API oibvh_status oibvh_ctx_set_tuning(oibvh_ctx* ctx, const oibvh_tuning* tuning);
OIBVH_API oibvh_status oibvh_ctx_get_tuning(const oibvh_ctx* ctx, oibvh_tuning* tuning);

/* Mesh::transform (mesh.cpp:189-214) for every tree of the scene at once:
 * mats holds 16 floats per tree (column-major like glm::mat4) in
 * oibvh_scene_add_tree order, and each tree then behaves as after
 * oibvh_tree_set_transform with its matrix. OIBVH_PTR_HOST: staged in stream
 * order before return (one small launch per tree); OIBVH_PTR_DEVICE: one
 * copy, in stream order, into a table the scene owns; OIBVH_PTR_DEVICE_BORROW:
 * read in place by every later build, refit and detect of these trees (no
 * device work; the caller keeps the memory unchanged until that work is
 * done). A tree in several scenes keeps the matrix of the last call. */
OIBVH_API oibvh_status oibvh_scene_set_transforms(oibvh_scene* scene, const float* mats, uint32_t ptr_kind);

/* Build every tree of the scene (OibvhTree::build on each, oibvhTree.cu:237-388;
 * no reference counterpart as one call): up to four consecutive trees of one
 * Morton width share every launch of the build chain, the sort treating each
 * tree's keys as a segment of its own; a build whose launches repeat the
 * previous call's is replayed from a captured hipGraph. Results are those of
 * oibvh_tree_build on every tree. Trees of a scene must be built before
 * oibvh_scene_add_tree, so this rebuilds (after new positions or matrices). */
OIBVH_API oibvh_status oibvh_scene_build(oibvh_scene* scene);

/* Refit every tree of the scene (OibvhTree::refit on each, oibvhTree.cu:193-235)
 * with one launch per stage for up to four trees instead of one per tree.
 * Results are those of oibvh_tree_refit on every tree with its current
 * positions and transform. */
OIBVH_API oibvh_status oibvh_scene_refit(oibvh_scene* scene);

/* One frame of the reference's loop (main.cpp:284: every tree refit, then
 * Scene::detectCollision) in one call: tree i first borrows device vertex
 * memory positions[i] (as oibvh_tree_set_positions with
 * OIBVH_PTR_DEVICE_BORROW; a NULL entry, or positions == NULL, keeps the
 * current positions), then oibvh_scene_refit, then oibvh_scene_detect with
 * entry_level, expand_levels and flags.
 *
 * With OIBVH_FRAME_OVERLAP (scenes of two or more trees, or one tree in self
 * mode), frames whose positions are known ahead (a prescribed animation) overlap:
 * each tree keeps four node buffers, frame f's refit writes the one frame
 * f - 4's detect read, and the detects run on two more streams of the context
 * (alternating by frame), so frame f + 1's refit, and its detect, run while
 * frame f's detect does. The call first waits on the host for frame f - 4's
 * detect to end, so the host stays at most four frames ahead of the GPU.
 * Results and errors are those of the same frames without the flag.
 * The caller keeps frame f's position memory (and model matrices) unchanged
 * until frame f's results are read or a later call that waits for it; every
 * other entry point first waits for the overlapped detects queued before it
 * (on the GPU, not the host). */
OIBVH_API oibvh_status oibvh_scene_frame(oibvh_scene* scene, const float* const* positions, uint32_t entry_level,
                                         uint32_t expand_levels, uint32_t flags);

/* Timing of a scene's work with HIP events on the context's stream (no
 * reference counterpart; the reference times with cudaEvents around its own
 * kernels, oibvhTree.cu:300-355, scene.cu:262-433). With every = n > 0, every
 * n-th oibvh_scene_refit has its (first) stage-1 reduce launch (k_reduce_w or
 * k_reduce) record an event pair as it starts and completes (the kernel's own
 * duration), and every n-th detect records one around its launches; every = 0
 * stops. A detect's event pair costs the GPU a few microseconds of pipeline
 * drain, so a sample keeps the timed work close to untimed work.
 * oibvh_scene_get_timing waits for the recorded events, sums their durations
 * and starts the counts afresh. */
typedef struct {
    double refit_ms;   /* summed stage-1 reduce durations */
    double detect_ms;  /* summed detect durations (first launch to last) */
    uint32_t refits;
    uint32_t detects;
} oibvh_scene_timing;
OIBVH_API oibvh_status oibvh_scene_set_timing(oibvh_scene* scene, uint32_t every);
OIBVH_API oibvh_status oibvh_scene_get_timing(oibvh_scene* scene, oibvh_scene_timing* out);

/* BVTT front length of each level of the last detect (level 0 = the initial
 * front); *n receives the number of levels. Diagnostics; the reference prints
 * these (scene.cu:274-276). */
OIBVH_API oibvh_status oibvh_scene_get_level_counts(oibvh_scene* scene, uint32_t* out, uint64_t cap, uint32_t* n);

/* ---------------------------------------------------------------- batches
 * No reference counterpart. Many independent scenes in one object: every
 * stage (build, refit, each BVTT step, narrow phase) is one launch over all
 * meshes of all scenes. Meshes of one scene are tested against each other
 * (and each against itself with OIBVH_DETECT_SELF), meshes of different
 * scenes never. Per mesh, results equal those of an oibvh_tree over the same
 * (transformed) vertices, and per scene those of an oibvh_scene over its
 * meshes in batch order. */

OIBVH_API oibvh_status oibvh_batch_create(oibvh_ctx* ctx, oibvh_batch** out);
OIBVH_API void oibvh_batch_destroy(oibvh_batch* batch);

/* Add a mesh to scene `scene` (any number; scenes need not be contiguous):
 * V vertices xyz and F faces idx in host memory, copied. *index receives the
 * mesh's batch index (insertion order). Only before the first build. */
OIBVH_API oibvh_status oibvh_batch_add_mesh(oibvh_batch* batch, uint32_t scene, const float* xyz, uint32_t V,
                                            const uint32_t* idx, uint32_t F, uint32_t* index);

/* Add a rigid instance of mesh `source` to scene `scene`: it shares the
 * source's vertex positions and faces and has its own model matrix, tree and
 * leaf order. Only before the first build. */
OIBVH_API oibvh_status oibvh_batch_add_instance(oibvh_batch* batch, uint32_t scene, uint32_t source,
                                                uint32_t* index);

OIBVH_API oibvh_status oibvh_batch_mesh_count(const oibvh_batch* batch, uint32_t* n);

/* Model matrices of all meshes: mesh_count x 16 floats, column-major like
 * glm::mat4, applied on the fly as oibvh_tree_set_transform does. Host memory
 * is read before return; device memory (OIBVH_PTR_DEVICE) is copied in stream
 * order (OIBVH_PTR_DEVICE_BORROW too), so a caller can keep every frame's
 * matrices in HBM. */
OIBVH_API oibvh_status oibvh_batch_set_transforms(oibvh_batch* batch, const float* mats, uint32_t ptr_kind);

/* Morton key width (30 or 63) of the next build of every mesh, as
 * oibvh_tree_set_morton_bits. */
OIBVH_API oibvh_status oibvh_batch_set_morton_bits(oibvh_batch* batch, uint32_t bits);

/* Device address of a mesh's vertex positions inside the batch (V*3 floats;
 * shared with its instances). Callers may rewrite them in stream order
 * (deforming meshes); the next refit or frame uses them. */
OIBVH_API oibvh_status oibvh_batch_positions_device(oibvh_batch* batch, uint32_t mesh, float** dptr);

/* New vertex positions for one mesh (V*3 floats), as oibvh_tree_set_positions:
 * host memory is copied before return, OIBVH_PTR_DEVICE is copied in stream
 * order into the batch's own buffer, OIBVH_PTR_DEVICE_BORROW makes the mesh
 * read the caller's memory in place from then on (switched in stream order,
 * so captured frames follow it). Instances of the mesh keep their source's
 * vertices. The next build, refit or frame uses them. */
OIBVH_API oibvh_status oibvh_batch_set_positions(oibvh_batch* batch, uint32_t mesh, const float* xyz,
                                                 uint32_t ptr_kind);

/* Build every tree (OibvhTree::build per mesh, oibvhTree.cu:237-388). */
OIBVH_API oibvh_status oibvh_batch_build(oibvh_batch* batch);
/* Refit every tree (OibvhTree::refit per mesh, oibvhTree.cu:193-235). */
OIBVH_API oibvh_status oibvh_batch_refit(oibvh_batch* batch);
/* Scene::detectCollision for every scene (scene.cu:225-446); flags as for
 * oibvh_scene_detect. */
OIBVH_API oibvh_status oibvh_batch_detect(oibvh_batch* batch, uint32_t entry_level, uint32_t expand_levels,
                                          uint32_t flags);
/* One frame, refit + detect, replayed from a hipGraph captured at the first
 * call and again whenever the parameters, the tuning or a device buffer
 * change. With OIBVH_DETECT_ASYNC it returns after the launch.
 *
 * With OIBVH_FRAME_OVERLAP, frames whose inputs are known ahead (prescribed
 * motion) overlap: the batch keeps a second node buffer, matrix table and
 * descriptor tables; frame f refits into one (a graph on the context's
 * stream) while frame f - 1's detect reads the other (a graph on one of the
 * context's two side streams, which alternate by frame). After an overlapped frame, oibvh_batch_set_transforms and
 * oibvh_batch_set_positions with OIBVH_PTR_DEVICE_BORROW stage the next
 * overlapped frame's inputs without waiting for this frame's detect;
 * set_positions with a copy, and every other call, first wait for the queued
 * detects. Results and errors equal those of the same frames without the
 * flag; the refit of such a frame is capped at tuning.refit_blocks
 * workgroups (0: uncapped). */
OIBVH_API oibvh_status oibvh_batch_frame(oibvh_batch* batch, uint32_t entry_level, uint32_t expand_levels,
                                         uint32_t flags);

/* HIP-event timing as oibvh_scene_set_timing / oibvh_scene_get_timing: every
 * n-th refit (oibvh_batch_refit or a frame's refit, which then runs without
 * its graph) has its stage-1 launch record an event pair as it starts and
 * completes; every n-th detect records a pair around its launches. */
OIBVH_API oibvh_status oibvh_batch_set_timing(oibvh_batch* batch, uint32_t every);
OIBVH_API oibvh_status oibvh_batch_get_timing(oibvh_batch* batch, oibvh_scene_timing* out);

OIBVH_API oibvh_status oibvh_batch_get_pair_count(oibvh_batch* batch, uint32_t* count);
/* Sorted (meshA, meshB, origTriA, origTriB) with batch mesh indices,
 * meshA <= meshB, 4 words each; cap in pairs. */
OIBVH_API oibvh_status oibvh_batch_get_pairs(oibvh_batch* batch, uint32_t* out, uint64_t cap, uint32_t* n);
/* as oibvh_scene_pairs_device, with batch mesh indices */
OIBVH_API oibvh_status oibvh_batch_pairs_device(oibvh_batch* batch, const uint32_t** dptr, uint32_t* n);
/* as oibvh_scene_get_pair_vertices, over the batch's meshes */
OIBVH_API oibvh_status oibvh_batch_get_pair_vertices(oibvh_batch* batch, float* out, uint64_t cap,
                                                     uint32_t ptr_kind, uint32_t* n);
OIBVH_API oibvh_status oibvh_batch_get_stats(oibvh_batch* batch, oibvh_detect_stats* stats);
OIBVH_API oibvh_status oibvh_batch_get_level_counts(oibvh_batch* batch, uint32_t* out, uint64_t cap, uint32_t* n);
/* One mesh's tree nodes (as oibvh_tree_get_nodes) and leaf permutation. */
OIBVH_API oibvh_status oibvh_batch_get_nodes(oibvh_batch* batch, uint32_t mesh, float* out, uint64_t cap,
                                             uint32_t* n);
OIBVH_API oibvh_status oibvh_batch_get_perm(oibvh_batch* batch, uint32_t mesh, uint32_t* out, uint64_t cap);

/* ---------------------------------------------------------------- no allocation */

/* Largest F that oibvh_build_noalloc accepts. */
#define OIBVH_NOALLOC_MAX_FACES (1u << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the build of oibvh_tree_build (OibvhTree::build,
 * oibvhTree.cu:237-388; 30-bit keys, no model matrix) of the mesh xyz (V*3
 * floats) / idx (F*3 vertex ids, each < V), all caller-owned device memory,
 * writing the tree's nodes (N*6 floats, N = 2F + popcount(next_pow2(F) - F) - 1,
 * real order as oibvh_tree_get_nodes), the leaf -> original face permutation
 * (F words, as oibvh_tree_get_perm) and the faces in leaf order (F*3 words, as
 * oibvh_tree_get_sorted_faces). Results equal those of oibvh_tree_create with
 * the same data plus oibvh_tree_build, bit for bit.
 *
 * Allocates nothing: the build's scratch (sort keys and values, digit
 * histograms, look-back tiles, mesh-box partials; about 13.7 MB) is a fixed
 * device array the library holds on every GPU it runs on until the process
 * ends. F above OIBVH_NOALLOC_MAX_FACES is refused (OIBVH_ERR_CAPACITY).
 * Enqueued on the context's stream; nothing waits. All calls on one GPU share
 * that GPU's array: calls that overlap on one GPU (from different streams or
 * contexts of the device) race; calls on one stream are ordered. */
OIBVH_API oibvh_status oibvh_build_noalloc(oibvh_ctx* ctx, const float* xyz, uint32_t V, const uint32_t* idx,
                                           uint32_t F, float* nodes, uint32_t* perm, uint32_t* sorted_faces);

/* ---------------------------------------------------------------- mesh files */

/* Wavefront OBJ positions and triangles, the mesh ingest the reference gets
 * from assimp (Model::loadModel, model.cpp:40-186, with aiProcess_Triangulate).
 * 'v' lines give positions (file order), 'f' lines polygons (corners i, i/t,
 * i//n, i/t/n; 1-based or negative indices), fan-triangulated as (c1, cj,
 * cj+1); every other statement is skipped and all groups form one mesh.
 * Host only, no device work. xyz: 3 floats per position (xyz_cap floats),
 * idx: 3 indices per triangle (idx_cap words). *V and *F always receive the
 * counts; with null or short buffers the call returns OIBVH_ERR_CAPACITY
 * (call once to size, once to read). ctx may be NULL; with a context, parse
 * errors (OIBVH_ERR_INVALID) leave their file:line text in oibvh_last_error. */
OIBVH_API oibvh_status oibvh_obj_read(oibvh_ctx* ctx, const char* path, float* xyz, uint64_t xyz_cap,
                                      uint32_t* idx, uint64_t idx_cap, uint32_t* V, uint32_t* F);

#ifdef __cplusplus
}
#endif
#endif /* OIBVH_H */
