// oibvh.hpp -- C++ facade with the reference's class interface over the C ABI.
//
// OibvhTree mirrors include/cuda/oibvhTree.cuh:41-174 and Scene mirrors
// include/cuda/scene.cuh:26-182 (minus the OpenGL draw paths). Differences a
// maintainer switching over should know:
//   * failures throw oibvh::Error (the reference asserts or does nothing);
//   * refit() reads the positions of the Mesh it was constructed with, as the
//     reference does (oibvhTree.cu:196-199);
//   * getIntTriPairs() returns the colliding pairs (the reference keeps them
//     private), with original face ids, sorted;
//   * one process-wide context (device + stream) per GPU replaces
//     cudaSetDevice + the default stream.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "oibvh.h"

namespace oibvh {

class Error : public std::runtime_error {
public:
    Error(oibvh_status s, const std::string& m) : std::runtime_error(m), status(s) {}
    oibvh_status status;
};

// DeviceType (scene.cuh:12-24)
enum class DeviceType { CPU = -1, GPU0, GPU1, GPU2, GPU3, GPU4, GPU5, GPU6, GPU7, GPU8 };

// The parts of the reference Mesh (mesh.h:171-187) the path reads: vertex
// positions (3 floats each) and triangle indices (3 per face).
struct Mesh {
    std::vector<float> positions;
    std::vector<uint32_t> indices;
    uint32_t vertexCount() const { return (uint32_t)(positions.size() / 3); }
    uint32_t faceCount() const { return (uint32_t)(indices.size() / 3); }
};

// Model(path) (model.cpp:40-186) for the collision path: positions and
// triangles of a Wavefront OBJ file (oibvh_obj_read)
inline std::shared_ptr<Mesh> loadObj(const std::string& path)
{
    uint32_t V = 0, F = 0;
    oibvh_status s = oibvh_obj_read(nullptr, path.c_str(), nullptr, 0, nullptr, 0, &V, &F);
    if (s != OIBVH_OK && s != OIBVH_ERR_CAPACITY) throw Error(s, "loadObj: cannot read " + path);
    auto m = std::make_shared<Mesh>();
    m->positions.resize(3ull * V);
    m->indices.resize(3ull * F);
    s = oibvh_obj_read(nullptr, path.c_str(), m->positions.data(), m->positions.size(), m->indices.data(),
                       m->indices.size(), &V, &F);
    if (s != OIBVH_OK) throw Error(s, "loadObj: cannot read " + path);
    return m;
}

// int_tri_pair_node_t (utils.h:57-61) with original face ids
struct IntTriPair {
    uint32_t bvhIndex[2];
    uint32_t triIndex[2];
};

class Context {
public:
    explicit Context(int device) : device_(device)
    {
        oibvh_status s = oibvh_ctx_create(device, nullptr, &h_);
        if (s != OIBVH_OK) throw Error(s, "oibvh_ctx_create(" + std::to_string(device) + ") failed");
    }
    ~Context() { oibvh_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    oibvh_ctx* handle() const { return h_; }
    int device() const { return device_; }
    void check(oibvh_status s, const char* what) const
    {
        if (s != OIBVH_OK) throw Error(s, std::string(what) + ": " + oibvh_last_error(h_));
    }
    static std::shared_ptr<Context> get(int device)
    {
        static std::mutex mu;
        static std::map<int, std::weak_ptr<Context>> reg;
        std::lock_guard<std::mutex> lk(mu);
        std::shared_ptr<Context> c = reg[device].lock();
        if (!c) {
            c = std::make_shared<Context>(device);
            reg[device] = c;
        }
        return c;
    }

private:
    int device_;
    oibvh_ctx* h_ = nullptr;
};

class OibvhTree {
public:
    OibvhTree() = delete;

    // OibvhTree(std::shared_ptr<Mesh>) (oibvhTree.cuh:50); `device` picks the GPU
    explicit OibvhTree(const std::shared_ptr<Mesh>& mesh, int device = 0)
        : ctx_(Context::get(device)), mesh_(mesh)
    {
        ctx_->check(oibvh_tree_create(ctx_->handle(), mesh->positions.data(), mesh->vertexCount(),
                                      mesh->indices.data(), mesh->faceCount(), OIBVH_PTR_HOST, &h_),
                    "OibvhTree");
    }

    // OibvhTree(std::shared_ptr<OibvhTree> other, std::shared_ptr<Mesh>) (oibvhTree.cuh:57)
    OibvhTree(const std::shared_ptr<OibvhTree>& other, const std::shared_ptr<Mesh>& mesh)
        : ctx_(other->ctx_), mesh_(mesh)
    {
        ctx_->check(oibvh_tree_create_like(other->h_, mesh->positions.data(), mesh->vertexCount(), OIBVH_PTR_HOST,
                                           &h_),
                    "OibvhTree(copy)");
    }

    ~OibvhTree() { oibvh_tree_destroy(h_); }
    OibvhTree(const OibvhTree&) = delete;
    OibvhTree& operator=(const OibvhTree&) = delete;

    // build() (oibvhTree.cu:237-388)
    void build() { ctx_->check(oibvh_tree_build(h_), "OibvhTree::build"); }

    // refit() (oibvhTree.cu:193-235): takes the mesh's current positions
    void refit()
    {
        ctx_->check(oibvh_tree_refit(h_, mesh_->positions.data(), OIBVH_PTR_HOST), "OibvhTree::refit");
    }

    unsigned int getDepth() const
    {
        uint32_t d = 0;
        ctx_->check(oibvh_tree_get_depth(h_, &d), "getDepth");
        return d;
    }

    unsigned int getPrimCount() const
    {
        uint32_t n = 0;
        ctx_->check(oibvh_tree_get_prim_count(h_, &n), "getPrimCount");
        return n;
    }

    // m_aabbTree: 6 floats per node, real order
    std::vector<float> aabbTree() const
    {
        uint32_t n = 0;
        ctx_->check(oibvh_tree_get_node_count(h_, &n), "node count");
        std::vector<float> out(6ull * n);
        ctx_->check(oibvh_tree_get_nodes(h_, out.data(), n, &n), "aabbTree");
        return out;
    }

    // m_faces after build: faces in leaf order
    std::vector<uint32_t> sortedFaces() const
    {
        std::vector<uint32_t> out(3ull * getPrimCount());
        ctx_->check(oibvh_tree_get_sorted_faces(h_, out.data(), getPrimCount()), "sortedFaces");
        return out;
    }

    // Mesh::transform(glm::mat4) (mesh.cpp:189-214) applied on the fly: later
    // build/refit/detect see mat4 * (p, 1). mat4 = 16 floats, column-major
    // (glm::value_ptr of a glm::mat4); nullptr restores the identity.
    void setTransform(const float* mat4) { ctx_->check(oibvh_tree_set_transform(h_, mat4), "setTransform"); }
    // Morton key width of later builds: 30 (the reference's) or 63
    void setMortonBits(uint32_t bits) { ctx_->check(oibvh_tree_set_morton_bits(h_, bits), "setMortonBits"); }

    // leaf position -> original face id
    std::vector<uint32_t> sortedTriIndex() const
    {
        std::vector<uint32_t> out(getPrimCount());
        ctx_->check(oibvh_tree_get_perm(h_, out.data(), out.size()), "sortedTriIndex");
        return out;
    }

    oibvh_tree* handle() const { return h_; }
    const std::shared_ptr<Context>& context() const { return ctx_; }

private:
    std::shared_ptr<Context> ctx_;
    std::shared_ptr<Mesh> mesh_;
    oibvh_tree* h_ = nullptr;
};

class Scene {
public:
    // Scene() (scene.cuh:32); `device` picks the GPU
    explicit Scene(int device = 0) : ctx_(Context::get(device))
    {
        ctx_->check(oibvh_scene_create(ctx_->handle(), &h_), "Scene");
    }
    ~Scene() { oibvh_scene_destroy(h_); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    // addOibvhTree (scene.cuh:44); the tree must be built
    void addOibvhTree(std::shared_ptr<OibvhTree> tree)
    {
        ctx_->check(oibvh_scene_add_tree(h_, tree->handle()), "Scene::addOibvhTree");
        trees_.push_back(std::move(tree));
    }

    // detectCollision (scene.cuh:53-55); DeviceType::CPU is the reference's
    // empty TODO (scene.cu:187-190) and is rejected here
    void detectCollision(const DeviceType deviceType = DeviceType::GPU0, const unsigned int entryLevel = 0,
                         const unsigned int expandLevels = 1, uint32_t flags = 0)
    {
        if (deviceType == DeviceType::CPU || (int)deviceType != ctx_->device())
            throw Error(OIBVH_ERR_INVALID, "detectCollision: scene lives on GPU" + std::to_string(ctx_->device()));
        ctx_->check(oibvh_scene_detect(h_, entryLevel, expandLevels, flags), "Scene::detectCollision");
    }

    // one frame of main.cpp's loop: tree i borrows device vertices
    // positions[i] (nullptr keeps its positions), every tree is refit, then
    // detectCollision
    void frame(const float* const* positions, const unsigned int entryLevel = 0, const unsigned int expandLevels = 1,
               uint32_t flags = 0)
    {
        ctx_->check(oibvh_scene_frame(h_, positions, entryLevel, expandLevels, flags), "Scene::frame");
    }

    // getIntTriPairCount (scene.cuh:61)
    unsigned int getIntTriPairCount() const
    {
        uint32_t n = 0;
        ctx_->check(oibvh_scene_get_pair_count(h_, &n), "getIntTriPairCount");
        return n;
    }

    std::vector<IntTriPair> getIntTriPairs() const
    {
        const uint32_t n = getIntTriPairCount();
        std::vector<IntTriPair> out(n);
        uint32_t got = 0;
        if (n)
            ctx_->check(oibvh_scene_get_pairs(h_, reinterpret_cast<uint32_t*>(out.data()), n, &got),
                        "getIntTriPairs");
        out.resize(got);
        return out;
    }

    // Scene::convertToVertexArray (scene.cu:69-93): 6 vertices (xyz) per
    // pair, triangle A's then triangle B's, in getIntTriPairs order
    std::vector<float> convertToVertexArray() const
    {
        const uint32_t n = getIntTriPairCount();
        std::vector<float> out(18ull * n);
        uint32_t got = 0;
        if (n)
            ctx_->check(oibvh_scene_get_pair_vertices(h_, out.data(), out.size(), OIBVH_PTR_HOST, &got),
                        "convertToVertexArray");
        out.resize(18ull * got);
        return out;
    }

    oibvh_detect_stats stats() const
    {
        oibvh_detect_stats s{};
        ctx_->check(oibvh_scene_get_stats(h_, &s), "stats");
        return s;
    }

private:
    std::shared_ptr<Context> ctx_;
    oibvh_scene* h_ = nullptr;
    std::vector<std::shared_ptr<OibvhTree>> trees_;
};

// Many independent scenes built, refit and queried together (no reference
// counterpart): every stage is one launch over all meshes of all scenes.
class Batch {
public:
    explicit Batch(int device = 0) : ctx_(Context::get(device))
    {
        ctx_->check(oibvh_batch_create(ctx_->handle(), &h_), "Batch");
    }
    ~Batch() { oibvh_batch_destroy(h_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // returns the mesh's batch index
    uint32_t addMesh(uint32_t scene, const Mesh& mesh)
    {
        uint32_t idx = 0;
        ctx_->check(oibvh_batch_add_mesh(h_, scene, mesh.positions.data(), mesh.vertexCount(), mesh.indices.data(),
                                         mesh.faceCount(), &idx),
                    "Batch::addMesh");
        return idx;
    }
    uint32_t addInstance(uint32_t scene, uint32_t source)
    {
        uint32_t idx = 0;
        ctx_->check(oibvh_batch_add_instance(h_, scene, source, &idx), "Batch::addInstance");
        return idx;
    }
    // meshCount x 16 floats, column-major
    void setTransforms(const float* mats, uint32_t ptrKind = OIBVH_PTR_HOST)
    {
        ctx_->check(oibvh_batch_set_transforms(h_, mats, ptrKind), "Batch::setTransforms");
    }
    void setMortonBits(uint32_t bits) { ctx_->check(oibvh_batch_set_morton_bits(h_, bits), "Batch::setMortonBits"); }
    void build() { ctx_->check(oibvh_batch_build(h_), "Batch::build"); }
    void refit() { ctx_->check(oibvh_batch_refit(h_), "Batch::refit"); }
    void detectCollision(unsigned int entryLevel = 0, unsigned int expandLevels = 1, uint32_t flags = 0)
    {
        ctx_->check(oibvh_batch_detect(h_, entryLevel, expandLevels, flags), "Batch::detectCollision");
    }
    // refit + detect replayed from a captured hipGraph
    void frame(unsigned int entryLevel = 0, unsigned int expandLevels = 1, uint32_t flags = 0)
    {
        ctx_->check(oibvh_batch_frame(h_, entryLevel, expandLevels, flags), "Batch::frame");
    }
    unsigned int getIntTriPairCount() const
    {
        uint32_t n = 0;
        ctx_->check(oibvh_batch_get_pair_count(h_, &n), "Batch::getIntTriPairCount");
        return n;
    }
    // bvhIndex = batch mesh indices
    std::vector<IntTriPair> getIntTriPairs() const
    {
        const uint32_t n = getIntTriPairCount();
        std::vector<IntTriPair> out(n);
        uint32_t got = 0;
        if (n)
            ctx_->check(oibvh_batch_get_pairs(h_, reinterpret_cast<uint32_t*>(out.data()), n, &got),
                        "Batch::getIntTriPairs");
        out.resize(got);
        return out;
    }
    oibvh_batch* handle() const { return h_; }

private:
    std::shared_ptr<Context> ctx_;
    oibvh_batch* h_ = nullptr;
};

}  // namespace oibvh
