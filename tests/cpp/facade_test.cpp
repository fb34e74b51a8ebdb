// facade_test.cpp -- drives include/oibvh.hpp the way the reference demo drives
// its classes (main.cpp:127-151, 284): build tree 1, copy-construct tree 2
// over a second mesh with the same topology, refit it, add both to a Scene,
// detectCollision(GPU0, 4, 3), read the pairs and their triangles
// (convertToVertexArray, scene.cu:69-93); optionally load an OBJ file the way
// the demo loads its model (main.cpp:127).
//
// usage: facade_test <in.bin> <out.bin> [model.obj]
//   in:  uint32 V, uint32 F, float xyz1[3V], float xyz2[3V], uint32 faces[3F]
//   out: uint32 N, float nodes1[6N], float nodes2[6N], uint32 perm[F],
//        uint32 H, uint32 pairs[4H], uint32 depth, uint32 prim_count,
//        float vertices[18H], and with model.obj: uint32 V', uint32 F',
//        float xyz[3V'], uint32 faces[3F']
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "oibvh.hpp"

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return fread(p, sizeof(T), n, f) == n;
}
template <class T>
static void wr(FILE* f, const T* p, size_t n)
{
    fwrite(p, sizeof(T), n, f);
}

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 4) {
        fprintf(stderr, "usage: %s in.bin out.bin [model.obj]\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t V = 0, F = 0;
    if (!rd(in, &V, 1) || !rd(in, &F, 1)) return 2;
    auto m1 = std::make_shared<oibvh::Mesh>();
    auto m2 = std::make_shared<oibvh::Mesh>();
    m1->positions.resize(3ull * V);
    m2->positions.resize(3ull * V);
    m1->indices.resize(3ull * F);
    if (!rd(in, m1->positions.data(), 3ull * V) || !rd(in, m2->positions.data(), 3ull * V) ||
        !rd(in, m1->indices.data(), 3ull * F))
        return 2;
    fclose(in);
    m2->indices = m1->indices;
    try {
        auto t1 = std::make_shared<oibvh::OibvhTree>(m1);
        t1->build();
        auto t2 = std::make_shared<oibvh::OibvhTree>(t1, m2);
        t2->refit();
        oibvh::Scene scene;
        scene.addOibvhTree(t1);
        scene.addOibvhTree(t2);
        scene.detectCollision(oibvh::DeviceType::GPU0, 4, 3);
        std::vector<oibvh::IntTriPair> pairs = scene.getIntTriPairs();
        std::vector<float> n1 = t1->aabbTree(), n2 = t2->aabbTree();
        std::vector<uint32_t> perm = t1->sortedTriIndex();
        FILE* out = fopen(argv[2], "wb");
        if (!out) return 2;
        uint32_t N = (uint32_t)(n1.size() / 6), H = (uint32_t)pairs.size();
        wr(out, &N, 1);
        wr(out, n1.data(), n1.size());
        wr(out, n2.data(), n2.size());
        wr(out, perm.data(), perm.size());
        wr(out, &H, 1);
        wr(out, reinterpret_cast<const uint32_t*>(pairs.data()), 4ull * H);
        uint32_t depth = t1->getDepth(), prims = t1->getPrimCount();
        wr(out, &depth, 1);
        wr(out, &prims, 1);
        std::vector<float> verts = scene.convertToVertexArray();
        wr(out, verts.data(), verts.size());
        if (argc == 4) {
            std::shared_ptr<oibvh::Mesh> model = oibvh::loadObj(argv[3]);
            uint32_t mv = model->vertexCount(), mf = model->faceCount();
            wr(out, &mv, 1);
            wr(out, &mf, 1);
            wr(out, model->positions.data(), model->positions.size());
            wr(out, model->indices.data(), model->indices.size());
        }
        fclose(out);
        printf("facade ok: %u nodes, %u pairs\n", N, H);
    } catch (const oibvh::Error& e) {
        fprintf(stderr, "oibvh error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
