// Drop-in check: leven's compute interface used exactly as src/clipmap.cpp and
// src/main.cpp use it, through include/lvn_compute.hpp, with the application's own
// types mirrored from src/render_types.h:24-90, src/compute.h:16-31 and src/aabb.h.
// Prints one line per mesh: counts and an FNV-1a hash of the vertex, triangle and
// seam bytes, which tests/test_gpu_dropin.py compares with the same calls made
// through the C ABI.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "lvn_compute.hpp"

namespace app {

struct ivec3 { int x, y, z; };
struct vec4 { float x, y, z, w; };

struct MeshVertex { vec4 position, normal, colour; };     // src/render_types.h:24-38
struct MeshTriangle { int indices_[3]; };                   // src/render_types.h:40-45
constexpr int MAX_MESH_VERTICES = 14336;                    // src/render_types.h:63
constexpr int MAX_MESH_TRIANGLES = 28672;                   // src/render_types.h:68
struct MeshBuffer {                                         // src/render_types.h:70-90
    MeshVertex vertices[MAX_MESH_VERTICES];
    int numVertices = 0;
    MeshTriangle triangles[MAX_MESH_TRIANGLES];
    int numTriangles = 0;
};

struct CSGOperationInfo {                                   // src/compute.h:16-24
    int type = 0;
    int brushShape = 0;
    int material = 0;
    float rotateY = 0.f;
    vec4 origin;
    vec4 dimensions;
};
struct SeamNodeInfo { int localspaceMin[4]; vec4 position, normal; };  // src/compute.h:26-31
struct AABB { ivec3 min, max; };                                      // src/aabb.h

using MeshGenContext = lvn::MeshGenContext<MeshBuffer, SeamNodeInfo, MAX_MESH_VERTICES, MAX_MESH_TRIANGLES>;

}  // namespace app

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

static int report(const char* tag, int r, const app::MeshBuffer& mb, const std::vector<app::SeamNodeInfo>& seams) {
    if (r != 0) {
        std::printf("%s error %d\n", tag, r);
        return r;
    }
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, mb.vertices, sizeof(app::MeshVertex) * (size_t)mb.numVertices);
    h = fnv(h, mb.triangles, sizeof(app::MeshTriangle) * (size_t)mb.numTriangles);
    h = fnv(h, seams.data(), sizeof(app::SeamNodeInfo) * seams.size());
    std::printf("%s %d %d %zu %016llx\n", tag, mb.numVertices, mb.numTriangles, seams.size(), (unsigned long long)h);
    return 0;
}

int main() {
    // the heightfield of the plain configuration: d = y - 24 fbm2(x, z), 6 octaves, f = 1/128
    lvn_field_desc field{};
    field.kind = LVN_FIELD_HEIGHTFIELD;
    field.height_octaves = 6;
    field.height_frequency = 1.f / 128.f;
    field.height_lacunarity = 2.f;
    field.height_gain = 0.5f;
    field.height_amplitude = 24.f;
    if (lvn::Compute_Initialise(1, 0, 2, &field) != 0) return 1;    // src/main.cpp:419
    app::MeshGenContext* ctx = app::MeshGenContext::create(64);     // src/clipmap.cpp:849-850
    if (!ctx) return 1;
    std::printf("voxels %d\n", ctx->voxelsPerChunk());

    static app::MeshBuffer mb;
    std::vector<app::SeamNodeInfo> seams;
    const app::ivec3 min = {0, -128, 0};
    const int size = 256;
    if (report("fresh", ctx->generateChunkMesh(min, size, &mb, seams), mb, seams)) return 1;  // src/clipmap.cpp:378

    bool hasEdges = false;
    ctx->isChunkEmpty(min, size, hasEdges);                         // inverted meaning kept
    std::printf("isChunkEmpty %d\n", hasEdges ? 1 : 0);

    // a stored sphere (src/clipmap.cpp:1618-1643 units: voxel origin + 0.5), applied on
    // the next generation of the chunk after its octree is freed
    app::CSGOperationInfo add;
    add.type = 0;
    add.brushShape = 1;
    add.material = 2;
    add.origin = {32.5f, -4.5f, 32.5f, 0.f};
    add.dimensions = {7.f, 7.f, 7.f, 0.f};
    const app::AABB bb = {{96, -52, 96}, {164, 16, 164}};
    lvn::Compute_StoreCSGOperation(add, bb);                        // src/clipmap.cpp:1718
    ctx->freeChunkOctree(min, size);                                // src/clipmap.cpp:651
    if (report("stored", ctx->generateChunkMesh(min, size, &mb, seams), mb, seams)) return 1;

    // an explicit batch (src/clipmap.cpp:1689,1702): a subtracted box
    std::vector<app::CSGOperationInfo> batch(1);
    batch[0].type = 1;
    batch[0].brushShape = 0;
    batch[0].material = LVN_MATERIAL_AIR;
    batch[0].origin = {20.5f, -2.5f, 44.5f, 0.f};
    batch[0].dimensions = {5.f, 6.f, 4.f, 0.f};
    if (ctx->applyCSGOperations(batch, min, size) != 0) return 1;
    ctx->freeChunkOctree(min, size);
    if (report("applied", ctx->generateChunkMesh(min, size, &mb, seams), mb, seams)) return 1;

    // the seam between this chunk and its +x neighbour (GenerateClipmapSeamMesh,
    // src/clipmap.cpp:573-611, with the two lists its gathering finds)
    std::vector<app::SeamNodeInfo> nbSeams;
    const app::ivec3 nbMin = {min.x + size, min.y, min.z};
    if (report("neighbour", ctx->generateChunkMesh(nbMin, size, &mb, nbSeams), mb, nbSeams)) return 1;
    std::vector<lvn_seam_source> sources = {app::MeshGenContext::seamSource(min, size, 0, seams),
                                            app::MeshGenContext::seamSource(nbMin, size, 4, nbSeams)};
    const app::vec4 colour = {1.f, 0.5f, 0.25f, 0.f};
    const std::vector<app::SeamNodeInfo> none;
    if (report("seam", ctx->generateSeamMesh(min, size, sources, colour, &mb), mb, none)) return 1;

    delete ctx;
    lvn::Compute_ClearCSGOperations();                              // src/volume.cpp:149
    lvn::Compute_Shutdown();                                        // src/main.cpp:599
    return 0;
}
