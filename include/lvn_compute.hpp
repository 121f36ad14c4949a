// lvn_compute.hpp — C++ drop-in for leven's compute interface (src/compute.h:35-77)
// over the C ABI in lvn_mesher.h.
//
// Header-only.  The application's own types are template parameters, so the
// header carries no glm or renderer dependency:
//   IVec3       anything with int members x, y, z            (glm::ivec3)
//   MeshBufferT anything with vertices[], numVertices, triangles[], numTriangles,
//               whose elements are layout-compatible with lvn_vertex / lvn_triangle
//               and whose arrays hold kMaxVertices / kMaxTriangles elements (MeshBuffer,
//               src/render_types.h:70-90)
//   SeamNodeT   48-byte SeamNodeInfo (src/compute.h:26-31)
//   CSGOpT      48-byte CSGOperationInfo (src/compute.h:16-24)
// A leven build replaces src/compute.h with a few lines that define its structs and
// alias these templates (INTEGRATION.md shows them).
#ifndef LVN_COMPUTE_HPP
#define LVN_COMPUTE_HPP

#include <cstring>
#include <vector>

#include "lvn_mesher.h"

namespace lvn {

// The reference's non-OpenCL failure code (src/compute.h:12); LVN_ERR_* codes pass through.
constexpr int kComputeError = -99999;

// leven's own density function (DensityFunc with Terrain, cl/noise.cl:205-268).
inline lvn_field_desc leven_terrain_field() {
    lvn_field_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = LVN_FIELD_LEVEN_TERRAIN;
    return d;
}

// Compute_Initialise(noiseSeed, defaultMaterial, numCSGBrushes), src/compute.cpp:195-234.
// numCSGBrushes is fixed at 2 (cube, sphere) by the brush set; other values are rejected.
inline int Compute_Initialise(int noiseSeed, unsigned defaultMaterial, int numCSGBrushes,
                              const lvn_field_desc* field = nullptr) {
    if (numCSGBrushes != 2) return LVN_ERR_INVALID;
    const lvn_field_desc leven = leven_terrain_field();
    return lvn_init(field ? field : &leven, noiseSeed, defaultMaterial);
}
inline int Compute_Shutdown() { return lvn_shutdown(); }
inline int Compute_SetNoiseSeed(int noiseSeed) { return lvn_set_noise_seed(noiseSeed); }

template <class CSGOpT, class AABBT>
int Compute_StoreCSGOperation(const CSGOpT& op, const AABBT& aabb) {
    static_assert(sizeof(CSGOpT) == sizeof(lvn_csg_op), "CSGOperationInfo must be 48 bytes");
    lvn_csg_op o;
    std::memcpy(&o, &op, sizeof(o));
    const int32_t mn[3] = {aabb.min.x, aabb.min.y, aabb.min.z};
    const int32_t mx[3] = {aabb.max.x, aabb.max.y, aabb.max.z};
    return lvn_store_csg_op(&o, mn, mx);
}
inline int Compute_ClearCSGOperations() { return lvn_clear_csg_ops(); }

// Compute_MeshGenContext (src/compute.h:46-77).
template <class MeshBufferT, class SeamNodeT, int kMaxVertices, int kMaxTriangles>
class MeshGenContext {
public:
    static MeshGenContext* create(int voxelsPerChunk, unsigned flags = 0) {
        lvn_ctx* c = nullptr;
        if (lvn_ctx_create(voxelsPerChunk, flags, &c) != LVN_OK) return nullptr;
        return new MeshGenContext(c);
    }
    ~MeshGenContext() { lvn_ctx_destroy(ctx_); }

    int voxelsPerChunk() const { return lvn_ctx_voxels_per_chunk(ctx_); }

    template <class CSGOpT, class IVec3>
    int applyCSGOperations(const std::vector<CSGOpT>& ops, const IVec3& min, int size) {
        static_assert(sizeof(CSGOpT) == sizeof(lvn_csg_op), "CSGOperationInfo must be 48 bytes");
        const int32_t m[3] = {min.x, min.y, min.z};
        return lvn_apply_csg(ctx_, reinterpret_cast<const lvn_csg_op*>(ops.data()), (int32_t)ops.size(), m, size);
    }

    template <class IVec3>
    int freeChunkOctree(const IVec3& min, int size) {
        const int32_t m[3] = {min.x, min.y, min.z};
        return lvn_free_chunk(ctx_, m, size);
    }

    // Keeps the reference's meaning: isEmpty is set to "the field has edges"
    // (src/compute_density_field.cpp:285,296).
    template <class IVec3>
    int isChunkEmpty(const IVec3& min, int size, bool& isEmpty) {
        const int32_t m[3] = {min.x, min.y, min.z};
        int32_t has = 0;
        const int r = lvn_chunk_has_surface(ctx_, m, size, &has);
        isEmpty = has != 0;
        return r;
    }

    template <class IVec3>
    int generateChunkMesh(const IVec3& min, int size, MeshBufferT* meshBuffer, std::vector<SeamNodeT>& seams) {
        static_assert(sizeof(SeamNodeT) == sizeof(lvn_seam_node), "SeamNodeInfo must be 48 bytes");
        seams.clear();
        const int32_t m[3] = {min.x, min.y, min.z};
        int32_t nv = 0, nt = 0, ns = 0;
        // seam nodes land in the context's staging list (sized once for the largest
        // chunk), then only the ns written are copied into the caller's vector
        const int r = lvn_generate_chunk_mesh(ctx_, m, size, reinterpret_cast<lvn_vertex*>(&meshBuffer->vertices[0]),
                                              kMaxVertices, &nv, reinterpret_cast<int32_t*>(&meshBuffer->triangles[0]),
                                              kMaxTriangles, &nt, seamStaging_.data(), (int32_t)seamStaging_.size(),
                                              &ns);
        if (r != LVN_OK) return r;
        meshBuffer->numVertices = nv;
        meshBuffer->numTriangles = nt;
        seams.resize((size_t)ns);
        if (ns) std::memcpy(seams.data(), seamStaging_.data(), sizeof(lvn_seam_node) * (size_t)ns);
        return LVN_OK;
    }

    // The body of GenerateClipmapSeamMesh (src/clipmap.cpp:573-611) after its neighbour
    // gathering: the seam mesh of the node at (min, size) from the gathered lists (one
    // seamSource per list SelectSeamNodes would have filtered, src/clipmap.cpp:594-595).
    // Writes into the caller's MeshBuffer; LVN_ERR_CAPACITY when the seam does not fit.
    template <class IVec3, class Vec3>
    int generateSeamMesh(const IVec3& min, int size, const std::vector<lvn_seam_source>& sources, const Vec3& colour,
                         MeshBufferT* meshBuffer) {
        const int32_t m[3] = {min.x, min.y, min.z};
        const float c[3] = {colour.x, colour.y, colour.z};
        int32_t nv = 0, nt = 0;
        const int r = lvn_generate_seam_mesh(ctx_, m, size, sources.data(), (int32_t)sources.size(), c,
                                             reinterpret_cast<lvn_vertex*>(&meshBuffer->vertices[0]), kMaxVertices,
                                             &nv, reinterpret_cast<int32_t*>(&meshBuffer->triangles[0]), kMaxTriangles,
                                             &nt);
        if (r != LVN_OK) return r;
        meshBuffer->numVertices = nv;
        meshBuffer->numTriangles = nt;
        return LVN_OK;
    }

    // A clipmap node's seam node list (generateChunkMesh's output) found under candidate
    // neighbour `neighbour` (CHILD_MIN_OFFSETS index) of the host.
    template <class IVec3>
    static lvn_seam_source seamSource(const IVec3& min, int size, int neighbour, const std::vector<SeamNodeT>& nodes) {
        static_assert(sizeof(SeamNodeT) == sizeof(lvn_seam_node), "SeamNodeInfo must be 48 bytes");
        lvn_seam_source s;
        s.min[0] = min.x;
        s.min[1] = min.y;
        s.min[2] = min.z;
        s.size = size;
        s.neighbour = neighbour;
        s.count = (int32_t)nodes.size();
        s.nodes = reinterpret_cast<const lvn_seam_node*>(nodes.data());
        return s;
    }

    lvn_ctx* handle() const { return ctx_; }

private:
    // A chunk's seam nodes are its active voxels with a coordinate at 0 or N - 1: at most
    // N^3 - (N - 2)^3 <= 6 N^2 of them.
    explicit MeshGenContext(lvn_ctx* c)
        : ctx_(c), seamStaging_((size_t)6 * lvn_ctx_voxels_per_chunk(c) * lvn_ctx_voxels_per_chunk(c)) {}
    lvn_ctx* ctx_;
    std::vector<lvn_seam_node> seamStaging_;
};

}  // namespace lvn

#endif  // LVN_COMPUTE_HPP
