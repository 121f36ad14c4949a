// lvn_seams.hip — seam stitching on the GPU.
//
// Replaces leven's CPU seam path for a clipmap node (GenerateClipmapSeamMesh,
// src/clipmap.cpp:573-611): SelectSeamNodes / FilterSeamNode (:508-569), the
// bottom-up seam octree (Octree_ConstructUpwards, src/octree.cpp:22-148) and its
// mesh (Octree_GenerateMesh, src/octree.cpp:520-550: GenerateVertexIndices
// :196-231, ContourCellProc / FaceProc / EdgeProc / ProcessEdge :235-474).
//
// Two forms of one pass: each step below runs as one launch over the elements of all
// seams of a request (k_sg_*: every seam spread over the whole GPU), or one workgroup
// stitches one seam (k_seam, for requests of small seams only).  The
// reference builds the octree with a hash map per level and contours it by recursion;
// here each phase is data-parallel:
//   1. select: every offered seam node is filtered and its leaf key
//      (level << 48 | Morton code in the root) inserted into an open-addressing
//      table in HBM (a repeated node keeps the last offer, as the reference's
//      child-slot assignment does);
//   2. ancestors: each leaf inserts its chain of parents, stopping at the first
//      one already present, so every internal node is created exactly once;
//   3. links, subtree leaf counts (bottom-up, level by level) and depth-first leaf
//      ranks (top-down prefix over the eight children, level by level): the rank is
//      the vertex index GenerateVertexIndices assigns;
//   4. contouring as a breadth-first expansion of the recursion's calls: each
//      round replaces every pending cell / face / edge call by the calls it makes,
//      in call order, at positions given by prefix sums over the calls; a finished
//      quad is carried forward in place.  The list therefore stays in depth-first
//      call order and its final content is the reference's triangle order.
// Everything is integer bookkeeping; vertex data is copied from the seam nodes.

#include "lvn_runtime.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <thread>
#include <vector>

namespace lvn {

// ---------------------------------------------------------------------------
// Device records
// ---------------------------------------------------------------------------

struct SeamSrcDev {  // one offered seam node list
    const lvn_seam_node* nodes;
    int count;
    int neighbour;
    int size;
    int min[3];
};

struct SeamJobDev {  // one seam
    int host_min[3];
    int host_size;
    float colour[4];
    int src_begin, src_count;
    int vpc;
    int leaf_cap, node_cap, list_cap, hash_cap;
    int pad;
    uint64_t scratch;  // byte offset of the seam's scratch
    uint64_t v_off;    // vertex offset in the unpacked output (leaf_cap vertices)
    uint64_t t_off;    // triangle offset in the unpacked output (2 * list_cap triangles)
};

struct SeamResDev {
    int nv, nt, status, rounds;
};

enum : int { kNodeInternal = 1, kNodeLeaf = 2 };
enum : int { kCallCell = 0, kCallFace = 1, kCallEdge = 2, kCallQuad = 3 };

struct SNode {
    int min[3];
    int size;
    int parent, level, mat, count;
    int offset, type;
    uint64_t code;
    uint64_t payload;  // leaves: (source + 1) << 32 | node index
};

// What contouring reads of a node, in one word beside its eight child links: the LOD0
// chunk of its min relative to the root's (3 x 14 bits; ChunkMinForPosition,
// src/volume.cpp:30-34, masks with CLIPMAP_LEAF_SIZE = 256), internal flag, log2 size,
// corner signs, and whether its whole region lies in that one chunk.
constexpr uint64_t kShapeChunkMask = (1ull << 42) - 1;
__device__ inline uint64_t node_shape(const int* world_min, const int* host_min, int size, bool internal,
                                      int corners) {
    uint64_t key = 0;
    bool one_chunk = true;
    for (int a = 0; a < 3; ++a) {
        key |= (uint64_t)(((world_min[a] >> 8) - (host_min[a] >> 8)) & 0x3fff) << (14 * a);
        one_chunk &= (world_min[a] >> 8) == ((world_min[a] + size - 1) >> 8);
    }
    return key | (uint64_t)internal << 42 | (uint64_t)(__builtin_ctz((unsigned)size) & 31) << 43 |
           (uint64_t)(corners & 0xff) << 48 | (uint64_t)one_chunk << 56;
}

// A pending call of the recursion in 16 bytes: four node indices (-1 = none; the first is
// always a node), the first carrying meta = kind | dir << 2 | flip << 4 in its top 5 bits.
// Node indices stay below 2^27 (checked when a seam's caps are set).
constexpr int kCallNodeBits = 27;
constexpr int kCallNodeMask = (1 << kCallNodeBits) - 1;
struct alignas(16) Call {
    int w[4];
    __device__ int node(int j) const { return j == 0 ? (w[0] & kCallNodeMask) : w[j]; }
    __device__ int meta() const { return (int)((unsigned)w[0] >> kCallNodeBits); }
};
__device__ inline Call make_call(int a, int b, int c, int d, int meta) {
    return Call{{a | meta << kCallNodeBits, b, c, d}};
}

constexpr int kSeamThreads = 1024;
constexpr int kSeamWaves = kSeamThreads / 64;
constexpr int kCallsPerThread = 2;  // pending calls one thread expands per round tile (1024 x 2 timed best)
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kMaxRounds = 64;

// Dual contouring tables (src/contour_constants.h); child i of a node sits at
// (bit 2, bit 1, bit 0) of i (CHILD_MIN_OFFSETS, src/volume_constants.h:24-35).
constexpr int kEdgeCorners[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                     {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};
constexpr int kCellFaces[12][3] = {{0, 4, 0}, {1, 5, 0}, {2, 6, 0}, {3, 7, 0}, {0, 2, 1}, {4, 6, 1},
                                   {1, 3, 1}, {5, 7, 1}, {0, 1, 2}, {2, 3, 2}, {4, 5, 2}, {6, 7, 2}};
constexpr int kCellEdges[6][5] = {{0, 1, 2, 3, 0}, {4, 5, 6, 7, 0}, {0, 4, 1, 5, 1},
                                  {2, 6, 3, 7, 1}, {0, 2, 4, 6, 2}, {1, 3, 5, 7, 2}};
constexpr int kFaceFaces[3][4][3] = {{{4, 0, 0}, {5, 1, 0}, {6, 2, 0}, {7, 3, 0}},
                                     {{2, 0, 1}, {6, 4, 1}, {3, 1, 1}, {7, 5, 1}},
                                     {{1, 0, 2}, {3, 2, 2}, {5, 4, 2}, {7, 6, 2}}};
constexpr int kFaceEdges[3][4][6] = {
    {{1, 4, 0, 5, 1, 1}, {1, 6, 2, 7, 3, 1}, {0, 4, 6, 0, 2, 2}, {0, 5, 7, 1, 3, 2}},
    {{0, 2, 3, 0, 1, 0}, {0, 6, 7, 4, 5, 0}, {1, 2, 0, 6, 4, 2}, {1, 3, 1, 7, 5, 2}},
    {{1, 1, 0, 3, 2, 0}, {1, 5, 4, 7, 6, 0}, {0, 1, 5, 0, 4, 1}, {0, 3, 7, 2, 6, 1}}};
constexpr int kEdgeEdges[3][2][5] = {{{3, 2, 1, 0, 0}, {7, 6, 5, 4, 0}},
                                     {{5, 1, 4, 0, 1}, {7, 3, 6, 2, 1}},
                                     {{6, 4, 2, 0, 2}, {7, 5, 3, 1, 2}}};
constexpr int kProcessEdge[3][4] = {{3, 2, 1, 0}, {7, 5, 6, 4}, {11, 10, 9, 8}};
constexpr int kFaceEdgeOrder[2][4] = {{0, 0, 1, 1}, {0, 1, 0, 1}};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// Values that other waves changed with atomics (which act in L2) are read past the
// CU's L1 with an agent-scope atomic load.
template <class T>
__device__ inline T load_l2(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint64_t morton(int x, int y, int z, int levels) {
    uint64_t c = 0;
    for (int k = 0; k < levels; ++k)
        c |= (uint64_t)(((x >> k) & 1) << 2 | ((y >> k) & 1) << 1 | ((z >> k) & 1)) << (3 * k);
    return c;
}
__device__ inline void unmorton(uint64_t c, int levels, int* xyz) {
    xyz[0] = xyz[1] = xyz[2] = 0;
    for (int k = 0; k < levels; ++k) {
        const int b = (int)((c >> (3 * k)) & 7);
        xyz[0] |= ((b >> 2) & 1) << k;
        xyz[1] |= ((b >> 1) & 1) << k;
        xyz[2] |= (b & 1) << k;
    }
}
__device__ inline int hash_home(uint64_t key, int mask) {
    return (int)((key * 0x9E3779B97F4A7C15ull) >> 40) & mask;
}
// Insert `key`; returns its slot (-1 when the table is full), *fresh = inserted now.
__device__ inline int hash_insert(uint64_t* keys, int mask, uint64_t key, bool* fresh) {
    int s = hash_home(key, mask);
    for (int probe = 0; probe <= mask; ++probe) {
        const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(keys + s), kEmptyKey, key);
        if (prev == kEmptyKey || prev == key) {
            *fresh = prev == kEmptyKey;
            return s;
        }
        s = (s + 1) & mask;
    }
    return -1;
}
__device__ inline int hash_find(const uint64_t* keys, int mask, uint64_t key) {
    int s = hash_home(key, mask);
    for (int probe = 0; probe <= mask; ++probe) {
        const uint64_t k = load_l2(keys + s);
        if (k == key) return s;
        if (k == kEmptyKey) return -1;
        s = (s + 1) & mask;
    }
    return -1;
}
__device__ inline bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }


// FilterSeamNode, src/clipmap.cpp:508-542.
__device__ inline bool filter_seam_node(int neighbour, const int* b, const int* mn, const int* mx) {
    switch (neighbour) {
    case 0: return mx[0] == b[0] || mx[1] == b[1] || mx[2] == b[2];
    case 1: return mn[2] == b[2];
    case 2: return mn[1] == b[1];
    case 3: return mn[1] == b[1] || mn[2] == b[2];
    case 4: return mn[0] == b[0];
    case 5: return mn[0] == b[0] || mn[2] == b[2];
    case 6: return mn[0] == b[0] || mn[1] == b[1];
    case 7: return mn[0] == b[0] && mn[1] == b[1] && mn[2] == b[2];
    }
    return false;
}

__device__ inline bool same_chunk(uint64_t a, uint64_t b) { return ((a ^ b) & kShapeChunkMask) == 0; }

__device__ inline void set_status(int* s, int code) { atomicCAS(s, 0, code); }

// SelectSeamNodes' test for one offered node (src/clipmap.cpp:552-567) with the node's
// leaf min (src/clipmap.cpp:402): FilterSeamNode and the host's seam AABB (min, 2 * size).
__device__ inline bool select_leaf(const SeamJobDev& J, const SeamSrcDev& src, const lvn_seam_node& sn, int* mn) {
    const int leaf_size = src.size / J.vpc;
    const int filter_size = src.size / (J.vpc * 4);
    const int root_size = 2 * J.host_size;
    const int bounds[3] = {J.host_min[0] + J.host_size, J.host_min[1] + J.host_size, J.host_min[2] + J.host_size};
    int mx[3];
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
        mn[a] = sn.local_min_and_material[a] * leaf_size + src.min[a];
        mx[a] = mn[a] + filter_size * 4;
        inside &= mn[a] >= J.host_min[a] && mn[a] < J.host_min[a] + root_size;
    }
    return inside && filter_seam_node(src.neighbour, bounds, mn, mx);
}

__device__ inline int wave_inclusive(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Exclusive workgroup scan of one value per thread (wave scans, then a scan of the
// wave totals in `sw`); *total = sum.
__device__ inline int block_exclusive(int v, int* sw, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int incl = wave_inclusive(v);
    if (lane == 63) sw[w] = incl;
    __syncthreads();
    if (w == 0) {
        int x = lane < kSeamWaves ? sw[lane] : 0;
        x = wave_inclusive(x);
        if (lane < kSeamWaves) sw[lane] = x;
    }
    __syncthreads();
    *total = sw[kSeamWaves - 1];
    return (w ? sw[w - 1] : 0) + incl - v;
}

// The calls one pending call makes, in the reference's call order (write = false only
// counts them).  Calls that would return at once are not made: a face with a missing
// node, both nodes in one chunk or no branch; an edge with a missing node or all four
// nodes in one chunk.  An edge of four leaves is resolved on the spot
// (ContourProcessEdge) into a quad or nothing.  Directions are template arguments so
// that every table index is a constant.
struct Expander {
    const int* C;        // child links, 8 per node
    const uint64_t* H;   // node shapes
    Call* dst;
    int at;
    int k;
    int more;
    bool write;

    __device__ __forceinline__ void push(int a, int b, int c, int d, int meta) {
        if (write) {
            dst[at + k] = make_call(a, b, c, d, meta);
            if ((meta & 3) != kCallQuad) more = 1;
        }
        ++k;
    }
    __device__ bool internal(int i) const { return (H[i] >> 42) & 1; }
    // A cell call on a node inside one chunk makes only calls on nodes of that chunk, which
    // the chunk test returns from: it emits nothing and is not made.
    __device__ bool cell_emits(int i) const { return ((H[i] >> 42) & 1) && !((H[i] >> 56) & 1); }
    __device__ int child(int i, int k) const { return C[8 * i + k]; }

    __device__ __forceinline__ void face(int a, int b, int dir) {
        if (a < 0 || b < 0) return;
        if (same_chunk(H[a], H[b])) return;
        if (!internal(a) && !internal(b)) return;
        push(a, b, -1, -1, kCallFace | dir << 2);
    }
    template <int D>
    __device__ __forceinline__ void edge_d(int n0, int n1, int n2, int n3) {
        if (n0 < 0 || n1 < 0 || n2 < 0 || n3 < 0) return;
        const uint64_t h[4] = {H[n0], H[n1], H[n2], H[n3]};
        if (same_chunk(h[0], h[1]) && same_chunk(h[0], h[2]) && same_chunk(h[0], h[3])) return;
        if (((h[0] | h[1] | h[2] | h[3]) >> 42) & 1) {
            push(n0, n1, n2, n3, kCallEdge | D << 2);
            return;
        }
        // ContourProcessEdge, src/octree.cpp:235-282
        int min_size = 0x7fffffff;
        bool flip = false, change = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = kProcessEdge[D][i];
            const int corners = (int)(h[i] >> 48) & 0xff;
            const int size = (int)(h[i] >> 43) & 31;  // log2: the order of sizes is kept
            const int m0 = (corners >> kEdgeCorners[e][0]) & 1;
            const int m1 = (corners >> kEdgeCorners[e][1]) & 1;
            if (size < min_size) {
                min_size = size;
                flip = m1 != 1;
                change = m0 != m1;
            }
        }
        if (change) push(n0, n1, n2, n3, kCallQuad | (flip ? 16 : 0));
    }
    __device__ __forceinline__ void edge(int n0, int n1, int n2, int n3, int dir) {
        if (dir == 0) edge_d<0>(n0, n1, n2, n3);
        else if (dir == 1) edge_d<1>(n0, n1, n2, n3);
        else edge_d<2>(n0, n1, n2, n3);
    }
    // ContourFaceProc, src/octree.cpp:342-418
    // Child rows are read whether or not the node is internal (a leaf's row is all -1),
    // so they are in flight together with the node shapes.
    template <int D>
    __device__ __forceinline__ void face_call(const Call& c) {
        int r0[8], r1[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            r0[k] = child(c.node(0), k);
            r1[k] = child(c.node(1), k);
        }
        const bool b0 = internal(c.node(0)), b1 = internal(c.node(1));
        int ch0[8], ch1[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ch0[k] = b0 ? r0[k] : c.node(0);
            ch1[k] = b1 ? r1[k] : c.node(1);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) face(ch0[kFaceFaces[D][i][0]], ch1[kFaceFaces[D][i][1]], kFaceFaces[D][i][2]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int* order = kFaceEdgeOrder[kFaceEdges[D][i][0]];
            const int s0 = order[0] ? ch1[kFaceEdges[D][i][1]] : ch0[kFaceEdges[D][i][1]];
            const int s1 = order[1] ? ch1[kFaceEdges[D][i][2]] : ch0[kFaceEdges[D][i][2]];
            const int s2 = order[2] ? ch1[kFaceEdges[D][i][3]] : ch0[kFaceEdges[D][i][3]];
            const int s3 = order[3] ? ch1[kFaceEdges[D][i][4]] : ch0[kFaceEdges[D][i][4]];
            edge(s0, s1, s2, s3, kFaceEdges[D][i][5]);
        }
    }
    // ContourEdgeProc, src/octree.cpp:286-338
    template <int D>
    __device__ __forceinline__ void edge_call(const Call& c) {
        int r[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) r[i][j] = child(c.node(j), kEdgeEdges[D][i][j]);
        bool br[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) br[j] = internal(c.node(j));
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            int sub[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) sub[j] = br[j] ? r[i][j] : c.node(j);
            edge(sub[0], sub[1], sub[2], sub[3], kEdgeEdges[D][i][4]);
        }
    }
    __device__ __forceinline__ void expand(const Call& c) {
        const int kind = c.meta() & 3, dir = (c.meta() >> 2) & 3;
        if (kind == kCallQuad) {
            push(c.node(0), c.node(1), c.node(2), c.node(3), c.meta());
        } else if (kind == kCallCell) {  // ContourCellProc, src/octree.cpp:422-474
            int ch[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) ch[i] = child(c.node(0), i);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (ch[i] >= 0 && cell_emits(ch[i])) push(ch[i], -1, -1, -1, kCallCell);
#pragma unroll
            for (int i = 0; i < 12; ++i) face(ch[kCellFaces[i][0]], ch[kCellFaces[i][1]], kCellFaces[i][2]);
#pragma unroll
            for (int i = 0; i < 6; ++i)
                edge(ch[kCellEdges[i][0]], ch[kCellEdges[i][1]], ch[kCellEdges[i][2]], ch[kCellEdges[i][3]],
                     kCellEdges[i][4]);
        } else if (kind == kCallFace) {
            if (dir == 0) face_call<0>(c);
            else if (dir == 1) face_call<1>(c);
            else face_call<2>(c);
        } else {
            if (dir == 0) edge_call<0>(c);
            else if (dir == 1) edge_call<1>(c);
            else edge_call<2>(c);
        }
    }
};

// ---------------------------------------------------------------------------
// The seam kernel: one workgroup per seam
// ---------------------------------------------------------------------------

struct SeamScratch {
    uint64_t* keys;
    uint64_t* pick;
    int* vals;
    int* leaf;
    SNode* nodes;
    int* ch;           // 8 child links per node
    uint64_t* shape;   // node shapes
    int* order;        // node indices bucketed by level
    Call* list0;       // pending calls, two rounds alternating
    Call* list1;
};

__host__ __device__ inline uint64_t align_up(uint64_t v) { return (v + 255) & ~255ull; }
__host__ __device__ inline uint64_t seam_scratch_bytes(int hash_cap, int node_cap, int list_cap) {
    return align_up(8ull * hash_cap) * 2 + align_up(4ull * hash_cap) * 2 + align_up(sizeof(SNode) * (uint64_t)node_cap) +
           align_up(32ull * node_cap) + align_up(8ull * node_cap) + align_up(4ull * node_cap) +
           2 * align_up(sizeof(Call) * (uint64_t)list_cap);
}
__device__ inline SeamScratch seam_scratch(char* base, const SeamJobDev& J) {
    SeamScratch s;
    char* p = base;
    s.keys = reinterpret_cast<uint64_t*>(p);
    p += align_up(8ull * J.hash_cap);
    s.pick = reinterpret_cast<uint64_t*>(p);
    p += align_up(8ull * J.hash_cap);
    s.vals = reinterpret_cast<int*>(p);
    p += align_up(4ull * J.hash_cap);
    s.leaf = reinterpret_cast<int*>(p);
    p += align_up(4ull * J.hash_cap);
    s.nodes = reinterpret_cast<SNode*>(p);
    p += align_up(sizeof(SNode) * (uint64_t)J.node_cap);
    s.ch = reinterpret_cast<int*>(p);
    p += align_up(32ull * J.node_cap);
    s.shape = reinterpret_cast<uint64_t*>(p);
    p += align_up(8ull * J.node_cap);
    s.order = reinterpret_cast<int*>(p);
    p += align_up(4ull * J.node_cap);
    s.list0 = reinterpret_cast<Call*>(p);
    p += align_up(sizeof(Call) * (uint64_t)J.list_cap);
    s.list1 = reinterpret_cast<Call*>(p);
    return s;
}

__global__ __launch_bounds__(kSeamThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_seam(const SeamJobDev* __restrict__ jobs,
                                                       const SeamSrcDev* __restrict__ srcs, char* scratch,
                                                       lvn_vertex* out_v, int* out_t, SeamResDev* res) {
    const SeamJobDev J = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    SeamScratch S = seam_scratch(scratch + J.scratch, J);
    SNode* N = S.nodes;
    const int mask = J.hash_cap - 1;
    const int root_size = 2 * J.host_size;
    const int root_levels = __builtin_ctz((unsigned)root_size);

    __shared__ int s_nodes, s_status, s_maxlev, s_n, s_more, s_carry;
    __shared__ int s_wave[kSeamWaves];
    __shared__ int s_lev[33], s_fill[32];
    if (tid == 0) {
        s_nodes = 0;
        s_status = 0;
        s_maxlev = 0;
    }
    for (int i = tid; i < J.hash_cap; i += kSeamThreads) {
        S.keys[i] = kEmptyKey;
        S.pick[i] = 0;
        S.vals[i] = -1;
        S.leaf[i] = 0;
    }
    __syncthreads();

    // 1. SelectSeamNodes (src/clipmap.cpp:546-569) over every offered list; the seam
    //    node -> leaf conversion of src/clipmap.cpp:398-412.
    for (int si = 0; si < J.src_count; ++si) {
        const SeamSrcDev src = srcs[J.src_begin + si];
        const int leaf_size = src.size / J.vpc;
        for (int j = tid; j < src.count; j += kSeamThreads) {
            int mn[3];
            if (!select_leaf(J, src, src.nodes[j], mn)) continue;
            bool aligned = pow2(leaf_size) && leaf_size < root_size;
            for (int a = 0; a < 3; ++a) aligned &= ((mn[a] - J.host_min[a]) & (leaf_size - 1)) == 0;
            if (!aligned) {
                set_status(&s_status, LVN_ERR_INVALID);
                continue;
            }
            const int level = root_levels - __builtin_ctz((unsigned)leaf_size);
            const uint64_t key = (uint64_t)level << 48 |
                                 morton((mn[0] - J.host_min[0]) / leaf_size, (mn[1] - J.host_min[1]) / leaf_size,
                                        (mn[2] - J.host_min[2]) / leaf_size, level);
            bool fresh;
            const int slot = hash_insert(S.keys, mask, key, &fresh);
            if (slot < 0) {
                set_status(&s_status, LVN_ERR_OVERFLOW);
                continue;
            }
            atomicMax(reinterpret_cast<unsigned long long*>(S.pick + slot),
                      (unsigned long long)((uint64_t)(si + 1) << 32 | (uint32_t)j));
            atomicMax(&s_maxlev, level);
        }
    }
    __syncthreads();

    // leaves: one node per distinct key, carrying the last offer
    for (int i = tid; i < J.hash_cap; i += kSeamThreads) {
        const uint64_t key = load_l2(S.keys + i);
        if (key == kEmptyKey) continue;
        const int idx = atomicAdd(&s_nodes, 1);
        if (idx >= J.node_cap) {
            set_status(&s_status, LVN_ERR_OVERFLOW);
            continue;
        }
        const uint64_t p = load_l2(S.pick + i);
        const SeamSrcDev src = srcs[J.src_begin + (int)(p >> 32) - 1];
        const lvn_seam_node& sn = src.nodes[(uint32_t)p];
        SNode n;
        n.size = src.size / J.vpc;
        for (int a = 0; a < 3; ++a) n.min[a] = sn.local_min_and_material[a] * n.size + src.min[a];
        n.parent = -1;
        n.level = (int)(key >> 48);
        n.mat = sn.local_min_and_material[3];
        n.count = 1;
        n.offset = 0;
        n.type = kNodeLeaf;
        n.code = key & ((1ull << 48) - 1);
        n.payload = p;
        N[idx] = n;
        for (int c = 0; c < 8; ++c) S.ch[8 * idx + c] = -1;
        S.shape[idx] = node_shape(n.min, J.host_min, n.size, false, n.mat);
        S.vals[i] = idx;
        S.leaf[i] = 1;
    }
    __syncthreads();
    const int nleaves = min(s_nodes, J.node_cap);

    // 2. ancestors (ConstructParents, src/octree.cpp:29-88): a leaf inserts its parent
    //    chain up to the first parent already present.  A parent key held by a leaf
    //    means two selected nodes overlap.
    if (s_status == 0) {
        for (int i = tid; i < nleaves; i += kSeamThreads) {
            const int level = N[i].level;
            const uint64_t code = N[i].code;
            for (int l = level - 1; l >= 0; --l) {
                const uint64_t c = code >> (3 * (level - l));
                bool fresh;
                const int slot = hash_insert(S.keys, mask, (uint64_t)l << 48 | c, &fresh);
                if (slot < 0) {
                    set_status(&s_status, LVN_ERR_OVERFLOW);
                    break;
                }
                if (!fresh) {
                    if (S.leaf[slot]) set_status(&s_status, LVN_ERR_INVALID);
                    break;
                }
                const int idx = atomicAdd(&s_nodes, 1);
                if (idx >= J.node_cap) {
                    set_status(&s_status, LVN_ERR_OVERFLOW);
                    break;
                }
                SNode n;
                n.size = root_size >> l;
                int xyz[3];
                unmorton(c, l, xyz);
                for (int a = 0; a < 3; ++a) n.min[a] = J.host_min[a] + xyz[a] * n.size;
                n.parent = -1;
                n.level = l;
                n.mat = 0;
                n.count = 0;
                n.offset = 0;
                n.type = kNodeInternal;
                n.code = c;
                n.payload = 0;
                N[idx] = n;
                for (int k = 0; k < 8; ++k) S.ch[8 * idx + k] = -1;
                S.shape[idx] = node_shape(n.min, J.host_min, n.size, true, 0);
                S.vals[slot] = idx;
            }
        }
    }
    __syncthreads();
    const int nnodes = min(s_nodes, J.node_cap);
    const bool ok = s_status == 0 && nleaves > 0;

    // 3. links, leaf counts, depth-first ranks
    if (ok) {
        for (int i = tid; i < nnodes; i += kSeamThreads) {
            if (N[i].level == 0) continue;
            const uint64_t c = N[i].code;
            const int slot = hash_find(S.keys, mask, (uint64_t)(N[i].level - 1) << 48 | (c >> 3));
            const int p = slot < 0 ? -1 : S.vals[slot];
            if (p < 0) {  // every parent was inserted: not reached
                set_status(&s_status, LVN_ERR_OVERFLOW);
                continue;
            }
            N[i].parent = p;
            S.ch[8 * p + (c & 7)] = i;
        }
        // bucket the nodes by level
        for (int i = tid; i < 33; i += kSeamThreads) s_lev[i] = 0;
    }
    __syncthreads();
    if (ok) {
        for (int i = tid; i < nnodes; i += kSeamThreads) atomicAdd(&s_lev[N[i].level + 1], 1);
    }
    __syncthreads();
    if (ok && tid == 0) {
        for (int l = 1; l < 33; ++l) s_lev[l] += s_lev[l - 1];  // s_lev[l] = first node of level l
        for (int l = 0; l < 32; ++l) s_fill[l] = s_lev[l];
    }
    __syncthreads();
    if (ok) {
        for (int i = tid; i < nnodes; i += kSeamThreads) S.order[atomicAdd(&s_fill[N[i].level], 1)] = i;
    }
    __syncthreads();
    // leaf counts, deepest level first: each node adds its count to its parent once
    if (ok && s_status == 0) {
        for (int l = s_maxlev; l >= 1; --l) {
            for (int k = s_lev[l] + tid; k < s_lev[l + 1]; k += kSeamThreads) {
                const int i = S.order[k];
                atomicAdd(&N[N[i].parent].count, load_l2(&N[i].count));
            }
            __syncthreads();
        }
    }
    __syncthreads();
    int root = -1;
    if (ok && s_status == 0) {
        const int slot = hash_find(S.keys, mask, 0ull);
        root = slot < 0 ? -1 : S.vals[slot];
    }
    const bool linked = ok && s_status == 0 && root >= 0;
    if (linked) {
        // depth-first ranks, top level first: children in order 0..7 after their parent's offset
        for (int l = 0; l < s_maxlev; ++l) {
            for (int k = s_lev[l] + tid; k < s_lev[l + 1]; k += kSeamThreads) {
                const int i = S.order[k];
                if (N[i].type != kNodeInternal) continue;
                int off = N[i].offset;
                for (int c = 0; c < 8; ++c) {
                    const int ch = S.ch[8 * i + c];
                    if (ch < 0) continue;
                    N[ch].offset = off;
                    off += load_l2(&N[ch].count);
                }
            }
            __syncthreads();
        }
        // GenerateVertexIndices (src/octree.cpp:196-231): vertex = rank of the leaf
        for (int i = tid; i < nleaves; i += kSeamThreads) {
            const uint64_t p = N[i].payload;
            const lvn_seam_node& sn = srcs[J.src_begin + (int)(p >> 32) - 1].nodes[(uint32_t)p];
            lvn_vertex v;
            v.pos[0] = sn.pos[0];
            v.pos[1] = sn.pos[1];
            v.pos[2] = sn.pos[2];
            v.pos[3] = 0.f;
            v.normal[0] = sn.normal[0];
            v.normal[1] = sn.normal[1];
            v.normal[2] = sn.normal[2];
            v.normal[3] = 0.f;
            v.colour[0] = J.colour[0];
            v.colour[1] = J.colour[1];
            v.colour[2] = J.colour[2];
            v.colour[3] = (float)(N[i].mat >> 8);
            out_v[J.v_off + N[i].offset] = v;
        }
    }

    // 4. contouring rounds
    if (tid == 0) {
        s_n = 0;
        if (linked && N[root].type == kNodeInternal && ((S.shape[root] >> 56) & 1) == 0) {
            S.list0[0] = make_call(root, -1, -1, -1, kCallCell);
            s_n = 1;
        }
    }
    __syncthreads();
    int cur = 0, rounds = 0;
    bool more = true;
    for (;;) {
        __syncthreads();
        if (s_status != 0 || s_n == 0 || !more) break;
        const int n = s_n;
        const Call* A = cur ? S.list1 : S.list0;  // (no dynamic index into a private array)
        Call* B = cur ? S.list0 : S.list1;
        __syncthreads();  // every thread has read s_n before it changes
        if (tid == 0) {
            s_carry = 0;
            s_more = 0;
        }
        __syncthreads();
        for (int t0 = 0; t0 < n; t0 += kSeamThreads * kCallsPerThread) {
            // thread t expands calls i0 .. i0 + kCallsPerThread - 1, in order
            const int i0 = t0 + tid * kCallsPerThread;
            int cnt = 0;
            for (int k = 0; k < kCallsPerThread && i0 + k < n; ++k) {
                Expander e{S.ch, S.shape, B, 0, 0, 0, false};
                e.expand(A[i0 + k]);
                cnt += e.k;
            }
            int total;
            const int excl = block_exclusive(cnt, s_wave, &total);
            const int base = s_carry;
            if (cnt > 0 && base + excl + cnt <= J.list_cap) {
                int at = base + excl;
                bool m = false;
                for (int k = 0; k < kCallsPerThread && i0 + k < n; ++k) {
                    Expander w{S.ch, S.shape, B, at, 0, 0, true};
                    w.expand(A[i0 + k]);
                    at += w.k;
                    m |= w.more != 0;
                }
                if (m) s_more = 1;
            }
            __syncthreads();
            if (tid == 0) s_carry = base + total;
            __syncthreads();
        }
        if (tid == 0) {
            if (s_carry > J.list_cap || rounds + 1 >= kMaxRounds) set_status(&s_status, LVN_ERR_OVERFLOW);
            s_n = s_carry;
        }
        __syncthreads();
        more = s_more != 0;
        cur ^= 1;
        ++rounds;
    }

    // triangles of the quads, in list order (ContourProcessEdge's two triangles)
    const int nq = s_status == 0 && linked ? s_n : 0;
    const Call* Q = cur ? S.list1 : S.list0;
    for (int i = tid; i < nq; i += kSeamThreads) {
        const Call c = Q[i];
        int v[4];
        for (int j = 0; j < 4; ++j) v[j] = N[c.node(j)].offset;
        int* t = out_t + 3 * (J.t_off + 2 * (uint64_t)i);
        if (c.meta() & 16) {
            t[0] = v[0]; t[1] = v[3]; t[2] = v[1];
            t[3] = v[0]; t[4] = v[2]; t[5] = v[3];
        } else {
            t[0] = v[0]; t[1] = v[1]; t[2] = v[3];
            t[3] = v[0]; t[4] = v[3]; t[5] = v[2];
        }
    }
    if (tid == 0) {
        SeamResDev r;
        r.status = s_status;
        r.nv = s_status == 0 && linked ? nleaves : 0;
        r.nt = 2 * nq;
        r.rounds = rounds;
        res[blockIdx.x] = r;
    }
}

// Selected seam nodes of each seam, to size its scratch: workgroup (s, b) of gridDim.y per
// seam counts every gridDim.y-th 256-node slice of seam s's lists into counts[s] (zeroed
// by the caller), so one big seam is counted by many CUs.
__global__ __launch_bounds__(256) void k_seam_count(const SeamJobDev* __restrict__ jobs,
                                                    const SeamSrcDev* __restrict__ srcs, int* counts) {
    const SeamJobDev J = jobs[blockIdx.x];
    __shared__ int s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    int mine = 0;
    const int stride = gridDim.y * blockDim.x;
    for (int si = 0; si < J.src_count; ++si) {
        const SeamSrcDev src = srcs[J.src_begin + si];
        for (int j = blockIdx.y * blockDim.x + threadIdx.x; j < src.count; j += stride) {
            int mn[3];
            mine += select_leaf(J, src, src.nodes[j], mn) ? 1 : 0;
        }
    }
    atomicAdd(&s_count, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(&counts[blockIdx.x], s_count);
}

// Exclusive offsets of every seam's vertices and triangles in the packed output
// (one workgroup); offs[2n], offs[2n+1] = totals.
__global__ __launch_bounds__(1024) void k_seam_offsets(const SeamResDev* res, int n, uint64_t* offs) {
    __shared__ unsigned long long bv[1024], bt[1024];
    __shared__ unsigned long long cv, ct;
    if (threadIdx.x == 0) cv = ct = 0;
    __syncthreads();
    for (int b0 = 0; b0 < n; b0 += 1024) {
        const int i = b0 + threadIdx.x;
        const unsigned long long v = i < n ? res[i].nv : 0, t = i < n ? res[i].nt : 0;
        bv[threadIdx.x] = v;
        bt[threadIdx.x] = t;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const unsigned long long uv = threadIdx.x >= off ? bv[threadIdx.x - off] : 0;
            const unsigned long long ut = threadIdx.x >= off ? bt[threadIdx.x - off] : 0;
            __syncthreads();
            bv[threadIdx.x] += uv;
            bt[threadIdx.x] += ut;
            __syncthreads();
        }
        if (i < n) {
            offs[2 * i] = cv + bv[threadIdx.x] - v;
            offs[2 * i + 1] = ct + bt[threadIdx.x] - t;
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            cv += bv[1023];
            ct += bt[1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        offs[2 * n] = cv;
        offs[2 * n + 1] = ct;
    }
}

// Copies each seam's vertices and triangles into the packed output (gridDim.y workgroups
// per seam).
__global__ __launch_bounds__(256) void k_seam_pack(const SeamJobDev* jobs, const SeamResDev* res,
                                                   const uint64_t* offs, const lvn_vertex* out_v, const int* out_t,
                                                   lvn_vertex* pv, int* pt) {
    const SeamJobDev& J = jobs[blockIdx.x];
    const SeamResDev r = res[blockIdx.x];
    const uint64_t ov = offs[2 * blockIdx.x], ot = offs[2 * blockIdx.x + 1];
    const int first = blockIdx.y * blockDim.x + threadIdx.x, stride = gridDim.y * blockDim.x;
    for (int i = first; i < r.nv; i += stride) pv[ov + i] = out_v[J.v_off + i];
    for (int i = first; i < 3 * r.nt; i += stride) pt[3 * ot + i] = out_t[3 * J.t_off + i];
}

// Workgroups per seam for the count and pack kernels: one CU's worth is too little for a
// request of few, large seams.
inline int per_seam_blocks(int m) { return std::max(1, std::min(256, 4096 / std::max(m, 1))); }

// ---------------------------------------------------------------------------
// The seam pass as grid-wide steps over all seams of a launch
// ---------------------------------------------------------------------------
//
// The steps of one seam above, each run as one launch over the elements of every seam.
// Work is cut into tiles of 1024 elements that belong to one seam (offered nodes of one
// list, hash slots, leaves, nodes, the internal nodes of one level, pending calls); a tile list
// is made on the device from each seam's counts, so lists whose length only the device
// knows (a level's nodes, a round's calls) need no read-back.  A seam's counts, level
// starts, call-list state and status live in its SeamRun record.  Counters shared by a
// tile's threads are reserved once per tile (LDS histograms, workgroup scans) or once per
// wave, so the per-seam atomics stay few.

constexpr int kSgTile = 1024;
constexpr int kSgThreads = 256;
constexpr int kSgPer = kSgTile / kSgThreads;
constexpr int kSgGrid = 2048;  // workgroups of the tile launches
constexpr int kSgStage = 3072;  // calls a round tile stages in LDS (48 KB)
constexpr int kSgLevels = 33;

struct SeamRun {
    int nodes, nleaves, nnodes, status;
    int root, linked, rounds, stage;  // stage: calls reserved in the staging list this round
    int n, more;              // pending calls; 1 = calls still expand (2 = one was written this round)
    int tbase[8], ntiles;     // this seam's tiles in the current round's list (first position per sub-list)
    int lev[kSgLevels + 1];   // internal nodes per level (lev[l + 1]), then level starts
    int fill[kSgLevels];      // placement cursors per level
};

// Tile list sizes and counters of one grid pass, in one int array.
enum : int {
    kSgCntHash = 0,
    kSgCntOffered = 1,
    kSgCntLeaf = 2,
    kSgCntNode = 3,
    kSgCntFinal = 4,
    kSgCntLevel = 8,                          // + level
    kSgCntRound = kSgCntLevel + kSgLevels,        // + round: tiles of the round's list
    kSgCntActive = kSgCntRound + kMaxRounds + 1,  // + round: seams whose calls still expand
    kSgCounters = kSgCntActive + kMaxRounds,
};

struct SeamGrid {
    const SeamJobDev* jobs;
    const SeamSrcDev* srcs;
    char* scratch;
    SeamRun* run;
    int* cnt;
    lvn_vertex* out_v;
    int* out_t;
    int m;
    int spread;  // tiles of each seam dealt over every XCD (requests of few seams)
};

__device__ __forceinline__ SeamScratch sg_scratch(const SeamGrid& g, int s) {
    return seam_scratch(g.scratch + g.jobs[s].scratch, g.jobs[s]);
}

// The grid pass's nodes, in the node region of a seam's scratch as three dense arrays:
// level << 48 | Morton code; (leaf count, depth-first leaf rank); for leaves the offer,
// (source + 1) << 32 | index.  Leaves are nodes 0 .. nleaves - 1, internal nodes follow.
struct SgNodes {
    uint64_t* key;
    int2* co;
    uint64_t* offer;
};
static_assert(sizeof(SNode) >= 24, "the grid pass's node arrays fit the node region");
__device__ __forceinline__ SgNodes sg_nodes(const SeamScratch& S, int node_cap) {
    SgNodes q;
    q.key = reinterpret_cast<uint64_t*>(S.nodes);
    q.co = reinterpret_cast<int2*>(q.key + node_cap);
    q.offer = reinterpret_cast<uint64_t*>(q.co + node_cap);
    return q;
}
constexpr uint64_t kCodeMask = (1ull << 48) - 1;

// Loop over the tiles of a list: tl = (seam, first element, source, 0); padding tiles
// (seam -1) are skipped by the whole workgroup.  gridDim.x is a multiple of 8, so a
// workgroup keeps to the tiles of its XCD.
#define LVN_SG_TILES(tiles, ntiles)                                \
    for (int tk = blockIdx.x; tk < (ntiles); tk += gridDim.x)      \
        if (const int4 tl = (tiles)[tk]; tl.x >= 0)

// Exclusive scan over the workgroup; *total = the sum.
__device__ __forceinline__ int sg_scan(int v, int* total) {
    __shared__ int wsum[kSgThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = wave_inclusive(v);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kSgThreads / 64; ++i) {
        base += i < w ? wsum[i] : 0;
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// `cnt` consecutive indices of *counter for each thread, reserved with one atomic per
// workgroup; returns the thread's first.
__device__ __forceinline__ int sg_reserve(int cnt, int* counter) {
    __shared__ int base;
    int total;
    const int excl = sg_scan(cnt, &total);
    if (threadIdx.x == 0) base = total ? atomicAdd(counter, total) : 0;
    __syncthreads();
    const int b = base;
    __syncthreads();
    return b + excl;
}

__device__ __forceinline__ int lane_rank(uint64_t mask) {
    return __popcll(mask & ((1ull << (threadIdx.x & 63)) - 1));
}

// Tile lists, placed by XCD.  Workgroups go to the eight XCDs in turn (workgroup b to XCD
// b mod 8), and every tile launch loops over its list with a stride that is a multiple of
// 8, so tile k runs on XCD k mod 8.  A list is eight interleaved sub-lists, one per XCD
// (sub-list x at k = 8p + x).  Local placement puts all of seam s's tiles in sub-list
// s mod 8: its hash table, nodes and call lists are then touched by one XCD and share its
// L2 (seams come largest first, so the sub-lists are about equally long).  A request of few
// seams would leave XCDs idle that way, so it spreads them: tile t of seam s goes to
// sub-list (s + t) mod 8.  Shorter sub-lists are padded with empty tiles (seam -1).  A
// seam's tiles keep their order: its tiles in one sub-list are consecutive, from base[x].
constexpr int kXcds = 8;
constexpr int kSpreadBelow = 64;  // requests of fewer seams spread each seam over every XCD

// Reserves seam s's cnt tiles in the sub-lists (fill = their lengths so far, in LDS):
// base[x] = the seam's first position in sub-list x.
__device__ __forceinline__ void sg_place(bool spread, int s, int cnt, int* fill, int* base) {
    if (!spread) {
        base[s & (kXcds - 1)] = cnt ? atomicAdd(&fill[s & (kXcds - 1)], cnt) : 0;
        return;
    }
#pragma unroll
    for (int x = 0; x < kXcds; ++x) {
        const int nx = cnt / kXcds + (((x - s) & (kXcds - 1)) < (cnt & (kXcds - 1)) ? 1 : 0);
        base[x] = nx ? atomicAdd(&fill[x], nx) : 0;
    }
}
// List index of tile t of seam s.
__device__ __forceinline__ int sg_tile_index(bool spread, int s, int t, const int* base) {
    if (!spread) return (base[s & (kXcds - 1)] + t) * kXcds + (s & (kXcds - 1));
    const int x = (s + t) & (kXcds - 1);
    return (base[x] + t / kXcds) * kXcds + x;
}

// Pads the sub-lists of `out` to the longest (fill[x] = length of sub-list x); returns the
// list's length.  Called by every thread of the workgroup.
__device__ int sg_pad(int4* out, const int* fill) {
    int longest = 0;
#pragma unroll
    for (int x = 0; x < kXcds; ++x) longest = max(longest, fill[x]);
    for (int k = threadIdx.x; k < kXcds * longest; k += blockDim.x)
        if (k / kXcds >= fill[k % kXcds]) out[k] = make_int4(-1, 0, 0, 0);
    return kXcds * longest;
}

// Tile lists (one workgroup of 1024 threads, a thread per seam; workgroup b of a level
// list makes level b's list at tiles + b * stride).  what: 0 a level's nodes, 1 a round's
// pending calls of the seams still expanding, 2 the final quads, 3 hash slots, 4 leaves,
// 5 nodes, 6 offered nodes (one list per source).
__global__ __launch_bounds__(1024) void k_sg_tiles(SeamGrid g, int what, int4* tiles, int stride, int* total) {
    __shared__ int fill[kXcds];
    const int l = blockIdx.x;
    int4* out = tiles + (size_t)l * stride;
    if (threadIdx.x < kXcds) fill[threadIdx.x] = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < g.m; s += blockDim.x) {
        const SeamRun& R = g.run[s];
        const SeamJobDev& J = g.jobs[s];
        auto nt = [](int n) { return (n + kSgTile - 1) / kSgTile; };
        int cnt = 0;
        switch (what) {
        case 0: cnt = R.status == 0 ? nt(R.lev[l + 1] - R.lev[l]) : 0; break;
        case 1: cnt = R.status == 0 && R.linked && R.more ? nt(R.n) : 0; break;
        case 2: cnt = R.status == 0 && R.linked ? nt(R.n) : 0; break;
        case 3: cnt = nt(J.hash_cap); break;
        case 4: cnt = nt(J.leaf_cap); break;
        case 5: cnt = nt(J.node_cap); break;
        default:
            for (int si = 0; si < J.src_count; ++si) cnt += nt(g.srcs[J.src_begin + si].count);
        }
        int base[kXcds];
        sg_place(g.spread, s, cnt, fill, base);
        if (what == 6) {
            int t = 0;
            for (int si = 0; si < J.src_count; ++si)
                for (int j = 0; j < g.srcs[J.src_begin + si].count; j += kSgTile)
                    out[sg_tile_index(g.spread, s, t++, base)] = make_int4(s, j, si, 0);
        } else {
            for (int t = 0; t < cnt; ++t) out[sg_tile_index(g.spread, s, t, base)] = make_int4(s, t * kSgTile, 0, 0);
            if (what == 1) {
                for (int x = 0; x < kXcds; ++x) g.run[s].tbase[x] = base[x];
                g.run[s].ntiles = cnt;
            }
        }
    }
    __syncthreads();
    const int n = sg_pad(out, fill);
    if (threadIdx.x == 0) total[l] = n;
}

// Seam records and round counters cleared; every hash slot emptied.
__global__ __launch_bounds__(kSgThreads) void k_sg_init(SeamGrid g, const int4* tiles) {
    for (int s = blockIdx.x * kSgThreads + threadIdx.x; s < g.m; s += gridDim.x * kSgThreads) {
        SeamRun& R = g.run[s];
        R.nodes = R.nleaves = R.nnodes = R.status = 0;
        R.root = -1;
        R.linked = R.rounds = R.stage = 0;
        R.n = R.more = 0;
        for (int x = 0; x < 8; ++x) R.tbase[x] = 0;
        R.ntiles = 0;
        for (int k = 0; k <= kSgLevels; ++k) R.lev[k] = 0;
        for (int k = 0; k < kSgLevels; ++k) R.fill[k] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kMaxRounds) g.cnt[kSgCntActive + threadIdx.x] = 0;
    LVN_SG_TILES(tiles, g.cnt[kSgCntHash]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamScratch S = sg_scratch(g, tl.x);
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            if (i >= J.hash_cap) continue;
            S.keys[i] = kEmptyKey;
            S.pick[i] = 0;
            S.vals[i] = -1;
            S.leaf[i] = 0;
        }
    }
}

// 1. SelectSeamNodes (src/clipmap.cpp:546-569) over the offered nodes of one list.
__global__ __launch_bounds__(kSgThreads) void k_sg_select(SeamGrid g, const int4* tiles) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntOffered]) {
        const int s = tl.x, si = tl.z;
        const SeamJobDev& J = g.jobs[s];
        SeamRun& R = g.run[s];
        SeamScratch S = sg_scratch(g, s);
        const SeamSrcDev src = g.srcs[J.src_begin + si];
        const int leaf_size = src.size / J.vpc;
        const int root_size = 2 * J.host_size;
        const int root_levels = __builtin_ctz((unsigned)root_size);
        const int mask = J.hash_cap - 1;
        for (int r = 0; r < kSgPer; ++r) {
            const int j = tl.y + threadIdx.x + r * kSgThreads;
            if (j >= src.count) continue;
            int mn[3];
            if (!select_leaf(J, src, src.nodes[j], mn)) continue;
            bool aligned = pow2(leaf_size) && leaf_size < root_size;
            for (int a = 0; a < 3; ++a) aligned &= ((mn[a] - J.host_min[a]) & (leaf_size - 1)) == 0;
            if (!aligned) {
                set_status(&R.status, LVN_ERR_INVALID);
                continue;
            }
            const int level = root_levels - __builtin_ctz((unsigned)leaf_size);
            const uint64_t key = (uint64_t)level << 48 |
                                 morton((mn[0] - J.host_min[0]) / leaf_size, (mn[1] - J.host_min[1]) / leaf_size,
                                        (mn[2] - J.host_min[2]) / leaf_size, level);
            bool fresh;
            const int slot = hash_insert(S.keys, mask, key, &fresh);
            if (slot < 0) {
                set_status(&R.status, LVN_ERR_OVERFLOW);
                continue;
            }
            atomicMax(reinterpret_cast<unsigned long long*>(S.pick + slot),
                      (unsigned long long)((uint64_t)(si + 1) << 32 | (uint32_t)j));
        }
    }
}

// Leaves: one node per distinct key, carrying the last offer.
__global__ __launch_bounds__(kSgThreads) void k_sg_leaves(SeamGrid g, const int4* tiles) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntHash]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, J.node_cap);
        const int i0 = tl.y + threadIdx.x * kSgPer;
        uint64_t key[kSgPer];
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < kSgPer; ++r) {
            key[r] = i0 + r < J.hash_cap ? S.keys[i0 + r] : kEmptyKey;
            cnt += key[r] != kEmptyKey;
        }
        int idx = sg_reserve(cnt, &R.nodes);
#pragma unroll
        for (int r = 0; r < kSgPer; ++r) {
            if (key[r] == kEmptyKey) continue;
            const int i = i0 + r;
            if (idx >= J.node_cap) {
                set_status(&R.status, LVN_ERR_OVERFLOW);
                continue;
            }
            const uint64_t p = S.pick[i];
            const SeamSrcDev src = g.srcs[J.src_begin + (int)(p >> 32) - 1];
            const lvn_seam_node& sn = src.nodes[(uint32_t)p];
            const int size = src.size / J.vpc;
            int mn[3];
            for (int a = 0; a < 3; ++a) mn[a] = sn.local_min_and_material[a] * size + src.min[a];
            Q.key[idx] = key[r];
            Q.co[idx] = make_int2(1, 0);
            Q.offer[idx] = p;
            for (int c = 0; c < 8; ++c) S.ch[8 * idx + c] = -1;
            S.shape[idx] = node_shape(mn, J.host_min, size, false, sn.local_min_and_material[3]);
            S.vals[i] = idx;
            S.leaf[i] = 1;
            ++idx;
        }
    }
}

// Per seam between steps.  0: the leaves are counted; 1: the ancestors too; 2: level
// counts become level starts and the root is found; 3: the contouring list starts with
// the root's cell call.
__global__ void k_sg_seam(SeamGrid g, int step) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.m) return;
    const SeamJobDev& J = g.jobs[s];
    SeamRun& R = g.run[s];
    if (step == 0) {
        R.nleaves = min(R.nodes, J.node_cap);
    } else if (step == 1) {
        R.nnodes = R.nleaves ? min(R.nodes, J.node_cap) : 0;
    } else if (step == 2) {
        for (int l = 1; l <= kSgLevels; ++l) R.lev[l] += R.lev[l - 1];
        for (int l = 0; l < kSgLevels; ++l) R.fill[l] = R.lev[l];
        if (R.status == 0 && R.nleaves > 0) {
            SeamScratch S = sg_scratch(g, s);
            const int slot = hash_find(S.keys, J.hash_cap - 1, 0ull);
            R.root = slot < 0 ? -1 : S.vals[slot];
        }
    } else {
        R.linked = R.status == 0 && R.nleaves > 0 && R.root >= 0;
        R.n = 0;
        if (R.linked) {
            SeamScratch S = sg_scratch(g, s);
            if (R.root >= R.nleaves && ((S.shape[R.root] >> 56) & 1) == 0) {  // an internal root
                S.list0[0] = make_call(R.root, -1, -1, -1, kCallCell);
                R.n = 1;
            }
        }
        R.more = R.n > 0;
    }
}

// 2. Ancestors (ConstructParents, src/octree.cpp:29-88): a leaf inserts its parent chain
// up to the first parent already present, a wave's lanes one level per step so that the
// wave reserves the new nodes with one atomic.  A parent key held by a leaf means two
// selected nodes overlap.
__global__ __launch_bounds__(kSgThreads) void k_sg_ancestors(SeamGrid g, const int4* tiles) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntLeaf]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        if (R.status != 0) continue;
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, J.node_cap);
        const int mask = J.hash_cap - 1;
        const int root_size = 2 * J.host_size;
        const int nleaves = R.nleaves;
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            const bool leaf = i < nleaves;
            const uint64_t lk = leaf ? Q.key[i] : 0;
            const int level = (int)(lk >> 48);
            const uint64_t code = lk & kCodeMask;
            int l = level - 1;
            bool act = leaf && l >= 0;
            while (__ballot(act)) {
                bool fresh = false;
                int slot = -1;
                uint64_t c = 0;
                if (act) {
                    c = code >> (3 * (level - l));
                    slot = hash_insert(S.keys, mask, (uint64_t)l << 48 | c, &fresh);
                    if (slot < 0) {
                        set_status(&R.status, LVN_ERR_OVERFLOW);
                        act = false;
                    } else if (!fresh) {
                        if (load_l2(S.leaf + slot)) set_status(&R.status, LVN_ERR_INVALID);
                        act = false;
                    }
                }
                const uint64_t fm = __ballot(fresh);
                if (!fm) continue;
                const int leader = __ffsll((unsigned long long)fm) - 1;
                int base = 0;
                if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(&R.nodes, __popcll(fm));
                base = __shfl(base, leader, 64);
                if (!fresh) continue;
                const int idx = base + lane_rank(fm);
                if (idx >= J.node_cap) {
                    set_status(&R.status, LVN_ERR_OVERFLOW);
                    act = false;
                    continue;
                }
                const int size = root_size >> l;
                int xyz[3], mn[3];
                unmorton(c, l, xyz);
                for (int a = 0; a < 3; ++a) mn[a] = J.host_min[a] + xyz[a] * size;
                Q.key[idx] = (uint64_t)l << 48 | c;
                Q.co[idx] = make_int2(0, 0);
                for (int q = 0; q < 8; ++q) S.ch[8 * idx + q] = -1;
                S.shape[idx] = node_shape(mn, J.host_min, size, true, 0);
                S.vals[slot] = idx;
                --l;
                act = l >= 0;
            }
        }
    }
}

// 3. Links to parents, and the nodes of each level counted (an LDS histogram per tile).
__global__ __launch_bounds__(kSgThreads) void k_sg_links(SeamGrid g, const int4* tiles) {
    __shared__ int hist[kSgLevels];
    __shared__ int skip;
    LVN_SG_TILES(tiles, g.cnt[kSgCntNode]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        // the status can change while the tile runs (set by other tiles): one thread
        // decides for the workgroup, whose barriers follow
        if (threadIdx.x == 0) skip = load_l2(&R.status) != 0 || tl.y >= R.nnodes;
        if (threadIdx.x < kSgLevels) hist[threadIdx.x] = 0;
        __syncthreads();
        if (skip) {
            __syncthreads();
            continue;
        }
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, J.node_cap);
        const int nnodes = R.nnodes;
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            if (i >= nnodes) continue;
            const uint64_t k = Q.key[i];
            const int level = (int)(k >> 48);
            if (i >= R.nleaves) atomicAdd(&hist[level], 1);  // internal nodes (created after the leaves)
            if (level == 0) continue;
            const uint64_t c = k & kCodeMask;
            const int slot = hash_find(S.keys, J.hash_cap - 1, (uint64_t)(level - 1) << 48 | (c >> 3));
            const int p = slot < 0 ? -1 : S.vals[slot];
            if (p < 0) {  // every parent was inserted: not reached
                set_status(&R.status, LVN_ERR_OVERFLOW);
                continue;
            }
            S.ch[8 * p + (c & 7)] = i;
        }
        __syncthreads();
        if (threadIdx.x < kSgLevels && hist[threadIdx.x]) atomicAdd(&R.lev[threadIdx.x + 1], hist[threadIdx.x]);
        __syncthreads();
    }
}

// Internal nodes (indices nleaves .. nnodes - 1) bucketed by level: a tile reserves its
// share of each level once.
__global__ __launch_bounds__(kSgThreads) void k_sg_order(SeamGrid g, const int4* tiles) {
    __shared__ int hist[kSgLevels], base[kSgLevels];
    LVN_SG_TILES(tiles, g.cnt[kSgCntNode]) {
        SeamRun& R = g.run[tl.x];
        if (R.status != 0 || tl.y >= R.nnodes || tl.y + kSgTile <= R.nleaves) continue;
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, g.jobs[tl.x].node_cap);
        if (threadIdx.x < kSgLevels) hist[threadIdx.x] = 0;
        __syncthreads();
        const int nnodes = R.nnodes, nleaves = R.nleaves;
        int lv[kSgPer];
#pragma unroll
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            lv[r] = i >= nleaves && i < nnodes ? (int)(Q.key[i] >> 48) : -1;
            if (lv[r] >= 0) atomicAdd(&hist[lv[r]], 1);
        }
        __syncthreads();
        if (threadIdx.x < kSgLevels) {
            base[threadIdx.x] = hist[threadIdx.x] ? atomicAdd(&R.fill[threadIdx.x], hist[threadIdx.x]) : 0;
            hist[threadIdx.x] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSgPer; ++r)
            if (lv[r] >= 0) S.order[base[lv[r]] + atomicAdd(&hist[lv[r]], 1)] = tl.y + threadIdx.x + r * kSgThreads;
        __syncthreads();
    }
}

// One level's internal nodes: leaf counts gathered from the children (a leaf counts 1;
// deepest level first), or depth-first ranks handed to the children in child order (top
// level first).
__global__ __launch_bounds__(kSgThreads) void k_sg_level(SeamGrid g, const int4* tiles, int l, int ranks) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntLevel + l]) {
        SeamRun& R = g.run[tl.x];
        if (ranks && !R.linked) continue;
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, g.jobs[tl.x].node_cap);
        const int first = R.lev[l], end = R.lev[l + 1];
        for (int r = 0; r < kSgPer; ++r) {
            const int k = first + tl.y + threadIdx.x + r * kSgThreads;
            if (k >= end) continue;
            const int i = S.order[k];
            int ch[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) ch[c] = S.ch[8 * i + c];
            int cnt[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) cnt[c] = ch[c] >= 0 ? Q.co[ch[c]].x : 0;
            if (!ranks) {
                int sum = 0;
#pragma unroll
                for (int c = 0; c < 8; ++c) sum += cnt[c];
                Q.co[i].x = sum;
            } else {
                int off = Q.co[i].y;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    if (ch[c] < 0) continue;
                    Q.co[ch[c]].y = off;
                    off += cnt[c];
                }
            }
        }
    }
}

// GenerateVertexIndices (src/octree.cpp:196-231): vertex = rank of the leaf.
__global__ __launch_bounds__(kSgThreads) void k_sg_vertices(SeamGrid g, const int4* tiles) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntLeaf]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        if (!R.linked) continue;
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, J.node_cap);
        const int nleaves = R.nleaves;
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            if (i >= nleaves) continue;
            const uint64_t p = Q.offer[i];
            const lvn_seam_node& sn = g.srcs[J.src_begin + (int)(p >> 32) - 1].nodes[(uint32_t)p];
            lvn_vertex v;
            v.pos[0] = sn.pos[0];
            v.pos[1] = sn.pos[1];
            v.pos[2] = sn.pos[2];
            v.pos[3] = 0.f;
            v.normal[0] = sn.normal[0];
            v.normal[1] = sn.normal[1];
            v.normal[2] = sn.normal[2];
            v.normal[3] = 0.f;
            v.colour[0] = J.colour[0];
            v.colour[1] = J.colour[1];
            v.colour[2] = J.colour[2];
            v.colour[3] = (float)(sn.local_min_and_material[3] >> 8);
            g.out_v[J.v_off + Q.co[i].y] = v;
        }
    }
}

// 4. One contouring round: every pending call replaced by the calls it makes.  Thread t
// of a tile expands calls t * 4 .. t * 4 + 3 of it: it counts their calls, the workgroup
// scans the counts, reserves the tile's block of the staging list (list1) with one
// atomic and writes the calls at once, their nodes still in cache.  The blocks are put
// in call order by k_sg_round_copy after k_sg_round_next has scanned each seam's tile
// counts in tile order.
__global__ __launch_bounds__(kSgThreads) void k_sg_round(SeamGrid g, const int4* tiles, int round, int* tsum,
                                                          int* tbase) {
    __shared__ int base;
    __shared__ Call stage[kSgStage];
    LVN_SG_TILES(tiles, g.cnt[kSgCntRound + round]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        SeamScratch S = sg_scratch(g, tl.x);
        const Call* A = S.list0;
        const int n = R.n;
        const int i0 = tl.y + threadIdx.x * kSgPer;
        int cnt = 0;
        for (int q = 0; q < kSgPer && i0 + q < n; ++q) {
            Expander e{S.ch, S.shape, S.list1, 0, 0, 0, false};
            e.expand(A[i0 + q]);
            cnt += e.k;
        }
        int total;
        const int at = sg_scan(cnt, &total);
        if (threadIdx.x == 0) {
            base = total ? atomicAdd(&R.stage, total) : 0;
            tsum[tk] = total;
            tbase[tk] = base;
        }
        __syncthreads();
        const int b0 = base;
        __syncthreads();  // base is rewritten by the next tile
        if (total == 0 || b0 + total > J.list_cap) continue;  // (the seam's total then exceeds its cap)
        // the tile's block is written from LDS in whole lines when it fits there
        const bool in_lds = total <= kSgStage;
        Call* dst = in_lds ? stage : S.list1 + b0;
        int w_at = at;
        bool m = false;
        for (int q = 0; q < kSgPer && i0 + q < n; ++q) {
            Expander w{S.ch, S.shape, dst, w_at, 0, 0, true};
            w.expand(A[i0 + q]);
            w_at += w.k;
            m |= w.more != 0;
        }
        if (m) R.more = 2;
        if (in_lds) {
            __syncthreads();
            Call* out = S.list1 + b0;
            for (int i = threadIdx.x; i < total; i += kSgThreads) out[i] = stage[i];
            __syncthreads();
        }
    }
}

// After a round, per seam (one workgroup of 1024 threads, a thread per seam): the tile
// counts of the round scanned in tile order into the blocks' places in the next list, the
// seam's round closed (the next list's length; the seam stays in the rounds while a
// written call still expands, counted in the round's active counter; an overflow of the
// list or of the round limit ends it), and the next round's tile list made.
__global__ __launch_bounds__(1024) void k_sg_round_next(SeamGrid g, int round, const int* tsum, int* toff,
                                                         int4* next) {
    __shared__ int fill[kXcds];
    if (threadIdx.x < kXcds) fill[threadIdx.x] = 0;
    __syncthreads();
    int active = 0;
    for (int s = threadIdx.x; s < g.m; s += blockDim.x) {
        SeamRun& R = g.run[s];
        const SeamJobDev& J = g.jobs[s];
        if (R.status == 0 && R.linked && R.more && R.ntiles > 0) {
            int sum = 0;
            for (int t = 0; t < R.ntiles; ++t) {
                const int k = sg_tile_index(g.spread, s, t, R.tbase);
                const int v = tsum[k];
                toff[k] = sum;
                sum += v;
            }
            R.rounds += 1;
            R.stage = 0;
            if (sum > J.list_cap || R.rounds >= kMaxRounds) {
                set_status(&R.status, LVN_ERR_OVERFLOW);
            } else {
                R.n = sum;
                R.more = R.more == 2;
                active += R.more;
            }
        }
        const int cnt = R.status == 0 && R.linked && R.more ? (R.n + kSgTile - 1) / kSgTile : 0;
        int base[kXcds];
        sg_place(g.spread, s, cnt, fill, base);
        for (int t = 0; t < cnt; ++t) next[sg_tile_index(g.spread, s, t, base)] = make_int4(s, t * kSgTile, 0, 0);
        for (int x = 0; x < kXcds; ++x) R.tbase[x] = base[x];
        R.ntiles = cnt;
    }
    __syncthreads();
    const int n = sg_pad(next, fill);
    if (threadIdx.x == 0) g.cnt[kSgCntRound + round + 1] = n;
    if (active) atomicAdd(&g.cnt[kSgCntActive + round], active);
}

// The round's blocks copied from the staging list into the seam's list in call order.
__global__ __launch_bounds__(kSgThreads) void k_sg_round_copy(SeamGrid g, const int4* tiles, int round,
                                                               const int* tsum, const int* tbase, const int* toff) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntRound + round]) {
        if (g.run[tl.x].status != 0) continue;
        SeamScratch S = sg_scratch(g, tl.x);
        const int n = tsum[tk];
        const Call* src = S.list1 + tbase[tk];
        Call* dst = S.list0 + toff[tk];
        for (int i = threadIdx.x; i < n; i += kSgThreads) dst[i] = src[i];
    }
}

// Triangles of the quads, in list order (ContourProcessEdge's two triangles), and each
// seam's result.
__global__ __launch_bounds__(kSgThreads) void k_sg_triangles(SeamGrid g, const int4* tiles, SeamResDev* res) {
    LVN_SG_TILES(tiles, g.cnt[kSgCntFinal]) {
        const SeamJobDev& J = g.jobs[tl.x];
        SeamRun& R = g.run[tl.x];
        SeamScratch S = sg_scratch(g, tl.x);
        const SgNodes Q = sg_nodes(S, J.node_cap);
        const Call* quads = S.list0;
        const int n = R.n;
        for (int r = 0; r < kSgPer; ++r) {
            const int i = tl.y + threadIdx.x + r * kSgThreads;
            if (i >= n) continue;
            const Call c = quads[i];
            int v[4];
            for (int j = 0; j < 4; ++j) v[j] = Q.co[c.node(j)].y;
            int* t = g.out_t + 3 * (J.t_off + 2 * (uint64_t)i);
            if (c.meta() & 16) {
                t[0] = v[0]; t[1] = v[3]; t[2] = v[1];
                t[3] = v[0]; t[4] = v[2]; t[5] = v[3];
            } else {
                t[0] = v[0]; t[1] = v[1]; t[2] = v[3];
                t[3] = v[0]; t[4] = v[3]; t[5] = v[2];
            }
        }
    }
    for (int s = blockIdx.x * kSgThreads + threadIdx.x; s < g.m; s += gridDim.x * kSgThreads) {
        const SeamRun& R = g.run[s];
        const bool ok = R.status == 0 && R.linked;
        SeamResDev r;
        r.status = R.status;
        r.nv = ok ? R.nleaves : 0;
        r.nt = ok ? 2 * R.n : 0;
        r.rounds = R.rounds;
        res[s] = r;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return LVN_OK;
        const size_t want = std::max(need, bytes + bytes / 2);  // grow by half at least
        if (p) LVN_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return LVN_ERR_NOMEM;
        }
        bytes = want;
        return LVN_OK;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return LVN_OK;
        const size_t want = std::max(need, bytes + bytes / 2);  // grow by half at least
        if (p) LVN_HIP(hipHostFree(p));
        p = nullptr;
        bytes = 0;
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return LVN_ERR_NOMEM;
        }
        bytes = want;
        return LVN_OK;
    }
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

// Where a seam's mesh of the last call lies in device memory: the unpacked output of
// the attempt that finished it.
struct SeamLoc {
    int attempt = -1;
    int nv = 0, nt = 0;
    uint64_t v_off = 0, t_off = 0;
};

// Seam-stitching buffers of a context, grown on demand and kept for reuse.  Each retry
// attempt writes its own output buffers, so every finished seam stays addressable.
struct SeamState {
    DevBuf in, jobs, srcs, scratch, res, offs, pv, pt, counts;
    DevBuf g_run, g_tiles, g_tsum, g_cnt;  // grid pass: seam records, tile lists, round scans, counters
    HostBuf h_in, h_jobs, h_srcs, h_res, h_pv, h_pt, h_counts, h_active;
    std::vector<std::unique_ptr<DevBuf>> out_v, out_t;  // per attempt
    std::vector<SeamLoc> loc;                            // per seam of the last call
    std::vector<hipEvent_t> ev;                          // ends of the download chunks
    ~SeamState() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

void destroy_seam_state(SeamState* s) { delete s; }

namespace {

bool host_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Runs f(lo, hi) over [0, n) split across host threads when `bytes` of copying justify
// it: the seam lists and meshes of a clipmap are gigabytes, more than one core copies
// at PCIe speed.
template <class F>
void parallel_ranges(int n, size_t bytes, F&& f) {
    const int hw = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const int parts = (int)std::min<size_t>((size_t)std::min(hw, std::max(n, 1)), std::max<size_t>(1, bytes >> 24));
    if (parts <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int step = (n + parts - 1) / parts;
    for (int p = 1; p < parts; ++p) {
        const int lo = p * step, hi = std::min(n, lo + step);
        if (lo < hi) pool.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    f(0, std::min(n, step));
    for (auto& t : pool) t.join();
}

// Ends of consecutive ranges of [0, n) holding about `target` bytes each (bytes(k) of
// element k), at most 32 ranges: host copies of one range overlap the DMA of the next.
template <class F>
std::vector<int> byte_chunks(int n, F&& bytes, size_t target) {
    size_t total = 0;
    for (int k = 0; k < n; ++k) total += bytes(k);
    target = std::max(target, total / 32 + 1);
    std::vector<int> ends;
    size_t acc = 0;
    for (int k = 0; k < n; ++k) {
        acc += bytes(k);
        if (acc >= target) {
            ends.push_back(k + 1);
            acc = 0;
        }
    }
    if (ends.empty() || ends.back() != n) ends.push_back(n);
    return ends;
}
constexpr size_t kCopyChunk = 128u << 20;

// Caps of one seam from its selected nodes; `grow` scales them after an overflow.  A
// seam's octree has about 1.4 nodes per leaf (the leaves lie on faces) and at most ~3
// quads per leaf.
// Returns false when the caps exceed what the pass indexes (node indices below 2^27, call
// lists and hash tables below 2^30 entries).
bool seam_caps(SeamJobDev& J, int selected, int grow) {
    const int64_t leaf = std::max(selected, 1);
    const int64_t node = (2 * leaf + 64) * grow, list = (6 * leaf + 1024) * grow;
    int64_t h = 1;
    while (h < node + node / 2) h <<= 1;
    if (node > kCallNodeMask || list >= (1 << 30) || h >= (1 << 30)) return false;
    J.leaf_cap = (int)leaf;
    J.node_cap = (int)node;
    J.list_cap = (int)list;
    J.hash_cap = (int)h;
    return true;
}

// Device location and length of a source given as count = -1 without a list: a cached
// chunk of the context, or a staged chunk of a batch.
using SourceResolver = int (*)(void* owner, const lvn_seam_source& q, SeamSrcDev& d);

int resolve_cached(void* owner, const lvn_seam_source& q, SeamSrcDev& d) {
    Ctx* c = static_cast<Ctx*>(owner);
    auto it = c->meshes.find(make_key(q.min, q.size));
    if (it == c->meshes.end()) return LVN_ERR_STATE;
    const MeshEntry& m = it->second;
    d.nodes = reinterpret_cast<const lvn_seam_node*>(m.d + 48 * (size_t)m.V + 12 * (size_t)m.T);
    d.count = m.S;
    return LVN_OK;
}

int resolve_batch(void* owner, const lvn_seam_source& q, SeamSrcDev& d) {
    Batch* b = static_cast<Batch*>(owner);
    auto it = b->slot_of.find(make_key(q.min, q.size));
    if (it == b->slot_of.end()) return LVN_ERR_STATE;
    const Ctx& c = b->ctx;
    const Counters& k = c.h_counters[it->second];  // read back by the caller of run_seams
    if (k.status) return LVN_ERR_OVERFLOW;
    d.nodes = reinterpret_cast<const lvn_seam_node*>(c.d_arena + (uint64_t)it->second * c.L.slot_bytes + c.L.seams);
    d.count = k.seams;
    return LVN_OK;
}

// The grid-wide seam pass over m seams whose jobs (caps, scratch and output offsets) are
// in device memory as d_jobs and on the host as h_jobs.  Launch sequence: tile lists of
// the sizes the host knows, clear, select, leaves, ancestors, links, level buckets, leaf
// counts (deepest level first), ranks (top level first), vertices, contouring rounds,
// triangles.  The rounds run in batches without a read-back; after a batch the host reads
// how many seams still expand.  A call descends at least one level per round, so the
// first batch (deepest possible level + 3 rounds) normally finishes every seam.
int run_seam_grid(Ctx* c, SeamState& st, int m, const SeamJobDev* h_jobs, const std::vector<SeamSrcDev>& srcs,
                  const SeamJobDev* d_jobs, lvn_vertex* ov, int* ot, SeamResDev* d_res) {
    // list lengths: eight interleaved sub-lists, seam k in sub-list k mod 8, each padded to
    // the longest (bounds from the seams' caps)
    const bool spread = m < kSpreadBelow;
    auto nt = [](int64_t n) { return (n + kSgTile - 1) / kSgTile; };
    int64_t xh[kXcds] = {}, xo[kXcds] = {}, xl[kXcds] = {}, xn[kXcds] = {}, xr[kXcds] = {};
    auto add = [spread](int64_t* xs, int s, int64_t cnt) {  // sub-list lengths as sg_place makes them
        if (!spread) {
            xs[s & (kXcds - 1)] += cnt;
            return;
        }
        for (int x = 0; x < kXcds; ++x) xs[x] += cnt / kXcds + (((x - s) & (kXcds - 1)) < cnt % kXcds ? 1 : 0);
    };
    int lmax = 0;
    for (int k = 0; k < m; ++k) {
        const SeamJobDev& J = h_jobs[k];
        add(xh, k, nt(J.hash_cap));
        add(xl, k, nt(J.leaf_cap));
        add(xn, k, nt(J.node_cap));
        add(xr, k, nt(J.list_cap));
        const int root_size = 2 * J.host_size, root_levels = __builtin_ctz((unsigned)root_size);
        int64_t offered = 0;
        for (int si = 0; si < J.src_count; ++si) {
            const SeamSrcDev& q = srcs[J.src_begin + si];
            offered += nt(q.count);
            const int leaf = q.size / J.vpc;
            if (host_pow2(leaf) && leaf < root_size) lmax = std::max(lmax, root_levels - __builtin_ctz((unsigned)leaf));
        }
        add(xo, k, offered);
    }
    auto padded = [](const int64_t* v) { return kXcds * std::max<int64_t>(1, *std::max_element(v, v + kXcds)); };
    const int64_t th = padded(xh), to = padded(xo), tl = padded(xl), tn = padded(xn), tr = padded(xr);
    lmax = std::min(lmax, kSgLevels - 1);
    const int64_t o_off = th, o_leaf = o_off + to, o_node = o_leaf + tl, o_lev = o_node + tn,
                  o_round = o_lev + (lmax + 1) * tn, o_fin = o_round + 2 * tr, total = o_fin + tr;
    int r;
    if ((r = st.g_run.ensure(m * sizeof(SeamRun))) || (r = st.g_tiles.ensure(total * sizeof(int4))) ||
        (r = st.g_tsum.ensure(3 * tr * sizeof(int))) || (r = st.g_cnt.ensure(kSgCounters * sizeof(int))) ||
        (r = st.h_active.ensure(sizeof(int))))
        return r;
    int* cnt = static_cast<int*>(st.g_cnt.p);
    int* tsum = static_cast<int*>(st.g_tsum.p);
    int* toff = tsum + tr;
    int* tbase = toff + tr;
    int4* T = static_cast<int4*>(st.g_tiles.p);
    SeamGrid g{d_jobs, static_cast<const SeamSrcDev*>(st.srcs.p), static_cast<char*>(st.scratch.p),
               static_cast<SeamRun*>(st.g_run.p), cnt, ov, ot, m, spread ? 1 : 0};
    const hipStream_t sm = c->stream;
    // tile launches: a multiple of 8 workgroups (see k_sg_tiles), looping over their lists
    static_assert(kSgGrid % kXcds == 0, "tile launches keep to one XCD per workgroup");
    auto grid = [](int64_t tiles) { return dim3((unsigned)std::min<int64_t>(tiles, kSgGrid)); };
    const dim3 per_seam((m + kSgThreads - 1) / kSgThreads), block(kSgThreads), dev_grid(kSgGrid);
#define LVN_SG(...)                              \
    do {                                          \
        hipLaunchKernelGGL(__VA_ARGS__);          \
        LVN_HIP(hipGetLastError());               \
    } while (0)
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 3, T, 0, cnt + kSgCntHash);
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 6, T + o_off, 0, cnt + kSgCntOffered);
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 4, T + o_leaf, 0, cnt + kSgCntLeaf);
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 5, T + o_node, 0, cnt + kSgCntNode);
    LVN_SG(k_sg_init, dev_grid, block, 0, sm, g, T);
    LVN_SG(k_sg_select, grid(to), block, 0, sm, g, T + o_off);
    LVN_SG(k_sg_leaves, grid(th), block, 0, sm, g, T);
    LVN_SG(k_sg_seam, per_seam, block, 0, sm, g, 0);
    LVN_SG(k_sg_ancestors, grid(tl), block, 0, sm, g, T + o_leaf);
    LVN_SG(k_sg_seam, per_seam, block, 0, sm, g, 1);
    LVN_SG(k_sg_links, grid(tn), block, 0, sm, g, T + o_node);
    LVN_SG(k_sg_seam, per_seam, block, 0, sm, g, 2);
    LVN_SG(k_sg_order, grid(tn), block, 0, sm, g, T + o_node);
    LVN_SG(k_sg_tiles, dim3(lmax + 1), dim3(1024), 0, sm, g, 0, T + o_lev, (int)tn, cnt + kSgCntLevel);
    for (int l = lmax - 1; l >= 1; --l) LVN_SG(k_sg_level, dev_grid, block, 0, sm, g, T + o_lev + l * tn, l, 0);
    LVN_SG(k_sg_seam, per_seam, block, 0, sm, g, 3);
    for (int l = 0; l < lmax; ++l) LVN_SG(k_sg_level, dev_grid, block, 0, sm, g, T + o_lev + l * tn, l, 1);
    LVN_SG(k_sg_vertices, grid(tl), block, 0, sm, g, T + o_leaf);
    // round r reads tile list r & 1 and makes list (r + 1) & 1
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 1, T + o_round, 0, cnt + kSgCntRound);
    int round = 0;
    for (int batch = lmax + 3; round < kMaxRounds; batch = 2) {
        for (int k = 0; k < batch && round < kMaxRounds; ++k, ++round) {
            int4* cur = T + o_round + (round & 1) * tr;
            int4* nxt = T + o_round + ((round + 1) & 1) * tr;
            LVN_SG(k_sg_round, dev_grid, block, 0, sm, g, cur, round, tsum, tbase);
            LVN_SG(k_sg_round_next, dim3(1), dim3(1024), 0, sm, g, round, tsum, toff, nxt);
            LVN_SG(k_sg_round_copy, dev_grid, block, 0, sm, g, cur, round, tsum, tbase, toff);
        }
        LVN_HIP(hipMemcpyAsync(st.h_active.p, cnt + kSgCntActive + round - 1, sizeof(int), hipMemcpyDeviceToHost, sm));
        LVN_HIP(hipStreamSynchronize(sm));
        if (*static_cast<const int*>(st.h_active.p) == 0) break;
    }
    LVN_SG(k_sg_tiles, dim3(1), dim3(1024), 0, sm, g, 2, T + o_fin, 0, cnt + kSgCntFinal);
    LVN_SG(k_sg_triangles, dev_grid, block, 0, sm, g, T + o_fin, d_res);
#undef LVN_SG
    return LVN_OK;
}

// Which seam pass runs a call: the grid-wide steps (measured on the clipmap's seams: its
// largest seam alone 3.1 against 7.8 ms, 512 seams 10.2 against 11.6 ms, all 1,798 seams
// 32.6 against 33.4 ms) unless every seam is small, where one workgroup per seam has the
// shorter fixed cost (a seam of 6,440 offered nodes: 0.20 against 0.38 ms; 36,606: 1.17
// against 0.80 ms).  LVN_SEAM_PATH=grid or =workgroup in the environment overrides the
// choice.
bool seam_pass_grid(const std::vector<SeamJobDev>& jobs, const std::vector<SeamSrcDev>& srcs) {
    if (const char* e = getenv("LVN_SEAM_PATH")) {
        if (!strcmp(e, "grid")) return true;
        if (!strcmp(e, "workgroup")) return false;
    }
    int64_t largest = 0;
    for (const SeamJobDev& J : jobs) {
        int64_t offered = 0;
        for (int si = 0; si < J.src_count; ++si) offered += srcs[J.src_begin + si].count;
        largest = std::max(largest, offered);
    }
    return largest >= 16384;
}

int run_seams(Ctx* c, int n, const lvn_seam_req* reqs, const lvn_seam_source* sources, lvn_seam_out* outs,
              SourceResolver resolve, void* owner) {
    if (!c->seams) c->seams = new SeamState();
    SeamState& st = *c->seams;
    // upload every distinct host list once; cached chunks are read where they lie
    std::map<const lvn_seam_node*, std::pair<uint64_t, size_t>> uploaded;  // offset, nodes
    std::vector<SeamSrcDev> srcs;
    size_t in_nodes = 0;
    int total_sources = 0;
    for (int i = 0; i < n; ++i) total_sources = std::max(total_sources, reqs[i].first_source + reqs[i].n_sources);
    for (int s = 0; s < total_sources; ++s) {
        const lvn_seam_source& q = sources[s];
        if (q.nodes && q.count > 0) {
            auto it = uploaded.find(q.nodes);
            if (it == uploaded.end()) {
                uploaded[q.nodes] = {in_nodes, (size_t)q.count};
                in_nodes += (size_t)q.count;
            } else if ((size_t)q.count > it->second.second) {  // the same list offered with a longer count
                it->second = {in_nodes, (size_t)q.count};
                in_nodes += (size_t)q.count;
            }
        }
    }
    int r = st.in.ensure(std::max<size_t>(in_nodes, 1) * sizeof(lvn_seam_node));
    if (!r) r = st.h_in.ensure(std::max<size_t>(in_nodes, 1) * sizeof(lvn_seam_node));
    if (r) return r;
    auto* h_in = static_cast<lvn_seam_node*>(st.h_in.p);
    auto* d_in = static_cast<lvn_seam_node*>(st.in.p);
    {  // lists staged in pinned memory in chunks, each uploaded while the next is staged
        std::vector<std::pair<const lvn_seam_node*, std::pair<uint64_t, size_t>>> lists(uploaded.begin(),
                                                                                         uploaded.end());
        std::sort(lists.begin(), lists.end(), [](const auto& a, const auto& b) { return a.second.first < b.second.first; });
        const auto ends = byte_chunks((int)lists.size(), [&](int k) { return lists[k].second.second * sizeof(lvn_seam_node); },
                                      kCopyChunk);
        int k0 = 0;
        for (int k1 : ends) {
            if (k1 == k0) continue;
            size_t bytes = 0;
            for (int k = k0; k < k1; ++k) bytes += lists[k].second.second * sizeof(lvn_seam_node);
            parallel_ranges(k1 - k0, bytes, [&](int lo, int hi) {
                for (int k = k0 + lo; k < k0 + hi; ++k)
                    std::memcpy(h_in + lists[k].second.first, lists[k].first,
                                lists[k].second.second * sizeof(lvn_seam_node));
            });
            const uint64_t first = lists[k0].second.first;
            const uint64_t end = lists[k1 - 1].second.first + lists[k1 - 1].second.second;
            LVN_HIP(hipMemcpyAsync(d_in + first, h_in + first, (end - first) * sizeof(lvn_seam_node),
                                   hipMemcpyHostToDevice, c->stream));
            k0 = k1;
        }
    }
    srcs.resize(std::max(total_sources, 1));
    for (int s = 0; s < total_sources; ++s) {
        const lvn_seam_source& q = sources[s];
        SeamSrcDev& d = srcs[s];
        d.neighbour = q.neighbour;
        d.size = q.size;
        for (int a = 0; a < 3; ++a) d.min[a] = q.min[a];
        if (q.nodes || q.count >= 0) {
            d.nodes = q.nodes && q.count > 0 ? d_in + uploaded[q.nodes].first : nullptr;
            d.count = q.nodes ? q.count : 0;
        } else {  // count < 0: seam nodes already in device memory
            const int rr = resolve(owner, q, d);
            if (rr) return rr;
        }
    }
    if ((r = st.srcs.ensure(srcs.size() * sizeof(SeamSrcDev))) || (r = st.h_srcs.ensure(srcs.size() * sizeof(SeamSrcDev))))
        return r;
    std::memcpy(st.h_srcs.p, srcs.data(), srcs.size() * sizeof(SeamSrcDev));
    LVN_HIP(hipMemcpyAsync(st.srcs.p, st.h_srcs.p, srcs.size() * sizeof(SeamSrcDev), hipMemcpyHostToDevice, c->stream));

    // the selected nodes of each seam (one counting pass) size its scratch
    std::vector<SeamJobDev> base(n);
    for (int i = 0; i < n; ++i) {
        const lvn_seam_req& q = reqs[i];
        SeamJobDev& J = base[i];
        std::memset(&J, 0, sizeof J);
        for (int a = 0; a < 3; ++a) J.host_min[a] = q.host_min[a];
        J.host_size = q.host_size;
        for (int a = 0; a < 4; ++a) J.colour[a] = q.colour[a];
        J.src_begin = q.first_source;
        J.src_count = q.n_sources;
        J.vpc = c->N;
    }
    if ((r = st.jobs.ensure(n * sizeof(SeamJobDev))) || (r = st.h_jobs.ensure(n * sizeof(SeamJobDev))) ||
        (r = st.counts.ensure(n * sizeof(int))) || (r = st.h_counts.ensure(n * sizeof(int))))
        return r;
    std::memcpy(st.h_jobs.p, base.data(), n * sizeof(SeamJobDev));
    LVN_HIP(hipMemcpyAsync(st.jobs.p, st.h_jobs.p, n * sizeof(SeamJobDev), hipMemcpyHostToDevice, c->stream));
    LVN_HIP(hipMemsetAsync(st.counts.p, 0, n * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_seam_count, dim3(n, per_seam_blocks(n)), dim3(256), 0, c->stream,
                       static_cast<const SeamJobDev*>(st.jobs.p),
                       static_cast<const SeamSrcDev*>(st.srcs.p), static_cast<int*>(st.counts.p));
    LVN_HIP(hipGetLastError());
    LVN_HIP(hipMemcpyAsync(st.h_counts.p, st.counts.p, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> selected(static_cast<const int*>(st.h_counts.p), static_cast<const int*>(st.h_counts.p) + n);

    // largest seams first: a seam is one workgroup, the biggest set the pass's length
    std::vector<int> grow(n, 1), todo(n);
    for (int i = 0; i < n; ++i) {
        todo[i] = i;
        outs[i].status = LVN_ERR_OVERFLOW;
        outs[i].nv = outs[i].nt = 0;
    }
    std::stable_sort(todo.begin(), todo.end(), [&](int a, int b) { return selected[a] > selected[b]; });
    st.loc.assign(n, SeamLoc{});
    bool host_out = false;
    for (int i = 0; i < n; ++i) host_out |= outs[i].v != nullptr || outs[i].tri != nullptr;
    for (int attempt = 0; attempt < 4 && !todo.empty(); ++attempt) {
        if ((int)st.out_v.size() <= attempt) {
            st.out_v.emplace_back(new DevBuf());
            st.out_t.emplace_back(new DevBuf());
        }
        DevBuf& ov = *st.out_v[attempt];
        DevBuf& ot = *st.out_t[attempt];
        std::vector<SeamJobDev> jobs;
        {  // a seam whose caps outgrow the pass stays LVN_ERR_OVERFLOW
            std::vector<int> fit;
            for (int i : todo) {
                SeamJobDev J = base[i];
                if (!seam_caps(J, selected[i], grow[i])) continue;
                fit.push_back(i);
                jobs.push_back(J);
            }
            todo.swap(fit);
        }
        const int m = (int)todo.size();
        if (m == 0) break;
        uint64_t scratch = 0, v_off = 0, t_off = 0;
        for (int k = 0; k < m; ++k) {
            SeamJobDev& J = jobs[k];
            J.scratch = scratch;
            J.v_off = v_off;
            J.t_off = t_off;
            scratch += seam_scratch_bytes(J.hash_cap, J.node_cap, J.list_cap);
            v_off += (uint64_t)J.leaf_cap;
            t_off += 2 * (uint64_t)J.list_cap;
        }
        if ((r = st.jobs.ensure(m * sizeof(SeamJobDev))) || (r = st.h_jobs.ensure(m * sizeof(SeamJobDev))) ||
            (r = st.scratch.ensure(scratch)) || (r = ov.ensure(std::max<uint64_t>(v_off, 1) * sizeof(lvn_vertex))) ||
            (r = ot.ensure(std::max<uint64_t>(t_off, 1) * 12)) || (r = st.res.ensure(m * sizeof(SeamResDev))) ||
            (r = st.h_res.ensure(m * sizeof(SeamResDev) + (2 * (size_t)m + 2) * 8)) ||
            (r = st.offs.ensure((2 * (size_t)m + 2) * 8)))
            return r;
        std::memcpy(st.h_jobs.p, jobs.data(), m * sizeof(SeamJobDev));
        LVN_HIP(hipMemcpyAsync(st.jobs.p, st.h_jobs.p, m * sizeof(SeamJobDev), hipMemcpyHostToDevice, c->stream));
        const auto* d_jobs = static_cast<const SeamJobDev*>(st.jobs.p);
        auto* d_res = static_cast<SeamResDev*>(st.res.p);
        auto* d_offs = static_cast<uint64_t*>(st.offs.p);
        if (seam_pass_grid(jobs, srcs)) {
            if ((r = run_seam_grid(c, st, m, jobs.data(), srcs, d_jobs, static_cast<lvn_vertex*>(ov.p),
                                   static_cast<int*>(ot.p), d_res)))
                return r;
        } else {
            hipLaunchKernelGGL(k_seam, dim3(m), dim3(kSeamThreads), 0, c->stream, d_jobs,
                               static_cast<const SeamSrcDev*>(st.srcs.p), static_cast<char*>(st.scratch.p),
                               static_cast<lvn_vertex*>(ov.p), static_cast<int*>(ot.p), d_res);
            LVN_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_seam_offsets, dim3(1), dim3(1024), 0, c->stream, d_res, m, d_offs);
        LVN_HIP(hipGetLastError());
        char* h_res = static_cast<char*>(st.h_res.p);
        LVN_HIP(hipMemcpyAsync(h_res, d_res, m * sizeof(SeamResDev), hipMemcpyDeviceToHost, c->stream));
        LVN_HIP(hipMemcpyAsync(h_res + m * sizeof(SeamResDev), d_offs, (2 * (size_t)m + 2) * 8, hipMemcpyDeviceToHost,
                               c->stream));
        LVN_HIP(hipStreamSynchronize(c->stream));
        const auto* rs = reinterpret_cast<const SeamResDev*>(h_res);
        const auto* offs = reinterpret_cast<const uint64_t*>(h_res + m * sizeof(SeamResDev));
        const uint64_t tv = offs[2 * m], tt = offs[2 * m + 1];
        std::vector<int> again;
        for (int k = 0; k < m; ++k) {
            const int i = todo[k];
            lvn_seam_out& o = outs[i];
            if (rs[k].status == LVN_ERR_OVERFLOW) {
                grow[i] *= 4;
                again.push_back(i);
                continue;
            }
            o.status = rs[k].status;
            o.nv = rs[k].nv;
            o.nt = rs[k].nt;
            if (o.status) continue;
            SeamLoc& l = st.loc[i];
            l.attempt = attempt;
            l.nv = o.nv;
            l.nt = o.nt;
            l.v_off = jobs[k].v_off;
            l.t_off = jobs[k].t_off;
            if ((o.v && o.nv > o.v_cap) || (o.tri && o.nt > o.tri_cap)) o.status = LVN_ERR_CAPACITY;
        }
        if (host_out) {
            // every seam's mesh packed back to back, read back in chunks of seams; each chunk
            // is copied into the callers' buffers while the next one is in flight
            if ((r = st.pv.ensure(std::max<uint64_t>(tv, 1) * sizeof(lvn_vertex))) ||
                (r = st.pt.ensure(std::max<uint64_t>(tt, 1) * 12)) ||
                (r = st.h_pv.ensure(std::max<uint64_t>(tv, 1) * sizeof(lvn_vertex))) ||
                (r = st.h_pt.ensure(std::max<uint64_t>(tt, 1) * 12)))
                return r;
            hipLaunchKernelGGL(k_seam_pack, dim3(m, per_seam_blocks(m)), dim3(256), 0, c->stream, d_jobs, d_res, d_offs,
                               static_cast<const lvn_vertex*>(ov.p), static_cast<const int*>(ot.p),
                               static_cast<lvn_vertex*>(st.pv.p), static_cast<int*>(st.pt.p));
            LVN_HIP(hipGetLastError());
            auto seam_bytes = [&](int k) {
                return (offs[2 * k + 2] - offs[2 * k]) * sizeof(lvn_vertex) + (offs[2 * k + 3] - offs[2 * k + 1]) * 12;
            };
            const auto ends = byte_chunks(m, seam_bytes, kCopyChunk);
            while (st.ev.size() < ends.size()) {
                hipEvent_t e;
                LVN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                st.ev.push_back(e);
            }
            auto* hv = static_cast<lvn_vertex*>(st.h_pv.p);
            auto* ht = static_cast<int*>(st.h_pt.p);
            const auto* dv = static_cast<const lvn_vertex*>(st.pv.p);
            const auto* dt = static_cast<const int*>(st.pt.p);
            int k0 = 0;
            for (size_t ci = 0; ci < ends.size(); ++ci) {
                const int k1 = ends[ci];
                const uint64_t v0 = offs[2 * k0], v1 = offs[2 * k1], t0 = offs[2 * k0 + 1], t1 = offs[2 * k1 + 1];
                if (v1 > v0)
                    LVN_HIP(hipMemcpyAsync(hv + v0, dv + v0, (v1 - v0) * sizeof(lvn_vertex), hipMemcpyDeviceToHost,
                                           c->stream));
                if (t1 > t0)
                    LVN_HIP(hipMemcpyAsync(ht + 3 * t0, dt + 3 * t0, (t1 - t0) * 12, hipMemcpyDeviceToHost, c->stream));
                LVN_HIP(hipEventRecord(st.ev[ci], c->stream));
                k0 = k1;
            }
            k0 = 0;
            for (size_t ci = 0; ci < ends.size(); ++ci) {
                const int k1 = ends[ci];
                LVN_HIP(hipEventSynchronize(st.ev[ci]));
                size_t bytes = 0;
                for (int k = k0; k < k1; ++k) bytes += seam_bytes(k);
                parallel_ranges(k1 - k0, bytes, [&](int lo, int hi) {
                    for (int k = k0 + lo; k < k0 + hi; ++k) {
                        const lvn_seam_out& o = outs[todo[k]];
                        if (rs[k].status != LVN_OK || o.status != LVN_OK) continue;
                        if (o.nv && o.v) std::memcpy(o.v, hv + offs[2 * k], o.nv * sizeof(lvn_vertex));
                        if (o.nt && o.tri) std::memcpy(o.tri, ht + 3 * offs[2 * k + 1], (size_t)o.nt * 12);
                    }
                });
                k0 = k1;
            }
        }
        todo.swap(again);
    }
    int first = LVN_OK;
    for (int i = 0; i < n && !first; ++i) first = outs[i].status;
    return first;
}

int check_seam_args(int n, const lvn_seam_req* reqs, const lvn_seam_source* sources) {
    if (n < 0 || (n > 0 && !reqs)) return LVN_ERR_INVALID;
    for (int i = 0; i < n; ++i) {
        const lvn_seam_req& q = reqs[i];
        if (!host_pow2(q.host_size) || q.host_size > (1 << 20) || q.first_source < 0 || q.n_sources < 0)
            return LVN_ERR_INVALID;
        if (q.n_sources > 0 && !sources) return LVN_ERR_INVALID;
        for (int s = 0; s < q.n_sources; ++s) {
            const lvn_seam_source& src = sources[q.first_source + s];
            if (src.neighbour < 0 || src.neighbour > 7 || src.size <= 0 || (src.nodes && src.count < 0) ||
                (!src.nodes && src.count > 0))
                return LVN_ERR_INVALID;
        }
    }
    return LVN_OK;
}

}  // namespace
}  // namespace lvn

using namespace lvn;

extern "C" {

int lvn_generate_seam_meshes(lvn_ctx* ctx, int32_t n, const lvn_seam_req* reqs, const lvn_seam_source* sources,
                             lvn_seam_out* outs) {
    if (!ctx || (n > 0 && !outs)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    const int r = check_seam_args(n, reqs, sources);
    if (r) return r;
    if (n == 0) return LVN_OK;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    return run_seams(c, n, reqs, sources, outs, resolve_cached, c);
}

int lvn_batch_seam_meshes(lvn_batch* bh, int32_t n, const lvn_seam_req* reqs, const lvn_seam_source* sources,
                          lvn_seam_out* outs) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || (n > 0 && !outs)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    const int r = check_seam_args(n, reqs, sources);
    if (r) return r;
    if (n == 0) return LVN_OK;
    Ctx* c = &b->ctx;
    if (b->n > 0) {  // seam counts of the staged chunks, after the batch's last run
        LVN_HIP(hipMemcpyAsync(c->h_counters, c->d_counters, sizeof(Counters) * b->n, hipMemcpyDeviceToHost, c->stream));
        LVN_HIP(hipStreamSynchronize(c->stream));
    }
    return run_seams(c, n, reqs, sources, outs, resolve_batch, b);
}

int lvn_batch_seam_device_mesh(lvn_batch* bh, int32_t i, const lvn_vertex** v, const int32_t** tri, int32_t* nv,
                               int32_t* nt) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b) return LVN_ERR_INVALID;
    return lvn_seam_device_mesh(reinterpret_cast<lvn_ctx*>(&b->ctx), i, v, tri, nv, nt);
}

int lvn_seam_device_mesh(lvn_ctx* ctx, int32_t i, const lvn_vertex** v, const int32_t** tri, int32_t* nv,
                         int32_t* nt) {
    if (!ctx || !v || !tri || !nv || !nt) return LVN_ERR_INVALID;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c->seams || i < 0 || i >= (int)c->seams->loc.size()) return LVN_ERR_INVALID;
    const SeamLoc& l = c->seams->loc[i];
    if (l.attempt < 0) return LVN_ERR_STATE;
    *v = static_cast<const lvn_vertex*>(c->seams->out_v[l.attempt]->p) + l.v_off;
    *tri = static_cast<const int32_t*>(c->seams->out_t[l.attempt]->p) + 3 * l.t_off;
    *nv = l.nv;
    *nt = l.nt;
    return LVN_OK;
}

int lvn_generate_seam_mesh(lvn_ctx* ctx, const int32_t host_min[3], int32_t host_size,
                           const lvn_seam_source* sources, int32_t n_sources, const float colour[3], lvn_vertex* v,
                           int32_t v_cap, int32_t* nv, int32_t* tri, int32_t tri_cap, int32_t* nt) {
    if (!host_min || !nv || !nt || !colour || n_sources < 0 || v_cap < 0 || tri_cap < 0) return LVN_ERR_INVALID;
    lvn_seam_req q;
    std::memset(&q, 0, sizeof q);
    for (int a = 0; a < 3; ++a) q.host_min[a] = host_min[a];
    q.host_size = host_size;
    for (int a = 0; a < 3; ++a) q.colour[a] = colour[a];
    q.first_source = 0;
    q.n_sources = n_sources;
    lvn_seam_out o;
    std::memset(&o, 0, sizeof o);
    o.v = v;
    o.tri = tri;
    o.v_cap = v ? v_cap : 0;
    o.tri_cap = tri ? tri_cap : 0;
    const int r = lvn_generate_seam_meshes(ctx, 1, &q, sources, &o);
    *nv = o.nv;
    *nt = o.nt;
    return r;
}

}  // extern "C"
