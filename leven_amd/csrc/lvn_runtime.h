// lvn_runtime.h — host-side state of the chunk mesher runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "lvn_kernels.h"

#include <cstdio>

// Returns LVN_ERR_DEVICE from the calling function when a HIP call fails, after
// printing the call and the runtime's message.
#define LVN_HIP(call)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            fprintf(stderr, "lvn: HIP error %s at %s:%d (%s)\n", hipGetErrorString(e_), __FILE__,     \
                    __LINE__, #call);                                                                 \
            return LVN_ERR_DEVICE;                                                                    \
        }                                                                                             \
    } while (0)

namespace lvn {

// Seam-stitching buffers of a context (lvn_seams.hip).
struct SeamState;
void destroy_seam_state(SeamState* s);
// Simplifier buffers of a context (lvn_simplify.hip).
struct SimpState;
void destroy_simp_state(SimpState* s);

// Global state set by lvn_init / lvn_set_noise_seed / lvn_store_csg_op
// (the reference keeps the same as process globals: src/compute.cpp:168-191,
// src/compute_density_field.cpp:23-24).
struct Global {
    std::atomic<bool> init{false};
    lvn_field_desc desc{};
    FieldParams fp{};
    int32_t seed = 0;
    uint32_t defmat = 0;
    std::vector<uint8_t> perm;
    uint8_t* d_perm = nullptr;
    float4* d_grads = nullptr;
    std::vector<lvn_csg_op> ops;
    std::vector<std::array<int32_t, 6>> aabbs;
    uint64_t generation = 0;  // bumped when field, table pointers or default material change
    int live = 0;             // contexts, batches and servers alive (each holds the tables)
    std::mutex mu;            // guards the tables' contents and lifetime and `live`
};
extern Global g_state;

// Contexts, batches and servers keep the device tables alive: lvn_shutdown frees them
// at once when none exists, else the last one's destruction does.
void retain_global();
void release_global();

// Cache key: ivec4(min, size) as in src/compute_local.h:125,138.
using Key = std::tuple<int32_t, int32_t, int32_t, int32_t>;
inline Key make_key(const int32_t* min, int32_t size) { return Key(min[0], min[1], min[2], size); }

// Device-resident density-field state (GPUDensityField, src/compute_local.h:113-123):
// final materials and the sorted edge list with its hermite data.
// A cached density field: its materials and edge list, in the context's field pool.
struct FieldEntry {
    uint8_t* d_mat = nullptr;
    int* d_ids = nullptr;
    float4* d_data = nullptr;
    int edges = 0;
    int cap = 0;
    int last_op = 0;
};

// Device memory for cached fields, handed out from large blocks: the reference keeps
// a field until shutdown (src/compute_local.h:28-72), so entries are never freed one
// by one and a bump allocator replaces a hipMalloc per cached chunk.
struct DevicePool {
    std::vector<void*> blocks;
    char* cur = nullptr;
    size_t left = 0;
    void* take(size_t bytes);  // 256-byte aligned, or null when the device is out of memory
    void release();            // frees every block
    ~DevicePool();
};

// Power-of-two size classes over a DevicePool with per-class free lists: cached meshes
// come and go (freeChunkOctree), so their blocks are recycled.
struct ClassPool {
    DevicePool base;
    std::map<size_t, std::vector<char*>> free_blocks;
    char* take(size_t bytes, size_t* cap);
    void give(char* p, size_t cap) { free_blocks[cap].push_back(p); }
};

// A finished chunk (the reference caches its octree on the device; the mesh is a
// pure function of it, so the mesh is what is kept, on the device): vertices,
// triangles and seam nodes back to back in one block of the context's mesh pool.
struct MeshEntry {
    char* d = nullptr;
    size_t cap = 0;             // block size class
    int V = 0, T = 0, S = 0;    // exported vertices (0 without triangles), triangles, seam nodes
};

struct Ctx {
    bool retained = false;  // holds a reference on the global tables
    int N = 0;
    uint32_t flags = 0;
    Geometry G{};
    Caps caps{};
    SlotLayout L{};
    int slots = 0;
    hipStream_t stream = nullptr;
    char* d_arena = nullptr;
    ChunkDesc* d_chunks = nullptr;
    Counters* d_counters = nullptr;
    ChunkDesc* h_chunks = nullptr;
    Counters* h_counters = nullptr;
    DevOp* d_ops = nullptr;
    DevOp* h_ops = nullptr;
    int* d_op_index = nullptr;
    int* h_op_index = nullptr;
    int ops_cap = 0, idx_cap = 0;
    int* d_queue = nullptr;
    int* d_item_prefix = nullptr;
    int cus = 0;
    std::map<Key, FieldEntry> fields;
    std::map<Key, MeshEntry> meshes;
    DevicePool field_pool;
    ClassPool mesh_pool;
    hipGraphExec_t exec = nullptr;       // the one-chunk pipeline, captured on first use
    hipGraph_t graph = nullptr;
    const void* graph_ops = nullptr;     // op arrays the graph was captured with
    const void* graph_index = nullptr;
    uint64_t graph_gen = 0;              // global generation the graph was captured with
    float* h_verts = nullptr;   // pinned read-back staging for one chunk's mesh (caps sized)
    SeamState* seams = nullptr; // seam stitching buffers, made on first use
    SimpState* simp = nullptr;  // simplifier buffers, made on first use

    int allocate(int n, uint32_t flags, int nslots, int edge_cap, int voxel_cap);
    int ensure_ops(int nops, int nidx);
    KArgs kargs() const;
    ~Ctx();
};

// A batch of chunk slots processed by one launch per stage (the batched form).
struct Batch {
    Ctx ctx;
    int n = 0;                       // staged chunks
    hipGraphExec_t exec = nullptr;   // captured pipeline for `graph_n` chunks
    hipGraph_t graph = nullptr;
    int graph_n = -1;
    const void* graph_ops = nullptr;   // op arrays the graph was captured with
    const void* graph_index = nullptr;
    uint64_t graph_gen = 0;
    unsigned long long* d_work = nullptr;
    lvn_tuning tuning = {0, 2, 8192, 8192, 0, 1, 0, {0}};  // ysplit, cross_lanes, height_lines 0 = by batch size
    std::map<Key, int> slot_of;      // staged chunk (min, size) -> slot
    // lvn_batch_simplify's per-slot (vertices, triangles), valid until the pipeline runs again
    std::vector<int32_t> simp_counts;
    bool simplified = false;
    ~Batch();
};

// Launch-shape defaults shared by every form (the batched form overrides them per batch).
constexpr int kDefaultPersistPerCu = 2;
constexpr int kDefaultDensityPerCu = 1;
// Batches from this many chunks evaluate x/z-edge line heights once per line by default.
constexpr int kHeightLinesMinChunks = 64;
constexpr int kDefaultRowBlocks = 8192;
constexpr int kDefaultVoxelBlocks = 8192;

DevOp make_dev_op(const lvn_csg_op& op);
// Kernel arguments of one arena with the default launch shapes of the one-chunk forms
// (the tables and field of the global state).
KArgs make_kargs(char* arena, const SlotLayout& L, const Geometry& G, const Caps& caps, const ChunkDesc* chunks,
                 Counters* counters, const DevOp* ops, const int* op_index, uint32_t flags, int* queue,
                 int* item_prefix, int cus);
Geometry make_geometry(int n);
Caps make_caps(const Geometry& G, int edge_cap, int voxel_cap);
SlotLayout make_layout(const Geometry& G, const Caps& c);
void fill_chunk_desc(ChunkDesc& cd, const int32_t* min, int32_t size, int N);
// A chunk size the pipeline accepts for N voxels per chunk: a positive multiple of
// N * LEAF_SIZE_SCALE (sampleScale = size / (N * 4) >= 1, src/compute_density_field.cpp:149).
inline bool valid_chunk_size(int32_t size, int N) { return size >= N * kLeafSizeScale && size % (N * kLeafSizeScale) == 0; }
bool aabb_overlaps(const int32_t* mn, int32_t size, const std::array<int32_t, 6>& bb);
// Launch shapes within their documented ranges (F = density samples per axis).
bool tuning_valid(const lvn_tuning& t, int F);

}  // namespace lvn
