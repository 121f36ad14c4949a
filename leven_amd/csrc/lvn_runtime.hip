// lvn_runtime.hip — host runtime and C ABI of the chunk mesher.
//
// Mirrors leven's compute host layer (src/compute.cpp, src/compute_density_field.cpp,
// src/compute_octree.cpp, src/compute_csg.cpp): global noise seed, stored CSG
// operations, and per-context caches of density fields and finished chunks
// keyed by (min, size).  Every stage runs on the GPU; the host only sequences
// launches on one stream and copies results out (two synchronisations per
// chunk instead of the reference's eight or more).

#include "lvn_kernels.h"
#include "lvn_runtime.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace lvn {

Global g_state;

// ---------------------------------------------------------------------------
// Permutation table (CreateNoisePermutationLookupImage, src/compute_density_field.cpp:90-125)
// ---------------------------------------------------------------------------

namespace {

// Ken Perlin's reference permutation; the reference's perm[512] holds it, then
// the same list without its first six entries, then six zeros
// (src/compute_density_field.cpp:28-52).
const uint8_t kPerlin[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  223, 183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

// Perlin's 16 gradients (src/compute_density_field.cpp:60-63).
const int kGrad3[16][3] = {{0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1}, {1, 0, 1},  {1, 0, -1},
                           {-1, 0, 1}, {-1, 0, -1}, {1, 1, 0},  {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0},
                           {1, 0, -1}, {-1, 0, -1}, {0, -1, 1}, {0, 1, 1}};

// Shuffle with a fixed generator (the reference's std::shuffle +
// default_random_engine is library-defined): Fisher-Yates from index 511 down,
// j = splitmix64 % (i + 1), state = (uint32)seed.
void build_perm_table(int32_t seed, uint8_t* out) {
    int sh[512];
    for (int i = 0; i < 256; ++i) sh[i] = kPerlin[i];
    for (int i = 0; i < 250; ++i) sh[256 + i] = kPerlin[6 + i];
    for (int i = 506; i < 512; ++i) sh[i] = 0;
    uint64_t st = (uint64_t)(uint32_t)seed;
    for (int i = 511; i > 0; --i) {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        std::swap(sh[i], sh[z % (uint64_t)(i + 1)]);
    }
    for (int row = 0; row < 256; ++row)
        for (int col = 0; col < 256; ++col) {
            // NoiseHash (src/compute_density_field.cpp:68-86): one-at-a-time hash of the key bytes.
            const uint32_t key = (((uint32_t)row << 24) | ((uint32_t)col << 16)) ^ (uint32_t)seed;
            uint32_t h = 0;
            for (int b = 0; b < 4; ++b) {
                h += (key >> (8 * b)) & 0xffu;
                h += h << 10;
                h ^= h >> 6;
            }
            h += h << 3;
            h ^= h >> 11;
            h += h << 15;
            out[row * 256 + col] = (uint8_t)sh[h & 0x1ffu];
        }
}

// UNORM8 decode of the gradient bytes g*64+64, then *4-1 (cl/simplex.cl:124).  Entry j
// serves every stored table byte w = value + 1 with w & 15 == j: the gradient of value
// w - 1, i.e. (j + 15) & 15.
void build_grads(float4* out) {
    for (int w = 0; w < kGradEntries; ++w) {
        const int i = (w + 15) & 15;
        float c[3];
        for (int k = 0; k < 3; ++k) c[k] = ((float)(kGrad3[i][k] * 64 + 64) / 255.0f) * 4.0f - 1.0f;
        out[w] = make_float4(c[0], c[1], c[2], 0.f);
    }
}

FieldParams make_field_params(const lvn_field_desc& d) {
    FieldParams f;
    f.kind = d.kind;
    f.hoct = d.height_octaves;
    f.hfreq = d.height_frequency;
    f.hlac = d.height_lacunarity;
    f.hgain = d.height_gain;
    f.hamp = d.height_amplitude;
    f.coct = d.cave_octaves;
    f.cfreq = d.cave_frequency;
    f.clac = d.cave_lacunarity;
    f.cgain = d.cave_gain;
    f.cthr = d.cave_threshold;
    // |snoise3| <= 1.1863: its value is 32 times the sum of the four corner terms of one
    // simplex cell, each at most (0.6 - r^2)^4 * r * |g| at its corner distance r
    // (|g| <= 1.4253 for the decoded gradients); the largest sum over a cell, found on a
    // 480-step grid of the cell plus a Lipschitz margin (tests/test_oracle.py recomputes
    // it), gives the bound.  1.19 leaves room for FP32 rounding.
    double amp = 1.0, sum = 0.0;
    for (int i = 0; i < d.cave_octaves; ++i) {
        sum += std::fabs(amp);
        amp *= (double)d.cave_gain;
    }
    f.cave_cut = (float)((1.19 * sum + 1e-3) * 1.001 - (double)d.cave_threshold);
    return f;
}

lvn_field_desc default_field() {
    lvn_field_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = LVN_FIELD_HEIGHTFIELD;
    d.height_octaves = 6;
    d.height_frequency = 1.f / 128.f;
    d.height_lacunarity = 2.f;
    d.height_gain = 0.5f;
    d.height_amplitude = 24.f;
    d.cave_octaves = 5;
    d.cave_frequency = 1.f / 64.f;
    d.cave_lacunarity = 2.f;
    d.cave_gain = 0.5f;
    d.cave_threshold = 0.3f;
    return d;
}

void lod_colour(int size, float* rgb) {  // ColourForMinLeafSize, src/clipmap.cpp:329-352
    switch (size / kClipmapLeafSize) {
    case 1: rgb[0] = 0.3f; rgb[1] = 0.1f; rgb[2] = 0.f; break;
    case 2: rgb[0] = 0.f; rgb[1] = 0.f; rgb[2] = 0.5f; break;
    case 4: rgb[0] = 0.f; rgb[1] = 0.5f; rgb[2] = 0.5f; break;
    case 8: rgb[0] = 0.5f; rgb[1] = 0.f; rgb[2] = 0.5f; break;
    case 16: rgb[0] = 0.f; rgb[1] = 0.5f; rgb[2] = 0.f; break;
    default: rgb[0] = 0.5f; rgb[1] = 0.f; rgb[2] = 0.f; break;
    }
}

inline uint64_t align256(uint64_t v) { return (v + 255) & ~(uint64_t)255; }

}  // namespace

DevOp make_dev_op(const lvn_csg_op& op) {
    DevOp d;
    std::memset(&d, 0, sizeof(d));
    d.type = op.type;
    d.brush = op.brush;
    d.material = op.material;
    d.ox = op.origin[0];
    d.oy = op.origin[1];
    d.oz = op.origin[2];
    d.dx = op.dims[0];
    d.dy = op.dims[1];
    d.dz = op.dims[2];
    d.c = cosf(op.rotate_y);
    d.s = sinf(op.rotate_y);
    d.radius = op.brush == 1 ? std::fabs(op.dims[0])
                             : std::sqrt(op.dims[0] * op.dims[0] + op.dims[1] * op.dims[1] + op.dims[2] * op.dims[2]);
    return d;
}

Geometry make_geometry(int n) {
    Geometry G;
    G.N = n;
    G.H = n + 1;
    G.F = n + 2;
    G.depth = 0;
    while ((1 << G.depth) < n) ++G.depth;
    G.shift = G.depth + 1;
    G.field_samples = G.F * G.F * G.F;
    G.edge_slots = 3 * G.H * G.H * G.H;
    G.field_rows = G.F * G.F;
    G.edge_rows = G.H * G.H;
    G.voxels = n * n * n;
    G.voxel_segs = std::max(1, G.voxels / 64);
    G.magic_h = (uint32_t)((((uint64_t)1 << 32) + G.H - 1) / G.H);
    G.magic_f = (uint32_t)((((uint64_t)1 << 32) + G.F - 1) / G.F);
    G.col_groups = (G.F * G.F + 63) / 64;
    return G;
}

Caps make_caps(const Geometry& G, int edge_cap, int voxel_cap) {
    Caps c;
    c.edges = edge_cap > 0 ? std::min(edge_cap, G.edge_slots) : G.edge_slots;
    c.voxels = voxel_cap > 0 ? std::min(voxel_cap, G.voxels) : G.voxels;
    c.quads = 3 * c.voxels;
    c.seams = c.voxels;
    c.seam_segs = (c.voxels + 63) / 64;
    return c;
}

SlotLayout make_layout(const Geometry& G, const Caps& c) {
    SlotLayout L;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = align256(o + bytes);
        return at;
    };
    const uint64_t fs = (uint64_t)G.field_samples;
    L.density = take(fs * 4);
    L.mat = take(fs);
    L.orig = take(fs);
    L.solid = take((uint64_t)G.field_rows * 16);
    L.planes = take((uint64_t)G.col_groups * G.F * 8);
    L.e_rows = take((uint64_t)G.edge_rows * 32);
    L.e_count = take((uint64_t)G.edge_rows * 4);
    L.e_base = take((uint64_t)G.edge_rows * 4);
    L.e_count_y = take((uint64_t)G.edge_rows * 4);
    L.e_base_y = take((uint64_t)G.edge_rows * 4);
    L.e_ids = take((uint64_t)c.edges * 4);
    L.e_data = take((uint64_t)c.edges * 16);
    L.e_order = take((uint64_t)c.edges * 4);
    L.h_lines = take((uint64_t)2 * G.H * G.H * kLineSamples * 4);
    L.h_cols = take((uint64_t)G.F * G.F * 4);
    L.v_mask = take((uint64_t)G.voxel_segs * 8);
    L.v_count = take((uint64_t)G.voxel_segs * 4);
    L.v_base = take((uint64_t)G.voxel_segs * 4);
    L.v_code = take((uint64_t)c.voxels * 4);
    L.v_emask = take((uint64_t)c.voxels * 4);
    L.v_mat = take((uint64_t)c.voxels * 4);
    L.q_count = take((uint64_t)c.seam_segs * 4);
    L.q_base = take((uint64_t)c.seam_segs * 4);
    L.tris = take((uint64_t)c.quads * 24);
    L.s_mask = take((uint64_t)c.seam_segs * 8);
    L.s_count = take((uint64_t)c.seam_segs * 4);
    L.s_base = take((uint64_t)c.seam_segs * 4);
    L.seams = take((uint64_t)c.seams * 48);
    L.verts = take((uint64_t)c.voxels * 48);
    L.slot_bytes = o;
    return L;
}

void fill_chunk_desc(ChunkDesc& cd, const int32_t* min, int32_t size, int N) {
    std::memset(&cd, 0, sizeof(cd));
    for (int i = 0; i < 3; ++i) {
        cd.min[i] = min[i];
        cd.off[i] = min[i] / kLeafSizeScale;  // LeafScaleVec, src/compute.cpp:560-568
    }
    cd.size = size;
    cd.scale = size / (N * kLeafSizeScale);  // src/compute_density_field.cpp:149
    lod_colour(size, cd.colour);
}

bool aabb_overlaps(const int32_t* mn, int32_t size, const std::array<int32_t, 6>& bb) {  // src/aabb.h:26-35
    return !(mn[0] + size < bb[0] || mn[1] + size < bb[1] || mn[2] + size < bb[2] || mn[0] > bb[3] ||
             mn[1] > bb[4] || mn[2] > bb[5]);
}

int ensure_global() { return g_state.init ? LVN_OK : LVN_ERR_STATE; }

static void free_tables_locked() {
    if (g_state.d_perm) (void)hipFree(g_state.d_perm);
    if (g_state.d_grads) (void)hipFree(g_state.d_grads);
    g_state.d_perm = nullptr;
    g_state.d_grads = nullptr;
}

void retain_global() {
    std::lock_guard<std::mutex> lock(g_state.mu);
    ++g_state.live;
}

void release_global() {
    std::lock_guard<std::mutex> lock(g_state.mu);
    if (--g_state.live == 0 && !g_state.init) free_tables_locked();
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

int Ctx::allocate(int n, uint32_t fl, int nslots, int edge_cap, int voxel_cap) {
    if (!retained) {
        retain_global();
        retained = true;
    }
    N = n;
    flags = fl;
    G = make_geometry(n);
    caps = make_caps(G, edge_cap, voxel_cap);
    L = make_layout(G, caps);
    slots = nslots;
    LVN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    LVN_HIP(hipMalloc(&d_arena, L.slot_bytes * (uint64_t)slots));
    LVN_HIP(hipMalloc(&d_chunks, sizeof(ChunkDesc) * slots));
    LVN_HIP(hipMalloc(&d_counters, sizeof(Counters) * slots));
    LVN_HIP(hipHostMalloc(&h_chunks, sizeof(ChunkDesc) * slots, hipHostMallocDefault));
    LVN_HIP(hipHostMalloc(&h_counters, sizeof(Counters) * slots, hipHostMallocDefault));
    LVN_HIP(hipMalloc(&d_queue, sizeof(int) * kNumQueues));
    LVN_HIP(hipMalloc(&d_item_prefix, sizeof(int) * (slots + 1)));
    int dev = 0;
    LVN_HIP(hipGetDevice(&dev));
    LVN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return LVN_OK;
}

int Ctx::ensure_ops(int nops, int nidx) {
    if (nops > ops_cap) {
        if (d_ops) LVN_HIP(hipFree(d_ops));
        if (h_ops) LVN_HIP(hipHostFree(h_ops));
        ops_cap = std::max(nops, 2 * ops_cap);
        LVN_HIP(hipMalloc(&d_ops, sizeof(DevOp) * ops_cap));
        LVN_HIP(hipHostMalloc(&h_ops, sizeof(DevOp) * ops_cap, hipHostMallocDefault));
    }
    if (nidx > idx_cap) {
        if (d_op_index) LVN_HIP(hipFree(d_op_index));
        if (h_op_index) LVN_HIP(hipHostFree(h_op_index));
        idx_cap = std::max(nidx, 2 * idx_cap);
        LVN_HIP(hipMalloc(&d_op_index, sizeof(int) * idx_cap));
        LVN_HIP(hipHostMalloc(&h_op_index, sizeof(int) * idx_cap, hipHostMallocDefault));
    }
    return LVN_OK;
}

KArgs make_kargs(char* arena, const SlotLayout& L, const Geometry& G, const Caps& caps, const ChunkDesc* chunks,
                 Counters* counters, const DevOp* ops, const int* op_index, uint32_t flags, int* queue,
                 int* item_prefix, int cus) {
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.arena = arena;
    a.L = L;
    a.G = G;
    a.caps = caps;
    a.chunks = chunks;
    a.counters = counters;
    a.ops = ops;
    a.op_index = op_index;
    a.F = g_state.fp;
    a.perm = g_state.d_perm;
    a.grads = g_state.d_grads;
    a.defmat = (int)g_state.defmat;
    a.fp16 = (flags & LVN_CTX_FP16_DENSITY) ? 1 : 0;
    a.ysplit = 1;
    a.precise = (flags & LVN_CTX_PRECISE_QEF) ? 1 : 0;
    a.queue = queue;
    a.item_prefix = item_prefix;
    a.cus = cus;
    a.persist_per_cu = kDefaultPersistPerCu;
    a.density_per_cu = kDefaultDensityPerCu;
    a.height_lines = 0;
    a.row_blocks = kDefaultRowBlocks;
    a.voxel_blocks = kDefaultVoxelBlocks;
    return a;
}

KArgs Ctx::kargs() const {
    return make_kargs(d_arena, L, G, caps, d_chunks, d_counters, d_ops, d_op_index, flags, d_queue, d_item_prefix,
                      cus);
}

Ctx::~Ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_verts) (void)hipHostFree(h_verts);
    if (seams) destroy_seam_state(seams);
    if (simp) destroy_simp_state(simp);
    if (d_arena) (void)hipFree(d_arena);
    if (d_chunks) (void)hipFree(d_chunks);
    if (d_counters) (void)hipFree(d_counters);
    if (d_ops) (void)hipFree(d_ops);
    if (d_op_index) (void)hipFree(d_op_index);
    if (d_queue) (void)hipFree(d_queue);
    if (d_item_prefix) (void)hipFree(d_item_prefix);
    if (h_chunks) (void)hipHostFree(h_chunks);
    if (h_counters) (void)hipHostFree(h_counters);
    if (h_ops) (void)hipHostFree(h_ops);
    if (h_op_index) (void)hipHostFree(h_op_index);
    if (stream) (void)hipStreamDestroy(stream);
    field_pool.release();
    mesh_pool.base.release();
    if (retained) release_global();  // after every stream of this context has drained
}

void* DevicePool::take(size_t bytes) {
    bytes = align256(bytes);
    if (bytes > left) {
        const size_t block = std::max(bytes, (size_t)64 << 20);
        void* p = nullptr;
        if (hipMalloc(&p, block) != hipSuccess) return nullptr;
        blocks.push_back(p);
        cur = static_cast<char*>(p);
        left = block;
    }
    void* out = cur;
    cur += bytes;
    left -= bytes;
    return out;
}

void DevicePool::release() {
    for (void* p : blocks) (void)hipFree(p);
    blocks.clear();
    cur = nullptr;
    left = 0;
}

DevicePool::~DevicePool() { release(); }

char* ClassPool::take(size_t bytes, size_t* cap) {
    size_t c = (size_t)64 << 10;
    while (c < bytes) c <<= 1;
    *cap = c;
    auto& list = free_blocks[c];
    if (!list.empty()) {
        char* p = list.back();
        list.pop_back();
        return p;
    }
    return static_cast<char*>(base.take(c));
}

// Upload one chunk descriptor + its op batch to slot 0.
static int upload_single(Ctx* c, const ChunkDesc& cd_in, const std::vector<lvn_csg_op>& ops) {
    const int n = (int)ops.size();
    int r = c->ensure_ops(std::max(n, 1), std::max(n, 1));
    if (r) return r;
    for (int i = 0; i < n; ++i) {
        c->h_ops[i] = make_dev_op(ops[i]);
        c->h_op_index[i] = i;
    }
    c->h_chunks[0] = cd_in;
    c->h_chunks[0].op_start = 0;
    c->h_chunks[0].op_count = n;
    if (n) {
        LVN_HIP(hipMemcpyAsync(c->d_ops, c->h_ops, sizeof(DevOp) * n, hipMemcpyHostToDevice, c->stream));
        LVN_HIP(hipMemcpyAsync(c->d_op_index, c->h_op_index, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    }
    LVN_HIP(hipMemcpyAsync(c->d_chunks, c->h_chunks, sizeof(ChunkDesc), hipMemcpyHostToDevice, c->stream));
    launch_clear_counters(c->kargs(), 1, c->stream);
    return LVN_OK;
}

// Cached mode of a chunk descriptor: materials and edge list of the cached field.
static void set_cached(ChunkDesc& cd, const FieldEntry* cached) {
    if (!cached) return;
    cd.mode = kModeCached;
    cd.old_edges = cached->edges;
    cd.old_ids = (uint64_t)(uintptr_t)cached->d_ids;
    cd.old_data = (uint64_t)(uintptr_t)cached->d_data;
    cd.old_mat = (uint64_t)(uintptr_t)cached->d_mat;
}

// Launch shapes of the one-chunk forms: split density columns 6 ways.
static KArgs single_args(const Ctx* c) {
    KArgs a = c->kargs();
    a.ysplit = 6;
    return a;
}

// The whole pipeline for slot 0 as one graph launch (captured once, again when the op
// arrays move).
static int run_pipeline_graph(Ctx* c) {
    if (!c->exec || c->graph_ops != c->d_ops || c->graph_index != c->d_op_index ||
        c->graph_gen != g_state.generation) {
        if (c->exec) LVN_HIP(hipGraphExecDestroy(c->exec));
        if (c->graph) LVN_HIP(hipGraphDestroy(c->graph));
        c->exec = nullptr;
        c->graph = nullptr;
        LVN_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_pipeline(single_args(c), 1, c->stream);
        LVN_HIP(hipStreamEndCapture(c->stream, &c->graph));
        LVN_HIP(hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
        c->graph_ops = c->d_ops;
        c->graph_index = c->d_op_index;
        c->graph_gen = g_state.generation;
    }
    LVN_HIP(hipGraphLaunch(c->exec, c->stream));
    return LVN_OK;
}

// Field stages for slot 0: density (fresh) or cached materials, CSG batch, edges, crossings.
static int run_field(Ctx* c, const int32_t* min, int32_t size, const std::vector<lvn_csg_op>& ops,
                     const FieldEntry* cached) {
    ChunkDesc cd;
    fill_chunk_desc(cd, min, size, c->N);
    set_cached(cd, cached);
    int r = upload_single(c, cd, ops);
    if (r) return r;
    const KArgs a = single_args(c);
    launch_density(a, 1, c->stream);  // or the cached field's materials
    if (!ops.empty()) launch_csg(a, 1, c->stream);
    launch_edge_rows(a, 1, c->stream);
    launch_scan(a, 1, kScanEdges, c->stream);
    launch_edge_compact(a, 1, c->stream);
    launch_edge_cross(a, 1, c->stream);
    LVN_HIP(hipGetLastError());
    return LVN_OK;
}

static int run_mesh(Ctx* c) {
    KArgs a = c->kargs();
    launch_voxel_rows(a, 1, c->stream);
    launch_scan(a, 1, kScanVoxels, c->stream);
    launch_voxel_compact(a, 1, c->stream);
    launch_qef(a, 1, c->stream);
    launch_scan(a, 1, kScanQuadsSeams, c->stream);
    launch_emit(a, 1, c->stream);
    LVN_HIP(hipGetLastError());
    return LVN_OK;
}

static int read_counters(Ctx* c, Counters* out) {
    LVN_HIP(hipMemcpyAsync(c->h_counters, c->d_counters, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    *out = c->h_counters[0];
    return LVN_OK;
}

// Copy slot 0's materials + edge list into a cache entry (StoreDensityField,
// src/compute_density_field.cpp:303-308).
static int store_field(Ctx* c, const Key& key, int edges, int last_op) {
    FieldEntry& f = c->fields[key];
    if (!f.d_mat) f.d_mat = static_cast<uint8_t*>(c->field_pool.take(c->G.field_samples));
    if (!f.d_ids || edges > f.cap) {  // a grown edge list moves to new pool space (the old stays with the pool)
        f.cap = std::max(edges + edges / 4, 64);
        f.d_ids = static_cast<int*>(c->field_pool.take(sizeof(int) * (size_t)f.cap));
        f.d_data = static_cast<float4*>(c->field_pool.take(16 * (size_t)f.cap));
    }
    if (!f.d_mat || !f.d_ids || !f.d_data) return LVN_ERR_NOMEM;
    char* slot = c->d_arena;
    LVN_HIP(hipMemcpyAsync(f.d_mat, slot + c->L.mat, c->G.field_samples, hipMemcpyDeviceToDevice, c->stream));
    if (edges) {
        LVN_HIP(hipMemcpyAsync(f.d_ids, slot + c->L.e_ids, sizeof(int) * edges, hipMemcpyDeviceToDevice, c->stream));
        LVN_HIP(hipMemcpyAsync(f.d_data, slot + c->L.e_data, 16 * (size_t)edges, hipMemcpyDeviceToDevice, c->stream));
    }
    f.edges = edges;
    f.last_op = last_op;
    return LVN_OK;  // the copies run on the context's stream before any later use of the slot
}

static std::vector<lvn_csg_op> stored_ops_for(const int32_t* min, int32_t size, int from) {
    std::vector<lvn_csg_op> out;
    for (size_t i = (size_t)std::max(from, 0); i < g_state.ops.size(); ++i)
        if (aabb_overlaps(min, size, g_state.aabbs[i])) out.push_back(g_state.ops[i]);
    return out;
}

// LoadDensityField (src/compute_density_field.cpp:235-274) into slot 0:
// cached or fresh field, then the stored ops recorded since the field's last
// update that overlap it; the field is cached only when ops were applied.
static int load_field(Ctx* c, const int32_t* min, int32_t size, Counters* cnt) {
    const Key key = make_key(min, size);
    auto it = c->fields.find(key);
    const FieldEntry* cached = it == c->fields.end() ? nullptr : &it->second;
    const int from = cached ? cached->last_op : 0;
    const std::vector<lvn_csg_op> ops = stored_ops_for(min, size, from);
    int r = run_field(c, min, size, ops, cached);
    if (r) return r;
    r = read_counters(c, cnt);
    if (r) return r;
    if (cnt->status) return LVN_ERR_OVERFLOW;
    if (!ops.empty()) {
        r = store_field(c, key, cnt->edges, (int)g_state.ops.size());
        if (r) return r;
    }
    return LVN_OK;
}

// Keep the mesh of slot 0 in the context's mesh pool (device to device, asynchronous).
static int cache_mesh(Ctx* c, const Counters& k, MeshEntry* m) {
    m->V = k.quads > 0 ? k.voxels : 0;  // the export rule, src/compute_octree.cpp:227-232
    m->T = 2 * k.quads;
    m->S = k.seams;
    const size_t bv = 48 * (size_t)m->V, bt = 12 * (size_t)m->T, bs = 48 * (size_t)m->S;
    m->d = c->mesh_pool.take(std::max<size_t>(bv + bt + bs, 1), &m->cap);
    if (!m->d) return LVN_ERR_NOMEM;
    const char* slot = c->d_arena;
    if (bv) LVN_HIP(hipMemcpyAsync(m->d, slot + c->L.verts, bv, hipMemcpyDeviceToDevice, c->stream));
    if (bt) LVN_HIP(hipMemcpyAsync(m->d + bv, slot + c->L.tris, bt, hipMemcpyDeviceToDevice, c->stream));
    if (bs) LVN_HIP(hipMemcpyAsync(m->d + bv + bt, slot + c->L.seams, bs, hipMemcpyDeviceToDevice, c->stream));
    return LVN_OK;
}

// ExportMeshBuffer (src/compute_octree.cpp:326-347): a cached mesh into the caller's
// buffers through pinned staging (one read-back, one host copy per array).
static int export_mesh(Ctx* c, const MeshEntry& m, lvn_vertex* v, int32_t* tri, lvn_seam_node* s) {
    const size_t bv = 48 * (size_t)m.V, bt = 12 * (size_t)m.T, bs = 48 * (size_t)m.S;
    if (bv + bt + bs == 0) return LVN_OK;
    if (!c->h_verts) {  // staging sized for the largest mesh a slot can hold
        LVN_HIP(hipHostMalloc(&c->h_verts, 48 * (size_t)c->caps.voxels + 24 * (size_t)c->caps.quads +
                                               48 * (size_t)c->caps.seams, hipHostMallocDefault));
    }
    char* h = reinterpret_cast<char*>(c->h_verts);
    LVN_HIP(hipMemcpyAsync(h, m.d, bv + bt + bs, hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    if (bv && v) std::memcpy(v, h, bv);
    if (bt && tri) std::memcpy(tri, h + bv, bt);
    if (bs && s) std::memcpy(s, h + bv + bt, bs);
    return LVN_OK;
}

Batch::~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_work) (void)hipFree(d_work);
}

// Stage chunk descriptors and the AABB-filtered stored-op lists of `n` requests.
static int stage_batch(Batch* b, int n, const lvn_chunk_req* reqs) {
    Ctx* c = &b->ctx;
    std::vector<int> index;
    index.reserve((size_t)n * 4);
    const int nops = (int)g_state.ops.size();
    int r = c->ensure_ops(std::max(nops, 1), 1);
    if (r) return r;
    for (int i = 0; i < nops; ++i) c->h_ops[i] = make_dev_op(g_state.ops[i]);
    for (int i = 0; i < n; ++i) {
        ChunkDesc& cd = c->h_chunks[i];
        fill_chunk_desc(cd, reqs[i].min, reqs[i].size, c->N);
        cd.op_start = (int)index.size();
        for (int k = 0; k < nops; ++k)
            if (aabb_overlaps(reqs[i].min, reqs[i].size, g_state.aabbs[k])) index.push_back(k);
        cd.op_count = (int)index.size() - cd.op_start;
    }
    r = c->ensure_ops(std::max(nops, 1), std::max((int)index.size(), 1));
    if (r) return r;
    for (int i = 0; i < nops; ++i) c->h_ops[i] = make_dev_op(g_state.ops[i]);
    if (!index.empty()) std::memcpy(c->h_op_index, index.data(), index.size() * sizeof(int));
    if (nops) LVN_HIP(hipMemcpyAsync(c->d_ops, c->h_ops, sizeof(DevOp) * nops, hipMemcpyHostToDevice, c->stream));
    if (!index.empty())
        LVN_HIP(hipMemcpyAsync(c->d_op_index, c->h_op_index, sizeof(int) * index.size(), hipMemcpyHostToDevice,
                               c->stream));
    LVN_HIP(hipMemcpyAsync(c->d_chunks, c->h_chunks, sizeof(ChunkDesc) * n, hipMemcpyHostToDevice, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    b->n = n;
    b->simplified = false;
    b->slot_of.clear();
    for (int i = 0; i < n; ++i) b->slot_of[make_key(reqs[i].min, reqs[i].size)] = i;
    return LVN_OK;
}

// The batch's launch shapes with the automatic column split resolved: small batches
// split the density columns to fill the GPU.
static lvn_tuning effective_tuning(const Batch* b) {
    lvn_tuning t = b->tuning;
    if (t.density_ysplit == 0) t.density_ysplit = b->n >= 64 ? 1 : 6;
    if (t.cross_lanes == 0) t.cross_lanes = default_cross_lanes(b->n);
    if (t.density_per_cu == 0) t.density_per_cu = kDefaultDensityPerCu;
    if (t.height_lines == 0) t.height_lines = b->n >= kHeightLinesMinChunks ? 2 : 1;
    return t;
}

// A batch exports meshes, never its density field: for a field without the cave term
// it leaves the density unstored (KArgs::dens_cols), and chunks without brushes leave
// their materials unstored (KArgs::mat_free).
static int density_from_columns(const KArgs& a) {
    return a.F.kind != LVN_FIELD_TERRAIN_CAVES && !a.fp16 ? 1 : 0;
}

static KArgs batch_args(const Batch* b, const lvn_tuning& t) {
    KArgs a = b->ctx.kargs();
    a.dens_cols = density_from_columns(a);
    a.mat_free = 1;
    a.ysplit = t.density_ysplit;
    a.persist_per_cu = t.persistent_per_cu;
    a.density_per_cu = t.density_per_cu;
    a.height_lines = t.height_lines == 2 ? 1 : 0;
    a.row_blocks = t.row_blocks;
    a.voxel_blocks = t.voxel_blocks;
    a.cross_lanes = t.cross_lanes;
    return a;
}

static KArgs batch_args(const Batch* b) { return batch_args(b, effective_tuning(b)); }

// One pass of the pipeline without the graph, each stage bracketed by HIP events.
static int time_stages(Batch* b, const KArgs& a, float* stage_ms) {
    Ctx* c = &b->ctx;
    b->simplified = false;
    hipEvent_t ev[kNumStages + 1];
    for (auto& e : ev) LVN_HIP(hipEventCreate(&e));
    launch_clear_counters(a, std::max(b->n, 1), c->stream);
    for (int i = 0; i < kNumStages; ++i) {
        LVN_HIP(hipEventRecord(ev[i], c->stream));
        if (b->n) launch_stage(a, b->n, i, c->stream);
    }
    LVN_HIP(hipEventRecord(ev[kNumStages], c->stream));
    LVN_HIP(hipEventSynchronize(ev[kNumStages]));
    LVN_HIP(hipGetLastError());
    for (int i = 0; i < kNumStages; ++i) LVN_HIP(hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]));
    for (auto& e : ev) LVN_HIP(hipEventDestroy(e));
    return LVN_OK;
}

bool tuning_valid(const lvn_tuning& t, int F) {
    return t.density_ysplit >= 0 && t.density_ysplit <= F && t.persistent_per_cu >= 1 &&
           t.persistent_per_cu <= 2 && t.row_blocks >= 1 && t.row_blocks <= (1 << 20) && t.voxel_blocks >= 1 &&
           t.voxel_blocks <= (1 << 20) &&
           (t.cross_lanes == 0 || t.cross_lanes == 1 || t.cross_lanes == 2 || t.cross_lanes == 4 ||
            t.cross_lanes == 8 || t.cross_lanes == kCrossSampleParallel) &&
           t.density_per_cu >= 0 && t.density_per_cu <= 2 && t.height_lines >= 0 && t.height_lines <= 2;
}

static bool valid_tuning(const Batch* b, const lvn_tuning& t) { return tuning_valid(t, b->ctx.G.F); }

static void drop_graph(Batch* b) { b->graph_n = -1; }

}  // namespace lvn

using namespace lvn;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace lvn {
// The kernels' copy of the permutation table: values plus one, rows of a wrap column,
// texels 0..255 and a second wrap column, and a wrap row.
static int upload_perm_table() {
    std::vector<uint8_t> pad(kPermPaddedBytes, 0);
    auto stored = [](int x, int y) { return (uint8_t)((g_state.perm[((y & 255) << 8) | (x & 255)] + 1) & 255); };
    for (int y = 0; y < kPermRows; ++y) {
        pad[y * kPermPitch] = stored(0, y);
        for (int x = 0; x <= 256; ++x) pad[y * kPermPitch + 1 + x] = stored(x, y);
    }
    LVN_HIP(hipMemcpy(g_state.d_perm, pad.data(), kPermPaddedBytes, hipMemcpyHostToDevice));
    return LVN_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free single-chunk form (lvn_generate_chunk_mesh_noalloc)
// ---------------------------------------------------------------------------

// The device area of the allocation-free form: one chunk slot at the form's caps, its
// descriptor, counters, op list and queue heads.  Each GPU holds its own copy, at its own
// address; host code uses that address (noalloc_ptr), never this name.
static __device__ __align__(256) char scratch[LVN_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    char* ptr = nullptr;  // address of `scratch` on the device
    int cus = 0;
};
static NoAllocDevice noalloc_dev[kMaxDevices];  // guarded by g_state.mu
// host staging of one call (guarded by g_state.mu)
static DevOp noalloc_h_ops[LVN_NOALLOC_MAX_OPS];
static int noalloc_h_index[LVN_NOALLOC_MAX_OPS];

// Locates `scratch` on the current device once (called with g_state.mu held).
static int noalloc_setup_locked(int* dev_out) {
    int dev = 0;
    LVN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return LVN_ERR_DEVICE;
    NoAllocDevice& d = noalloc_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        LVN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        LVN_HIP(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev));
        d.ptr = static_cast<char*>(p);
    }
    *dev_out = dev;
    return LVN_OK;
}

// Clears n16 16-byte words (the area keeps the previous call's contents).
__global__ void k_noalloc_clear(uint4* p, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Offsets of the parts of the area for one chunk of N voxels per side.
struct NoAllocLayout {
    Geometry G;
    Caps caps;
    SlotLayout L;
    uint64_t chunk, counters, ops, index, queue, prefix, bytes;
};

static NoAllocLayout noalloc_layout(int n) {
    NoAllocLayout t;
    t.G = make_geometry(n);
    t.caps = make_caps(t.G, LVN_NOALLOC_MAX_EDGES, LVN_NOALLOC_MAX_VOXELS);
    t.L = make_layout(t.G, t.caps);
    uint64_t o = t.L.slot_bytes;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = align256(o + bytes);
        return at;
    };
    t.chunk = take(sizeof(ChunkDesc));
    t.counters = take(sizeof(Counters));
    t.ops = take(sizeof(DevOp) * LVN_NOALLOC_MAX_OPS);
    t.index = take(sizeof(int) * LVN_NOALLOC_MAX_OPS);
    t.queue = take(sizeof(int) * kNumQueues);
    t.prefix = take(sizeof(int) * 2);
    t.bytes = o;
    return t;
}
}  // namespace lvn

extern "C" {

int lvn_init(const lvn_field_desc* field, int32_t noise_seed, uint32_t default_material) {
    if (default_material > 255) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_state.mu);
    const lvn_field_desc d = field ? *field : default_field();
    if (d.kind < 0 || d.kind > 2 || d.height_octaves < 0 || d.cave_octaves < 0)
        return LVN_ERR_INVALID;  // nothing changed
    g_state.desc = d;
    g_state.fp = make_field_params(g_state.desc);
    g_state.defmat = default_material;
    // kernels still in flight on any stream read the tables: let them finish first
    LVN_HIP(hipDeviceSynchronize());
    if (!g_state.d_perm) {
        LVN_HIP(hipMalloc(&g_state.d_perm, kPermPaddedBytes));
        LVN_HIP(hipMalloc(&g_state.d_grads, sizeof(float4) * kGradEntries));
    }
    g_state.perm.resize(65536);
    build_perm_table(noise_seed, g_state.perm.data());
    float4 grads[kGradEntries];
    build_grads(grads);
    int r = upload_perm_table();
    if (r) return r;
    LVN_HIP(hipMemcpy(g_state.d_grads, grads, sizeof(grads), hipMemcpyHostToDevice));
    g_state.seed = noise_seed;
    g_state.init = true;
    ++g_state.generation;
    int dev = 0;
    return noalloc_setup_locked(&dev);  // the allocation-free form's area on this device
}

int lvn_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_state.mu);
    g_state.ops.clear();
    g_state.aabbs.clear();
    g_state.init = false;
    ++g_state.generation;
    // Every launching entry point now returns LVN_ERR_STATE; contexts, batches and
    // servers still alive keep the tables until the last of them is destroyed.
    if (g_state.live == 0) free_tables_locked();
    return LVN_OK;
}

int lvn_set_noise_seed(int32_t seed) {
    std::lock_guard<std::mutex> lock(g_state.mu);
    if (!g_state.init) return LVN_ERR_STATE;
    // Work already enqueued (an asynchronous lvn_batch_run, server batches) runs with the
    // old table, as on the reference's in-order queue: drain it before rewriting the table.
    LVN_HIP(hipDeviceSynchronize());
    build_perm_table(seed, g_state.perm.data());
    const int r = upload_perm_table();
    if (r) return r;
    g_state.seed = seed;
    return LVN_OK;
}

int lvn_store_csg_op(const lvn_csg_op* op, const int32_t aabb_min[3], const int32_t aabb_max[3]) {
    if (!op || !aabb_min || !aabb_max) return LVN_ERR_INVALID;
    if (op->material < 0 || op->material > 255 || op->type < 0 || op->type > 1 || op->brush < 0 || op->brush > 1)
        return LVN_ERR_INVALID;
    g_state.ops.push_back(*op);
    g_state.aabbs.push_back({aabb_min[0], aabb_min[1], aabb_min[2], aabb_max[0], aabb_max[1], aabb_max[2]});
    return LVN_OK;
}

int lvn_clear_csg_ops(void) {
    g_state.ops.clear();
    g_state.aabbs.clear();
    return LVN_OK;
}

int lvn_get_perm_table(uint8_t* out) {
    if (!out) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    std::memcpy(out, g_state.perm.data(), 65536);
    return LVN_OK;
}

int lvn_ctx_create(int32_t voxels_per_chunk, uint32_t flags, lvn_ctx** out) {
    if (!out) return LVN_ERR_INVALID;
    *out = nullptr;
    if (!g_state.init) return LVN_ERR_STATE;
    const int n = voxels_per_chunk;
    if (n < 4 || n > 64 || (n & (n - 1))) return LVN_ERR_INVALID;
    Ctx* c = new Ctx;
    const int r = c->allocate(n, flags, 1, 0, 0);
    if (r) {
        delete c;
        return r == LVN_ERR_DEVICE ? LVN_ERR_NOMEM : r;
    }
    *out = reinterpret_cast<lvn_ctx*>(c);
    return LVN_OK;
}

void lvn_ctx_destroy(lvn_ctx* ctx) { delete reinterpret_cast<Ctx*>(ctx); }

int lvn_ctx_voxels_per_chunk(const lvn_ctx* ctx) {
    return ctx ? reinterpret_cast<const Ctx*>(ctx)->N : 0;
}

int lvn_free_chunk(lvn_ctx* ctx, const int32_t min[3], int32_t size) {
    if (!ctx || !min) return LVN_ERR_INVALID;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    auto it = c->meshes.find(make_key(min, size));  // Compute_FreeChunkOctree erases the octree only
    if (it != c->meshes.end()) {
        c->mesh_pool.give(it->second.d, it->second.cap);
        c->meshes.erase(it);
    }
    return LVN_OK;
}

int lvn_chunk_has_surface(lvn_ctx* ctx, const int32_t min[3], int32_t size, int32_t* out) {
    if (!ctx || !min || !out) return LVN_ERR_INVALID;
    if (!valid_chunk_size(size, reinterpret_cast<Ctx*>(ctx)->N)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    const Key key = make_key(min, size);
    auto it = c->fields.find(key);
    if (it != c->fields.end()) {
        *out = it->second.edges > 0;
        return LVN_OK;
    }
    // Compute_ChunkIsEmpty (src/compute_density_field.cpp:278-299): default field, no
    // stored ops, and the field is cached (with no ops recorded as applied).
    int r = run_field(c, min, size, {}, nullptr);
    if (r) return r;
    Counters k;
    r = read_counters(c, &k);
    if (r) return r;
    if (k.status) return LVN_ERR_OVERFLOW;
    r = store_field(c, key, k.edges, 0);
    if (r) return r;
    *out = k.edges > 0;
    return LVN_OK;
}

int lvn_apply_csg(lvn_ctx* ctx, const lvn_csg_op* ops, int32_t n, const int32_t min[3], int32_t size) {
    if (!ctx || !min || n < 0 || (n > 0 && !ops)) return LVN_ERR_INVALID;
    if (!valid_chunk_size(size, reinterpret_cast<Ctx*>(ctx)->N)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    for (int i = 0; i < n; ++i)
        if (ops[i].material < 0 || ops[i].material > 255 || ops[i].type < 0 || ops[i].type > 1 ||
            ops[i].brush < 0 || ops[i].brush > 1)
            return LVN_ERR_INVALID;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    // Compute_ApplyCSGOperations (src/compute_csg.cpp:224-242): LoadDensityField, apply, store.
    Counters k;
    int r;
    const Key key = make_key(min, size);
    auto fit = c->fields.find(key);
    if (fit != c->fields.end() && stored_ops_for(min, size, fit->second.last_op).empty()) {
        // the cached field already is what LoadDensityField would produce (no stored op
        // newer than it overlaps the chunk): loading it again would rebuild the same
        // materials and edge list
        fit->second.last_op = (int)g_state.ops.size();
    } else {
        r = load_field(c, min, size, &k);
        if (r) return r;
        r = store_field(c, key, k.edges, (int)g_state.ops.size());  // the field of the load step (slot 0)
        if (r) return r;
    }
    if (n == 0) return LVN_OK;
    std::vector<lvn_csg_op> batch(ops, ops + n);
    r = run_field(c, min, size, batch, &c->fields[key]);
    if (r) return r;
    r = read_counters(c, &k);
    if (r) return r;
    if (k.status) return LVN_ERR_OVERFLOW;
    return store_field(c, key, k.edges, (int)g_state.ops.size() + n);
}

int lvn_generate_chunk_mesh(lvn_ctx* ctx, const int32_t min[3], int32_t size, lvn_vertex* v, int32_t v_cap,
                            int32_t* nv, int32_t* tri, int32_t tri_cap, int32_t* nt, lvn_seam_node* s,
                            int32_t s_cap, int32_t* ns) {
    if (!ctx || !min || !nv || !nt || !ns) return LVN_ERR_INVALID;
    if (!valid_chunk_size(size, reinterpret_cast<Ctx*>(ctx)->N)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    *nv = *nt = *ns = 0;
    const Key key = make_key(min, size);
    auto it = c->meshes.find(key);
    const MeshEntry* m = nullptr;
    if (it != c->meshes.end()) {
        m = &it->second;  // LoadOctree cache hit (src/compute_octree.cpp:157-163)
    } else {
        // LoadDensityField + ConstructOctreeFromField + GenerateMeshFromOctree as one
        // graph launch over slot 0, then the field is cached (when ops were applied)
        // and the mesh read back.
        auto fit = c->fields.find(key);
        const FieldEntry* cached = fit == c->fields.end() ? nullptr : &fit->second;
        const std::vector<lvn_csg_op> ops = stored_ops_for(min, size, cached ? cached->last_op : 0);
        ChunkDesc cd;
        fill_chunk_desc(cd, min, size, c->N);
        set_cached(cd, cached);
        int r = upload_single(c, cd, ops);
        if (r) return r;
        r = run_pipeline_graph(c);
        if (r) return r;
        Counters k;
        r = read_counters(c, &k);
        if (r) return r;
        if (k.status) return LVN_ERR_OVERFLOW;
        if (!ops.empty()) {
            r = store_field(c, key, k.edges, (int)g_state.ops.size());
            if (r) return r;
        }
        if (k.edges == 0) return LVN_OK;  // no octree, nothing cached (src/compute_octree.cpp:167-171)
        MeshEntry fresh;
        r = cache_mesh(c, k, &fresh);
        if (r) return r;
        m = &(c->meshes[key] = fresh);
    }
    *nv = m->V;
    *nt = m->T;
    *ns = m->S;
    if (m->V > v_cap || m->T > tri_cap || m->S > s_cap) return LVN_ERR_CAPACITY;
    return export_mesh(c, *m, v, tri, s);
}

int lvn_generate_chunk_mesh_noalloc(int32_t n, uint32_t flags, const int32_t min[3], int32_t size, lvn_vertex* v,
                                    int32_t v_cap, int32_t* nv, int32_t* tri, int32_t tri_cap, int32_t* nt,
                                    lvn_seam_node* s, int32_t s_cap, int32_t* ns) {
    if (!min || !nv || !nt || !ns) return LVN_ERR_INVALID;
    if (n < 4 || n > 64 || (n & (n - 1)) || (flags & ~(uint32_t)(LVN_CTX_FP16_DENSITY | LVN_CTX_PRECISE_QEF)))
        return LVN_ERR_INVALID;
    if (!valid_chunk_size(size, n)) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_state.mu);  // one call at a time: every call uses the one area
    if (!g_state.init) return LVN_ERR_STATE;
    *nv = *nt = *ns = 0;
    int dev = 0;
    int r = noalloc_setup_locked(&dev);
    if (r) return r;
    const NoAllocDevice& d = noalloc_dev[dev];
    const NoAllocLayout t = noalloc_layout(n);
    if (t.bytes > LVN_NOALLOC_SCRATCH_BYTES) return LVN_ERR_INVALID;  // not reached for n <= 64
    // the stored ops overlapping the chunk, in store order (LoadDensityField's filter,
    // src/compute_density_field.cpp:255-263), converted as for every other form
    int nops = 0;
    for (size_t i = 0; i < g_state.ops.size(); ++i) {
        if (!aabb_overlaps(min, size, g_state.aabbs[i])) continue;
        if (nops == LVN_NOALLOC_MAX_OPS) return LVN_ERR_OVERFLOW;
        noalloc_h_ops[nops] = make_dev_op(g_state.ops[i]);
        noalloc_h_index[nops] = nops;
        ++nops;
    }
    static ChunkDesc cd;  // host staging, guarded by g_state.mu
    static Counters k;
    fill_chunk_desc(cd, min, size, n);
    cd.op_start = 0;
    cd.op_count = nops;
    char* base = d.ptr;
    hipStream_t st = nullptr;
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1024), dim3(256), 0, st, reinterpret_cast<uint4*>(base), t.bytes / 16);
    LVN_HIP(hipGetLastError());
    if (nops) {
        LVN_HIP(hipMemcpyAsync(base + t.ops, noalloc_h_ops, sizeof(DevOp) * nops, hipMemcpyHostToDevice, st));
        LVN_HIP(hipMemcpyAsync(base + t.index, noalloc_h_index, sizeof(int) * nops, hipMemcpyHostToDevice, st));
    }
    LVN_HIP(hipMemcpyAsync(base + t.chunk, &cd, sizeof(ChunkDesc), hipMemcpyHostToDevice, st));
    KArgs a = make_kargs(base, t.L, t.G, t.caps, reinterpret_cast<const ChunkDesc*>(base + t.chunk),
                         reinterpret_cast<Counters*>(base + t.counters), reinterpret_cast<const DevOp*>(base + t.ops),
                         reinterpret_cast<const int*>(base + t.index), flags, reinterpret_cast<int*>(base + t.queue),
                         reinterpret_cast<int*>(base + t.prefix), d.cus);
    a.ysplit = 6;  // the one-chunk forms' density column split (single_args)
    launch_pipeline(a, 1, st);
    LVN_HIP(hipGetLastError());
    LVN_HIP(hipMemcpyAsync(&k, base + t.counters, sizeof(Counters), hipMemcpyDeviceToHost, st));
    LVN_HIP(hipStreamSynchronize(st));
    if (k.status) return LVN_ERR_OVERFLOW;
    if (k.edges == 0) return LVN_OK;  // no octree (src/compute_octree.cpp:167-171)
    const int V = k.quads > 0 ? k.voxels : 0, T = 2 * k.quads, S = k.seams;  // src/compute_octree.cpp:227-232
    *nv = V;
    *nt = T;
    *ns = S;
    if (V > v_cap || T > tri_cap || S > s_cap) return LVN_ERR_CAPACITY;
    if (V && v) LVN_HIP(hipMemcpyAsync(v, base + t.L.verts, sizeof(lvn_vertex) * (size_t)V, hipMemcpyDefault, st));
    if (T && tri) LVN_HIP(hipMemcpyAsync(tri, base + t.L.tris, 12 * (size_t)T, hipMemcpyDefault, st));
    if (S && s) LVN_HIP(hipMemcpyAsync(s, base + t.L.seams, sizeof(lvn_seam_node) * (size_t)S, hipMemcpyDefault, st));
    LVN_HIP(hipStreamSynchronize(st));
    return LVN_OK;
}

int lvn_batch_create(int32_t voxels_per_chunk, int32_t max_chunks, int32_t edge_cap, int32_t voxel_cap,
                     uint32_t flags, lvn_batch** out) {
    if (!out) return LVN_ERR_INVALID;
    *out = nullptr;
    if (!g_state.init) return LVN_ERR_STATE;
    const int n = voxels_per_chunk;
    if (n < 4 || n > 64 || (n & (n - 1)) || max_chunks < 1 || edge_cap < 0 || voxel_cap < 0) return LVN_ERR_INVALID;
    Batch* b = new Batch;
    int r = b->ctx.allocate(n, flags, max_chunks, edge_cap, voxel_cap);
    if (!r && hipMalloc(&b->d_work, sizeof(unsigned long long) * 8) != hipSuccess) r = LVN_ERR_NOMEM;
    if (r) {
        delete b;
        return r == LVN_ERR_DEVICE ? LVN_ERR_NOMEM : r;
    }
    *out = reinterpret_cast<lvn_batch*>(b);
    return LVN_OK;
}

void lvn_batch_destroy(lvn_batch* b) { delete reinterpret_cast<Batch*>(b); }

int lvn_batch_set_chunks(lvn_batch* bh, int32_t n, const lvn_chunk_req* reqs) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || n < 0 || (n > 0 && !reqs) || n > b->ctx.slots) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    for (int i = 0; i < n; ++i)
        if (!valid_chunk_size(reqs[i].size, b->ctx.N)) return LVN_ERR_INVALID;
    return stage_batch(b, n, reqs);
}

int lvn_batch_run(lvn_batch* bh) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;  // the library was shut down
    if (b->n == 0) return LVN_OK;
    Ctx* c = &b->ctx;
    if (b->graph_n != b->n || b->graph_ops != c->d_ops || b->graph_index != c->d_op_index ||
        b->graph_gen != g_state.generation) {
        if (b->exec) LVN_HIP(hipGraphExecDestroy(b->exec));
        if (b->graph) LVN_HIP(hipGraphDestroy(b->graph));
        b->exec = nullptr;
        b->graph = nullptr;
        const KArgs a = batch_args(b);
        LVN_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_pipeline(a, b->n, c->stream);
        LVN_HIP(hipStreamEndCapture(c->stream, &b->graph));
        LVN_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_n = b->n;
        b->graph_ops = c->d_ops;
        b->graph_index = c->d_op_index;
        b->graph_gen = g_state.generation;
    }
    b->simplified = false;
    LVN_HIP(hipGraphLaunch(b->exec, c->stream));
    return LVN_OK;
}

int lvn_batch_wait(lvn_batch* bh) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b) return LVN_ERR_INVALID;
    LVN_HIP(hipStreamSynchronize(b->ctx.stream));
    return LVN_OK;
}

const char* lvn_batch_stage_name(int32_t stage) {
    return stage >= 0 && stage < kNumStages ? kStageNames[stage] : nullptr;
}

int lvn_batch_run_timed(lvn_batch* bh, float* stage_ms, int32_t n_stages) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !stage_ms || n_stages < kNumStages) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;  // the library was shut down
    return time_stages(b, batch_args(b), stage_ms);
}

int lvn_batch_get_tuning(lvn_batch* bh, lvn_tuning* out) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !out) return LVN_ERR_INVALID;
    *out = effective_tuning(b);
    return LVN_OK;
}

int lvn_batch_set_tuning(lvn_batch* bh, const lvn_tuning* t) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !t || !valid_tuning(b, *t)) return LVN_ERR_INVALID;
    b->tuning = *t;
    drop_graph(b);
    return LVN_OK;
}

int lvn_batch_autotune(lvn_batch* bh, int32_t reps, lvn_tuning* out) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || reps < 1 || reps > 100) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;  // the library was shut down
    lvn_tuning best = effective_tuning(b);
    if (b->n > 0) {
        struct Param {
            int32_t lvn_tuning::*field;
            std::vector<int32_t> values;
            std::vector<int> stages;  // the stages whose time the parameter drives
        };
        const std::vector<Param> params = {
            {&lvn_tuning::density_ysplit, {1, 2, 3, 6}, {0}},
            {&lvn_tuning::density_per_cu, {1, 2}, {0}},
            {&lvn_tuning::persistent_per_cu, {1, 2}, {5}},
            {&lvn_tuning::height_lines, {1, 2}, {5}},
            {&lvn_tuning::cross_lanes, {1, 2, 4, 8, kCrossSampleParallel}, {5}},
            {&lvn_tuning::row_blocks, {1024, 2048, 4096, 8192, 16384, 32768}, {2, 4, 6, 8}},
            {&lvn_tuning::voxel_blocks, {1024, 2048, 4096, 8192, 16384, 32768}, {9, 11}},
        };
        float ms[kNumStages];
        for (const Param& p : params) {
            int32_t pick = best.*p.field;
            float pick_ms = 0.f;
            bool first = true;
            for (int32_t v : p.values) {
                lvn_tuning cand = best;
                cand.*p.field = v;
                if (!valid_tuning(b, cand)) continue;
                const KArgs a = batch_args(b, cand);
                int r = time_stages(b, a, ms);  // untimed warm-up of this shape
                if (r) return r;
                float total = 0.f;
                for (int k = 0; k < reps; ++k) {
                    r = time_stages(b, a, ms);
                    if (r) return r;
                    for (int st : p.stages) total += ms[st];
                }
                if (first || total < pick_ms) {
                    pick = v;
                    pick_ms = total;
                    first = false;
                }
            }
            best.*p.field = pick;
        }
        b->tuning = best;
        drop_graph(b);
    }
    if (out) *out = best;
    return LVN_OK;
}

int lvn_batch_work(lvn_batch* bh, uint64_t* work) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !work) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;  // the library was shut down
    Ctx* c = &b->ctx;
    KArgs a = batch_args(b);
    a.work = b->d_work;
    LVN_HIP(hipMemsetAsync(b->d_work, 0, sizeof(unsigned long long) * 8, c->stream));
    b->simplified = false;
    if (b->n) launch_pipeline(a, b->n, c->stream);
    unsigned long long h[8];
    LVN_HIP(hipMemcpyAsync(h, b->d_work, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 7; ++i) work[i] = h[i];
    return LVN_OK;
}

int lvn_batch_counts(lvn_batch* bh, int32_t* counts) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !counts) return LVN_ERR_INVALID;
    Ctx* c = &b->ctx;
    if (b->n == 0) return LVN_OK;
    LVN_HIP(hipMemcpyAsync(c->h_counters, c->d_counters, sizeof(Counters) * b->n, hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < b->n; ++i) {
        const Counters& k = c->h_counters[i];
        counts[6 * i + 0] = k.edges;
        counts[6 * i + 1] = k.voxels;
        counts[6 * i + 2] = 2 * k.quads;
        counts[6 * i + 3] = k.seams;
        counts[6 * i + 4] = k.status ? LVN_ERR_OVERFLOW : LVN_OK;
        counts[6 * i + 5] = k.resolved;
    }
    return LVN_OK;
}

int lvn_batch_get_mesh(lvn_batch* bh, int32_t i, lvn_vertex* v, int32_t v_cap, int32_t* nv, int32_t* tri,
                       int32_t tri_cap, int32_t* nt, lvn_seam_node* s, int32_t s_cap, int32_t* ns) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || i < 0 || i >= b->n || !nv || !nt || !ns) return LVN_ERR_INVALID;
    Ctx* c = &b->ctx;
    Counters k;
    LVN_HIP(hipMemcpyAsync(&c->h_counters[i], c->d_counters + i, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    k = c->h_counters[i];
    if (k.status) return LVN_ERR_OVERFLOW;
    const bool simp = b->simplified && (size_t)(2 * i + 1) < b->simp_counts.size();
    const int V = simp ? b->simp_counts[2 * i] : k.edges > 0 && k.quads > 0 ? k.voxels : 0;
    const int T = simp ? b->simp_counts[2 * i + 1] : k.edges > 0 ? 2 * k.quads : 0;
    const int S = k.edges > 0 ? k.seams : 0;
    *nv = V;
    *nt = T;
    *ns = S;
    if (V > v_cap || T > tri_cap || S > s_cap) return LVN_ERR_CAPACITY;
    const char* slot = c->d_arena + (uint64_t)i * c->L.slot_bytes;
    if (V && v) LVN_HIP(hipMemcpy(v, slot + c->L.verts, sizeof(lvn_vertex) * (size_t)V, hipMemcpyDeviceToHost));
    if (T && tri) LVN_HIP(hipMemcpy(tri, slot + c->L.tris, 12 * (size_t)T, hipMemcpyDeviceToHost));
    if (S && s) LVN_HIP(hipMemcpy(s, slot + c->L.seams, sizeof(lvn_seam_node) * (size_t)S, hipMemcpyDeviceToHost));
    return LVN_OK;
}

int lvn_debug_chunk_stages(lvn_ctx* ctx, const int32_t min[3], int32_t size, const lvn_csg_op* ops, int32_t nops,
                           lvn_stage_dump* out) {
    if (!ctx || !min || !out || nops < 0 || (nops > 0 && !ops)) return LVN_ERR_INVALID;
    if (!valid_chunk_size(size, reinterpret_cast<Ctx*>(ctx)->N)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    std::vector<lvn_csg_op> batch(ops, ops + nops);
    int r = run_field(c, min, size, batch, nullptr);
    if (r) return r;
    r = run_mesh(c);
    if (r) return r;
    Counters k;
    r = read_counters(c, &k);
    if (r) return r;
    out->num_edges = k.edges;
    out->num_voxels = k.voxels;
    char* slot = c->d_arena;
    const size_t fs = (size_t)c->G.field_samples;
    if (out->density) {
        if (c->flags & LVN_CTX_FP16_DENSITY) {
            std::vector<__half> tmp(fs);
            LVN_HIP(hipMemcpy(tmp.data(), slot + c->L.density, fs * 2, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < fs; ++i) out->density[i] = __half2float(tmp[i]);
        } else {
            LVN_HIP(hipMemcpy(out->density, slot + c->L.density, fs * 4, hipMemcpyDeviceToHost));
        }
    }
    if (out->materials) LVN_HIP(hipMemcpy(out->materials, slot + c->L.mat, fs, hipMemcpyDeviceToHost));
    const int ne = std::min(k.edges, out->edge_cap);
    if (out->edge_ids && ne > 0) LVN_HIP(hipMemcpy(out->edge_ids, slot + c->L.e_ids, 4 * (size_t)ne, hipMemcpyDeviceToHost));
    if (out->edge_data && ne > 0) LVN_HIP(hipMemcpy(out->edge_data, slot + c->L.e_data, 16 * (size_t)ne, hipMemcpyDeviceToHost));
    const int nvx = std::min(k.voxels, out->voxel_cap);
    if (nvx > 0) {
        if (out->voxel_codes) LVN_HIP(hipMemcpy(out->voxel_codes, slot + c->L.v_code, 4 * (size_t)nvx, hipMemcpyDeviceToHost));
        if (out->voxel_edge_masks) {
            // bits 12-14 hold the voxel's quad flags
            LVN_HIP(hipMemcpy(out->voxel_edge_masks, slot + c->L.v_emask, 4 * (size_t)nvx, hipMemcpyDeviceToHost));
            for (int i = 0; i < nvx; ++i) out->voxel_edge_masks[i] &= 0xfff;
        }
        if (out->voxel_materials) LVN_HIP(hipMemcpy(out->voxel_materials, slot + c->L.v_mat, 4 * (size_t)nvx, hipMemcpyDeviceToHost));
        // position and normal are the first two float4 of each 48-byte vertex
        if (out->voxel_positions && nvx)
            LVN_HIP(hipMemcpy2D(out->voxel_positions, 16, slot + c->L.verts, 48, 16, (size_t)nvx, hipMemcpyDeviceToHost));
        if (out->voxel_normals && nvx)
            LVN_HIP(hipMemcpy2D(out->voxel_normals, 16, slot + c->L.verts + 16, 48, 16, (size_t)nvx, hipMemcpyDeviceToHost));
    }
    if (k.status) return LVN_ERR_OVERFLOW;
    if (k.edges > out->edge_cap || k.voxels > out->voxel_cap) return LVN_ERR_CAPACITY;
    return LVN_OK;
}

}  // extern "C"
