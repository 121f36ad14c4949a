// lvn_kernels.hip — the chunk-mesher stages as gfx950 kernels.
//
// Every kernel takes the chunk from blockIdx.y (or, in the persistent kernels,
// from a work item), so one launch serves one chunk (plain form) or a whole
// batch (batched form).  Compaction is deterministic: a classification kernel
// writes 64-bit masks and counts per row (edges: x/y/z masks of a hermite row;
// voxels, quads, seams: one mask per 64 elements), one workgroup per chunk scans
// the counts, and element i lands at base[row] + popcount(mask bits below i).
// The element order is therefore the reference's linear order everywhere
// (cl/density_field.cl:52-53, cl/octree.cl:153, cl/octree.cl:398-439).

#include "lvn_device.hpp"
#include "lvn_kernels.h"

#include <hip/hip_fp16.h>

#include <algorithm>

namespace lvn {

namespace {

// Batch entry `chunk` lives in arena slot slot_map[chunk] (identity without a map).
template <typename T>
__device__ __forceinline__ T* slot_ptr(const KArgs& a, int chunk, uint64_t off) {
    const uint64_t slot = a.slot_map ? (uint64_t)a.slot_map[chunk] : (uint64_t)chunk;
    return reinterpret_cast<T*>(a.arena + slot * a.L.slot_bytes + off);
}

__device__ __forceinline__ int field_index(const Geometry& G, int x, int y, int z) {
    return x + G.F * (y + G.F * z);
}

// Division by H or F (runtime values) as a multiply-high: exact for n < 2^21
// since (n * ceil(2^32/d)) / 2^32 - n/d < n / 2^32 < 1/d for d < 2^11.
__device__ __forceinline__ int div_magic(int n, uint32_t magic) { return (int)__umulhi((uint32_t)n, magic); }

// Voxel v = x + N (y + N z) with N = 2^depth.
__device__ __forceinline__ void decode_voxel(const Geometry& G, int v, int& x, int& y, int& z) {
    x = v & (G.N - 1);
    y = (v >> G.depth) & (G.N - 1);
    z = v >> (2 * G.depth);
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// Dense lookup of a compacted element through its segment's mask word m (bit `lane`)
// and base: its index, or -1.
__device__ __forceinline__ int dense_index_of(uint64_t m, int base, int lane) {
    return ((m >> lane) & 1ull) ? base + __popcll(m & lanemask_lt(lane)) : -1;
}

// Work split of a chunk over the x dimension of its grid (blockIdx.y is the chunk).
__device__ __forceinline__ int wave_in_chunk() { return (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); }
__device__ __forceinline__ int waves_per_chunk() { return (int)(gridDim.x * (blockDim.x >> 6)); }
__device__ __forceinline__ int thread_in_chunk() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
__device__ __forceinline__ int threads_per_chunk() { return (int)(gridDim.x * blockDim.x); }

// Bits x + 1 of a 66-bit sample row (lo = x 0..63, hi = x 64, 65) moved to bit x.
__device__ __forceinline__ uint64_t shift_next(uint64_t lo, uint64_t hi) { return (lo >> 1) | (hi << 63); }

// Index of edge (x, y, z, axis) in the chunk's compacted edge list, or -1 when the
// edge has no sign change, from its edge row r = y + H z, already loaded: the row's four
// words w (x, y, z edge bits of x < 64; bit 0/1/2 of w[3]: the x, y, z edge at x = 64)
// and its base.  The list is in slot order s = 3 (x + H y + H^2 z) + axis, so the index
// is the row's base plus the edges of lower x in the row, plus the lower axes at the
// same x.
__device__ __forceinline__ int edge_index_in_row(const uint64_t w[4], int base, int x, int axis) {
    const uint64_t ex = w[0], ey = w[1], ez = w[2];
    if (x < 64) {
        const uint64_t bits = axis == 0 ? ex : axis == 1 ? ey : ez;
        if (!((bits >> x) & 1ull)) return -1;
        const uint64_t lt = lanemask_lt(x);
        int k = __popcll(ex & lt) + __popcll(ey & lt) + __popcll(ez & lt);
        if (axis > 0) k += (int)((ex >> x) & 1ull);
        if (axis > 1) k += (int)((ey >> x) & 1ull);
        return base + k;
    }
    const uint64_t hi = w[3];
    if (!((hi >> axis) & 1ull)) return -1;
    int k = __popcll(ex) + __popcll(ey) + __popcll(ez);
    if (axis > 0) k += (int)(hi & 1ull);
    if (axis > 1) k += (int)((hi >> 1) & 1ull);
    return base + k;
}

// FP16 store that keeps the sign of every negative sample (an underflow to -0
// would turn a solid sample into air under the reference's `density < 0` test,
// cl/density_field.cl:34).
__device__ __forceinline__ __half to_half_keep_sign(float d) {
    __half h = __float2half_rn(d);
    if (d < 0.f && !(__half2float(h) < 0.f)) h = __ushort_as_half((unsigned short)0x8001u);
    return h;
}

// cl/shared_constants.cl:4-9 / cl/octree.cl:227-232
__constant__ int kEdgeMap[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                    {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};

}  // namespace

// ---------------------------------------------------------------------------
// Density + default material (GenerateDefaultField, cl/density_field.cl:11-37)
// ---------------------------------------------------------------------------
// One thread per (x, z) column segment: the height term depends on x and z
// only, so it is evaluated once and reused down the column.  A wave takes 64
// consecutive columns (group g, column x + F z = 64 g + lane) over one y-segment.
__device__ __forceinline__ bool chunk_mat_free(const KArgs& a, const ChunkDesc& cd) {
    return a.mat_free && cd.op_count == 0 && cd.mode != kModeCached;
}

template <bool kCount>
__device__ __forceinline__ void density_column(const KArgs& a, const NoiseLds& T, int chunk, int seg, int g) {
    const Geometry& G = a.G;
    const int F = G.F;
    const int lane = threadIdx.x & 63;
    const bool valid = g * 64 + lane < F * F;     // the last group is partial
    const int col = valid ? g * 64 + lane : F * F - 1;
    const int seg_len = (F + a.ysplit - 1) / a.ysplit;
    const ChunkDesc cd = a.chunks[chunk];  // a copy: the loops below store through pointers that could alias it
    const int z = div_magic(col, G.magic_f);
    const int x = col - z * F;
    const int y0 = seg * seg_len;
    const int y1 = min(F, y0 + seg_len);
    uint8_t* mat = slot_ptr<uint8_t>(a, chunk, a.L.mat);
    uint8_t* orig = slot_ptr<uint8_t>(a, chunk, a.L.orig);
    // the materials before CSG are read only to find edges a brush touched: chunks
    // without ops skip them (uniform over a 64-column item, which lies in one chunk)
    const bool keep_orig = cd.op_count > 0;
    if (cd.mode == kModeCached) {  // a cached field: its materials, no density
        const uint8_t* old = reinterpret_cast<const uint8_t*>(cd.old_mat);
        for (int y = y0; y < y1 && valid; ++y) {
            const int idx = field_index(G, x, y, z);
            const uint8_t m = old[idx];
            mat[idx] = m;
            if (keep_orig) orig[idx] = m;
        }
        return;
    }
    // A chunk without brushes in a batch stores no materials (uniform over the wave): its
    // "not air" bits go to the planes, one ballot per y, and lane j keeps the word of
    // y = y0 + 64 k + j until the wave stores them together.
    const bool planes = chunk_mat_free(a, cd);
    uint64_t* P = slot_ptr<uint64_t>(a, chunk, a.L.planes) + (long)g * F;
    const float wx = (float)(x * cd.scale + cd.off[0]);
    const float wz = (float)(z * cd.scale + cd.off[2]);
    const float H = height_term(a.F, T, wx, wz);
    if (seg == 0 && valid) slot_ptr<float>(a, chunk, a.L.h_cols)[col] = H;  // read by the crossing kernels
    float* dens = slot_ptr<float>(a, chunk, a.L.density);
    __half* dens16 = slot_ptr<__half>(a, chunk, a.L.density);
    unsigned caves = 0;
    uint64_t word = 0;
    for (int y = y0; y < y1; ++y) {
        const float wy = (float)(y * cd.scale + cd.off[1]);
        const float d = combine(a.F, T, wy - H, wx, wy, wz);
        if constexpr (kCount) caves += (a.F.kind == LVN_FIELD_TERRAIN_CAVES && !((wy - H) > a.F.cave_cut)) ? 1u : 0u;
        const int idx = field_index(G, x, y, z);
        const uint8_t m = d < 0.f ? (uint8_t)a.defmat : (uint8_t)kAir;
        if (valid) {
            if (a.fp16)
                dens16[idx] = to_half_keep_sign(d);
            else if (!a.dens_cols)
                dens[idx] = d;
            if (!planes) {
                mat[idx] = m;
                if (keep_orig) orig[idx] = m;
            }
        }
        if (planes) {
            const uint64_t bits = __ballot(valid && m != (uint8_t)kAir);
            const int j = (y - y0) & 63;
            if (lane == j) word = bits;
            if (j == 63 || y == y1 - 1) {
                if (lane <= j) P[y - j + lane] = word;
            }
        }
    }
    if constexpr (kCount) {
        if (valid) {
            atomicAdd(&a.work[0], 1ull);
            atomicAdd(&a.work[1], (unsigned long long)caves);
            atomicAdd(&a.work[5], (unsigned long long)(y1 > y0 ? y1 - y0 : 0));
        }
    }
}

// Wave-level work queue: lane 0 takes the next 64-element item, all lanes get it.
__device__ __forceinline__ int next_item(int* queue) {
    int item = 0;
    if ((threadIdx.x & 63) == 0) item = atomicAdd(queue, 1);
    return __shfl(item, 0);
}

// Items of a persistent kernel shared within the workgroup: workgroup b owns items b,
// b + gridDim.x, ... (`total` items in all; interleaved, so every workgroup gets items
// of every chunk and their totals even out) and its waves take them through the LDS
// counter `next` (zeroed before the workgroup's first barrier).  No global atomic per
// item: one shared counter serves only about 77,000 grabs per ms on MI355X, and each
// grab is a round trip to memory.  Returns the item, or -1 when the workgroup's are done.
__device__ __forceinline__ int wg_next_item(int* next, int total) {
    const int mine = (total - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
    int j = 0;
    if ((threadIdx.x & 63) == 0) j = atomicAdd(next, 1);
    j = __shfl(j, 0);
    return j < mine ? (int)blockIdx.x + (int)gridDim.x * j : -1;
}

// Persistent: the grid fills the machine once (one or two 1024-thread workgroups per
// CU, a launch-shape parameter), each workgroup loads the noise tables once, and its
// waves take 64-column items from the workgroup's share (wg_next_item: 1.01-1.04 ms ->
// 0.99-1.00 ms against one global queue).
// kCount adds the work counters of lvn_batch_work (a separate instantiation, so
// the timed kernels carry no counting code).
// At most 80 SGPRs and 64 VGPRs: two 1024-thread workgroups then fit a CU (8 waves per
// SIMD; some SGPRs spill into VGPR lanes), which the tuner takes where the pass is a
// chain of height-term lookups per column (the reference's terrain field: 0.69 -> 0.57 ms
// on its clipmap batch); the cave field's pass takes the same time either way.
template <bool kCount>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_num_sgpr(80), amdgpu_waves_per_eu(8, 8))) void k_density(
    KArgs a, int chunks) {
    __shared__ NoiseLds T;
    __shared__ int next_local;
    if (threadIdx.x == 0) next_local = 0;
    load_noise_tables(T, a.perm, a.grads);  // ends with a barrier
    const int groups = a.G.col_groups;
    const int items_per_chunk = groups * a.ysplit;  // (y-segment, column group) pairs
    const int total = items_per_chunk * chunks;
    for (;;) {
        const int item = wg_next_item(&next_local, total);
        if (item < 0) break;
        const int chunk = item / items_per_chunk;
        const int r = item - chunk * items_per_chunk;
        const int seg = r / groups;
        density_column<kCount>(a, T, chunk, seg, r - seg * groups);
    }
}

// ---------------------------------------------------------------------------
// CSG material override (CSG_HermiteIndices + BrushMaterial,
// cl/apply_csg_operation.cl:114-140,179-210): ops in order, the last op whose
// brush contains a sample sets its material.  A chunk's samples are split into
// kCsgSlabs z-slabs, one workgroup each, so the chunk with the most brushes is spread
// over several CUs; each workgroup walks the ops in order over the (x, y) columns of
// each op's bounding box within its slab.  Every sample belongs to one workgroup and
// sees the ops in order (a barrier after each op), as in a single pass.
// ---------------------------------------------------------------------------
constexpr int kCsgSlabs = 8;

__global__ __launch_bounds__(256) void k_csg(KArgs a) {
    const int chunk = blockIdx.y;
    const ChunkDesc cd = a.chunks[chunk];  // a copy: the loops below store through pointers that could alias it
    const int F = a.G.F;
    const int zs = (F + kCsgSlabs - 1) / kCsgSlabs;
    const int z0 = (int)blockIdx.x * zs, z1 = min(F, z0 + zs);  // this workgroup's slab
    if (cd.op_count == 0 || z0 >= z1) return;                    // uniform, before any barrier
    uint8_t* mat = slot_ptr<uint8_t>(a, chunk, a.L.mat);
    const float inv = 1.f / (float)cd.scale;
    for (int k = 0; k < cd.op_count; ++k) {
        const DevOp op = a.ops[a.op_index[cd.op_start + k]];
        const float r = op.radius + 1.f;
        int lo[3], hi[3];
        const float o[3] = {op.ox, op.oy, op.oz};
        for (int c = 0; c < 3; ++c) {
            lo[c] = max(0, (int)floorf((o[c] - r - (float)cd.off[c]) * inv));
            hi[c] = min(F - 1, (int)ceilf((o[c] + r - (float)cd.off[c]) * inv));
        }
        lo[2] = max(lo[2], z0);
        hi[2] = min(hi[2], z1 - 1);
        const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
        if (nx > 0 && ny > 0 && hi[2] >= lo[2]) {
            const uint8_t newm = op.type == 0 ? (uint8_t)op.material : (uint8_t)kAir;
            for (int i = threadIdx.x; i < nx * ny; i += blockDim.x) {
                const int y = lo[1] + i / nx;
                const int x = lo[0] + (i - (y - lo[1]) * nx);
                const float px = (float)(x * cd.scale + cd.off[0]);
                const float py = (float)(y * cd.scale + cd.off[1]);
                for (int z = lo[2]; z <= hi[2]; ++z) {
                    const float pz = (float)(z * cd.scale + cd.off[2]);
                    if (brush_density(op, px, py, pz) <= 0.f) mat[field_index(a.G, x, y, z)] = newm;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Edge classification (FindFieldEdges, cl/density_field.cl:41-76) as bit rows.
// k_solid_rows packs the post-CSG "not air" test of every sample into rows of
// F bits (y, z fixed); k_edge_rows turns three neighbouring rows into the x, y
// and z edge masks of a hermite row with a few 64-bit operations, so a chunk's
// 3 H^3 edge slots are classified by H^2 threads.
// ---------------------------------------------------------------------------
constexpr int kRowsPerBlock = 16;  // sample rows staged per wave

__global__ __launch_bounds__(256) void k_solid_rows(KArgs a) {
    __shared__ uint32_t stage[4][kRowsPerBlock * 66 / 4];
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const int F = G.F;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t* mat = slot_ptr<const uint32_t>(a, chunk, a.L.mat);
    uint64_t* S = slot_ptr<uint64_t>(a, chunk, a.L.solid);
    if (chunk_mat_free(a, a.chunks[chunk])) {  // uniform over the workgroup
        // Row (y, z) is bits [F z, F z + F) of plane y: a 64-bit window of two or three
        // consecutive group words, then F - 64 bits above it (none when F <= 64).
        const uint64_t* P = slot_ptr<const uint64_t>(a, chunk, a.L.planes);
        const uint64_t lo_mask = F >= 64 ? ~0ull : ((1ull << F) - 1ull);
        const uint64_t hi_mask = F > 64 ? ((1ull << (F - 64)) - 1ull) : 0ull;
        for (int r = thread_in_chunk(); r < G.field_rows; r += threads_per_chunk()) {
            const int z = div_magic(r, G.magic_f), y = r - z * F;
            const int t0 = F * z, g0 = t0 >> 6, sh = t0 & 63;
            const uint64_t w0 = P[g0 * F + y];
            const uint64_t w1 = g0 + 1 < G.col_groups ? P[(g0 + 1) * F + y] : 0ull;
            const uint64_t w2 = g0 + 2 < G.col_groups ? P[(g0 + 2) * F + y] : 0ull;
            const uint64_t lo = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
            const uint64_t hi = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
            S[2 * r] = lo & lo_mask;
            S[2 * r + 1] = hi & hi_mask;
        }
        return;
    }
    const uint8_t* bytes = reinterpret_cast<const uint8_t*>(stage[wave]);
    const int blocks = (G.field_rows + kRowsPerBlock - 1) / kRowsPerBlock;
    const int words_total = G.field_samples >> 2;  // F is even, so F^3 is a multiple of 8
    const int block_words = (kRowsPerBlock * F) >> 2;
    for (int b = wave_in_chunk(); b < blocks; b += waves_per_chunk()) {
        // block b = rows [16 b, 16 b + 16) = bytes [16 F b, 16 F (b + 1)): 4-byte aligned
        const int w0 = b * block_words;
        for (int j = lane; j < block_words; j += 64)
            if (w0 + j < words_total) stage[wave][j] = mat[w0 + j];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        // The block's rows are 2 * rows consecutive words of S: lane 2k keeps row k's low
        // word and lane 2k + 1 its high word, and one store writes them all.
        const int rows = min(kRowsPerBlock, G.field_rows - b * kRowsPerBlock);
        uint64_t word = 0;
        for (int k = 0; k < rows; ++k) {
            const bool s0 = lane < F && bytes[k * F + lane] != kAir;
            const bool s1 = lane + 64 < F && bytes[k * F + lane + 64] != kAir;
            const uint64_t lo = __ballot(s0), hi = __ballot(s1);
            if (lane == 2 * k) word = lo;
            if (lane == 2 * k + 1) word = hi;
        }
        if (lane < 2 * rows) S[2 * b * kRowsPerBlock + lane] = word;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
}

__global__ __launch_bounds__(256) void k_edge_rows(KArgs a) {
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const int H = G.H, F = G.F;
    const uint64_t* S = slot_ptr<const uint64_t>(a, chunk, a.L.solid);
    uint64_t* E = slot_ptr<uint64_t>(a, chunk, a.L.e_rows);
    int* count = slot_ptr<int>(a, chunk, a.L.e_count);
    int* count_y = slot_ptr<int>(a, chunk, a.L.e_count_y);
    const uint64_t mlo = H >= 64 ? ~0ull : ((1ull << H) - 1ull);  // x in [0, H)
    const uint64_t mhi = H > 64 ? 1ull : 0ull;                     // x = 64
    for (int r = thread_in_chunk(); r < G.edge_rows; r += threads_per_chunk()) {
        const int z = div_magic(r, G.magic_h), y = r - z * H;
        const int s0 = y + F * z;
        const uint64_t alo = S[2 * s0], ahi = S[2 * s0 + 1];                   // (y, z)
        const uint64_t blo = S[2 * (s0 + 1)], bhi = S[2 * (s0 + 1) + 1];       // (y + 1, z)
        const uint64_t clo = S[2 * (s0 + F)], chi = S[2 * (s0 + F) + 1];       // (y, z + 1)
        const uint64_t xlo = (alo ^ shift_next(alo, ahi)) & mlo;
        const uint64_t xhi = (ahi ^ (ahi >> 1)) & mhi;
        const uint64_t ylo = (alo ^ blo) & mlo, yhi = (ahi ^ bhi) & mhi;
        const uint64_t zlo = (alo ^ clo) & mlo, zhi = (ahi ^ chi) & mhi;
        const uint64_t hi = xhi | (yhi << 1) | (zhi << 2);
        E[4 * r] = xlo;
        E[4 * r + 1] = ylo;
        E[4 * r + 2] = zlo;
        E[4 * r + 3] = hi;
        count[r] = __popcll(xlo) + __popcll(ylo) + __popcll(zlo) + __popcll(hi);
        count_y[r] = __popcll(ylo) + (int)((hi >> 1) & 1ull);
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of segment counts, one workgroup per chunk (replaces the
// recursive Scan / ExclusiveScan with its blocking readback, src/compute.cpp:344-396).
// ---------------------------------------------------------------------------
// Exclusive scan of count[0, n) into base[0, n) by one 1024-thread workgroup;
// returns the total to every thread.
__device__ __forceinline__ int block_scan(const int* count, int* base, int n, int* buf) {
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b0 = min(n, tid * per), b1 = min(n, b0 + per);
    int sum = 0;
    for (int i = b0; i < b1; ++i) sum += count[i];
    __syncthreads();
    buf[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = tid >= off ? buf[tid - off] : 0;
        __syncthreads();
        buf[tid] += v;
        __syncthreads();
    }
    int run = buf[tid] - sum;
    for (int i = b0; i < b1; ++i) {
        base[i] = run;
        run += count[i];
    }
    return buf[1023];
}

// Edge (kScanEdges) or voxel (kScanVoxels) scan: the stage is a template parameter
// (with a runtime selector hipcc, ROCm 7.2, gfx950, kept one stage's counter
// address for the other's total).
template <int KIND>
__global__ __launch_bounds__(1024) void k_scan(KArgs a) {
    __shared__ int buf[1024];
    const int chunk = blockIdx.y;
    Counters* cnt = a.counters + chunk;
    if constexpr (KIND == kScanEdges) {
        const int t = block_scan(slot_ptr<const int>(a, chunk, a.L.e_count), slot_ptr<int>(a, chunk, a.L.e_base),
                                 a.G.edge_rows, buf);
        const int ty = block_scan(slot_ptr<const int>(a, chunk, a.L.e_count_y), slot_ptr<int>(a, chunk, a.L.e_base_y),
                                  a.G.edge_rows, buf);
        if (threadIdx.x == 0) {
            cnt->edges = min(t, a.caps.edges);
            cnt->edges_y = min(ty, a.caps.edges);
            if (t > a.caps.edges) atomicOr(&cnt->status, kStatusEdgeOverflow);
        }
    } else {
        const int t = block_scan(slot_ptr<const int>(a, chunk, a.L.v_count), slot_ptr<int>(a, chunk, a.L.v_base),
                                 a.G.voxel_segs, buf);
        if (threadIdx.x == 0) {
            cnt->voxels = min(t, a.caps.voxels);
            if (t > a.caps.voxels) atomicOr(&cnt->status, kStatusVoxelOverflow);
        }
    }
}

// ---------------------------------------------------------------------------
// Edge compaction (CompactEdges, cl/density_field.cl:80-92): the edge id is
// ((x | y << S | z << 2S) << 2) | axis (cl/density_field.cl:63,74).
// ---------------------------------------------------------------------------
// The ids of every hermite row's edges in slot order (x, then axis), and each edge's
// crossing position (y-edges first, then x and z edges).  A wave takes 64 consecutive
// rows r = y + H z and loads their words together (lane j: row j).  Where its live rows
// are dense (kDenseRow edges or more on average) it walks them in order with lane x
// holding position x of the row: an edge's rank in the row is the popcount of the three
// edge words below its lane (lower x) plus the lower axes at its own x, so each store
// instruction writes one contiguous stretch of the lists.  Where they are sparse, each
// lane walks its own row's few edges (fewer, cheaper iterations; their stores scatter
// over the rows' stretches).  With the per-lane walk alone, lines of the clipmap batch's
// dense rows left L2 half-written: 47 bytes reached HBM per 8 bytes of list, and the
// stage took 3.3 ms.  Both walks write the same lists.
constexpr int kDenseRow = 16;

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(256) void k_edge_compact(KArgs a) {
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const uint64_t* E = slot_ptr<const uint64_t>(a, chunk, a.L.e_rows);
    const int* count = slot_ptr<const int>(a, chunk, a.L.e_count);
    const int* base = slot_ptr<const int>(a, chunk, a.L.e_base);
    const int* base_y = slot_ptr<const int>(a, chunk, a.L.e_base_y);
    int* ids = slot_ptr<int>(a, chunk, a.L.e_ids);
    int* order = slot_ptr<int>(a, chunk, a.L.e_order);
    const int S = G.shift, cap = a.caps.edges, ey_total = a.counters[chunk].edges_y;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt(lane);
    // rank k in id order; crossing position: y-edges first, then x and z edges
    auto put = [&](int k, int ky, int id, bool is_y) {
        if (k < cap) {
            ids[k] = id;
            const int pos = is_y ? ky : ey_total + (k - ky);
            if (pos < cap) order[pos] = k;
        }
    };
    const int groups = (G.edge_rows + 63) >> 6;
    for (int g = wave_in_chunk(); g < groups; g += waves_per_chunk()) {
        const int rl = (g << 6) + lane;
        int cnt = 0, kb = 0, kyb = 0;
        uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
        if (rl < G.edge_rows) {
            cnt = count[rl];
            kb = base[rl];
            kyb = base_y[rl];
            w0 = E[4 * rl];
            w1 = E[4 * rl + 1];
            w2 = E[4 * rl + 2];
            w3 = E[4 * rl + 3];
        }
        uint64_t live = __ballot(cnt != 0);
        if (!live) continue;
        if (wave_sum(cnt) < kDenseRow * __popcll(live)) {  // sparse rows: one lane per row
            if (cnt == 0) continue;
            const int z = div_magic(rl, G.magic_h), y = rl - z * G.H;
            const int row_id = (y << S) | (z << (2 * S));
            int k = kb, ky = kyb;
            uint64_t any = w0 | w1 | w2;
            while (any) {
                const int x = (int)__builtin_ctzll(any);
                any &= any - 1ull;
                const int id = (x | row_id) << 2;
                if ((w0 >> x) & 1ull) put(k++, ky, id, false);
                if ((w1 >> x) & 1ull) put(k++, ky++, id | 1, true);
                if ((w2 >> x) & 1ull) put(k++, ky, id | 2, false);
            }
            const int id64 = (64 | row_id) << 2;  // x = 64 (N = 64 only)
            for (int axis = 0; axis < 3; ++axis) {
                if (!((w3 >> axis) & 1ull)) continue;
                put(k++, ky, id64 | axis, axis == 1);
                ky += axis == 1 ? 1 : 0;
            }
            continue;
        }
        while (live) {
            const int j = (int)__builtin_ctzll(live);
            live &= live - 1ull;
            const int r = (g << 6) + j;
            const int z = div_magic(r, G.magic_h), y = r - z * G.H;
            const int row_id = (y << S) | (z << (2 * S));
            const uint64_t ex = readlane64(w0, j), ey = readlane64(w1, j), ez = readlane64(w2, j);
            const uint64_t hi = readlane64(w3, j);
            const int k0 = __builtin_amdgcn_readlane(kb, j), ky0 = __builtin_amdgcn_readlane(kyb, j);
            const int bx = (int)((ex >> lane) & 1ull), by = (int)((ey >> lane) & 1ull), bz = (int)((ez >> lane) & 1ull);
            int k = k0 + __popcll(ex & lt) + __popcll(ey & lt) + __popcll(ez & lt);
            int ky = ky0 + __popcll(ey & lt);
            const int id = (lane | row_id) << 2;
            if (bx) put(k, ky, id, false);
            k += bx;
            if (by) put(k, ky, id | 1, true);
            k += by;
            ky += by;
            if (bz) put(k, ky, id | 2, false);
            if (hi != 0 && lane == 0) {  // x = 64 (N = 64 only): after every x < 64
                int kh = k0 + __popcll(ex) + __popcll(ey) + __popcll(ez), kyh = ky0 + __popcll(ey);
                const int id64 = (64 | row_id) << 2;
                for (int axis = 0; axis < 3; ++axis) {
                    if (!((hi >> axis) & 1ull)) continue;
                    put(kh, kyh, id64 | axis, axis == 1);
                    ++kh;
                    kyh += axis == 1 ? 1 : 0;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Edge crossing + normal
// ---------------------------------------------------------------------------

// Heights along the lines of x- and z-edges.  An x-edge (x, y, z) samples the height
// term at (x + k/16, z) and a z-edge at (x, z + k/16): the same values for every y, and
// caves give a line about three crossing edges.  So the interior samples k = 1..15 of
// every line holding a crossing x- or z-edge are evaluated once, before the crossing
// kernels, by the expressions those kernels use (bit-identical values), and read back
// there.  Line (sel, x, z), sel 0 for x-edges and 1 for z-edges, starts at:
__device__ __forceinline__ int hline_index(int H, int sel, int x, int z) { return ((sel * H + z) * H + x) * kLineSamples; }

// Position of the j-th set bit (from 0) of m; m has more than j set bits.
__device__ __forceinline__ int nth_bit(uint64_t m, int j) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (j >= c) {
            j -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// Item = (chunk, sel, z).  Items are short (a few hundred height evaluations), and a
// shared atomic queue serves only about 77,000 grabs per ms, which would bound this pass
// at 0.86 ms for 512 chunks; a static stride over the waves leaves many idle (the items
// of one wave then come from one layer of chunks, all sky or all caves).  So workgroup b
// takes items b, b + gridDim.x, ... (chunks of every layer) and its waves share them
// through an LDS counter.  The lines
// needed are the OR over y of the chunk's edge rows (x-edge or z-edge words); the lanes
// run over their samples only.  Chunks whose edges take no density crossings (the
// cached-field form) are skipped.
template <bool kCount>
__global__ __launch_bounds__(1024) void k_height_lines(KArgs a, int chunks) {
    __shared__ NoiseLds T;
    __shared__ int next_local;
    if (threadIdx.x == 0) next_local = 0;
    load_noise_tables(T, a.perm, a.grads);  // ends with a barrier
    const Geometry& G = a.G;
    const int H = G.H;
    const int per_chunk = 2 * H;
    const int total = per_chunk * chunks;
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int item = wg_next_item(&next_local, total);
        if (item < 0) break;
        const int chunk = item / per_chunk;
        const int r = item - chunk * per_chunk;
        const int sel = r / H, z = r - sel * H;
        const ChunkDesc cd = a.chunks[chunk];  // a copy: the sample loop's stores could alias it
        if (cd.mode == kModeCached) continue;  // uniform over the wave
        const uint64_t* E = slot_ptr<const uint64_t>(a, chunk, a.L.e_rows);
        const int word = sel == 0 ? 0 : 2;  // x-edge or z-edge bits of x < 64
        uint64_t lo = 0, hi = 0;
        for (int y = lane; y < H; y += 64) {
            const int row = y + H * z;
            lo |= E[4 * row + word];
            hi |= (E[4 * row + 3] >> word) & 1ull;  // bit 0: x-edge, bit 2: z-edge at x = 64
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            lo |= __shfl_xor(lo, m);
            hi |= __shfl_xor(hi, m);
        }
        float* out = slot_ptr<float>(a, chunk, a.L.h_lines) + hline_index(H, sel, 0, z);
        const int n_lo = __popcll(lo), n = n_lo + (int)(hi & 1ull);  // hi is set only when H > 64
        unsigned nh = 0;
        for (int p = lane; p < n * kLineSamples; p += 64) {
            const int j = p / kLineSamples, k = p - j * kLineSamples + 1;
            const int x = j < n_lo ? nth_bit(lo, j) : 64;
            const int w0x = x * cd.scale + cd.off[0], w0z = z * cd.scale + cd.off[2];
            const float p0x = (float)w0x, p0z = (float)w0z;
            const float p1x = (float)(w0x + (sel == 0 ? cd.scale : 0));
            const float p1z = (float)(w0z + (sel == 1 ? cd.scale : 0));
            const float cur = (float)k * 0.0625f;  // = the sequential scan's running sum (exact)
            out[x * kLineSamples + (k - 1)] = height_term(a.F, T, p0x + (p1x - p0x) * cur, p0z + (p1z - p0z) * cur);
            ++nh;
        }
        if constexpr (kCount) atomicAdd(&a.work[2], (unsigned long long)nh);
    }
}

// The normal of FindEdgeIntersectionInfo (cl/density_field.cl:131-150): central
// differences with h = 0.001 at the arg-min sample t of the edge p0 -> p1, whose height
// term `Hbest` is reused when it was evaluated there.
__device__ __forceinline__ float4 crossing_normal(const FieldParams& Fp, const NoiseLds& T, int axis, bool ends_known,
                                                  float tbest, float Hbest, float p0x, float p0y, float p0z, float p1x,
                                                  float p1y, float p1z, unsigned& nh, unsigned& nc) {
    const bool caves = Fp.kind == LVN_FIELD_TERRAIN_CAVES;
    const float px = p0x + (p1x - p0x) * tbest;
    const float py = p0y + (p1y - p0y) * tbest;
    const float pz = p0z + (p1z - p0z) * tbest;
    const float h = 0.001f;
    // The height at the arg-min sample is only known when it was evaluated there.
    const bool hb_known = axis == 1 || !(ends_known && (tbest == 0.f || tbest == 1.f));
    const float H0 = hb_known ? Hbest : height_term(Fp, T, px, pz);
    const float Hxp = height_term(Fp, T, px + h, pz);
    const float Hxm = height_term(Fp, T, px - h, pz);
    const float Hzp = height_term(Fp, T, px, pz + h);
    const float Hzm = height_term(Fp, T, px, pz - h);
    const float dx = combine(Fp, T, py - Hxp, px + h, py, pz) - combine(Fp, T, py - Hxm, px - h, py, pz);
    const float dy = combine(Fp, T, (py + h) - H0, px, py + h, pz) - combine(Fp, T, (py - h) - H0, px, py - h, pz);
    const float dz = combine(Fp, T, py - Hzp, px, py, pz + h) - combine(Fp, T, py - Hzm, px, py, pz - h);
    nh += hb_known ? 4u : 5u;
    if (caves) {
        nc += !((py - Hxp) > Fp.cave_cut) + !((py - Hxm) > Fp.cave_cut) + !(((py + h) - H0) > Fp.cave_cut) +
              !(((py - h) - H0) > Fp.cave_cut) + !((py - Hzp) > Fp.cave_cut) + !((py - Hzm) > Fp.cave_cut);
    }
    const float3 n = normalize_or_zero(dx, dy, dz);
    return make_float4(n.x, n.y, n.z, tbest);
}

// FindEdgeIntersectionInfo, cl/density_field.cl:96-151: first arg-min of |d|
// over t = k/16, then a central-difference normal with h = 0.001.  The height
// term of samples that share (x, z) is evaluated once: per y-edge, and per x- or
// z-edge line (`hl`, k_height_lines); the two end samples are the stored field
// samples when the field is FP32.
__device__ __forceinline__ float4 density_crossing(const KArgs& a, const NoiseLds& T, const ChunkDesc& cd, int x, int y, int z,
                                   int axis, bool ends_known, float d_start, float d_end, const float* hl, unsigned& nh,
                                   unsigned& nc) {
    const FieldParams& Fp = a.F;
    const bool caves = Fp.kind == LVN_FIELD_TERRAIN_CAVES;
    const int w0x = x * cd.scale + cd.off[0], w0y = y * cd.scale + cd.off[1], w0z = z * cd.scale + cd.off[2];
    const float p0x = (float)w0x, p0y = (float)w0y, p0z = (float)w0z;
    const float p1x = (float)(w0x + (axis == 0 ? cd.scale : 0));
    const float p1y = (float)(w0y + (axis == 1 ? cd.scale : 0));
    const float p1z = (float)(w0z + (axis == 2 ? cd.scale : 0));
    const float colH = axis == 1 ? height_term(Fp, T, p0x, p0z) : 0.f;
    nh += axis == 1 ? 1u : 0u;
    float minV = 3.402823466e+38f, tbest = 0.f, Hbest = colH, cur = 0.f;
    const bool lines = a.height_lines != 0 && axis != 1;
    float Hn = lines ? hl[0] : 0.f;  // the next interior sample's line height, loaded a sample ahead
    for (int k = 0; k <= 16; ++k) {
        const float px = p0x + (p1x - p0x) * cur;
        const float py = p0y + (p1y - p0y) * cur;
        const float pz = p0z + (p1z - p0z) * cur;
        float d, H = colH;
        if (ends_known && k == 0) {
            d = fabsf(d_start);
        } else if (ends_known && k == 16) {
            d = fabsf(d_end);
        } else {
            if (axis != 1) {
                if (lines && k >= 1 && k <= kLineSamples) {
                    H = Hn;
                    if (k < kLineSamples) Hn = hl[k];
                } else {
                    H = height_term(Fp, T, px, pz);
                    ++nh;
                }
            }
            nc += (caves && !((py - H) > Fp.cave_cut)) ? 1u : 0u;
            d = fabsf(combine(Fp, T, py - H, px, py, pz));
        }
        if (d < minV) {
            tbest = cur;
            minV = d;
            Hbest = H;
        }
        cur = cur + 0.0625f;
    }
    return crossing_normal(Fp, T, axis, ends_known, tbest, Hbest, p0x, p0y, p0z, p1x, p1y, p1z, nh, nc);
}

// FindEdgeIntersectionInfo of the CSG program, cl/apply_csg_operation.cl:443-477,
// with BrushZeroCrossing (:86-110) and BrushNormal (:144-175).
__device__ __forceinline__ float4 brush_crossing(const KArgs& a, const ChunkDesc& cd, int x, int y, int z, int axis) {
    const int w0x = x * cd.scale + cd.off[0], w0y = y * cd.scale + cd.off[1], w0z = z * cd.scale + cd.off[2];
    const float p0x = (float)w0x, p0y = (float)w0y, p0z = (float)w0z;
    const float p1x = (float)(w0x + (axis == 0 ? cd.scale : 0));
    const float p1y = (float)(w0y + (axis == 1 ? cd.scale : 0));
    const float p1z = (float)(w0z + (axis == 2 ? cd.scale : 0));
    const DevOp* ops = a.ops;
    const int* idx = a.op_index + cd.op_start;
    float minD = 3.402823466e+38f, crossing = 0.f;
    for (float t = 0.f; t <= 1.f; t = t + (1.f / 16.f)) {
        const float px = p0x + (p1x - p0x) * t;
        const float py = p0y + (p1y - p0y) * t;
        const float pz = p0z + (p1z - p0z) * t;
        for (int k = 0; k < cd.op_count; ++k) {
            const float d = fabsf(brush_density(ops[idx[k]], px, py, pz));
            if (d < minD) {
                crossing = t;
                minD = d;
            }
        }
    }
    const float px = p0x + (p1x - p0x) * crossing;
    const float py = p0y + (p1y - p0y) * crossing;
    const float pz = p0z + (p1z - p0z) * crossing;
    const float h = 0.001f;
    float3 n = make_float3(0.f, 0.f, 0.f);
    for (int k = 0; k < cd.op_count; ++k) {
        const DevOp op = ops[idx[k]];
        if (brush_density(op, px, py, pz) > 0.f) continue;
        const float dx = brush_density(op, px + h, py, pz) - brush_density(op, px - h, py, pz);
        const float dy = brush_density(op, px, py + h, pz) - brush_density(op, px, py - h, pz);
        const float dz = brush_density(op, px, py, pz + h) - brush_density(op, px, py, pz - h);
        const float flip = op.type == 0 ? 1.f : -1.f;
        const float3 u = normalize_or_zero(dx, dy, dz);
        n = make_float3(flip * u.x, flip * u.y, flip * u.z);
    }
    return make_float4(n.x, n.y, n.z, crossing);
}

// One compacted edge.  An edge with an endpoint whose material a brush changed
// takes brush data (src/compute_csg.cpp:97-217); otherwise the density crossing
// (fresh field) or its data from the cached edge list.
// The end samples of edge (x, y, z, axis), field indices i0 and i1: the stored field, or,
// for a field without the cave term whose density a batch does not store (dens_cols),
// y - H with H from the column heights: density_column's expression, bit for bit (its
// sample is combine(wy - H) = wy - H for such a field).
__device__ __forceinline__ float2 edge_ends(const KArgs& a, int chunk, const ChunkDesc& cd, int x, int y, int z, int axis,
                                           int i0, int i1) {
    if (a.dens_cols) {
        const float* hc = slot_ptr<const float>(a, chunk, a.L.h_cols);
        const int F = a.G.F;
        const float H0 = hc[x + F * z];
        const float H1 = axis == 1 ? H0 : hc[(x + (axis == 0 ? 1 : 0)) + F * (z + (axis == 2 ? 1 : 0))];
        const float wy0 = (float)(y * cd.scale + cd.off[1]);
        const float wy1 = (float)((y + (axis == 1 ? 1 : 0)) * cd.scale + cd.off[1]);
        return make_float2(wy0 - H0, wy1 - H1);
    }
    const float* dens = slot_ptr<const float>(a, chunk, a.L.density);
    return make_float2(dens[i0], dens[i1]);
}

template <bool kCount>
__device__ __forceinline__ void cross_edge(const KArgs& a, const NoiseLds& T, int chunk, int i) {
    const Geometry& G = a.G;
    const ChunkDesc& cd = a.chunks[chunk];
    const int id = slot_ptr<int>(a, chunk, a.L.e_ids)[i];
    const int axis = id & 3, hidx = id >> 2;
    const int msk = (1 << G.shift) - 1;
    const int x = hidx & msk, y = (hidx >> G.shift) & msk, z = (hidx >> (2 * G.shift)) & msk;
    if (x >= G.H || y >= G.H || z >= G.H || axis > 2) {  // never produced by k_edge_compact
        atomicOr(&a.counters[chunk].status, kStatusLookupMiss);
        return;
    }
    const int i0 = field_index(G, x, y, z);
    const int i1 = i0 + (axis == 0 ? 1 : axis == 1 ? G.F : G.F * G.F);
    const uint8_t* mat = slot_ptr<uint8_t>(a, chunk, a.L.mat);
    const uint8_t* orig = slot_ptr<uint8_t>(a, chunk, a.L.orig);
    float4 out;
    const bool touched = cd.op_count > 0 && (mat[i0] != orig[i0] || mat[i1] != orig[i1]);
    if (touched) {
        out = brush_crossing(a, cd, x, y, z, axis);
    } else if (cd.mode == kModeCached) {
        const int* oid = reinterpret_cast<const int*>(cd.old_ids);
        const float4* odata = reinterpret_cast<const float4*>(cd.old_data);
        int lo = 0, hi = cd.old_edges;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (oid[mid] < id) lo = mid + 1; else hi = mid;
        }
        if (lo < cd.old_edges && oid[lo] == id) {
            out = odata[lo];
        } else {
            out = make_float4(0.f, 0.f, 0.f, 0.f);
            atomicOr(&a.counters[chunk].status, kStatusLookupMiss);
        }
    } else {
        const bool ends = !a.fp16;
        const float2 de = ends ? edge_ends(a, chunk, cd, x, y, z, axis, i0, i1) : make_float2(0.f, 0.f);
        const float* hl = slot_ptr<const float>(a, chunk, a.L.h_lines) + hline_index(G.H, axis == 2 ? 1 : 0, x, z);
        unsigned nh = 0, nc = 0;
        out = density_crossing(a, T, cd, x, y, z, axis, ends, de.x, de.y, hl, nh, nc);
        if constexpr (kCount) {
            atomicAdd(&a.work[2], (unsigned long long)nh);
            atomicAdd(&a.work[3], (unsigned long long)nc);
        }
    }
    if constexpr (kCount) {
        atomicAdd(&a.work[6], 1ull);
        if (touched) atomicAdd(&a.work[4], (unsigned long long)(17 * cd.op_count));
    }
    slot_ptr<float4>(a, chunk, a.L.e_data)[i] = out;
}

// A chunk's edges are crossed in two runs, its y-edges (one height evaluation per
// edge) and its x and z edges (one per sample), so that a wave holds one kind and
// skips the per-sample height work when all its edges are y-edges.
__device__ __forceinline__ int edge_items_y(const Counters& c, int per_item) {
    return (min(c.edges_y, c.edges) + per_item - 1) / per_item;
}
__device__ __forceinline__ int edge_items_xz(const Counters& c, int per_item) {
    return (c.edges - min(c.edges_y, c.edges) + per_item - 1) / per_item;
}
// Edge rank (index into e_ids / e_data) of item j's slot `e`, or -1 past the end.
__device__ __forceinline__ int edge_of_item(const KArgs& a, int chunk, int j, int e, int per_item) {
    const Counters& c = a.counters[chunk];
    const int ey = min(c.edges_y, c.edges), ny = (ey + per_item - 1) / per_item;
    const int pos = j < ny ? j * per_item + e : ey + (j - ny) * per_item + e;
    if (j < ny ? pos >= ey : pos >= c.edges) return -1;
    const int i = slot_ptr<const int>(a, chunk, a.L.e_order)[pos];
    return (unsigned)i < (unsigned)c.edges ? i : -1;  // entries past a capped list are not written
}

// Items of `per_item` edges (64 / lanes per edge) over all chunks: item_prefix[c] =
// sum over earlier chunks of the chunk's y-edge and x/z-edge items, item_prefix[chunks] = total.
__global__ __launch_bounds__(1024) void k_edge_items(KArgs a, int chunks, int per_item) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < chunks; base += 1024) {
        const int c = base + threadIdx.x;
        const int v = c < chunks ? edge_items_y(a.counters[c], per_item) + edge_items_xz(a.counters[c], per_item) : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int u = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += u;
            __syncthreads();
        }
        if (c < chunks) a.item_prefix[c] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.item_prefix[chunks] = carry;
}

// Persistent like k_density: waves pull 64-edge items of any chunk.  768-thread
// workgroups: at the kernel's register count (6 waves per SIMD) two of them fit a
// CU beside their two 64 KiB tables, where one 1024-thread workgroup would leave
// a third of the wave slots empty.
constexpr int kCrossThreads = 768;

template <bool kCount>
__global__ __attribute__((amdgpu_waves_per_eu(6, 8))) __launch_bounds__(kCrossThreads) void k_edge_cross(KArgs a,
                                                                                                          int chunks) {
    __shared__ NoiseLds T;
    load_noise_tables(T, a.perm, a.grads);
    const int total = a.item_prefix[chunks];
    const int lane = threadIdx.x & 63;
    // One global queue (sharing items within workgroups instead measured slower: clipmap
    // 93.2 -> 94.5 ms).  A grab takes up to 4 consecutive items, as many as still leave
    // every wave 8 grabs: one counter serves ~77,000 grabs per ms, which items of a few
    // microseconds (a heightfield's y-edges) would otherwise wait on, and the chunk of
    // an item is found by one search per grab, the grab's later items walking forward.
    const int waves = (int)(gridDim.x * (blockDim.x >> 6));
    const int per_grab = max(1, min(4, total / (8 * waves)));
    for (;;) {
        int first = 0;
        if (lane == 0) first = atomicAdd(&a.queue[1], per_grab);
        first = __shfl(first, 0);
        if (first >= total) break;
        const int end = min(first + per_grab, total);
        int lo = 0, hi = chunks;  // last chunk with item_prefix[c] <= first
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.item_prefix[mid] <= first) lo = mid; else hi = mid;
        }
        for (int item = first; item < end; ++item) {
            while (a.item_prefix[lo + 1] <= item) ++lo;  // item < item_prefix[chunks] stops it
            const int i = edge_of_item(a, lo, item - a.item_prefix[lo], lane, 64);
            if (i >= 0) cross_edge<kCount>(a, T, lo, i);
        }
    }
}

// Small batches: kG lanes share one edge, so a chunk's few hundred edge items become
// kG times as many waves with kG times shorter dependency chains.  Lane `sub`
// evaluates samples k = sub, sub + kG, ... of the 17; the group reduces to the first
// arg-min (smallest |d|, then smallest k, as the sequential scan decides), then the
// six normal samples are spread over the group.  Every value is computed by the
// same expression as in density_crossing, so results are identical.
template <bool kCount, int kG>
__device__ __forceinline__ void cross_edge_group(const KArgs& a, const NoiseLds& T, int chunk, int i, int sub) {
    const Geometry& G = a.G;
    const ChunkDesc& cd = a.chunks[chunk];
    const int id = slot_ptr<int>(a, chunk, a.L.e_ids)[i];
    const int axis = id & 3, hidx = id >> 2;
    const int msk = (1 << G.shift) - 1;
    const int x = hidx & msk, y = (hidx >> G.shift) & msk, z = (hidx >> (2 * G.shift)) & msk;
    if (x >= G.H || y >= G.H || z >= G.H || axis > 2) {  // never produced by k_edge_compact
        if (sub == 0) atomicOr(&a.counters[chunk].status, kStatusLookupMiss);
        return;
    }
    const int i0 = field_index(G, x, y, z);
    const int i1 = i0 + (axis == 0 ? 1 : axis == 1 ? G.F : G.F * G.F);
    const uint8_t* mat = slot_ptr<uint8_t>(a, chunk, a.L.mat);
    const uint8_t* orig = slot_ptr<uint8_t>(a, chunk, a.L.orig);
    const bool touched = cd.op_count > 0 && (mat[i0] != orig[i0] || mat[i1] != orig[i1]);
    if (touched || cd.mode == kModeCached) {  // brush data or the cached list: one lane
        if (sub == 0) cross_edge<kCount>(a, T, chunk, i);
        return;
    }
    const FieldParams& Fp = a.F;
    const bool caves = Fp.kind == LVN_FIELD_TERRAIN_CAVES;
    const bool ends = !a.fp16;
    const float2 de = ends ? edge_ends(a, chunk, cd, x, y, z, axis, i0, i1) : make_float2(0.f, 0.f);
    const float d_start = de.x, d_end = de.y;
    const int w0x = x * cd.scale + cd.off[0], w0y = y * cd.scale + cd.off[1], w0z = z * cd.scale + cd.off[2];
    const float p0x = (float)w0x, p0y = (float)w0y, p0z = (float)w0z;
    const float p1x = (float)(w0x + (axis == 0 ? cd.scale : 0));
    const float p1y = (float)(w0y + (axis == 1 ? cd.scale : 0));
    const float p1z = (float)(w0z + (axis == 2 ? cd.scale : 0));
    unsigned nh = 0, nc = 0;
    const float colH = axis == 1 ? height_term(Fp, T, p0x, p0z) : 0.f;
    nh += (axis == 1 && sub == 0) ? 1u : 0u;
    const float* hl = slot_ptr<const float>(a, chunk, a.L.h_lines) + hline_index(G.H, axis == 2 ? 1 : 0, x, z);
    float minV = 3.402823466e+38f, Hbest = colH;
    int kbest = 17;
    const int k_first = sub == 0 ? kG : sub;  // this lane's first interior sample
    const bool lines = a.height_lines != 0 && axis != 1;
    float Hn = (lines && k_first <= kLineSamples) ? hl[k_first - 1] : 0.f;
    for (int k = sub; k <= 16; k += kG) {
        const float cur = (float)k * 0.0625f;  // = the sequential scan's running sum (exact)
        const float px = p0x + (p1x - p0x) * cur;
        const float py = p0y + (p1y - p0y) * cur;
        const float pz = p0z + (p1z - p0z) * cur;
        float d, H = colH;
        if (ends && k == 0) {
            d = fabsf(d_start);
        } else if (ends && k == 16) {
            d = fabsf(d_end);
        } else {
            if (axis != 1) {
                if (lines && k >= 1 && k <= kLineSamples) {
                    H = Hn;
                    if (k + kG <= kLineSamples) Hn = hl[k + kG - 1];
                } else {
                    H = height_term(Fp, T, px, pz);
                    ++nh;
                }
            }
            nc += (caves && !((py - H) > Fp.cave_cut)) ? 1u : 0u;
            d = fabsf(combine(Fp, T, py - H, px, py, pz));
        }
        if (d < minV) {
            minV = d;
            kbest = k;
            Hbest = H;
        }
    }
#pragma unroll
    for (int m = 1; m < kG; m <<= 1) {
        const float ov = __shfl_xor(minV, m, kG);
        const int ok = __shfl_xor(kbest, m, kG);
        const float oh = __shfl_xor(Hbest, m, kG);
        if (ov < minV || (ov == minV && ok < kbest)) {
            minV = ov;
            kbest = ok;
            Hbest = oh;
        }
    }
    const float tbest = (float)kbest * 0.0625f;
    const float px = p0x + (p1x - p0x) * tbest;
    const float py = p0y + (p1y - p0y) * tbest;
    const float pz = p0z + (p1z - p0z) * tbest;
    const float h = 0.001f;
    const bool hb_known = axis == 1 || !(ends && (kbest == 0 || kbest == 16));
    constexpr int kR = (6 + kG - 1) / kG;
    float vals[kR];
#pragma unroll
    for (int r = 0; r < kR; ++r) {
        vals[r] = 0.f;
        const int j = r * kG + sub;  // normal sample j: +x, -x, +y, -y, +z, -z
        if (j >= 6) continue;
        const float sx = j == 0 ? h : j == 1 ? -h : 0.f;
        const float sz = j == 4 ? h : j == 5 ? -h : 0.f;
        float H;
        if (j == 2 || j == 3) {
            H = hb_known ? Hbest : height_term(Fp, T, px, pz);
            nh += (!hb_known && j == 2) ? 1u : 0u;
        } else {
            H = height_term(Fp, T, j < 2 ? px + sx : px, j < 2 ? pz : pz + sz);
            ++nh;
        }
        const float qx = j < 2 ? px + sx : px;
        const float qy = j == 2 ? py + h : j == 3 ? py - h : py;
        const float qz = j >= 4 ? pz + sz : pz;
        const float aa = qy - H;
        nc += (caves && !(aa > Fp.cave_cut)) ? 1u : 0u;
        vals[r] = combine(Fp, T, aa, qx, qy, qz);
    }
    float v[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = __shfl(vals[j / kG], j % kG, kG);
    const float3 n = normalize_or_zero(v[0] - v[1], v[2] - v[3], v[4] - v[5]);
    if constexpr (kCount) {
        atomicAdd(&a.work[2], (unsigned long long)nh);
        atomicAdd(&a.work[3], (unsigned long long)nc);
        if (sub == 0) atomicAdd(&a.work[6], 1ull);
    }
    if (sub == 0) slot_ptr<float4>(a, chunk, a.L.e_data)[i] = make_float4(n.x, n.y, n.z, tbest);
}

template <bool kCount, int kG>
__global__ __attribute__((amdgpu_waves_per_eu(6, 8))) __launch_bounds__(kCrossThreads) void k_edge_cross_group(
    KArgs a, int chunks) {
    __shared__ NoiseLds T;
    load_noise_tables(T, a.perm, a.grads);
    constexpr int kPerItem = 64 / kG;
    const int total = a.item_prefix[chunks];
    const int lane = threadIdx.x & 63;
    for (;;) {
        const int item = next_item(&a.queue[1]);
        if (item >= total) break;
        int lo = 0, hi = chunks;  // last chunk with item_prefix[c] <= item
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.item_prefix[mid] <= item) lo = mid; else hi = mid;
        }
        const int i = edge_of_item(a, lo, item - a.item_prefix[lo], lane / kG, kPerItem);
        if (i >= 0) cross_edge_group<kCount, kG>(a, T, lo, i, lane & (kG - 1));
    }
}

// Sample-parallel search.  Lanes of different edges evaluate points a voxel or more
// apart (32 voxels in a coarse chunk), so their noise-table reads hit unrelated LDS
// addresses and conflict; lanes of one edge evaluate points a fraction of a voxel apart
// and mostly read the same table entries.  A wave takes a 64-edge item: each lane first
// loads its own edge (edges with brush data or a cached list are finished there by the
// one-lane path, cross_edge); the search then runs four edges at a time, 16 lanes each
// (lane s evaluates interior sample s + 1, lane 15 the end samples, read from the stored
// field when it is FP32), reduced over the row to the first arg-min (smallest |d|, then
// smallest k: the sequential scan's rule); finally each lane computes its own edge's
// normal from the search's result (crossing_normal).  Every value comes from the same
// expression as in density_crossing, so results are identical.  A y-edge's height term is
// its column's, stored by the density pass.
constexpr int kSpWaves = kCrossThreads / 64;

// v from the lane kN places further round its row of 16 (DPP row_ror).
template <int kN>
__device__ __forceinline__ int row_ror(int v) {
    return __builtin_amdgcn_update_dpp(v, v, 0x120 + kN, 0xf, 0xf, false);
}

// One step of the row's arg-min: keep the smaller (|d| bits, k) pair and its height.
template <int kN>
__device__ __forceinline__ void argmin_step(uint32_t& d, int& k, float& H) {
    const uint32_t od = (uint32_t)row_ror<kN>((int)d);
    const int ok = row_ror<kN>(k);
    const float oh = __int_as_float(row_ror<kN>(__float_as_int(H)));
    const bool take = od < d || (od == d && ok < k);
    d = take ? od : d;
    k = take ? ok : k;
    H = take ? oh : H;
}

template <bool kCount>
__global__ __attribute__((amdgpu_waves_per_eu(6, 8))) __launch_bounds__(kCrossThreads) void k_edge_cross_sp(KArgs a,
                                                                                                             int chunks) {
    __shared__ NoiseLds T;
    __shared__ int sp_k[kSpWaves][64];    // per wave: kbest of each edge of the item
    __shared__ float sp_h[kSpWaves][64];  // and the height term there
    load_noise_tables(T, a.perm, a.grads);
    const int total = a.item_prefix[chunks];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane >> 4, sub = lane & 15;
    const FieldParams& Fp = a.F;
    const bool caves = Fp.kind == LVN_FIELD_TERRAIN_CAVES;
    const bool ends = !a.fp16;
    const Geometry& G = a.G;
    unsigned nh = 0, nc = 0, ne = 0;
    for (;;) {
        const int item = next_item(&a.queue[1]);
        if (item >= total) break;
        int lo = 0, hi = chunks;  // last chunk with item_prefix[c] <= item
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.item_prefix[mid] <= item) lo = mid; else hi = mid;
        }
        const int chunk = lo;
        const ChunkDesc& cd = a.chunks[chunk];
        // ---- each lane loads its own edge
        int ei = edge_of_item(a, chunk, item - a.item_prefix[chunk], lane, 64);
        int exyz = 0;
        float eH = 0.f, eS = 0.f, eE = 0.f;
        if (ei >= 0) {
            const int id = slot_ptr<int>(a, chunk, a.L.e_ids)[ei];
            const int axis = id & 3, hidx = id >> 2, msk = (1 << G.shift) - 1;
            const int x = hidx & msk, y = (hidx >> G.shift) & msk, z = (hidx >> (2 * G.shift)) & msk;
            if (x >= G.H || y >= G.H || z >= G.H || axis > 2) {  // never produced by k_edge_compact
                atomicOr(&a.counters[chunk].status, kStatusLookupMiss);
                ei = -1;
            } else {
                const int i0 = field_index(G, x, y, z);
                const int i1 = i0 + (axis == 0 ? 1 : axis == 1 ? G.F : G.F * G.F);
                const uint8_t* mat = slot_ptr<uint8_t>(a, chunk, a.L.mat);
                const uint8_t* orig = slot_ptr<uint8_t>(a, chunk, a.L.orig);
                const bool touched = cd.op_count > 0 && (mat[i0] != orig[i0] || mat[i1] != orig[i1]);
                if (touched || cd.mode == kModeCached) {
                    cross_edge<kCount>(a, T, chunk, ei);
                    ei = -1;
                } else {
                    exyz = x | (y << 8) | (z << 16) | (axis << 24);
                    eH = axis == 1 ? slot_ptr<const float>(a, chunk, a.L.h_cols)[x + G.F * z] : 0.f;
                    if (ends) {
                        const float2 de = edge_ends(a, chunk, cd, x, y, z, axis, i0, i1);
                        eS = de.x;
                        eE = de.y;
                    }
                }
            }
        }
        // ---- search: four edges at a time, 16 lanes each
        for (int r = 0; r < 16; ++r) {
            const int src = 4 * r + row;
            // every lane takes part in the shuffles: a lane of a row that skips the round
            // may be the source of another row's edge
            const int i = __shfl(ei, src);
            const int xyz = __shfl(exyz, src);
            const float colH = __shfl(eH, src);
            const float dS = __shfl(eS, src), dE = __shfl(eE, src);
            if (i < 0) continue;  // uniform over the row
            const int x = xyz & 255, y = (xyz >> 8) & 255, z = (xyz >> 16) & 255, axis = xyz >> 24;
            const int w0x = x * cd.scale + cd.off[0], w0y = y * cd.scale + cd.off[1], w0z = z * cd.scale + cd.off[2];
            const float p0x = (float)w0x, p0y = (float)w0y, p0z = (float)w0z;
            const float p1x = (float)(w0x + (axis == 0 ? cd.scale : 0));
            const float p1y = (float)(w0y + (axis == 1 ? cd.scale : 0));
            const float p1z = (float)(w0z + (axis == 2 ? cd.scale : 0));
            const bool lines = a.height_lines != 0 && axis != 1;
            // the lane's candidate: (|d| bits, k, H); none = (FLT_MAX, 17, colH)
            uint32_t md = 0x7f7fffffu;
            int mk = 17;
            float mh = colH;
            const int nk = sub < 15 ? 1 : 2;
            for (int q = 0; q < nk; ++q) {
                const int k = sub < 15 ? sub + 1 : 16 * q;
                const float cur = (float)k * 0.0625f;  // = the sequential scan's running sum (exact)
                float d, H = colH;
                if (ends && k == 0) {
                    d = fabsf(dS);
                } else if (ends && k == 16) {
                    d = fabsf(dE);
                } else {
                    const float px = p0x + (p1x - p0x) * cur;
                    const float py = p0y + (p1y - p0y) * cur;
                    const float pz = p0z + (p1z - p0z) * cur;
                    if (axis != 1) {
                        if (lines && k >= 1 && k <= kLineSamples) {
                            H = slot_ptr<const float>(a, chunk, a.L.h_lines)[hline_index(G.H, axis == 2 ? 1 : 0, x, z) +
                                                                             (k - 1)];
                        } else {
                            H = height_term(Fp, T, px, pz);
                            ++nh;
                        }
                    }
                    nc += (caves && !((py - H) > Fp.cave_cut)) ? 1u : 0u;
                    d = fabsf(combine(Fp, T, py - H, px, py, pz));
                }
                // the scan takes a sample only when |d| < its running minimum (from FLT_MAX):
                // lane 15 tries sample 0 before sample 16
                if (d < __uint_as_float(md)) {
                    md = __float_as_uint(d);
                    mk = k;
                    mh = H;
                }
            }
            // |d| >= 0, so its bits order like its values
            argmin_step<1>(md, mk, mh);
            argmin_step<2>(md, mk, mh);
            argmin_step<4>(md, mk, mh);
            argmin_step<8>(md, mk, mh);
            if (sub == 0) {
                sp_k[wave][src] = mk == 17 ? 0 : mk;  // no sample taken: the scan keeps t = 0
                sp_h[wave][src] = mh;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // ---- normal: each lane its own edge
        if (ei >= 0) {
            const int kb = sp_k[wave][lane];
            const float hb = sp_h[wave][lane];
            const int x = exyz & 255, y = (exyz >> 8) & 255, z = (exyz >> 16) & 255, axis = exyz >> 24;
            const int w0x = x * cd.scale + cd.off[0], w0y = y * cd.scale + cd.off[1], w0z = z * cd.scale + cd.off[2];
            const float p1x = (float)(w0x + (axis == 0 ? cd.scale : 0));
            const float p1y = (float)(w0y + (axis == 1 ? cd.scale : 0));
            const float p1z = (float)(w0z + (axis == 2 ? cd.scale : 0));
            slot_ptr<float4>(a, chunk, a.L.e_data)[ei] =
                crossing_normal(Fp, T, axis, ends, (float)kb * 0.0625f, hb, (float)w0x, (float)w0y, (float)w0z, p1x, p1y,
                                p1z, nh, nc);
            ++ne;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if constexpr (kCount) {
        atomicAdd(&a.work[2], (unsigned long long)nh);
        atomicAdd(&a.work[3], (unsigned long long)nc);
        atomicAdd(&a.work[6], (unsigned long long)ne);
    }
}

// ---------------------------------------------------------------------------
// Active voxels (FindActiveVoxels + CompactVoxels, cl/octree.cl:142-223)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int corner_materials(const KArgs& a, const uint8_t* mat, int x, int y, int z, int* cm) {
    const Geometry& G = a.G;
    int corners = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cm[i] = mat[field_index(G, x + ((i >> 2) & 1), y + ((i >> 1) & 1), z + (i & 1))];
        corners |= (cm[i] == kAir ? 0 : 1) << i;
    }
    return corners;
}

// The same from the solid rows, for a chunk that stores no materials: a sample's material
// is then the default material where it is not air (density < 0) and air elsewhere.
__device__ __forceinline__ int corner_materials_rows(const KArgs& a, const uint64_t* S, int x, int y, int z, int* cm) {
    const int F = a.G.F;
    int corners = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int cx = x + ((i >> 2) & 1), s = (y + ((i >> 1) & 1)) + F * (z + (i & 1));
        const int bit = (int)((S[2 * s + (cx >> 6)] >> (cx & 63)) & 1ull);
        cm[i] = bit ? a.defmat : kAir;
        corners |= bit << i;
    }
    return corners;
}

// FindDominantMaterial, cl/octree.cl:78-138.  The reference sorts the eight corner
// materials by insertion; any correct sort yields the same sequence, and Batcher's
// odd-even merge network (19 compare-exchanges, min/max pairs) does it without the
// insertion sort's data-dependent loops, which diverge across a wave.
__device__ __forceinline__ void sort_exchange(int& a, int& b) {
    const int lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

__device__ __forceinline__ int dominant_material(const int* m) {
    int d[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = m[i];
    sort_exchange(d[0], d[1]); sort_exchange(d[2], d[3]); sort_exchange(d[4], d[5]); sort_exchange(d[6], d[7]);
    sort_exchange(d[0], d[2]); sort_exchange(d[1], d[3]); sort_exchange(d[4], d[6]); sort_exchange(d[5], d[7]);
    sort_exchange(d[1], d[2]); sort_exchange(d[5], d[6]);
    sort_exchange(d[0], d[4]); sort_exchange(d[1], d[5]); sort_exchange(d[2], d[6]); sort_exchange(d[3], d[7]);
    sort_exchange(d[2], d[4]); sort_exchange(d[3], d[5]);
    sort_exchange(d[1], d[2]); sort_exchange(d[3], d[4]); sort_exchange(d[5], d[6]);
    int current = d[0], count = 1, maxCount = 0, maxMaterial = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        const int v = d[i];
        if (v == kAir || v == kNone) continue;
        if (current != v) {
            if (count > maxCount) {
                maxCount = count;
                maxMaterial = current;
            }
            current = v;
            count = 1;
        } else {
            ++count;
        }
    }
    if (count > maxCount) maxMaterial = current;
    return maxMaterial;
}

// A voxel is active when its 8 corners are not all solid or all air: over a
// voxel row, the AND and OR of the four sample rows around it and their x + 1
// shifts.  Segment `seg` of 64 voxels holds 64 / N voxel rows.
__global__ __launch_bounds__(256) void k_voxel_rows(KArgs a) {
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const int N = G.N, F = G.F;
    const uint64_t* S = slot_ptr<const uint64_t>(a, chunk, a.L.solid);
    uint64_t* vmask = slot_ptr<uint64_t>(a, chunk, a.L.v_mask);
    int* vcount = slot_ptr<int>(a, chunk, a.L.v_count);
    const int per_seg = N >= 64 ? 1 : 64 / N;
    const uint64_t nmask = N >= 64 ? ~0ull : ((1ull << N) - 1ull);
    for (int seg = thread_in_chunk(); seg < G.voxel_segs; seg += threads_per_chunk()) {
        uint64_t m = 0;
        for (int k = 0; k < per_seg; ++k) {
            const int vr = seg * per_seg + k;  // voxel row y + N z
            const int y = vr & (N - 1), z = vr >> G.depth;
            const int s0 = y + F * z;
            const int rows[4] = {s0, s0 + 1, s0 + F, s0 + F + 1};
            uint64_t all = ~0ull, any = 0ull;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint64_t lo = S[2 * rows[c]], hi = S[2 * rows[c] + 1];
                const uint64_t nx = shift_next(lo, hi);
                all &= lo & nx;
                any |= lo | nx;
            }
            m |= ((any & ~all) & nmask) << (k * N);
        }
        vmask[seg] = m;
        vcount[seg] = __popcll(m);
    }
}

// The octree codes of every segment's active voxels in linear order; the corner
// materials, edge mask and dominant material are found by the QEF kernel, which reads
// the corners anyway.  A wave takes 64 consecutive segments and loads their masks and
// bases together (lane j: segment j).  Where its live segments are dense (kDenseRow
// voxels or more on average) it walks them in order with lane b holding the segment's
// voxel b: its rank is the popcount of the mask below b, so each segment's codes are one
// contiguous store.  Where they are sparse, each lane walks its own segment (one lane
// per segment alone scattered the dense segments' stores: 4.4 bytes reached HBM per
// byte of codes on the clipmap batch).  Both walks write the same codes.
__global__ __launch_bounds__(256) void k_voxel_compact(KArgs a) {
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const uint64_t* vmask = slot_ptr<const uint64_t>(a, chunk, a.L.v_mask);
    const int* vbase = slot_ptr<const int>(a, chunk, a.L.v_base);
    int* vcode = slot_ptr<int>(a, chunk, a.L.v_code);
    const int cap = a.caps.voxels;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt(lane);
    const int groups = (G.voxel_segs + 63) >> 6;
    for (int g = wave_in_chunk(); g < groups; g += waves_per_chunk()) {
        const int sl = (g << 6) + lane;
        uint64_t ml = 0;
        int bl = 0;
        if (sl < G.voxel_segs) {
            ml = vmask[sl];
            bl = vbase[sl];
        }
        uint64_t live = __ballot(ml != 0);
        if (!live) continue;
        if (wave_sum(__popcll(ml)) < kDenseRow * __popcll(live)) {  // sparse: one lane per segment
            int i = bl;
            while (ml && i < cap) {
                const int v = (sl << 6) + (int)__builtin_ctzll(ml);
                ml &= ml - 1ull;
                int x, y, z;
                decode_voxel(G, v, x, y, z);
                vcode[i++] = (int)code_for_position(x, y, z, G.depth);
            }
            continue;
        }
        while (live) {
            const int j = (int)__builtin_ctzll(live);
            live &= live - 1ull;
            const int seg = (g << 6) + j;
            const uint64_t m = readlane64(ml, j);
            const int i = __builtin_amdgcn_readlane(bl, j) + __popcll(m & lt);
            if (((m >> lane) & 1ull) && i < cap) {
                int x, y, z;
                decode_voxel(G, (seg << 6) + lane, x, y, z);
                vcode[i] = (int)code_for_position(x, y, z, G.depth);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// QEF build + solve (CreateLeafNodes + SolveQEFs, cl/octree.cl:236-331,
// cl/qef.cl).  Edge data is found through the edge segment masks instead of a
// cuckoo hash (cl/cuckoo.cl:73-104).
// ---------------------------------------------------------------------------

struct Q4 {
    float x, y, z, w;
};

__device__ __forceinline__ float dot4(const Q4& a, const Q4& b) {
    return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

// Both square roots have arguments >= 1 and the cosine's divisor lies in [1, sqrt 2]
// (|tn| <= 1), so they take the short correctly rounded forms; the other two
// divisions see any magnitude (tau overflows to inf for a tiny apq).
__device__ __forceinline__ void givens(float app, float apq, float aqq, float& c, float& s) {  // cl/qef.cl:31-42
    const float tau = (aqq - app) / (2.f * apq);
    const float stt = sqrt_ge1(1.f + tau * tau);
    const float tn = 1.f / ((tau >= 0.f) ? (tau + stt) : (tau - stt));
    c = recip_1_2(sqrt_ge1(1.f + tn * tn));
    s = tn * c;
}

// svd_rotate, cl/qef.cl:58-86, for the pair (A, B); U/W are the other two
// off-diagonal entries rotated with it (vtav[0][3-b], vtav[1-a][2]).
#define LVN_ROTATE(AA, BB, AB, U, W, VA0, VB0, VA1, VB1, VA2, VB2)                   \
    if (AB != 0.f) {                                                                   \
        float c, s;                                                                    \
        givens(AA, AB, BB, c, s);                                                      \
        const float cc = c * c, ss = s * s;                                            \
        const float mx = ((2.f * c) * s) * AB;                                         \
        const float u0 = AA, w0 = BB;                                                  \
        AA = (cc * u0 - mx) + ss * w0;                                                 \
        BB = (ss * u0 + mx) + cc * w0;                                                 \
        const float u1 = U, w1 = W;                                                    \
        U = c * u1 - s * w1;                                                           \
        W = s * u1 + c * w1;                                                           \
        AB = 0.f;                                                                      \
        float uu = VA0, ww = VB0;                                                      \
        VA0 = c * uu - s * ww;                                                         \
        VB0 = s * uu + c * ww;                                                         \
        uu = VA1; ww = VB1;                                                            \
        VA1 = c * uu - s * ww;                                                         \
        VB1 = s * uu + c * ww;                                                         \
        uu = VA2; ww = VB2;                                                            \
        VA2 = c * uu - s * ww;                                                         \
        VB2 = s * uu + c * ww;                                                         \
    }

__device__ __forceinline__ float invdet(float x) {  // svd_invdet, cl/qef.cl:107-109
    return (fabsf(x) < 0.1f || fabsf(1.f / x) < 0.1f) ? 0.f : 1.f / x;
}

__device__ __forceinline__ void qef_fast(const KArgs& a, int chunk, const ChunkDesc& cd, int3 p, int emask, float4& out,
                                         float4& nrm) {
    const Geometry& G = a.G;
    const uint64_t* erows = slot_ptr<const uint64_t>(a, chunk, a.L.e_rows);
    const int* eb = slot_ptr<const int>(a, chunk, a.L.e_base);
    const float4* ed = slot_ptr<float4>(a, chunk, a.L.e_data);
    const float fscale = (float)cd.scale;

    // The voxel's edges start at its corners, so they lie in four edge rows, (y or y + 1,
    // z or z + 1): those rows and bases are loaded together and every edge's index is
    // found from them, so only the hermite-data loads remain in the accumulation loop
    // (loading those ahead as well was slower: 99 VGPRs, 4 waves per SIMD).
    uint64_t rw[4][4];
    int rb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int r = (p.y + (j & 1)) + G.H * (p.z + (j >> 1));
#pragma unroll
        for (int q = 0; q < 4; ++q) rw[j][q] = erows[4 * r + q];
        rb[j] = eb[r];
    }
    int ke[12];
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const int e0 = kEdgeMap[e][0];
        const int j = ((e0 >> 1) & 1) + 2 * (e0 & 1);  // row of the edge's lower end
        const int k = ((emask >> e) & 1) ? edge_index_in_row(rw[j], rb[j], p.x + ((e0 >> 2) & 1), e >> 2) : -1;
        ke[e] = k >= 0 && k < a.caps.edges ? k : -1;
    }

    // The edge's point is a0 + (a1 - a0) t per component (cl/octree.cl:266-272), scaled.  Its
    // end corners differ by exactly 1 along the edge's axis and by exactly 0 across it, and t
    // is a finite crossing in [0, 1], so (a1 - a0) t is t along the axis and +0 across it:
    // the across components are fscale * a0 (six per voxel, formed once) and the w lane is
    // +0, which leaves the dot product's fourth term +0 and makes the w sums count edges.
    const float sx0 = fscale * (float)p.x, sx1 = fscale * (float)(p.x + 1);
    const float sy0 = fscale * (float)p.y, sy1 = fscale * (float)(p.y + 1);
    const float sz0 = fscale * (float)p.z, sz1 = fscale * (float)(p.z + 1);
    float ata0 = 0.f, ata1 = 0.f, ata2 = 0.f, ata3 = 0.f, ata4 = 0.f, ata5 = 0.f;
    Q4 atb = {0.f, 0.f, 0.f, 0.f}, mass = {0.f, 0.f, 0.f, 0.f}, nsum = {0.f, 0.f, 0.f, 0.f};
    int count = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        if (ke[e] < 0) continue;
        const int e0 = kEdgeMap[e][0], axis = e >> 2;
        const int hx = p.x + ((e0 >> 2) & 1), hy = p.y + ((e0 >> 1) & 1), hz = p.z + (e0 & 1);
        const float4 d = ed[ke[e]];
        Q4 pos, n;
        pos.x = axis == 0 ? fscale * ((float)hx + d.w) : ((e0 >> 2) & 1) ? sx1 : sx0;
        pos.y = axis == 1 ? fscale * ((float)hy + d.w) : ((e0 >> 1) & 1) ? sy1 : sy0;
        pos.z = axis == 2 ? fscale * ((float)hz + d.w) : (e0 & 1) ? sz1 : sz0;
        n.x = d.x;
        n.y = d.y;
        n.z = d.z;
        // qef_add_point, cl/qef.cl:170-191
        ata0 = ata0 + n.x * n.x;
        ata1 = ata1 + n.x * n.y;
        ata2 = ata2 + n.x * n.z;
        ata3 = ata3 + n.y * n.y;
        ata4 = ata4 + n.y * n.z;
        ata5 = ata5 + n.z * n.z;
        const float b = (((pos.x * n.x + pos.y * n.y) + pos.z * n.z) + 0.f);
        atb.x = atb.x + n.x * b;
        atb.y = atb.y + n.y * b;
        atb.z = atb.z + n.z * b;
        mass.x = mass.x + pos.x;
        mass.y = mass.y + pos.y;
        mass.z = mass.z + pos.z;
        nsum.x = nsum.x + n.x;
        nsum.y = nsum.y + n.y;
        nsum.z = nsum.z + n.z;
        ++count;
    }
    mass.w = (float)count;  // the sums of 1 per edge
    nsum.w = (float)count;
    // qef_create_from_points tail (cl/qef.cl:302), then qef_solve's divide by max(w, 1) (:244).
    {
        const float w = mass.w;
        mass.x = mass.x / w;
        mass.y = mass.y / w;
        mass.z = mass.z / w;
        mass.w = mass.w / w;
        const float mw = fmaxf(mass.w, 1.f);
        mass.x = mass.x / mw;
        mass.y = mass.y / mw;
        mass.z = mass.z / mw;
        mass.w = mass.w / mw;
    }
    Q4 amp;
    {
        const Q4 row = {ata0, ata1, ata2, 0.f};
        amp.x = atb.x - dot4(row, mass);
        amp.y = atb.y - ((ata1 * mass.x + ata3 * mass.y) + ata4 * mass.z);
        amp.z = atb.z - ((ata2 * mass.x + ata4 * mass.y) + ata5 * mass.z);
        amp.w = atb.w - 0.f;
    }
    // svd_solve_sym, 10 cyclic sweeps (cl/qef.cl:88-105).
    float v00 = ata0, v01 = ata1, v02 = ata2, v11 = ata3, v12 = ata4, v22 = ata5;
    float V00 = 1.f, V01 = 0.f, V02 = 0.f, V10 = 0.f, V11 = 1.f, V12 = 0.f, V20 = 0.f, V21 = 0.f, V22 = 1.f;
    for (int sweep = 0; sweep < 10; ++sweep) {
        LVN_ROTATE(v00, v11, v01, v02, v12, V00, V01, V10, V11, V20, V21)
        LVN_ROTATE(v00, v22, v02, v01, v12, V00, V02, V10, V12, V20, V22)
        LVN_ROTATE(v11, v22, v12, v01, v02, V01, V02, V11, V12, V21, V22)
    }
    const float d0 = invdet(v00), d1 = invdet(v11), d2 = invdet(v22);
    const float Vm[3][3] = {{V00, V01, V02}, {V10, V11, V12}, {V20, V21, V22}};
    float x[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        Q4 row;
        row.x = ((Vm[r][0] * d0) * Vm[0][0] + (Vm[r][1] * d1) * Vm[0][1]) + (Vm[r][2] * d2) * Vm[0][2];
        row.y = ((Vm[r][0] * d0) * Vm[1][0] + (Vm[r][1] * d1) * Vm[1][1]) + (Vm[r][2] * d2) * Vm[1][2];
        row.z = ((Vm[r][0] * d0) * Vm[2][0] + (Vm[r][1] * d1) * Vm[2][1]) + (Vm[r][2] * d2) * Vm[2][2];
        row.w = 0.f;
        x[r] = dot4(row, amp);
    }
    out.x = ((x[0] + mass.x) * 4.f) + (float)cd.min[0];
    out.y = ((x[1] + mass.y) * 4.f) + (float)cd.min[1];
    out.z = ((x[2] + mass.z) * 4.f) + (float)cd.min[2];
    out.w = 1.f;
    const float nw = nsum.w;
    nrm = make_float4(nsum.x / nw, nsum.y / nw, nsum.z / nw, 0.f);
}

// ---------------------------------------------------------------------------
// Precision form of the QEF (LVN_CTX_PRECISE_QEF).  The normal equations are
// formed about the mass point, A = sum n n^T and r = sum n (n . (p - m)) (the
// same least-squares problem as ATb - ATA m, without its cancellation), with
// every product split exactly by an FMA (TwoProduct) and summed with Neumaier
// compensation.  The Jacobi pseudo-inverse runs in FP32; a vertex whose
// condition number exceeds 1e6, or with an eigenvalue within 1e-3 (relative) of
// the truncation thresholds 0.1 / 10, is re-solved entirely in FP64.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void neumaier(float& s, float& c, float x) {
    const float t = s + x;
    c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

__device__ __forceinline__ void add_product(float& s, float& c, float a, float b) {
    const float p = a * b;
    const float e = fmaf(a, b, -p);  // exact error of the product
    neumaier(s, c, p);
    c += e;
}

template <typename T>
__device__ __forceinline__ void jacobi_rotate(T& aa, T& bb, T& ab, T& u, T& w, T& va0, T& vb0, T& va1, T& vb1, T& va2,
                                              T& vb2) {
    if (ab == (T)0) return;
    const T tau = (bb - aa) / ((T)2 * ab);
    const T stt = sqrt((T)1 + tau * tau);
    const T tn = (T)1 / ((tau >= (T)0) ? (tau + stt) : (tau - stt));
    const T c = (T)1 / sqrt((T)1 + tn * tn);
    const T s = tn * c;
    const T cc = c * c, ss = s * s, mx = (((T)2 * c) * s) * ab;
    const T u0 = aa, w0 = bb;
    aa = (cc * u0 - mx) + ss * w0;
    bb = (ss * u0 + mx) + cc * w0;
    const T u1 = u, w1 = w;
    u = c * u1 - s * w1;
    w = s * u1 + c * w1;
    ab = (T)0;
    T x = va0, y = vb0;
    va0 = c * x - s * y;
    vb0 = s * x + c * y;
    x = va1; y = vb1;
    va1 = c * x - s * y;
    vb1 = s * x + c * y;
    x = va2; y = vb2;
    va2 = c * x - s * y;
    vb2 = s * x + c * y;
}

// 10 cyclic Jacobi sweeps + truncated pseudo-inverse applied to r (cl/qef.cl:88-144).
template <typename T>
__device__ __forceinline__ void qef_pinv_solve(const T A[6], const T r[3], T x[3], T sig[3]) {
    T v00 = A[0], v01 = A[1], v02 = A[2], v11 = A[3], v12 = A[4], v22 = A[5];
    T V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 10; ++sweep) {
        jacobi_rotate<T>(v00, v11, v01, v02, v12, V[0][0], V[0][1], V[1][0], V[1][1], V[2][0], V[2][1]);
        jacobi_rotate<T>(v00, v22, v02, v01, v12, V[0][0], V[0][2], V[1][0], V[1][2], V[2][0], V[2][2]);
        jacobi_rotate<T>(v11, v22, v12, v01, v02, V[0][1], V[0][2], V[1][1], V[1][2], V[2][1], V[2][2]);
    }
    sig[0] = v00;
    sig[1] = v11;
    sig[2] = v22;
    const T tol = (T)0.1;
    T d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        d[i] = (fabs(sig[i]) < tol || fabs((T)1 / sig[i]) < tol) ? (T)0 : (T)1 / sig[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T acc = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
            acc += ((V[i][0] * d[0]) * V[j][0] + (V[i][1] * d[1]) * V[j][1] + (V[i][2] * d[2]) * V[j][2]) * r[j];
        x[i] = acc;
    }
}

// The truncated pseudo-inverse keeps eigenvalues in [0.1, 10] (svd_invdet, cl/qef.cl:107-109),
// so the retained ones have a condition number of at most 100 and FP32 solves them well;
// what FP32 can get wrong is which eigenvalues are retained.  FP32 Jacobi's eigenvalues
// are within a small multiple of eps * |A|_F of the exact ones (backward stable, about 30
// rotations here): an eigenvalue that close to 0.1 or 10 is re-solved in FP64.
constexpr float kThresholdWindow = 256.f * 5.9604645e-8f;  // 256 eps, times |A|_F

__device__ __forceinline__ bool near_threshold(float s, float w) {
    const float a = fabsf(s);
    return fabsf(a - 0.1f) <= w || fabsf(a - 10.f) <= w;
}

__device__ __forceinline__ void qef_precise(const KArgs& a, int chunk, const ChunkDesc& cd, int3 p, int emask, float4& out,
                                            float4& nrm) {
    const Geometry& G = a.G;
    const uint64_t* erows = slot_ptr<const uint64_t>(a, chunk, a.L.e_rows);
    const int* eb = slot_ptr<const int>(a, chunk, a.L.e_base);
    const float4* ed = slot_ptr<float4>(a, chunk, a.L.e_data);
    const float fscale = (float)cd.scale;

    // edge indices from the voxel's four edge rows, loaded together (as in qef_fast)
    uint64_t rw[4][4];
    int rb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int r = (p.y + (j & 1)) + G.H * (p.z + (j >> 1));
#pragma unroll
        for (int q = 0; q < 4; ++q) rw[j][q] = erows[4 * r + q];
        rb[j] = eb[r];
    }
    int ke[12];
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        const int e0 = kEdgeMap[e][0];
        const int j = ((e0 >> 1) & 1) + 2 * (e0 & 1);
        const int k = ((emask >> e) & 1) ? edge_index_in_row(rw[j], rb[j], p.x + ((e0 >> 2) & 1), e >> 2) : -1;
        ke[e] = k >= 0 && k < a.caps.edges ? k : -1;
    }

    float px[12], py[12], pz[12], nx[12], ny[12], nz[12];
    bool ok[12];
    float cnt = 0.f;
    float sx = 0.f, cx = 0.f, sy = 0.f, cy = 0.f, sz = 0.f, cz = 0.f;
    float nsx = 0.f, nsy = 0.f, nsz = 0.f;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        ok[e] = false;
        px[e] = py[e] = pz[e] = nx[e] = ny[e] = nz[e] = 0.f;
        if (ke[e] < 0) continue;
        const int e0 = kEdgeMap[e][0], e1 = kEdgeMap[e][1];
        const int hx = p.x + ((e0 >> 2) & 1), hy = p.y + ((e0 >> 1) & 1), hz = p.z + (e0 & 1);
        const float4 d = ed[ke[e]];
        const float a0x = (float)hx, a0y = (float)hy, a0z = (float)hz;
        const float a1x = (float)(p.x + ((e1 >> 2) & 1)), a1y = (float)(p.y + ((e1 >> 1) & 1)),
                    a1z = (float)(p.z + (e1 & 1));
        ok[e] = true;
        px[e] = fscale * (a0x + (a1x - a0x) * d.w);
        py[e] = fscale * (a0y + (a1y - a0y) * d.w);
        pz[e] = fscale * (a0z + (a1z - a0z) * d.w);
        nx[e] = d.x;
        ny[e] = d.y;
        nz[e] = d.z;
        cnt += 1.f;
        neumaier(sx, cx, px[e]);
        neumaier(sy, cy, py[e]);
        neumaier(sz, cz, pz[e]);
        nsx = nsx + d.x;
        nsy = nsy + d.y;
        nsz = nsz + d.z;
    }
    const float mx = (sx + cx) / cnt, my = (sy + cy) / cnt, mz = (sz + cz) / cnt;
    float As[6] = {0, 0, 0, 0, 0, 0}, Ac[6] = {0, 0, 0, 0, 0, 0}, rs[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < 12; ++e) {
        if (!ok[e]) continue;
        add_product(As[0], Ac[0], nx[e], nx[e]);
        add_product(As[1], Ac[1], nx[e], ny[e]);
        add_product(As[2], Ac[2], nx[e], nz[e]);
        add_product(As[3], Ac[3], ny[e], ny[e]);
        add_product(As[4], Ac[4], ny[e], nz[e]);
        add_product(As[5], Ac[5], nz[e], nz[e]);
        const float b = fmaf(nx[e], px[e] - mx, fmaf(ny[e], py[e] - my, nz[e] * (pz[e] - mz)));
        add_product(rs[0], rc[0], nx[e], b);
        add_product(rs[1], rc[1], ny[e], b);
        add_product(rs[2], rc[2], nz[e], b);
    }
    float A[6], r[3], x[3], sig[3];
#pragma unroll
    for (int i = 0; i < 6; ++i) A[i] = As[i] + Ac[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] = rs[i] + rc[i];
    qef_pinv_solve<float>(A, r, x, sig);
    const float normf = sqrtf((A[0] * A[0] + A[3] * A[3] + A[5] * A[5]) + 2.f * (A[1] * A[1] + A[2] * A[2] + A[4] * A[4]));
    const float w = kThresholdWindow * normf;
    if (near_threshold(sig[0], w) || near_threshold(sig[1], w) || near_threshold(sig[2], w)) {
        double n = 0, mxd = 0, myd = 0, mzd = 0;
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            if (!ok[e]) continue;
            n += 1.0;
            mxd += (double)px[e];
            myd += (double)py[e];
            mzd += (double)pz[e];
        }
        mxd /= n;
        myd /= n;
        mzd /= n;
        double Ad[6] = {0, 0, 0, 0, 0, 0}, rd[3] = {0, 0, 0}, xd[3], sd[3];
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            if (!ok[e]) continue;
            const double a0 = nx[e], a1 = ny[e], a2 = nz[e];
            Ad[0] += a0 * a0;
            Ad[1] += a0 * a1;
            Ad[2] += a0 * a2;
            Ad[3] += a1 * a1;
            Ad[4] += a1 * a2;
            Ad[5] += a2 * a2;
            const double b = a0 * ((double)px[e] - mxd) + a1 * ((double)py[e] - myd) + a2 * ((double)pz[e] - mzd);
            rd[0] += a0 * b;
            rd[1] += a1 * b;
            rd[2] += a2 * b;
        }
        qef_pinv_solve<double>(Ad, rd, xd, sd);
        out = make_float4((float)((xd[0] + mxd) * 4.0 + (double)cd.min[0]),
                          (float)((xd[1] + myd) * 4.0 + (double)cd.min[1]),
                          (float)((xd[2] + mzd) * 4.0 + (double)cd.min[2]), 1.f);
        atomicAdd(&a.counters[chunk].resolved, 1);
    } else {
        out = make_float4(((x[0] + mx) * 4.f) + (float)cd.min[0], ((x[1] + my) * 4.f) + (float)cd.min[1],
                          ((x[2] + mz) * 4.f) + (float)cd.min[2], 1.f);
    }
    // the vertex normal is the reference's unnormalised mean of the edge normals
    const float nw = cnt;
    nrm = make_float4(nsx / nw, nsy / nw, nsz / nw, 0.f);
}

// ---------------------------------------------------------------------------
// Per active voxel, one pass: corner materials and edge mask (CreateLeafNodes,
// cl/octree.cl:236-283), dominant material (cl/octree.cl:78-138), the QEF vertex
// written straight into the vertex buffer (GenerateMeshVertexBuffer,
// cl/octree.cl:475-496), the voxel's quad flags (GenerateMesh, cl/octree.cl:398-439)
// and seam flag (FindSeamNodes, cl/octree.cl:500-518).  Per 64-voxel segment the
// quad count and the seam mask go to the quad and seam scans.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int quad_flag(const Geometry& G, int3 p, int corners, int axis) {
    const int pa = axis == 0 ? p.y : axis == 1 ? p.z : p.x;
    const int pb = axis == 0 ? p.z : axis == 1 ? p.x : p.y;
    if (pa == G.N - 1 || pb == G.N - 1) return 0;  // isEdgeVoxel, cl/octree.cl:411-417
    const int edge = axis * 4 + 3;
    const int m1 = (corners >> kEdgeMap[edge][0]) & 1, m2 = (corners >> kEdgeMap[edge][1]) & 1;
    return m1 != m2;
}

template <bool kPrecise>
__global__ __launch_bounds__(256) void k_qef(KArgs a) {
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const int V = a.counters[chunk].voxels;
    const ChunkDesc cd = a.chunks[chunk];
    const int* vcode = slot_ptr<const int>(a, chunk, a.L.v_code);
    const uint8_t* mat = slot_ptr<const uint8_t>(a, chunk, a.L.mat);
    const uint64_t* S = slot_ptr<const uint64_t>(a, chunk, a.L.solid);
    const bool rows = chunk_mat_free(a, cd);  // corners from the solid rows
    int* vemask = slot_ptr<int>(a, chunk, a.L.v_emask);
    int* vmat = slot_ptr<int>(a, chunk, a.L.v_mat);
    float4* verts = slot_ptr<float4>(a, chunk, a.L.verts);
    int* qcount = slot_ptr<int>(a, chunk, a.L.q_count);
    uint64_t* smask = slot_ptr<uint64_t>(a, chunk, a.L.s_mask);
    int* scount = slot_ptr<int>(a, chunk, a.L.s_count);
    const int lane = threadIdx.x & 63;
    const int N = G.N;
    for (int base = wave_in_chunk() << 6; base < V; base += waves_per_chunk() << 6) {
        const int vi = base + lane;
        int qbits = 0, seam = 0;
        if (vi < V) {
            const int3 p = position_for_code((uint32_t)vcode[vi], G.depth);
            int cm[8];
            const int corners = rows ? corner_materials_rows(a, S, p.x, p.y, p.z, cm)
                                     : corner_materials(a, mat, p.x, p.y, p.z, cm);
            int emask = 0;
#pragma unroll
            for (int e = 0; e < 12; ++e)
                emask |= (((corners >> kEdgeMap[e][0]) & 1) != ((corners >> kEdgeMap[e][1]) & 1)) << e;
            const int m = (dominant_material(cm) << 8) | corners;
            float4 pos, nrm;
            if constexpr (kPrecise)
                qef_precise(a, chunk, cd, p, emask, pos, nrm);
            else
                qef_fast(a, chunk, cd, p, emask, pos, nrm);
#pragma unroll
            for (int axis = 0; axis < 3; ++axis) qbits |= quad_flag(G, p, corners, axis) << axis;
            seam = p.x == 0 || p.x == N - 1 || p.y == 0 || p.y == N - 1 || p.z == 0 || p.z == N - 1;
            vemask[vi] = emask | (qbits << 12);
            vmat[vi] = m;
            verts[3 * vi] = pos;
            verts[3 * vi + 1] = nrm;
            verts[3 * vi + 2] = make_float4(cd.colour[0], cd.colour[1], cd.colour[2], (float)(m >> 8));
        }
        const int nq = __popc(qbits);
        const uint64_t b0 = __ballot(nq & 1), b1 = __ballot(nq & 2), sm = __ballot(seam);
        if (lane == 0) {
            const int seg = base >> 6;
            qcount[seg] = __popcll(b0) + 2 * __popcll(b1);
            smask[seg] = sm;
            scount[seg] = __popcll(sm);
        }
    }
}

// Quad and seam scans of one chunk in one workgroup (the segment count is the
// chunk's voxel count, known only on the device).
__global__ __launch_bounds__(1024) void k_scan_quads_seams(KArgs a) {
    __shared__ int buf[1024];
    const int chunk = blockIdx.y;
    Counters* cnt = a.counters + chunk;
    const int n = (cnt->voxels + 63) >> 6;
    const int q = block_scan(slot_ptr<const int>(a, chunk, a.L.q_count), slot_ptr<int>(a, chunk, a.L.q_base), n, buf);
    if (threadIdx.x == 0) {
        cnt->quads = min(q, a.caps.quads);
        if (q > a.caps.quads) atomicOr(&cnt->status, kStatusQuadOverflow);
    }
    const int sn = block_scan(slot_ptr<const int>(a, chunk, a.L.s_count), slot_ptr<int>(a, chunk, a.L.s_base), n, buf);
    if (threadIdx.x == 0) {
        cnt->seams = min(sn, a.caps.seams);
        if (sn > a.caps.seams) atomicOr(&cnt->status, kStatusSeamOverflow);
    }
}

// cl/octree.cl:376-381
__constant__ int kEdgeNodeOffsets[3][4][3] = {{{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}},
                                              {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}},
                                              {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0}}};

// Quads (GenerateMesh + CompactMeshTriangles, cl/octree.cl:335-464) in (voxel,
// axis) order, and seam nodes (ExtractSeamNodeInfo, cl/octree.cl:529-551) in
// voxel order.  A segment's quads fill one contiguous run of the index buffer
// (at most 3 per voxel): each wave builds its run in LDS and copies it out with
// 8-byte stores of consecutive addresses, so every store instruction writes whole
// lines instead of 64 scattered 12-byte pieces.
constexpr int kEmitWaves = 4;                  // waves per k_emit workgroup
constexpr int kEmitStageInts = 64 * 3 * 6;     // one segment's quads, 6 indices each
__global__ __launch_bounds__(64 * kEmitWaves) void k_emit(KArgs a) {
    __shared__ int2 stage_all[kEmitWaves][kEmitStageInts / 2];
    int* stage = reinterpret_cast<int*>(stage_all[threadIdx.x >> 6]);
    const int chunk = blockIdx.y;
    const Geometry& G = a.G;
    const int V = a.counters[chunk].voxels;
    const int* vcode = slot_ptr<const int>(a, chunk, a.L.v_code);
    const int* vemask = slot_ptr<const int>(a, chunk, a.L.v_emask);
    const int* vmat = slot_ptr<const int>(a, chunk, a.L.v_mat);
    const uint64_t* vm = slot_ptr<const uint64_t>(a, chunk, a.L.v_mask);
    const int* vb = slot_ptr<const int>(a, chunk, a.L.v_base);
    const int* qbase = slot_ptr<const int>(a, chunk, a.L.q_base);
    const uint64_t* smask = slot_ptr<const uint64_t>(a, chunk, a.L.s_mask);
    const int* sbase = slot_ptr<const int>(a, chunk, a.L.s_base);
    const float4* verts = slot_ptr<const float4>(a, chunk, a.L.verts);
    int* tris = slot_ptr<int>(a, chunk, a.L.tris);
    float4* seams = slot_ptr<float4>(a, chunk, a.L.seams);
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt(lane);
    for (int base = wave_in_chunk() << 6; base < V; base += waves_per_chunk() << 6) {
        const int vi = base + lane, seg = base >> 6;
        // the voxel's own words, loaded together (a lane past the end reads the last voxel)
        const int vc = min(vi, V - 1);
        const int em = vemask[vc], code = vcode[vc], m = vmat[vc], qb = qbase[seg];
        const uint64_t sm = smask[seg];
        const int qbits = vi < V ? (em >> 12) & 7 : 0;
        const int nq = __popc(qbits);
        const uint64_t b0 = __ballot(nq & 1), b1 = __ballot(nq & 2);
        const bool seam = (sm >> lane) & 1ull;
        const int run = __popcll(b0) + 2 * __popcll(b1);  // the segment's quads
        if (vi < V && (qbits || seam)) {
            const int3 p = position_for_code((uint32_t)code, G.depth);
            int qi = __popcll(b0 & lt) + 2 * __popcll(b1 & lt);  // within the segment's run
            // The nine neighbour lookups (three per axis) are loaded together, before any is
            // used: one memory round trip instead of one per axis and step.  An index is
            // clamped into the grid, so an axis without a quad reads a valid, unused word (a
            // quad's neighbours all lie in the grid: quad_flag excludes the last layers).
            uint64_t nw[3][3];
            int nb[3][3], nl[3][3];
            if (G.N == 64) {
                // A segment is the voxel row (y, z): the nine neighbours (kEdgeNodeOffsets) lie
                // in the rows (y, z), (y + 1, z), (y, z + 1) and (y + 1, z + 1), at x or x + 1.
                const int r = p.y + 64 * p.z, rmax = 64 * 64 - 1;
                const int r10 = min(r + 1, rmax), r01 = min(r + 64, rmax), r11 = min(r + 65, rmax);
                const uint64_t m00 = vm[r], m10 = vm[r10], m01 = vm[r01], m11 = vm[r11];
                const int b00 = vb[r], b10 = vb[r10], b01 = vb[r01], b11 = vb[r11];
                const int x0 = p.x, x1 = (p.x + 1) & 63;
                nw[0][0] = m01; nb[0][0] = b01; nl[0][0] = x0;  // (0, 0, 1)
                nw[0][1] = m10; nb[0][1] = b10; nl[0][1] = x0;  // (0, 1, 0)
                nw[0][2] = m11; nb[0][2] = b11; nl[0][2] = x0;  // (0, 1, 1)
                nw[1][0] = m00; nb[1][0] = b00; nl[1][0] = x1;  // (1, 0, 0)
                nw[1][1] = m01; nb[1][1] = b01; nl[1][1] = x0;  // (0, 0, 1)
                nw[1][2] = m01; nb[1][2] = b01; nl[1][2] = x1;  // (1, 0, 1)
                nw[2][0] = m10; nb[2][0] = b10; nl[2][0] = x0;  // (0, 1, 0)
                nw[2][1] = m00; nb[2][1] = b00; nl[2][1] = x1;  // (1, 0, 0)
                nw[2][2] = m10; nb[2][2] = b10; nl[2][2] = x1;  // (1, 1, 0)
            } else {
                const int last = G.N * G.N * G.N - 1;
                #pragma unroll
                for (int axis = 0; axis < 3; ++axis) {
                    #pragma unroll
                    for (int n = 1; n < 4; ++n) {
                        const int x = p.x + kEdgeNodeOffsets[axis][n][0];
                        const int y = p.y + kEdgeNodeOffsets[axis][n][1];
                        const int z = p.z + kEdgeNodeOffsets[axis][n][2];
                        const int i = min(x + G.N * (y + G.N * z), last);
                        nw[axis][n - 1] = vm[i >> 6];
                        nb[axis][n - 1] = vb[i >> 6];
                        nl[axis][n - 1] = i & 63;
                    }
                }
            }
            #pragma unroll
            for (int axis = 0; axis < 3; ++axis) {
                if (!((qbits >> axis) & 1)) continue;
                int nodes[4];
                nodes[0] = vi;
                bool ok = true;
                #pragma unroll
                for (int n = 1; n < 4; ++n) {
                    nodes[n] = dense_index_of(nw[axis][n - 1], nb[axis][n - 1], nl[axis][n - 1]);
                    ok = ok && nodes[n] >= 0;
                }
                int3* out = reinterpret_cast<int3*>(stage + 6 * qi);  // two 12-byte LDS stores
                if (!ok) {  // cannot happen: all four voxels around a sign-changing interior edge are active
                    atomicOr(&a.counters[chunk].status, kStatusLookupMiss);
                    out[0] = out[1] = make_int3(0, 0, 0);
                } else {
                    const int edge = axis * 4 + 3;
                    const int flip = ((m >> kEdgeMap[edge][0]) & 1) != 0 ? 1 : 0;  // ProcessEdge, cl/octree.cl:335-372
                    if (flip) {
                        out[0] = make_int3(nodes[0], nodes[3], nodes[1]);
                        out[1] = make_int3(nodes[0], nodes[2], nodes[3]);
                    } else {
                        out[0] = make_int3(nodes[0], nodes[1], nodes[3]);
                        out[1] = make_int3(nodes[0], nodes[3], nodes[2]);
                    }
                }
                ++qi;
            }
            if (seam) {
                const int si = sbase[seg] + __popcll(sm & lt);
                if (si < a.caps.seams) {
                    const float4 pos = verts[3 * vi], nrm = verts[3 * vi + 1];
                    reinterpret_cast<int4*>(seams)[3 * si] = make_int4(p.x, p.y, p.z, m);
                    seams[3 * si + 1] = pos;
                    seams[3 * si + 2] = nrm;
                }
            }
        }
        // The run's LDS stores are complete before any lane of the wave reads them back.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Copy the run to quads [qb, qb + run), clipped at the quad capacity; 6 indices
        // per quad make it a whole number of 8-byte pairs at an 8-byte aligned address.
        const int keep = max(0, min(run, a.caps.quads - qb));
        int2* dst = reinterpret_cast<int2*>(tris + 6 * (long)qb);
        const int2* src = reinterpret_cast<const int2*>(stage);
        for (int i = lane; i < 3 * keep; i += 64) dst[i] = src[i];
        // ... and every lane has read the run before the next segment overwrites it.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

static inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Zero the per-chunk counters (a kernel rather than a memset node, so that a
// captured graph clears them correctly on every replay).
__global__ __launch_bounds__(256) void k_clear_counters(Counters* c, int n, int* queue) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = Counters{0, 0, 0, 0, 0, 0, 0, 0};
    if (i < kNumQueues) queue[i] = 0;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

void launch_clear_counters(const KArgs& a, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_counters, dim3(ceil_div(std::max(n, kNumQueues), 256)), dim3(256), 0, s,
                       a.counters, n, a.queue);
}

// per_cu workgroups per CU (at most two: the LDS table limit), fewer when the work is
// small.
static unsigned persistent_blocks(const KArgs& a, int per_cu, long items) {
    return (unsigned)std::max(1L, std::min((long)per_cu * a.cus, (items + 15) / 16));
}

void launch_density(const KArgs& a, int chunks, hipStream_t s) {
    const long items = (long)a.G.col_groups * a.ysplit * chunks;
    // One workgroup per CU by default (density_per_cu); the tuner tries two, which the
    // kernel's register caps let a CU hold at once.
    const unsigned blocks = persistent_blocks(a, a.density_per_cu, items);
    if (a.work)
        hipLaunchKernelGGL(k_density<true>, dim3(blocks), dim3(1024), 0, s, a, chunks);
    else
        hipLaunchKernelGGL(k_density<false>, dim3(blocks), dim3(1024), 0, s, a, chunks);
}
void launch_csg(const KArgs& a, int chunks, hipStream_t s) {
    hipLaunchKernelGGL(k_csg, dim3(kCsgSlabs, chunks), dim3(256), 0, s, a);
}
// Workgroups per chunk for a kernel that loops over `threads` work items of 256-thread
// workgroups: all of them for a few chunks, about `target` workgroups over a large batch.
static unsigned chunk_blocks(int chunks, long threads, int target) {
    const long full = std::max(1L, (threads + 255) / 256);
    const long share = std::max(1L, ((long)target + chunks - 1) / chunks);
    return (unsigned)std::min(full, share);
}

void launch_edge_rows(const KArgs& a, int chunks, hipStream_t s) {
    const long solid_waves = (a.G.field_rows + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(k_solid_rows, dim3(chunk_blocks(chunks, 64L * solid_waves, a.row_blocks), chunks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_edge_rows, dim3(chunk_blocks(chunks, a.G.edge_rows, a.row_blocks), chunks), dim3(256), 0, s, a);
}
void launch_scan(const KArgs& a, int chunks, ScanKind kind, hipStream_t s) {
    const dim3 grid(1, chunks), block(1024);
    switch (kind) {
    case kScanEdges: hipLaunchKernelGGL(k_scan<kScanEdges>, grid, block, 0, s, a); break;
    case kScanVoxels: hipLaunchKernelGGL(k_scan<kScanVoxels>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_scan_quads_seams, grid, block, 0, s, a); break;
    }
}
void launch_edge_compact(const KArgs& a, int chunks, hipStream_t s) {
    const long groups = (a.G.edge_rows + 63) / 64;
    hipLaunchKernelGGL(k_edge_compact, dim3(chunk_blocks(chunks, 64L * groups, a.row_blocks), chunks), dim3(256), 0, s, a);
}
// Lanes per edge of the crossing stage: the tuned value, else by batch size (a
// typical chunk has a few hundred 64-edge items; kG multiplies the waves).
int default_cross_lanes(int chunks) { return chunks <= 1 ? 8 : chunks <= 3 ? 4 : chunks <= 9 ? 2 : 1; }

static int cross_lanes(const KArgs& a, int chunks) {
    return a.cross_lanes > 0 ? a.cross_lanes : default_cross_lanes(chunks);
}

template <int kG>
static void launch_cross_group(const KArgs& a, int chunks, hipStream_t s) {
    const dim3 grid(persistent_blocks(a, a.persist_per_cu, 512L * kG * chunks));
    if (a.work)
        hipLaunchKernelGGL((k_edge_cross_group<true, kG>), grid, dim3(kCrossThreads), 0, s, a, chunks);
    else
        hipLaunchKernelGGL((k_edge_cross_group<false, kG>), grid, dim3(kCrossThreads), 0, s, a, chunks);
}

void launch_edge_cross(const KArgs& a, int chunks, hipStream_t s) {
    if (a.height_lines) {
        const long line_items = 2L * a.G.H * chunks;
        const dim3 line_grid(persistent_blocks(a, 2, line_items));  // 63 SGPRs: two workgroups per CU stay resident
        if (a.work)
            hipLaunchKernelGGL(k_height_lines<true>, line_grid, dim3(1024), 0, s, a, chunks);
        else
            hipLaunchKernelGGL(k_height_lines<false>, line_grid, dim3(1024), 0, s, a, chunks);
    }
    const int g = cross_lanes(a, chunks);
    hipLaunchKernelGGL(k_edge_items, dim3(1), dim3(1024), 0, s, a, chunks, g == kCrossSampleParallel ? 64 : 64 / g);
    // the item count is on the device; size the grid for 512 edge items per chunk
    switch (g) {
    case 2: launch_cross_group<2>(a, chunks, s); return;
    case 4: launch_cross_group<4>(a, chunks, s); return;
    case 8: launch_cross_group<8>(a, chunks, s); return;
    case kCrossSampleParallel: {
        const dim3 grid(persistent_blocks(a, a.persist_per_cu, 512L * chunks));
        if (a.work)
            hipLaunchKernelGGL(k_edge_cross_sp<true>, grid, dim3(kCrossThreads), 0, s, a, chunks);
        else
            hipLaunchKernelGGL(k_edge_cross_sp<false>, grid, dim3(kCrossThreads), 0, s, a, chunks);
        return;
    }
    default: break;
    }
    const dim3 grid(persistent_blocks(a, a.persist_per_cu, 512L * chunks));
    if (a.work)
        hipLaunchKernelGGL(k_edge_cross<true>, grid, dim3(kCrossThreads), 0, s, a, chunks);
    else
        hipLaunchKernelGGL(k_edge_cross<false>, grid, dim3(kCrossThreads), 0, s, a, chunks);
}
void launch_voxel_rows(const KArgs& a, int chunks, hipStream_t s) {
    hipLaunchKernelGGL(k_voxel_rows, dim3(chunk_blocks(chunks, a.G.voxel_segs, a.row_blocks), chunks), dim3(256), 0, s, a);
}
void launch_voxel_compact(const KArgs& a, int chunks, hipStream_t s) {
    hipLaunchKernelGGL(k_voxel_compact, dim3(chunk_blocks(chunks, a.G.voxel_segs, a.row_blocks), chunks), dim3(256), 0, s, a);
}
// The kernels below loop up to the chunk's voxel count, known only on the device.
void launch_qef(const KArgs& a, int chunks, hipStream_t s) {
    const dim3 grid(chunk_blocks(chunks, a.caps.voxels, a.voxel_blocks), chunks);
    if (a.precise)
        hipLaunchKernelGGL(k_qef<true>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_qef<false>, grid, dim3(256), 0, s, a);
}
void launch_emit(const KArgs& a, int chunks, hipStream_t s) {
    hipLaunchKernelGGL(k_emit, dim3(chunk_blocks(chunks, a.caps.voxels, a.voxel_blocks), chunks), dim3(64 * kEmitWaves), 0,
                       s, a);
}

const char* const kStageNames[kNumStages] = {
    "density", "csg", "edge_rows", "scan_edges", "edge_compact", "edge_cross",
    "voxel_rows", "scan_voxels", "voxel_compact", "qef", "scan_quads_seams", "emit"};

void launch_stage(const KArgs& a, int chunks, int i, hipStream_t s) {
    switch (i) {
    case 0: launch_density(a, chunks, s); break;
    case 1: launch_csg(a, chunks, s); break;
    case 2: launch_edge_rows(a, chunks, s); break;
    case 3: launch_scan(a, chunks, kScanEdges, s); break;
    case 4: launch_edge_compact(a, chunks, s); break;
    case 5: launch_edge_cross(a, chunks, s); break;
    case 6: launch_voxel_rows(a, chunks, s); break;
    case 7: launch_scan(a, chunks, kScanVoxels, s); break;
    case 8: launch_voxel_compact(a, chunks, s); break;
    case 9: launch_qef(a, chunks, s); break;
    case 10: launch_scan(a, chunks, kScanQuadsSeams, s); break;
    default: launch_emit(a, chunks, s); break;
    }
}

void launch_pipeline(const KArgs& a, int chunks, hipStream_t s) {
    launch_clear_counters(a, chunks, s);
    for (int i = 0; i < kNumStages; ++i) launch_stage(a, chunks, i, s);
}

}  // namespace lvn
