// lvn_simplify.hip — post-mesh simplification on the GPU (ngMeshSimplifier,
// src/ng_mesh_simplify.cpp:438-539) over many meshes at once.
//
// The reference's iterations are parallel in shape: a random fraction of the candidate
// edges is evaluated, every vertex keeps its cheapest valid candidate (the first in edge
// order on ties), and an edge collapses only when both its vertices chose it, so the
// collapses of one iteration are independent.  Here every step is one launch over all
// meshes of a group, cut into tiles of 1024 elements that never straddle two meshes
// (vertex, triangle, edge and draw tiles, listed by the host): the candidate edges sorted
// by (max, min) index through per-vertex buckets, the per-vertex choice as a 64-bit atomic
// minimum of (cost bits, edge index) (costs are never negative, so their bits order like
// their values), and order-preserving compaction of triangles and edges by segmented scans
// (tile sums, a per-mesh scan of the tile sums, then each tile's own scan).  Each mesh
// keeps its own counts, iteration number and stop state; a mesh that has stopped passes
// through the compactions unchanged.  The arithmetic of each candidate (the two-plane QEF,
// its cost) is the one include/lvn_mesher.h states (lvn_simplify_options), so results are
// identical to the sequential restatement.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstddef>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "lvn_kernels.h"
#include "lvn_runtime.h"

namespace lvn {

constexpr int kCollapseMaxDegree = 16;  // COLLAPSE_MAX_DEGREE, src/ng_mesh_simplify.cpp:17
constexpr int kSimpTile = 1024;         // elements per tile
constexpr int kSimpThreads = 256;       // threads per tile workgroup, 4 consecutive elements each
constexpr int kSimpPer = kSimpTile / kSimpThreads;

// One mesh to simplify: read from (vin, tin), written to (vout, tout), which may be the
// same arrays.
struct SimpJob {
    const lvn_vertex* vin;
    const int32_t* tin;
    lvn_vertex* vout;
    int32_t* tout;
    int32_t nv, nt;
    int32_t result;  // the caller's index of this mesh
    float max_error, max_edge_size;  // this mesh's limits (leven scales them by the chunk's voxel size)
    float offset[4];
};

// A mesh of a group: its ranges in the group's arrays, its first tile of each kind, and
// its state.
struct SMesh {
    int vbase, nv, tbase, nt0, ebase, ecap, dbase, dcap;
    int vt0, vtn, tt0, ttn, et0, etn, dt0, dtn;
    int target;  // (int)(nt0 * target_percentage)
    float max_error, max_edge;
    float off[4];
    const lvn_vertex* vin;
    const int32_t* tin;
    lvn_vertex* vout;
    int32_t* tout;
    // state
    int T;        // live triangles (the list keeps its slots; a dropped triangle is marked dead)
    int E, En;    // candidate edges, and their count after this iteration's compaction
    int ecur;     // which edge array holds the mesh's edges (0: edge_a, 1: edge_b)
    int iter, done, active, draws, dlim, accepted;
    int valid, coll, coll_it, drop_it;  // valid draws, collapses (all, this iteration), triangles dropped
    int tpass, ecompact;                // this iteration remaps triangles / compacts edges
    int deg_in;                         // degenerate input triangles (the first RemoveTriangles drops them)
    int last, fail, nvo, Tf, one, dummy;
};

// The group's device arrays (V vertices, T triangles, E edge slots, D draw slots in all).
struct SG {
    SMesh* mesh;
    int n_mesh;
    const int2* vt;  // tiles: (mesh, first element)
    const int2* tt;
    const int2* et;
    const int2* dt;
    int nvt, ntt, net, ndt;
    float4* wpos;  // [V] working vertices, positions relative to the mesh's offset
    float4* wnrm;
    float4* wcol;
    int* vtc;    // [V] live triangles per vertex
    unsigned long long* key;  // [V] (cost bits << 32) | edge index of the vertex's choice
    int* tgt;                 // [V] collapse target, -1 none
    int* bcnt;                // [V] bucket sizes; later the used-vertex flags
    int* boff;                // [V] bucket offsets
    int* bnd;                 // [V] boundary marks
    int* ucnt;                // [V] bucket cursors; later the new vertex indices
    int* tri;                 // [3 T] the triangles, remapped in place; a dead one has index -1 first
    int* bucket;              // [3 T] the edge buckets while building
    int2* edge_a;             // [E] candidate edges (min, max), two arrays per mesh's ecur
    int2* edge_b;
    int* pick;    // [D] drawn edge index; after validity the edge index or -1
    float4* dpos; // [D] collapse position / normal of a valid draw
    float4* dnrm;
    int* tsum;    // [tiles] per-tile counts / offsets
    int* toff;
    const uint32_t* u;  // std::mt19937 outputs for opt.seed
    int n_u;
    lvn_simplify_options opt;
};

namespace {

inline uint64_t align256(uint64_t v) { return (v + 255) & ~(uint64_t)255; }

// glm::dot of two vec4: (x x' + y y') + (z z' + w w').
__device__ __forceinline__ float dot_glm(float4 a, float4 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + a.w * b.w); }

// The SSE solver's values (src/qef_simd.h), one F4 per __m128, lane for lane.
struct F4 {
    float v[4];
};
struct M4 {
    F4 row[4];
};

__device__ __forceinline__ F4 f4(float x, float y, float z, float w) { return F4{{x, y, z, w}}; }
__device__ __forceinline__ F4 f4(float4 a) { return F4{{a.x, a.y, a.z, a.w}}; }

// vec4_dot, src/qef_simd.h:101-110: (x x' + y y') + (w w' + z z').
__device__ __forceinline__ float vec4_dot(const F4& a, const F4& b) {
    const float m0 = a.v[0] * b.v[0], m1 = a.v[1] * b.v[1], m2 = a.v[2] * b.v[2], m3 = a.v[3] * b.v[3];
    return (m0 + m1) + (m3 + m2);
}

// vec4_mul_m4x4, src/qef_simd.h:114-121.
__device__ __forceinline__ F4 vec4_mul_m4x4(const F4& a, const M4& B) {
    F4 r;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        r.v[c] = ((a.v[0] * B.row[0].v[c] + a.v[1] * B.row[1].v[c]) + a.v[2] * B.row[2].v[c]) + a.v[3] * B.row[3].v[c];
    return r;
}

// svd_rotate, src/qef_simd.h:278-319.
__device__ __forceinline__ void svd_rotate(M4& t, M4& V, float c, float s, int a, int b) {
    if (t.row[a].v[b] == 0.f) return;
    const F4 u = f4(V.row[0].v[a], V.row[1].v[a], V.row[2].v[a], t.row[0].v[3 - b]);
    const F4 w = f4(V.row[0].v[b], V.row[1].v[b], V.row[2].v[b], t.row[1 - a].v[2]);
    F4 x, y;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        x.v[l] = c * u.v[l] - s * w.v[l];
        y.v[l] = s * u.v[l] + c * w.v[l];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) V.row[r].v[a] = x.v[r];
    t.row[0].v[3 - b] = x.v[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) V.row[r].v[b] = y.v[r];
    t.row[1 - a].v[2] = y.v[3];
    t.row[a].v[b] = 0.f;
}

// qef_solve_from_points_4d_interleaved over the two endpoints (src/qef_simd.h:403-533,
// count 2): the solved position in *out, returns the error |ATb - x ATA|^2 (x relative to
// the mass point).  The Givens cosine is the IEEE 1 / sqrt where the reference takes
// _mm_rsqrt_ps; rotateq_xy (:228-274) and the pseudo-inverse's product with V (:371-382)
// are computed as written.
__device__ float qef_two(float4 p0, float4 n0, float4 p1, float4 n1, float4* out) {
    M4 A;
#pragma unroll
    for (int r = 0; r < 4; ++r) A.row[r] = f4(0.f, 0.f, 0.f, 0.f);
    F4 atb = f4(0.f, 0.f, 0.f, 0.f), acc = f4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 2; ++k) {  // qef_simd_add, :403-422
        const F4 p = f4(k == 0 ? p0 : p1), n = f4(k == 0 ? n0 : n1);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int l = 0; l < 4; ++l) A.row[r].v[l] = A.row[r].v[l] + n.v[r] * n.v[l];
        const float d = vec4_dot(p, n);
        const F4 dd = f4(d, d, d, 0.f);
#pragma unroll
        for (int l = 0; l < 4; ++l) atb.v[l] = atb.v[l] + dd.v[l] * n.v[l];
#pragma unroll
        for (int l = 0; l < 4; ++l) acc.v[l] = acc.v[l] + p.v[l];
    }
    F4 mp;  // qef_simd_solve, :436-455
#pragma unroll
    for (int l = 0; l < 4; ++l) mp.v[l] = acc.v[l] / acc.v[3];
    const F4 pm = vec4_mul_m4x4(mp, A);
    F4 rhs;
#pragma unroll
    for (int l = 0; l < 4; ++l) rhs.v[l] = atb.v[l] - pm.v[l];
    // svd_solve_ATA_ATb, :323-398
    M4 V, t = A;
    V.row[0] = f4(1.f, 0.f, 0.f, 0.f);
    V.row[1] = f4(0.f, 1.f, 0.f, 0.f);
    V.row[2] = f4(0.f, 0.f, 1.f, 0.f);
    V.row[3] = f4(0.f, 0.f, 0.f, 0.f);
    for (int sweep = 0; sweep < 5; ++sweep) {  // SVD_NUM_SWEEPS
        F4 c, s;
        {  // givens_coeffs_sym, :163-224
            const F4 pp = f4(t.row[0].v[0], t.row[0].v[0], t.row[1].v[1], 0.f);
            const F4 pq = f4(t.row[0].v[1], t.row[0].v[2], t.row[1].v[2], 0.f);
            const F4 qq = f4(t.row[1].v[1], t.row[2].v[2], t.row[2].v[2], 0.f);
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float tau = (qq.v[l] - pp.v[l]) / (pq.v[l] * 2.f);
                const float stt = sqrtf(tau * tau + 1.f);
                const float tn = 1.f / (tau >= 0.f ? tau + stt : tau - stt);
                const float cl = 1.f / sqrtf(1.f + tn * tn);
                const float sl = tn * cl;
                const bool zero = pq.v[l] == 0.f;
                c.v[l] = zero ? 1.f : cl;
                s.v[l] = zero ? 0.f : sl;
            }
        }
        {  // rotateq_xy, :228-274
            const F4 u = f4(t.row[0].v[0], t.row[0].v[0], t.row[1].v[1], 0.f);
            const F4 v = f4(t.row[1].v[1], t.row[2].v[2], t.row[2].v[2], 0.f);
            const F4 a = f4(t.row[0].v[1], t.row[0].v[2], t.row[1].v[2], 0.f);
            F4 x, y;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const float cc = c.v[l] * c.v[l], ss = s.v[l] * s.v[l];
                const float mx = ((2.f * c.v[l]) * s.v[l]) * a.v[l];
                const float x0 = cc * u.v[l];
                x.v[l] = (x0 - mx) + ss * v.v[l];
                y.v[l] = (x0 + mx) + (cc + v.v[l]);
            }
            t.row[0].v[0] = x.v[1];
            t.row[1].v[1] = x.v[2];
            t.row[0].v[1] = y.v[0];
            t.row[0].v[2] = y.v[1];
            t.row[1].v[2] = y.v[2];
        }
        svd_rotate(t, V, c.v[0], s.v[0], 0, 1);
        svd_rotate(t, V, c.v[1], s.v[1], 0, 2);
        svd_rotate(t, V, c.v[2], s.v[2], 1, 2);
    }
    const F4 sigma = f4(t.row[0].v[0], t.row[1].v[1], t.row[2].v[2], 0.f);
    F4 inv;
#pragma unroll
    for (int l = 0; l < 4; ++l) {  // svd_invdet, :353-364
        const float ax = fabsf(sigma.v[l]);
        const float r = 1.f / sigma.v[l];
        const float ar = fabsf(r);
        const float mn = ax < ar ? ax : ar;
        inv.v[l] = mn >= 0.1f ? r : 0.f;
    }
    M4 o, pinv;  // svd_pseudoinverse, :368-382
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int l = 0; l < 4; ++l) o.row[r].v[l] = V.row[r].v[l] * inv.v[l];
    o.row[3] = f4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int r = 0; r < 4; ++r) pinv.row[r] = vec4_mul_m4x4(o.row[r], V);
    const F4 x = vec4_mul_m4x4(rhs, pinv);
    const F4 xa = vec4_mul_m4x4(x, A);  // qef_simd_calc_error, :426-432
    F4 e;
#pragma unroll
    for (int l = 0; l < 4; ++l) e.v[l] = atb.v[l] - xa.v[l];
    *out = make_float4(x.v[0] + mp.v[0], x.v[1] + mp.v[1], x.v[2] + mp.v[2], x.v[3] + mp.v[3]);
    return vec4_dot(e, e);
}


// ---- scans ---------------------------------------------------------------------------

// Exclusive scan of v over the workgroup's threads; *total = the sum over them.
__device__ __forceinline__ int block_scan(int v, int* total) {
    __shared__ int wsum[kSimpThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kSimpThreads / 64; ++i) {
        const int s = wsum[i];
        base += i < w ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// An order-preserving pass over one kind of tile: Op::limit(mesh) elements take part,
// Op::count(i, item) is element i's output count (item keeps what the writes need),
// Op::emit(i, pos, item) writes them from pos on.
// Three launches: per-tile sums, a scan of each mesh's tile sums (its total goes to
// Field; meshes whose Gate is 0 are left alone), then each tile's own scan and the writes.
template <class Op>
__global__ __launch_bounds__(kSimpThreads) void k_tile_count(SG g, const int2* tiles, int ntiles) {
    for (int k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int2 tl = tiles[k];
        const SMesh& M = g.mesh[tl.x];
        const int lim = Op::limit(M);
        const int i0 = tl.y + threadIdx.x * kSimpPer;
        int s = 0;
        if (tl.y < lim) {
#pragma unroll
            for (int r = 0; r < kSimpPer; ++r) {
                typename Op::Item item;
                if (i0 + r < lim) s += Op::count(g, M, i0 + r, item);
            }
        }
        int tot;
        block_scan(s, &tot);
        if (threadIdx.x == 0) g.tsum[k] = tot;
    }
}

template <int SMesh::*First, int SMesh::*Count, int SMesh::*Field, int SMesh::*Gate>
__global__ __launch_bounds__(kSimpThreads) void k_mesh_scan(SG g) {
    const int m = blockIdx.x * (kSimpThreads / 64) + (threadIdx.x >> 6);
    if (m >= g.n_mesh) return;  // whole waves
    SMesh& M = g.mesh[m];
    if (!(M.*Gate)) return;
    const int t0 = M.*First, tn = M.*Count, lane = threadIdx.x & 63;
    int carry = 0;
    for (int c = 0; c < tn; c += 64) {
        const int v = c + lane < tn ? g.tsum[t0 + c + lane] : 0;
        int x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (c + lane < tn) g.toff[t0 + c + lane] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
    if (lane == 0) M.*Field = carry;
}

template <class Op>
__global__ __launch_bounds__(kSimpThreads) void k_tile_emit(SG g, const int2* tiles, int ntiles) {
    for (int k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const int2 tl = tiles[k];
        const SMesh& M = g.mesh[tl.x];
        const int lim = Op::limit(M);
        if (tl.y >= lim) continue;
        const int i0 = tl.y + threadIdx.x * kSimpPer;
        int c[kSimpPer], s = 0;
        typename Op::Item item[kSimpPer];
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r) {
            c[r] = i0 + r < lim ? Op::count(g, M, i0 + r, item[r]) : 0;
            s += c[r];
        }
        int tot;
        int pos = block_scan(s, &tot);
        if (tot == 0) continue;
        pos += g.toff[k];
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r)
            if (c[r]) {
                Op::emit(g, M, i0 + r, pos, item[r]);
                pos += c[r];
            }
    }
}

// ---- the passes' element rules -------------------------------------------------------

// Bucket offsets: vertex i's bucket holds the smaller ends of the triangle sides whose
// larger end is i.
struct BucketOp {
    using Item = int;
    static __device__ int limit(const SMesh& M) { return M.nv; }
    static __device__ int count(const SG& g, const SMesh& M, int i, Item&) { return g.bcnt[M.vbase + i]; }
    static __device__ void emit(const SG& g, const SMesh& M, int i, int pos, const Item&) {
        g.boff[M.vbase + i] = pos;
        g.ucnt[M.vbase + i] = pos;
    }
};

// BuildCandidateEdges (src/ng_mesh_simplify.cpp:118-172): in (max, min) key order, each
// run of equal sides seen more than once whose ends are not boundary, except the last run
// of all, which the reference's loop never examines.
struct EdgeOp {
    template <class F>
    static __device__ int walk(const SG& g, const SMesh& M, int i, F f) {
        const int n = g.bcnt[M.vbase + i];
        if (n == 0) return 0;
        const int* b = g.bucket + 3 * (int64_t)M.tbase + g.boff[M.vbase + i];
        const bool bi = g.bnd[M.vbase + i] != 0;
        int kept = 0, k = 0;
        while (k < n) {
            int e = k + 1;
            while (e < n && b[e] == b[k]) ++e;
            if (!(i == M.last && e == n) && e - k > 1 && !bi && !g.bnd[M.vbase + b[k]]) {
                f(kept, b[k]);
                ++kept;
            }
            k = e;
        }
        return kept;
    }
    using Item = int;
    static __device__ int limit(const SMesh& M) { return M.nv; }
    // the count pass keeps each vertex's count (the bucket cursors are free by now) for
    // the write pass, which then walks each bucket once
    static __device__ int count(const SG& g, const SMesh& M, int i, Item&) {
        const int kept = walk(g, M, i, [](int, int) {});
        g.ucnt[M.vbase + i] = kept;
        return kept;
    }
    static __device__ void emit(const SG& g, const SMesh& M, int i, int pos, const Item&) {
        int2* out = g.edge_a + M.ebase + pos;
        walk(g, M, i, [&](int q, int mn) { out[q] = make_int2(mn, i); });
    }
};

// The write pass of the candidate edges: counts kept by EdgeOp's count pass.
struct EdgeEmitOp {
    using Item = int;
    static __device__ int limit(const SMesh& M) { return M.nv; }
    static __device__ int count(const SG& g, const SMesh& M, int i, Item&) { return g.ucnt[M.vbase + i]; }
    static __device__ void emit(const SG& g, const SMesh& M, int i, int pos, const Item& it) {
        EdgeOp::emit(g, M, i, pos, it);
    }
};

__device__ __forceinline__ int remap(const SG& g, const SMesh& M, int v) {
    const int t = g.tgt[M.vbase + v];
    return t != -1 ? t : v;
}

__device__ __forceinline__ const int2* edges_of(const SG& g, const SMesh& M) {
    return (M.ecur ? g.edge_b : g.edge_a) + M.ebase;
}

// The draws of FindValidCollapses (:193-207): draw j is the j-th value of
// std::uniform_int_distribution<int>(0, E - 1) over a std::mt19937 seeded anew, as
// libstdc++ maps it (Lemire: the high word of u * E, drawing again while the low word is
// below 2^32 mod E).  Over the engine's outputs u[k] (a table): the accepted ones, in
// order, are the draws.
struct DrawOp {
    using Item = int;
    static __device__ int limit(const SMesh& M) { return M.dlim; }
    static __device__ int count(const SG& g, const SMesh& M, int k, Item& hi) {
        const uint32_t E = (uint32_t)M.E;
        const uint64_t prod = (uint64_t)g.u[k] * E;
        hi = (int)(prod >> 32);
        return (uint32_t)prod >= (0u - E) % E ? 1 : 0;
    }
    static __device__ void emit(const SG& g, const SMesh& M, int, int pos, const Item& hi) {
        if (pos < M.draws) g.pick[M.dbase + pos] = hi;
    }
};

// RemoveEdges (src/ng_mesh_simplify.cpp:353-381): remapped, collapsed ones dropped, in
// order, into the mesh's other edge array.  Only meshes with a collapse this iteration:
// for the others the step keeps every edge as it is.
struct EdgeKeepOp {
    using Item = int2;
    static __device__ int limit(const SMesh& M) { return M.ecompact ? M.E : 0; }
    static __device__ int count(const SG& g, const SMesh& M, int i, Item& r) {
        const int2 e = edges_of(g, M)[i];
        r = make_int2(remap(g, M, e.x), remap(g, M, e.y));
        return r.x != r.y ? 1 : 0;
    }
    static __device__ void emit(const SG& g, const SMesh& M, int, int pos, const Item& r) {
        ((M.ecur ? g.edge_a : g.edge_b) + M.ebase)[pos] = r;
    }
};

// CompactVertices (src/ng_mesh_simplify.cpp:385-430) and the offset restored
// (:531-537): the used vertices in order, written to the output.
struct UsedOp {
    using Item = int;
    static __device__ int limit(const SMesh& M) { return M.nv; }
    static __device__ int count(const SG& g, const SMesh& M, int i, Item&) { return g.bcnt[M.vbase + i]; }
    static __device__ void emit(const SG& g, const SMesh& M, int i, int pos, const Item&) {
        const int v = M.vbase + i;
        g.ucnt[v] = pos;
        const float4 p = g.wpos[v], n = g.wnrm[v], c = g.wcol[v];
        lvn_vertex o;
        o.pos[0] = p.x + M.off[0];
        o.pos[1] = p.y + M.off[1];
        o.pos[2] = p.z + M.off[2];
        o.pos[3] = p.w + M.off[3];
        o.normal[0] = n.x;
        o.normal[1] = n.y;
        o.normal[2] = n.z;
        o.normal[3] = n.w;
        o.colour[0] = c.x;
        o.colour[1] = c.y;
        o.colour[2] = c.z;
        o.colour[3] = c.w;
        M.vout[pos] = o;
    }
};

// The live triangles in order, with the compacted vertex indices (:516-530).
struct TriFinalOp {
    using Item = int3;
    static __device__ int limit(const SMesh& M) { return M.nt0; }
    static __device__ int count(const SG& g, const SMesh& M, int t, Item& q) {
        const int* p = g.tri + 3 * (int64_t)(M.tbase + t);
        q = make_int3(p[0], p[1], p[2]);
        return q.x >= 0 ? 1 : 0;
    }
    static __device__ void emit(const SG& g, const SMesh& M, int, int pos, const Item& q) {
        M.tout[3 * pos] = g.ucnt[M.vbase + q.x];
        M.tout[3 * pos + 1] = g.ucnt[M.vbase + q.y];
        M.tout[3 * pos + 2] = g.ucnt[M.vbase + q.z];
    }
};

// ---- set-up --------------------------------------------------------------------------

#define LVN_TILES(tiles, ntiles)                                             \
    for (int k_ = blockIdx.x; k_ < (ntiles); k_ += gridDim.x)                \
        for (int tl_m = (tiles)[k_].x, tl_s = (tiles)[k_].y, once_ = 1; once_; once_ = 0)

// Working copy of the vertices, positions relative to the offset (:459-462).
__global__ __launch_bounds__(kSimpThreads) void k_init_vertices(SG g) {
    LVN_TILES(g.vt, g.nvt) {
        const SMesh& M = g.mesh[tl_m];
        for (int r = 0; r < kSimpPer; ++r) {
            const int i = tl_s + threadIdx.x + r * kSimpThreads;
            if (i >= M.nv) continue;
            const lvn_vertex v = M.vin[i];
            const int q = M.vbase + i;
            g.wpos[q] = make_float4(v.pos[0] - M.off[0], v.pos[1] - M.off[1], v.pos[2] - M.off[2], v.pos[3] - M.off[3]);
            g.wnrm[q] = make_float4(v.normal[0], v.normal[1], v.normal[2], v.normal[3]);
            g.wcol[q] = make_float4(v.colour[0], v.colour[1], v.colour[2], v.colour[3]);
            g.vtc[q] = 0;
            g.bcnt[q] = 0;
            g.bnd[q] = 0;
        }
    }
}

// Vertex-index window of a tile's triangles held in LDS: a tile's 1024 triangles use a
// few thousand consecutive vertex indices in a chunk mesh, so counts gather there first
// and each vertex takes one device atomic per tile instead of one per use.
constexpr int kSimpWindow = 8192;

__device__ __forceinline__ void tile_window(const SG& g, const SMesh& M, int tl_s, int tri[kSimpPer][3], int* lo,
                                            int* hi) {
    __shared__ int s_lo, s_hi;
    if (threadIdx.x == 0) {
        s_lo = INT_MAX;
        s_hi = -1;
    }
    __syncthreads();
    int l = INT_MAX, h = -1;
#pragma unroll
    for (int r = 0; r < kSimpPer; ++r) {
        const int t = tl_s + threadIdx.x + r * kSimpThreads;
        if (t < M.nt0) {
            const int* p = g.tri + 3 * (int64_t)(M.tbase + t);
            tri[r][0] = p[0];
            tri[r][1] = p[1];
            tri[r][2] = p[2];
            l = min(l, min(tri[r][0], min(tri[r][1], tri[r][2])));
            h = max(h, max(tri[r][0], max(tri[r][1], tri[r][2])));
        } else {
            tri[r][0] = -1;
        }
    }
    if (h >= 0) {
        atomicMin(&s_lo, l);
        atomicMax(&s_hi, h);
    }
    __syncthreads();
    *lo = s_lo;
    *hi = s_hi;
    __syncthreads();
}

// The triangles, their per-vertex counts (:470-482) and the bucket sizes.
__global__ __launch_bounds__(kSimpThreads) void k_init_triangles(SG g) {
    __shared__ unsigned int hist[kSimpWindow];  // triangles << 16 | sides whose larger end it is
    LVN_TILES(g.tt, g.ntt) {
        SMesh& M = g.mesh[tl_m];
        int top = -1, deg = 0;
        int tri[kSimpPer][3];
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r) {
            const int t = tl_s + threadIdx.x + r * kSimpThreads;
            if (t >= M.nt0) continue;
            int* q = g.tri + 3 * (int64_t)(M.tbase + t);
            q[0] = M.tin[3 * t];
            q[1] = M.tin[3 * t + 1];
            q[2] = M.tin[3 * t + 2];
        }
        int lo, hi;
        tile_window(g, M, tl_s, tri, &lo, &hi);
        const bool local = hi >= 0 && hi - lo < kSimpWindow;
        if (local)
            for (int e = threadIdx.x; e <= hi - lo; e += kSimpThreads) hist[e] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r) {
            if (tri[r][0] < 0) continue;
            const int a = tri[r][0], b = tri[r][1], c = tri[r][2];
            if (local) {
                atomicAdd(&hist[a - lo], 0x10000u);
                atomicAdd(&hist[b - lo], 0x10000u);
                atomicAdd(&hist[c - lo], 0x10000u);
                atomicAdd(&hist[max(a, b) - lo], 1u);
                atomicAdd(&hist[max(b, c) - lo], 1u);
                atomicAdd(&hist[max(a, c) - lo], 1u);
            } else {
                atomicAdd(&g.vtc[M.vbase + a], 1);
                atomicAdd(&g.vtc[M.vbase + b], 1);
                atomicAdd(&g.vtc[M.vbase + c], 1);
                atomicAdd(&g.bcnt[M.vbase + max(a, b)], 1);
                atomicAdd(&g.bcnt[M.vbase + max(b, c)], 1);
                atomicAdd(&g.bcnt[M.vbase + max(a, c)], 1);
            }
            top = max(top, max(a, max(b, c)));
            deg += (a == b || a == c || b == c) ? 1 : 0;
        }
        __syncthreads();
        if (local)
            for (int e = threadIdx.x; e <= hi - lo; e += kSimpThreads) {
                const unsigned int v = hist[e];
                if (v >> 16) atomicAdd(&g.vtc[M.vbase + lo + e], (int)(v >> 16));
                if (v & 0xffffu) atomicAdd(&g.bcnt[M.vbase + lo + e], (int)(v & 0xffffu));
            }
        if (top >= 0) atomicMax(&M.last, top);
        int tot;
        block_scan(deg, &tot);
        if (threadIdx.x == 0 && tot) atomicAdd(&M.deg_in, tot);
    }
}

// Each side's smaller end into the bucket of its larger end: per tile, the sides of each
// vertex are counted and ranked in LDS, then one device atomic reserves the vertex's slots.
__global__ __launch_bounds__(kSimpThreads) void k_bucket_scatter(SG g) {
    __shared__ int hist[kSimpWindow];
    LVN_TILES(g.tt, g.ntt) {
        const SMesh& M = g.mesh[tl_m];
        int* bucket = g.bucket + 3 * (int64_t)M.tbase;
        int tri[kSimpPer][3];
        int lo, hi;
        tile_window(g, M, tl_s, tri, &lo, &hi);
        const bool local = hi >= 0 && hi - lo < kSimpWindow;
        if (!local) {
#pragma unroll
            for (int r = 0; r < kSimpPer; ++r) {
                if (tri[r][0] < 0) continue;
                const int a = tri[r][0], b = tri[r][1], c = tri[r][2];
                bucket[atomicAdd(&g.ucnt[M.vbase + max(a, b)], 1)] = min(a, b);
                bucket[atomicAdd(&g.ucnt[M.vbase + max(b, c)], 1)] = min(b, c);
                bucket[atomicAdd(&g.ucnt[M.vbase + max(a, c)], 1)] = min(a, c);
            }
            continue;
        }
        for (int e = threadIdx.x; e <= hi - lo; e += kSimpThreads) hist[e] = 0;
        __syncthreads();
        int rank[kSimpPer][3];
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r) {
            if (tri[r][0] < 0) continue;
            const int a = tri[r][0], b = tri[r][1], c = tri[r][2];
            rank[r][0] = atomicAdd(&hist[max(a, b) - lo], 1);
            rank[r][1] = atomicAdd(&hist[max(b, c) - lo], 1);
            rank[r][2] = atomicAdd(&hist[max(a, c) - lo], 1);
        }
        __syncthreads();
        for (int e = threadIdx.x; e <= hi - lo; e += kSimpThreads) {
            const int n = hist[e];
            if (n) hist[e] = atomicAdd(&g.ucnt[M.vbase + lo + e], n);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSimpPer; ++r) {
            if (tri[r][0] < 0) continue;
            const int a = tri[r][0], b = tri[r][1], c = tri[r][2];
            bucket[hist[max(a, b) - lo] + rank[r][0]] = min(a, b);
            bucket[hist[max(b, c) - lo] + rank[r][1]] = min(b, c);
            bucket[hist[max(a, c) - lo] + rank[r][2]] = min(a, c);
        }
        __syncthreads();
    }
}

// Each bucket sorted; a side seen once marks both ends as boundary (:140-166).
__global__ __launch_bounds__(kSimpThreads) void k_bucket_sort(SG g) {
    LVN_TILES(g.vt, g.nvt) {
        const SMesh& M = g.mesh[tl_m];
        for (int r = 0; r < kSimpPer; ++r) {
            const int i = tl_s + threadIdx.x + r * kSimpThreads;
            if (i >= M.nv) continue;
            const int n = g.bcnt[M.vbase + i];
            if (n == 0) continue;
            int* b = g.bucket + 3 * (int64_t)M.tbase + g.boff[M.vbase + i];
            for (int k = 1; k < n; ++k) {
                const int x = b[k];
                int j = k - 1;
                while (j >= 0 && b[j] > x) {
                    b[j + 1] = b[j];
                    --j;
                }
                b[j + 1] = x;
            }
            int k = 0;
            while (k < n) {
                int e = k + 1;
                while (e < n && b[e] == b[k]) ++e;
                if (!(i == M.last && e == n) && e - k == 1) {
                    g.bnd[M.vbase + b[k]] = 1;
                    g.bnd[M.vbase + i] = 1;
                }
                k = e;
            }
        }
    }
}

// ---- one iteration (:490-512) --------------------------------------------------------

// Draw positions scanned for `draws` draws: enough for the redraws (about draws * E / 2^32).
__device__ __forceinline__ int draw_limit(int draws, int dcap) { return min(draws + draws / 16 + 64, dcap); }

// Per mesh: take the last iteration's edge compaction; stop a mesh whose last iteration
// found no valid collapse (the loop's break), or changed nothing (then every later
// iteration would repeat it until the limit), or that met its triangle target or its
// iteration limit; count the meshes still going into *active.
__global__ void k_iter_begin(SG g, int* active) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.n_mesh) return;
    SMesh& M = g.mesh[m];
    if (M.ecompact) {
        M.E = M.En;
        M.ecur ^= 1;
    }
    M.ecompact = 0;
    M.tpass = 0;
    if (M.active) {
        if (M.valid == 0) {
            M.done = 1;
        } else if (M.coll_it == 0 && M.drop_it == 0) {
            M.iter = g.opt.max_iterations;
            M.done = 1;
        }
    }
    const bool go = !M.done && !M.fail && M.T > M.target && M.iter < g.opt.max_iterations;
    if (!go) M.done = 1;
    M.active = go ? 1 : 0;
    if (go) M.iter += 1;
    M.draws = go ? (int)((float)M.E * g.opt.edge_fraction) : 0;
    M.dlim = M.draws > 0 ? draw_limit(M.draws, M.dcap) : 0;
    M.accepted = 0;
    M.valid = 0;
    M.coll_it = 0;
    M.drop_it = 0;
    if (go) atomicAdd(active, 1);
}

__global__ __launch_bounds__(kSimpThreads) void k_iter_reset(SG g) {
    LVN_TILES(g.vt, g.nvt) {
        const SMesh& M = g.mesh[tl_m];
        if (!M.active) continue;
        for (int r = 0; r < kSimpPer; ++r) {
            const int i = tl_s + threadIdx.x + r * kSimpThreads;
            if (i >= M.nv) continue;
            g.key[M.vbase + i] = ~0ull;
            g.tgt[M.vbase + i] = -1;
        }
    }
}

// The candidate's checks and cost (:213-275); each end keeps its cheapest valid candidate,
// the first in edge order on ties.
__global__ __launch_bounds__(kSimpThreads) void k_valid(SG g) {
    LVN_TILES(g.dt, g.ndt) {
        SMesh& M = g.mesh[tl_m];
        const int n = min(M.draws, M.accepted);
        if (tl_s >= M.draws) continue;
        if (M.accepted < M.draws && tl_s == 0 && threadIdx.x == 0) M.fail = 1;  // too few draw positions
        const int2* edges = edges_of(g, M);
        int valid = 0;
        for (int r = 0; r < kSimpPer; ++r) {
            const int j = tl_s + threadIdx.x + r * kSimpThreads;
            if (j >= n) continue;
            const int i = g.pick[M.dbase + j];
            const int2 e = edges[i];
            const int va = M.vbase + e.x, vb = M.vbase + e.y;
            int ok = 0;
            const float4 pa = g.wpos[va], pb = g.wpos[vb], na = g.wnrm[va], nb = g.wnrm[vb];
            const float cos_angle = dot_glm(na, nb);
            const float4 d = make_float4(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z, pb.w - pa.w);
            const int degree = g.vtc[va] + g.vtc[vb];
            if (!(cos_angle < g.opt.min_angle_cosine) && !(dot_glm(d, d) > M.max_edge * M.max_edge) &&
                !((double)fabsf(g.wcol[va].w - g.wcol[vb].w) > 1e-3) && !(degree > kCollapseMaxDegree)) {
                float4 q;
                float error = qef_two(pa, na, pb, nb, &q);
                if (error > 0.f) error = 1.f / error;
                const int penalty = max(0, degree - 10);
                error = error + (float)penalty * (M.max_error * 0.1f);
                if (!(error > M.max_error)) {
                    ok = 1;
                    g.dnrm[M.dbase + j] = make_float4((na.x - nb.x) * 0.5f, (na.y - nb.y) * 0.5f,
                                                      (na.z - nb.z) * 0.5f, (na.w - nb.w) * 0.5f);
                    g.dpos[M.dbase + j] = make_float4(q.x, q.y, q.z, 1.f);
                    if (error < FLT_MAX) {
                        const unsigned long long k2 =
                            ((unsigned long long)__float_as_uint(error) << 32) | (unsigned long long)(uint32_t)i;
                        atomicMin(&g.key[va], k2);
                        atomicMin(&g.key[vb], k2);
                    }
                }
            }
            g.pick[M.dbase + j] = ok ? i : -1;
            valid += ok;
        }
        int tot;
        block_scan(valid, &tot);
        if (threadIdx.x == 0 && tot) atomicAdd(&M.valid, tot);
    }
}

// CollapseEdges (:281-304): an edge both of whose ends chose it collapses onto its
// smaller end.
__global__ __launch_bounds__(kSimpThreads) void k_collapse(SG g) {
    LVN_TILES(g.dt, g.ndt) {
        SMesh& M = g.mesh[tl_m];
        const int n = min(M.draws, M.accepted);
        if (tl_s >= n) continue;
        const int2* edges = edges_of(g, M);
        int coll = 0;
        for (int r = 0; r < kSimpPer; ++r) {
            const int j = tl_s + threadIdx.x + r * kSimpThreads;
            if (j >= n) continue;
            const int i = g.pick[M.dbase + j];
            if (i < 0) continue;
            const int2 e = edges[i];
            const int va = M.vbase + e.x, vb = M.vbase + e.y;
            if ((uint32_t)(g.key[va] & 0xffffffffull) == (uint32_t)i && (uint32_t)(g.key[vb] & 0xffffffffull) == (uint32_t)i) {
                coll += atomicExch(&g.tgt[vb], e.x) != e.x;  // an edge drawn twice collapses once
                g.wpos[va] = g.dpos[M.dbase + j];
                g.wnrm[va] = g.dnrm[M.dbase + j];
            }
        }
        int tot;
        block_scan(coll, &tot);
        if (threadIdx.x == 0 && tot) {
            atomicAdd(&M.coll, tot);
            atomicAdd(&M.coll_it, tot);
        }
    }
}

// Per mesh: which of RemoveTriangles / RemoveEdges can change anything this iteration.
// Both run after a valid draw; without a collapse the edges stay as they are, and the
// triangles too unless this is the first pass over input holding degenerate triangles.
__global__ void k_iter_mid(SG g) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.n_mesh) return;
    SMesh& M = g.mesh[m];
    const bool run = M.active && M.valid > 0;
    M.ecompact = run && M.coll_it > 0 ? 1 : 0;
    M.tpass = run && (M.coll_it > 0 || (M.iter == 1 && M.deg_in > 0)) ? 1 : 0;
}

// RemoveTriangles (:308-349) in place: a triangle with a collapsed vertex is remapped and
// moves its counts from the old vertices to the new ones, or is dropped (marked dead) when
// the remap makes it degenerate; degenerate input triangles are dropped in the first pass.
// The survivors keep their order, so the final compaction gives the reference's list.
__global__ __launch_bounds__(kSimpThreads) void k_tri_remap(SG g) {
    LVN_TILES(g.tt, g.ntt) {
        SMesh& M = g.mesh[tl_m];
        if (!M.tpass) continue;
        const bool first = M.iter == 1;
        int drop = 0;
        for (int r = 0; r < kSimpPer; ++r) {
            const int t = tl_s + threadIdx.x + r * kSimpThreads;
            if (t >= M.nt0) continue;
            int* p = g.tri + 3 * (int64_t)(M.tbase + t);
            const int a = p[0];
            if (a < 0) continue;
            const int b = p[1], c = p[2];
            const int a2 = remap(g, M, a), b2 = remap(g, M, b), c2 = remap(g, M, c);
            const bool touched = a2 != a || b2 != b || c2 != c;
            const bool degen = a2 == b2 || a2 == c2 || b2 == c2;
            if (!touched && !(first && degen)) continue;
            atomicSub(&g.vtc[M.vbase + a], 1);
            atomicSub(&g.vtc[M.vbase + b], 1);
            atomicSub(&g.vtc[M.vbase + c], 1);
            if (degen) {
                p[0] = -1;
                ++drop;
                continue;
            }
            p[0] = a2;
            p[1] = b2;
            p[2] = c2;
            atomicAdd(&g.vtc[M.vbase + a2], 1);
            atomicAdd(&g.vtc[M.vbase + b2], 1);
            atomicAdd(&g.vtc[M.vbase + c2], 1);
        }
        int tot;
        block_scan(drop, &tot);
        if (threadIdx.x == 0 && tot) {
            atomicSub(&M.T, tot);
            atomicAdd(&M.drop_it, tot);
        }
    }
}

// ---- the end ---------------------------------------------------------------------------

__global__ __launch_bounds__(kSimpThreads) void k_used_clear(SG g) {
    LVN_TILES(g.vt, g.nvt) {
        const SMesh& M = g.mesh[tl_m];
        for (int r = 0; r < kSimpPer; ++r) {
            const int i = tl_s + threadIdx.x + r * kSimpThreads;
            if (i < M.nv) g.bcnt[M.vbase + i] = 0;
        }
    }
}

__global__ __launch_bounds__(kSimpThreads) void k_used_mark(SG g) {
    LVN_TILES(g.tt, g.ntt) {
        const SMesh& M = g.mesh[tl_m];
        for (int r = 0; r < kSimpPer; ++r) {
            const int t = tl_s + threadIdx.x + r * kSimpThreads;
            if (t >= M.nt0) continue;
            const int* p = g.tri + 3 * (int64_t)(M.tbase + t);
            if (p[0] < 0) continue;
            g.bcnt[M.vbase + p[0]] = 1;
            g.bcnt[M.vbase + p[1]] = 1;
            g.bcnt[M.vbase + p[2]] = 1;
        }
    }
}

}  // namespace

// The simplifier's device buffers, grown on demand and kept for reuse.
struct SimpState {
    char* d_scratch = nullptr;
    uint64_t scratch_bytes = 0;
    uint32_t* d_u = nullptr;  // std::mt19937 outputs for u_seed
    int u_len = 0;
    uint32_t u_seed = 0;
    int* d_active = nullptr;
    int* h_active = nullptr;
    lvn_vertex* d_v = nullptr;  // a host mesh staged for lvn_simplify_mesh
    int32_t* d_t = nullptr;
    int64_t v_cap = 0, t_cap = 0;
    ~SimpState() {
        if (d_scratch) (void)hipFree(d_scratch);
        if (d_u) (void)hipFree(d_u);
        if (d_active) (void)hipFree(d_active);
        if (h_active) (void)hipHostFree(h_active);
        if (d_v) (void)hipFree(d_v);
        if (d_t) (void)hipFree(d_t);
    }
};

void destroy_simp_state(SimpState* s) { delete s; }

// Engine outputs kept beyond the most draws one iteration can take: room for the redraws.
constexpr int kDrawMargin = 4096;

static int64_t tiles_of(int64_t n) { return (n + kSimpTile - 1) / kSimpTile; }
static int ecap_of(int nt) { return (int)(3 * (int64_t)nt / 2 + 1); }  // a kept side occurs at least twice
static int dcap_of(int ecap, float frac) {  // the most draws, and room for their redraws
    const int d = (int)((float)ecap * frac) + 1;
    return d + d / 16 + 64;
}

// Scratch bytes of one mesh (vertex, triangle, edge and draw arrays, tile tables).
static uint64_t mesh_bytes(int nv, int nt, float frac) {
    const int64_t e = ecap_of(nt), d = dcap_of((int)e, frac);
    const int64_t tiles = tiles_of(nv) + tiles_of(nt) + tiles_of(e) + tiles_of(d);
    return (uint64_t)nv * 80 + (uint64_t)nt * 24 + (uint64_t)e * 16 + (uint64_t)d * 36 + (uint64_t)tiles * 16 + 1024;
}

template <typename T>
static int grow(T*& p, int64_t& cap, int64_t want) {
    if (want <= cap) return LVN_OK;
    if (p) LVN_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, sizeof(T) * (size_t)want) != hipSuccess) {
        p = nullptr;
        return LVN_ERR_NOMEM;
    }
    cap = want;
    return LVN_OK;
}

static int ensure_u(SimpState& st, uint32_t seed, int len) {
    if (st.d_u && st.u_seed == seed && st.u_len >= len) return LVN_OK;
    std::vector<uint32_t> u((size_t)len);
    std::mt19937 prng;
    prng.seed(seed);
    for (auto& x : u) x = (uint32_t)prng();
    if (st.d_u) LVN_HIP(hipFree(st.d_u));
    st.d_u = nullptr;
    st.u_len = 0;
    if (hipMalloc(&st.d_u, sizeof(uint32_t) * u.size()) != hipSuccess) {
        st.d_u = nullptr;
        return LVN_ERR_NOMEM;
    }
    LVN_HIP(hipMemcpy(st.d_u, u.data(), sizeof(uint32_t) * u.size(), hipMemcpyHostToDevice));
    st.u_len = len;
    st.u_seed = seed;
    return LVN_OK;
}

static dim3 tile_grid(int ntiles, int cus) { return dim3((unsigned)std::max(1, std::min(ntiles, 8 * std::max(cus, 1)))); }
static dim3 mesh_grid(int n) { return dim3((unsigned)((n + 255) / 256)); }

// Simplify one group of meshes (all of them resident in scratch at once); their final
// state lands in `out` (one SMesh per job, in job order).
static int run_group(SimpState& st, const SimpJob* jobs, int n, const lvn_simplify_options& opt, int cus,
                     hipStream_t stream, std::vector<SMesh>& out) {
    std::vector<SMesh> mesh((size_t)n);
    std::vector<int2> vt, tt, et, dt;
    int64_t V = 0, T = 0, E = 0, D = 0;
    int max_d = 1;
    for (int m = 0; m < n; ++m) {
        const SimpJob& j = jobs[m];
        SMesh& M = mesh[m];
        std::memset(&M, 0, sizeof(M));
        M.vbase = (int)V;
        M.nv = j.nv;
        M.tbase = (int)T;
        M.nt0 = j.nt;
        M.ebase = (int)E;
        M.ecap = ecap_of(j.nt);
        M.dbase = (int)D;
        M.dcap = dcap_of(M.ecap, opt.edge_fraction);
        max_d = std::max(max_d, M.dcap);
        auto add = [](std::vector<int2>& tl, int m2, int count, int& first, int& num) {
            first = (int)tl.size();
            for (int s = 0; s < count; s += kSimpTile) tl.push_back(make_int2(m2, s));
            num = (int)tl.size() - first;
        };
        add(vt, m, M.nv, M.vt0, M.vtn);
        add(tt, m, M.nt0, M.tt0, M.ttn);
        add(et, m, M.ecap, M.et0, M.etn);
        add(dt, m, M.dcap, M.dt0, M.dtn);
        M.target = (int)((float)j.nt * opt.target_percentage);
        M.max_error = j.max_error;
        M.max_edge = j.max_edge_size;
        for (int a = 0; a < 4; ++a) M.off[a] = j.offset[a];
        M.vin = j.vin;
        M.tin = j.tin;
        M.vout = j.vout;
        M.tout = j.tout;
        M.T = j.nt;
        M.last = -1;
        M.one = 1;
        V += M.nv;
        T += M.nt0;
        E += M.ecap;
        D += M.dcap;
    }
    int r = ensure_u(st, opt.seed, max_d + kDrawMargin);
    if (r) return r;
    if (!st.d_active) {
        LVN_HIP(hipMalloc(&st.d_active, sizeof(int)));
        LVN_HIP(hipHostMalloc(&st.h_active, sizeof(int), hipHostMallocDefault));
    }
    // one scratch block: arrays 256-byte aligned
    const int64_t ntile_max = std::max<int64_t>({(int64_t)vt.size(), (int64_t)tt.size(), (int64_t)et.size(), 1});
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = align256(o + bytes);
        return at;
    };
    const uint64_t o_mesh = take(sizeof(SMesh) * n), o_vt = take(8 * vt.size()), o_tt = take(8 * tt.size()),
                   o_et = take(8 * et.size()), o_dt = take(8 * dt.size());
    const uint64_t o_wpos = take(16 * V), o_wnrm = take(16 * V), o_wcol = take(16 * V), o_vtc = take(4 * V),
                   o_key = take(8 * V), o_tgt = take(4 * V), o_bcnt = take(4 * V), o_boff = take(4 * V),
                   o_bnd = take(4 * V), o_ucnt = take(4 * V);
    const uint64_t o_tri = take(12 * T), o_bucket = take(12 * T), o_ea = take(8 * E), o_eb = take(8 * E);
    const uint64_t o_pick = take(4 * D), o_dpos = take(16 * D), o_dnrm = take(16 * D);
    const uint64_t o_tsum = take(4 * ntile_max), o_toff = take(4 * ntile_max);
    if (o > st.scratch_bytes) {
        if (st.d_scratch) LVN_HIP(hipFree(st.d_scratch));
        st.d_scratch = nullptr;
        st.scratch_bytes = 0;
        if (hipMalloc(&st.d_scratch, o) != hipSuccess) {
            st.d_scratch = nullptr;
            return LVN_ERR_NOMEM;
        }
        st.scratch_bytes = o;
    }
    char* s = st.d_scratch;
    LVN_HIP(hipMemcpyAsync(s + o_mesh, mesh.data(), sizeof(SMesh) * n, hipMemcpyHostToDevice, stream));
    LVN_HIP(hipMemcpyAsync(s + o_vt, vt.data(), 8 * vt.size(), hipMemcpyHostToDevice, stream));
    LVN_HIP(hipMemcpyAsync(s + o_tt, tt.data(), 8 * tt.size(), hipMemcpyHostToDevice, stream));
    LVN_HIP(hipMemcpyAsync(s + o_et, et.data(), 8 * et.size(), hipMemcpyHostToDevice, stream));
    LVN_HIP(hipMemcpyAsync(s + o_dt, dt.data(), 8 * dt.size(), hipMemcpyHostToDevice, stream));
    SG g;
    g.mesh = reinterpret_cast<SMesh*>(s + o_mesh);
    g.n_mesh = n;
    g.vt = reinterpret_cast<const int2*>(s + o_vt);
    g.tt = reinterpret_cast<const int2*>(s + o_tt);
    g.et = reinterpret_cast<const int2*>(s + o_et);
    g.dt = reinterpret_cast<const int2*>(s + o_dt);
    g.nvt = (int)vt.size();
    g.ntt = (int)tt.size();
    g.net = (int)et.size();
    g.ndt = (int)dt.size();
    g.wpos = reinterpret_cast<float4*>(s + o_wpos);
    g.wnrm = reinterpret_cast<float4*>(s + o_wnrm);
    g.wcol = reinterpret_cast<float4*>(s + o_wcol);
    g.vtc = reinterpret_cast<int*>(s + o_vtc);
    g.key = reinterpret_cast<unsigned long long*>(s + o_key);
    g.tgt = reinterpret_cast<int*>(s + o_tgt);
    g.bcnt = reinterpret_cast<int*>(s + o_bcnt);
    g.boff = reinterpret_cast<int*>(s + o_boff);
    g.bnd = reinterpret_cast<int*>(s + o_bnd);
    g.ucnt = reinterpret_cast<int*>(s + o_ucnt);
    g.tri = reinterpret_cast<int*>(s + o_tri);
    g.bucket = reinterpret_cast<int*>(s + o_bucket);
    g.edge_a = reinterpret_cast<int2*>(s + o_ea);
    g.edge_b = reinterpret_cast<int2*>(s + o_eb);
    g.pick = reinterpret_cast<int*>(s + o_pick);
    g.dpos = reinterpret_cast<float4*>(s + o_dpos);
    g.dnrm = reinterpret_cast<float4*>(s + o_dnrm);
    g.tsum = reinterpret_cast<int*>(s + o_tsum);
    g.toff = reinterpret_cast<int*>(s + o_toff);
    g.u = st.d_u;
    g.n_u = st.u_len;
    g.opt = opt;
    const dim3 B(kSimpThreads), GV = tile_grid(g.nvt, cus), GT = tile_grid(g.ntt, cus), GE = tile_grid(g.net, cus),
               GD = tile_grid(g.ndt, cus), GM = mesh_grid(n), GW((unsigned)((n + 3) / 4));
    // candidate edges
    hipLaunchKernelGGL(k_init_vertices, GV, B, 0, stream, g);
    hipLaunchKernelGGL(k_init_triangles, GT, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_count<BucketOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    hipLaunchKernelGGL((k_mesh_scan<&SMesh::vt0, &SMesh::vtn, &SMesh::dummy, &SMesh::one>), GW, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_emit<BucketOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    hipLaunchKernelGGL(k_bucket_scatter, GT, B, 0, stream, g);
    hipLaunchKernelGGL(k_bucket_sort, GV, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_count<EdgeOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    hipLaunchKernelGGL((k_mesh_scan<&SMesh::vt0, &SMesh::vtn, &SMesh::E, &SMesh::one>), GW, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_emit<EdgeEmitOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    LVN_HIP(hipGetLastError());
    for (int it = 0; it < opt.max_iterations; ++it) {
        LVN_HIP(hipMemsetAsync(st.d_active, 0, sizeof(int), stream));
        hipLaunchKernelGGL(k_iter_begin, GM, dim3(256), 0, stream, g, st.d_active);
        LVN_HIP(hipMemcpyAsync(st.h_active, st.d_active, sizeof(int), hipMemcpyDeviceToHost, stream));
        LVN_HIP(hipStreamSynchronize(stream));
        if (*st.h_active == 0) break;  // every mesh has stopped
        hipLaunchKernelGGL(k_iter_reset, GV, B, 0, stream, g);
        hipLaunchKernelGGL(k_tile_count<DrawOp>, GD, B, 0, stream, g, g.dt, g.ndt);
        hipLaunchKernelGGL((k_mesh_scan<&SMesh::dt0, &SMesh::dtn, &SMesh::accepted, &SMesh::active>), GW, B, 0,
                           stream, g);
        hipLaunchKernelGGL(k_tile_emit<DrawOp>, GD, B, 0, stream, g, g.dt, g.ndt);
        hipLaunchKernelGGL(k_valid, GD, B, 0, stream, g);
        hipLaunchKernelGGL(k_collapse, GD, B, 0, stream, g);
        hipLaunchKernelGGL(k_iter_mid, GM, dim3(256), 0, stream, g);
        hipLaunchKernelGGL(k_tri_remap, GT, B, 0, stream, g);
        hipLaunchKernelGGL(k_tile_count<EdgeKeepOp>, GE, B, 0, stream, g, g.et, g.net);
        hipLaunchKernelGGL((k_mesh_scan<&SMesh::et0, &SMesh::etn, &SMesh::En, &SMesh::ecompact>), GW, B, 0, stream,
                           g);
        hipLaunchKernelGGL(k_tile_emit<EdgeKeepOp>, GE, B, 0, stream, g, g.et, g.net);
        LVN_HIP(hipGetLastError());
    }
    // the compacted vertices and the live triangles written to the outputs
    hipLaunchKernelGGL(k_used_clear, GV, B, 0, stream, g);
    hipLaunchKernelGGL(k_used_mark, GT, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_count<UsedOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    hipLaunchKernelGGL((k_mesh_scan<&SMesh::vt0, &SMesh::vtn, &SMesh::nvo, &SMesh::one>), GW, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_emit<UsedOp>, GV, B, 0, stream, g, g.vt, g.nvt);
    hipLaunchKernelGGL(k_tile_count<TriFinalOp>, GT, B, 0, stream, g, g.tt, g.ntt);
    hipLaunchKernelGGL((k_mesh_scan<&SMesh::tt0, &SMesh::ttn, &SMesh::Tf, &SMesh::one>), GW, B, 0, stream, g);
    hipLaunchKernelGGL(k_tile_emit<TriFinalOp>, GT, B, 0, stream, g, g.tt, g.ntt);
    LVN_HIP(hipGetLastError());
    out.resize((size_t)n);
    LVN_HIP(hipMemcpyAsync(out.data(), s + o_mesh, sizeof(SMesh) * n, hipMemcpyDeviceToHost, stream));
    LVN_HIP(hipStreamSynchronize(stream));
    for (const SMesh& M : out)
        if (M.fail || M.Tf != M.T) return LVN_ERR_DEVICE;  // too few draw positions for the redraws
    return LVN_OK;
}

// Simplify jobs[0, n) on `stream`, in groups whose scratch fits the budget; res gets
// (vertices, triangles, iterations, collapses) per job.
static int run_simplify(SimpState& st, const std::vector<SimpJob>& jobs, const lvn_simplify_options& opt, int cus,
                        hipStream_t stream, std::vector<int4>& res) {
    res.assign(jobs.size(), make_int4(0, 0, 0, 0));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 8ull << 30;
    const uint64_t budget = std::max<uint64_t>(1ull << 30, std::min<uint64_t>(32ull << 30, free_b / 3));
    size_t first = 0;
    std::vector<SMesh> out;
    while (first < jobs.size()) {
        size_t last = first;
        uint64_t bytes = 0;
        while (last < jobs.size()) {
            const uint64_t b = mesh_bytes(jobs[last].nv, jobs[last].nt, opt.edge_fraction);
            if (last > first && bytes + b > budget) break;
            bytes += b;
            ++last;
        }
        const int r = run_group(st, jobs.data() + first, (int)(last - first), opt, cus, stream, out);
        if (r) return r;
        for (size_t k = first; k < last; ++k) {
            const SMesh& M = out[k - first];
            res[k] = make_int4(M.nvo, M.T, M.iter, M.coll);
        }
        first = last;
    }
    return LVN_OK;
}

static bool valid_options(const lvn_simplify_options* o) {
    return o && o->edge_fraction >= 0.f && o->edge_fraction <= 1.f && o->max_iterations >= 0 &&
           o->max_iterations <= 1000000 && o->target_percentage >= 0.f && o->target_percentage <= 1.f;
}

static SimpState* simp_state(Ctx* c) {
    if (!c->simp) c->simp = new (std::nothrow) SimpState();
    return c->simp;
}

}  // namespace lvn

using namespace lvn;

extern "C" {

int lvn_simplify_mesh(lvn_ctx* ctx, lvn_vertex* vertices, int32_t* num_vertices, int32_t* triangles,
                      int32_t* num_triangles, const float offset[4], const lvn_simplify_options* options,
                      int32_t* stats) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !num_vertices || !num_triangles || !offset || !valid_options(options)) return LVN_ERR_INVALID;
    const int32_t nv = *num_vertices, nt = *num_triangles;
    if (nv < 0 || nt < 0 || (nv > 0 && !vertices) || (nt > 0 && !triangles)) return LVN_ERR_INVALID;
    if (nt > (INT32_MAX - 2) / 3) return LVN_ERR_INVALID;
    for (int64_t i = 0; i < 3 * (int64_t)nt; ++i)
        if (triangles[i] < 0 || triangles[i] >= nv) return LVN_ERR_INVALID;
    if (stats) stats[0] = stats[1] = 0;
    if (!g_state.init) return LVN_ERR_STATE;  // the library was shut down
    if (nt < 100 || nv < 100) return LVN_OK;  // unchanged (src/ng_mesh_simplify.cpp:443-446)
    SimpState* st = simp_state(c);
    if (!st) return LVN_ERR_NOMEM;
    int r = grow(st->d_v, st->v_cap, nv);
    if (!r) r = grow(st->d_t, st->t_cap, 3 * (int64_t)nt);
    if (r) return r;
    LVN_HIP(hipMemcpyAsync(st->d_v, vertices, sizeof(lvn_vertex) * (size_t)nv, hipMemcpyHostToDevice, c->stream));
    LVN_HIP(hipMemcpyAsync(st->d_t, triangles, 12 * (size_t)nt, hipMemcpyHostToDevice, c->stream));
    std::vector<SimpJob> jobs(1);
    SimpJob& j = jobs[0];
    std::memset(&j, 0, sizeof(j));
    j.vin = st->d_v;
    j.tin = st->d_t;
    j.vout = st->d_v;
    j.tout = st->d_t;
    j.nv = nv;
    j.nt = nt;
    j.max_error = options->max_error;
    j.max_edge_size = options->max_edge_size;
    for (int a = 0; a < 4; ++a) j.offset[a] = offset[a];
    std::vector<int4> res;
    r = run_simplify(*st, jobs, *options, c->cus, c->stream, res);
    if (r) return r;
    LVN_HIP(hipMemcpy(vertices, st->d_v, sizeof(lvn_vertex) * (size_t)res[0].x, hipMemcpyDeviceToHost));
    LVN_HIP(hipMemcpy(triangles, st->d_t, 12 * (size_t)res[0].y, hipMemcpyDeviceToHost));
    *num_vertices = res[0].x;
    *num_triangles = res[0].y;
    if (stats) {
        stats[0] = res[0].z;
        stats[1] = res[0].w;
    }
    return LVN_OK;
}

int lvn_batch_simplify(lvn_batch* bh, const lvn_simplify_options* options, int32_t* results) {
    Batch* b = reinterpret_cast<Batch*>(bh);
    if (!b || !valid_options(options)) return LVN_ERR_INVALID;
    if (!g_state.init) return LVN_ERR_STATE;
    if (b->n == 0) return LVN_OK;
    Ctx* c = &b->ctx;
    SimpState* st = simp_state(c);
    if (!st) return LVN_ERR_NOMEM;
    LVN_HIP(hipMemcpyAsync(c->h_counters, c->d_counters, sizeof(Counters) * b->n, hipMemcpyDeviceToHost, c->stream));
    LVN_HIP(hipStreamSynchronize(c->stream));
    std::vector<int32_t> counts((size_t)2 * b->n, 0);
    if (b->simplified) counts = b->simp_counts;  // a second pass simplifies the simplified meshes
    std::vector<SimpJob> jobs;
    jobs.reserve(b->n);
    for (int i = 0; i < b->n; ++i) {
        const Counters& k = c->h_counters[i];
        if (!b->simplified) {
            counts[2 * i] = k.status ? 0 : k.edges > 0 && k.quads > 0 ? k.voxels : 0;
            counts[2 * i + 1] = k.status ? 0 : k.edges > 0 ? 2 * k.quads : 0;
        }
        if (results) {
            results[4 * i] = counts[2 * i];
            results[4 * i + 1] = counts[2 * i + 1];
            results[4 * i + 2] = k.status ? LVN_ERR_OVERFLOW : 0;
            results[4 * i + 3] = 0;
        }
        // overflowed chunks (lvn_batch_get_mesh reports them) and meshes under 100 vertices
        // or triangles (src/ng_mesh_simplify.cpp:443-446) stay as they are
        if (k.status || counts[2 * i] < 100 || counts[2 * i + 1] < 100) continue;
        const ChunkDesc& cd = c->h_chunks[i];
        SimpJob j;
        std::memset(&j, 0, sizeof(j));
        char* slot = c->d_arena + (uint64_t)i * c->L.slot_bytes;
        j.vin = reinterpret_cast<const lvn_vertex*>(slot + c->L.verts);
        j.tin = reinterpret_cast<const int32_t*>(slot + c->L.tris);
        j.vout = reinterpret_cast<lvn_vertex*>(slot + c->L.verts);
        j.tout = reinterpret_cast<int32_t*>(slot + c->L.tris);
        j.nv = counts[2 * i];
        j.nt = counts[2 * i + 1];
        j.result = i;
        // ConstructClipmapNodeData, src/clipmap.cpp:455-462: centre offset, limits scaled by
        // the voxel size LEAF_SIZE_SCALE * (size / CLIPMAP_LEAF_SIZE)
        const float leaf = (float)(kLeafSizeScale * cd.scale);
        j.max_error = options->max_error * leaf;
        j.max_edge_size = options->max_edge_size * leaf;
        const float half = (float)cd.size / 2.f;
        for (int a = 0; a < 3; ++a) j.offset[a] = (float)cd.min[a] + half;
        j.offset[3] = 0.f;
        jobs.push_back(j);
    }
    std::vector<int4> res;
    const int r = run_simplify(*st, jobs, *options, c->cus, c->stream, res);
    if (r) return r;
    for (size_t q = 0; q < jobs.size(); ++q) {
        const int i = jobs[q].result;
        counts[2 * i] = res[q].x;
        counts[2 * i + 1] = res[q].y;
        if (results) {
            results[4 * i] = res[q].x;
            results[4 * i + 1] = res[q].y;
            results[4 * i + 2] = res[q].z;
            results[4 * i + 3] = res[q].w;
        }
    }
    b->simp_counts.swap(counts);
    b->simplified = true;
    return LVN_OK;
}

}  // extern "C"
