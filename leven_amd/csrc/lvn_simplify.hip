// lvn_simplify.hip — post-mesh simplification on the GPU (ngMeshSimplifier,
// src/ng_mesh_simplify.cpp:438-525), one 1024-thread workgroup per mesh.
//
// The reference's iterations are already parallel in shape: a random fraction of the
// candidate edges is evaluated, every vertex keeps its cheapest valid candidate (the
// first in edge order on ties), and an edge collapses only when both its vertices chose
// it, so the collapses of one iteration are independent.  Here each step is a pass of
// the workgroup over the mesh in device memory: candidate edges sorted by (max, min)
// index through per-vertex buckets, the per-vertex choice as a 64-bit atomic minimum of
// (cost bits, edge index) (costs are never negative, so their bits order like their
// values), order-preserving compaction of triangles and edges by workgroup scans.  The
// arithmetic of each candidate (the two-plane QEF, its cost) follows the operation order
// stated in include/lvn_mesher.h (lvn_simplify_options), so results are identical to the
// sequential restatement.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "lvn_kernels.h"
#include "lvn_runtime.h"

namespace lvn {

constexpr int kSimpThreads = 1024;
constexpr int kCollapseMaxDegree = 16;  // COLLAPSE_MAX_DEGREE, src/ng_mesh_simplify.cpp:17

// One mesh: read from (vin, tin), written to (vout, tout), which may be the same arrays.
struct SimpJob {
    const lvn_vertex* vin;
    const int32_t* tin;
    lvn_vertex* vout;
    int32_t* tout;
    int32_t nv, nt;
    int32_t result;  // the caller's index of this mesh (results are written in job order)
    float max_error, max_edge_size;  // this mesh's limits (leven scales them by the chunk's voxel size)
    int32_t pad;
    float offset[4];
};

// Byte offsets of a workgroup's scratch for meshes of at most V vertices and T triangles.
struct SimpLayout {
    uint64_t wpos, wnrm, wcol, vtc, key, tgt, bcnt, boff, bcur, bnd, ucnt, uoff, bucket;
    uint64_t edge_a, edge_b, cpos, cnrm, flags, tri_a, tri_b, bytes;
    int32_t max_v, max_t, max_e;
};

struct SimpArgs {
    const uint32_t* u;  // the first n_u outputs of std::mt19937 seeded with opt.seed
    int n_u;
    const SimpJob* jobs;
    int n_jobs;
    int* queue;
    int* results;
    char* scratch;
    SimpLayout L;
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

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total) by the whole workgroup.
__device__ int wg_scan(const int* in, int* out, int n, int* lds) {
    const int tid = threadIdx.x;
    const int per = (n + kSimpThreads - 1) / kSimpThreads;
    const int b0 = min(n, tid * per), b1 = min(n, b0 + per);
    int sum = 0;
    for (int i = b0; i < b1; ++i) sum += in[i];
    __syncthreads();
    lds[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kSimpThreads; off <<= 1) {
        const int v = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += v;
        __syncthreads();
    }
    int run = lds[tid] - sum;
    for (int i = b0; i < b1; ++i) {
        const int c = in[i];
        out[i] = run;
        run += c;
    }
    const int total = lds[kSimpThreads - 1];
    if (tid == 0) out[n] = total;
    __syncthreads();
    return total;
}

template <typename T>
__device__ __forceinline__ T* at(char* base, uint64_t off) {
    return reinterpret_cast<T*>(base + off);
}

__device__ void simplify_one(const SimpArgs& s, const SimpJob& J, int job, char* w) {
    __shared__ int lds[kSimpThreads];
    __shared__ int sh_E, sh_T, sh_valid, sh_last, sh_iter, sh_coll, sh_redraw, sh_fail;
    const SimpLayout& L = s.L;
    const lvn_simplify_options& o = s.opt;
    const int tid = threadIdx.x;
    const int nv = J.nv, nt = J.nt;
    int* res = s.results + 4 * job;
    // too small to simplify: unchanged (src/ng_mesh_simplify.cpp:443-446)
    if (nt < 100 || nv < 100) {
        if (J.vout != J.vin)
            for (int i = tid; i < nv; i += kSimpThreads) J.vout[i] = J.vin[i];
        if (J.tout != J.tin)
            for (int i = tid; i < 3 * nt; i += kSimpThreads) J.tout[i] = J.tin[i];
        if (tid == 0) {
            res[0] = nv;
            res[1] = nt;
            res[2] = 0;
            res[3] = 0;
        }
        __syncthreads();
        return;
    }
    float4* wpos = at<float4>(w, L.wpos);
    float4* wnrm = at<float4>(w, L.wnrm);
    float4* wcol = at<float4>(w, L.wcol);
    int* vtc = at<int>(w, L.vtc);
    unsigned long long* key = at<unsigned long long>(w, L.key);
    int* tgt = at<int>(w, L.tgt);
    int* bcnt = at<int>(w, L.bcnt);
    int* boff = at<int>(w, L.boff);
    int* bcur = at<int>(w, L.bcur);
    int* bnd = at<int>(w, L.bnd);
    int* ucnt = at<int>(w, L.ucnt);
    int* uoff = at<int>(w, L.uoff);
    int* bucket = at<int>(w, L.bucket);
    int2* ea = at<int2>(w, L.edge_a);
    int2* eb = at<int2>(w, L.edge_b);
    float4* cpos = at<float4>(w, L.cpos);
    float4* cnrm = at<float4>(w, L.cnrm);
    int* flags = at<int>(w, L.flags);
    int* ta = at<int>(w, L.tri_a);
    int* tb = at<int>(w, L.tri_b);
    const float4 off = make_float4(J.offset[0], J.offset[1], J.offset[2], J.offset[3]);

    // ---- working copy: positions relative to the offset, per-vertex state
    if (tid == 0) {
        sh_last = -1;
        sh_redraw = 0;
        sh_fail = 0;
    }
    for (int i = tid; i < nv; i += kSimpThreads) {
        const lvn_vertex v = J.vin[i];
        wpos[i] = make_float4(v.pos[0] - off.x, v.pos[1] - off.y, v.pos[2] - off.z, v.pos[3] - off.w);
        wnrm[i] = make_float4(v.normal[0], v.normal[1], v.normal[2], v.normal[3]);
        wcol[i] = make_float4(v.colour[0], v.colour[1], v.colour[2], v.colour[3]);
        vtc[i] = 0;
        bcnt[i] = 0;
        bnd[i] = 0;
    }
    for (int i = tid; i < 3 * nt; i += kSimpThreads) ta[i] = J.tin[i];
    __syncthreads();
    // ---- BuildCandidateEdges (src/ng_mesh_simplify.cpp:118-172): the 3T triangle edges in
    // buckets by their larger index, each bucket sorted by the smaller: the order of the
    // reference's sort of (min | max << 32)
    for (int t = tid; t < nt; t += kSimpThreads) {
        const int i0 = ta[3 * t], i1 = ta[3 * t + 1], i2 = ta[3 * t + 2];
        atomicAdd(&vtc[i0], 1);
        atomicAdd(&vtc[i1], 1);
        atomicAdd(&vtc[i2], 1);
        atomicAdd(&bcnt[max(i0, i1)], 1);
        atomicAdd(&bcnt[max(i1, i2)], 1);
        atomicAdd(&bcnt[max(i0, i2)], 1);
    }
    __syncthreads();
    wg_scan(bcnt, boff, nv, lds);
    for (int i = tid; i < nv; i += kSimpThreads) {
        bcur[i] = boff[i];
        if (bcnt[i] > 0) atomicMax(&sh_last, i);
    }
    __syncthreads();
    for (int t = tid; t < nt; t += kSimpThreads) {
        const int i0 = ta[3 * t], i1 = ta[3 * t + 1], i2 = ta[3 * t + 2];
        bucket[atomicAdd(&bcur[max(i0, i1)], 1)] = min(i0, i1);
        bucket[atomicAdd(&bcur[max(i1, i2)], 1)] = min(i1, i2);
        bucket[atomicAdd(&bcur[max(i0, i2)], 1)] = min(i0, i2);
    }
    __syncthreads();
    const int last = sh_last;
    // sort each bucket; a run of one marks both ends as boundary; the last run of the last
    // bucket is never examined (the reference's loop ends before it)
    for (int m = tid; m < nv; m += kSimpThreads) {
        int* b = bucket + boff[m];
        const int n = bcnt[m];
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
            if (!(m == last && e == n) && e - k == 1) {
                bnd[b[k]] = 1;
                bnd[m] = 1;
            }
            k = e;
        }
    }
    __syncthreads();
    for (int m = tid; m < nv; m += kSimpThreads) {
        const int* b = bucket + boff[m];
        const int n = bcnt[m];
        int kept = 0, k = 0;
        while (k < n) {
            int e = k + 1;
            while (e < n && b[e] == b[k]) ++e;
            if (!(m == last && e == n) && e - k > 1 && !bnd[b[k]] && !bnd[m]) ++kept;
            k = e;
        }
        ucnt[m] = kept;
    }
    __syncthreads();
    const int E0 = wg_scan(ucnt, uoff, nv, lds);
    for (int m = tid; m < nv; m += kSimpThreads) {
        const int* b = bucket + boff[m];
        const int n = bcnt[m];
        int pos = uoff[m], k = 0;
        while (k < n) {
            int e = k + 1;
            while (e < n && b[e] == b[k]) ++e;
            if (!(m == last && e == n) && e - k > 1 && !bnd[b[k]] && !bnd[m]) ea[pos++] = make_int2(b[k], m);
            k = e;
        }
    }
    if (tid == 0) {
        sh_E = E0;
        sh_T = nt;
        sh_iter = 0;
        sh_coll = 0;
    }
    __syncthreads();

    // ---- the iterations (src/ng_mesh_simplify.cpp:490-512)
    const int target_tris = (int)((float)nt * o.target_percentage);
    for (;;) {
        const int T = sh_T, E = sh_E;
        if (!(T > target_tris && sh_iter < o.max_iterations)) break;
        const int it = sh_iter;
        __syncthreads();
        if (tid == 0) {
            sh_iter = it + 1;
            sh_valid = 0;
        }
        for (int i = tid; i < nv; i += kSimpThreads) {
            key[i] = ~0ull;
            tgt[i] = -1;
        }
        __syncthreads();
        // FindValidCollapses (:181-277).  Draw j is the reference's j-th
        // std::uniform_int_distribution<int>(0, E - 1) value from a std::mt19937 seeded anew
        // in every iteration, as libstdc++ maps it (Lemire: the high word of u * E, drawing
        // again while the low word is below 2^32 mod E): the engine's outputs u come from a
        // table, so only the rare redraw makes the draws sequential.
        const int draws = (int)((float)E * o.edge_fraction);
        const uint32_t En = (uint32_t)E;
        const uint32_t thr = draws > 0 ? (0u - En) % En : 0u;
        for (int j = tid; j < draws; j += kSimpThreads) {
            const uint64_t prod = (uint64_t)s.u[j] * En;
            if ((uint32_t)prod < thr) sh_redraw = 1;
            flags[j] = (int)(prod >> 32);
        }
        __syncthreads();
        if (sh_redraw) {  // every thread reads the flag before thread 0 clears it
            __syncthreads();
            if (tid == 0) {
                int k = 0;
                for (int j = 0; j < draws; ++j) {
                    uint64_t prod = 0;
                    do {
                        prod = k < s.n_u ? (uint64_t)s.u[k] * En : 0;
                        ++k;
                    } while ((uint32_t)prod < thr && k <= s.n_u);
                    flags[j] = (int)(prod >> 32);
                }
                if (k > s.n_u) sh_fail = 1;  // the table ran out (the host sizes it with a wide margin)
                sh_redraw = 0;
            }
            __syncthreads();
            if (sh_fail) break;
        }
        int valid = 0;
        for (int j = tid; j < draws; j += kSimpThreads) {
            const int i = flags[j];
            const int2 e = ea[i];
            int ok = 0;
            const float4 pa = wpos[e.x], pb = wpos[e.y], na = wnrm[e.x], nb = wnrm[e.y];
            const float cos_angle = dot_glm(na, nb);
            const float4 d = make_float4(pb.x - pa.x, pb.y - pa.y, pb.z - pa.z, pb.w - pa.w);
            const int degree = vtc[e.x] + vtc[e.y];
            if (!(cos_angle < o.min_angle_cosine) && !(dot_glm(d, d) > J.max_edge_size * J.max_edge_size) &&
                !((double)fabsf(wcol[e.x].w - wcol[e.y].w) > 1e-3) && !(degree > kCollapseMaxDegree)) {
                float4 q;
                float error = qef_two(pa, na, pb, nb, &q);
                if (error > 0.f) error = 1.f / error;
                const int penalty = max(0, degree - 10);
                error = error + (float)penalty * (J.max_error * 0.1f);
                if (!(error > J.max_error)) {
                    ok = 1;
                    cnrm[i] = make_float4((na.x - nb.x) * 0.5f, (na.y - nb.y) * 0.5f, (na.z - nb.z) * 0.5f,
                                          (na.w - nb.w) * 0.5f);
                    cpos[i] = make_float4(q.x, q.y, q.z, 1.f);
                    if (error < FLT_MAX) {
                        const unsigned long long k2 =
                            ((unsigned long long)__float_as_uint(error) << 32) | (unsigned long long)(uint32_t)i;
                        atomicMin(&key[e.x], k2);
                        atomicMin(&key[e.y], k2);
                    }
                }
            }
            flags[j] = ok ? i : -1;
            valid += ok;
        }
        if (valid) atomicAdd(&sh_valid, valid);
        __syncthreads();
        if (sh_valid == 0) break;
        // CollapseEdges (:281-304): both vertices chose the edge
        int coll = 0;
        for (int j = tid; j < draws; j += kSimpThreads) {
            const int i = flags[j];
            if (i < 0) continue;
            const int2 e = ea[i];
            if ((uint32_t)(key[e.x] & 0xffffffffull) == (uint32_t)i && (uint32_t)(key[e.y] & 0xffffffffull) == (uint32_t)i) {
                coll += atomicExch(&tgt[e.y], e.x) != e.x;  // an edge drawn twice collapses once
                wpos[e.x] = cpos[i];
                wnrm[e.x] = cnrm[i];
            }
        }
        if (coll) atomicAdd(&sh_coll, coll);
        __syncthreads();
        // RemoveTriangles (:308-349), in order
        for (int i = tid; i < nv; i += kSimpThreads) vtc[i] = 0;
        for (int t = tid; t < T; t += kSimpThreads) {
            int i0 = ta[3 * t], i1 = ta[3 * t + 1], i2 = ta[3 * t + 2];
            int g = tgt[i0];
            if (g != -1) i0 = g;
            g = tgt[i1];
            if (g != -1) i1 = g;
            g = tgt[i2];
            if (g != -1) i2 = g;
            ta[3 * t] = i0;
            ta[3 * t + 1] = i1;
            ta[3 * t + 2] = i2;
            flags[t] = (i0 == i1 || i0 == i2 || i1 == i2) ? 0 : 1;
        }
        __syncthreads();
        const int T2 = wg_scan(flags, ucnt, T, lds);  // ucnt (per-vertex counts) is free after the edge build
        for (int t = tid; t < T; t += kSimpThreads) {
            if (!flags[t]) continue;
            const int q = ucnt[t];
            const int i0 = ta[3 * t], i1 = ta[3 * t + 1], i2 = ta[3 * t + 2];
            tb[3 * q] = i0;
            tb[3 * q + 1] = i1;
            tb[3 * q + 2] = i2;
            atomicAdd(&vtc[i0], 1);
            atomicAdd(&vtc[i1], 1);
            atomicAdd(&vtc[i2], 1);
        }
        __syncthreads();
        // RemoveEdges (:353-381), in order
        for (int i = tid; i < E; i += kSimpThreads) {
            int2 e = ea[i];
            int g = tgt[e.x];
            if (g != -1) e.x = g;
            g = tgt[e.y];
            if (g != -1) e.y = g;
            ea[i] = e;
            flags[i] = e.x != e.y ? 1 : 0;
        }
        __syncthreads();
        const int E2 = wg_scan(flags, ucnt, E, lds);
        for (int i = tid; i < E; i += kSimpThreads)
            if (flags[i]) eb[ucnt[i]] = ea[i];
        __syncthreads();
        {
            int* tt = ta;
            ta = tb;
            tb = tt;
            int2* te = ea;
            ea = eb;
            eb = te;
        }
        if (tid == 0) {
            sh_T = T2;
            sh_E = E2;
        }
        __syncthreads();
    }
    __syncthreads();
    // ---- CompactVertices (:385-430) and the offset restored
    const int T = sh_T;
    for (int i = tid; i < nv; i += kSimpThreads) flags[i] = 0;
    __syncthreads();
    for (int i = tid; i < 3 * T; i += kSimpThreads) flags[ta[i]] = 1;
    __syncthreads();
    const int nvo = wg_scan(flags, ucnt, nv, lds);
    for (int i = tid; i < nv; i += kSimpThreads) {
        if (!flags[i]) continue;
        const float4 p = wpos[i], n = wnrm[i], c = wcol[i];
        lvn_vertex v;
        v.pos[0] = p.x + off.x;
        v.pos[1] = p.y + off.y;
        v.pos[2] = p.z + off.z;
        v.pos[3] = p.w + off.w;
        v.normal[0] = n.x;
        v.normal[1] = n.y;
        v.normal[2] = n.z;
        v.normal[3] = n.w;
        v.colour[0] = c.x;
        v.colour[1] = c.y;
        v.colour[2] = c.z;
        v.colour[3] = c.w;
        J.vout[ucnt[i]] = v;
    }
    for (int i = tid; i < 3 * T; i += kSimpThreads) J.tout[i] = ucnt[ta[i]];
    if (tid == 0) {
        res[0] = nvo;
        res[1] = T;
        res[2] = sh_fail ? -1 : sh_iter;
        res[3] = sh_coll;
    }
    __syncthreads();
}

// Persistent: workgroups take meshes from a queue (the host lists them largest first),
// each with its own scratch; job j's results (vertices, triangles, iterations, collapses)
// go to results[4 j ..].
__global__ __launch_bounds__(kSimpThreads) void k_simplify(SimpArgs s) {
    __shared__ int next;
    char* w = s.scratch + (uint64_t)blockIdx.x * s.L.bytes;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(s.queue, 1);
        __syncthreads();
        const int j = next;
        __syncthreads();
        if (j >= s.n_jobs) break;
        simplify_one(s, s.jobs[j], j, w);
    }
}

__global__ void k_simplify_reset(int* queue) { *queue = 0; }

}  // namespace

SimpLayout simplify_layout(int max_v, int max_t) {
    SimpLayout L;
    std::memset(&L, 0, sizeof(L));
    const uint64_t V = (uint64_t)std::max(max_v, 1), T = (uint64_t)std::max(max_t, 1);
    const uint64_t K = 3 * T, E = K / 2 + 1;  // a kept edge occurs at least twice among the 3T
    const uint64_t M = std::max(V, std::max(T, E)) + 1;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.wpos = take(16 * V);
    L.wnrm = take(16 * V);
    L.wcol = take(16 * V);
    L.vtc = take(4 * V);
    L.key = take(8 * V);
    L.tgt = take(4 * V);
    L.bcnt = take(4 * V);
    L.boff = take(4 * (V + 1));
    L.bcur = take(4 * V);
    L.bnd = take(4 * V);
    L.ucnt = take(4 * M);
    L.uoff = take(4 * (V + 1));
    L.bucket = take(4 * K);
    L.edge_a = take(8 * E);
    L.edge_b = take(8 * E);
    L.cpos = take(16 * E);
    L.cnrm = take(16 * E);
    L.flags = take(4 * M);
    L.tri_a = take(4 * K);
    L.tri_b = take(4 * K);
    L.bytes = o;
    L.max_v = (int32_t)V;
    L.max_t = (int32_t)T;
    L.max_e = (int32_t)E;
    return L;
}

// The simplifier's device buffers, grown on demand and kept for reuse.
struct SimpState {
    SimpJob* d_jobs = nullptr;
    int jobs_cap = 0;
    int* d_results = nullptr;
    int* h_results = nullptr;
    int results_cap = 0;
    int* d_queue = nullptr;
    char* d_scratch = nullptr;
    uint64_t scratch_bytes = 0;
    uint32_t* d_u = nullptr;  // std::mt19937 outputs for `u_seed`
    int u_len = 0;
    uint32_t u_seed = 0;
    lvn_vertex* d_v = nullptr;  // a host mesh staged for lvn_simplify_mesh
    int32_t* d_t = nullptr;
    int64_t v_cap = 0, t_cap = 0;
    ~SimpState() {
        if (d_jobs) (void)hipFree(d_jobs);
        if (d_results) (void)hipFree(d_results);
        if (h_results) (void)hipHostFree(h_results);
        if (d_queue) (void)hipFree(d_queue);
        if (d_scratch) (void)hipFree(d_scratch);
        if (d_u) (void)hipFree(d_u);
        if (d_v) (void)hipFree(d_v);
        if (d_t) (void)hipFree(d_t);
    }
};

void destroy_simp_state(SimpState* s) { delete s; }

// Scratch of all workgroups together stays within this many bytes (fewer workgroups
// for large meshes).
constexpr uint64_t kSimpScratchBudget = 8ull << 30;
// Engine outputs kept beyond the most draws one iteration can take: room for the redraws.
constexpr int kDrawMargin = 4096;

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

// Simplify jobs[0, n) on `stream` with up to `groups` persistent workgroups; the results
// (4 ints per job, in job order) land in st.h_results and the call waits for them.
static int run_simplify(SimpState& st, std::vector<SimpJob>& jobs, const lvn_simplify_options& opt, int groups,
                        hipStream_t stream) {
    const int n = (int)jobs.size();
    if (n == 0) return LVN_OK;
    int max_v = 1, max_t = 1;
    for (const SimpJob& j : jobs) {
        max_v = std::max(max_v, j.nv);
        max_t = std::max(max_t, j.nt);
    }
    // largest meshes first: each is one workgroup's, the largest set the pass's length
    std::stable_sort(jobs.begin(), jobs.end(), [](const SimpJob& a, const SimpJob& b) { return a.nt > b.nt; });
    const SimpLayout L = simplify_layout(max_v, max_t);
    groups = (int)std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)groups, (uint64_t)n, kSimpScratchBudget / L.bytes}));
    // the engine's outputs: enough for the most draws of any iteration plus redraws
    const int64_t max_draws = (int64_t)((double)L.max_e * (double)opt.edge_fraction) + 1;
    const int u_len = (int)std::min<int64_t>(max_draws + kDrawMargin, INT32_MAX);
    if (!st.d_u || st.u_seed != opt.seed || st.u_len < u_len) {
        std::vector<uint32_t> u((size_t)u_len);
        std::mt19937 prng;
        prng.seed(opt.seed);
        for (auto& x : u) x = (uint32_t)prng();
        if (st.d_u) LVN_HIP(hipFree(st.d_u));
        st.d_u = nullptr;
        st.u_len = 0;
        if (hipMalloc(&st.d_u, sizeof(uint32_t) * u.size()) != hipSuccess) {
            st.d_u = nullptr;
            return LVN_ERR_NOMEM;
        }
        LVN_HIP(hipMemcpy(st.d_u, u.data(), sizeof(uint32_t) * u.size(), hipMemcpyHostToDevice));
        st.u_len = u_len;
        st.u_seed = opt.seed;
    }
    if (n > st.jobs_cap) {
        if (st.d_jobs) LVN_HIP(hipFree(st.d_jobs));
        st.d_jobs = nullptr;
        st.jobs_cap = 0;
        if (hipMalloc(&st.d_jobs, sizeof(SimpJob) * n) != hipSuccess) {
            st.d_jobs = nullptr;
            return LVN_ERR_NOMEM;
        }
        st.jobs_cap = n;
    }
    if (n > st.results_cap) {
        if (st.d_results) LVN_HIP(hipFree(st.d_results));
        if (st.h_results) LVN_HIP(hipHostFree(st.h_results));
        st.d_results = nullptr;
        st.h_results = nullptr;
        st.results_cap = 0;
        if (hipMalloc(&st.d_results, sizeof(int) * 4 * n) != hipSuccess) {
            st.d_results = nullptr;
            return LVN_ERR_NOMEM;
        }
        LVN_HIP(hipHostMalloc(&st.h_results, sizeof(int) * 4 * n, hipHostMallocDefault));
        st.results_cap = n;
    }
    if (!st.d_queue) LVN_HIP(hipMalloc(&st.d_queue, sizeof(int)));
    const uint64_t need = L.bytes * (uint64_t)groups;
    if (need > st.scratch_bytes) {
        if (st.d_scratch) LVN_HIP(hipFree(st.d_scratch));
        st.d_scratch = nullptr;
        st.scratch_bytes = 0;
        if (hipMalloc(&st.d_scratch, need) != hipSuccess) {
            st.d_scratch = nullptr;
            return LVN_ERR_NOMEM;
        }
        st.scratch_bytes = need;
    }
    LVN_HIP(hipMemcpyAsync(st.d_jobs, jobs.data(), sizeof(SimpJob) * n, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_simplify_reset, dim3(1), dim3(1), 0, stream, st.d_queue);
    SimpArgs a;
    a.u = st.d_u;
    a.n_u = st.u_len;
    a.jobs = st.d_jobs;
    a.n_jobs = n;
    a.queue = st.d_queue;
    a.results = st.d_results;
    a.scratch = st.d_scratch;
    a.L = L;
    a.opt = opt;
    hipLaunchKernelGGL(k_simplify, dim3(groups), dim3(kSimpThreads), 0, stream, a);
    LVN_HIP(hipGetLastError());
    LVN_HIP(hipMemcpyAsync(st.h_results, st.d_results, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, stream));
    LVN_HIP(hipStreamSynchronize(stream));
    for (int i = 0; i < n; ++i)
        if (st.h_results[4 * i + 2] < 0) return LVN_ERR_DEVICE;  // the engine-output table ran out
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
    j.result = 0;
    j.max_error = options->max_error;
    j.max_edge_size = options->max_edge_size;
    for (int a = 0; a < 4; ++a) j.offset[a] = offset[a];
    r = run_simplify(*st, jobs, *options, 1, c->stream);
    if (r) return r;
    const int* res = st->h_results;
    LVN_HIP(hipMemcpy(vertices, st->d_v, sizeof(lvn_vertex) * (size_t)res[0], hipMemcpyDeviceToHost));
    LVN_HIP(hipMemcpy(triangles, st->d_t, 12 * (size_t)res[1], hipMemcpyDeviceToHost));
    *num_vertices = res[0];
    *num_triangles = res[1];
    if (stats) {
        stats[0] = res[2];
        stats[1] = res[3];
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
        if (k.status) continue;  // overflowed: lvn_batch_get_mesh reports it
        if (!b->simplified) {
            counts[2 * i] = k.edges > 0 && k.quads > 0 ? k.voxels : 0;
            counts[2 * i + 1] = k.edges > 0 ? 2 * k.quads : 0;
        }
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
    if (results)
        for (int i = 0; i < b->n; ++i) {
            results[4 * i] = counts[2 * i];
            results[4 * i + 1] = counts[2 * i + 1];
            results[4 * i + 2] = c->h_counters[i].status ? LVN_ERR_OVERFLOW : 0;
            results[4 * i + 3] = 0;
        }
    int r = run_simplify(*st, jobs, *options, 2 * std::max(c->cus, 1), c->stream);
    if (r) return r;
    for (int q = 0; q < (int)jobs.size(); ++q) {
        const int i = jobs[q].result;
        const int* res = st->h_results + 4 * q;
        counts[2 * i] = res[0];
        counts[2 * i + 1] = res[1];
        if (results) {
            results[4 * i] = res[0];
            results[4 * i + 1] = res[1];
            results[4 * i + 2] = res[2];
            results[4 * i + 3] = res[3];
        }
    }
    b->simp_counts.swap(counts);
    b->simplified = true;
    return LVN_OK;
}

}  // extern "C"
