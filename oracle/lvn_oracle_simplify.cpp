// lvn_oracle_simplify.cpp — CPU oracle of the post-mesh simplifier.
//
// TEST INFRASTRUCTURE ONLY (see lvn_oracle.h).  A sequential restatement of leven's
// ngMeshSimplifier (src/ng_mesh_simplify.cpp:438-539) and of the SSE QEF it calls
// (qef_solve_from_points_4d_interleaved, src/qef_simd.h:403-533), lane for lane.  The
// random candidate edges are drawn exactly as the reference's code draws them when built
// with libstdc++: a std::mt19937 seeded anew in every iteration, and
// std::uniform_int_distribution<int>(0, E - 1), which libstdc++ (GCC 11 and later) maps
// with Lemire's multiply-and-reject method (MSVC, which leven was built with, maps the
// same engine output differently).  The one change of arithmetic: the Givens cosine's
// _mm_rsqrt_ps (an approximation whose bits the platform chooses, src/qef_simd.h:206)
// is the IEEE 1 / sqrt.  rotateq_xy's slips (src/qef_simd.h:228-274: row 0's diagonal
// stored twice, y formed from cc * u and cc + v, y stored into the off-diagonal entries,
// the third diagonal never rotated) and the pseudo-inverse's product with V where V^T is
// meant (:371-382) are kept: they are what the reference computes.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "lvn_oracle.h"

namespace {

constexpr int kCollapseMaxDegree = 16;  // COLLAPSE_MAX_DEGREE, src/ng_mesh_simplify.cpp:17

// The SSE solver's 4-lane values: F4 is one __m128, M4 one Mat4x4 (rows).
struct F4 {
    float v[4];
};
struct M4 {
    F4 row[4];
};

inline F4 f4(const float* p) { return F4{{p[0], p[1], p[2], p[3]}}; }
inline F4 f4(float x, float y, float z, float w) { return F4{{x, y, z, w}}; }

// vec4_dot, src/qef_simd.h:101-110: (x x' + y y') + (w w' + z z').
inline float vec4_dot(const F4& a, const F4& b) {
    const float m0 = a.v[0] * b.v[0], m1 = a.v[1] * b.v[1], m2 = a.v[2] * b.v[2], m3 = a.v[3] * b.v[3];
    return (m0 + m1) + (m3 + m2);
}
// glm::dot of two vec4: (x x' + y y') + (z z' + w w').
inline float glm_dot4(const F4& a, const F4& b) {
    return (a.v[0] * b.v[0] + a.v[1] * b.v[1]) + (a.v[2] * b.v[2] + a.v[3] * b.v[3]);
}

// vec4_mul_m4x4, src/qef_simd.h:114-121: the row vector a times B.
inline F4 vec4_mul_m4x4(const F4& a, const M4& B) {
    F4 r;
    for (int c = 0; c < 4; ++c)
        r.v[c] = ((a.v[0] * B.row[0].v[c] + a.v[1] * B.row[1].v[c]) + a.v[2] * B.row[2].v[c]) + a.v[3] * B.row[3].v[c];
    return r;
}

// givens_coeffs_sym, src/qef_simd.h:163-224: the three pairs' (c, s) from one vtav.
void givens_coeffs_sym(F4& c, F4& s, const M4& t) {
    const F4 pp = f4(t.row[0].v[0], t.row[0].v[0], t.row[1].v[1], 0.f);
    const F4 pq = f4(t.row[0].v[1], t.row[0].v[2], t.row[1].v[2], 0.f);
    const F4 qq = f4(t.row[1].v[1], t.row[2].v[2], t.row[2].v[2], 0.f);
    for (int l = 0; l < 4; ++l) {
        const float tau = (qq.v[l] - pp.v[l]) / (pq.v[l] * 2.f);
        const float stt = std::sqrt(tau * tau + 1.f);
        const float tn = 1.f / (tau >= 0.f ? tau + stt : tau - stt);
        const float cl = 1.f / std::sqrt(1.f + tn * tn);  // _mm_rsqrt_ps in the reference
        const float sl = tn * cl;
        const bool zero = pq.v[l] == 0.f;
        c.v[l] = zero ? 1.f : cl;
        s.v[l] = zero ? 0.f : sl;
    }
}

// rotateq_xy, src/qef_simd.h:228-274, as written.
void rotateq_xy(M4& t, const F4& c, const F4& s) {
    const F4 u = f4(t.row[0].v[0], t.row[0].v[0], t.row[1].v[1], 0.f);
    const F4 v = f4(t.row[1].v[1], t.row[2].v[2], t.row[2].v[2], 0.f);
    const F4 a = f4(t.row[0].v[1], t.row[0].v[2], t.row[1].v[2], 0.f);
    F4 x, y;
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

// svd_rotate, src/qef_simd.h:278-319.
void svd_rotate(M4& t, M4& V, float c, float s, int a, int b) {
    if (t.row[a].v[b] == 0.f) return;
    const F4 u = f4(V.row[0].v[a], V.row[1].v[a], V.row[2].v[a], t.row[0].v[3 - b]);
    const F4 w = f4(V.row[0].v[b], V.row[1].v[b], V.row[2].v[b], t.row[1 - a].v[2]);
    F4 x, y;
    for (int l = 0; l < 4; ++l) {
        x.v[l] = c * u.v[l] - s * w.v[l];
        y.v[l] = s * u.v[l] + c * w.v[l];
    }
    for (int r = 0; r < 3; ++r) V.row[r].v[a] = x.v[r];
    t.row[0].v[3 - b] = x.v[3];
    for (int r = 0; r < 3; ++r) V.row[r].v[b] = y.v[r];
    t.row[1 - a].v[2] = y.v[3];
    t.row[a].v[b] = 0.f;
}

// svd_solve_ATA_ATb, src/qef_simd.h:323-398 (svd_solve_sym with 5 sweeps, svd_invdet,
// svd_pseudoinverse): x = ATb times the pseudo-inverse.
F4 svd_solve_ATA_ATb(const M4& ATA, const F4& ATb) {
    M4 V{};
    V.row[0] = f4(1.f, 0.f, 0.f, 0.f);
    V.row[1] = f4(0.f, 1.f, 0.f, 0.f);
    V.row[2] = f4(0.f, 0.f, 1.f, 0.f);
    V.row[3] = f4(0.f, 0.f, 0.f, 0.f);
    M4 t = ATA;
    for (int i = 0; i < 5; ++i) {  // SVD_NUM_SWEEPS
        F4 c, s;
        givens_coeffs_sym(c, s, t);
        rotateq_xy(t, c, s);
        svd_rotate(t, V, c.v[0], s.v[0], 0, 1);
        svd_rotate(t, V, c.v[1], s.v[1], 0, 2);
        svd_rotate(t, V, c.v[2], s.v[2], 1, 2);
    }
    const F4 sigma = f4(t.row[0].v[0], t.row[1].v[1], t.row[2].v[2], 0.f);
    F4 inv;
    for (int l = 0; l < 4; ++l) {  // svd_invdet: min(|x|, |1/x|) >= 0.1 ? 1/x : 0
        const float ax = std::fabs(sigma.v[l]);
        const float r = 1.f / sigma.v[l];
        const float ar = std::fabs(r);
        const float mn = ax < ar ? ax : ar;  // _mm_min_ps
        inv.v[l] = mn >= 0.1f ? r : 0.f;
    }
    M4 o{};
    for (int r = 0; r < 3; ++r)
        for (int l = 0; l < 4; ++l) o.row[r].v[l] = V.row[r].v[l] * inv.v[l];
    o.row[3] = f4(0.f, 0.f, 0.f, 0.f);
    M4 p;
    for (int r = 0; r < 4; ++r) p.row[r] = vec4_mul_m4x4(o.row[r], V);
    return vec4_mul_m4x4(ATb, p);
}

// qef_solve_from_points over the two endpoints (src/qef_simd.h:403-478, count 2): the
// solved position in out, the error |ATb - x ATA|^2 with x relative to the mass point.
float qef_two(const float* p0, const float* n0, const float* p1, const float* n1, float out[4]) {
    M4 ATA{};
    F4 ATb = f4(0.f, 0.f, 0.f, 0.f), acc = f4(0.f, 0.f, 0.f, 0.f);
    const float* ps[2] = {p0, p1};
    const float* ns[2] = {n0, n1};
    for (int k = 0; k < 2; ++k) {  // qef_simd_add, src/qef_simd.h:403-422
        const F4 p = f4(ps[k]), n = f4(ns[k]);
        for (int r = 0; r < 3; ++r)
            for (int l = 0; l < 4; ++l) ATA.row[r].v[l] = ATA.row[r].v[l] + n.v[r] * n.v[l];
        const float d = vec4_dot(p, n);
        const F4 dd = f4(d, d, d, 0.f);
        for (int l = 0; l < 4; ++l) ATb.v[l] = ATb.v[l] + dd.v[l] * n.v[l];
        for (int l = 0; l < 4; ++l) acc.v[l] = acc.v[l] + p.v[l];
    }
    // qef_simd_solve, src/qef_simd.h:436-455
    F4 mp;
    for (int l = 0; l < 4; ++l) mp.v[l] = acc.v[l] / acc.v[3];
    const F4 pm = vec4_mul_m4x4(mp, ATA);
    F4 r;
    for (int l = 0; l < 4; ++l) r.v[l] = ATb.v[l] - pm.v[l];
    F4 x = svd_solve_ATA_ATb(ATA, r);
    const F4 xa = vec4_mul_m4x4(x, ATA);  // qef_simd_calc_error, :426-432
    F4 e;
    for (int l = 0; l < 4; ++l) e.v[l] = ATb.v[l] - xa.v[l];
    const float err = vec4_dot(e, e);
    for (int l = 0; l < 4; ++l) out[l] = x.v[l] + mp.v[l];
    return err;
}

struct Edge {
    uint32_t mn, mx;
};

}  // namespace

extern "C" float orc_simplify_qef(const float p0[4], const float n0[4], const float p1[4], const float n1[4],
                                  float out[4]) {
    return qef_two(p0, n0, p1, n1, out);
}

extern "C" int orc_simplify_mesh(const lvn_vertex* vin, int32_t nv, const int32_t* tin, int32_t nt,
                                 const float offset[4], const lvn_simplify_options* o, lvn_vertex* vout,
                                 int32_t* nv_out, int32_t* tout, int32_t* nt_out, int32_t* stats) {
    if (!vin || !tin || !offset || !o || !vout || !tout || !nv_out || !nt_out) return LVN_ERR_INVALID;
    if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0;
    // too small to simplify: unchanged (src/ng_mesh_simplify.cpp:443-446)
    if (nt < 100 || nv < 100) {
        std::memcpy(vout, vin, sizeof(lvn_vertex) * (size_t)nv);
        std::memcpy(tout, tin, sizeof(int32_t) * 3 * (size_t)nt);
        *nv_out = nv;
        *nt_out = nt;
        return LVN_OK;
    }
    std::vector<lvn_vertex> verts(vin, vin + nv);
    for (auto& v : verts)
        for (int a = 0; a < 4; ++a) v.pos[a] = v.pos[a] - offset[a];
    std::vector<int32_t> tris(tin, tin + 3 * (size_t)nt);
    std::vector<int> vtc(nv, 0);
    for (int t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) vtc[tris[3 * t + k]] += 1;

    // BuildCandidateEdges, src/ng_mesh_simplify.cpp:118-172
    std::vector<uint64_t> keys;
    keys.reserve(3 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        const int i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
        auto key = [](int a, int b) { return ((uint64_t)(uint32_t)std::max(a, b) << 32) | (uint32_t)std::min(a, b); };
        keys.push_back(key(i0, i1));
        keys.push_back(key(i1, i2));
        keys.push_back(key(i0, i2));
    }
    std::sort(keys.begin(), keys.end(), [](uint64_t l, uint64_t r) {  // Edge::idx_: min in the low word
        return l < r;
    });
    std::vector<Edge> filtered;
    std::vector<char> boundary(nv, 0);
    {
        uint64_t prev = keys[0];
        int count = 1;
        for (size_t idx = 1; idx < keys.size(); ++idx) {
            const uint64_t cur = keys[idx];
            if (cur != prev) {
                const Edge e{(uint32_t)(prev & 0xffffffffu), (uint32_t)(prev >> 32)};
                if (count == 1) {
                    boundary[e.mn] = 1;
                    boundary[e.mx] = 1;
                } else {
                    filtered.push_back(e);
                }
                count = 1;
            } else {
                ++count;
            }
            prev = cur;
        }
        // the last run is never examined (the reference's loop ends without it)
    }
    std::vector<Edge> edges;
    for (const Edge& e : filtered)
        if (!boundary[e.mn] && !boundary[e.mx]) edges.push_back(e);

    std::vector<float> cpos(4 * edges.size()), cnrm(4 * edges.size());
    std::vector<int> valid, edge_id(nv), target(nv);
    const int target_tris = (int)((float)nt * o->target_percentage);
    int iterations = 0, collapsed = 0;
    while ((int)tris.size() / 3 > target_tris && iterations++ < o->max_iterations) {
        std::fill(edge_id.begin(), edge_id.end(), -1);
        std::fill(target.begin(), target.end(), -1);
        valid.clear();
        // FindValidCollapses, src/ng_mesh_simplify.cpp:176-277
        const int E = (int)edges.size();
        const int draws = (int)((float)E * o->edge_fraction);
        std::vector<int> picks;
        picks.reserve(draws);
        if (draws > 0) {  // :193-207, the generator seeded anew in every iteration
            std::mt19937 prng;
            prng.seed(o->seed);
            std::uniform_int_distribution<int> dist(0, E - 1);
            for (int j = 0; j < draws; ++j) picks.push_back(dist(prng));
        }
        std::sort(picks.begin(), picks.end());
        std::vector<float> min_cost(nv, FLT_MAX);
        for (int i : picks) {
            const Edge& e = edges[i];
            const lvn_vertex& a = verts[e.mn];
            const lvn_vertex& b = verts[e.mx];
            const float cos_angle = glm_dot4(f4(a.normal), f4(b.normal));
            if (cos_angle < o->min_angle_cosine) continue;
            const F4 d = f4(b.pos[0] - a.pos[0], b.pos[1] - a.pos[1], b.pos[2] - a.pos[2], b.pos[3] - a.pos[3]);
            const float edge_size = glm_dot4(d, d);
            if (edge_size > o->max_edge_size * o->max_edge_size) continue;
            if ((double)std::fabs(a.colour[3] - b.colour[3]) > 1e-3) continue;
            const int degree = vtc[e.mn] + vtc[e.mx];
            if (degree > kCollapseMaxDegree) continue;
            float pos[4];
            float error = qef_two(a.pos, a.normal, b.pos, b.normal, pos);
            if (error > 0.f) error = 1.f / error;
            const int penalty = std::max(0, degree - 10);
            error = error + (float)penalty * (o->max_error * 0.1f);
            if (error > o->max_error) continue;
            valid.push_back(i);
            for (int k = 0; k < 4; ++k) cnrm[4 * i + k] = (a.normal[k] - b.normal[k]) * 0.5f;
            cpos[4 * i] = pos[0];
            cpos[4 * i + 1] = pos[1];
            cpos[4 * i + 2] = pos[2];
            cpos[4 * i + 3] = 1.f;
            if (error < min_cost[e.mn]) {
                min_cost[e.mn] = error;
                edge_id[e.mn] = i;
            }
            if (error < min_cost[e.mx]) {
                min_cost[e.mx] = error;
                edge_id[e.mx] = i;
            }
        }
        if (valid.empty()) break;
        // CollapseEdges, src/ng_mesh_simplify.cpp:281-304
        for (int i : valid) {
            const Edge& e = edges[i];
            if (edge_id[e.mn] == i && edge_id[e.mx] == i) {
                if (target[e.mx] != (int)e.mn) ++collapsed;
                target[e.mx] = (int)e.mn;
                for (int k = 0; k < 4; ++k) verts[e.mn].pos[k] = cpos[4 * i + k];
                for (int k = 0; k < 4; ++k) verts[e.mn].normal[k] = cnrm[4 * i + k];
            }
        }
        // RemoveTriangles, src/ng_mesh_simplify.cpp:308-349
        std::fill(vtc.begin(), vtc.end(), 0);
        std::vector<int32_t> kept;
        kept.reserve(tris.size());
        for (size_t t = 0; t < tris.size() / 3; ++t) {
            int idx[3];
            for (int k = 0; k < 3; ++k) {
                idx[k] = tris[3 * t + k];
                const int g = target[idx[k]];
                if (g != -1) idx[k] = g;
            }
            if (idx[0] == idx[1] || idx[0] == idx[2] || idx[1] == idx[2]) continue;
            for (int k = 0; k < 3; ++k) {
                vtc[idx[k]] += 1;
                kept.push_back(idx[k]);
            }
        }
        tris.swap(kept);
        // RemoveEdges, src/ng_mesh_simplify.cpp:353-381
        std::vector<Edge> ke;
        ke.reserve(edges.size());
        for (Edge e : edges) {
            int g = target[e.mn];
            if (g != -1) e.mn = (uint32_t)g;
            g = target[e.mx];
            if (g != -1) e.mx = (uint32_t)g;
            if (e.mn != e.mx) ke.push_back(e);
        }
        edges.swap(ke);
    }
    // CompactVertices, src/ng_mesh_simplify.cpp:385-430, and the offset restored
    std::vector<int> remap(nv, -1);
    std::vector<char> used(nv, 0);
    for (int32_t i : tris) used[i] = 1;
    int nvo = 0;
    for (int i = 0; i < nv; ++i) {
        if (!used[i]) continue;
        remap[i] = nvo;
        lvn_vertex v = verts[i];
        for (int a = 0; a < 4; ++a) v.pos[a] = v.pos[a] + offset[a];
        vout[nvo++] = v;
    }
    for (size_t k = 0; k < tris.size(); ++k) tout[k] = remap[tris[k]];
    *nv_out = nvo;
    *nt_out = (int32_t)(tris.size() / 3);
    if (stats) {
        stats[0] = std::min(iterations, o->max_iterations);  // the loop test counts one past the last
        stats[1] = collapsed;
        stats[2] = (int32_t)edges.size();
        stats[3] = (int32_t)filtered.size();
    }
    return LVN_OK;
}
