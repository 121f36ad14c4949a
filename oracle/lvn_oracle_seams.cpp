// lvn_oracle_seams.cpp — CPU oracle of seam stitching.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the seam tests and bench.py's
// cpu_baseline seam leg.  Never linked into the product library.
//
// A pointer-free restatement of leven's CPU seam path, kept in the reference's own
// shape (bottom-up octree with a hash map per level, recursive contouring) so that
// it can be read against it line by line:
//   SelectSeamNodes / FilterSeamNode      src/clipmap.cpp:508-569
//   seam node -> octree leaf               src/clipmap.cpp:398-416
//   Octree_ConstructUpwards / Parents      src/octree.cpp:22-148
//   GenerateVertexIndices                  src/octree.cpp:196-231
//   ContourProcessEdge / EdgeProc /
//   FaceProc / CellProc                    src/octree.cpp:235-474
//   Octree_GenerateMesh                    src/octree.cpp:520-550
// Integer bookkeeping only; vertex data is copied, never computed.

#include "lvn_oracle.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <atomic>
#include <unordered_map>
#include <vector>

namespace {

constexpr int kLeafSizeScale = 4;      // src/volume_constants.h:7-8
constexpr int kClipmapLeafSize = 256;  // src/volume_constants.h:15 (4 * 64)

// Dual contouring tables of src/contour_constants.h (Ju et al.'s cell/face/edge
// recursion).  Child i of a node sits at offset (bit 2, bit 1, bit 0) of i in units
// of the child size (CHILD_MIN_OFFSETS, src/volume_constants.h:24-35).
const int kEdgeCorners[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                 {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};
const int kCellFaces[12][3] = {{0, 4, 0}, {1, 5, 0}, {2, 6, 0}, {3, 7, 0}, {0, 2, 1}, {4, 6, 1},
                               {1, 3, 1}, {5, 7, 1}, {0, 1, 2}, {2, 3, 2}, {4, 5, 2}, {6, 7, 2}};
const int kCellEdges[6][5] = {{0, 1, 2, 3, 0}, {4, 5, 6, 7, 0}, {0, 4, 1, 5, 1},
                              {2, 6, 3, 7, 1}, {0, 2, 4, 6, 2}, {1, 3, 5, 7, 2}};
const int kFaceFaces[3][4][3] = {{{4, 0, 0}, {5, 1, 0}, {6, 2, 0}, {7, 3, 0}},
                                 {{2, 0, 1}, {6, 4, 1}, {3, 1, 1}, {7, 5, 1}},
                                 {{1, 0, 2}, {3, 2, 2}, {5, 4, 2}, {7, 6, 2}}};
const int kFaceEdges[3][4][6] = {{{1, 4, 0, 5, 1, 1}, {1, 6, 2, 7, 3, 1}, {0, 4, 6, 0, 2, 2}, {0, 5, 7, 1, 3, 2}},
                                 {{0, 2, 3, 0, 1, 0}, {0, 6, 7, 4, 5, 0}, {1, 2, 0, 6, 4, 2}, {1, 3, 1, 7, 5, 2}},
                                 {{1, 1, 0, 3, 2, 0}, {1, 5, 4, 7, 6, 0}, {0, 1, 5, 0, 4, 1}, {0, 3, 7, 2, 6, 1}}};
const int kEdgeEdges[3][2][5] = {{{3, 2, 1, 0, 0}, {7, 6, 5, 4, 0}},
                                 {{5, 1, 4, 0, 1}, {7, 3, 6, 2, 1}},
                                 {{6, 4, 2, 0, 2}, {7, 5, 3, 1, 2}}};
const int kProcessEdge[3][4] = {{3, 2, 1, 0}, {7, 5, 6, 4}, {11, 10, 9, 8}};
const int kFaceEdgeOrder[2][4] = {{0, 0, 1, 1}, {0, 1, 0, 1}};

enum NodeType { kNodeNone = 0, kNodeInternal = 1, kNodeLeaf = 2 };

struct Node {
    int type = kNodeNone;
    int min[3] = {0, 0, 0};
    int size = 0;
    int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    const lvn_seam_node* info = nullptr;  // leaves
    int index = -1;                       // leaves: vertex index
};

// ChunkMinForPosition, src/volume.cpp:30-34.
inline void chunk_min(const int* p, int* out) {
    const int mask = ~(kClipmapLeafSize - 1);
    for (int a = 0; a < 3; ++a) out[a] = p[a] & mask;
}
inline bool same_chunk(const Node& a, const Node& b) {
    int ca[3], cb[3];
    chunk_min(a.min, ca);
    chunk_min(b.min, cb);
    return ca[0] == cb[0] && ca[1] == cb[1] && ca[2] == cb[2];
}

// FilterSeamNode, src/clipmap.cpp:508-542.
bool filter_seam_node(int neighbour, const int* bounds, const int* mn, const int* mx) {
    switch (neighbour) {
    case 0: return mx[0] == bounds[0] || mx[1] == bounds[1] || mx[2] == bounds[2];
    case 1: return mn[2] == bounds[2];
    case 2: return mn[1] == bounds[1];
    case 3: return mn[1] == bounds[1] || mn[2] == bounds[2];
    case 4: return mn[0] == bounds[0];
    case 5: return mn[0] == bounds[0] || mn[2] == bounds[2];
    case 6: return mn[0] == bounds[0] || mn[1] == bounds[1];
    case 7: return mn[0] == bounds[0] && mn[1] == bounds[1] && mn[2] == bounds[2];
    }
    return false;
}

struct SeamTree {
    std::vector<Node> nodes;
    int root_min[3];
    int root_size;
    bool overlap = false;

    // ConstructParents, src/octree.cpp:29-88.  Parents are keyed by HashOctreeMin of
    // their root-relative min (:22-26).  A child slot claimed twice keeps the last
    // claim, as the reference's assignment does: a seam node offered under two
    // neighbours (a coarse node next to several of the host's neighbour regions) is
    // one leaf.  A slot claimed by a leaf and by a parent built from smaller leaves
    // means two offered nodes overlap, which no clipmap produces; that is flagged.
    std::vector<int> construct_parents(const std::vector<int>& in, int parent_size) {
        std::unordered_map<uint64_t, int> parents;
        std::vector<int> order;
        for (int ni : in) {
            int pos[3];
            for (int a = 0; a < 3; ++a) {
                const int local = nodes[ni].min[a] - root_min[a];
                pos[a] = nodes[ni].min[a] - local % parent_size;
            }
            const uint64_t key = (uint64_t)(uint32_t)(pos[0] - root_min[0]) |
                                 ((uint64_t)(uint32_t)(pos[1] - root_min[1]) << 20) |
                                 ((uint64_t)(uint32_t)(pos[2] - root_min[2]) << 40);
            auto it = parents.find(key);
            int pi;
            if (it == parents.end()) {
                Node p;
                p.type = kNodeInternal;
                std::memcpy(p.min, pos, sizeof pos);
                p.size = parent_size;
                nodes.push_back(p);
                pi = (int)nodes.size() - 1;
                parents.emplace(key, pi);
                order.push_back(pi);
            } else {
                pi = it->second;
            }
            bool found = false;
            for (int i = 0; i < 8; ++i) {
                const int h = parent_size / 2;
                if (pos[0] + h * ((i >> 2) & 1) == nodes[ni].min[0] && pos[1] + h * ((i >> 1) & 1) == nodes[ni].min[1] &&
                    pos[2] + h * (i & 1) == nodes[ni].min[2]) {
                    const int prev = nodes[pi].child[i];
                    if (prev >= 0 && (nodes[prev].type != nodes[ni].type || nodes[ni].type != kNodeLeaf)) overlap = true;
                    nodes[pi].child[i] = ni;
                    found = true;
                    break;
                }
            }
            if (!found) overlap = true;  // misaligned: the reference asserts
        }
        return order;
    }

    // Octree_ConstructUpwards, src/octree.cpp:92-148.  Returns the root, or -1.
    int construct_upwards(std::vector<int> list) {
        if (list.empty()) return -1;
        std::stable_sort(list.begin(), list.end(), [&](int a, int b) { return nodes[a].size < nodes[b].size; });
        while (nodes[list.front()].size != nodes[list.back()].size) {
            const int size = nodes[list.front()].size;
            size_t run = 1;
            while (nodes[list[run]].size == size) ++run;
            std::vector<int> head(list.begin(), list.begin() + run);
            std::vector<int> next = construct_parents(head, size * 2);
            next.insert(next.end(), list.begin() + run, list.end());
            list.swap(next);
        }
        for (int parent_size = nodes[list.front()].size * 2; parent_size <= root_size; parent_size *= 2)
            list = construct_parents(list, parent_size);
        if (list.size() != 1) {
            overlap = true;
            return -1;
        }
        return list.front();
    }
};

struct Mesher {
    SeamTree& t;
    const float* colour;
    std::vector<lvn_vertex> verts;
    std::vector<int32_t> tris;
    std::vector<int32_t> leaf_min;  // per vertex: leaf min xyz, size

    // GenerateVertexIndices, src/octree.cpp:196-231.
    void vertex_indices(int ni) {
        if (ni < 0) return;
        Node& n = t.nodes[ni];
        bool has_children = false;
        for (int i = 0; i < 8; ++i) has_children |= n.child[i] >= 0;
        if (has_children) {
            for (int i = 0; i < 8; ++i) vertex_indices(n.child[i]);
            return;
        }
        n.index = (int)verts.size();
        lvn_vertex v;
        const lvn_seam_node& s = *n.info;
        v.pos[0] = s.pos[0]; v.pos[1] = s.pos[1]; v.pos[2] = s.pos[2]; v.pos[3] = 0.f;
        v.normal[0] = s.normal[0]; v.normal[1] = s.normal[1]; v.normal[2] = s.normal[2]; v.normal[3] = 0.f;
        v.colour[0] = colour[0]; v.colour[1] = colour[1]; v.colour[2] = colour[2];
        v.colour[3] = (float)(s.local_min_and_material[3] >> 8);
        verts.push_back(v);
        leaf_min.insert(leaf_min.end(), {n.min[0], n.min[1], n.min[2], n.size});
    }

    // ContourProcessEdge, src/octree.cpp:235-282.
    void process_edge(const int* nd, int dir) {
        int min_size = 0x7fffffff, min_index = 0, idx[4] = {-1, -1, -1, -1};
        bool flip = false, change[4] = {false, false, false, false};
        for (int i = 0; i < 4; ++i) {
            const Node& n = t.nodes[nd[i]];
            if (n.type == kNodeInternal) continue;
            const int e = kProcessEdge[dir][i];
            const int corners = n.info->local_min_and_material[3] & 0xff;
            const int m0 = (corners >> kEdgeCorners[e][0]) & 1;
            const int m1 = (corners >> kEdgeCorners[e][1]) & 1;
            if (n.size < min_size) {
                min_size = n.size;
                min_index = i;
                flip = m1 != 1;
            }
            idx[i] = n.index;
            change[i] = m0 != m1;
        }
        if (!change[min_index]) return;
        const int a[6] = {idx[0], idx[1], idx[3], idx[0], idx[3], idx[2]};
        const int b[6] = {idx[0], idx[3], idx[1], idx[0], idx[2], idx[3]};
        tris.insert(tris.end(), flip ? b : a, (flip ? b : a) + 6);
    }

    // ContourEdgeProc, src/octree.cpp:286-338: only edges whose four nodes span more
    // than one LOD0 chunk (by node min) are contoured.
    void edge_proc(const int* nd, int dir) {
        for (int i = 0; i < 4; ++i)
            if (nd[i] < 0) return;
        bool spans = false;
        for (int i = 1; i < 4; ++i) spans |= !same_chunk(t.nodes[nd[0]], t.nodes[nd[i]]);
        if (!spans) return;
        bool branch[4], any = false;
        for (int i = 0; i < 4; ++i) any |= (branch[i] = t.nodes[nd[i]].type == kNodeInternal);
        if (!any) {
            process_edge(nd, dir);
            return;
        }
        for (int i = 0; i < 2; ++i) {
            int sub[4];
            for (int j = 0; j < 4; ++j)
                sub[j] = branch[j] ? t.nodes[nd[j]].child[kEdgeEdges[dir][i][j]] : nd[j];
            edge_proc(sub, kEdgeEdges[dir][i][4]);
        }
    }

    // ContourFaceProc, src/octree.cpp:342-418.
    void face_proc(const int* nd, int dir) {
        if (nd[0] < 0 || nd[1] < 0) return;
        if (same_chunk(t.nodes[nd[0]], t.nodes[nd[1]])) return;
        const bool branch[2] = {t.nodes[nd[0]].type == kNodeInternal, t.nodes[nd[1]].type == kNodeInternal};
        if (!branch[0] && !branch[1]) return;
        for (int i = 0; i < 4; ++i) {
            int sub[2];
            for (int j = 0; j < 2; ++j) sub[j] = branch[j] ? t.nodes[nd[j]].child[kFaceFaces[dir][i][j]] : nd[j];
            face_proc(sub, kFaceFaces[dir][i][2]);
        }
        for (int i = 0; i < 4; ++i) {
            const int* order = kFaceEdgeOrder[kFaceEdges[dir][i][0]];
            int sub[4];
            for (int j = 0; j < 4; ++j) {
                const int src = nd[order[j]];
                sub[j] = branch[order[j]] ? t.nodes[src].child[kFaceEdges[dir][i][1 + j]] : src;
            }
            edge_proc(sub, kFaceEdges[dir][i][5]);
        }
    }

    // ContourCellProc, src/octree.cpp:422-474.
    void cell_proc(int ni) {
        if (ni < 0 || t.nodes[ni].type == kNodeLeaf) return;
        const Node n = t.nodes[ni];
        for (int i = 0; i < 8; ++i) cell_proc(n.child[i]);
        for (int i = 0; i < 12; ++i) {
            const int sub[2] = {n.child[kCellFaces[i][0]], n.child[kCellFaces[i][1]]};
            face_proc(sub, kCellFaces[i][2]);
        }
        for (int i = 0; i < 6; ++i) {
            const int sub[4] = {n.child[kCellEdges[i][0]], n.child[kCellEdges[i][1]], n.child[kCellEdges[i][2]],
                                n.child[kCellEdges[i][3]]};
            edge_proc(sub, kCellEdges[i][4]);
        }
    }
};

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// One seam: returns LVN_OK, LVN_ERR_INVALID or LVN_ERR_CAPACITY (counts always set).
int seam_mesh(const int32_t* host_min, int32_t host_size, const lvn_seam_source* sources, int32_t n_sources,
              const float* colour, int32_t vpc, lvn_vertex* v, int32_t v_cap, int32_t* nv, int32_t* tri,
              int32_t tri_cap, int32_t* nt, int32_t* leaf_min = nullptr) {
    *nv = 0;
    *nt = 0;
    if (host_size <= 0 || vpc <= 0 || n_sources < 0 || (n_sources > 0 && !sources)) return LVN_ERR_INVALID;
    SeamTree t;
    std::memcpy(t.root_min, host_min, sizeof t.root_min);
    t.root_size = host_size * 2;
    std::vector<int> leaves;
    const int bounds[3] = {host_min[0] + host_size, host_min[1] + host_size, host_min[2] + host_size};
    for (int s = 0; s < n_sources; ++s) {
        const lvn_seam_source& src = sources[s];
        if (src.neighbour < 0 || src.neighbour > 7 || src.count < 0 || (src.count > 0 && !src.nodes))
            return LVN_ERR_INVALID;
        const int leaf_size = src.size / vpc;                     // src/clipmap.cpp:400
        const int filter_size = src.size / (vpc * kLeafSizeScale);  // src/clipmap.cpp:555
        for (int j = 0; j < src.count; ++j) {
            const lvn_seam_node& info = src.nodes[j];
            Node n;
            n.type = kNodeLeaf;
            n.size = leaf_size;
            for (int a = 0; a < 3; ++a) n.min[a] = info.local_min_and_material[a] * leaf_size + src.min[a];
            const int mx[3] = {n.min[0] + filter_size * kLeafSizeScale, n.min[1] + filter_size * kLeafSizeScale,
                               n.min[2] + filter_size * kLeafSizeScale};
            bool inside = true;
            for (int a = 0; a < 3; ++a)
                inside &= n.min[a] >= host_min[a] && n.min[a] < host_min[a] + t.root_size;
            if (!filter_seam_node(src.neighbour, bounds, n.min, mx) || !inside) continue;
            // the octree needs power-of-two leaves aligned in the root
            if (!pow2(leaf_size) || leaf_size >= t.root_size || t.root_size % leaf_size) return LVN_ERR_INVALID;
            for (int a = 0; a < 3; ++a)
                if ((n.min[a] - host_min[a]) % leaf_size) return LVN_ERR_INVALID;
            n.info = &info;
            t.nodes.push_back(n);
            leaves.push_back((int)t.nodes.size() - 1);
        }
    }
    if (!pow2(t.root_size)) return LVN_ERR_INVALID;
    const int root = t.construct_upwards(leaves);
    if (t.overlap) return LVN_ERR_INVALID;
    Mesher m{t, colour, {}, {}, {}};
    if (root >= 0) {
        m.vertex_indices(root);
        m.cell_proc(root);
    }
    *nv = (int32_t)m.verts.size();
    *nt = (int32_t)(m.tris.size() / 3);
    if (*nv > v_cap || *nt > tri_cap) return LVN_ERR_CAPACITY;
    if (v && *nv) std::memcpy(v, m.verts.data(), sizeof(lvn_vertex) * m.verts.size());
    if (tri && *nt) std::memcpy(tri, m.tris.data(), sizeof(int32_t) * m.tris.size());
    if (leaf_min && *nv) std::memcpy(leaf_min, m.leaf_min.data(), sizeof(int32_t) * m.leaf_min.size());
    return LVN_OK;
}

}  // namespace

extern "C" int orc_seam_mesh(const int32_t host_min[3], int32_t host_size, const lvn_seam_source* sources,
                             int32_t n_sources, const float colour[3], int32_t voxels_per_chunk, lvn_vertex* v,
                             int32_t v_cap, int32_t* nv, int32_t* tri, int32_t tri_cap, int32_t* nt) {
    return seam_mesh(host_min, host_size, sources, n_sources, colour, voxels_per_chunk, v, v_cap, nv, tri, tri_cap,
                     nt);
}

extern "C" int orc_seam_mesh_leaves(const int32_t host_min[3], int32_t host_size, const lvn_seam_source* sources,
                                    int32_t n_sources, int32_t voxels_per_chunk, int32_t cap, int32_t* n,
                                    int32_t* leaf_min_size) {
    const float colour[3] = {0.f, 0.f, 0.f};
    int32_t nt = 0;
    std::vector<lvn_vertex> v((size_t)std::max(cap, 0));
    std::vector<int32_t> t;
    int rc = seam_mesh(host_min, host_size, sources, n_sources, colour, voxels_per_chunk, v.data(), cap, n, nullptr,
                       0x7fffffff, &nt, nullptr);
    if (rc != LVN_OK) return rc;
    return seam_mesh(host_min, host_size, sources, n_sources, colour, voxels_per_chunk, v.data(), cap, n, nullptr,
                     0x7fffffff, &nt, leaf_min_size);
}

// Position-weighted checksum of a mesh's bytes as 32-bit words (vertices, then
// triangles): sum of word[i] * (i + 1) mod 2^64.
static uint64_t mesh_checksum(const lvn_vertex* v, int32_t nv, const int32_t* t, int32_t nt) {
    uint64_t c = 0, i = 0;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(v);
    for (size_t k = 0; k < (size_t)nv * (sizeof(lvn_vertex) / 4); ++k) c += (uint64_t)w[k] * ++i;
    w = reinterpret_cast<const uint32_t*>(t);
    for (size_t k = 0; k < 3 * (size_t)nt; ++k) c += (uint64_t)w[k] * ++i;
    return c;
}

// Many seams on `threads` host threads (the CPU baseline); totals += (vertices, triangles,
// failed seams); per_seam (optional, 4 per seam) = status, vertices, triangles, checksum.
extern "C" double orc_seam_meshes(int32_t n, const lvn_seam_req* reqs, const lvn_seam_source* sources,
                                  int32_t voxels_per_chunk, int32_t threads, int64_t* totals, int64_t* per_seam) {
    std::atomic<int> next{0};
    std::atomic<int64_t> verts{0}, tris{0}, failed{0};
    auto worker = [&]() {
        std::vector<lvn_vertex> v;
        std::vector<int32_t> t;
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) break;
            int32_t nv = 0, nt = 0;
            const lvn_seam_req& r = reqs[i];
            int rc = seam_mesh(r.host_min, r.host_size, sources + r.first_source, r.n_sources, r.colour,
                               voxels_per_chunk, v.data(), (int32_t)v.size(), &nv, t.data(),
                               (int32_t)(t.size() / 3), &nt);
            if (rc == LVN_ERR_CAPACITY) {
                v.resize(nv);
                t.resize(3 * (size_t)nt);
                rc = seam_mesh(r.host_min, r.host_size, sources + r.first_source, r.n_sources, r.colour,
                               voxels_per_chunk, v.data(), nv, &nv, t.data(), nt, &nt);
            }
            if (rc != LVN_OK) failed.fetch_add(1);
            verts.fetch_add(nv);
            tris.fetch_add(nt);
            if (per_seam) {
                per_seam[4 * (size_t)i] = rc;
                per_seam[4 * (size_t)i + 1] = nv;
                per_seam[4 * (size_t)i + 2] = nt;
                per_seam[4 * (size_t)i + 3] = rc == LVN_OK ? (int64_t)mesh_checksum(v.data(), nv, t.data(), nt) : 0;
            }
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (totals) {
        totals[0] += verts.load();
        totals[1] += tris.load();
        totals[2] += failed.load();
    }
    return secs;
}
