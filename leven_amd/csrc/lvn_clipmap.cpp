// lvn_clipmap.cpp — request producer of the chunk mesher: leven's camera-driven clipmap
// (src/clipmap.cpp), whose updates decide which chunks to mesh, which meshes have gone
// stale, which chunks turned out empty and which seams to regenerate.  Host code only:
// the meshing itself is the caller's (the single-chunk context, a batch or a server).
//
// The node tree differs from the reference's in one respect, stated where it matters:
// leven builds one power-of-two cube around the world bounds and allocates every node
// down to the leaf chunks up front (ConstructChildren, src/clipmap.cpp:715-735; 2.4M
// nodes for its largest world), so its selection also meshes chunks outside the world
// bounds.  Here the roots are the LOD_MAX_NODE_SIZE cells tiling the world box
// [world_min, world_max) from world_min, and a node's children are created the first
// time the selection descends into it.  Inside the world the selection visits the same
// nodes in the same order and applies the same rule.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_set>
#include <vector>

#include "../../include/lvn_mesher.h"

namespace lvn {

namespace {

constexpr int kNumLods = 6;  // NUM_LODS, src/clipmap.cpp:28
// LOD_ACTIVE_DISTANCES in leaf-chunk sizes, src/clipmap.cpp:33-41
constexpr float kLodDistance[kNumLods] = {0.f, 1.5f, 3.5f, 5.5f, 7.5f, 13.5f};
constexpr int kLeafSizeScale = 4;  // src/volume_constants.h:7-8

// CHILD_MIN_OFFSETS, src/volume_constants.h:24-35: x = bit 2, y = bit 1, z = bit 0.
inline int child_offset(int i, int axis) { return (i >> (2 - axis)) & 1; }

struct Node {
    int32_t min[3];
    int32_t size;
    int32_t child = -1;  // first of its 8 children in the pool; -1 until created
    int32_t seams = 0;   // seam nodes of its current mesh
    bool active = false, invalidated = false, empty = false;
    // the empty state its children take when they are created: leven's full tree would
    // have marked them with this node (PropagateEmptyStateDownward)
    bool empty_below = false;
};

// AABB(min, size).pointIsInside, src/aabb.h: half-open.
inline bool inside(const int32_t* mn, int32_t size, const int32_t* p) {
    for (int a = 0; a < 3; ++a)
        if (p[a] < mn[a] || p[a] >= mn[a] + size) return false;
    return true;
}

// AABB::overlaps, src/aabb.h: closed boxes.
inline bool overlaps(const int32_t* mn, int32_t size, const int32_t* bmin, const int32_t* bmax) {
    return !(mn[0] + size < bmin[0] || mn[1] + size < bmin[1] || mn[2] + size < bmin[2] || mn[0] > bmax[0] ||
             mn[1] > bmax[1] || mn[2] > bmax[2]);
}

}  // namespace

struct Clipmap {
    int N = 64, leaf = 256, lod_max = 8192;
    int32_t wmin[3], wmax[3];
    std::vector<Node> nodes;
    std::vector<int> roots;
    std::vector<int> lists[5];  // to mesh, released, selected, active (on request), touched by the last edit
    std::vector<lvn_seam_req> seam_reqs;
    std::vector<lvn_seam_source> seam_srcs;
    std::mutex mu;

    int make_children(int n);
    void select(int n, bool parent_active, const float* cam);
    void release(int n);
    void propagate_empty(int n);
    void find_active(int n, const int32_t* rmin, int32_t rsize, std::vector<int>& out) const;
    bool in_world(const int32_t* p) const;
    void active_under(const int32_t* rmin, int32_t rsize, std::vector<int>& out) const;
    void touched(int n, const int32_t* bmin, const int32_t* bmax, std::vector<int>& out);
    void collect_active(int n, std::vector<int>& out) const;
};

// ConstructChildren (src/clipmap.cpp:715-735), one level, on first use.
int Clipmap::make_children(int n) {
    const int first = (int)nodes.size();
    const Node parent = nodes[n];
    const int h = parent.size / 2;
    for (int i = 0; i < 8; ++i) {
        Node c;
        for (int a = 0; a < 3; ++a) c.min[a] = parent.min[a] + h * child_offset(i, a);
        c.size = h;
        c.empty = c.empty_below = parent.empty_below;
        nodes.push_back(c);
    }
    nodes[n].child = first;
    return first;
}

// SelectActiveClipmapNodes, src/clipmap.cpp:1084-1126, with DistanceToNode (:1055-1065)
// in float as the reference computes it.
void Clipmap::select(int n, bool parent_active, const float* cam) {
    bool node_active = false;
    {
        Node& nd = nodes[n];
        if (!parent_active && nd.size <= lod_max) {
            int idx = 0;
            while ((leaf << idx) < nd.size) ++idx;  // glm::log2 of size / leaf (a power of two)
            const float d = kLodDistance[idx] * (float)leaf;
            float dd[3];
            for (int a = 0; a < 3; ++a) {
                const float lo = (float)nd.min[a], hi = (float)(nd.min[a] + nd.size);
                dd[a] = std::max(std::max(lo - cam[a], 0.f), cam[a] - hi);
            }
            const float dist = std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
            if (dist >= d) {
                lists[2].push_back(n);
                node_active = true;
            }
        }
        if (nd.active && !node_active) nd.invalidated = true;
        if (nd.size <= leaf) return;
        // nodes never created below a selected node hold no state to invalidate
        if (nd.child < 0 && (parent_active || node_active)) return;
    }
    const int first = nodes[n].child >= 0 ? nodes[n].child : make_children(n);
    for (int i = 0; i < 8; ++i) select(first + i, parent_active || node_active, cam);
}

// ReleaseInvalidatedNodes + ReleaseClipmapNodeData, src/clipmap.cpp:1130-1151,644-675.
void Clipmap::release(int n) {
    Node& nd = nodes[n];
    if (nd.invalidated) {
        nd.active = false;
        nd.seams = 0;
        nd.invalidated = false;
        lists[1].push_back(n);
    }
    if (nd.child >= 0)
        for (int i = 0; i < 8; ++i) release(nodes[n].child + i);
}

// PropagateEmptyStateDownward, src/clipmap.cpp:147-161.
void Clipmap::propagate_empty(int n) {
    nodes[n].empty = true;
    if (nodes[n].child >= 0) {
        for (int i = 0; i < 8; ++i) propagate_empty(nodes[n].child + i);
    } else {
        nodes[n].empty_below = true;  // children made later start empty
    }
}

// FindActiveNodes, src/clipmap.cpp:1449-1473, against the box (rmin, rsize).
void Clipmap::find_active(int n, const int32_t* rmin, int32_t rsize, std::vector<int>& out) const {
    const Node& nd = nodes[n];
    if (!(inside(nd.min, nd.size, rmin) || inside(rmin, rsize, nd.min))) return;
    if (nd.active) {
        out.push_back(n);
    } else if (nd.size > leaf && nd.child >= 0) {
        for (int i = 0; i < 8; ++i) find_active(nd.child + i, rmin, rsize, out);
    }
}

// Clipmap::findNode(min, size) finds a node wherever the point lies in the tree.
bool Clipmap::in_world(const int32_t* p) const {
    for (int r : roots)
        if (inside(nodes[r].min, nodes[r].size, p)) return true;
    return false;
}

void Clipmap::active_under(const int32_t* rmin, int32_t rsize, std::vector<int>& out) const {
    for (int r : roots) find_active(r, rmin, rsize, out);
}

// FindNodesInsideAABB, src/clipmap.cpp:1571-1597: post-order, nodes up to LOD_MAX_NODE_SIZE.
// Below a node whose uncreated children would be empty, the children along the box are
// made, so that the edit clears the empty mark of exactly the nodes it overlaps.
void Clipmap::touched(int n, const int32_t* bmin, const int32_t* bmax, std::vector<int>& out) {
    if (!overlaps(nodes[n].min, nodes[n].size, bmin, bmax)) return;
    if (nodes[n].child < 0 && nodes[n].empty_below && nodes[n].size > leaf) make_children(n);
    if (nodes[n].child >= 0)
        for (int i = 0; i < 8; ++i) touched(nodes[n].child + i, bmin, bmax, out);
    if (nodes[n].size <= lod_max) out.push_back(n);
}

void Clipmap::collect_active(int n, std::vector<int>& out) const {
    if (nodes[n].active) out.push_back(n);
    if (nodes[n].child >= 0)
        for (int i = 0; i < 8; ++i) collect_active(nodes[n].child + i, out);
}

}  // namespace lvn

using namespace lvn;

extern "C" {

int lvn_clipmap_create(int32_t voxels_per_chunk, const int32_t world_min[3], const int32_t world_max[3],
                       lvn_clipmap** out) {
    if (!out || !world_min || !world_max) return LVN_ERR_INVALID;
    *out = nullptr;
    const int n = voxels_per_chunk;
    if (n < 4 || n > 64 || (n & (n - 1))) return LVN_ERR_INVALID;
    const int64_t leaf = (int64_t)kLeafSizeScale * n, lod_max = leaf << (kNumLods - 1);
    int64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (world_max[a] <= world_min[a]) return LVN_ERR_INVALID;
        cells *= ((int64_t)world_max[a] - world_min[a] + lod_max - 1) / lod_max;
        if ((int64_t)world_max[a] + lod_max > INT32_MAX || (int64_t)world_min[a] - lod_max < INT32_MIN)
            return LVN_ERR_INVALID;
    }
    if (cells > (1 << 20)) return LVN_ERR_INVALID;
    Clipmap* c = new Clipmap;
    c->N = n;
    c->leaf = (int)leaf;
    c->lod_max = (int)lod_max;
    for (int a = 0; a < 3; ++a) {
        c->wmin[a] = world_min[a];
        c->wmax[a] = world_max[a];
    }
    // roots in x, then y, then z order, each a LOD_MAX_NODE_SIZE cell
    for (int64_t x = world_min[0]; x < world_max[0]; x += lod_max)
        for (int64_t y = world_min[1]; y < world_max[1]; y += lod_max)
            for (int64_t z = world_min[2]; z < world_max[2]; z += lod_max) {
                Node r;
                r.min[0] = (int32_t)x;
                r.min[1] = (int32_t)y;
                r.min[2] = (int32_t)z;
                r.size = (int32_t)lod_max;
                c->roots.push_back((int)c->nodes.size());
                c->nodes.push_back(r);
            }
    *out = reinterpret_cast<lvn_clipmap*>(c);
    return LVN_OK;
}

void lvn_clipmap_destroy(lvn_clipmap* ch) { delete reinterpret_cast<Clipmap*>(ch); }

int lvn_clipmap_update(lvn_clipmap* ch, const float camera[3], const float* frustum, int32_t counts[3]) {
    Clipmap* c = reinterpret_cast<Clipmap*>(ch);
    if (!c || !camera || !counts) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    for (int k = 0; k < 3; ++k) c->lists[k].clear();
    c->seam_reqs.clear();
    c->seam_srcs.clear();
    for (int r : c->roots) c->select(r, false, camera);
    for (int r : c->roots) c->release(r);
    // the filter of Clipmap::update (src/clipmap.cpp:1212-1249): selected nodes neither
    // active nor empty, inside the frustum (AABBInsideFrustum, src/frustum.cpp:30-52);
    // when none is inside, the ones outside it
    std::vector<int> reserve;
    for (int n : c->lists[2]) {
        const Node& nd = c->nodes[n];
        if (nd.active || nd.empty) continue;
        bool in = true;
        if (frustum) {
            for (int p = 0; p < 6 && in; ++p) {
                const float* pl = frustum + 4 * p;
                int outside = 0;
                for (int k = 0; k < 8; ++k) {
                    const float x = (float)(nd.min[0] + ((k & 1) ? nd.size : 0));
                    const float y = (float)(nd.min[1] + ((k & 2) ? nd.size : 0));
                    const float z = (float)(nd.min[2] + ((k & 4) ? nd.size : 0));
                    // glm::dot(p, vec4(corner, 1)): (p.x x + p.y y) + (p.z z + p.w 1)
                    outside += ((pl[0] * x + pl[1] * y) + (pl[2] * z + pl[3] * 1.f)) < 0.f ? 1 : 0;
                }
                in = outside < 8;
            }
        }
        (in ? c->lists[0] : reserve).push_back(n);
    }
    if (c->lists[0].empty()) c->lists[0] = reserve;
    counts[0] = (int32_t)c->lists[0].size();
    counts[1] = (int32_t)c->lists[1].size();
    counts[2] = (int32_t)c->lists[2].size();
    return LVN_OK;
}

int lvn_clipmap_list(lvn_clipmap* ch, int32_t which, lvn_chunk_req* out, int32_t cap, int32_t* n) {
    Clipmap* c = reinterpret_cast<Clipmap*>(ch);
    if (!c || !n || which < 0 || which > LVN_CLIPMAP_TOUCHED || cap < 0 || (cap > 0 && !out)) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (which == LVN_CLIPMAP_ACTIVE) {
        c->lists[3].clear();
        for (int r : c->roots) c->collect_active(r, c->lists[3]);
    }
    const std::vector<int>& l = c->lists[which];
    *n = (int32_t)l.size();
    if ((int32_t)l.size() > cap) return LVN_ERR_CAPACITY;
    for (size_t k = 0; k < l.size(); ++k) {
        const Node& nd = c->nodes[l[k]];
        for (int a = 0; a < 3; ++a) out[k].min[a] = nd.min[a];
        out[k].size = nd.size;
    }
    return LVN_OK;
}

int lvn_clipmap_commit(lvn_clipmap* ch, int32_t n, const int32_t* results, int32_t counts[2]) {
    Clipmap* c = reinterpret_cast<Clipmap*>(ch);
    if (!c || !counts || n < 0 || (n > 0 && !results)) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if (n != (int32_t)c->lists[0].size()) return LVN_ERR_INVALID;
    c->seam_reqs.clear();
    c->seam_srcs.clear();
    // Clipmap::update, src/clipmap.cpp:1254-1287: a chunk with a mesh or seam nodes becomes
    // active; one with neither (also one whose meshing failed: ConstructClipmapNodeData
    // leaves it without data) becomes empty, with every node below it
    std::vector<int> constructed, empty;
    for (int k = 0; k < n; ++k) {
        Node& nd = c->nodes[c->lists[0][k]];
        const int status = results[3 * k], tris = results[3 * k + 1], seams = results[3 * k + 2];
        const bool ok = status == LVN_OK;
        if (ok && (tris > 0 || seams > 0)) {
            nd.active = true;
            nd.seams = seams;
            constructed.push_back(c->lists[0][k]);
        } else {
            empty.push_back(c->lists[0][k]);
        }
    }
    for (int e : empty) c->propagate_empty(e);
    // the seams to regenerate (src/clipmap.cpp:1306-1324): every active node found under
    // the node's eight neighbour regions on its negative sides, the node included, in
    // first-found order
    std::vector<int> hosts, found;
    std::unordered_set<int> seen;
    for (int m : constructed) {
        for (int i = 0; i < 8; ++i) {
            int32_t nmin[3];
            for (int a = 0; a < 3; ++a) nmin[a] = c->nodes[m].min[a] - child_offset(i, a) * c->nodes[m].size;
            if (!c->in_world(nmin)) continue;
            found.clear();
            c->active_under(nmin, c->nodes[m].size, found);
            for (int h : found)
                if (seen.insert(h).second) hosts.push_back(h);
        }
    }
    // each host's sources as GenerateClipmapSeamMesh gathers them (src/clipmap.cpp:579-597):
    // for neighbour i at host.min + host.size * CHILD_MIN_OFFSETS[i], the active nodes
    // under that region (with seam nodes), their lists where they lie (count = -1)
    for (int h : hosts) {
        const Node& host = c->nodes[h];
        lvn_seam_req q;
        std::memset(&q, 0, sizeof(q));
        for (int a = 0; a < 3; ++a) q.host_min[a] = host.min[a];
        q.host_size = host.size;
        q.colour[0] = q.colour[1] = q.colour[2] = 1.f;
        q.first_source = (int32_t)c->seam_srcs.size();
        for (int i = 0; i < 8; ++i) {
            int32_t nmin[3];
            for (int a = 0; a < 3; ++a) nmin[a] = host.min[a] + child_offset(i, a) * host.size;
            if (!c->in_world(nmin)) continue;
            found.clear();
            c->active_under(nmin, host.size, found);
            for (int s : found) {
                const Node& src = c->nodes[s];
                if (src.seams <= 0) continue;
                lvn_seam_source d;
                std::memset(&d, 0, sizeof(d));
                for (int a = 0; a < 3; ++a) d.min[a] = src.min[a];
                d.size = src.size;
                d.neighbour = i;
                d.count = -1;
                d.nodes = nullptr;
                c->seam_srcs.push_back(d);
            }
        }
        q.n_sources = (int32_t)c->seam_srcs.size() - q.first_source;
        c->seam_reqs.push_back(q);
    }
    counts[0] = (int32_t)c->seam_reqs.size();
    counts[1] = (int32_t)c->seam_srcs.size();
    return LVN_OK;
}

int lvn_clipmap_seams(lvn_clipmap* ch, lvn_seam_req* reqs, int32_t req_cap, lvn_seam_source* sources,
                      int32_t src_cap) {
    Clipmap* c = reinterpret_cast<Clipmap*>(ch);
    if (!c || req_cap < 0 || src_cap < 0 || (req_cap > 0 && !reqs) || (src_cap > 0 && !sources))
        return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    if ((int32_t)c->seam_reqs.size() > req_cap || (int32_t)c->seam_srcs.size() > src_cap) return LVN_ERR_CAPACITY;
    std::copy(c->seam_reqs.begin(), c->seam_reqs.end(), reqs);
    std::copy(c->seam_srcs.begin(), c->seam_srcs.end(), sources);
    return LVN_OK;
}

int lvn_clipmap_edit(lvn_clipmap* ch, const int32_t aabb_min[3], const int32_t aabb_max[3], int32_t* n) {
    Clipmap* c = reinterpret_cast<Clipmap*>(ch);
    if (!c || !aabb_min || !aabb_max || !n) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(c->mu);
    // processCSGOperationsImpl, src/clipmap.cpp:1655-1705: every node the op's bounds
    // overlap is invalidated (released and, while selected, meshed again by the next
    // update) and loses its empty mark
    c->lists[LVN_CLIPMAP_TOUCHED].clear();
    for (int r : c->roots) c->touched(r, aabb_min, aabb_max, c->lists[LVN_CLIPMAP_TOUCHED]);
    for (int t : c->lists[LVN_CLIPMAP_TOUCHED]) {
        c->nodes[t].invalidated = true;
        c->nodes[t].empty = false;
    }
    *n = (int32_t)c->lists[LVN_CLIPMAP_TOUCHED].size();
    return LVN_OK;
}

}  // extern "C"
