// lvn_server.hip — serving form: continuous batching over device-resident chunk slots.
//
// The batch arena's slots hold each served chunk's field and mesh in HBM; the
// number of slots is what the HBM budget admits.  A scheduler thread takes
// submitted requests, answers cache hits at once, gives each miss a free slot
// (evicting the least-recently-used finished chunk), and launches batches of
// misses on one of several lanes (stream + descriptors + counters + work
// queues) as soon as a lane is free, so batches overlap on the device and new
// requests join the next batch instead of waiting for a queue to drain.

#include "lvn_kernels.h"
#include "lvn_runtime.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace lvn {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    int64_t ticket;
    lvn_request req;
    double submit;
};

struct Job {
    int64_t ticket;
    Key key;
    double submit;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    ChunkDesc* d_chunks = nullptr;
    ChunkDesc* h_chunks = nullptr;
    Counters* d_counters = nullptr;
    Counters* h_counters = nullptr;
    int* d_slot_map = nullptr;
    int* h_slot_map = nullptr;
    int* d_queue = nullptr;
    int* d_item_prefix = nullptr;
    DevOp* d_ops = nullptr;
    DevOp* h_ops = nullptr;
    int4* d_boxes = nullptr;  // per op: AABB min, AABB max
    int4* h_boxes = nullptr;
    int ops_cap = 0;
    int* d_op_index = nullptr;  // per chunk of the batch: its overlapping ops (k_op_index)
    int64_t idx_cap = 0;
    bool busy = false;
    std::vector<Job> jobs;
    std::vector<Key> keys;
    std::vector<uint64_t> versions;
};

struct Entry {
    int slot = -1;
    bool valid = false;       // the slot holds this chunk's current mesh
    bool in_flight = false;   // a lane is meshing it
    uint64_t version = 0;     // bumped by every edit overlapping the chunk
    uint64_t last_use = 0;
    int nv = 0, nt = 0, ns = 0, status = 0;
    std::vector<Job> waiters; // requests for this chunk that arrived while it was in flight
    std::vector<int64_t> tickets;  // completed mesh requests served by the current mesh
};

}  // namespace

struct Server {
    Ctx arena;
    int max_batch = 64;
    std::vector<Lane> lanes;

    std::mutex mu;                 // incoming, completed, outstanding, stop
    std::condition_variable cv;
    std::deque<Pending> incoming;
    std::deque<lvn_result> completed;
    int64_t next_ticket = 1;
    int64_t outstanding = 0;
    bool stop = false;
    int error = LVN_OK;

    std::mutex tuning_mu;          // launch shapes (lvn_server_set_tuning), read at each launch
    lvn_tuning tuning = {0, kDefaultPersistPerCu, kDefaultRowBlocks, kDefaultVoxelBlocks, 0, kDefaultDensityPerCu, 0, {0}};

    std::mutex cache_mu;           // cache, ticket_key and stats (read by lvn_server_get_mesh / _stats)
    std::map<Key, Entry> cache;
    // completed mesh request -> its chunk, while the chunk's mesh is the one it was
    // served; dropped when that mesh is evicted or invalidated, so the map holds at most
    // the requests served by resident meshes
    std::unordered_map<int64_t, Key> ticket_key;

    // scheduler-thread state
    std::vector<int> free_slots;
    std::deque<Job> waiting;       // misses not yet admitted
    std::vector<lvn_csg_op> ops;
    std::vector<std::array<int32_t, 6>> aabbs;
    uint64_t tick = 0;
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    std::thread worker;

    int init_lane(Lane& L);
    void run();
    void finish(int64_t ticket, int kind, int status, const Entry* e, bool hit, double submit);
    void complete_mesh(Entry& e, const Key& key, const Job& j, bool hit);  // with cache_mu held
    void forget_tickets(Entry& e);                                          // with cache_mu held
    void admit_and_launch();
    void poll_lanes();
    void handle(const Pending& p);
    int launch(Lane& L, const std::vector<Job>& batch, const std::vector<int>& slots);
    ~Server();
};

// The stored CSG ops of each chunk of a lane's batch, in stored order: op k is listed for
// chunk i when its AABB overlaps the chunk's (aabb_overlaps, src/aabb.h:26-35: touching
// counts).  One workgroup per chunk, ballots compact the hits in order; chunk i's list
// goes to index[i * nops ..], its start and length into the chunk's descriptor.
constexpr int kOpIndexThreads = 256;
__global__ __launch_bounds__(kOpIndexThreads) void k_op_index(ChunkDesc* chunks, const int4* boxes, int nops,
                                                              int* index) {
    __shared__ int wsum[kOpIndexThreads / 64];
    __shared__ int carry;
    ChunkDesc& cd = chunks[blockIdx.x];
    const int m0 = cd.min[0], m1 = cd.min[1], m2 = cd.min[2], size = cd.size;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int* out = index + (size_t)blockIdx.x * nops;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < nops; b += kOpIndexThreads) {
        const int k = b + threadIdx.x;
        bool hit = false;
        if (k < nops) {
            const int4 lo = boxes[2 * k], hi = boxes[2 * k + 1];
            hit = !(m0 + size < lo.x || m1 + size < lo.y || m2 + size < lo.z || m0 > hi.x || m1 > hi.y || m2 > hi.z);
        }
        const uint64_t mask = __ballot(hit);
        if (lane == 0) wsum[w] = __popcll(mask);
        __syncthreads();
        int base = carry, total = 0;
#pragma unroll
        for (int i = 0; i < kOpIndexThreads / 64; ++i) {
            base += i < w ? wsum[i] : 0;
            total += wsum[i];
        }
        if (hit) out[base + __popcll(mask & ((1ull << lane) - 1))] = k;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cd.op_start = blockIdx.x * nops;
        cd.op_count = carry;
    }
}

#define LVN_TRY(call)                                    \
    do {                                                 \
        if ((call) != hipSuccess) return LVN_ERR_DEVICE; \
    } while (0)

int Server::init_lane(Lane& L) {
    const int B = max_batch;
    LVN_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    LVN_TRY(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    LVN_TRY(hipMalloc(&L.d_chunks, sizeof(ChunkDesc) * B));
    LVN_TRY(hipHostMalloc(&L.h_chunks, sizeof(ChunkDesc) * B, hipHostMallocDefault));
    LVN_TRY(hipMalloc(&L.d_counters, sizeof(Counters) * B));
    LVN_TRY(hipHostMalloc(&L.h_counters, sizeof(Counters) * B, hipHostMallocDefault));
    LVN_TRY(hipMalloc(&L.d_slot_map, sizeof(int) * B));
    LVN_TRY(hipHostMalloc(&L.h_slot_map, sizeof(int) * B, hipHostMallocDefault));
    LVN_TRY(hipMalloc(&L.d_queue, sizeof(int) * kNumQueues));
    LVN_TRY(hipMalloc(&L.d_item_prefix, sizeof(int) * (B + 1)));
    // op lists sized up front (1,024 stored ops, op lists of a full batch with up to 256
    // stored ops) so that a lane's first batches allocate nothing; more ops grow them
    L.ops_cap = 1024;
    L.idx_cap = std::max<int64_t>(4096, 256 * (int64_t)B);
    LVN_TRY(hipMalloc(&L.d_ops, sizeof(DevOp) * L.ops_cap));
    LVN_TRY(hipHostMalloc(&L.h_ops, sizeof(DevOp) * L.ops_cap, hipHostMallocDefault));
    LVN_TRY(hipMalloc(&L.d_boxes, sizeof(int4) * 2 * L.ops_cap));
    LVN_TRY(hipHostMalloc(&L.h_boxes, sizeof(int4) * 2 * L.ops_cap, hipHostMallocDefault));
    LVN_TRY(hipMalloc(&L.d_op_index, sizeof(int) * L.idx_cap));
    return LVN_OK;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mu);
        stop = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
    for (Lane& L : lanes) {
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.d_chunks) (void)hipFree(L.d_chunks);
        if (L.h_chunks) (void)hipHostFree(L.h_chunks);
        if (L.d_counters) (void)hipFree(L.d_counters);
        if (L.h_counters) (void)hipHostFree(L.h_counters);
        if (L.d_slot_map) (void)hipFree(L.d_slot_map);
        if (L.h_slot_map) (void)hipHostFree(L.h_slot_map);
        if (L.d_queue) (void)hipFree(L.d_queue);
        if (L.d_item_prefix) (void)hipFree(L.d_item_prefix);
        if (L.d_ops) (void)hipFree(L.d_ops);
        if (L.h_ops) (void)hipHostFree(L.h_ops);
        if (L.d_boxes) (void)hipFree(L.d_boxes);
        if (L.h_boxes) (void)hipHostFree(L.h_boxes);
        if (L.d_op_index) (void)hipFree(L.d_op_index);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
}

void Server::finish(int64_t ticket, int kind, int status, const Entry* e, bool hit, double submit) {
    lvn_result r;
    std::memset(&r, 0, sizeof(r));
    r.ticket = ticket;
    r.kind = kind;
    r.status = status;
    if (e) {
        r.nv = e->nv;
        r.nt = e->nt;
        r.ns = e->ns;
    }
    r.cache_hit = hit ? 1 : 0;
    r.submit_s = submit;
    r.complete_s = now_s();
    std::lock_guard<std::mutex> lock(mu);
    completed.push_back(r);
    --outstanding;
    cv.notify_all();
}

void Server::complete_mesh(Entry& e, const Key& key, const Job& j, bool hit) {
    ticket_key[j.ticket] = key;
    e.tickets.push_back(j.ticket);
    finish(j.ticket, LVN_REQ_MESH, e.status, &e, hit, j.submit);
}

void Server::forget_tickets(Entry& e) {
    for (int64_t t : e.tickets) ticket_key.erase(t);
    e.tickets.clear();
}

// An edit stores the op and invalidates every chunk whose AABB it overlaps; a mesh
// request is a cache hit, joins an in-flight mesh of the same chunk, or waits.
void Server::handle(const Pending& p) {
    if (p.req.kind == LVN_REQ_EDIT) {
        ops.push_back(p.req.op);
        aabbs.push_back({p.req.aabb_min[0], p.req.aabb_min[1], p.req.aabb_min[2], p.req.aabb_max[0],
                         p.req.aabb_max[1], p.req.aabb_max[2]});
        {
            std::lock_guard<std::mutex> lock(cache_mu);
            for (auto& kv : cache) {
                const int32_t mn[3] = {std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)};
                if (!aabb_overlaps(mn, std::get<3>(kv.first), aabbs.back())) continue;
                Entry& e = kv.second;
                ++e.version;
                if (e.valid) {
                    e.valid = false;
                    forget_tickets(e);
                    if (!e.in_flight) {
                        free_slots.push_back(e.slot);
                        e.slot = -1;
                    }
                }
            }
        }
        finish(p.ticket, LVN_REQ_EDIT, LVN_OK, nullptr, false, p.submit);
        return;
    }
    const Key key = make_key(p.req.min, p.req.size);
    {
        std::lock_guard<std::mutex> lock(cache_mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            Entry& e = it->second;
            if (e.valid && !e.in_flight) {
                e.last_use = ++tick;
                ++stats[3];
                complete_mesh(e, key, Job{p.ticket, key, p.submit}, true);
                return;
            }
            if (e.in_flight) {
                e.waiters.push_back(Job{p.ticket, key, p.submit});
                return;
            }
        }
    }
    waiting.push_back(Job{p.ticket, key, p.submit});
}

int Server::launch(Lane& L, const std::vector<Job>& batch, const std::vector<int>& slots) {
    const int n = (int)batch.size();
    // the tables stay allocated while this server lives, but after lvn_shutdown no new
    // work starts; holding the global lock orders the enqueue against reseeding
    std::lock_guard<std::mutex> glock(g_state.mu);
    if (!g_state.init) return LVN_ERR_STATE;
    const int nops = (int)ops.size();
    for (int i = 0; i < n; ++i) {
        const int32_t mn[3] = {std::get<0>(batch[i].key), std::get<1>(batch[i].key), std::get<2>(batch[i].key)};
        ChunkDesc& cd = L.h_chunks[i];
        fill_chunk_desc(cd, mn, std::get<3>(batch[i].key), arena.N);
        cd.op_start = cd.op_count = 0;  // set by k_op_index
        L.h_slot_map[i] = slots[i];
    }
    if (nops > L.ops_cap) {
        LVN_TRY(hipFree(L.d_ops));
        LVN_TRY(hipHostFree(L.h_ops));
        LVN_TRY(hipFree(L.d_boxes));
        LVN_TRY(hipHostFree(L.h_boxes));
        L.d_ops = nullptr;
        L.h_ops = nullptr;
        L.d_boxes = L.h_boxes = nullptr;
        L.ops_cap = std::max(2 * L.ops_cap, nops);
        LVN_TRY(hipMalloc(&L.d_ops, sizeof(DevOp) * L.ops_cap));
        LVN_TRY(hipHostMalloc(&L.h_ops, sizeof(DevOp) * L.ops_cap, hipHostMallocDefault));
        LVN_TRY(hipMalloc(&L.d_boxes, sizeof(int4) * 2 * L.ops_cap));
        LVN_TRY(hipHostMalloc(&L.h_boxes, sizeof(int4) * 2 * L.ops_cap, hipHostMallocDefault));
    }
    if ((int64_t)n * nops > L.idx_cap) {
        LVN_TRY(hipFree(L.d_op_index));
        L.d_op_index = nullptr;
        L.idx_cap = std::max(2 * L.idx_cap, (int64_t)n * nops);
        LVN_TRY(hipMalloc(&L.d_op_index, sizeof(int) * L.idx_cap));
    }
    for (int k = 0; k < nops; ++k) {
        L.h_ops[k] = make_dev_op(ops[k]);
        const auto& bb = aabbs[k];
        L.h_boxes[2 * k] = make_int4(bb[0], bb[1], bb[2], 0);
        L.h_boxes[2 * k + 1] = make_int4(bb[3], bb[4], bb[5], 0);
    }
    if (nops) {
        LVN_TRY(hipMemcpyAsync(L.d_ops, L.h_ops, sizeof(DevOp) * nops, hipMemcpyHostToDevice, L.stream));
        LVN_TRY(hipMemcpyAsync(L.d_boxes, L.h_boxes, sizeof(int4) * 2 * nops, hipMemcpyHostToDevice, L.stream));
    }
    LVN_TRY(hipMemcpyAsync(L.d_chunks, L.h_chunks, sizeof(ChunkDesc) * n, hipMemcpyHostToDevice, L.stream));
    LVN_TRY(hipMemcpyAsync(L.d_slot_map, L.h_slot_map, sizeof(int) * n, hipMemcpyHostToDevice, L.stream));
    if (nops) {
        hipLaunchKernelGGL(k_op_index, dim3(n), dim3(kOpIndexThreads), 0, L.stream, L.d_chunks,
                           static_cast<const int4*>(L.d_boxes), nops, L.d_op_index);
        LVN_TRY(hipGetLastError());
    }
    KArgs a = arena.kargs();
    a.chunks = L.d_chunks;
    a.counters = L.d_counters;
    a.slot_map = L.d_slot_map;
    a.queue = L.d_queue;
    a.item_prefix = L.d_item_prefix;
    a.ops = L.d_ops;
    a.op_index = L.d_op_index;
    lvn_tuning t;
    {
        std::lock_guard<std::mutex> tl(tuning_mu);
        t = tuning;
    }
    // served chunks export meshes only: a field without the cave term leaves its density
    // unstored (KArgs::dens_cols)
    a.dens_cols = a.F.kind != LVN_FIELD_TERRAIN_CAVES && !a.fp16 ? 1 : 0;
    a.mat_free = 1;  // and chunks without brushes their materials
    // zero fields: the automatic choice by batch size
    a.ysplit = t.density_ysplit ? t.density_ysplit : n >= 32 ? 1 : 6;
    a.height_lines = t.height_lines ? (t.height_lines == 2 ? 1 : 0) : (n >= kHeightLinesMinChunks ? 1 : 0);
    a.persist_per_cu = t.persistent_per_cu;
    a.density_per_cu = t.density_per_cu ? t.density_per_cu : kDefaultDensityPerCu;
    a.row_blocks = t.row_blocks;
    a.voxel_blocks = t.voxel_blocks;
    a.cross_lanes = t.cross_lanes;
    launch_pipeline(a, n, L.stream);
    LVN_TRY(hipGetLastError());
    LVN_TRY(hipMemcpyAsync(L.h_counters, L.d_counters, sizeof(Counters) * n, hipMemcpyDeviceToHost, L.stream));
    LVN_TRY(hipEventRecord(L.done, L.stream));
    L.busy = true;
    L.jobs = batch;
    std::lock_guard<std::mutex> lock(cache_mu);
    ++stats[1];
    stats[2] += n;
    stats[7] = std::max<int64_t>(stats[7], n);
    return LVN_OK;
}

// Give waiting misses slots (free, else the least-recently-used finished chunk) and
// launch them on free lanes, oldest first.
void Server::admit_and_launch() {
    for (Lane& L : lanes) {
        if (L.busy || waiting.empty()) continue;
        std::vector<Job> batch;
        std::vector<int> slots;
        {
            std::lock_guard<std::mutex> lock(cache_mu);
            while (!waiting.empty() && (int)batch.size() < max_batch) {
                Job j = waiting.front();
                Entry& e = cache[j.key];
                if (e.in_flight) {  // an earlier request in this pass took it
                    e.waiters.push_back(j);
                    waiting.pop_front();
                    continue;
                }
                if (e.valid) {  // meshed while this request waited
                    e.last_use = ++tick;
                    ++stats[3];
                    complete_mesh(e, j.key, j, true);
                    waiting.pop_front();
                    continue;
                }
                int slot = -1;
                if (e.slot >= 0) {
                    slot = e.slot;
                } else if (!free_slots.empty()) {
                    slot = free_slots.back();
                    free_slots.pop_back();
                } else {
                    Entry* victim = nullptr;
                    for (auto& kv : cache)
                        if (kv.second.valid && !kv.second.in_flight && kv.second.slot >= 0 &&
                            (!victim || kv.second.last_use < victim->last_use))
                            victim = &kv.second;
                    if (!victim) break;  // every slot is in flight: admit later
                    slot = victim->slot;
                    victim->slot = -1;
                    victim->valid = false;
                    forget_tickets(*victim);
                    ++stats[4];
                }
                e.slot = slot;
                e.in_flight = true;
                batch.push_back(j);
                slots.push_back(slot);
                waiting.pop_front();
            }
            L.versions.clear();
            for (const Job& j : batch) L.versions.push_back(cache[j.key].version);
        }
        if (batch.empty()) continue;
        const int r = launch(L, batch, slots);
        if (r != LVN_OK) {
            std::lock_guard<std::mutex> lock(mu);
            error = r;
        }
    }
}

void Server::poll_lanes() {
    for (Lane& L : lanes) {
        if (!L.busy) continue;
        if (hipEventQuery(L.done) != hipSuccess) continue;
        std::vector<Job> requeue;
        {
            std::lock_guard<std::mutex> lock(cache_mu);
            for (size_t i = 0; i < L.jobs.size(); ++i) {
                const Job& j = L.jobs[i];
                Entry& e = cache[j.key];
                e.in_flight = false;
                if (e.version != L.versions[i]) {
                    // an edit arrived while this chunk was meshed: mesh it again
                    ++stats[5];
                    free_slots.push_back(e.slot);
                    e.slot = -1;
                    requeue.push_back(j);
                    for (const Job& w : e.waiters) requeue.push_back(w);
                    e.waiters.clear();
                    continue;
                }
                const Counters& k = L.h_counters[i];
                e.valid = true;
                e.status = k.status ? LVN_ERR_OVERFLOW : LVN_OK;
                e.nv = k.edges > 0 && k.quads > 0 ? k.voxels : 0;
                e.nt = k.edges > 0 ? 2 * k.quads : 0;
                e.ns = k.edges > 0 ? k.seams : 0;
                e.last_use = ++tick;
                complete_mesh(e, j.key, j, false);
                for (const Job& w : e.waiters) complete_mesh(e, j.key, w, true);
                e.waiters.clear();
            }
            stats[6] = 0;
            for (auto& kv : cache) stats[6] += kv.second.valid ? 1 : 0;
        }
        for (auto it = requeue.rbegin(); it != requeue.rend(); ++it) waiting.push_front(*it);
        L.busy = false;
        L.jobs.clear();
    }
}

void Server::run() {
    for (;;) {
        std::deque<Pending> batch;
        {
            std::unique_lock<std::mutex> lock(mu);
            const bool busy = std::any_of(lanes.begin(), lanes.end(), [](const Lane& L) { return L.busy; });
            if (incoming.empty() && waiting.empty() && !busy) {
                if (stop) return;
                cv.wait_for(lock, std::chrono::milliseconds(2));
                continue;
            }
            batch.swap(incoming);
        }
        for (const Pending& p : batch) handle(p);
        poll_lanes();
        admit_and_launch();
        // while batches are in flight, poll their events without sleeping: a sleep's
        // granularity (tens of microseconds) would hold finished lanes idle
        const bool busy = std::any_of(lanes.begin(), lanes.end(), [](const Lane& L) { return L.busy; });
        if (busy) std::this_thread::yield();
    }
}

}  // namespace lvn

using namespace lvn;

extern "C" {

int lvn_server_create(int32_t voxels_per_chunk, uint64_t hbm_budget, int32_t max_batch, int32_t lanes,
                      int32_t edge_cap, int32_t voxel_cap, uint32_t flags, lvn_server** out) {
    if (!out) return LVN_ERR_INVALID;
    *out = nullptr;
    if (!g_state.init) return LVN_ERR_STATE;
    const int n = voxels_per_chunk;
    if (n < 4 || n > 64 || (n & (n - 1)) || max_batch < 1 || lanes < 1 || lanes > 8 || edge_cap < 0 ||
        voxel_cap < 0)
        return LVN_ERR_INVALID;
    const SlotLayout L = make_layout(make_geometry(n), make_caps(make_geometry(n), edge_cap, voxel_cap));
    const uint64_t nslots = hbm_budget / L.slot_bytes;
    if (nslots < (uint64_t)max_batch) return LVN_ERR_NOMEM;
    Server* s = new Server;
    s->max_batch = max_batch;
    int r = s->arena.allocate(n, flags, (int)std::min<uint64_t>(nslots, 1 << 20), edge_cap, voxel_cap);
    if (!r) {
        s->lanes.resize((size_t)lanes);
        for (Lane& ln : s->lanes)
            if (!r) r = s->init_lane(ln);
    }
    if (r) {
        delete s;
        return r == LVN_ERR_DEVICE ? LVN_ERR_NOMEM : r;
    }
    for (int i = s->arena.slots - 1; i >= 0; --i) s->free_slots.push_back(i);
    s->stats[0] = s->arena.slots;
    s->ops = g_state.ops;
    s->aabbs = g_state.aabbs;
    s->worker = std::thread([s] { s->run(); });
    *out = reinterpret_cast<lvn_server*>(s);
    return LVN_OK;
}

void lvn_server_destroy(lvn_server* sh) { delete reinterpret_cast<Server*>(sh); }

static bool valid_request(const Server* s, const lvn_request* req) {
    if (req->kind == LVN_REQ_MESH)
        return req->size >= s->arena.N * kLeafSizeScale && req->size % (s->arena.N * kLeafSizeScale) == 0;
    if (req->kind == LVN_REQ_EDIT) {
        const lvn_csg_op& o = req->op;
        return o.material >= 0 && o.material <= 255 && o.type >= 0 && o.type <= 1 && o.brush >= 0 && o.brush <= 1;
    }
    return false;
}

int lvn_server_submit(lvn_server* sh, const lvn_request* req, int64_t* ticket) {
    return lvn_server_submit_many(sh, 1, req, ticket);
}

int lvn_server_submit_many(lvn_server* sh, int32_t n, const lvn_request* reqs, int64_t* tickets) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || n < 0 || (n > 0 && (!reqs || !tickets))) return LVN_ERR_INVALID;
    for (int32_t i = 0; i < n; ++i)
        if (!valid_request(s, &reqs[i])) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->error) return s->error;
    const double t = now_s();
    for (int32_t i = 0; i < n; ++i) {
        tickets[i] = s->next_ticket++;
        s->incoming.push_back(Pending{tickets[i], reqs[i], t});
        ++s->outstanding;
    }
    s->cv.notify_all();
    return LVN_OK;
}

int lvn_server_poll(lvn_server* sh, lvn_result* out, int32_t cap, int32_t* n) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !n || cap < 0 || (cap > 0 && !out)) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->mu);
    int k = 0;
    while (k < cap && !s->completed.empty()) {
        out[k++] = s->completed.front();
        s->completed.pop_front();
    }
    *n = k;
    return s->error;
}

int lvn_server_drain(lvn_server* sh) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s) return LVN_ERR_INVALID;
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv.wait(lock, [s] { return s->outstanding == 0 || s->error != LVN_OK; });
    return s->error;
}

int lvn_server_get_mesh(lvn_server* sh, int64_t ticket, lvn_vertex* v, int32_t v_cap, int32_t* nv, int32_t* tri,
                        int32_t tri_cap, int32_t* nt, lvn_seam_node* sn, int32_t s_cap, int32_t* ns) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !nv || !nt || !ns) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->cache_mu);
    auto tk = s->ticket_key.find(ticket);
    if (tk == s->ticket_key.end()) {
        // not completed yet, an edit, or its chunk's mesh has since been evicted or
        // invalidated: a state error for a ticket this server issued
        std::lock_guard<std::mutex> lk(s->mu);
        return ticket >= 1 && ticket < s->next_ticket ? LVN_ERR_STATE : LVN_ERR_INVALID;
    }
    auto it = s->cache.find(tk->second);
    if (it == s->cache.end() || !it->second.valid || it->second.in_flight || it->second.slot < 0) return LVN_ERR_STATE;
    const Entry& e = it->second;
    *nv = e.nv;
    *nt = e.nt;
    *ns = e.ns;
    if (e.status) return e.status;
    if (e.nv > v_cap || e.nt > tri_cap || e.ns > s_cap) return LVN_ERR_CAPACITY;
    const Ctx& c = s->arena;
    const char* slot = c.d_arena + (uint64_t)e.slot * c.L.slot_bytes;
    if (e.nv && v && hipMemcpy(v, slot + c.L.verts, sizeof(lvn_vertex) * (size_t)e.nv, hipMemcpyDeviceToHost))
        return LVN_ERR_DEVICE;
    if (e.nt && tri && hipMemcpy(tri, slot + c.L.tris, 12 * (size_t)e.nt, hipMemcpyDeviceToHost)) return LVN_ERR_DEVICE;
    if (e.ns && sn && hipMemcpy(sn, slot + c.L.seams, sizeof(lvn_seam_node) * (size_t)e.ns, hipMemcpyDeviceToHost))
        return LVN_ERR_DEVICE;
    return LVN_OK;
}

int lvn_server_set_tuning(lvn_server* sh, const lvn_tuning* t) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !t || !tuning_valid(*t, s->arena.G.F)) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->tuning_mu);
    s->tuning = *t;
    return LVN_OK;
}

int lvn_server_stats(lvn_server* sh, int64_t* stats) {
    Server* s = reinterpret_cast<Server*>(sh);
    if (!s || !stats) return LVN_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->cache_mu);
    for (int i = 0; i < 8; ++i) stats[i] = s->stats[i];
    return LVN_OK;
}

}  // extern "C"
