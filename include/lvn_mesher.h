/*
 * lvn_mesher.h — C ABI of the MI355X (gfx950) dual-contouring chunk mesher.
 *
 * Drop-in replacement for leven's OpenCL compute module.  Every entry point
 * names the reference interface it replaces (paths relative to the leven
 * source tree, leven/src/...).  Plain pointers and sizes only; no HIP or torch
 * types cross this boundary.  All functions return LVN_OK (0) or a negative
 * LVN_ERR_* code, like the reference's CL_SUCCESS / negative CL error
 * convention (src/compute.h:12, src/compute_local.h:194-204).
 *
 * Threading: one lvn_ctx per host thread.  The global state set by lvn_init /
 * lvn_set_noise_seed / lvn_store_csg_op is shared by all contexts and must not
 * be changed while another thread uses a context (the reference is likewise
 * single-threaded: src/volume.cpp:68-113).
 */
#ifndef LVN_MESHER_H
#define LVN_MESHER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes -------------------------------------------------------- */
#define LVN_OK              0
#define LVN_ERR_INVALID    (-1)   /* bad argument (null pointer, unsupported size, material > 255) */
#define LVN_ERR_NOMEM      (-2)   /* device allocation failed */
#define LVN_ERR_CAPACITY   (-3)   /* caller buffer too small; the needed counts are still written */
#define LVN_ERR_DEVICE     (-4)   /* HIP runtime error */
#define LVN_ERR_STATE      (-5)   /* lvn_init not called / library shut down */
#define LVN_ERR_OVERFLOW   (-6)   /* a chunk exceeded a per-chunk device capacity */

/* ---- materials (src/volume_materials.h:8-9) ------------------------------ */
#define LVN_MATERIAL_NONE  200
#define LVN_MATERIAL_AIR   201

/* ---- POD types, byte-compatible with the reference ------------------------ */

/* = CSGOperationInfo (src/compute.h:16-24) = CSGOperation (cl/apply_csg_operation.cl:5-14); 48 bytes.
 * Units as the reference caller produces them (src/clipmap.cpp:1628-1632): voxel units. */
typedef struct lvn_csg_op {
    int32_t type;        /* 0 = add, 1 = subtract */
    int32_t brush;       /* 0 = cube (RenderShape_Cube), 1 = sphere (RenderShape_Sphere) */
    int32_t material;    /* 0..255; subtract ops conventionally carry LVN_MATERIAL_AIR */
    float   rotate_y;    /* radians, applied to the cube's xz plane (pR, cl/apply_csg_operation.cl:57-60) */
    float   origin[4];   /* xyz used */
    float   dims[4];     /* cube: half extents xyz; sphere: radius in dims[0] */
} lvn_csg_op;

/* = MeshVertex (src/render_types.h:24-38); 48 bytes. */
typedef struct lvn_vertex {
    float pos[4];        /* world units, w = 1 */
    float normal[4];     /* mean of edge normals (not renormalised), w = 0 */
    float colour[4];     /* LOD colour rgb, a = dominant material */
} lvn_vertex;

/* = MeshTriangle (src/render_types.h:42-58); 12 bytes. */
typedef struct lvn_triangle {
    int32_t idx[3];
} lvn_triangle;

/* = SeamNodeInfo (src/compute.h:26-31); 48 bytes. */
typedef struct lvn_seam_node {
    int32_t local_min_and_material[4];  /* voxel xyz in the chunk, w = (dominant << 8) | corner mask */
    float   pos[4];
    float   normal[4];
} lvn_seam_node;

/* Density function selector.  The reference compiles its density function in
 * (DensityFunc, cl/noise.cl:225-268); here it is a runtime descriptor. */
#define LVN_FIELD_HEIGHTFIELD     0   /* d = y - amp * fbm2(x, z) */
#define LVN_FIELD_TERRAIN_CAVES   1   /* d = max(y - amp * fbm2(x, z), fbm3(x, y, z) - threshold) */
#define LVN_FIELD_LEVEN_TERRAIN   2   /* leven's own: d = y - 900 * Terrain(x, z) (cl/noise.cl:205-223) */

typedef struct lvn_field_desc {
    int32_t kind;
    int32_t height_octaves;     /* BasicFractal (cl/noise.cl:8-36) over snoise2 */
    float   height_frequency;
    float   height_lacunarity;
    float   height_gain;
    float   height_amplitude;   /* voxels */
    int32_t cave_octaves;       /* the same fBm over snoise3 (cl/simplex.cl:159-230) */
    float   cave_frequency;
    float   cave_lacunarity;
    float   cave_gain;
    float   cave_threshold;
    int32_t reserved;
} lvn_field_desc;

/* Context flags. */
#define LVN_CTX_FP16_DENSITY  1u   /* store the density field as sign-preserving FP16 */
#define LVN_CTX_PRECISE_QEF   2u   /* compensated FP32 QEF sums + FP64 re-solve of ill-conditioned QEFs */

typedef struct lvn_ctx lvn_ctx;

/* ---- global state (src/compute.h:35-40) ---------------------------------- */

/* Replaces Compute_Initialise(noiseSeed, defaultMaterial, numCSGBrushes), src/compute.cpp:195-234.
 * field == NULL selects the 64^3 heightfield of the plain CPU workload
 * (6 octaves, f = 1/128, lacunarity 2, gain 0.5, amplitude 24). */
int lvn_init(const lvn_field_desc* field, int32_t noise_seed, uint32_t default_material);
/* Replaces Compute_Shutdown, src/compute.cpp:238-241. */
int lvn_shutdown(void);
/* Replaces Compute_SetNoiseSeed, src/compute_density_field.cpp:129-134. */
int lvn_set_noise_seed(int32_t noise_seed);
/* Replaces Compute_StoreCSGOperation(op, AABB), src/compute_density_field.cpp:312-317.
 * aabb_min/aabb_max are world units (AABB, src/aabb.h). */
int lvn_store_csg_op(const lvn_csg_op* op, const int32_t aabb_min[3], const int32_t aabb_max[3]);
/* Replaces Compute_ClearCSGOperations, src/compute_density_field.cpp:321-326. */
int lvn_clear_csg_ops(void);
/* Copies the 256x256 permutation table in use (row = texel y) into out[65536]. */
int lvn_get_perm_table(uint8_t* out);

/* ---- mesh generation context (src/compute.h:46-77) ------------------------ */

/* Replaces Compute_MeshGenContext::create(voxelsPerChunk), src/compute.cpp:605-610.
 * voxels_per_chunk: power of two in [4, 64]. */
int  lvn_ctx_create(int32_t voxels_per_chunk, uint32_t flags, lvn_ctx** out);
void lvn_ctx_destroy(lvn_ctx* ctx);
/* Replaces Compute_MeshGenContext::voxelsPerChunk, src/compute.cpp:612-615. */
int  lvn_ctx_voxels_per_chunk(const lvn_ctx* ctx);

/* Replaces Compute_MeshGenContext::applyCSGOperations, src/compute.cpp:617-623
 * (-> Compute_ApplyCSGOperations, src/compute_csg.cpp:224-242).  min/size in world units. */
int lvn_apply_csg(lvn_ctx* ctx, const lvn_csg_op* ops, int32_t n, const int32_t min[3], int32_t size);
/* Replaces Compute_MeshGenContext::freeChunkOctree, src/compute_octree.cpp:379-387. */
int lvn_free_chunk(lvn_ctx* ctx, const int32_t min[3], int32_t size);
/* Replaces Compute_MeshGenContext::isChunkEmpty, src/compute_density_field.cpp:278-299.
 * *out = 1 when the chunk's field has at least one sign-changing edge.  (The reference
 * returns that same value under the name isEmpty; the C++ wrapper keeps its name.) */
int lvn_chunk_has_surface(lvn_ctx* ctx, const int32_t min[3], int32_t size, int32_t* out);
/* Replaces Compute_MeshGenContext::generateChunkMesh, src/compute_octree.cpp:351-375.
 * Writes up to the caps; returns LVN_ERR_CAPACITY (with the needed counts in
 * *nv / *nt / *ns) when any output does not fit.  tri holds 3 indices per triangle. */
int lvn_generate_chunk_mesh(lvn_ctx* ctx, const int32_t min[3], int32_t size,
                            lvn_vertex* v, int32_t v_cap, int32_t* nv,
                            int32_t* tri, int32_t tri_cap, int32_t* nt,
                            lvn_seam_node* s, int32_t s_cap, int32_t* ns);

/* ---- seam stitching (GenerateClipmapSeamMesh, src/clipmap.cpp:573-611) ---------- */

/* One clipmap node's seam nodes offered to a host node's seam mesh: the node's min
 * and size in world units, its list as lvn_generate_chunk_mesh returned it, and the
 * candidate neighbour it was found under: neighbour i of the host lies at
 * host_min + host_size * CHILD_MIN_OFFSETS[i] (src/volume_constants.h:24-35), i.e.
 * x = bit 2, y = bit 1, z = bit 0 of i; 0 is the host node itself. */
typedef struct lvn_seam_source {
    int32_t min[3];
    int32_t size;
    int32_t neighbour;           /* 0..7 */
    int32_t count;               /* -1 with nodes == NULL: the seam nodes of the context's cached
                                    chunk at (min, size), read where they lie in device memory */
    const lvn_seam_node* nodes;  /* host memory */
} lvn_seam_source;

/* Replaces the body of GenerateClipmapSeamMesh (src/clipmap.cpp:573-611): SelectSeamNodes
 * (:546-569, FilterSeamNode :508-542), Octree_ConstructUpwards (src/octree.cpp:90-148)
 * and Octree_GenerateMesh (src/octree.cpp:520-550: GenerateVertexIndices :196-231,
 * ContourCellProc / FaceProc / EdgeProc / ProcessEdge :235-474) over the selected nodes,
 * in a root of size 2 * host_size at host_min.  Vertices: one per selected seam node in
 * the octree's depth-first order, pos.w = 0, normal.w = 0, colour = (colour.rgb,
 * material >> 8).  Triangles: 3 indices each, in the reference's contouring order.
 * The reference discards a seam mesh without triangles; here *nv still counts its
 * vertices.  LVN_ERR_INVALID when two selected nodes overlap or a node is not aligned
 * to its size in the root (the reference asserts); LVN_ERR_CAPACITY with the needed
 * counts when an output does not fit. */
int lvn_generate_seam_mesh(lvn_ctx* ctx, const int32_t host_min[3], int32_t host_size,
                           const lvn_seam_source* sources, int32_t n_sources, const float colour[3],
                           lvn_vertex* v, int32_t v_cap, int32_t* nv,
                           int32_t* tri, int32_t tri_cap, int32_t* nt);

/* Many seams in one launch sequence, one workgroup per seam.  reqs[i] takes sources
 * [first_source, first_source + n_sources) of `sources`; a source list shared by
 * several seams (the same nodes pointer) is uploaded once.  Per-seam results go to
 * outs[i] (status LVN_OK, LVN_ERR_INVALID or LVN_ERR_CAPACITY as above); the call
 * returns the first failing status.  No reference counterpart: the reference stitches
 * one seam per call. */
typedef struct lvn_seam_req {
    int32_t host_min[3];
    int32_t host_size;
    float   colour[4];           /* rgb used */
    int32_t first_source;
    int32_t n_sources;
} lvn_seam_req;

typedef struct lvn_seam_out {
    lvn_vertex* v;
    int32_t*    tri;
    int32_t     v_cap, tri_cap;  /* tri_cap in triangles */
    int32_t     nv, nt;
    int32_t     status;
    int32_t     reserved;
} lvn_seam_out;

int lvn_generate_seam_meshes(lvn_ctx* ctx, int32_t n, const lvn_seam_req* reqs,
                             const lvn_seam_source* sources, lvn_seam_out* outs);

/* Device-resident results: a seam passed with v == NULL and tri == NULL is not copied
 * out (its counts still are); this returns device pointers to seam i's vertices and
 * triangles from the last lvn_generate_seam_meshes / lvn_generate_seam_mesh call on ctx,
 * valid until the next such call or lvn_ctx_destroy.  With cached sources (count -1) a
 * clipmap's seams are stitched without crossing PCIe.  No reference counterpart: for
 * callers that render or post-process on the GPU. */
int lvn_seam_device_mesh(lvn_ctx* ctx, int32_t i, const lvn_vertex** v, const int32_t** tri,
                         int32_t* nv, int32_t* nt);

/* ---- batched form: many chunks per launch, device-resident outputs ---------- */

/* One chunk request: min in world units, size = 256 * 2^LOD (src/clipmap.cpp:378). */
typedef struct lvn_chunk_req {
    int32_t min[3];
    int32_t size;
} lvn_chunk_req;

typedef struct lvn_batch lvn_batch;

/* A batch of up to max_chunks chunk slots in HBM.  edge_cap / voxel_cap bound the
 * compacted edges / active voxels of one chunk (0 = the worst case for the chunk
 * size); a chunk that exceeds them reports LVN_ERR_OVERFLOW in its status.
 * No reference counterpart: the reference meshes one chunk per call
 * (GenerateMeshDataForNode, src/clipmap.cpp:356-427). */
int  lvn_batch_create(int32_t voxels_per_chunk, int32_t max_chunks, int32_t edge_cap, int32_t voxel_cap,
                      uint32_t flags, lvn_batch** out);
void lvn_batch_destroy(lvn_batch* b);
/* Stage n chunk requests; each chunk takes the stored CSG ops whose AABB overlaps it,
 * in store order (LoadDensityField's filter, src/compute_density_field.cpp:255-263).
 * Each chunk is meshed as a fresh chunk (no field or octree cache). */
int  lvn_batch_set_chunks(lvn_batch* b, int32_t n, const lvn_chunk_req* reqs);
/* Enqueue the whole pipeline for the staged chunks: one launch per stage covering
 * every chunk, replayed from a captured hipGraph.  Asynchronous. */
int  lvn_batch_run(lvn_batch* b);
int  lvn_batch_wait(lvn_batch* b);
/* Same as lvn_batch_run + wait without the graph, timing each stage with HIP events
 * on the batch stream: stage_ms[0 .. LVN_NUM_STAGES - 1] in pipeline order
 * (see lvn_batch_stage_name); n_stages >= LVN_NUM_STAGES. */
#define LVN_NUM_STAGES 12

/* Launch shapes of the batched pipeline (no counterpart in the reference, whose
 * NDRanges are fixed).  Results never depend on them, only time does. */
typedef struct lvn_tuning {
    int32_t density_ysplit;     /* y segments per (x, z) column work item of the density stage, 1..F */
    int32_t persistent_per_cu;  /* persistent workgroups per CU in the edge-crossing stage, 1..2 */
    int32_t row_blocks;         /* workgroups over the batch for the edge / voxel bookkeeping kernels */
    int32_t voxel_blocks;       /* workgroups over the batch for the per-voxel kernels (QEF, emit) */
    int32_t cross_lanes;        /* lanes sharing one edge in the edge-crossing stage: 1, 2, 4 or 8, or 16:
                                   sample-parallel (16 lanes per edge in the search, 6 in the normal) */
    int32_t density_per_cu;     /* persistent workgroups per CU in the density stage, 1..2 (0: 1) */
    int32_t height_lines;       /* heights of x/z-edge samples: 1 per sample, 2 once per line (0: 2 for
                                   batches of 64 chunks or more, else 1) */
    int32_t reserved[1];
} lvn_tuning;

int lvn_batch_get_tuning(lvn_batch* b, lvn_tuning* out);
/* Set the launch shapes (validated; the captured graph is rebuilt on the next run). */
int lvn_batch_set_tuning(lvn_batch* b, const lvn_tuning* t);
/* Autotuner: times candidate shapes on the staged chunks (one untimed and `reps`
 * timed passes per candidate, per-stage HIP events), one parameter at a time, keeps
 * the fastest for the stages each parameter drives, and writes the result to *out. */
int lvn_batch_autotune(lvn_batch* b, int32_t reps, lvn_tuning* out);
int  lvn_batch_run_timed(lvn_batch* b, float* stage_ms, int32_t n_stages);
const char* lvn_batch_stage_name(int32_t stage);
/* Per chunk: 6 ints = edges, active voxels, triangles, seam nodes, status, and the
 * vertices re-solved in FP64 by the precise QEF (LVN_CTX_PRECISE_QEF).  Waits. */
int  lvn_batch_counts(lvn_batch* b, int32_t* counts);
/* Copy chunk i's mesh out with generateChunkMesh's export rules (vertices only when
 * triangles exist, src/compute_octree.cpp:227-232). */
int  lvn_batch_get_mesh(lvn_batch* b, int32_t i, lvn_vertex* v, int32_t v_cap, int32_t* nv,
                        int32_t* tri, int32_t tri_cap, int32_t* nt, lvn_seam_node* s, int32_t s_cap, int32_t* ns);
/* Runs the pipeline once more with work counters on and returns the algorithmic work
 * the kernels performed: work[0] height-term and work[1] cave-term evaluations in the
 * density stage, work[2] / work[3] the same in the edge-crossing stage, work[4] brush
 * evaluations in the edge-crossing stage, work[5] density samples, work[6] edges. */
int  lvn_batch_work(lvn_batch* b, uint64_t* work);

/* The seams of a batch's chunks, stitched where they lie: as lvn_generate_seam_meshes,
 * but a source with count = -1 and no list is the batch's staged chunk at (min, size),
 * read from its slot after the last lvn_batch_run (LVN_ERR_STATE when no staged chunk
 * matches, LVN_ERR_OVERFLOW when that chunk overflowed).  Meshing a clipmap selection and
 * its seams is then two launches over device memory.  lvn_batch_seam_device_mesh gives
 * the device results of seams passed without host buffers.  No reference counterpart. */
int lvn_batch_seam_meshes(lvn_batch* b, int32_t n, const lvn_seam_req* reqs,
                          const lvn_seam_source* sources, lvn_seam_out* outs);
int lvn_batch_seam_device_mesh(lvn_batch* b, int32_t i, const lvn_vertex** v, const int32_t** tri,
                               int32_t* nv, int32_t* nt);

/* ---- serving form: continuous batching with device-resident chunk state ---- */

#define LVN_REQ_MESH  0   /* mesh the chunk (min, size) */
#define LVN_REQ_EDIT  1   /* store a CSG op and invalidate the chunks its AABB overlaps */

typedef struct lvn_request {
    int32_t    kind;
    int32_t    min[3];       /* mesh: chunk min, world units */
    int32_t    size;         /* mesh: chunk size, world units */
    int32_t    aabb_min[3];  /* edit: op AABB, world units (CalcCSGOperationBounds) */
    int32_t    aabb_max[3];
    int32_t    pad;
    lvn_csg_op op;           /* edit */
} lvn_request;

typedef struct lvn_result {
    int64_t ticket;
    int32_t kind;
    int32_t status;          /* LVN_OK or a negative LVN_ERR_* */
    int32_t nv, nt, ns;      /* mesh: exported vertices, triangles, seam nodes */
    int32_t cache_hit;       /* 1 when served from the device-resident cache */
    double  submit_s;        /* host steady clock, seconds */
    double  complete_s;
} lvn_result;

typedef struct lvn_server lvn_server;

/* A server whose chunk slots (each holding one chunk's field and mesh in HBM) fit in
 * hbm_budget bytes.  A scheduler thread admits waiting requests into free slots (evicting
 * least-recently-used finished chunks), forms batches of up to max_batch chunks and keeps
 * up to `lanes` batches in flight on separate streams.  Each chunk is meshed as a fresh
 * chunk with the stored ops overlapping it, so a result does not depend on what else is
 * in flight.  No reference counterpart: the reference meshes chunks one at a time on one
 * task thread (src/volume.cpp:68-113). */
int  lvn_server_create(int32_t voxels_per_chunk, uint64_t hbm_budget, int32_t max_batch, int32_t lanes,
                       int32_t edge_cap, int32_t voxel_cap, uint32_t flags, lvn_server** out);
void lvn_server_destroy(lvn_server* s);
/* Non-blocking.  An edit completes when it is recorded; later mesh requests see it. */
int  lvn_server_submit(lvn_server* s, const lvn_request* req, int64_t* ticket);
/* n requests at once (one lock, one wake-up of the scheduler), tickets[i] for reqs[i];
 * nothing is submitted when one of them is invalid. */
int  lvn_server_submit_many(lvn_server* s, int32_t n, const lvn_request* reqs, int64_t* tickets);
/* Moves up to cap completed results into out (completion order); *n = number moved. */
int  lvn_server_poll(lvn_server* s, lvn_result* out, int32_t cap, int32_t* n);
/* Blocks until every submitted request has completed. */
int  lvn_server_drain(lvn_server* s);
/* Copies the current mesh of a completed mesh request's chunk (LVN_ERR_STATE if it has
 * since been evicted or invalidated by an edit). */
int  lvn_server_get_mesh(lvn_server* s, int64_t ticket, lvn_vertex* v, int32_t v_cap, int32_t* nv,
                         int32_t* tri, int32_t tri_cap, int32_t* nt, lvn_seam_node* sn, int32_t s_cap, int32_t* ns);
/* Launch shapes of the server's batches, as lvn_batch_set_tuning (for instance the result
 * of lvn_batch_autotune on a batch of the expected chunks; zero density_ysplit,
 * cross_lanes, density_per_cu or height_lines = the automatic choice by batch size).
 * Applies to batches launched after the call; meshes do not depend on it.  No counterpart
 * in the reference (its chunks are meshed one by one). */
int  lvn_server_set_tuning(lvn_server* s, const lvn_tuning* t);
/* stats[0..7] = slots, batches launched, chunks meshed, cache hits, evictions,
 * results re-meshed because an edit arrived in flight, cached chunks, max batch seen. */
int  lvn_server_stats(lvn_server* s, int64_t* stats);

/* ---- post-mesh simplification (ngMeshSimplifier, src/ng_mesh_simplify.cpp) ---- */

/* = MeshSimplificationOptions (src/ng_mesh_simplify.h:6-31), plus the generator seed (the
 * reference seeds with 42, src/ng_mesh_simplify.cpp:195-196).  The result is what the
 * reference computes when built with libstdc++, with one change of arithmetic.  Its random
 * candidate edges: a std::mt19937 seeded anew in every iteration and
 * std::uniform_int_distribution<int>(0, E - 1), whose mapping the C++ library chooses;
 * libstdc++ (GCC 11 and later) takes Lemire's method (the high word of u * E, drawing again
 * while the low word is below 2^32 mod E), and so does this library; MSVC's, which leven
 * was built with, differs.  The change: the SSE QEF's Givens cosine (src/qef_simd.h:206)
 * takes _mm_rsqrt_ps, an approximation whose bits depend on the processor; here it is the
 * IEEE 1 / sqrt.  The rest of that solver is computed lane for lane as written, including
 * rotateq_xy's slips (:228-274: row 0's diagonal stored twice, y formed from cc * u and
 * cc + v and stored into the off-diagonal entries, the third diagonal never rotated) and the
 * pseudo-inverse's product with V where V^T is meant (:371-382). */
typedef struct lvn_simplify_options {
    float    edge_fraction;      /* 0.125: share of the candidate edges drawn per iteration */
    int32_t  max_iterations;     /* 10 */
    float    target_percentage;  /* 0.05: stop at this share of the input triangles */
    float    max_error;          /* collapse cost limit (cost = 1 / QEF error + degree penalty) */
    float    max_edge_size;      /* world units */
    float    min_angle_cosine;   /* of the two endpoint normals */
    uint32_t seed;               /* 42 */
    int32_t  reserved;
} lvn_simplify_options;

/* Replaces ngMeshSimplifier (src/ng_mesh_simplify.cpp:438-539) on one host mesh, in place:
 * vertices[0 .. *num_vertices), triangles[0 .. 3 * *num_triangles) (indices), positions
 * taken relative to offset (leven passes the chunk centre, w = 0) while simplifying.  A mesh
 * under 100 vertices or 100 triangles is left as it is.  stats (NULL or 2 ints) =
 * iterations run, edges collapsed.  Runs on the context's stream and waits. */
int lvn_simplify_mesh(lvn_ctx* ctx, lvn_vertex* vertices, int32_t* num_vertices, int32_t* triangles,
                      int32_t* num_triangles, const float offset[4], const lvn_simplify_options* options,
                      int32_t* stats);
/* Simplifies every staged chunk's mesh in its slot, as ConstructClipmapNodeData does
 * (src/clipmap.cpp:453-464): offset = the chunk centre min + size / 2, and max_error and
 * max_edge_size given per voxel (leven's meshMaxError and meshMaxEdgeLen, src/options.h)
 * scaled by the chunk's voxel size LEAF_SIZE_SCALE * (size / (LEAF_SIZE_SCALE * N)).  One
 * workgroup per chunk, every chunk in one launch.  lvn_batch_get_mesh then returns the
 * simplified meshes until the pipeline runs again; a second call simplifies them again.
 * results (NULL or 4 ints per chunk) = vertices, triangles, iterations run, edges collapsed
 * (iterations = LVN_ERR_OVERFLOW for a chunk that overflowed, which is left alone). */
int lvn_batch_simplify(lvn_batch* b, const lvn_simplify_options* options, int32_t* results);

/* ---- request producer: the camera-driven clipmap (src/clipmap.cpp) ----------- */

/* leven's clipmap as a request producer: which chunks to mesh as the camera moves, which
 * meshes went stale, which chunks are empty, which seams to regenerate.  Host-side state
 * only; the caller meshes with a context, a batch or a server.  Six LODs: chunk sizes
 * 4 * voxels_per_chunk * 2^lod world units, LOD_ACTIVE_DISTANCES as in src/clipmap.cpp:33-41.
 * The roots are the LOD_MAX_NODE_SIZE cells tiling [world_min, world_max) from world_min
 * (leven: one power-of-two cube around the world bounds, src/clipmap.cpp:797-818); a node's
 * children are made when the selection first descends into it. */
typedef struct lvn_clipmap lvn_clipmap;

#define LVN_CLIPMAP_TO_MESH   0   /* chunks the last update wants meshed, in selection order */
#define LVN_CLIPMAP_RELEASED  1   /* nodes the last update released: their meshes and octrees are stale */
#define LVN_CLIPMAP_SELECTED  2   /* every node the last update selected */
#define LVN_CLIPMAP_ACTIVE    3   /* every node holding a mesh or seam nodes now */
#define LVN_CLIPMAP_TOUCHED   4   /* nodes the last edit touched */

/* Replaces Clipmap::initialise / constructTree (src/clipmap.cpp:797-852); world units. */
int  lvn_clipmap_create(int32_t voxels_per_chunk, const int32_t world_min[3], const int32_t world_max[3],
                        lvn_clipmap** out);
void lvn_clipmap_destroy(lvn_clipmap* c);
/* Replaces the selection half of Clipmap::update (src/clipmap.cpp:1176-1252):
 * SelectActiveClipmapNodes (:1084-1126) around `camera` (world units), ReleaseInvalidatedNodes
 * (:1130-1151) and the frustum filter (AABBInsideFrustum, src/frustum.cpp:30-52) with six
 * planes (a, b, c, d), a box being outside when all 8 corners give a x + b y + c z + d < 0;
 * frustum == NULL keeps every node.  counts = {chunks to mesh, nodes released, nodes selected};
 * lvn_clipmap_list reads the lists until the next update. */
int lvn_clipmap_update(lvn_clipmap* c, const float camera[3], const float* frustum, int32_t counts[3]);
/* Copies list `which` (LVN_CLIPMAP_*) as chunk requests; LVN_ERR_CAPACITY with *n set when it
 * does not fit. */
int lvn_clipmap_list(lvn_clipmap* c, int32_t which, lvn_chunk_req* out, int32_t cap, int32_t* n);
/* Replaces the rest of Clipmap::update (src/clipmap.cpp:1254-1342) once the last update's
 * chunks are meshed: results[3 i .. 3 i + 2] = status, triangles, seam nodes of chunk i of the
 * to-mesh list.  A chunk with triangles or seam nodes becomes active; one with neither, or
 * whose meshing failed, becomes empty with everything below it (:1280-1287).  Then the seams
 * to regenerate (:1306-1324), each host with its sources gathered as GenerateClipmapSeamMesh
 * does (:579-597), count = -1 sources: the lists where the chunks' meshes lie, ready for
 * lvn_batch_seam_meshes / lvn_generate_seam_meshes.  counts = {seam hosts, sources}. */
int lvn_clipmap_commit(lvn_clipmap* c, int32_t n, const int32_t* results, int32_t counts[2]);
int lvn_clipmap_seams(lvn_clipmap* c, lvn_seam_req* reqs, int32_t req_cap, lvn_seam_source* sources,
                      int32_t src_cap);
/* Replaces the node half of Clipmap::processCSGOperationsImpl (src/clipmap.cpp:1655-1705):
 * every node up to LOD_MAX_NODE_SIZE that the op's AABB (world units) overlaps is invalidated
 * (released and re-meshed by the next update) and loses its empty mark; *n = nodes touched
 * (list LVN_CLIPMAP_TOUCHED). */
int lvn_clipmap_edit(lvn_clipmap* c, const int32_t aabb_min[3], const int32_t aabb_max[3], int32_t* n);

/* ---- allocation-free single-chunk form ------------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device memory:
 * code whose work is recorded under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc.  Meshes the chunk (min, size) fresh with the stored CSG ops
 * that overlap it and exports it with generateChunkMesh's rules, exactly as
 * lvn_generate_chunk_mesh does for a chunk its context has not cached, but needs no
 * context and calls no allocator: the chunk's working arrays live in one fixed device
 * area of the library (LVN_NOALLOC_SCRATCH_BYTES on each GPU it runs on, held until
 * the process ends), located per device by lvn_init or by the first call on a device.
 * Limits: voxels_per_chunk a power of two in [4, 64]; a chunk with more than
 * LVN_NOALLOC_MAX_EDGES compacted edges or LVN_NOALLOC_MAX_VOXELS active voxels, or
 * more than LVN_NOALLOC_MAX_OPS overlapping stored ops, returns LVN_ERR_OVERFLOW with
 * nothing written.  v / tri / s may be host or device memory.  The call is synchronous
 * and holds the library's global lock, so calls from one process never overlap; the
 * area keeps the last call's contents, and each call clears it before use. */
#define LVN_NOALLOC_MAX_EDGES     131072
#define LVN_NOALLOC_MAX_VOXELS    65536
#define LVN_NOALLOC_MAX_OPS       1024
#define LVN_NOALLOC_SCRATCH_BYTES (17u << 20)
int lvn_generate_chunk_mesh_noalloc(int32_t voxels_per_chunk, uint32_t flags, const int32_t min[3], int32_t size,
                                    lvn_vertex* v, int32_t v_cap, int32_t* nv,
                                    int32_t* tri, int32_t tri_cap, int32_t* nt,
                                    lvn_seam_node* s, int32_t s_cap, int32_t* ns);

/* ---- test hook: intermediate stages of one chunk --------------------------- */

/* Buffers for lvn_debug_chunk_stages.  F = voxels_per_chunk + 2, H = voxels_per_chunk + 1.
 * Any pointer may be NULL. Counts are always written. */
typedef struct lvn_stage_dump {
    float*   density;            /* F^3, x fastest (FP16 contexts: widened) */
    uint8_t* materials;          /* F^3 */
    int32_t* edge_ids;           /* edge_cap: ((x | y<<S | z<<2S) << 2) | axis, ascending */
    float*   edge_data;          /* 4 per edge: normal.xyz, t */
    int32_t  edge_cap;
    int32_t  num_edges;
    int32_t* voxel_codes;        /* voxel_cap each, ascending x + N*y + N*N*z order */
    int32_t* voxel_edge_masks;
    int32_t* voxel_materials;    /* (dominant << 8) | corner mask */
    float*   voxel_positions;    /* 4 per voxel: solved vertex, world units, w = 1 */
    float*   voxel_normals;      /* 4 per voxel */
    int32_t  voxel_cap;
    int32_t  num_voxels;
} lvn_stage_dump;

/* Runs the plain pipeline for one fresh chunk (no caches) with the given CSG
 * batch and copies out each stage.  Not a reference interface: exists so that
 * parity tests can compare every stage. */
int lvn_debug_chunk_stages(lvn_ctx* ctx, const int32_t min[3], int32_t size,
                           const lvn_csg_op* ops, int32_t nops, lvn_stage_dump* out);

#ifdef __cplusplus
}
#endif

#endif /* LVN_MESHER_H */
