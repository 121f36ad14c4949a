// lvn_kernels.h — launch interface of the gfx950 kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "lvn_shared.h"

namespace lvn {

// Arguments common to every stage kernel (passed by value).
struct KArgs {
    char* arena;             // chunk slot 0; slot i starts at arena + i * L.slot_bytes
    SlotLayout L;
    Geometry G;
    Caps caps;
    const ChunkDesc* chunks;
    Counters* counters;
    const DevOp* ops;
    const int* op_index;
    FieldParams F;
    const uint8_t* perm;     // 65,536 bytes
    const float4* grads;     // decoded gradients, one per stored table byte
    int defmat;
    int fp16;                // density stored as sign-preserving FP16
    int ysplit;              // density kernel: y segments per column
    int precise;             // QEF: compensated FP32 sums + FP64 re-solve
    unsigned long long* work;  // optional work counters (see lvn_batch_work), else null
    int* queue;              // kNumQueues work-queue heads, zeroed by launch_clear_counters
    int* item_prefix;        // chunks + 1 edge-item offsets (k_edge_items)
    int cus;                 // compute units of the device
    int persist_per_cu;      // persistent workgroups per CU (edge crossing)
    int density_per_cu;      // persistent workgroups per CU (density)
    int height_lines;        // 1: x/z-edge sample heights evaluated once per line (k_height_lines)
    const int* slot_map;     // batch entry -> arena slot, or null for the identity
    int row_blocks;          // workgroups over the batch for the bookkeeping kernels
    int voxel_blocks;        // workgroups over the batch for the per-voxel kernels
    int cross_lanes;         // lanes per edge in the crossing stage (1, 2, 4, 8; 16 = sample-parallel), 0 = by batch size
    int dens_cols;           // 1: a field without the cave term stores no density; crossings form
                             // their end samples as y - H from the column heights (h_cols)
    int mat_free;            // 1: a fresh chunk without brushes stores no materials; the density
                             // pass writes its solid bits as planes, which k_solid_rows turns
                             // into rows and the QEF reads corners from
};

// Global work-queue heads (queue[1]: the edge-crossing items; the density and
// line-height passes share their items within workgroups instead; the rest are spare).
constexpr int kNumQueues = 8;

enum ScanKind { kScanEdges = 0, kScanVoxels = 1, kScanQuadsSeams = 2 };

void launch_clear_counters(const KArgs& a, int n, hipStream_t s);

constexpr int kNumStages = LVN_NUM_STAGES;
extern const char* const kStageNames[kNumStages];
// Launch stage `i` (0..kNumStages-1) of the pipeline.
void launch_stage(const KArgs& a, int chunks, int i, hipStream_t s);

// Stage launchers: `chunks` chunk slots [0, chunks) processed by one launch each.
void launch_density(const KArgs& a, int chunks, hipStream_t s);
void launch_csg(const KArgs& a, int chunks, hipStream_t s);
void launch_edge_rows(const KArgs& a, int chunks, hipStream_t s);
void launch_scan(const KArgs& a, int chunks, ScanKind kind, hipStream_t s);
void launch_edge_compact(const KArgs& a, int chunks, hipStream_t s);
void launch_edge_cross(const KArgs& a, int chunks, hipStream_t s);
void launch_voxel_rows(const KArgs& a, int chunks, hipStream_t s);
void launch_voxel_compact(const KArgs& a, int chunks, hipStream_t s);
void launch_qef(const KArgs& a, int chunks, hipStream_t s);
void launch_emit(const KArgs& a, int chunks, hipStream_t s);

// Lanes per edge in the crossing stage when the launch shapes leave it open.
int default_cross_lanes(int chunks);
// cross_lanes value of the sample-parallel crossing kernel (16 lanes per edge in the
// search, 6 in the normal).
constexpr int kCrossSampleParallel = 16;

// Whole pipeline for chunk slots [0, chunks), counters cleared first.
void launch_pipeline(const KArgs& a, int chunks, hipStream_t s);

}  // namespace lvn
