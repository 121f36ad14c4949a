// lvn_shared.h — POD types shared by the host runtime and the gfx950 kernels.
#pragma once

#include <stdint.h>

#include "../../include/lvn_mesher.h"

namespace lvn {

// The permutation table in device memory and LDS.  Each byte holds its texel's value
// plus one (mod 256).  Rows of 258 bytes: texel (x, y) at y * 258 + 1 + x for x, y in
// 0..256 (x or y = 256 repeating 0), and byte y * 258 holds texel (0, y) once more.
// 257 rows, padded to 16 bytes.  The gradient table has 16 entries indexed by the stored
// byte's low four bits: entry j holds gradient (j - 1) & 15, the gradient of every value
// w - 1 with w & 15 == j.
constexpr int kPermPitch = 258;
constexpr int kPermRows = 257;
constexpr int kPermPaddedBytes = ((kPermPitch * kPermRows + 15) / 16) * 16;
constexpr int kGradEntries = 16;


// Interior samples (k = 1..15 of the 17) of an x- or z-edge line whose heights are
// evaluated once per line (k_height_lines).
constexpr int kLineSamples = 15;

constexpr int kAir = LVN_MATERIAL_AIR;    // src/volume_materials.h:9
constexpr int kNone = LVN_MATERIAL_NONE;  // src/volume_materials.h:8
constexpr int kLeafSizeScale = 4;         // src/volume_constants.h:7-8
constexpr int kClipmapLeafSize = 256;     // src/volume_constants.h:14-15

// Device copy of lvn_field_desc plus the cave-term cut (see combine()).
struct FieldParams {
    int kind;
    int hoct;
    float hfreq, hlac, hgain, hamp;
    int coct;
    float cfreq, clac, cgain, cthr;
    float cave_cut;
};

// Device form of one CSG operation: cos/sin of rotate_y precomputed on the
// host, and the op's bounding radius for work culling.
struct DevOp {
    int type, brush, material, pad;
    float ox, oy, oz, radius;
    float dx, dy, dz, pad2;
    float c, s, pad3, pad4;
};

// Geometry of a chunk (src/compute.cpp:245-278): N voxels per side,
// H = N + 1 hermite points, F = N + 2 field samples, index shift S = log2(N) + 1.
struct Geometry {
    int N, H, F, shift, depth;
    int field_samples;   // F^3
    int edge_slots;      // 3 H^3
    int field_rows;      // F^2 rows (y, z) of F samples: one solid bit per sample
    int edge_rows;       // H^2 rows (y, z) of H hermite points: x, y, z edge bit masks
    int voxels;          // N^3
    int voxel_segs;      // N^3 / 64 (>= 1)
    uint32_t magic_h;    // ceil(2^32 / H): n / H == umulhi(n, magic_h) for n < 2^21
    uint32_t magic_f;    // ceil(2^32 / F)
    int col_groups;      // ceil(F^2 / 64): groups of 64 field columns (x, z), column x + F z
};

// Per-slot capacities of the compacted arrays.
struct Caps {
    int edges;     // E
    int voxels;    // V
    int quads;     // Q (<= 3 V)
    int seams;     // <= V
    int seam_segs;     // ceil(V / 64): voxel segments of the quad and seam scans
};

// Byte offsets of every array inside one chunk slot of the arena.
struct SlotLayout {
    uint64_t slot_bytes;
    uint64_t density, mat, orig;
    uint64_t solid;                       // field_rows x 2 words: bit x = sample (x, y, z) is not air
    uint64_t planes;                      // col_groups x F words, [g][y] bit c: column 64 g + c of plane y
                                          // is not air (chunks whose materials are not stored)
    uint64_t e_rows, e_count, e_base, e_ids, e_data;  // e_rows: edge_rows x {x, y, z, x = 64} words
    uint64_t e_count_y, e_base_y, e_order;  // y-edges per row and their scan; crossing order (y-edges first)
    uint64_t h_lines;                     // heights at k/16 (k = 1..15) along x- and z-edge lines, 2 x H x H x 15 f32
    uint64_t h_cols;                      // the height term of each field column (x, z), F x F f32 (fresh fields)
    uint64_t v_mask, v_count, v_base, v_code, v_emask, v_mat;
    uint64_t q_count, q_base, tris;
    uint64_t s_mask, s_count, s_base, seams, verts;
};

// Per-chunk request descriptor.
struct ChunkDesc {
    int min[3];
    int size;
    int off[3];      // min / 4 (C truncation), voxel units
    int scale;       // size / (N * 4)
    int op_start;    // into the op-index array
    int op_count;
    int mode;        // kModeFresh / kModeCached
    int old_edges;   // cached mode: number of edges in the old list
    float colour[4];
    uint64_t old_ids;   // cached mode: device pointer to the old sorted edge ids
    uint64_t old_data;  // cached mode: device pointer to the old edge data (float4)
    uint64_t old_mat;   // cached mode: device pointer to the cached materials (F^3 bytes)
};

constexpr int kModeFresh = 0;   // density from the field; untouched edges take density crossings
constexpr int kModeCached = 1;  // materials from old_mat; untouched edges keep their data from the old list

// Per-chunk counters written by the kernels.
struct Counters {
    int edges, voxels, quads, seams;
    int status;      // 0 ok, else LVN_ERR_OVERFLOW bits
    int resolved;    // precise QEF: vertices re-solved in FP64
    int edges_y;     // y-edges among `edges` (capped like edges)
    int pad;
};

constexpr int kStatusEdgeOverflow = 1;
constexpr int kStatusVoxelOverflow = 2;
constexpr int kStatusQuadOverflow = 4;
constexpr int kStatusSeamOverflow = 8;
constexpr int kStatusLookupMiss = 16;

}  // namespace lvn
