// lvn_device.hpp — device-side arithmetic of the chunk mesher (gfx950).
//
// Every function follows the operation order of the reference formula it
// cites, with no contraction (the library is compiled with -ffp-contract=off),
// correctly rounded division and sqrt (HIP's default for FP32), and float
// constants in place of the reference's unsuffixed double literals.  The
// permutation table lives in LDS (64 KiB of bytes, row = texel y) together
// with the decoded gradient vectors.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lvn_shared.h"

namespace lvn {

// LDS-resident noise tables, built by the host.  The 256 x 256 permutation table is
// stored in rows of 258 bytes with a wrap row: texel (x, y) at y * 258 + 1 + x for x, y
// in 0..256 (256 repeating 0), so a corner's texel, whose coordinates are the cell's
// wrapped ones plus 0 or 1, is one add away from the cell's base index with no wrap
// masks per corner (REPEAT addressing, cl/simplex.cl:59).  Each byte holds the texel's
// value plus one, mod 256, and byte y * 258 holds texel (0, y): snoise3's second read,
// at x = value with 255 wrapping to 0, is then row + stored byte with no compare.  The
// 16-entry gradient table is indexed by the stored byte's low four bits (rotated by one
// entry, so no subtraction is needed).  The gradients are stored as (x, y) pairs, 8 bytes
// an entry, and z components at the same 8-byte stride, the two blocks on bank rows: the
// 16 pairs fill the 32 banks once and the 16 z values take the even banks once, so a
// wave's reads never meet two addresses in one bank (equal gradients share an address;
// whole 16-byte entries put entries 8 apart in the same banks).  The gradient blocks
// come first: their offsets then fit the LDS instructions' 16-bit immediate, and a
// corner's two reads share one address, its index times 8.  snoise2 reads a pair,
// snoise3 a pair and a z.  66,688 bytes per workgroup.
struct alignas(128) NoiseLds {
    float2 gxy[kGradEntries];  // entry j: gradient (j - 1) & 15, components x and y
    float2 gz[kGradEntries];   // entry j: (z component, unused)
    uint8_t perm[kPermPaddedBytes];
};
static_assert(offsetof(NoiseLds, gz) == 128 && offsetof(NoiseLds, perm) == 256,
              "gradient blocks on bank rows, ahead of the permutation table");

// Cooperative copy of the tables into LDS (all threads of the block call it).
__device__ __forceinline__ void load_noise_tables(NoiseLds& lds, const uint8_t* __restrict__ perm,
                                                  const float4* __restrict__ grads) {
    const uint4* src = reinterpret_cast<const uint4*>(perm);
    uint4* dst = reinterpret_cast<uint4*>(lds.perm);
    for (int i = threadIdx.x; i < kPermPaddedBytes / 16; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < kGradEntries; i += blockDim.x) {
        const float4 g = grads[i];
        lds.gxy[i] = make_float2(g.x, g.y);
        lds.gz[i] = make_float2(g.z, 0.f);
    }
    __syncthreads();
}

// Correctly rounded square root of x >= 1 (and of +inf): the hardware square root,
// within 1 ulp, replaced by its lower or upper neighbour when the FMA residual
// x - neighbour * root says that one is nearer.  These are the steps of the
// compiler's IEEE sqrtf without its range scaling (taken below 2^-96 only) and its
// zero / infinity select (+inf already comes out as +inf: both residuals are NaN).
__device__ __forceinline__ float sqrt_ge1(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const float hi = __int_as_float(__float_as_int(s) + 1);
    const float r = fmaf(-lo, s, x) <= 0.f ? lo : s;
    return fmaf(-hi, s, x) > 0.f ? hi : r;
}

// Correctly rounded 1 / d for d in [1, 2): the hardware reciprocal refined by one
// Newton step, then the quotient corrected twice by its FMA residual.  These are the
// steps of the compiler's IEEE division 1.f / d with its operand scaling and its
// special-case fixup left out: for this range neither changes a value.
__device__ __forceinline__ float recip_1_2(float d) {
    const float r0 = __builtin_amdgcn_rcpf(d);
    const float r1 = fmaf(fmaf(-d, r0, 1.f), r0, r0);
    const float q1 = fmaf(fmaf(-d, r1, 1.f), r1, r1);
    return fmaf(fmaf(-d, q1, 1.f), r1, q1);
}

// One snoise2 corner (cl/simplex.cl:123-153).  The squared distance and the
// gradient dot product are fused multiply-adds, as the restatement defines them
// (the reference's -cl-fast-relaxed-math build, src/compute.cpp:213, allows it).
// A corner beyond the radius (r < 0) contributes zero: r is clamped at +0, which
// gives that zero up to its sign (+0 times the dot product).  No result can see the
// sign: the corner sum of a point with a non-zero corner is unchanged by it, and
// every noise value feeds either fBm's fma chain, which starts at +0 (a zero noise
// value then adds +0 to a sum that is never -0), or the ridged form's fabsf.
__device__ __forceinline__ float corner2(const NoiseLds& T, int texel, float fx, float fy) {
    const float r = fmaxf(0.5f - fmaf(fy, fy, fx * fx), 0.f);
    const float2 g = T.gxy[T.perm[texel] & 15];
    const float t = r * r;
    return (t * t) * fmaf(g.y, fy, g.x * fx);
}

// snoise2, cl/simplex.cl:99-157.
__device__ __forceinline__ float snoise2(const NoiseLds& T, float px, float py) {
    const float F2 = 0.366025403784f;
    const float G2 = 0.211324865405f;
    const float C2 = 1.f - 2.f * G2;
    const float s = (px + py) * F2;
    const float pix = floorf(px + s);
    const float piy = floorf(py + s);
    const float t = (pix + piy) * G2;
    const float f0x = px - (pix - t);
    const float f0y = py - (piy - t);
    const int base = ((int)piy & 255) * kPermPitch + ((int)pix & 255) + 1;  // texel of the cell corner
    const int o1x = f0x > f0y ? 1 : 0;
    const int o1y = 1 - o1x;
    const float n0 = corner2(T, base, f0x, f0y);
    const float n1 = corner2(T, base + o1y * kPermPitch + o1x, (f0x - (float)o1x) + G2, (f0y - (float)o1y) + G2);
    const float n2 = corner2(T, base + kPermPitch + 1, f0x - C2, f0y - C2);
    return 70.f * ((n0 + n1) + n2);
}

// One snoise3 corner (cl/simplex.cl:181-226).  The second table read uses the
// first read's value as x texel; value 255 wraps to texel 0 (normalised REPEAT).
// `texel` is the corner's (x, y) texel, `zrow` the start of its z texel's row; the
// stored byte (value + 1, or 0 for 255) is the second read's offset in that row.
// r is clamped at +0 as in corner2.
__device__ __forceinline__ float corner3(const NoiseLds& T, int texel, int zrow, float fx, float fy, float fz) {
    const int w = T.perm[texel];
    const float r = fmaxf(0.6f - fmaf(fz, fz, fmaf(fy, fy, fx * fx)), 0.f);
    const int j = T.perm[zrow + w] & 15;
    const float2 g = T.gxy[j];
    const float t = r * r;
    return (t * t) * fmaf(T.gz[j].x, fz, fmaf(g.y, fy, g.x * fx));
}

// snoise3, cl/simplex.cl:159-230 (simplex ranking of :74-97).
__device__ __forceinline__ float snoise3(const NoiseLds& T, float px, float py, float pz) {
    const float F3 = 0.333333333333f;
    const float G3 = 0.166666666667f;
    const float C2 = 2.f * G3;
    const float C3 = 1.f - 3.f * G3;
    const float s = ((px + py) + pz) * F3;
    const float pix = floorf(px + s);
    const float piy = floorf(py + s);
    const float piz = floorf(pz + s);
    const float t = ((pix + piy) + piz) * G3;
    const float f0x = px - (pix - t);
    const float f0y = py - (piy - t);
    const float f0z = pz - (piz - t);
    const int base = ((int)piy & 255) * kPermPitch + ((int)pix & 255) + 1;  // (x, y) texel of the cell corner
    const int zrow = ((int)piz & 255) * kPermPitch;
    const int xy = f0x >= f0y ? 1 : 0;
    const int xz = f0x >= f0z ? 1 : 0;
    const int yz = f0y >= f0z ? 1 : 0;
    const int rx = xy + xz;
    const int ry = (1 - xy) + yz;
    const int rz = (1 - xz) + (1 - yz);
    const int a1x = rx >> 1, a1y = ry >> 1, a1z = rz >> 1;            // rank 2 -> 1
    const int a2x = rx > 0 ? 1 : 0, a2y = ry > 0 ? 1 : 0, a2z = rz > 0 ? 1 : 0;  // rank >= 1 -> 1
    // f0 - offset for offsets 0 and 1: f0 - 0.f is f0 itself (also for -0)
    const float g0x = f0x - 1.f, g0y = f0y - 1.f, g0z = f0z - 1.f;
    const float n0 = corner3(T, base, zrow, f0x, f0y, f0z);
    const float n1 = corner3(T, base + a1y * kPermPitch + a1x, zrow + a1z * kPermPitch, (a1x ? g0x : f0x) + G3,
                             (a1y ? g0y : f0y) + G3, (a1z ? g0z : f0z) + G3);
    const float n2 = corner3(T, base + a2y * kPermPitch + a2x, zrow + a2z * kPermPitch, (a2x ? g0x : f0x) + C2,
                             (a2y ? g0y : f0y) + C2, (a2z ? g0z : f0z) + C2);
    const float n3 = corner3(T, base + kPermPitch + 1, zrow + kPermPitch, f0x - C3, f0y - C3, f0z - C3);
    return 32.f * (((n0 + n1) + n2) + n3);
}

// BasicFractal, cl/noise.cl:8-36.
__device__ __forceinline__ float fbm2(const NoiseLds& T, int octaves, float freq, float lac, float gain, float x,
                                      float y) {
    float px = x * freq, py = y * freq, amp = 1.f, n = 0.f;
    for (int i = 0; i < octaves; ++i) {
        n = fmaf(snoise2(T, px, py), amp, n);
        px = px * lac;
        py = py * lac;
        amp = amp * gain;
    }
    return n;
}

// The same fractal over snoise3.
__device__ __forceinline__ float fbm3(const NoiseLds& T, int octaves, float freq, float lac, float gain, float x,
                                      float y, float z) {
    float px = x * freq, py = y * freq, pz = z * freq, amp = 1.f, n = 0.f;
    for (int i = 0; i < octaves; ++i) {
        n = fmaf(snoise3(T, px, py, pz), amp, n);
        px = px * lac;
        py = py * lac;
        pz = pz * lac;
        amp = amp * gain;
    }
    return n;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// RidgedMultiFractal, cl/noise.cl:42-81 (pow(f, -1) as 1 / f).
__device__ __forceinline__ float ridged(const NoiseLds& T, int octaves, float lac, float gain, float offset, float x,
                                        float y) {
    float px = x, py = y;
    float signal = offset - fabsf(snoise2(T, px, py));
    signal = signal * signal;
    float n = signal, freq = 1.f;
    for (int i = 0; i < octaves; ++i) {
        px = px * lac;
        py = py * lac;
        const float weight = clamp01(signal * gain);
        signal = offset - fabsf(snoise2(T, px, py));
        signal = signal * weight;
        const float expo = 1.f / freq;
        freq = freq * lac;
        n = n + signal * expo;
    }
    return n * (1.f / (float)octaves);
}

// Terrain, cl/noise.cl:205-223.
__device__ __forceinline__ float leven_terrain(const NoiseLds& T, float x, float z) {
    const float qx = x * (1.f / 2000.f);
    const float qy = z * (1.f / 2000.f);
    const float r = clamp01(0.8f * ridged(T, 7, 2.114352f, 1.5241f, 1.f, qx, qy));
    float billow = 0.6f * fbm2(T, 4, 0.24f, 1.8754f, 0.433f, -4.33f * qx, 7.98f * qy);
    billow = (0.5f * billow) + 0.5f;
    const float n = billow * r;
    float b2 = 0.6f * fbm2(T, 2, 0.63f, 2.2f, 0.15f, qx, qy);
    b2 = (b2 * 0.5f) + 0.5f;
    return n + b2;
}

// Height term H(x, z) of every field kind: density = y - H (cl/noise.cl:265-266).
__device__ __forceinline__ float height_term(const FieldParams& F, const NoiseLds& T, float x, float z) {
    if (F.kind == LVN_FIELD_LEVEN_TERRAIN) return 900.f * leven_terrain(T, x, z);
    return F.hamp * fbm2(T, F.hoct, F.hfreq, F.hlac, F.hgain, x, z);
}

// Cave term of LVN_FIELD_TERRAIN_CAVES.
__device__ __forceinline__ float cave_term(const FieldParams& F, const NoiseLds& T, float x, float y, float z) {
    return fbm3(T, F.coct, F.cfreq, F.clac, F.cgain, x, y, z) - F.cthr;
}

// density = max(a, cave) with a = y - H.  The cave term is bounded by F.cave_cut
// (|snoise3| <= 1.19 times the sum of octave amplitudes, minus the threshold);
// when a exceeds it, max(a, cave) == a exactly and the cave term is skipped.
__device__ __forceinline__ float combine(const FieldParams& F, const NoiseLds& T, float a, float x, float y, float z) {
    if (F.kind != LVN_FIELD_TERRAIN_CAVES || a > F.cave_cut) return a;
    return fmaxf(a, cave_term(F, T, x, y, z));
}

__device__ __forceinline__ float density(const FieldParams& F, const NoiseLds& T, float x, float y, float z) {
    const float a = y - height_term(F, T, x, z);
    return combine(F, T, a, x, y, z);
}

// ---------------------------------------------------------------------------
// CSG brushes (cl/apply_csg_operation.cl:42-81)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

// Density_Sphere / Density_Cuboid (pR rotation of xz, then the box SDF
// length(max(d,0)) + max(min(d,0)) with d = |p| - half extents).
__device__ __forceinline__ float brush_density(const DevOp& op, float x, float y, float z) {
    const float lx = x - op.ox;
    const float ly = y - op.oy;
    const float lz = z - op.oz;
    if (op.brush == 1) return len3(lx, ly, lz) - op.dx;
    const float rx = op.c * lx + op.s * lz;
    const float rz = op.c * lz + op.s * (-lx);
    const float ex = fabsf(rx) - op.dx;
    const float ey = fabsf(ly) - op.dy;
    const float ez = fabsf(rz) - op.dz;
    const float outside = len3(fmaxf(ex, 0.f), fmaxf(ey, 0.f), fmaxf(ez, 0.f));
    const float inside = fmaxf(fmaxf(fminf(ex, 0.f), fminf(ey, 0.f)), fminf(ez, 0.f));
    return outside + inside;
}

// normalize with zero vectors passed through.
__device__ __forceinline__ float3 normalize_or_zero(float x, float y, float z) {
    if (x == 0.f && y == 0.f && z == 0.f) return make_float3(x, y, z);
    const float l = len3(x, y, z);
    return make_float3(x / l, y / l, z / l);
}

// ---------------------------------------------------------------------------
// Octree codes (cl/octree.cl:34-72)
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t code_for_position(int x, int y, int z, int depth) {
    uint32_t code = 1;
    for (int d = depth - 1; d >= 0; --d)
        code = (code << 3) | (uint32_t)((((x >> d) & 1) << 2) | (((y >> d) & 1) << 1) | ((z >> d) & 1));
    return code;
}

__device__ __forceinline__ int3 position_for_code(uint32_t code, int max_depth) {
    const int node_depth = (32 - __clz(code)) / 3;
    int3 p = make_int3(0, 0, 0);
    for (int i = max_depth - node_depth; i < max_depth; ++i) {
        const uint32_t c = code & 7u;
        code >>= 3;
        p.x |= (int)((c >> 2) & 1u) << i;
        p.y |= (int)((c >> 1) & 1u) << i;
        p.z |= (int)(c & 1u) << i;
    }
    return p;
}

}  // namespace lvn
