// lvn_oracle.cpp — CPU oracle of the dual-contouring chunk mesher.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg.  Never linked into the product library.
//
// A scalar restatement of leven's OpenCL pipeline.  Each function names the
// reference lines it follows (paths relative to leven/).  Arithmetic contract:
// see lvn_oracle.h.  Built with -ffp-contract=off and without -ffast-math.

#include "lvn_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kAir = LVN_MATERIAL_AIR;    // src/volume_materials.h:9
constexpr int kNone = LVN_MATERIAL_NONE;  // src/volume_materials.h:8
constexpr int kLeafSizeScale = 4;         // src/volume_constants.h:7-8
constexpr int kClipmapLeafSize = 256;     // src/volume_constants.h:14-15 (4 * 64)

// ---------------------------------------------------------------------------
// A1 — permutation table (src/compute_density_field.cpp:28-125)
// ---------------------------------------------------------------------------

// Ken Perlin's reference permutation ("Improved Noise", 2002).  leven's perm[512]
// (src/compute_density_field.cpp:28-52) holds it once, then the same list without
// its first six entries, and its last six elements are zero-initialised.
const uint8_t kPerlin[256] = {
    151, 160, 137, 91,  90,  15,  131, 13,  201, 95,  96,  53,  194, 233, 7,   225, 140, 36,  103, 30,  69,  142,
    8,   99,  37,  240, 21,  10,  23,  190, 6,   148, 247, 120, 234, 75,  0,   26,  197, 62,  94,  252, 219, 203,
    117, 35,  11,  32,  57,  177, 33,  88,  237, 149, 56,  87,  174, 20,  125, 136, 171, 168, 68,  175, 74,  165,
    71,  134, 139, 48,  27,  166, 77,  146, 158, 231, 83,  111, 229, 122, 60,  211, 133, 230, 220, 105, 92,  41,
    55,  46,  245, 40,  244, 102, 143, 54,  65,  25,  63,  161, 1,   216, 80,  73,  209, 76,  132, 187, 208, 89,
    18,  169, 200, 196, 135, 130, 116, 188, 159, 86,  164, 100, 109, 198, 173, 186, 3,   64,  52,  217, 226, 250,
    124, 123, 5,   202, 38,  147, 118, 126, 255, 82,  85,  212, 207, 206, 59,  227, 47,  16,  58,  17,  182, 189,
    28,  42,  223, 183, 170, 213, 119, 248, 152, 2,   44,  154, 163, 70,  221, 153, 101, 155, 167, 43,  172, 9,
    129, 22,  39,  253, 19,  98,  108, 110, 79,  113, 224, 232, 178, 185, 112, 104, 218, 246, 97,  228, 251, 34,
    242, 193, 238, 210, 144, 12,  191, 179, 162, 241, 81,  51,  145, 235, 249, 14,  239, 107, 49,  192, 214, 31,
    181, 199, 106, 157, 184, 84,  204, 176, 115, 121, 50,  45,  127, 4,   150, 254, 138, 236, 205, 93,  222, 114,
    67,  29,  24,  72,  243, 141, 128, 195, 78,  66,  215, 61,  156, 180};

// Perlin's 16 gradient directions (src/compute_density_field.cpp:60-63).
const int kGrad3[16][3] = {{0, 1, 1},  {0, 1, -1},  {0, -1, 1}, {0, -1, -1}, {1, 0, 1},  {1, 0, -1},
                           {-1, 0, 1}, {-1, 0, -1}, {1, 1, 0},  {1, -1, 0},  {-1, 1, 0}, {-1, -1, 0},
                           {1, 0, -1}, {-1, 0, -1}, {0, -1, 1}, {0, 1, 1}};

void base_perm512(int32_t* out) {
    for (int i = 0; i < 256; ++i) out[i] = kPerlin[i];
    for (int i = 0; i < 250; ++i) out[256 + i] = kPerlin[6 + i];
    for (int i = 506; i < 512; ++i) out[i] = 0;
}

// The reference shuffles with std::shuffle + std::default_random_engine(seed)
// (src/compute_density_field.cpp:98), whose output is library-defined.  The
// restatement fixes a documented generator: Fisher-Yates from the top index
// down, j = splitmix64() % (i + 1), state seeded with (uint32)seed.
uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void shuffled_perm512(int32_t seed, int32_t* out) {
    base_perm512(out);
    uint64_t s = (uint64_t)(uint32_t)seed;
    for (int i = 511; i > 0; --i) {
        const uint64_t j = splitmix64(s) % (uint64_t)(i + 1);
        std::swap(out[i], out[j]);
    }
}

// Bob Jenkins' one-at-a-time hash.
uint32_t oat_hash(const uint8_t* bytes, int n) {
    uint32_t h = 0;
    for (int i = 0; i < n; ++i) {
        h += bytes[i];
        h += h << 10;
        h ^= h >> 6;
    }
    h += h << 3;
    h ^= h >> 11;
    h += h << 15;
    return h;
}

// NoiseHash: one-at-a-time over the 4 little-endian bytes of
// ((x << 24) | (y << 16)) ^ seed (src/compute_density_field.cpp:68-86).
uint32_t noise_hash(int32_t x, int32_t y, int32_t seed) {
    const uint32_t key = (((uint32_t)x << 24) | ((uint32_t)y << 16)) ^ (uint32_t)seed;
    const uint8_t b[4] = {(uint8_t)key, (uint8_t)(key >> 8), (uint8_t)(key >> 16), (uint8_t)(key >> 24)};
    return oat_hash(b, 4);
}

// UNORM8 texel decode of the gradient bytes grad3*64+64 ∈ {0, 64, 128}
// (src/compute_density_field.cpp:108-110, read as byte/255*4-1 in cl/simplex.cl:124).
float grad_value(int component) {
    const float byte = (float)(component * 64 + 64);
    const float unorm = byte / 255.0f;
    return unorm * 4.0f - 1.0f;
}

// ---------------------------------------------------------------------------
// A2/A3 — simplex noise (cl/simplex.cl)
// ---------------------------------------------------------------------------

// Texture read with NEAREST / REPEAT / normalised coordinates (cl/simplex.cl:59):
// the texel of integer lattice coordinate i is i mod 256 (exact integer wrap).
inline int texel(const uint8_t* perm, int tx, int ty) { return perm[((ty & 255) << 8) | (tx & 255)]; }

// snoise3 re-reads the table at x-coordinate perm/255 (cl/simplex.cl:182-183):
// the repeat wrap maps value 255 (coordinate 1.0) to texel 0, all others to themselves.
inline int perm_as_texel(int v) { return v == 255 ? 0 : v; }

struct GradTable {
    float v[3];
    GradTable() {
        v[0] = grad_value(-1);
        v[1] = grad_value(0);
        v[2] = grad_value(1);
    }
    float comp(int g, int c) const { return v[kGrad3[g & 15][c] + 1]; }
};
const GradTable kGrads;

// snoise2, cl/simplex.cl:99-157.
float snoise2(const uint8_t* perm, float px, float py) {
    const float F2 = 0.366025403784f;
    const float G2 = 0.211324865405f;
    const float C2 = 1.f - 2.f * G2;

    const float s = (px + py) * F2;
    const float pix = floorf(px + s);
    const float piy = floorf(py + s);
    const float t = (pix + piy) * G2;
    const float x0 = pix - t;
    const float y0 = piy - t;
    const float f0x = px - x0;
    const float f0y = py - y0;
    const int ix = (int)pix;
    const int iy = (int)piy;
    const int o1x = f0x > f0y ? 1 : 0;
    const int o1y = 1 - o1x;

    float n0 = 0.f, n1 = 0.f, n2 = 0.f;
    {
        const int g = texel(perm, ix, iy);
        float t0 = 0.5f - std::fma(f0y, f0y, f0x * f0x);
        if (!(t0 < 0.f)) {
            t0 = t0 * t0;
            n0 = (t0 * t0) * std::fma(kGrads.comp(g, 1), f0y, kGrads.comp(g, 0) * f0x);
        }
    }
    {
        const float fx = (f0x - (float)o1x) + G2;
        const float fy = (f0y - (float)o1y) + G2;
        const int g = texel(perm, ix + o1x, iy + o1y);
        float t1 = 0.5f - std::fma(fy, fy, fx * fx);
        if (!(t1 < 0.f)) {
            t1 = t1 * t1;
            n1 = (t1 * t1) * std::fma(kGrads.comp(g, 1), fy, kGrads.comp(g, 0) * fx);
        }
    }
    {
        const float fx = f0x - C2;
        const float fy = f0y - C2;
        const int g = texel(perm, ix + 1, iy + 1);
        float t2 = 0.5f - std::fma(fy, fy, fx * fx);
        if (!(t2 < 0.f)) {
            t2 = t2 * t2;
            n2 = (t2 * t2) * std::fma(kGrads.comp(g, 1), fy, kGrads.comp(g, 0) * fx);
        }
    }
    return 70.f * ((n0 + n1) + n2);
}

// One snoise3 corner: cl/simplex.cl:182-226.
inline float snoise3_corner(const uint8_t* perm, int tx, int ty, int tz, float fx, float fy, float fz) {
    const int p = texel(perm, tx, ty);
    const int g = texel(perm, perm_as_texel(p), tz);
    float t = 0.6f - std::fma(fz, fz, std::fma(fy, fy, fx * fx));
    if (t < 0.f) return 0.f;
    t = t * t;
    return (t * t) * std::fma(kGrads.comp(g, 2), fz, std::fma(kGrads.comp(g, 1), fy, kGrads.comp(g, 0) * fx));
}

// snoise3, cl/simplex.cl:159-230, with simplex() of :74-97.
float snoise3(const uint8_t* perm, float px, float py, float pz) {
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
    const int ix = (int)pix, iy = (int)piy, iz = (int)piz;

    // simplex(): rank of each component (ties favour x, then y).
    const int isXy = f0x >= f0y ? 1 : 0;
    const int isXz = f0x >= f0z ? 1 : 0;
    const int isYz = f0y >= f0z ? 1 : 0;
    const int rx = isXy + isXz;
    const int ry = (1 - isXy) + isYz;
    const int rz = (1 - isXz) + (1 - isYz);
    const int o2x = std::min(rx, 1), o2y = std::min(ry, 1), o2z = std::min(rz, 1);
    const int o1x = std::max(rx - 1, 0), o1y = std::max(ry - 1, 0), o1z = std::max(rz - 1, 0);

    const float n0 = snoise3_corner(perm, ix, iy, iz, f0x, f0y, f0z);
    const float n1 = snoise3_corner(perm, ix + o1x, iy + o1y, iz + o1z, (f0x - (float)o1x) + G3,
                                    (f0y - (float)o1y) + G3, (f0z - (float)o1z) + G3);
    const float n2 = snoise3_corner(perm, ix + o2x, iy + o2y, iz + o2z, (f0x - (float)o2x) + C2,
                                    (f0y - (float)o2y) + C2, (f0z - (float)o2z) + C2);
    const float n3 = snoise3_corner(perm, ix + 1, iy + 1, iz + 1, f0x - C3, f0y - C3, f0z - C3);
    return 32.f * (((n0 + n1) + n2) + n3);
}

// ---------------------------------------------------------------------------
// A4 — fractals and density functions (cl/noise.cl)
// ---------------------------------------------------------------------------

// BasicFractal, cl/noise.cl:8-36 (NOISE_SCALE = 1).
float fbm2(const uint8_t* perm, int octaves, float frequency, float lacunarity, float gain, float x, float y) {
    float px = x * frequency, py = y * frequency;
    float amplitude = 1.f, noise = 0.f;
    for (int i = 0; i < octaves; ++i) {
        noise = std::fma(snoise2(perm, px, py), amplitude, noise);
        px = px * lacunarity;
        py = py * lacunarity;
        amplitude = amplitude * gain;
    }
    return noise;
}

// The same fractal over snoise3 (the cave term of the 3D terrain workload).
float fbm3(const uint8_t* perm, int octaves, float frequency, float lacunarity, float gain, float x, float y,
           float z) {
    float px = x * frequency, py = y * frequency, pz = z * frequency;
    float amplitude = 1.f, noise = 0.f;
    for (int i = 0; i < octaves; ++i) {
        noise = std::fma(snoise3(perm, px, py, pz), amplitude, noise);
        px = px * lacunarity;
        py = py * lacunarity;
        pz = pz * lacunarity;
        amplitude = amplitude * gain;
    }
    return noise;
}

inline float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// RidgedMultiFractal, cl/noise.cl:42-81; pow(frequency, -1) restated as 1 / frequency.
float ridged_multifractal(const uint8_t* perm, int octaves, float lacunarity, float gain, float offset, float x,
                          float y) {
    float px = x, py = y;
    float signal = snoise2(perm, px, py);
    signal = fabsf(signal);
    signal = offset - signal;
    signal = signal * signal;
    float noise = signal, weight = 1.f, frequency = 1.f;
    for (int i = 0; i < octaves; ++i) {
        px = px * lacunarity;
        py = py * lacunarity;
        weight = clamp01(signal * gain);
        signal = snoise2(perm, px, py);
        signal = fabsf(signal);
        signal = offset - signal;
        signal = signal * weight;
        const float exponent = 1.f / frequency;
        frequency = frequency * lacunarity;
        noise = noise + signal * exponent;
    }
    return noise * (1.f / (float)octaves);
}

// Terrain, cl/noise.cl:205-223.
float leven_terrain(const uint8_t* perm, float x, float z) {
    const float qx = x * (1.f / 2000.f);
    const float qy = z * (1.f / 2000.f);
    float ridged = 0.8f * ridged_multifractal(perm, 7, 2.114352f, 1.5241f, 1.f, qx, qy);
    ridged = clamp01(ridged);
    float billow = 0.6f * fbm2(perm, 4, 0.24f, 1.8754f, 0.433f, -4.33f * qx, 7.98f * qy);
    billow = (0.5f * billow) + 0.5f;
    float noise = billow * ridged;
    float b2 = 0.6f * fbm2(perm, 2, 0.63f, 2.2f, 0.15f, qx, qy);
    b2 = (b2 * 0.5f) + 0.5f;
    return noise + b2;
}

// DensityFunc (cl/noise.cl:225-268) selected by the field descriptor.
float density(const lvn_field_desc& f, const uint8_t* perm, float x, float y, float z) {
    switch (f.kind) {
    case LVN_FIELD_LEVEN_TERRAIN:
        return y - (900.f * leven_terrain(perm, x, z));
    case LVN_FIELD_TERRAIN_CAVES: {
        const float a = y - (f.height_amplitude *
                             fbm2(perm, f.height_octaves, f.height_frequency, f.height_lacunarity, f.height_gain, x, z));
        const float b =
            fbm3(perm, f.cave_octaves, f.cave_frequency, f.cave_lacunarity, f.cave_gain, x, y, z) - f.cave_threshold;
        return fmaxf(a, b);
    }
    default:
        return y - (f.height_amplitude *
                    fbm2(perm, f.height_octaves, f.height_frequency, f.height_lacunarity, f.height_gain, x, z));
    }
}

// FP64 twin of the density function (same formulas, double arithmetic), used only
// to state the error of FP32 normals.
double snoise2_d(const uint8_t* perm, double px, double py) {
    const double F2 = 0.366025403784, G2 = 0.211324865405;
    const double s = (px + py) * F2;
    const double pix = floor(px + s), piy = floor(py + s);
    const double t = (pix + piy) * G2;
    const double f0x = px - (pix - t), f0y = py - (piy - t);
    const int ix = (int)pix, iy = (int)piy;
    const int o1x = f0x > f0y ? 1 : 0, o1y = 1 - o1x;
    const double fx[3] = {f0x, (f0x - o1x) + G2, f0x - (1.0 - 2.0 * G2)};
    const double fy[3] = {f0y, (f0y - o1y) + G2, f0y - (1.0 - 2.0 * G2)};
    const int tx[3] = {ix, ix + o1x, ix + 1}, ty[3] = {iy, iy + o1y, iy + 1};
    double n = 0.0;
    for (int c = 0; c < 3; ++c) {
        const int g = texel(perm, tx[c], ty[c]);
        double tt = 0.5 - (fx[c] * fx[c] + fy[c] * fy[c]);
        if (tt < 0.0) continue;
        tt *= tt;
        n += tt * tt * ((double)kGrads.comp(g, 0) * fx[c] + (double)kGrads.comp(g, 1) * fy[c]);
    }
    return 70.0 * n;
}

double snoise3_d(const uint8_t* perm, double px, double py, double pz) {
    const double F3 = 0.333333333333, G3 = 0.166666666667;
    const double s = (px + py + pz) * F3;
    const double pix = floor(px + s), piy = floor(py + s), piz = floor(pz + s);
    const double t = (pix + piy + piz) * G3;
    const double f0x = px - (pix - t), f0y = py - (piy - t), f0z = pz - (piz - t);
    const int ix = (int)pix, iy = (int)piy, iz = (int)piz;
    const int isXy = f0x >= f0y, isXz = f0x >= f0z, isYz = f0y >= f0z;
    const int rx = isXy + isXz, ry = (1 - isXy) + isYz, rz = (1 - isXz) + (1 - isYz);
    const int o2[3] = {std::min(rx, 1), std::min(ry, 1), std::min(rz, 1)};
    const int o1[3] = {std::max(rx - 1, 0), std::max(ry - 1, 0), std::max(rz - 1, 0)};
    const double fx[4] = {f0x, (f0x - o1[0]) + G3, (f0x - o2[0]) + 2.0 * G3, f0x - (1.0 - 3.0 * G3)};
    const double fy[4] = {f0y, (f0y - o1[1]) + G3, (f0y - o2[1]) + 2.0 * G3, f0y - (1.0 - 3.0 * G3)};
    const double fz[4] = {f0z, (f0z - o1[2]) + G3, (f0z - o2[2]) + 2.0 * G3, f0z - (1.0 - 3.0 * G3)};
    const int tx[4] = {ix, ix + o1[0], ix + o2[0], ix + 1};
    const int ty[4] = {iy, iy + o1[1], iy + o2[1], iy + 1};
    const int tz[4] = {iz, iz + o1[2], iz + o2[2], iz + 1};
    double n = 0.0;
    for (int c = 0; c < 4; ++c) {
        const int p = texel(perm, tx[c], ty[c]);
        const int g = texel(perm, perm_as_texel(p), tz[c]);
        double tt = 0.6 - (fx[c] * fx[c] + fy[c] * fy[c] + fz[c] * fz[c]);
        if (tt < 0.0) continue;
        tt *= tt;
        n += tt * tt *
             ((double)kGrads.comp(g, 0) * fx[c] + (double)kGrads.comp(g, 1) * fy[c] + (double)kGrads.comp(g, 2) * fz[c]);
    }
    return 32.0 * n;
}

double fbm2_d(const uint8_t* perm, int oct, double freq, double lac, double gain, double x, double y) {
    double px = x * freq, py = y * freq, amp = 1.0, n = 0.0;
    for (int i = 0; i < oct; ++i) {
        n += snoise2_d(perm, px, py) * amp;
        px *= lac;
        py *= lac;
        amp *= gain;
    }
    return n;
}

double fbm3_d(const uint8_t* perm, int oct, double freq, double lac, double gain, double x, double y, double z) {
    double px = x * freq, py = y * freq, pz = z * freq, amp = 1.0, n = 0.0;
    for (int i = 0; i < oct; ++i) {
        n += snoise3_d(perm, px, py, pz) * amp;
        px *= lac;
        py *= lac;
        pz *= lac;
        amp *= gain;
    }
    return n;
}

// RidgedMultiFractal (cl/noise.cl:42-81) and Terrain (cl/noise.cl:205-223) in double.
double ridged_multifractal_d(const uint8_t* perm, int oct, double lac, double gain, double offset, double x,
                             double y) {
    double px = x, py = y;
    double signal = offset - std::fabs(snoise2_d(perm, px, py));
    signal *= signal;
    double noise = signal, frequency = 1.0;
    for (int i = 0; i < oct; ++i) {
        px *= lac;
        py *= lac;
        const double weight = std::min(std::max(signal * gain, 0.0), 1.0);
        signal = (offset - std::fabs(snoise2_d(perm, px, py))) * weight;
        noise += signal / frequency;
        frequency *= lac;
    }
    return noise / oct;
}

double leven_terrain_d(const uint8_t* perm, double x, double z) {
    const double qx = x / 2000.0, qy = z / 2000.0;
    const double ridged =
        std::min(std::max(0.8 * ridged_multifractal_d(perm, 7, 2.114352, 1.5241, 1.0, qx, qy), 0.0), 1.0);
    const double billow = 0.5 * (0.6 * fbm2_d(perm, 4, 0.24, 1.8754, 0.433, -4.33 * qx, 7.98 * qy)) + 0.5;
    const double b2 = 0.5 * (0.6 * fbm2_d(perm, 2, 0.63, 2.2, 0.15, qx, qy)) + 0.5;
    return billow * ridged + b2;
}

double density_d(const lvn_field_desc& f, const uint8_t* perm, double x, double y, double z) {
    if (f.kind == LVN_FIELD_LEVEN_TERRAIN) return y - 900.0 * leven_terrain_d(perm, x, z);
    const double a = y - (double)f.height_amplitude * fbm2_d(perm, f.height_octaves, f.height_frequency,
                                                              f.height_lacunarity, f.height_gain, x, z);
    if (f.kind == LVN_FIELD_TERRAIN_CAVES) {
        const double b = fbm3_d(perm, f.cave_octaves, f.cave_frequency, f.cave_lacunarity, f.cave_gain, x, y, z) -
                         (double)f.cave_threshold;
        return std::max(a, b);
    }
    return a;
}

// ---------------------------------------------------------------------------
// A7 — CSG brushes (cl/apply_csg_operation.cl)
// ---------------------------------------------------------------------------

inline float length3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

// Density_Cuboid with pR (rotation of the xz plane) and fBox (box SDF:
// length(max(d, 0)) + max component of min(d, 0), d = |p| - b), cl/apply_csg_operation.cl:49-68.
// cos/sin of the angle are evaluated on the host (the reference evaluates them per sample).
float brush_density(const lvn_csg_op& op, float c, float s, float x, float y, float z) {
    const float lx = x - op.origin[0];
    const float ly = y - op.origin[1];
    const float lz = z - op.origin[2];
    if (op.brush == 1) {  // Density_Sphere, cl/apply_csg_operation.cl:42-45
        return length3(lx, ly, lz) - op.dims[0];
    }
    const float rx = c * lx + s * lz;
    const float rz = c * lz + s * (-lx);
    const float dx = fabsf(rx) - op.dims[0];
    const float dy = fabsf(ly) - op.dims[1];
    const float dz = fabsf(rz) - op.dims[2];
    const float outside = length3(fmaxf(dx, 0.f), fmaxf(dy, 0.f), fmaxf(dz, 0.f));
    const float inside = fmaxf(fmaxf(fminf(dx, 0.f), fminf(dy, 0.f)), fminf(dz, 0.f));
    return outside + inside;
}

struct Brush {
    lvn_csg_op op;
    float c, s;
    float eval(float x, float y, float z) const { return brush_density(op, c, s, x, y, z); }
};

std::vector<Brush> make_brushes(const lvn_csg_op* ops, int n) {
    std::vector<Brush> out((size_t)std::max(n, 0));
    for (int i = 0; i < n; ++i) {
        out[i].op = ops[i];
        out[i].c = cosf(ops[i].rotate_y);
        out[i].s = sinf(ops[i].rotate_y);
    }
    return out;
}

struct Vec3 {
    float x, y, z;
};

struct Vec3d {
    double x, y, z;
};

// normalize (see lvn_oracle.h): zero vectors are returned unchanged.
Vec3 normalize3(Vec3 v) {
    if (v.x == 0.f && v.y == 0.f && v.z == 0.f) return v;
    const float len = length3(v.x, v.y, v.z);
    return {v.x / len, v.y / len, v.z / len};
}

// ---------------------------------------------------------------------------
// Chunk geometry (src/compute.cpp:245-278)
// ---------------------------------------------------------------------------

struct Geometry {
    int N = 0, H = 0, F = 0, shift = 0, mask = 0, depth = 0;
    explicit Geometry(int n = 64) {
        N = n;
        H = n + 1;
        F = n + 2;
        depth = 0;
        while ((1 << depth) < n) ++depth;
        shift = depth + 1;
        mask = (1 << shift) - 1;
    }
};

const int kChild[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                          {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};   // cl/shared_constants.cl:12-23
const int kEdgeMap[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                             {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};  // cl/shared_constants.cl:4-9
const int kEdgeNodeOffsets[3][4][3] = {  // cl/octree.cl:376-381
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}},
    {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}},
    {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0}}};

uint32_t code_for_position(int x, int y, int z, int node_depth, int max_depth) {  // cl/octree.cl:34-47
    uint32_t code = 1;
    for (int d = max_depth - 1; d >= max_depth - node_depth; --d) {
        const uint32_t c = (uint32_t)((((x >> d) & 1) << 2) | (((y >> d) & 1) << 1) | ((z >> d) & 1));
        code = (code << 3) | c;
    }
    return code;
}

void position_for_code(uint32_t code, int max_depth, int* xyz) {  // cl/octree.cl:51-72
    int msb = 0;
    for (uint32_t c = code; c; c >>= 1) ++msb;
    const int node_depth = msb / 3;
    xyz[0] = xyz[1] = xyz[2] = 0;
    for (int i = max_depth - node_depth; i < max_depth; ++i) {
        const uint32_t c = code & 7u;
        code >>= 3;
        xyz[0] |= (int)((c >> 2) & 1u) << i;
        xyz[1] |= (int)((c >> 1) & 1u) << i;
        xyz[2] |= (int)(c & 1u) << i;
    }
}

// FindDominantMaterial with its insertion sort, cl/octree.cl:78-138.
int dominant_material(const int m[8]) {
    int d[8];
    for (int i = 0; i < 8; ++i) d[i] = m[i];
    for (int i = 1; i < 8; ++i) {
        const int tmp = d[i];
        int j = i;
        for (; j >= 1 && tmp < d[j - 1]; --j) d[j] = d[j - 1];
        d[j] = tmp;
    }
    int current = d[0], count = 1, maxCount = 0, maxMaterial = 0;
    for (int i = 1; i < 8; ++i) {
        const int v = d[i];
        if (v == kAir || v == kNone) continue;
        if (current != v) {
            if (count > maxCount) {
                maxCount = count;
                maxMaterial = current;
            }
            current = v;
            count = 1;
        } else {
            ++count;
        }
    }
    if (count > maxCount) maxMaterial = current;
    return maxMaterial;
}

// ---------------------------------------------------------------------------
// A10/A11 — QEF (cl/qef.cl, cl/octree.cl:236-331), templated on the arithmetic type
// ---------------------------------------------------------------------------

template <typename T>
struct V4 {
    T x = 0, y = 0, z = 0, w = 0;
};

template <typename T>
struct Qef {
    T ata[6] = {0, 0, 0, 0, 0, 0};
    V4<T> atb, mass;
};

template <typename T>
inline T dot4(const V4<T>& a, const V4<T>& b) {
    return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}

template <typename T>
void qef_add_point(Qef<T>& q, const V4<T>& n, const V4<T>& p) {  // cl/qef.cl:170-191
    q.ata[0] = q.ata[0] + n.x * n.x;
    q.ata[1] = q.ata[1] + n.x * n.y;
    q.ata[2] = q.ata[2] + n.x * n.z;
    q.ata[3] = q.ata[3] + n.y * n.y;
    q.ata[4] = q.ata[4] + n.y * n.z;
    q.ata[5] = q.ata[5] + n.z * n.z;
    const T b = dot4(p, n);
    q.atb.x = q.atb.x + n.x * b;
    q.atb.y = q.atb.y + n.y * b;
    q.atb.z = q.atb.z + n.z * b;
    q.mass.x = q.mass.x + p.x;
    q.mass.y = q.mass.y + p.y;
    q.mass.z = q.mass.z + p.z;
    q.mass.w = q.mass.w + (T)1;
}

template <typename T>
void qef_finish_mass(Qef<T>& q) {  // qef_create_from_points tail, cl/qef.cl:302
    const T w = q.mass.w;
    q.mass.x = q.mass.x / w;
    q.mass.y = q.mass.y / w;
    q.mass.z = q.mass.z / w;
    q.mass.w = q.mass.w / w;
}

template <typename T>
void givens_coeffs_sym(T app, T apq, T aqq, T& c, T& s) {  // cl/qef.cl:31-42
    if (apq == (T)0) {
        c = 1;
        s = 0;
        return;
    }
    const T tau = (aqq - app) / ((T)2 * apq);
    const T stt = std::sqrt((T)1 + tau * tau);
    const T tn = (T)1 / ((tau >= (T)0) ? (tau + stt) : (tau - stt));
    c = (T)1 / std::sqrt((T)1 + tn * tn);
    s = tn * c;
}

template <typename T>
void svd_rotate(T vtav[3][3], T v[3][3], int a, int b) {  // cl/qef.cl:44-86
    if (vtav[a][b] == (T)0) return;
    T c, s;
    givens_coeffs_sym(vtav[a][a], vtav[a][b], vtav[b][b], c, s);
    {
        const T cc = c * c, ss = s * s;
        const T mx = (((T)2 * c) * s) * vtav[a][b];
        const T u = vtav[a][a], w = vtav[b][b];
        vtav[a][a] = (cc * u - mx) + ss * w;
        vtav[b][b] = (ss * u + mx) + cc * w;
    }
    {
        const T u = vtav[0][3 - b], w = vtav[1 - a][2];
        vtav[0][3 - b] = c * u - s * w;
        vtav[1 - a][2] = s * u + c * w;
    }
    vtav[a][b] = 0;
    for (int r = 0; r < 3; ++r) {
        const T u = v[r][a], w = v[r][b];
        v[r][a] = c * u - s * w;
        v[r][b] = s * u + c * w;
    }
}

template <typename T>
T svd_invdet(T x, T tol) {  // cl/qef.cl:107-109
    return (std::fabs(x) < tol || std::fabs((T)1 / x) < tol) ? (T)0 : (T)1 / x;
}

// qef_solve (cl/qef.cl:239-256) followed by SolveQEFs' world transform is applied by the caller.
template <typename T>
V4<T> qef_solve(Qef<T> q) {
    const T mw = std::max(q.mass.w, (T)1);
    q.mass.x = q.mass.x / mw;
    q.mass.y = q.mass.y / mw;
    q.mass.z = q.mass.z / mw;
    q.mass.w = q.mass.w / mw;

    V4<T> amp;  // svd_vmul_sym, cl/qef.cl:146-152 (w untouched)
    {
        V4<T> row;
        row.x = q.ata[0];
        row.y = q.ata[1];
        row.z = q.ata[2];
        amp.x = dot4(row, q.mass);
        amp.y = (q.ata[1] * q.mass.x + q.ata[3] * q.mass.y) + q.ata[4] * q.mass.z;
        amp.z = (q.ata[2] * q.mass.x + q.ata[4] * q.mass.y) + q.ata[5] * q.mass.z;
    }
    amp.x = q.atb.x - amp.x;
    amp.y = q.atb.y - amp.y;
    amp.z = q.atb.z - amp.z;
    amp.w = q.atb.w - amp.w;

    // svd_solve_ATA_ATb, cl/qef.cl:127-144; svd_solve_sym :88-105 (10 sweeps).
    T vtav[3][3] = {{q.ata[0], q.ata[1], q.ata[2]}, {0, q.ata[3], q.ata[4]}, {0, 0, q.ata[5]}};
    T v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 10; ++i) {
        svd_rotate(vtav, v, 0, 1);
        svd_rotate(vtav, v, 0, 2);
        svd_rotate(vtav, v, 1, 2);
    }
    const T tol = (T)0.1f;  // PSUEDO_INVERSE_THRESHOLD, cl/qef.cl:21
    const T d[3] = {svd_invdet(vtav[0][0], tol), svd_invdet(vtav[1][1], tol), svd_invdet(vtav[2][2], tol)};
    T o[3][3];  // svd_pseudoinverse, cl/qef.cl:111-125
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[i][j] = ((v[i][0] * d[0]) * v[j][0] + (v[i][1] * d[1]) * v[j][1]) + (v[i][2] * d[2]) * v[j][2];
    V4<T> x;  // svd_mul_matrix_vec, cl/qef.cl:23-29
    for (int i = 0; i < 3; ++i) {
        V4<T> row;
        row.x = o[i][0];
        row.y = o[i][1];
        row.z = o[i][2];
        const T r = dot4(row, amp);
        if (i == 0) x.x = r;
        if (i == 1) x.y = r;
        if (i == 2) x.z = r;
    }
    x.w = 0;
    x.x = x.x + q.mass.x;
    x.y = x.y + q.mass.y;
    x.z = x.z + q.mass.z;
    x.w = x.w + q.mass.w;
    return x;
}

// ColourForMinLeafSize (src/clipmap.cpp:329-352), called with size / CLIPMAP_LEAF_SIZE
// (src/compute_octree.cpp:253).
void lod_colour(int size, float* rgb) {
    switch (size / kClipmapLeafSize) {
    case 1: rgb[0] = 0.3f; rgb[1] = 0.1f; rgb[2] = 0.f; break;
    case 2: rgb[0] = 0.f; rgb[1] = 0.f; rgb[2] = 0.5f; break;
    case 4: rgb[0] = 0.f; rgb[1] = 0.5f; rgb[2] = 0.5f; break;
    case 8: rgb[0] = 0.5f; rgb[1] = 0.f; rgb[2] = 0.5f; break;
    case 16: rgb[0] = 0.f; rgb[1] = 0.5f; rgb[2] = 0.f; break;
    default: rgb[0] = 0.5f; rgb[1] = 0.f; rgb[2] = 0.f; break;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Chunk state
// ---------------------------------------------------------------------------

struct orc_chunk {
    Geometry g;
    lvn_field_desc desc{};
    std::vector<uint8_t> perm;
    int min[3] = {0, 0, 0};
    int size = 0, scale = 1;
    int off[3] = {0, 0, 0};
    uint32_t defmat = 0;

    std::vector<float> density;   // F^3
    std::vector<uint8_t> mat;     // F^3
    std::vector<int32_t> edgeIds; // ascending
    std::vector<float> edgeData;  // 4 per edge
    std::vector<uint8_t> edgeBrush;  // 1 when the edge took brush data (CSG-touched)

    std::vector<int32_t> vcode, vmask, vmat;
    std::vector<float> vpos, vnorm;
    std::vector<double> vpos64, vnorm64;
    std::vector<int32_t> tris;
    int quads = 0;
    std::vector<lvn_vertex> verts;
    std::vector<lvn_seam_node> seams;

    explicit orc_chunk(int n) : g(n) {}

    int fi(int x, int y, int z) const { return x + g.F * (y + g.F * z); }
    float wx(int x) const { return (float)(x * scale + off[0]); }
    float wy(int y) const { return (float)(y * scale + off[1]); }
    float wz(int z) const { return (float)(z * scale + off[2]); }
    float D(float x, float y, float z) const { return ::density(desc, perm.data(), x, y, z); }

    void setup(const lvn_field_desc* d, const uint8_t* p, const int32_t* mn, int32_t sz, uint32_t dm) {
        desc = *d;
        perm.assign(p, p + 65536);
        for (int i = 0; i < 3; ++i) {
            min[i] = mn[i];
            off[i] = mn[i] / kLeafSizeScale;  // LeafScaleVec, src/compute.cpp:560-568 (C truncation)
        }
        size = sz;
        scale = sz / (g.N * kLeafSizeScale);  // src/compute_density_field.cpp:149
        defmat = dm;
    }

    // FindEdgeIntersectionInfo, cl/density_field.cl:96-151.
    void density_edge(int x, int y, int z, int axis, float* out) const {
        const int e[3] = {axis == 0, axis == 1, axis == 2};
        const int w0[3] = {x * scale + off[0], y * scale + off[1], z * scale + off[2]};
        const float p0[3] = {(float)w0[0], (float)w0[1], (float)w0[2]};
        const float p1[3] = {(float)(w0[0] + scale * e[0]), (float)(w0[1] + scale * e[1]),
                             (float)(w0[2] + scale * e[2])};
        float minValue = 3.402823466e+38f, currentT = 0.f, t = 0.f;
        for (int i = 0; i <= 16; ++i) {
            const float px = p0[0] + (p1[0] - p0[0]) * currentT;
            const float py = p0[1] + (p1[1] - p0[1]) * currentT;
            const float pz = p0[2] + (p1[2] - p0[2]) * currentT;
            const float d = fabsf(D(px, py, pz));
            if (d < minValue) {
                t = currentT;
                minValue = d;
            }
            currentT = currentT + 0.0625f;
        }
        const float px = p0[0] + (p1[0] - p0[0]) * t;
        const float py = p0[1] + (p1[1] - p0[1]) * t;
        const float pz = p0[2] + (p1[2] - p0[2]) * t;
        const float h = 0.001f;
        const float dx = D(px + h, py, pz) - D(px - h, py, pz);
        const float dy = D(px, py + h, pz) - D(px, py - h, pz);
        const float dz = D(px, py, pz + h) - D(px, py, pz - h);
        const Vec3 n = normalize3({dx, dy, dz});
        out[0] = n.x;
        out[1] = n.y;
        out[2] = n.z;
        out[3] = t;
    }

    // FindEdgeIntersectionInfo of the CSG program, cl/apply_csg_operation.cl:443-477,
    // with BrushZeroCrossing :86-110 and BrushNormal :144-175.
    void brush_edge(int x, int y, int z, int axis, const std::vector<Brush>& brushes, float* out) const {
        const int e[3] = {axis == 0, axis == 1, axis == 2};
        const int w0[3] = {x * scale + off[0], y * scale + off[1], z * scale + off[2]};
        const float p0[3] = {(float)w0[0], (float)w0[1], (float)w0[2]};
        const float p1[3] = {(float)(w0[0] + scale * e[0]), (float)(w0[1] + scale * e[1]),
                             (float)(w0[2] + scale * e[2])};
        float minDensity = 3.402823466e+38f, crossing = 0.f;
        for (float t = 0.f; t <= 1.f; t = t + (1.f / 16.f)) {
            const float px = p0[0] + (p1[0] - p0[0]) * t;
            const float py = p0[1] + (p1[1] - p0[1]) * t;
            const float pz = p0[2] + (p1[2] - p0[2]) * t;
            for (const Brush& b : brushes) {
                const float d = fabsf(b.eval(px, py, pz));
                if (d < minDensity) {
                    crossing = t;
                    minDensity = d;
                }
            }
        }
        const float px = p0[0] + (p1[0] - p0[0]) * crossing;
        const float py = p0[1] + (p1[1] - p0[1]) * crossing;
        const float pz = p0[2] + (p1[2] - p0[2]) * crossing;
        Vec3 normal = {0.f, 0.f, 0.f};
        const float h = 0.001f;
        for (const Brush& b : brushes) {
            if (b.eval(px, py, pz) > 0.f) continue;
            const float dx = b.eval(px + h, py, pz) - b.eval(px - h, py, pz);
            const float dy = b.eval(px, py + h, pz) - b.eval(px, py - h, pz);
            const float dz = b.eval(px, py, pz + h) - b.eval(px, py, pz - h);
            const float flip = b.op.type == 0 ? 1.f : -1.f;
            const Vec3 n = normalize3({dx, dy, dz});
            normal = {flip * n.x, flip * n.y, flip * n.z};
        }
        out[0] = normal.x;
        out[1] = normal.y;
        out[2] = normal.z;
        out[3] = crossing;
    }

    // FP64 twin of density_edge's normal at the same crossing t: central differences
    // with h = 0.001 of the FP64 density, normalised in FP64.
    Vec3d density_normal_f64(int x, int y, int z, int axis, float t) const {
        const double p0[3] = {(double)(x * scale + off[0]), (double)(y * scale + off[1]), (double)(z * scale + off[2])};
        double p[3] = {p0[0], p0[1], p0[2]};
        p[axis] += (double)scale * (double)t;
        const double h = 0.001;
        const double dx = density_d(desc, perm.data(), p[0] + h, p[1], p[2]) - density_d(desc, perm.data(), p[0] - h, p[1], p[2]);
        const double dy = density_d(desc, perm.data(), p[0], p[1] + h, p[2]) - density_d(desc, perm.data(), p[0], p[1] - h, p[2]);
        const double dz = density_d(desc, perm.data(), p[0], p[1], p[2] + h) - density_d(desc, perm.data(), p[0], p[1], p[2] - h);
        const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (len == 0.0) return {0.0, 0.0, 0.0};
        return {dx / len, dy / len, dz / len};
    }

    // FP64 twin of density_edge's crossing: the first arg-min of |d| over the 17 samples
    // t = k / 16 of the FP64 density.
    double crossing_f64(int x, int y, int z, int axis) const {
        const double p0[3] = {(double)(x * scale + off[0]), (double)(y * scale + off[1]), (double)(z * scale + off[2])};
        double best = HUGE_VAL, t = 0.0;
        for (int i = 0; i <= 16; ++i) {
            double p[3] = {p0[0], p0[1], p0[2]};
            p[axis] += (double)scale * (i / 16.0);
            const double d = std::fabs(density_d(desc, perm.data(), p[0], p[1], p[2]));
            if (d < best) {
                best = d;
                t = i / 16.0;
            }
        }
        return t;
    }

    Vec3d density_normal_f64_at(int x, int y, int z, int axis, double t) const {
        const double p0[3] = {(double)(x * scale + off[0]), (double)(y * scale + off[1]), (double)(z * scale + off[2])};
        double p[3] = {p0[0], p0[1], p0[2]};
        p[axis] += (double)scale * t;
        const double h = 0.001;
        const double dx = density_d(desc, perm.data(), p[0] + h, p[1], p[2]) - density_d(desc, perm.data(), p[0] - h, p[1], p[2]);
        const double dy = density_d(desc, perm.data(), p[0], p[1] + h, p[2]) - density_d(desc, perm.data(), p[0], p[1] - h, p[2]);
        const double dz = density_d(desc, perm.data(), p[0], p[1], p[2] + h) - density_d(desc, perm.data(), p[0], p[1], p[2] - h);
        const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (len == 0.0) return {0.0, 0.0, 0.0};
        return {dx / len, dy / len, dz / len};
    }

    // The whole vertex computation in FP64 on this chunk's topology: every density edge's
    // crossing and normal from the FP64 field (brush edges keep their data), then the QEF
    // in FP64.  pos4 / normal4 per vertex; *moved = edges whose FP64 crossing picks another
    // of the 17 samples than the FP32 one.
    void fp64_pipeline(double* pos4, double* normal4, int32_t* moved) const {
        const int H = g.H;
        std::vector<int32_t> slot((size_t)H * H * H * 3, -1);
        for (size_t k = 0; k < edgeIds.size(); ++k) {
            const int id = edgeIds[k], a = id & 3, h = id >> 2;
            const int hx = h & g.mask, hy = (h >> g.shift) & g.mask, hz = (h >> (2 * g.shift)) & g.mask;
            slot[(((size_t)hz * H + hy) * H + hx) * 3 + a] = (int32_t)k;
        }
        std::vector<double> et(edgeIds.size(), -1.0);
        std::vector<Vec3d> en(edgeIds.size());
        int32_t changed = 0;
        for (size_t k = 0; k < edgeIds.size(); ++k) {
            const int id = edgeIds[k], a = id & 3, h = id >> 2;
            const int hx = h & g.mask, hy = (h >> g.shift) & g.mask, hz = (h >> (2 * g.shift)) & g.mask;
            const float* ed = &edgeData[k * 4];
            if (edgeBrush[k]) {
                et[k] = ed[3];
                en[k] = Vec3d{ed[0], ed[1], ed[2]};
            } else {
                et[k] = crossing_f64(hx, hy, hz, a);
                en[k] = density_normal_f64_at(hx, hy, hz, a, et[k]);
                changed += et[k] != (double)ed[3];
            }
        }
        if (moved) *moved = changed;
        const int V = (int)vcode.size();
        for (int v = 0; v < V; ++v) {
            int p[3];
            position_for_code((uint32_t)vcode[v], g.depth, p);
            Qef<double> q;
            double n[3] = {0.0, 0.0, 0.0};
            double count = 0.0;
            for (int i = 0; i < 12; ++i) {
                if (!((vmask[v] >> i) & 1)) continue;
                const int e0 = kEdgeMap[i][0], axis = i / 4;
                const int hx = p[0] + kChild[e0][0], hy = p[1] + kChild[e0][1], hz = p[2] + kChild[e0][2];
                const int k = slot[(((size_t)hz * H + hy) * H + hx) * 3 + axis];
                if (k < 0) continue;
                V4<double> pd, nd;
                const double a0[3] = {(double)hx, (double)hy, (double)hz};
                pd.x = (double)scale * (a0[0] + (axis == 0 ? et[k] : 0.0));
                pd.y = (double)scale * (a0[1] + (axis == 1 ? et[k] : 0.0));
                pd.z = (double)scale * (a0[2] + (axis == 2 ? et[k] : 0.0));
                pd.w = 0.0;
                nd.x = en[k].x;
                nd.y = en[k].y;
                nd.z = en[k].z;
                nd.w = 0.0;
                qef_add_point(q, nd, pd);
                n[0] += en[k].x;
                n[1] += en[k].y;
                n[2] += en[k].z;
                count += 1.0;
            }
            qef_finish_mass(q);
            const V4<double> x = qef_solve(q);
            pos4[4 * v + 0] = x.x * 4.0 + (double)min[0];
            pos4[4 * v + 1] = x.y * 4.0 + (double)min[1];
            pos4[4 * v + 2] = x.z * 4.0 + (double)min[2];
            pos4[4 * v + 3] = 1.0;
            for (int c = 0; c < 3; ++c) normal4[4 * v + c] = count > 0 ? n[c] / count : 0.0;
            normal4[4 * v + 3] = 0.0;
        }
    }

    bool solid(int idx) const { return mat[idx] != kAir; }

    // Sign-changing edges of the hermite grid in ascending id order
    // (FindFieldEdges + scan + CompactEdges, cl/density_field.cl:41-92).
    template <typename DataFn>
    void rebuild_edges(DataFn data_for) {
        std::vector<int32_t> ids;
        std::vector<float> data;
        std::vector<uint8_t> brush;
        for (int z = 0; z < g.H; ++z)
            for (int y = 0; y < g.H; ++y)
                for (int x = 0; x < g.H; ++x) {
                    const bool s0 = solid(fi(x, y, z));
                    const int nb[3] = {fi(x + 1, y, z), fi(x, y + 1, z), fi(x, y, z + 1)};
                    for (int a = 0; a < 3; ++a) {
                        if (s0 == solid(nb[a])) continue;
                        const int32_t id = ((x | (y << g.shift) | (z << (2 * g.shift))) << 2) | a;
                        float d4[4];
                        brush.push_back(data_for(x, y, z, a, id, d4) ? 1 : 0);
                        ids.push_back(id);
                        data.insert(data.end(), d4, d4 + 4);
                    }
                }
        edgeIds.swap(ids);
        edgeData.swap(data);
        edgeBrush.swap(brush);
    }

    void default_materials() {
        mat.assign(density.size(), 0);
        for (size_t i = 0; i < density.size(); ++i) mat[i] = density[i] < 0.f ? (uint8_t)defmat : (uint8_t)kAir;
    }

    // CSG_HermiteIndices / BrushMaterial (cl/apply_csg_operation.cl:114-140,179-210):
    // material override in op order; returns which samples changed.
    std::vector<uint8_t> apply_brush_materials(const std::vector<Brush>& brushes) {
        std::vector<uint8_t> touched(mat.size(), 0);
        if (brushes.empty()) return touched;
        for (int z = 0; z < g.F; ++z)
            for (int y = 0; y < g.F; ++y)
                for (int x = 0; x < g.F; ++x) {
                    const int i = fi(x, y, z);
                    int m = mat[i];
                    const float px = wx(x), py = wy(y), pz = wz(z);
                    for (const Brush& b : brushes)
                        if (b.eval(px, py, pz) <= 0.f) m = b.op.type == 0 ? b.op.material : kAir;
                    touched[i] = m != mat[i];
                    mat[i] = (uint8_t)m;
                }
        return touched;
    }

    bool edge_touched(const std::vector<uint8_t>& touched, int x, int y, int z, int a) const {
        const int e[3] = {a == 0, a == 1, a == 2};
        return touched[fi(x, y, z)] || touched[fi(x + e[0], y + e[1], z + e[2])];
    }

    // Fresh field + one CSG batch.  The edge set after ApplyCSGOperations
    // (src/compute_csg.cpp:11-220) is every sign-changing edge of the final
    // materials; an edge with a changed endpoint takes brush data, every other
    // edge keeps the density data.  Order: ascending id (the reference appends
    // re-created edges after the survivors in a hash-race order; the set is equal).
    void generate(const std::vector<Brush>& brushes, const float* given_density) {
        const int F = g.F;
        density.assign((size_t)F * F * F, 0.f);
        for (int z = 0; z < F; ++z)
            for (int y = 0; y < F; ++y)
                for (int x = 0; x < F; ++x) {
                    const int i = fi(x, y, z);
                    density[i] = given_density ? given_density[i] : D(wx(x), wy(y), wz(z));  // cl/density_field.cl:11-37
                }
        default_materials();
        const std::vector<uint8_t> touched = apply_brush_materials(brushes);
        rebuild_edges([&](int x, int y, int z, int a, int32_t, float* d4) {
            if (!brushes.empty() && edge_touched(touched, x, y, z, a)) {
                brush_edge(x, y, z, a, brushes, d4);
                return true;
            }
            density_edge(x, y, z, a, d4);
            return false;
        });
    }

    void apply_csg(const std::vector<Brush>& brushes) {
        if (brushes.empty()) return;
        const std::vector<int32_t> oldIds = edgeIds;
        const std::vector<float> oldData = edgeData;
        const std::vector<uint8_t> oldBrush = edgeBrush;
        const std::vector<uint8_t> touched = apply_brush_materials(brushes);
        rebuild_edges([&](int x, int y, int z, int a, int32_t id, float* d4) {
            if (edge_touched(touched, x, y, z, a)) {
                brush_edge(x, y, z, a, brushes, d4);
                return true;
            }
            const auto it = std::lower_bound(oldIds.begin(), oldIds.end(), id);
            const size_t k = (size_t)(it - oldIds.begin());
            for (int c = 0; c < 4; ++c) d4[c] = oldData[4 * k + c];
            return oldBrush[k] != 0;
        });
    }

    // ConstructOctreeFromField, GenerateMeshFromOctree, GatherSeamNodesFromOctree
    // (src/compute_octree.cpp:25-150,189-322) with the kernels of cl/octree.cl.
    void build() {
        vcode.clear();
        vmask.clear();
        vmat.clear();
        vpos.clear();
        vnorm.clear();
        vpos64.clear();
        vnorm64.clear();
        tris.clear();
        verts.clear();
        seams.clear();
        quads = 0;
        if (edgeIds.empty()) return;  // src/compute_octree.cpp:39-44, LoadOctree :167-171
        const int N = g.N, H = g.H;

        // FindActiveVoxels + CompactVoxels, cl/octree.cl:142-223.
        std::vector<int32_t> voxelIndex((size_t)N * N * N, -1);
        for (int z = 0; z < N; ++z)
            for (int y = 0; y < N; ++y)
                for (int x = 0; x < N; ++x) {
                    int cm[8], corners = 0, edges = 0;
                    for (int i = 0; i < 8; ++i) {
                        cm[i] = mat[fi(x + kChild[i][0], y + kChild[i][1], z + kChild[i][2])];
                        corners |= (cm[i] == kAir ? 0 : 1) << i;
                    }
                    for (int i = 0; i < 12; ++i) {
                        const int a = (corners >> kEdgeMap[i][0]) & 1, b = (corners >> kEdgeMap[i][1]) & 1;
                        edges |= (a != b) << i;
                    }
                    if (corners == 0 || corners == 255) continue;
                    voxelIndex[x + N * (y + N * z)] = (int32_t)vcode.size();
                    vcode.push_back((int32_t)code_for_position(x, y, z, g.depth, g.depth));
                    vmask.push_back(edges);
                    vmat.push_back((dominant_material(cm) << 8) | corners);
                }
        const int V = (int)vcode.size();
        if (V == 0) return;  // src/compute_octree.cpp:67-75

        // Dense edge lookup (replaces the cuckoo table of src/compute_octree.cpp:107-109).
        std::vector<int32_t> slot((size_t)H * H * H * 3, -1);
        for (size_t k = 0; k < edgeIds.size(); ++k) {
            const int id = edgeIds[k], a = id & 3, h = id >> 2;
            const int hx = h & g.mask, hy = (h >> g.shift) & g.mask, hz = (h >> (2 * g.shift)) & g.mask;
            slot[(((size_t)hz * H + hy) * H + hx) * 3 + a] = (int32_t)k;
        }

        vpos.assign((size_t)V * 4, 0.f);
        vnorm.assign((size_t)V * 4, 0.f);
        vpos64.assign((size_t)V * 4, 0.0);
        vnorm64.assign((size_t)V * 4, 0.0);
        const float fscale = (float)scale;
        for (int v = 0; v < V; ++v) {  // CreateLeafNodes + SolveQEFs, cl/octree.cl:236-331
            int p[3];
            position_for_code((uint32_t)vcode[v], g.depth, p);
            Qef<float> q;
            Qef<double> q64;
            V4<float> normal;
            double n64[3] = {0.0, 0.0, 0.0};
            float count = 0.f;
            for (int i = 0; i < 12; ++i) {
                if (!((vmask[v] >> i) & 1)) continue;
                const int e0 = kEdgeMap[i][0], e1 = kEdgeMap[i][1], axis = i / 4;
                const int hx = p[0] + kChild[e0][0], hy = p[1] + kChild[e0][1], hz = p[2] + kChild[e0][2];
                const int k = slot[(((size_t)hz * H + hy) * H + hx) * 3 + axis];
                if (k < 0) continue;
                const float* ed = &edgeData[(size_t)k * 4];
                const float t = ed[3];
                V4<float> pos, nrm;
                const float a0[3] = {(float)(p[0] + kChild[e0][0]), (float)(p[1] + kChild[e0][1]),
                                     (float)(p[2] + kChild[e0][2])};
                const float a1[3] = {(float)(p[0] + kChild[e1][0]), (float)(p[1] + kChild[e1][1]),
                                     (float)(p[2] + kChild[e1][2])};
                pos.x = fscale * (a0[0] + (a1[0] - a0[0]) * t);
                pos.y = fscale * (a0[1] + (a1[1] - a0[1]) * t);
                pos.z = fscale * (a0[2] + (a1[2] - a0[2]) * t);
                pos.w = fscale * (0.f + (0.f - 0.f) * t);
                nrm.x = ed[0];
                nrm.y = ed[1];
                nrm.z = ed[2];
                nrm.w = 0.f;
                qef_add_point(q, nrm, pos);
                V4<double> pd, nd;
                pd.x = pos.x; pd.y = pos.y; pd.z = pos.z; pd.w = pos.w;
                nd.x = nrm.x; nd.y = nrm.y; nd.z = nrm.z; nd.w = 0.0;
                qef_add_point(q64, nd, pd);
                normal.x = normal.x + nrm.x;
                normal.y = normal.y + nrm.y;
                normal.z = normal.z + nrm.z;
                normal.w = (normal.w + nrm.w) + 1.f;
                count += 1.f;
                const Vec3d e64 = edgeBrush[k] ? Vec3d{ed[0], ed[1], ed[2]} : density_normal_f64(hx, hy, hz, axis, t);
                n64[0] += e64.x;
                n64[1] += e64.y;
                n64[2] += e64.z;
            }
            qef_finish_mass(q);
            qef_finish_mass(q64);
            const float nw = normal.w;
            vnorm[4 * v + 0] = normal.x / nw;
            vnorm[4 * v + 1] = normal.y / nw;
            vnorm[4 * v + 2] = normal.z / nw;
            vnorm[4 * v + 3] = 0.f;
            for (int c = 0; c < 3; ++c) vnorm64[4 * v + c] = n64[c] / (double)count;
            const V4<float> x = qef_solve(q);
            vpos[4 * v + 0] = (x.x * 4.f) + (float)min[0];
            vpos[4 * v + 1] = (x.y * 4.f) + (float)min[1];
            vpos[4 * v + 2] = (x.z * 4.f) + (float)min[2];
            vpos[4 * v + 3] = 1.f;
            const V4<double> x64 = qef_solve(q64);
            vpos64[4 * v + 0] = x64.x * 4.0 + (double)min[0];
            vpos64[4 * v + 1] = x64.y * 4.0 + (double)min[1];
            vpos64[4 * v + 2] = x64.z * 4.0 + (double)min[2];
            vpos64[4 * v + 3] = 1.0;
            (void)count;
        }

        // GenerateMesh + CompactMeshTriangles, cl/octree.cl:335-464.
        for (int v = 0; v < V; ++v) {
            int p[3];
            position_for_code((uint32_t)vcode[v], g.depth, p);
            for (int axis = 0; axis < 3; ++axis) {
                const int a = p[(axis + 1) % 3], b = p[(axis + 2) % 3];
                if (a == N - 1 || b == N - 1) continue;
                int nodes[4] = {v, -1, -1, -1};
                bool ok = true;
                for (int n = 1; n < 4; ++n) {
                    const int x = p[0] + kEdgeNodeOffsets[axis][n][0];
                    const int y = p[1] + kEdgeNodeOffsets[axis][n][1];
                    const int z = p[2] + kEdgeNodeOffsets[axis][n][2];
                    nodes[n] = voxelIndex[x + N * (y + N * z)];
                    ok = ok && nodes[n] >= 0;
                }
                if (!ok) continue;
                const int edge = axis * 4 + 3;
                const int corners = vmat[v] & 0xff;
                const int m1 = (corners >> kEdgeMap[edge][0]) & 1, m2 = (corners >> kEdgeMap[edge][1]) & 1;
                if (m1 == m2) continue;
                static const int kWinding[2][6] = {{0, 1, 3, 0, 3, 2}, {0, 3, 1, 0, 2, 3}};
                const int flip = m1 != 0 ? 1 : 0;
                for (int k = 0; k < 6; ++k) tris.push_back(nodes[kWinding[flip][k]]);
                ++quads;
            }
        }

        // GenerateMeshVertexBuffer, cl/octree.cl:475-487; exported only when triangles exist
        // (src/compute_octree.cpp:227-232 returns before the vertex count is set).
        if (quads > 0) {
            float rgb[3];
            lod_colour(size, rgb);
            verts.resize((size_t)V);
            for (int v = 0; v < V; ++v) {
                for (int c = 0; c < 4; ++c) {
                    verts[v].pos[c] = vpos[4 * v + c];
                    verts[v].normal[c] = vnorm[4 * v + c];
                }
                verts[v].colour[0] = rgb[0];
                verts[v].colour[1] = rgb[1];
                verts[v].colour[2] = rgb[2];
                verts[v].colour[3] = (float)(vmat[v] >> 8);
            }
        }

        // FindSeamNodes + ExtractSeamNodeInfo, cl/octree.cl:506-551.
        for (int v = 0; v < V; ++v) {
            int p[3];
            position_for_code((uint32_t)vcode[v], g.depth, p);
            bool seam = false;
            for (int c = 0; c < 3; ++c) seam = seam || p[c] == 0 || p[c] == N - 1;
            if (!seam) continue;
            lvn_seam_node s;
            s.local_min_and_material[0] = p[0];
            s.local_min_and_material[1] = p[1];
            s.local_min_and_material[2] = p[2];
            s.local_min_and_material[3] = vmat[v];
            for (int c = 0; c < 4; ++c) {
                s.pos[c] = vpos[4 * v + c];
                s.normal[c] = vnorm[4 * v + c];
            }
            seams.push_back(s);
        }
    }
};

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------

extern "C" {

void orc_base_perm512(int32_t* out) { base_perm512(out); }
void orc_shuffled_perm512(int32_t seed, int32_t* out) { shuffled_perm512(seed, out); }
uint32_t orc_noise_hash(int32_t x, int32_t y, int32_t seed) { return noise_hash(x, y, seed); }
uint32_t orc_oat_hash(const uint8_t* bytes, int32_t n) { return oat_hash(bytes, n); }
void orc_splitmix64(uint64_t state, int32_t n, uint64_t* out) {
    for (int i = 0; i < n; ++i) out[i] = splitmix64(state);
}
void orc_grad_values(float* out3) {
    for (int i = 0; i < 3; ++i) out3[i] = kGrads.v[i];
}

// CreateNoisePermutationLookupImage, src/compute_density_field.cpp:90-125 (channel a).
void orc_perm_table(int32_t seed, uint8_t* out) {
    int32_t sh[512];
    shuffled_perm512(seed, sh);
    for (int i = 0; i < 256; ++i)
        for (int j = 0; j < 256; ++j) out[i * 256 + j] = (uint8_t)sh[noise_hash(i, j, seed) & 0x1ffu];
}

float orc_snoise2(const uint8_t* perm, float x, float y) { return snoise2(perm, x, y); }
float orc_snoise3(const uint8_t* perm, float x, float y, float z) { return snoise3(perm, x, y, z); }
float orc_fbm2(const uint8_t* perm, int32_t o, float f, float l, float g, float x, float y) {
    return fbm2(perm, o, f, l, g, x, y);
}
float orc_fbm3(const uint8_t* perm, int32_t o, float f, float l, float g, float x, float y, float z) {
    return fbm3(perm, o, f, l, g, x, y, z);
}
float orc_ridged_multifractal(const uint8_t* perm, int32_t o, float l, float g, float offset, float x, float y) {
    return ridged_multifractal(perm, o, l, g, offset, x, y);
}
float orc_leven_terrain(const uint8_t* perm, float x, float z) { return leven_terrain(perm, x, z); }
float orc_density(const lvn_field_desc* d, const uint8_t* perm, float x, float y, float z) {
    return density(*d, perm, x, y, z);
}
double orc_density_f64(const lvn_field_desc* d, const uint8_t* perm, double x, double y, double z) {
    return density_d(*d, perm, x, y, z);
}
float orc_brush_density(const lvn_csg_op* op, float x, float y, float z) {
    return brush_density(*op, cosf(op->rotate_y), sinf(op->rotate_y), x, y, z);
}

uint32_t orc_code_for_position(int32_t x, int32_t y, int32_t z, int32_t nd, int32_t md) {
    return code_for_position(x, y, z, nd, md);
}
void orc_position_for_code(uint32_t code, int32_t md, int32_t* xyz) { position_for_code(code, md, xyz); }

void orc_qef_solve_points_f32(const float* positions, const float* normals, int32_t count, float* out4) {
    Qef<float> q;
    for (int i = 0; i < count; ++i) {
        V4<float> p, n;
        p.x = positions[4 * i]; p.y = positions[4 * i + 1]; p.z = positions[4 * i + 2]; p.w = positions[4 * i + 3];
        n.x = normals[4 * i]; n.y = normals[4 * i + 1]; n.z = normals[4 * i + 2]; n.w = normals[4 * i + 3];
        qef_add_point(q, n, p);
    }
    qef_finish_mass(q);
    const V4<float> x = qef_solve(q);
    out4[0] = x.x; out4[1] = x.y; out4[2] = x.z; out4[3] = x.w;
}

void orc_qef_solve_points_f64(const double* positions, const double* normals, int32_t count, double* out4) {
    Qef<double> q;
    for (int i = 0; i < count; ++i) {
        V4<double> p, n;
        p.x = positions[4 * i]; p.y = positions[4 * i + 1]; p.z = positions[4 * i + 2]; p.w = positions[4 * i + 3];
        n.x = normals[4 * i]; n.y = normals[4 * i + 1]; n.z = normals[4 * i + 2]; n.w = normals[4 * i + 3];
        qef_add_point(q, n, p);
    }
    qef_finish_mass(q);
    const V4<double> x = qef_solve(q);
    out4[0] = x.x; out4[1] = x.y; out4[2] = x.z; out4[3] = x.w;
}

orc_chunk* orc_chunk_new(int32_t n) {
    if (n < 2 || n > 64 || (n & (n - 1))) return nullptr;
    return new orc_chunk(n);
}
void orc_chunk_free(orc_chunk* c) { delete c; }

int orc_chunk_generate(orc_chunk* c, const lvn_field_desc* d, const uint8_t* perm, const int32_t min[3],
                       int32_t size, uint32_t dm, const lvn_csg_op* ops, int32_t nops) {
    c->setup(d, perm, min, size, dm);
    c->generate(make_brushes(ops, nops), nullptr);
    return 0;
}

int orc_chunk_generate_from_density(orc_chunk* c, const lvn_field_desc* d, const uint8_t* perm,
                                    const int32_t min[3], int32_t size, uint32_t dm, const float* dens) {
    c->setup(d, perm, min, size, dm);
    c->generate({}, dens);
    return 0;
}

int orc_chunk_apply_csg(orc_chunk* c, const lvn_csg_op* ops, int32_t nops) {
    c->apply_csg(make_brushes(ops, nops));
    return 0;
}

int orc_chunk_build(orc_chunk* c) {
    c->build();
    return 0;
}

void orc_chunk_counts(const orc_chunk* c, int32_t* k) {
    k[0] = (int32_t)c->edgeIds.size();
    k[1] = (int32_t)c->vcode.size();
    k[2] = c->quads;
    k[3] = (int32_t)c->verts.size();
    k[4] = 2 * c->quads;
    k[5] = (int32_t)c->seams.size();
}
void orc_chunk_get_density(const orc_chunk* c, float* out) { std::copy(c->density.begin(), c->density.end(), out); }
void orc_chunk_get_materials(const orc_chunk* c, uint8_t* out) { std::copy(c->mat.begin(), c->mat.end(), out); }
void orc_chunk_get_edges(const orc_chunk* c, int32_t* ids, float* data4) {
    if (ids) std::copy(c->edgeIds.begin(), c->edgeIds.end(), ids);
    if (data4) std::copy(c->edgeData.begin(), c->edgeData.end(), data4);
}
void orc_chunk_get_voxels(const orc_chunk* c, int32_t* codes, int32_t* masks, int32_t* mats) {
    if (codes) std::copy(c->vcode.begin(), c->vcode.end(), codes);
    if (masks) std::copy(c->vmask.begin(), c->vmask.end(), masks);
    if (mats) std::copy(c->vmat.begin(), c->vmat.end(), mats);
}
void orc_chunk_get_positions(const orc_chunk* c, float* pos4, float* normal4) {
    if (pos4) std::copy(c->vpos.begin(), c->vpos.end(), pos4);
    if (normal4) std::copy(c->vnorm.begin(), c->vnorm.end(), normal4);
}
void orc_chunk_get_positions_f64(const orc_chunk* c, double* pos4) {
    std::copy(c->vpos64.begin(), c->vpos64.end(), pos4);
}
void orc_chunk_get_normals_f64(const orc_chunk* c, double* normal4) {
    std::copy(c->vnorm64.begin(), c->vnorm64.end(), normal4);
}
void orc_chunk_fp64_pipeline(const orc_chunk* c, double* pos4, double* normal4, int32_t* moved) {
    c->fp64_pipeline(pos4, normal4, moved);
}
void orc_chunk_get_triangles(const orc_chunk* c, int32_t* tri3) { std::copy(c->tris.begin(), c->tris.end(), tri3); }
void orc_chunk_get_vertices(const orc_chunk* c, lvn_vertex* v) { std::copy(c->verts.begin(), c->verts.end(), v); }
void orc_chunk_get_seams(const orc_chunk* c, lvn_seam_node* s) { std::copy(c->seams.begin(), c->seams.end(), s); }

double orc_mesh_chunks(const lvn_field_desc* desc, const uint8_t* perm, int32_t n_vox, int32_t n,
                       const int32_t* mins, const int32_t* sizes, uint32_t dm, const lvn_csg_op* ops, int32_t nops,
                       const int32_t* aabbs, int32_t threads, double* seconds, int64_t* totals) {
    using clk = std::chrono::steady_clock;
    std::atomic<int> next{0};
    std::atomic<int64_t> sumE{0}, sumV{0}, sumT{0}, nonEmpty{0};
    const auto t0 = clk::now();
    auto worker = [&]() {
        orc_chunk c(n_vox);
        std::vector<lvn_csg_op> sel;
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n) break;
            const auto ts = clk::now();
            const int* mn = &mins[3 * i];
            const int sz = sizes[i];
            sel.clear();
            for (int k = 0; k < nops; ++k) {  // AABB::overlaps, src/aabb.h:26-35
                const int* bb = &aabbs[6 * k];
                const bool disjoint = mn[0] + sz < bb[0] || mn[1] + sz < bb[1] || mn[2] + sz < bb[2] ||
                                      mn[0] > bb[3] || mn[1] > bb[4] || mn[2] > bb[5];
                if (!disjoint) sel.push_back(ops[k]);
            }
            c.setup(desc, perm, mn, sz, dm);
            c.generate(make_brushes(sel.data(), (int)sel.size()), nullptr);
            c.build();
            if (seconds) seconds[i] = std::chrono::duration<double>(clk::now() - ts).count();
            sumE += (int64_t)c.edgeIds.size();
            sumV += (int64_t)c.vcode.size();
            sumT += 2 * (int64_t)c.quads;
            nonEmpty += c.quads > 0 ? 1 : 0;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < std::max(threads, 1); ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    if (totals) {
        totals[0] += sumE;
        totals[1] += sumV;
        totals[2] += sumT;
        totals[3] += nonEmpty;
    }
    return std::chrono::duration<double>(clk::now() - t0).count();
}

}  // extern "C"
