"""Pins of the CPU oracle (CPU only).

Each test checks an oracle function against something the oracle did not
compute: published test vectors, the reference's own fixtures, closed forms,
invariants of the mathematics, or brute force on small inputs.
"""
import math

import numpy as np
import pytest

from leven_amd import abi
from leven_amd.abi import CSG_OP_DTYPE, MATERIAL_AIR

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")

# grad3 of the reference (src/compute_density_field.cpp:60-63): Perlin's 12 cube edges + 4 repeats
GRAD3 = np.array([[0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1], [1, 0, 1], [1, 0, -1], [-1, 0, 1], [-1, 0, -1],
                  [1, 1, 0], [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [1, 0, -1], [-1, 0, -1], [0, -1, 1], [0, 1, 1]])
F2, G2 = 0.366025403784, 0.211324865405
F3, G3 = 0.333333333333, 0.166666666667


def test_one_at_a_time_published_vectors(oracle):
    # Bob Jenkins' one-at-a-time hash: published outputs.
    assert oracle.oat_hash(b"a") == 0xCA2E9442
    assert oracle.oat_hash(b"The quick brown fox jumps over the lazy dog") == 0x519E91F5
    # NoiseHash hashes the 4 little-endian key bytes (src/compute_density_field.cpp:68-86)
    for x, y, seed in [(0, 0, 0), (3, 7, 1), (255, 255, 93923590), (17, 200, -5)]:
        key = (((x << 24) | (y << 16)) ^ seed) & 0xFFFFFFFF
        assert oracle.noise_hash(x, y, seed) == oracle.oat_hash(key.to_bytes(4, "little"))
    assert oracle.noise_hash(0, 0, 0) == 0


def test_splitmix64_published_vectors(oracle):
    out = oracle.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_base_permutation_structure(oracle):
    base = oracle.base_perm512()
    assert sorted(base[:256].tolist()) == list(range(256))
    assert base[:6].tolist() == [151, 160, 137, 91, 90, 15]  # Perlin's reference permutation
    assert (base[256:506] == base[6:256]).all() and (base[506:] == 0).all()


def test_shuffle_is_permutation_and_seeded(oracle):
    base = np.sort(oracle.base_perm512())
    a, b = oracle.shuffled_perm512(1), oracle.shuffled_perm512(2)
    assert (np.sort(a) == base).all() and (np.sort(b) == base).all()
    assert (a != b).any() and (oracle.shuffled_perm512(1) == a).all()


def test_perm_table_lookup_rule(oracle):
    seed = 1234
    perm = oracle.perm_table(seed)
    sh = oracle.shuffled_perm512(seed)
    rng = np.random.default_rng(0)
    for i, j in rng.integers(0, 256, size=(200, 2)):
        assert perm[i * 256 + j] == sh[oracle.noise_hash(int(i), int(j), seed) & 0x1FF]


def test_gradient_decode_values(oracle):
    g = oracle.grad_values()
    assert g[0] == np.float32(-1.0)
    assert abs(float(g[1]) - (64 / 255 * 4 - 1)) <= 2e-7
    assert abs(float(g[2]) - (128 / 255 * 4 - 1)) <= 2e-7


def _grad(perm, tx, ty, comps):
    v = perm[(ty & 255) * 256 + (tx & 255)] & 15
    gv = {-1: -1.0, 0: 64 / 255 * 4 - 1, 1: 128 / 255 * 4 - 1}
    return np.array([gv[int(c)] for c in GRAD3[v][:comps]])


def test_snoise_zero_at_origin(oracle):
    perm = oracle.perm_table(7)
    assert oracle.snoise2(perm, 0.0, 0.0) == 0.0
    assert oracle.snoise3(perm, 0.0, 0.0, 0.0) == 0.0


def test_snoise2_single_corner_neighbourhood(oracle):
    """Within 0.05 of a lattice vertex only that vertex contributes:
    n = 70 (0.5 - |d|^2)^4 (g . d), g read at texel (i mod 256, j mod 256)."""
    perm = oracle.perm_table(11)
    rng = np.random.default_rng(1)
    for _ in range(300):
        i, j = (int(v) for v in rng.integers(-300, 300, 2))
        p0 = np.array([i - (i + j) * G2, j - (i + j) * G2])
        d = rng.uniform(-0.05, 0.05, 2)
        p = (p0 + d).astype(np.float32)
        dd = p.astype(np.float64) - p0
        g = _grad(perm, i, j, 2)
        expect = 70 * (0.5 - dd @ dd) ** 4 * (g @ dd)
        tol = 2e-5 + 16 * float(np.spacing(np.abs(p).max()))  # FP32 rounding of P - P0
        assert abs(oracle.snoise2(perm, float(p[0]), float(p[1])) - expect) < tol


def test_snoise3_single_corner_neighbourhood(oracle):
    """3D: only the nearby vertex contributes; its gradient comes from the double
    lookup table[k][perm(i, j)] with value 255 wrapping to texel 0."""
    perm = oracle.perm_table(12)
    rng = np.random.default_rng(2)
    hit255 = 0
    for _ in range(400):
        i, j, k = (int(v) for v in rng.integers(-200, 200, 3))
        p0 = np.array([i, j, k], dtype=np.float64) - (i + j + k) * G3
        d = rng.uniform(-0.05, 0.05, 3)
        p = (p0 + d).astype(np.float32)
        dd = p.astype(np.float64) - p0
        v = int(perm[(j & 255) * 256 + (i & 255)])
        hit255 += v == 255
        tx = 0 if v == 255 else v
        g = _grad(perm, tx, k, 3)
        expect = 32 * (0.6 - dd @ dd) ** 4 * (g @ dd)
        tol = 2e-5 + 16 * float(np.spacing(np.abs(p).max()))
        assert abs(oracle.snoise3(perm, *map(float, p)) - expect) < tol
    # make the 255 wrap rule itself observable on purpose
    ys, xs = np.nonzero(perm.reshape(256, 256) == 255)
    i, j, k = int(xs[0]), int(ys[0]), 5
    p = (np.array([i, j, k], dtype=np.float64) - (i + j + k) * G3 + 0.02).astype(np.float32)
    dd = p.astype(np.float64) - (np.array([i, j, k], dtype=np.float64) - (i + j + k) * G3)
    g = _grad(perm, 0, k, 3)
    assert abs(oracle.snoise3(perm, *map(float, p)) - 32 * (0.6 - dd @ dd) ** 4 * (g @ dd)) < 2e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_snoise_is_continuous(oracle, dim):
    """Simplex noise is C1: a wrong corner or offset shows as a jump at simplex faces."""
    perm = oracle.perm_table(3)
    rng = np.random.default_rng(dim)
    step = 1e-3
    for _ in range(4):
        o = rng.uniform(-50, 50, dim)
        u = rng.normal(size=dim)
        u /= np.linalg.norm(u)
        ts = np.arange(0, 20, step)
        pts = (o[None, :] + ts[:, None] * u[None, :]).astype(np.float32)
        fn = oracle.snoise2 if dim == 2 else oracle.snoise3
        vals = np.array([fn(perm, *map(float, q)) for q in pts])
        assert np.abs(np.diff(vals)).max() < 40 * step
        assert np.abs(vals).max() < 1.1


def test_fbm_composition(oracle):
    perm = oracle.perm_table(5)
    L = oracle.lib()
    for x, y in [(3.25, -7.5), (120.0, 64.0), (-900.5, 13.0)]:
        s1 = oracle.snoise2(perm, np.float32(x) * np.float32(1 / 128), np.float32(y) * np.float32(1 / 128))
        assert L.orc_fbm2(oracle.vptr(perm), 1, 1 / 128, 2.0, 0.5, x, y) == s1
        s2 = oracle.snoise2(perm, np.float32(x) * np.float32(1 / 64), np.float32(y) * np.float32(1 / 64))
        assert L.orc_fbm2(oracle.vptr(perm), 2, 1 / 128, 2.0, 0.5, x, y) == np.float32(s1) + np.float32(s2) * 0.5


def _f32(v):
    return float(np.float32(v))


def _fbm_expected(noise, octaves, frequency, lacunarity, gain, coords):
    """BasicFractal (cl/noise.cl:8-36): sum_i gain^i * noise(p_i), summed in float64; the
    sample points follow the reference's float32 steps (p *= frequency, then p *= lacunarity
    per octave), since the noise is a function of those float32 coordinates."""
    p = [np.float32(c) * np.float32(frequency) for c in coords]
    total = 0.0
    for i in range(octaves):
        total += float(gain) ** i * noise(*[float(c) for c in p])
        p = [c * np.float32(lacunarity) for c in p]
    return total


def _ridged_expected(noise, octaves, lacunarity, gain, offset, coords):
    """RidgedMultiFractal (cl/noise.cl:42-81) step by step in float64 from independently
    evaluated snoise2 values; RIDGED_MULTI_H = 1, so exponent_i = lacunarity^-i."""
    p = [np.float32(c) for c in coords]
    signal = (offset - abs(noise(*[float(c) for c in p]))) ** 2
    total = signal
    for i in range(1, octaves + 1):
        p = [c * np.float32(lacunarity) for c in p]
        weight = min(max(signal * gain, 0.0), 1.0)
        signal = (offset - abs(noise(*[float(c) for c in p]))) * weight
        total += signal * float(lacunarity) ** -(i - 1)
    return total / octaves


def _terrain_expected(noise, x, z):
    """Terrain (cl/noise.cl:205-223) in float64 from the pieces above."""
    q = (np.float32(x) * np.float32(1 / 2000), np.float32(z) * np.float32(1 / 2000))
    ridged = min(max(0.8 * _ridged_expected(noise, 7, 2.114352, 1.5241, 1.0, q), 0.0), 1.0)
    billow = 0.5 * (0.6 * _fbm_expected(noise, 4, 0.24, 1.8754, 0.433, (np.float32(-4.33) * q[0], np.float32(7.98) * q[1]))) + 0.5
    b2 = 0.5 * (0.6 * _fbm_expected(noise, 2, 0.63, 2.2, 0.15, q)) + 0.5
    return billow * ridged + b2


@pytest.mark.parametrize("seed", [1, -7, 93923590])
def test_fbm_octave_sums(oracle, seed):
    """fbm2 and fbm3 equal the octave sum built from independently evaluated simplex values
    at lacunarities and gains away from 2 and 1/2 (a dropped lacunarity multiply, a swapped
    gain or a frequency applied twice moves the value by O(0.1)); lacunarity 1 is the
    geometric series s * (1 - g^n) / (1 - g)."""
    perm = oracle.perm_table(seed)
    L = oracle.lib()
    rng = np.random.default_rng(abs(seed) % 1000)
    s2 = lambda x, y: oracle.snoise2(perm, x, y)  # noqa: E731
    s3 = lambda x, y, z: oracle.snoise3(perm, x, y, z)  # noqa: E731
    for octaves, freq, lac, gain in [(1, 1 / 64, 2.0, 0.5), (2, 1 / 128, 1.7, 0.37), (3, 0.013, 2.31, 0.8),
                                     (5, 1 / 64, 2.0, 0.5), (4, 0.24, 1.8754, 0.433)]:
        for x, y, z in rng.uniform(-400, 400, (12, 3)).astype(np.float32):
            x, y, z = float(x), float(y), float(z)
            got2 = L.orc_fbm2(oracle.vptr(perm), octaves, freq, lac, gain, x, y)
            want2 = _fbm_expected(s2, octaves, _f32(freq), _f32(lac), _f32(gain), (x, y))
            assert abs(got2 - want2) <= 2e-5, (octaves, lac, gain, x, y)
            got3 = L.orc_fbm3(oracle.vptr(perm), octaves, freq, lac, gain, x, y, z)
            want3 = _fbm_expected(s3, octaves, _f32(freq), _f32(lac), _f32(gain), (x, y, z))
            assert abs(got3 - want3) <= 2e-5, (octaves, lac, gain, x, y, z)
    for x, y, z in rng.uniform(-400, 400, (8, 3)).astype(np.float32):
        x, y, z = float(x), float(y), float(z)
        one = oracle.snoise3(perm, _f32(x * _f32(0.02)), _f32(y * _f32(0.02)), _f32(z * _f32(0.02)))
        got = L.orc_fbm3(oracle.vptr(perm), 6, 0.02, 1.0, 0.6, x, y, z)
        assert abs(got - one * (1 - 0.6 ** 6) / (1 - 0.6)) <= 1e-5


@pytest.mark.parametrize("seed", [2, -11])
def test_ridged_multifractal_steps(oracle, seed):
    """RidgedMultiFractal against the reference's steps (cl/noise.cl:42-81) in float64:
    octave counts 1, 2, 3 and leven's 7, lacunarities and gains of either side of 1.
    Gain 0 leaves only the first signal: (offset - |s(p)|)^2 / octaves, which pins the
    final 1/octaves."""
    perm = oracle.perm_table(seed)
    L = oracle.lib()
    rng = np.random.default_rng(abs(seed))
    s2 = lambda x, y: oracle.snoise2(perm, x, y)  # noqa: E731
    for octaves, lac, gain, offset in [(1, 2.114352, 1.5241, 1.0), (2, 1.6, 0.7, 0.8), (3, 2.5, 1.2, 1.0),
                                       (7, 2.114352, 1.5241, 1.0)]:
        for x, y in rng.uniform(-3, 3, (16, 2)).astype(np.float32):
            x, y = float(x), float(y)
            got = L.orc_ridged_multifractal(oracle.vptr(perm), octaves, lac, gain, offset, x, y)
            want = _ridged_expected(s2, octaves, _f32(lac), _f32(gain), _f32(offset), (x, y))
            assert abs(got - want) <= 1e-5 * max(1.0, abs(want)), (octaves, lac, gain, x, y)
    for octaves in (1, 3, 7):
        for x, y in rng.uniform(-3, 3, (4, 2)).astype(np.float32):
            x, y = float(x), float(y)
            got = L.orc_ridged_multifractal(oracle.vptr(perm), octaves, 2.0, 0.0, 1.0, x, y)
            assert abs(got - (1.0 - abs(oracle.snoise2(perm, x, y))) ** 2 / octaves) <= 1e-6


def test_leven_terrain_and_density_kinds(oracle):
    """Terrain (cl/noise.cl:205-223) assembled from the pinned pieces, and each density kind
    (DensityFunc, cl/noise.cl:225-268, and the workloads' swaps) assembled from
    independently evaluated terms: y - 900 Terrain(x, z); y - A fbm2(x, z); and
    max(y - A fbm2(x, z), fbm3(x, y, z) - threshold) with both branches taken."""
    perm = oracle.perm_table(3)
    L = oracle.lib()
    rng = np.random.default_rng(8)
    s2 = lambda x, y: oracle.snoise2(perm, x, y)  # noqa: E731
    for x, y, z in rng.uniform(-6000, 6000, (40, 3)).astype(np.float32):
        x, y, z = float(x), float(y), float(z)
        t = L.orc_leven_terrain(oracle.vptr(perm), x, z)
        assert abs(t - _terrain_expected(s2, x, z)) <= 1e-5, (x, z)
        d = oracle.density(abi.field_desc(kind=abi.FIELD_LEVEN_TERRAIN), perm, x, y, z)
        assert d == np.float32(np.float32(y) - np.float32(np.float32(900.0) * np.float32(t)))
    hdesc = abi.field_desc(height_octaves=4, height_frequency=0.011, height_lacunarity=1.9, height_gain=0.45,
                           height_amplitude=17.0)
    cdesc = abi.field_desc(kind=abi.FIELD_TERRAIN_CAVES, height_octaves=3, height_frequency=1 / 100,
                           height_lacunarity=2.2, height_gain=0.55, height_amplitude=20.0, cave_octaves=4,
                           cave_frequency=1 / 50, cave_lacunarity=1.8, cave_gain=0.45, cave_threshold=0.25)
    branches = set()
    for x, y, z in rng.uniform(-200, 200, (300, 3)).astype(np.float32):
        x, y, z = float(x), float(y) * 0.2, float(z)
        h = L.orc_fbm2(oracle.vptr(perm), 4, 0.011, 1.9, 0.45, x, z)
        assert oracle.density(hdesc, perm, x, y, z) == np.float32(y) - np.float32(17.0) * np.float32(h)
        a = np.float32(y) - np.float32(20.0) * np.float32(L.orc_fbm2(oracle.vptr(perm), 3, 1 / 100, 2.2, 0.55, x, z))
        b = np.float32(L.orc_fbm3(oracle.vptr(perm), 4, 1 / 50, 1.8, 0.45, x, y, z)) - np.float32(0.25)
        assert oracle.density(cdesc, perm, x, y, z) == max(a, b)
        branches.add(bool(a > b))
    assert branches == {True, False}


def test_density_matches_fp64_twin(oracle):
    perm = oracle.perm_table(2)
    rng = np.random.default_rng(3)
    for desc in (abi.field_desc(), abi.field_desc(kind=abi.FIELD_TERRAIN_CAVES, height_octaves=5),
                 abi.field_desc(kind=abi.FIELD_LEVEN_TERRAIN)):
        for x, y, z in rng.uniform(-300, 300, (200, 3)).astype(np.float32):
            d32 = oracle.density(desc, perm, float(x), float(y), float(z))
            d64 = oracle.lib().orc_density_f64(desc, oracle.vptr(perm), float(x), float(y), float(z))
            assert abs(d32 - d64) <= 1e-5 * max(abs(d64), 1.0)


def test_octree_codes_reference_fixtures(oracle):
    """The reference's octree key fixtures (src/testdata/octree_keys_*.cpp): node codes of
    depth d carry a '10' marker above 3d path bits (PositionForCode's depth rule
    floor(bitlen / 3) ignores the marker, cl/octree.cl:51-72).  Decoding then re-encoding
    reproduces the 3d path bits (CodeForPosition writes a '1' marker), and a node's parent
    code (k >> 3) decodes to the child's position with the child's level bit cleared."""
    keys = np.load(f"{GOLDEN}/octree_keys.npz")
    assert len(keys.files) == 12
    xyz = np.zeros(3, np.int32)
    pxyz = np.zeros(3, np.int32)
    L = oracle.lib()
    checked = 0
    for name in keys.files:
        arr = keys[name]
        assert len(np.unique(arr)) == len(arr)
        present = set(arr.tolist())
        for k in arr[:: max(1, len(arr) // 1000)]:
            k = int(k)
            depth = k.bit_length() // 3
            assert k >> (3 * depth) == 2
            L.orc_position_for_code(k, 7, oracle.vptr(xyz))
            assert ((xyz & ((1 << (7 - depth)) - 1)) == 0).all()
            code = L.orc_code_for_position(int(xyz[0]), int(xyz[1]), int(xyz[2]), depth, 7)
            assert code == (1 << (3 * depth)) | (k & ((1 << (3 * depth)) - 1))
            if depth > 2:
                assert (k >> 3) in present
                L.orc_position_for_code(k >> 3, 7, oracle.vptr(pxyz))
                assert (pxyz == (xyz & ~(1 << (7 - depth)))).all()
            checked += 1
    assert checked > 10000


# ---------------------------------------------------------------------------
# QEF
# ---------------------------------------------------------------------------

def np_qef(points, normals, tol=0.1):
    """FP64 least squares about the mass point with the reference's truncation rule
    (eigenvalues < 0.1 or > 10 dropped, cl/qef.cl:107-109)."""
    p = np.asarray(points, np.float64)
    n = np.asarray(normals, np.float64)
    mp = p.mean(axis=0)
    A = n.T @ n
    b = n.T @ np.einsum("ij,ij->i", n, p)
    w, V = np.linalg.eigh(A)
    inv = np.where((np.abs(w) < tol) | (np.abs(1 / np.where(w == 0, 1, w)) < tol), 0.0, 1 / np.where(w == 0, 1, w))
    pinv = (V * inv) @ V.T
    return pinv @ (b - A @ mp) + mp


def test_qef_five_plane_example(oracle):
    """The five-plane example of src/glsl_svd.cpp:500-514 against FP64 least squares."""
    nrm = np.array([[1, 1, 0], [1, 1, 0], [-1, 1, 0], [-1, 2, 1], [-1, 1, 0]], np.float64)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    pts = np.array([[1.0, 0.0, 0.3], [0.9, 0.1, -0.5], [-0.8, 0.2, 0.6], [-1.0, 0.0, 0.01], [-1.1, -0.1, -0.5]])
    p4 = np.hstack([pts, np.zeros((5, 1))])
    n4 = np.hstack([nrm, np.zeros((5, 1))])
    expect = np_qef(pts, nrm)
    got32 = oracle.qef_solve_f32(p4, n4)[:3]
    got64 = oracle.qef_solve_f64(p4, n4)[:3]
    assert np.abs(got64 - expect).max() < 1e-6
    assert np.abs(got32 - expect).max() < 1e-5


def test_qef_random_problems(oracle):
    rng = np.random.default_rng(9)
    for _ in range(300):
        k = int(rng.integers(2, 13))
        pts = rng.uniform(0, 1, (k, 3))
        base = rng.normal(size=3)
        nrm = base + rng.normal(scale=rng.choice([0.01, 0.3, 2.0]), size=(k, 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        expect = np_qef(pts, nrm)
        w = np.linalg.eigvalsh(nrm.T @ nrm)
        if np.min(np.abs(w - 0.1)) < 1e-3 or np.min(np.abs(w - 10)) < 1e-3:
            continue  # truncation decision within rounding of the threshold
        got = oracle.qef_solve_f64(np.hstack([pts, np.zeros((k, 1))]), np.hstack([nrm, np.zeros((k, 1))]))
        assert np.abs(got[:3] - expect).max() < 1e-6


# ---------------------------------------------------------------------------
# Chunk pipeline: closed forms, invariants, brute force
# ---------------------------------------------------------------------------

def plane_desc():
    return abi.field_desc(height_octaves=0)  # d = y exactly


@pytest.mark.parametrize("n", [8, 64])
def test_plane_closed_form(oracle, n):
    """d = y with the sheet at world y = 0: H^2 y-edges, N^2 voxels, (N-1)^2 quads,
    vertices at the voxel centres in x/z, t = 1, normal (0, 1, 0)."""
    perm = oracle.perm_table(1)
    c = oracle.Chunk(n)
    size = n * 4
    mn = (0, -4 * (n // 2), 0)
    c.generate(plane_desc(), perm, mn, size).build()
    k = c.counts()
    H = n + 1
    assert k["edges"] == H * H and k["voxels"] == n * n
    assert k["quads"] == (n - 1) ** 2 and k["triangles"] == 2 * (n - 1) ** 2 and k["vertices"] == n * n
    ids, data = c.edges()
    assert (ids & 3 == 1).all()
    assert (data[:, 3] == 1.0).all() and (data[:, :3] == [0, 1, 0]).all()
    pos, nrm = c.positions()
    codes, masks, mats = c.voxels()
    xyz = np.zeros(3, np.int32)
    depth = int(math.log2(n))
    for i in range(0, len(codes), max(1, len(codes) // 200)):
        oracle.lib().orc_position_for_code(int(codes[i]), depth, oracle.vptr(xyz))
        expect = (np.array([xyz[0] + 0.5, n // 2, xyz[2] + 0.5]) * 4 + np.array(mn))
        assert np.allclose(pos[i, :3], expect, atol=1e-4)
    assert (nrm[:, :3] == [0, 1, 0]).all()
    assert np.allclose(c.normals_f64()[:, :3], [0, 1, 0], atol=1e-12)  # the FP64 twin of the normals
    tri = c.triangles()
    # winding: corner 5 (lower end of the host's edge 7) is solid -> flipped order {0,3,1,0,2,3}
    assert (tri[0::2, 0] == tri[1::2, 0]).all()


def _sphere_op(cx, cy, cz, r, material=1, typ=0):
    op = np.zeros(1, CSG_OP_DTYPE)
    op["type"] = typ
    op["brush"] = 1
    op["material"] = material if typ == 0 else MATERIAL_AIR
    op["origin"] = [cx, cy, cz, 0]
    op["dims"] = [r, r, r, 0]
    return op


def test_sphere_brush_mesh_is_closed_manifold(oracle):
    """An added sphere in air meshes to a closed, consistently oriented surface of genus 0."""
    perm = oracle.perm_table(1)
    n = 32
    c = oracle.Chunk(n)
    # the plane d = y lies below the chunk (voxel y >= 8): everything air before the brush
    c.generate(plane_desc(), perm, (0, 4 * 8, 0), n * 4, ops=_sphere_op(16.3, 24.4, 15.7, 9.2)).build()
    tri = c.triangles()
    assert len(tri) > 100
    directed = {}
    for t in tri:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            directed[(a, b)] = directed.get((a, b), 0) + 1
    assert max(directed.values()) == 1
    assert all((b, a) in directed for (a, b) in directed)
    used = np.unique(tri)
    E = len(directed) // 2
    assert len(used) - E + len(tri) == 2
    # brush data: t is the first arg-min of | |p - o| - r | over 17 samples, normal radial
    ids, data = c.edges()
    shift = int(math.log2(n)) + 1
    m = (1 << shift) - 1
    o = np.array([16.3, 24.4, 15.7])
    checked = 0
    for eid, d in zip(ids, data):
        a, h = eid & 3, eid >> 2
        p0 = np.array([h & m, (h >> shift) & m, (h >> (2 * shift)) & m], np.float64) + [0, 8, 0]
        e = np.eye(3)[a]
        ts = np.arange(17) / 16.0
        dist = np.abs(np.linalg.norm(p0[None] + ts[:, None] * e[None] - o, axis=1) - 9.2)
        if np.sort(dist)[1] - dist.min() < 1e-4:
            continue  # near tie: FP32 may order it either way
        assert d[3] == ts[np.argmin(dist)]
        p = p0 + d[3] * e
        signed = np.linalg.norm(p - o) - 9.2
        if abs(signed) < 1e-4:
            continue
        if signed > 0:
            # BrushNormal only uses brushes containing the point (cl/apply_csg_operation.cl:153-158)
            assert (d[:3] == 0).all()
        else:
            radial = (p - o) / np.linalg.norm(p - o)
            assert np.abs(d[:3] - radial).max() < 5e-3  # FP32 central difference, h = 0.001, |p| < 50
        checked += 1
    assert checked > 100


def test_csg_material_override_semantics(oracle):
    """Add then subtract: the last containing op decides; re-materialing solid samples
    marks their edges as touched (brush data), untouched edges keep density data."""
    perm = oracle.perm_table(1)
    n = 16
    desc = plane_desc()
    base = oracle.Chunk(n).generate(desc, perm, (0, -32, 0), 64)
    ops = np.concatenate([_sphere_op(8, 0, 8, 5, material=3), _sphere_op(8, 0, 8, 2.5, typ=1)])
    c = oracle.Chunk(n).generate(desc, perm, (0, -32, 0), 64, ops=ops)
    F = n + 2
    m0 = base.materials().reshape(F, F, F)
    m1 = c.materials().reshape(F, F, F)
    zz, yy, xx = np.mgrid[0:F, 0:F, 0:F]
    wp = np.stack([xx, yy - 8, zz], -1).astype(np.float64)
    r = np.linalg.norm(wp - [8, 0, 8], axis=-1)
    expect = m0.copy()
    expect[r <= 5] = 3
    expect[r <= 2.5] = MATERIAL_AIR
    assert (m1 == expect).all()
    # edges untouched by the brushes keep the plane's density data exactly
    ids0, data0 = base.edges()
    ids1, data1 = c.edges()
    lookup = dict(zip(ids0.tolist(), range(len(ids0))))
    changed = (m0 != m1)
    shift = int(math.log2(n)) + 1
    m = (1 << shift) - 1
    kept = 0
    for eid, d in zip(ids1, data1):
        a, h = eid & 3, eid >> 2
        x, y, z = h & m, (h >> shift) & m, (h >> (2 * shift)) & m
        e = np.eye(3, dtype=int)[a]
        touched = changed[z, y, x] or changed[z + e[2], y + e[1], x + e[0]]
        if not touched:
            assert (d == data0[lookup[eid]]).all()
            kept += 1
    assert kept > 0


def test_density_crossing_brute_force(oracle):
    """t is the first arg-min of |d| over t = k/16 (evaluated with the oracle's density
    at each sample); the normal matches an FP64 central difference."""
    perm = oracle.perm_table(4)
    desc = abi.field_desc()
    n = 16
    c = oracle.Chunk(n).generate(desc, perm, (256, -64, 512), 64)
    ids, data = c.edges()
    shift = int(math.log2(n)) + 1
    m = (1 << shift) - 1
    L = oracle.lib()
    for eid, d in list(zip(ids, data))[:: max(1, len(ids) // 150)]:
        a, h = eid & 3, eid >> 2
        p0 = np.array([h & m, (h >> shift) & m, (h >> (2 * shift)) & m], np.float32) + np.float32([64, -16, 128])
        e = np.eye(3, dtype=np.float32)[a]
        vals = [abs(oracle.density(desc, perm, *map(float, p0 + e * np.float32(k / 16)))) for k in range(17)]
        assert d[3] == np.float32(int(np.argmin(vals)) / 16)
        p = p0.astype(np.float64) + e * float(d[3])
        grad = np.array([L.orc_density_f64(desc, oracle.vptr(perm), *(p + s * 1e-3 * np.eye(3)[i]))
                         - L.orc_density_f64(desc, oracle.vptr(perm), *(p - s * 1e-3 * np.eye(3)[i]))
                         for i in range(3) for s in [1]])
        grad /= np.linalg.norm(grad)
        assert np.abs(d[:3] - grad).max() < 2e-2  # FP32 central differences of O(10) values


def brute_force_dc(materials, n):
    """Independent restatement of the quad rule on a tiny grid: for every voxel
    (x-fastest order) and axis, the edge (axis*4+3) of the voxel with a sign change
    and both other coordinates < n-1 yields a quad of the 4 voxels around it."""
    F = n + 2
    solid = (materials.reshape(F, F, F) != MATERIAL_AIR)
    active = {}
    order = []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                c = solid[z:z + 2, y:y + 2, x:x + 2]
                if c.any() and not c.all():
                    active[(x, y, z)] = len(order)
                    order.append((x, y, z))
    quads = []
    for (x, y, z) in order:
        for axis in range(3):
            p = [x, y, z]
            if p[(axis + 1) % 3] == n - 1 or p[(axis + 2) % 3] == n - 1:
                continue
            # the edge along `axis` at the voxel's (+1, +1) corner in the other two axes
            lo = [x + 1, y + 1, z + 1]
            lo[axis] = p[axis]
            hi = list(lo)
            hi[axis] += 1
            s0, s1 = solid[lo[2], lo[1], lo[0]], solid[hi[2], hi[1], hi[0]]
            if s0 == s1:
                continue
            offs = {0: [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1)],
                    1: [(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)],
                    2: [(0, 0, 0), (0, 1, 0), (1, 0, 0), (1, 1, 0)]}[axis]
            nodes = [active[(x + o[0], y + o[1], z + o[2])] for o in offs]
            order_ = [0, 3, 1, 0, 2, 3] if s0 else [0, 1, 3, 0, 3, 2]
            quads.append([nodes[i] for i in order_])
    return np.array(quads, np.int32).reshape(-1, 3)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_topology_brute_force(oracle, seed):
    perm = oracle.perm_table(seed)
    n = 8
    desc = abi.field_desc(kind=abi.FIELD_TERRAIN_CAVES, height_octaves=5, height_amplitude=3.0,
                          cave_frequency=1 / 6.0)
    c = oracle.Chunk(n).generate(desc, perm, (seed * 32, -16, 0), n * 4).build()
    tri = c.triangles()
    expect = brute_force_dc(c.materials(), n)
    assert len(expect) > 0
    assert (tri == expect).all()


def test_snoise3_bound(oracle):
    """The cave early-out of the kernels (skip the cave term where y - H exceeds
    1.19 x the sum of octave amplitudes minus the threshold) relies on |snoise3| <= 1.1863.
    Recomputed here: 32 x |g|max x max over one simplex cell of the sum of the four corner
    terms (0.6 - r^2)^4 r, on a 480-step grid of the cell's skewed coordinates (ordered
    region x >= y >= z; the other five orderings are axis permutations with the same
    distances) plus the grid's Lipschitz margin; then checked against random samples."""
    G3 = 1.0 / 6.0
    n = 480
    h = 1.0 / n
    g = np.arange(n + 1) * h

    def phi(r2):
        return np.maximum(0.6 - r2, 0.0) ** 4 * np.sqrt(r2)

    best = 0.0
    for ix, sx in enumerate(g):
        sy, sz = np.meshgrid(g[:ix + 1], g[:ix + 1], indexing="ij")
        keep = sz <= sy
        sy, sz = sy[keep], sz[keep]
        t = (sx + sy + sz) * G3
        f = np.stack([sx - t, sy - t, sz - t], axis=-1)
        total = 0.0
        for off in ([0, 0, 0], [1 - G3, -G3, -G3], [1 - 2 * G3, 1 - 2 * G3, -2 * G3], [1 - 3 * G3] * 3):
            d = f - np.array(off)
            total = total + phi((d * d).sum(-1))
        best = max(best, float(total.max()))
    r = np.linspace(0.0, np.sqrt(0.6), 200001)
    lip = 4 * np.abs((0.6 - r * r) ** 3 * (0.6 - 9 * r * r)).max()  # |d/ds sum phi|, unskew norm 1
    one, zero = ((1 * 64 + 64) / 255.0) * 4 - 1, ((0 * 64 + 64) / 255.0) * 4 - 1
    gmax = np.sqrt(2 * one * one + zero * zero)
    bound = 32 * gmax * (best + lip * h * np.sqrt(3) / 2)
    assert bound <= 1.1863 + 1e-4
    # the decoded gradient components are the ones bounded here: Perlin's 16 directions
    # have two components of magnitude 1 and one 0, decoded as byte/255*4-1
    decoded = np.asarray(oracle.grad_values(), dtype=np.float64)  # values of -1, 0, +1
    one_o, zero_o = np.abs(decoded[[0, 2]]).max(), abs(decoded[1])
    assert np.sqrt(2 * one_o ** 2 + zero_o ** 2) <= gmax + 1e-6
    rng = np.random.default_rng(3)
    perm = oracle.perm_table(7)
    pts = rng.uniform(-300, 300, (20000, 3)).astype(np.float32)
    m = max(abs(oracle.snoise3(perm, float(x), float(y), float(z))) for x, y, z in pts)
    assert m < 1.0
