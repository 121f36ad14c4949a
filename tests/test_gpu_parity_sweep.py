"""Seeded random sweep against the oracle: every stage bit-exact (density, materials, edge
list and data, voxel set, codes, masks, materials, vertex positions and normals).

Each case draws a field kind (heightfield, heightfield + caves, leven's own terrain), its
fractal parameters (octaves 1-8, frequencies, lacunarities away from 2 so octave scaling is
inexact, gains, amplitudes, cave threshold), a noise seed of either sign, a chunk size
(N = 16, 32 or 64 voxels; a second sweep N = 4 and 8) and LOD (0-4), a position with the surface inside the chunk, and
0-6 brushes (spheres and Y-rotated boxes, add and subtract, several materials) through
the chunk (needs a GPU).  Each case is also meshed by a one-chunk batch from the stored
ops and compared with the oracle's mesh.
"""
import numpy as np
import pytest

from leven_amd import abi, mesher, workloads
from leven_amd.abi import CSG_OP_DTYPE, MATERIAL_AIR
from test_gpu_parity import compare_stages

pytestmark = pytest.mark.gpu


def _case(k, n_choice=(16, 32, 64), n_p=(0.4, 0.3, 0.3)):
    rng = np.random.default_rng(1000 + k)
    kind = int(rng.integers(0, 3))
    desc = abi.field_desc(kind=kind, height_octaves=int(rng.integers(1, 9)),
                          height_frequency=float(2.0 ** -rng.uniform(5, 9)),
                          height_lacunarity=float(rng.uniform(1.6, 2.4)), height_gain=float(rng.uniform(0.35, 0.65)),
                          height_amplitude=float(rng.uniform(8, 48)), cave_octaves=int(rng.integers(1, 7)),
                          cave_frequency=float(2.0 ** -rng.uniform(4, 7)),
                          cave_lacunarity=float(rng.uniform(1.7, 2.3)), cave_gain=float(rng.uniform(0.4, 0.6)),
                          cave_threshold=float(rng.uniform(-0.1, 0.5)))
    seed = int(rng.integers(-2 ** 31, 2 ** 31 - 1))
    n = int(rng.choice(list(n_choice), p=list(n_p)))
    lod = int(rng.integers(0, 5 if n < 64 else 3))
    size = n * 4 << lod
    scale = size // (n * 4)
    mx, mz = (int(v) * size for v in rng.integers(-64, 64, 2))
    return dict(k=k, desc=desc, seed=seed, n=n, size=size, scale=scale, mx=mx, mz=mz, nops=int(rng.integers(0, 7)),
                rng=rng)


def _surface_min(oracle, c, perm):
    """Chunk min (world units) whose y range holds the surface at the chunk's centre column."""
    n, scale = c["n"], c["scale"]
    cx = c["mx"] / 4 + n * scale / 2
    cz = c["mz"] / 4 + n * scale / 2
    if c["desc"].kind == abi.FIELD_TERRAIN_CAVES:
        h = 0.0
    else:  # density(x, 0, z) = -H for a pure height field
        h = -oracle.density(c["desc"], perm, cx, 0.0, cz)
    y0 = int(np.floor(h / scale)) - n // 2  # in voxels of this LOD
    return (c["mx"], y0 * scale * 4, c["mz"])


def _ops(c, mn):
    rng, n, scale = c["rng"], c["n"], c["scale"]
    ops = np.zeros(c["nops"], CSG_OP_DTYPE)
    lo = np.array(mn, np.float64) / 4
    for op in ops:
        add = rng.random() < 0.5
        op["type"] = 0 if add else 1
        op["brush"] = int(rng.integers(0, 2))
        op["material"] = int(rng.integers(1, 6)) if add else MATERIAL_AIR
        centre = lo + rng.uniform(0.15, 0.85, 3) * n * scale
        op["origin"] = [centre[0] + 0.5, centre[1] + 0.5, centre[2] + 0.5, 0.0]
        hi = n * scale / 4  # N = 4 at LOD 0 leaves a chunk too small for 1.5-voxel brushes
        r = rng.uniform(1.5, hi) if hi > 1.5 else rng.uniform(0.5, 1.5)
        lo_d = min(1.5, r)
        op["dims"] = [r, r, r, 0.0] if op["brush"] == 1 else [r, rng.uniform(lo_d, r), rng.uniform(lo_d, r), 0.0]
        op["rotate_y"] = float(rng.uniform(-np.pi, np.pi)) if op["brush"] == 0 else 0.0
    return ops


def _run_case(oracle, c):
    k = c["k"]
    perm = oracle.perm_table(c["seed"])
    mn = _surface_min(oracle, c, perm)
    ops = _ops(c, mn)
    print(f"case {k}: kind {c['desc'].kind} seed {c['seed']} N {c['n']} size {c['size']} min {mn} ops {len(ops)}",
          flush=True)
    mesher.init(c["desc"], c["seed"], 0)
    mesher.clear_csg_ops()
    ctx = mesher.Context(c["n"])
    g = ctx.debug_stages(mn, c["size"], ops)
    ch = oracle.Chunk(c["n"]).generate(c["desc"], perm, mn, c["size"], ops=ops).build()
    compare_stages(g, ch, c["n"], c["size"])
    ctx.close()
    # The batched form meshes the same chunk from the stored ops; for the fields without
    # the cave term it stores no density and takes the edges' end samples from the
    # column heights.
    for op in ops:
        bb = workloads.op_aabb(op)
        mesher.store_csg_op(op, bb[:3], bb[3:])
    b = mesher.Batch(c["n"], 1, edge_cap=0, voxel_cap=0)
    b.set_chunks(np.array([mn], np.int32), np.array([c["size"]], np.int32))
    b.run()
    b.wait()
    v, t, s = b.get_mesh(0)
    b.close()
    mesher.clear_csg_ops()
    assert (t == ch.triangles()).all(), "batched triangles"
    assert (v.view(np.uint32) == ch.vertices().view(np.uint32)).all(), "batched vertices"
    assert (s.view(np.uint32) == ch.seams().view(np.uint32)).all(), "batched seams"


@pytest.mark.parametrize("k", range(24))
def test_random_case_bit_exact(oracle, k):
    _run_case(oracle, _case(k))


@pytest.mark.parametrize("k", range(100, 112))
def test_small_chunk_case_bit_exact(oracle, k):
    """N = 4 and 8: partial LDS staging of sample rows, CSG slabs past the chunk's z
    range and single-segment voxel rows."""
    _run_case(oracle, _case(k, n_choice=(4, 8), n_p=(0.5, 0.5)))
