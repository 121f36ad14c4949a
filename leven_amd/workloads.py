"""Synthetic workloads of the chunk mesher (all seeds explicit).

Units follow the reference: chunk ``min``/``size`` in world units (4 world units
per LOD0 voxel, src/volume_constants.h:7-8, size = 256 * 2^LOD); CSG op origins
and extents in voxel units with the reference caller's +0.5 offset
(src/clipmap.cpp:1618-1643); op AABBs in world units as CalcCSGOperationBounds
computes them.

  plain_cpu     one 64^3 chunk of a 2D simplex-fBm heightfield (6 octaves)
  terrain_512   8x8x8 chunks of heightfield + 3D cave noise with 256 brushes
  precision     the terrain_512 field with 1024 Y-rotated boxes (FP16 + precise QEF)
  clipmap       6-LOD clipmap selection around a seeded camera (~2000 chunks), terrain_512's field
  clipmap_leven the same selection around a camera on the reference's own terrain field
  serving       Poisson request stream along a camera path with 5% CSG edits
"""
import math

import numpy as np

from .abi import (CSG_OP_DTYPE, CTX_PRECISE_QEF, FIELD_HEIGHTFIELD, FIELD_LEVEN_TERRAIN, FIELD_TERRAIN_CAVES,
                  MATERIAL_AIR, field_desc)

LEAF_SIZE_SCALE = 4
CHUNK_SIZE = 256  # world units of a LOD0 64^3 chunk
NUM_LODS = 6
LOD_MAX_NODE_SIZE = CHUNK_SIZE * (1 << (NUM_LODS - 1))
# src/clipmap.cpp:33-41
LOD_ACTIVE_DISTANCES = [0.0, CHUNK_SIZE * 1.5, CHUNK_SIZE * 3.5, CHUNK_SIZE * 5.5, CHUNK_SIZE * 7.5,
                        CHUNK_SIZE * 13.5]


def heightfield_desc():
    """2D simplex fBm: 6 octaves, f = 1/128, lacunarity 2, gain 0.5, amplitude 24 voxels."""
    return field_desc(FIELD_HEIGHTFIELD, height_octaves=6, height_frequency=1.0 / 128.0, height_lacunarity=2.0,
                      height_gain=0.5, height_amplitude=24.0)


def terrain_desc():
    """max(y - 24 fbm2(x, z), fbm3(x, y, z) - 0.3), 5 octaves each, f = 1/128 (2D) and 1/64 (3D)."""
    return field_desc(FIELD_TERRAIN_CAVES, height_octaves=5, height_frequency=1.0 / 128.0, height_lacunarity=2.0,
                      height_gain=0.5, height_amplitude=24.0, cave_octaves=5, cave_frequency=1.0 / 64.0,
                      cave_lacunarity=2.0, cave_gain=0.5, cave_threshold=0.3)


def op_aabb(op):
    """CalcCSGOperationBounds (src/clipmap.cpp:1634-1639), widened for Y-rotated boxes."""
    d = np.array(op["dims"][:3], dtype=np.float32)
    if op["brush"] == 0 and op["rotate_y"] != 0.0:
        r = math.sqrt(float(d[0]) ** 2 + float(d[2]) ** 2)
        d = np.array([r, d[1], r], dtype=np.float32)
    half = np.trunc(d * np.float32(LEAF_SIZE_SCALE)).astype(np.int64) + 2
    o = np.trunc((np.array(op["origin"][:3], dtype=np.float32) - np.float32(0.5)) * np.float32(LEAF_SIZE_SCALE))
    o = o.astype(np.int64)
    return np.concatenate([o - half, o + half]).astype(np.int32)


def make_ops(n, seed, spheres=None, rotated=False, extent=512, y_range=(-32.0, 32.0), size_range=(2.0, 12.0)):
    """n brushes: centres uniform over x, z in [0, extent) voxels and y in y_range; half extents /
    radii uniform in size_range voxels; add/subtract 50/50; add materials 1..3."""
    rng = np.random.default_rng(seed)
    ops = np.zeros(n, dtype=CSG_OP_DTYPE)
    n_sph = n // 2 if spheres is None else spheres
    shapes = np.array([1] * n_sph + [0] * (n - n_sph))
    rng.shuffle(shapes)
    for i in range(n):
        c = np.array([rng.uniform(0, extent), rng.uniform(*y_range), rng.uniform(0, extent)], dtype=np.float32)
        is_add = rng.integers(0, 2) == 0
        ops[i]["type"] = 0 if is_add else 1
        ops[i]["brush"] = shapes[i]
        ops[i]["material"] = int(rng.integers(1, 4)) if is_add else MATERIAL_AIR
        if shapes[i] == 1:
            r = rng.uniform(*size_range)
            ops[i]["dims"] = [r, r, r, 0.0]
        else:
            ops[i]["dims"] = [rng.uniform(*size_range), rng.uniform(*size_range), rng.uniform(*size_range), 0.0]
            if rotated:
                ops[i]["rotate_y"] = rng.uniform(0.0, 2.0 * math.pi)
        ops[i]["origin"] = [c[0] + 0.5, c[1] + 0.5, c[2] + 0.5, 0.0]
    aabbs = np.stack([op_aabb(op) for op in ops]) if n else np.zeros((0, 6), np.int32)
    return ops, aabbs


def plain_cpu():
    """One 64^3 LOD0 chunk of the heightfield; min = (0, -32, 0) voxels."""
    return dict(name="plain_cpu", desc=heightfield_desc(), seed=1, default_material=0,
                mins=np.array([[0, -32 * LEAF_SIZE_SCALE, 0]], np.int32), sizes=np.array([CHUNK_SIZE], np.int32),
                ops=np.zeros(0, CSG_OP_DTYPE), aabbs=np.zeros((0, 6), np.int32), flags=0)


def terrain_512():
    """8x8x8 LOD0 chunks over x, z in [0, 512) and y in [-256, 256) voxels, 256 brushes (seed 2)."""
    mins = np.array([[x * CHUNK_SIZE, y * CHUNK_SIZE, z * CHUNK_SIZE]
                     for z in range(8) for y in range(-4, 4) for x in range(8)], np.int32)
    ops, aabbs = make_ops(256, seed=2)
    return dict(name="terrain_512", desc=terrain_desc(), seed=2, default_material=0, mins=mins,
                sizes=np.full(len(mins), CHUNK_SIZE, np.int32), ops=ops, aabbs=aabbs, flags=0)


def precision():
    """terrain_512's chunks and field, 1024 seeded Y-rotated boxes; FP16 density + precise QEF."""
    w = terrain_512()
    ops, aabbs = make_ops(1024, seed=3, spheres=0, rotated=True)
    w.update(name="precision", ops=ops, aabbs=aabbs, flags=CTX_PRECISE_QEF)
    return w


def _distance_to_node(mn, size, cam):  # DistanceToNode, src/clipmap.cpp:1055-1065
    lo = np.asarray(mn, np.float64)
    hi = lo + size
    d = np.maximum(np.maximum(lo - cam, 0.0), cam - hi)
    return float(np.sqrt((d * d).sum()))


def clipmap_select(camera, world_min, world_max):
    """SelectActiveClipmapNodes (src/clipmap.cpp:1084-1126) over an octree of LOD_MAX_NODE_SIZE roots."""
    selected = []

    def visit(mn, size, parent_active):
        active = False
        if not parent_active and size <= LOD_MAX_NODE_SIZE:
            idx = int(math.log2(size // CHUNK_SIZE))
            if _distance_to_node(mn, size, camera) >= LOD_ACTIVE_DISTANCES[idx]:
                selected.append((tuple(mn), size))
                active = True
        if size > CHUNK_SIZE:
            h = size // 2
            for i in range(8):
                c = (mn[0] + h * ((i >> 2) & 1), mn[1] + h * ((i >> 1) & 1), mn[2] + h * (i & 1))
                visit(c, h, parent_active or active)

    R = LOD_MAX_NODE_SIZE
    for x in range(world_min[0], world_max[0], R):
        for y in range(world_min[1], world_max[1], R):
            for z in range(world_min[2], world_max[2], R):
                visit((x, y, z), R, False)
    return selected


def clipmap(seed=4):
    """6-LOD clipmap around a seeded camera near the terrain surface (~2000 chunks)."""
    rng = np.random.default_rng(seed)
    cam = np.array([rng.uniform(-1024, 1024), rng.uniform(-64, 64), rng.uniform(-1024, 1024)])
    R = LOD_MAX_NODE_SIZE
    # 28 x 2 x 28 root nodes: the rule keeps ~330 nodes of LOD 0-4 near the camera, the
    # rest of the ~1900 selected chunks are LOD-5 tiles (voxel size 32).  The selection is
    # the library's request producer (lvn_clipmap_update); clipmap_select above restates it.
    from . import mesher
    cm = mesher.Clipmap((-14 * R, -R, -14 * R), (14 * R, R, 14 * R))
    mins, sizes = cm.update(cam)["selected"]
    cm.close()
    ops, aabbs = make_ops(256, seed=2)
    return dict(name="clipmap", desc=terrain_desc(), seed=2, default_material=0, mins=mins, sizes=sizes,
                ops=ops, aabbs=aabbs, flags=0, camera=cam)


def leven_desc():
    """The reference's own field, DensityFunc = y - 900 Terrain(x, z) (cl/noise.cl:205-266)."""
    return field_desc(FIELD_LEVEN_TERRAIN)


LEVEN_SEED = 93923590
# Camera of clipmap_leven in world units: x, z drawn as clipmap()'s (seed 4); y sits 16 voxels
# above leven's terrain there, whose surface near the camera is the plane
# h = 575.5 + 0.30 (x - cx) + 0.40 (z - cz) voxels to within ~10 voxels over +-96 voxels
# (the field sampled at 3 x 3 points 64 voxels apart).
LEVEN_CAMERA = (907.379, 4.0 * 575.53 + 64.0, 23.199)
LEVEN_SURFACE = (575.53, 0.30, 0.40)


def clipmap_leven(seed=6):
    """The 6-LOD clipmap selection (~1,900 chunks at voxel sizes 1..32) around a camera just
    above the reference's own terrain field, with 256 brushes on the surface near the camera
    (radii 2-12 voxels, spheres and boxes, add/subtract 50/50)."""
    cam = np.array(LEVEN_CAMERA, np.float64)
    R = LOD_MAX_NODE_SIZE
    from . import mesher
    cm = mesher.Clipmap((-14 * R, -R, -14 * R), (14 * R, R, 14 * R))
    mins, sizes = cm.update(cam)["selected"]
    cm.close()
    # brushes within +-96 voxels of the camera, centred on the surface plane +-12 voxels
    cx, cz = cam[0] / LEAF_SIZE_SCALE, cam[2] / LEAF_SIZE_SCALE
    ops, _ = make_ops(256, seed=seed, extent=192, y_range=(-12.0, 12.0))
    h0, sx, sz = LEVEN_SURFACE
    for op in ops:
        x = op["origin"][0] + cx - 96.0
        z = op["origin"][2] + cz - 96.0
        y = op["origin"][1] + h0 + sx * (x - cx) + sz * (z - cz)
        op["origin"][:3] = [x, y, z]
    aabbs = np.stack([op_aabb(op) for op in ops])
    return dict(name="clipmap_leven", desc=leven_desc(), seed=LEVEN_SEED, default_material=0, mins=mins,
                sizes=sizes, ops=ops, aabbs=aabbs, flags=0, camera=cam)


def serving(seed=5, duration=2.0, rate=4000.0, edit_fraction=0.05, speed_voxels=10.0):
    """Poisson request stream (rate per second) along a seeded camera path at
    speed_voxels voxels/s over the terrain_512 field with its 256 brushes.  A mesh
    request picks a chunk uniformly from the 6-LOD clipmap selection around the camera
    at its arrival time; an edit (edit_fraction of requests) drops a brush near the
    camera and is followed, at the same time, by mesh requests for every selected chunk
    its AABB overlaps (the forced re-mesh).  Returns the event list sorted by time:
    (t, "mesh", (min, size)) or (t, "edit", (op, aabb))."""
    rng = np.random.default_rng(seed)
    base = terrain_512()
    cam0 = np.array([rng.uniform(512, 1536), 0.0, rng.uniform(512, 1536)])
    ang = rng.uniform(0, 2 * math.pi)
    vel = np.array([math.cos(ang), 0.0, math.sin(ang)]) * speed_voxels * LEAF_SIZE_SCALE
    R = LOD_MAX_NODE_SIZE
    step = 0.25
    selections = {}

    def selection(t):
        k = int(t / step)
        if k not in selections:
            from . import mesher
            cam = cam0 + vel * (k * step)
            c = (int(cam[0]) // R) * R, (int(cam[2]) // R) * R
            cm = mesher.Clipmap((c[0] - 2 * R, -R, c[1] - 2 * R), (c[0] + 2 * R, R, c[1] + 2 * R))
            mins, sizes = cm.update(cam)["selected"]
            cm.close()
            selections[k] = [(tuple(int(v) for v in m), int(sz)) for m, sz in zip(mins, sizes)]
        return selections[k]

    events = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        sel = selection(t)
        if rng.random() < edit_fraction:
            cam = cam0 + vel * t
            op = np.zeros(1, CSG_OP_DTYPE)[0]
            is_add = rng.random() < 0.5
            r = rng.uniform(2.0, 8.0)
            c = cam / LEAF_SIZE_SCALE + rng.uniform(-32, 32, 3) * np.array([1, 0.5, 1])
            op["type"] = 0 if is_add else 1
            op["brush"] = int(rng.integers(0, 2))
            op["material"] = int(rng.integers(1, 4)) if is_add else MATERIAL_AIR
            op["origin"] = [c[0] + 0.5, c[1] + 0.5, c[2] + 0.5, 0.0]
            op["dims"] = [r, r, r, 0.0] if op["brush"] == 1 else [r, rng.uniform(2, 8), rng.uniform(2, 8), 0.0]
            bb = op_aabb(op)
            events.append((t, "edit", (op, bb)))
            for mn, size in sel:
                if not (mn[0] + size < bb[0] or mn[1] + size < bb[1] or mn[2] + size < bb[2] or mn[0] > bb[3]
                        or mn[1] > bb[4] or mn[2] > bb[5]):
                    events.append((t, "mesh", (mn, size)))
        else:
            mn, size = sel[int(rng.integers(0, len(sel)))]
            events.append((t, "mesh", (mn, size)))
    return dict(name="serving", desc=base["desc"], seed=base["seed"], default_material=0, ops=base["ops"],
                aabbs=base["aabbs"], events=events, duration=duration, rate=rate, flags=0,
                mins=np.array([e[2][0] for e in events if e[1] == "mesh"], np.int32),
                sizes=np.array([e[2][1] for e in events if e[1] == "mesh"], np.int32))


# CHILD_MIN_OFFSETS, src/volume_constants.h:24-35: child / neighbour i at (bit 2, bit 1, bit 0) of i
CHILD_MIN_OFFSETS = [((i >> 2) & 1, (i >> 1) & 1, i & 1) for i in range(8)]


def seam_sources(nodes):
    """The seam node lists GenerateClipmapSeamMesh gathers for each active clipmap node
    (src/clipmap.cpp:579-597): for neighbour i at host.min + host.size * CHILD_MIN_OFFSETS[i],
    every active node that contains that point or whose min lies in the neighbour's region
    (FindActiveNodes, src/clipmap.cpp:1449-1481; the active nodes partition the meshed world).
    nodes: list of (min, size, seams) with seams a SEAM_DTYPE array.  Returns, per node, the
    list of source dicts {min, size, neighbour, nodes} for lvn_generate_seam_mesh."""
    index = {(tuple(int(v) for v in mn), int(size)): i for i, (mn, size, _) in enumerate(nodes)}
    sizes = sorted({int(size) for _, size, _ in nodes})
    inside = {}
    for i, (mn, size, _) in enumerate(nodes):
        for S in sizes:
            if S >= size:
                key = (tuple((int(v) // S) * S for v in mn), S)
                inside.setdefault(key, []).append(i)
    out = []
    for hmin, hsize, _ in nodes:
        hsize = int(hsize)
        srcs = []
        for i, off in enumerate(CHILD_MIN_OFFSETS):
            nmin = tuple(int(hmin[a]) + hsize * off[a] for a in range(3))
            found = set(inside.get((nmin, hsize), []))
            for S in sizes:
                if S > hsize:
                    j = index.get((tuple((v // S) * S for v in nmin), S))
                    if j is not None:
                        found.add(j)
            for j in sorted(found):
                mn, size, seams = nodes[j]
                if len(seams):
                    srcs.append(dict(min=tuple(int(v) for v in mn), size=int(size), neighbour=i, nodes=seams))
        out.append(srcs)
    return out


WORKLOADS = {"plain_cpu": plain_cpu, "terrain_512": terrain_512, "precision": precision, "clipmap": clipmap,
             "serving": serving}
