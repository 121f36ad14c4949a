"""ctypes mirrors of the POD types in include/lvn_mesher.h.

Byte layouts equal the reference's structs: CSGOperationInfo (src/compute.h:16-24),
MeshVertex (src/render_types.h:24-38), SeamNodeInfo (src/compute.h:26-31).
Importing this module loads nothing.
"""
import ctypes as C

import numpy as np

LVN_OK = 0
LVN_ERR_INVALID = -1
LVN_ERR_NOMEM = -2
LVN_ERR_CAPACITY = -3
LVN_ERR_DEVICE = -4
LVN_ERR_STATE = -5
LVN_ERR_OVERFLOW = -6

ERROR_NAMES = {
    LVN_OK: "LVN_OK",
    LVN_ERR_INVALID: "LVN_ERR_INVALID",
    LVN_ERR_NOMEM: "LVN_ERR_NOMEM",
    LVN_ERR_CAPACITY: "LVN_ERR_CAPACITY",
    LVN_ERR_DEVICE: "LVN_ERR_DEVICE",
    LVN_ERR_STATE: "LVN_ERR_STATE",
    LVN_ERR_OVERFLOW: "LVN_ERR_OVERFLOW",
}

MATERIAL_NONE = 200
MATERIAL_AIR = 201

FIELD_HEIGHTFIELD = 0
FIELD_TERRAIN_CAVES = 1
FIELD_LEVEN_TERRAIN = 2

CTX_FP16_DENSITY = 1
CTX_PRECISE_QEF = 2

# lists of the clipmap request producer (lvn_clipmap_list)
CLIPMAP_TO_MESH, CLIPMAP_RELEASED, CLIPMAP_SELECTED, CLIPMAP_ACTIVE, CLIPMAP_TOUCHED = 0, 1, 2, 3, 4

# limits of the allocation-free single-chunk form (lvn_generate_chunk_mesh_noalloc)
NOALLOC_MAX_EDGES = 131072
NOALLOC_MAX_VOXELS = 65536
NOALLOC_MAX_OPS = 1024
NOALLOC_SCRATCH_BYTES = 17 << 20


class CsgOp(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("brush", C.c_int32),
        ("material", C.c_int32),
        ("rotate_y", C.c_float),
        ("origin", C.c_float * 4),
        ("dims", C.c_float * 4),
    ]


class Vertex(C.Structure):
    _fields_ = [("pos", C.c_float * 4), ("normal", C.c_float * 4), ("colour", C.c_float * 4)]


class SeamNode(C.Structure):
    _fields_ = [
        ("local_min_and_material", C.c_int32 * 4),
        ("pos", C.c_float * 4),
        ("normal", C.c_float * 4),
    ]


class FieldDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("height_octaves", C.c_int32),
        ("height_frequency", C.c_float),
        ("height_lacunarity", C.c_float),
        ("height_gain", C.c_float),
        ("height_amplitude", C.c_float),
        ("cave_octaves", C.c_int32),
        ("cave_frequency", C.c_float),
        ("cave_lacunarity", C.c_float),
        ("cave_gain", C.c_float),
        ("cave_threshold", C.c_float),
        ("reserved", C.c_int32),
    ]


class StageDump(C.Structure):
    _fields_ = [
        ("density", C.POINTER(C.c_float)),
        ("materials", C.POINTER(C.c_uint8)),
        ("edge_ids", C.POINTER(C.c_int32)),
        ("edge_data", C.POINTER(C.c_float)),
        ("edge_cap", C.c_int32),
        ("num_edges", C.c_int32),
        ("voxel_codes", C.POINTER(C.c_int32)),
        ("voxel_edge_masks", C.POINTER(C.c_int32)),
        ("voxel_materials", C.POINTER(C.c_int32)),
        ("voxel_positions", C.POINTER(C.c_float)),
        ("voxel_normals", C.POINTER(C.c_float)),
        ("voxel_cap", C.c_int32),
        ("num_voxels", C.c_int32),
    ]


assert C.sizeof(CsgOp) == 48
assert C.sizeof(Vertex) == 48
assert C.sizeof(SeamNode) == 48
assert C.sizeof(FieldDesc) == 48

REQ_MESH = 0
REQ_EDIT = 1


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("min", C.c_int32 * 3),
        ("size", C.c_int32),
        ("aabb_min", C.c_int32 * 3),
        ("aabb_max", C.c_int32 * 3),
        ("pad", C.c_int32),
        ("op", CsgOp),
    ]


class Result(C.Structure):
    _fields_ = [
        ("ticket", C.c_int64),
        ("kind", C.c_int32),
        ("status", C.c_int32),
        ("nv", C.c_int32),
        ("nt", C.c_int32),
        ("ns", C.c_int32),
        ("cache_hit", C.c_int32),
        ("submit_s", C.c_double),
        ("complete_s", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("density_ysplit", C.c_int32),
        ("persistent_per_cu", C.c_int32),
        ("row_blocks", C.c_int32),
        ("voxel_blocks", C.c_int32),
        ("cross_lanes", C.c_int32),
        ("density_per_cu", C.c_int32),
        ("height_lines", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]


class SeamSource(C.Structure):
    _fields_ = [
        ("min", C.c_int32 * 3),
        ("size", C.c_int32),
        ("neighbour", C.c_int32),
        ("count", C.c_int32),
        ("nodes", C.c_void_p),
    ]


class SeamReq(C.Structure):
    _fields_ = [
        ("host_min", C.c_int32 * 3),
        ("host_size", C.c_int32),
        ("colour", C.c_float * 4),
        ("first_source", C.c_int32),
        ("n_sources", C.c_int32),
    ]


class SimplifyOptions(C.Structure):
    _fields_ = [
        ("edge_fraction", C.c_float),
        ("max_iterations", C.c_int32),
        ("target_percentage", C.c_float),
        ("max_error", C.c_float),
        ("max_edge_size", C.c_float),
        ("min_angle_cosine", C.c_float),
        ("seed", C.c_uint32),
        ("reserved", C.c_int32),
    ]


def simplify_options_per_voxel(max_error=5.0, max_edge_len=2.5, min_cos_angle=0.7, **kw):
    """Options for lvn_batch_simplify: leven's meshMaxError, meshMaxEdgeLen and meshMinCosAngle
    (src/options.h:14-16), which the library scales per chunk by its voxel size."""
    o = SimplifyOptions(edge_fraction=0.125, max_iterations=10, target_percentage=0.05, max_error=max_error,
                        max_edge_size=max_edge_len, min_angle_cosine=min_cos_angle, seed=42, reserved=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def simplify_options(chunk_size, max_error=5.0, max_edge_len=2.5, min_cos_angle=0.7, **kw):
    """MeshSimplificationOptions as leven sets them for a chunk of `chunk_size` world units
    (src/clipmap.cpp:453-462): error and edge limits scale with the voxel size
    4 * chunk_size / 256 (options.h defaults 5, 2.5, 0.7)."""
    leaf = 4.0 * (chunk_size // 256)
    o = SimplifyOptions(edge_fraction=0.125, max_iterations=10, target_percentage=0.05, max_error=max_error * leaf,
                        max_edge_size=max_edge_len * leaf, min_angle_cosine=min_cos_angle, seed=42, reserved=0)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class SeamOut(C.Structure):
    _fields_ = [
        ("v", C.c_void_p),
        ("tri", C.c_void_p),
        ("v_cap", C.c_int32),
        ("tri_cap", C.c_int32),
        ("nv", C.c_int32),
        ("nt", C.c_int32),
        ("status", C.c_int32),
        ("reserved", C.c_int32),
    ]


assert C.sizeof(Request) == 96
assert C.sizeof(Result) == 48
assert C.sizeof(SeamSource) == 32 and C.sizeof(SeamReq) == 40 and C.sizeof(SeamOut) == 40
assert C.sizeof(SimplifyOptions) == 32


def seam_sources(sources):
    """ctypes array of SeamSource from dicts {min, size, neighbour, nodes (SEAM_DTYPE array or None)}.
    The caller keeps the node arrays alive while the array is in use."""
    arr = (SeamSource * max(1, len(sources)))()
    for i, s in enumerate(sources):
        nodes = s["nodes"]
        arr[i].min[:] = [int(v) for v in s["min"]]
        arr[i].size = int(s["size"])
        arr[i].neighbour = int(s["neighbour"])
        if nodes is None:  # the context's cached chunk supplies them
            arr[i].count, arr[i].nodes = -1, None
            continue
        assert nodes.dtype == SEAM_DTYPE and nodes.flags["C_CONTIGUOUS"]
        arr[i].count = len(nodes)
        arr[i].nodes = nodes.ctypes.data if len(nodes) else None
    return arr

RESULT_DTYPE = np.dtype([("ticket", "<i8"), ("kind", "<i4"), ("status", "<i4"), ("nv", "<i4"), ("nt", "<i4"),
                         ("ns", "<i4"), ("cache_hit", "<i4"), ("submit_s", "<f8"), ("complete_s", "<f8")])

# numpy views of the same layouts
CSG_OP_DTYPE = np.dtype(
    [("type", "<i4"), ("brush", "<i4"), ("material", "<i4"), ("rotate_y", "<f4"),
     ("origin", "<f4", 4), ("dims", "<f4", 4)]
)
VERTEX_DTYPE = np.dtype([("pos", "<f4", 4), ("normal", "<f4", 4), ("colour", "<f4", 4)])
SEAM_DTYPE = np.dtype([("local_min_and_material", "<i4", 4), ("pos", "<f4", 4), ("normal", "<f4", 4)])
assert CSG_OP_DTYPE.itemsize == 48 and VERTEX_DTYPE.itemsize == 48 and SEAM_DTYPE.itemsize == 48


def field_desc(kind=FIELD_HEIGHTFIELD, height_octaves=6, height_frequency=1.0 / 128.0,
               height_lacunarity=2.0, height_gain=0.5, height_amplitude=24.0,
               cave_octaves=5, cave_frequency=1.0 / 64.0, cave_lacunarity=2.0, cave_gain=0.5,
               cave_threshold=0.3):
    return FieldDesc(kind, height_octaves, height_frequency, height_lacunarity, height_gain,
                     height_amplitude, cave_octaves, cave_frequency, cave_lacunarity, cave_gain,
                     cave_threshold, 0)


def ops_array(ops):
    """numpy structured array (CSG_OP_DTYPE) -> contiguous ctypes pointer-compatible buffer."""
    arr = np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
    return arr


def ptr(arr, ctype):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ctype))


def vptr(arr):
    """void pointer to a contiguous numpy array (or None).  The pointer object keeps the array
    alive, so a temporary array passed straight into a call is not freed before the call."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.c_void_p)
