"""ctypes binding of the CPU oracle (oracle/liblvn_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing measured
or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from leven_amd.abi import (CSG_OP_DTYPE, SEAM_DTYPE, VERTEX_DTYPE, FieldDesc, ptr, seam_sources,  # noqa: F401
                           vptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblvn_oracle.so")

_lib = None


def build(force=False):
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    cmd = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32, i32, u32, f64 = C.c_float, C.c_int32, C.c_uint32, C.c_double
        P = C.POINTER
        vp = C.c_void_p
        sig = {
            "orc_perm_table": (None, [i32, vp]),
            "orc_base_perm512": (None, [vp]),
            "orc_shuffled_perm512": (None, [i32, vp]),
            "orc_noise_hash": (u32, [i32, i32, i32]),
            "orc_oat_hash": (u32, [C.c_char_p, i32]),
            "orc_splitmix64": (None, [C.c_uint64, i32, vp]),
            "orc_grad_values": (None, [vp]),
            "orc_snoise2": (f32, [vp, f32, f32]),
            "orc_snoise3": (f32, [vp, f32, f32, f32]),
            "orc_fbm2": (f32, [vp, i32, f32, f32, f32, f32, f32]),
            "orc_fbm3": (f32, [vp, i32, f32, f32, f32, f32, f32, f32]),
            "orc_ridged_multifractal": (f32, [vp, i32, f32, f32, f32, f32, f32]),
            "orc_leven_terrain": (f32, [vp, f32, f32]),
            "orc_density": (f32, [P(FieldDesc), vp, f32, f32, f32]),
            "orc_density_f64": (f64, [P(FieldDesc), vp, f64, f64, f64]),
            "orc_brush_density": (f32, [vp, f32, f32, f32]),
            "orc_code_for_position": (u32, [i32, i32, i32, i32, i32]),
            "orc_position_for_code": (None, [u32, i32, vp]),
            "orc_qef_solve_points_f32": (None, [vp, vp, i32, vp]),
            "orc_qef_solve_points_f64": (None, [vp, vp, i32, vp]),
            "orc_chunk_new": (vp, [i32]),
            "orc_chunk_free": (None, [vp]),
            "orc_chunk_generate": (i32, [vp, P(FieldDesc), vp, vp, i32, u32, vp, i32]),
            "orc_chunk_generate_from_density": (i32, [vp, P(FieldDesc), vp, vp, i32, u32, vp]),
            "orc_chunk_apply_csg": (i32, [vp, vp, i32]),
            "orc_chunk_build": (i32, [vp]),
            "orc_chunk_counts": (None, [vp, vp]),
            "orc_chunk_get_density": (None, [vp, vp]),
            "orc_chunk_get_materials": (None, [vp, vp]),
            "orc_chunk_get_edges": (None, [vp, vp, vp]),
            "orc_chunk_get_voxels": (None, [vp, vp, vp, vp]),
            "orc_chunk_get_positions": (None, [vp, vp, vp]),
            "orc_chunk_get_positions_f64": (None, [vp, vp]),
            "orc_chunk_get_normals_f64": (None, [vp, vp]),
            "orc_chunk_fp64_pipeline": (None, [vp, vp, vp, vp]),
            "orc_chunk_get_triangles": (None, [vp, vp]),
            "orc_chunk_get_vertices": (None, [vp, vp]),
            "orc_chunk_get_seams": (None, [vp, vp]),
            "orc_mesh_chunks": (f64, [P(FieldDesc), vp, i32, i32, vp, vp, u32, vp, i32, vp, i32, vp, vp]),
            "orc_seam_mesh": (i32, [vp, i32, vp, i32, vp, i32, vp, i32, vp, vp, i32, vp]),
            "orc_seam_meshes": (f64, [i32, vp, vp, i32, i32, vp, vp]),
            "orc_seam_mesh_leaves": (i32, [vp, i32, vp, i32, i32, i32, vp, vp]),
            "orc_simplify_mesh": (i32, [vp, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp]),
            "orc_simplify_qef": (C.c_float, [vp, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def perm_table(seed):
    out = np.zeros(65536, dtype=np.uint8)
    lib().orc_perm_table(int(seed), vptr(out))
    return out


def base_perm512():
    out = np.zeros(512, dtype=np.int32)
    lib().orc_base_perm512(vptr(out))
    return out


def shuffled_perm512(seed):
    out = np.zeros(512, dtype=np.int32)
    lib().orc_shuffled_perm512(int(seed), vptr(out))
    return out


def oat_hash(data: bytes):
    return lib().orc_oat_hash(data, len(data))


def splitmix64(state, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_splitmix64(state, n, vptr(out))
    return out


def noise_hash(x, y, seed):
    return lib().orc_noise_hash(x, y, seed)


def grad_values():
    out = np.zeros(3, dtype=np.float32)
    lib().orc_grad_values(vptr(out))
    return out


def snoise2(perm, x, y):
    return lib().orc_snoise2(vptr(perm), x, y)


def snoise3(perm, x, y, z):
    return lib().orc_snoise3(vptr(perm), x, y, z)


def density(desc, perm, x, y, z):
    return lib().orc_density(C.byref(desc), vptr(perm), x, y, z)


def qef_solve_f32(positions, normals):
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 4)
    out = np.zeros(4, dtype=np.float32)
    lib().orc_qef_solve_points_f32(vptr(p), vptr(n), len(p), vptr(out))
    return out


def qef_solve_f64(positions, normals):
    p = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 4)
    n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 4)
    out = np.zeros(4, dtype=np.float64)
    lib().orc_qef_solve_points_f64(vptr(p), vptr(n), len(p), vptr(out))
    return out


class Chunk:
    """One chunk of the oracle pipeline; all arrays are numpy copies."""

    def __init__(self, voxels_per_chunk=64):
        self.n = voxels_per_chunk
        self.f = voxels_per_chunk + 2
        self._h = lib().orc_chunk_new(voxels_per_chunk)
        if not self._h:
            raise ValueError("unsupported voxels_per_chunk")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_chunk_free(self._h)
            self._h = None

    def generate(self, desc, perm, min_xyz, size, default_material=0, ops=None):
        mn = np.ascontiguousarray(min_xyz, dtype=np.int32)
        o = None if ops is None or len(ops) == 0 else np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
        lib().orc_chunk_generate(self._h, C.byref(desc), vptr(perm), vptr(mn), int(size),
                                 int(default_material), vptr(o), 0 if o is None else len(o))
        return self

    def generate_from_density(self, desc, perm, min_xyz, size, density, default_material=0):
        mn = np.ascontiguousarray(min_xyz, dtype=np.int32)
        d = np.ascontiguousarray(density, dtype=np.float32).reshape(-1)
        assert d.size == self.f ** 3
        lib().orc_chunk_generate_from_density(self._h, C.byref(desc), vptr(perm), vptr(mn), int(size),
                                              int(default_material), vptr(d))
        return self

    def apply_csg(self, ops):
        o = np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
        lib().orc_chunk_apply_csg(self._h, vptr(o), len(o))
        return self

    def build(self):
        lib().orc_chunk_build(self._h)
        return self

    def counts(self):
        k = np.zeros(6, dtype=np.int32)
        lib().orc_chunk_counts(self._h, vptr(k))
        return dict(edges=int(k[0]), voxels=int(k[1]), quads=int(k[2]), vertices=int(k[3]),
                    triangles=int(k[4]), seams=int(k[5]))

    def density(self):
        out = np.zeros(self.f ** 3, dtype=np.float32)
        lib().orc_chunk_get_density(self._h, vptr(out))
        return out

    def materials(self):
        out = np.zeros(self.f ** 3, dtype=np.uint8)
        lib().orc_chunk_get_materials(self._h, vptr(out))
        return out

    def edges(self):
        e = self.counts()["edges"]
        ids = np.zeros(e, dtype=np.int32)
        data = np.zeros((e, 4), dtype=np.float32)
        lib().orc_chunk_get_edges(self._h, vptr(ids), vptr(data))
        return ids, data

    def voxels(self):
        v = self.counts()["voxels"]
        codes = np.zeros(v, dtype=np.int32)
        masks = np.zeros(v, dtype=np.int32)
        mats = np.zeros(v, dtype=np.int32)
        lib().orc_chunk_get_voxels(self._h, vptr(codes), vptr(masks), vptr(mats))
        return codes, masks, mats

    def positions(self):
        v = self.counts()["voxels"]
        pos = np.zeros((v, 4), dtype=np.float32)
        nrm = np.zeros((v, 4), dtype=np.float32)
        lib().orc_chunk_get_positions(self._h, vptr(pos), vptr(nrm))
        return pos, nrm

    def positions_f64(self):
        v = self.counts()["voxels"]
        pos = np.zeros((v, 4), dtype=np.float64)
        lib().orc_chunk_get_positions_f64(self._h, vptr(pos))
        return pos

    def normals_f64(self):
        v = self.counts()["voxels"]
        nrm = np.zeros((v, 4), dtype=np.float64)
        lib().orc_chunk_get_normals_f64(self._h, vptr(nrm))
        return nrm

    def fp64_pipeline(self):
        """(positions, normals, moved crossings) of the vertex computation in FP64 end to end
        on this chunk's topology."""
        v = self.counts()["voxels"]
        pos = np.zeros((v, 4), dtype=np.float64)
        nrm = np.zeros((v, 4), dtype=np.float64)
        moved = C.c_int32(0)
        lib().orc_chunk_fp64_pipeline(self._h, vptr(pos), vptr(nrm), C.byref(moved))
        return pos, nrm, int(moved.value)

    def triangles(self):
        t = self.counts()["quads"] * 2
        out = np.zeros((t, 3), dtype=np.int32)
        lib().orc_chunk_get_triangles(self._h, vptr(out))
        return out

    def vertices(self):
        v = self.counts()["vertices"]
        out = np.zeros(v, dtype=VERTEX_DTYPE)
        lib().orc_chunk_get_vertices(self._h, vptr(out))
        return out

    def seams(self):
        s = self.counts()["seams"]
        out = np.zeros(s, dtype=SEAM_DTYPE)
        lib().orc_chunk_get_seams(self._h, vptr(out))
        return out


def mesh_chunks(desc, perm, voxels_per_chunk, mins, sizes, default_material=0, ops=None, aabbs=None,
                threads=1):
    """Time the oracle over many chunks (CPU baseline).  Returns (wall_s, per_chunk_s, totals)."""
    mins = np.ascontiguousarray(mins, dtype=np.int32).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
    n = len(sizes)
    o = None if ops is None or len(ops) == 0 else np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
    bb = None if o is None else np.ascontiguousarray(aabbs, dtype=np.int32).reshape(-1, 6)
    secs = np.zeros(n, dtype=np.float64)
    totals = np.zeros(4, dtype=np.int64)
    wall = lib().orc_mesh_chunks(C.byref(desc), vptr(perm), int(voxels_per_chunk), n, vptr(mins),
                                 vptr(sizes), int(default_material), vptr(o), 0 if o is None else len(o),
                                 vptr(bb), int(threads), vptr(secs), vptr(totals))
    return wall, secs, totals


def seam_mesh(host_min, host_size, sources, colour=(1.0, 1.0, 1.0), voxels_per_chunk=64):
    """Seam mesh of one host node from source dicts {min, size, neighbour, nodes}.
    Returns (status, vertices VERTEX_DTYPE, triangles int32 (T, 3))."""
    src = seam_sources(sources)
    mn = np.ascontiguousarray(host_min, dtype=np.int32)
    col = np.ascontiguousarray(colour, dtype=np.float32)
    nv, nt = C.c_int32(0), C.c_int32(0)
    rc = lib().orc_seam_mesh(vptr(mn), int(host_size), C.cast(src, C.c_void_p), len(sources), vptr(col),
                             int(voxels_per_chunk), None, 0, C.byref(nv), None, 0, C.byref(nt))
    if rc not in (0, -3):
        return rc, np.zeros(0, dtype=VERTEX_DTYPE), np.zeros((0, 3), dtype=np.int32)
    v = np.zeros(nv.value, dtype=VERTEX_DTYPE)
    t = np.zeros((nt.value, 3), dtype=np.int32)
    rc = lib().orc_seam_mesh(vptr(mn), int(host_size), C.cast(src, C.c_void_p), len(sources), vptr(col),
                             int(voxels_per_chunk), vptr(v), len(v), C.byref(nv), vptr(t), len(t), C.byref(nt))
    return rc, v, t


def seam_mesh_leaves(host_min, host_size, sources, voxels_per_chunk=64):
    """The leaf (min xyz, size) behind each vertex of the seam mesh, int32 (nv, 4)."""
    src = seam_sources(sources)
    mn = np.ascontiguousarray(host_min, dtype=np.int32)
    cap = max(1, sum(len(s["nodes"]) for s in sources))
    out = np.zeros((cap, 4), dtype=np.int32)
    n = C.c_int32(0)
    rc = lib().orc_seam_mesh_leaves(vptr(mn), int(host_size), C.cast(src, C.c_void_p), len(sources),
                                    int(voxels_per_chunk), cap, C.byref(n), vptr(out))
    assert rc == 0, rc
    return out[:n.value].copy()


def seam_meshes(reqs, sources, voxels_per_chunk=64, threads=1, per_seam=False):
    """Time the oracle over many seams (CPU baseline): reqs is a ctypes SeamReq array, sources
    a ctypes SeamSource array.  Returns (wall_s, totals [vertices, triangles, failed]), and
    with per_seam an int64 (n, 4) array of status, vertices, triangles and mesh_checksum."""
    totals = np.zeros(3, dtype=np.int64)
    each = np.zeros((len(reqs), 4), dtype=np.int64) if per_seam else None
    wall = lib().orc_seam_meshes(len(reqs), C.cast(reqs, C.c_void_p), C.cast(sources, C.c_void_p),
                                 int(voxels_per_chunk), int(threads), vptr(totals), vptr(each))
    return (wall, totals, each) if per_seam else (wall, totals)


def mesh_checksum(vertices, triangles):
    """The oracle's per-seam checksum of a mesh: the vertex bytes then the triangle bytes as
    32-bit words w, sum of w[i] * (i + 1) mod 2^64."""
    w = np.concatenate([np.frombuffer(np.ascontiguousarray(vertices).tobytes(), dtype=np.uint32),
                        np.frombuffer(np.ascontiguousarray(triangles, dtype=np.int32).tobytes(), dtype=np.uint32)])
    return int((w.astype(np.uint64) * np.arange(1, len(w) + 1, dtype=np.uint64)).sum(dtype=np.uint64))


def simplify_qef(p0, n0, p1, n1):
    """The simplifier's QEF on two endpoints (4-vectors): (position float32[4], error)."""
    a = [np.ascontiguousarray(x, dtype=np.float32).reshape(4) for x in (p0, n0, p1, n1)]
    out = np.zeros(4, dtype=np.float32)
    err = lib().orc_simplify_qef(vptr(a[0]), vptr(a[1]), vptr(a[2]), vptr(a[3]), vptr(out))
    return out, np.float32(err)


def simplify_mesh(vertices, triangles, offset, options):
    """ngMeshSimplifier restated: (vertices VERTEX_DTYPE, triangles int32 (T, 3), stats) of the
    simplified mesh; offset = the chunk centre (4 floats, w 0); options an abi.SimplifyOptions."""
    v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    off = np.ascontiguousarray(offset, dtype=np.float32).reshape(4)
    vo = np.zeros(max(len(v), 1), dtype=VERTEX_DTYPE)
    to = np.zeros((max(len(t), 1), 3), dtype=np.int32)
    nv, nt = C.c_int32(), C.c_int32()
    stats = np.zeros(4, dtype=np.int32)
    rc = lib().orc_simplify_mesh(vptr(v), len(v), vptr(t), len(t), vptr(off), C.byref(options), vptr(vo),
                                 C.byref(nv), vptr(to), C.byref(nt), vptr(stats))
    assert rc == 0
    return vo[:nv.value].copy(), to[:nt.value].copy(), stats
