"""Python binding of the C ABI in include/lvn_mesher.h (leven_amd/liblvn_mesher.so).

Mirrors the reference's compute interface (src/compute.h:35-77):
``init`` = Compute_Initialise, ``Context`` = Compute_MeshGenContext.  There is no
CPU fallback: a missing or unloadable library raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import (CSG_OP_DTYPE, ERROR_NAMES, LVN_ERR_CAPACITY, LVN_OK, REQ_EDIT, REQ_MESH, RESULT_DTYPE, SEAM_DTYPE,
                  VERTEX_DTYPE, CsgOp, FieldDesc, Request, SeamOut, SeamReq, SeamSource, SimplifyOptions, StageDump,
                  Tuning, ptr, seam_sources, vptr)

HERE = os.path.dirname(os.path.abspath(__file__))
# LVN_LIB_PATH selects another build of the library (e.g. to compare two builds)
LIB_PATH = os.environ.get("LVN_LIB_PATH") or os.path.join(HERE, "liblvn_mesher.so")

# every symbol declared in include/lvn_mesher.h
EXPORTS = [
    "lvn_init", "lvn_shutdown", "lvn_set_noise_seed", "lvn_store_csg_op", "lvn_clear_csg_ops",
    "lvn_get_perm_table", "lvn_ctx_create", "lvn_ctx_destroy", "lvn_ctx_voxels_per_chunk", "lvn_apply_csg",
    "lvn_free_chunk", "lvn_chunk_has_surface", "lvn_generate_chunk_mesh", "lvn_debug_chunk_stages",
    "lvn_batch_create", "lvn_batch_destroy", "lvn_batch_set_chunks", "lvn_batch_run", "lvn_batch_wait",
    "lvn_batch_run_timed", "lvn_batch_stage_name", "lvn_batch_counts", "lvn_batch_get_mesh", "lvn_batch_work",
    "lvn_batch_get_tuning", "lvn_batch_set_tuning", "lvn_batch_autotune",
    "lvn_server_create", "lvn_server_destroy", "lvn_server_submit", "lvn_server_poll", "lvn_server_drain",
    "lvn_server_get_mesh", "lvn_server_stats", "lvn_server_submit_many", "lvn_server_set_tuning", "lvn_generate_seam_mesh", "lvn_generate_seam_meshes",
    "lvn_seam_device_mesh", "lvn_batch_seam_meshes", "lvn_batch_seam_device_mesh",
    "lvn_generate_chunk_mesh_noalloc",
    "lvn_clipmap_create", "lvn_clipmap_destroy", "lvn_clipmap_update", "lvn_clipmap_list", "lvn_clipmap_commit",
    "lvn_clipmap_seams", "lvn_clipmap_edit", "lvn_simplify_mesh", "lvn_batch_simplify",
]

_lib = None


class LvnError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed: {ERROR_NAMES.get(code, code)}")
        self.code = code


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m leven_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        i32, u32 = C.c_int32, C.c_uint32
        vp = C.c_void_p
        sig = {
            "lvn_init": (i32, [C.POINTER(FieldDesc), i32, u32]),
            "lvn_shutdown": (i32, []),
            "lvn_set_noise_seed": (i32, [i32]),
            "lvn_store_csg_op": (i32, [vp, vp, vp]),
            "lvn_clear_csg_ops": (i32, []),
            "lvn_get_perm_table": (i32, [vp]),
            "lvn_ctx_create": (i32, [i32, u32, C.POINTER(vp)]),
            "lvn_ctx_destroy": (None, [vp]),
            "lvn_ctx_voxels_per_chunk": (i32, [vp]),
            "lvn_apply_csg": (i32, [vp, vp, i32, vp, i32]),
            "lvn_free_chunk": (i32, [vp, vp, i32]),
            "lvn_chunk_has_surface": (i32, [vp, vp, i32, C.POINTER(i32)]),
            "lvn_generate_chunk_mesh": (i32, [vp, vp, i32, vp, i32, C.POINTER(i32), vp, i32, C.POINTER(i32), vp,
                                              i32, C.POINTER(i32)]),
            "lvn_generate_chunk_mesh_noalloc": (i32, [i32, u32, vp, i32, vp, i32, C.POINTER(i32), vp, i32,
                                                      C.POINTER(i32), vp, i32, C.POINTER(i32)]),
            "lvn_debug_chunk_stages": (i32, [vp, vp, i32, vp, i32, C.POINTER(StageDump)]),
            "lvn_clipmap_create": (i32, [i32, vp, vp, C.POINTER(vp)]),
            "lvn_clipmap_destroy": (None, [vp]),
            "lvn_clipmap_update": (i32, [vp, vp, vp, vp]),
            "lvn_clipmap_list": (i32, [vp, i32, vp, i32, C.POINTER(i32)]),
            "lvn_clipmap_commit": (i32, [vp, i32, vp, vp]),
            "lvn_clipmap_seams": (i32, [vp, vp, i32, vp, i32]),
            "lvn_clipmap_edit": (i32, [vp, vp, vp, C.POINTER(i32)]),
            "lvn_simplify_mesh": (i32, [vp, vp, C.POINTER(i32), vp, C.POINTER(i32), vp, C.POINTER(SimplifyOptions),
                                        vp]),
            "lvn_batch_simplify": (i32, [vp, C.POINTER(SimplifyOptions), vp]),
            "lvn_batch_create": (i32, [i32, i32, i32, i32, u32, C.POINTER(vp)]),
            "lvn_batch_destroy": (None, [vp]),
            "lvn_batch_set_chunks": (i32, [vp, i32, vp]),
            "lvn_batch_run": (i32, [vp]),
            "lvn_batch_wait": (i32, [vp]),
            "lvn_batch_run_timed": (i32, [vp, vp, i32]),
            "lvn_batch_stage_name": (C.c_char_p, [i32]),
            "lvn_batch_counts": (i32, [vp, vp]),
            "lvn_batch_get_mesh": (i32, [vp, i32, vp, i32, C.POINTER(i32), vp, i32, C.POINTER(i32), vp, i32,
                                         C.POINTER(i32)]),
            "lvn_batch_work": (i32, [vp, vp]),
            "lvn_batch_get_tuning": (i32, [vp, vp]),
            "lvn_batch_set_tuning": (i32, [vp, vp]),
            "lvn_batch_autotune": (i32, [vp, i32, vp]),
            "lvn_server_create": (i32, [i32, C.c_uint64, i32, i32, i32, i32, u32, C.POINTER(vp)]),
            "lvn_server_destroy": (None, [vp]),
            "lvn_server_submit": (i32, [vp, C.POINTER(Request), C.POINTER(C.c_int64)]),
            "lvn_server_submit_many": (i32, [vp, i32, vp, vp]),
            "lvn_server_poll": (i32, [vp, vp, i32, C.POINTER(i32)]),
            "lvn_server_drain": (i32, [vp]),
            "lvn_server_get_mesh": (i32, [vp, C.c_int64, vp, i32, C.POINTER(i32), vp, i32, C.POINTER(i32), vp, i32,
                                          C.POINTER(i32)]),
            "lvn_server_stats": (i32, [vp, vp]),
            "lvn_server_set_tuning": (i32, [vp, vp]),
            "lvn_generate_seam_mesh": (i32, [vp, vp, i32, vp, i32, vp, vp, i32, C.POINTER(i32), vp, i32,
                                             C.POINTER(i32)]),
            "lvn_generate_seam_meshes": (i32, [vp, i32, vp, vp, vp]),
            "lvn_seam_device_mesh": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(i32), C.POINTER(i32)]),
            "lvn_batch_seam_meshes": (i32, [vp, i32, vp, vp, vp]),
            "lvn_batch_seam_device_mesh": (i32, [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(i32),
                                                 C.POINTER(i32)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code, what):
    if code != LVN_OK:
        raise LvnError(code, what)
    return code


def _i3(v):
    return np.ascontiguousarray(v, dtype=np.int32).reshape(3)


def init(field=None, seed=0, default_material=0):
    """Compute_Initialise(noiseSeed, defaultMaterial, ...) with a field descriptor."""
    check(lib().lvn_init(C.byref(field) if field is not None else None, int(seed), int(default_material)),
          "lvn_init")


def shutdown():
    check(lib().lvn_shutdown(), "lvn_shutdown")


def set_noise_seed(seed):
    check(lib().lvn_set_noise_seed(int(seed)), "lvn_set_noise_seed")


def store_csg_op(op, aabb_min, aabb_max):
    o = np.ascontiguousarray(op, dtype=CSG_OP_DTYPE).reshape(1)
    check(lib().lvn_store_csg_op(vptr(o), vptr(_i3(aabb_min)), vptr(_i3(aabb_max))), "lvn_store_csg_op")


def clear_csg_ops():
    check(lib().lvn_clear_csg_ops(), "lvn_clear_csg_ops")


def perm_table():
    out = np.zeros(65536, dtype=np.uint8)
    check(lib().lvn_get_perm_table(vptr(out)), "lvn_get_perm_table")
    return out


def generate_chunk_mesh_noalloc(voxels_per_chunk, min_xyz, size, flags=0, v_cap=None, tri_cap=None, s_cap=None):
    """lvn_generate_chunk_mesh_noalloc: one chunk meshed fresh in the library's fixed device
    area (no context, no allocation).  Returns (vertices, triangles int32[T,3], seams)."""
    n3 = voxels_per_chunk ** 3
    v_cap = n3 if v_cap is None else v_cap
    tri_cap = 6 * n3 if tri_cap is None else tri_cap
    s_cap = n3 if s_cap is None else s_cap
    v = np.zeros(max(v_cap, 1), dtype=VERTEX_DTYPE)
    t = np.zeros((max(tri_cap, 1), 3), dtype=np.int32)
    s = np.zeros(max(s_cap, 1), dtype=SEAM_DTYPE)
    nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
    rc = lib().lvn_generate_chunk_mesh_noalloc(int(voxels_per_chunk), int(flags), vptr(_i3(min_xyz)), int(size),
                                               vptr(v), v_cap, C.byref(nv), vptr(t), tri_cap, C.byref(nt), vptr(s),
                                               s_cap, C.byref(ns))
    if rc == LVN_ERR_CAPACITY:
        raise LvnError(rc, f"lvn_generate_chunk_mesh_noalloc (needs {nv.value} vertices, {nt.value} triangles, "
                           f"{ns.value} seam nodes)")
    check(rc, "lvn_generate_chunk_mesh_noalloc")
    return v[:nv.value].copy(), t[:nt.value].copy(), s[:ns.value].copy()


def _seam_request(hosts, sources_per_host):
    """ctypes SeamReq and SeamSource arrays (plus the flat source list that keeps the node
    arrays alive) for hosts = [(min, size, colour)] and their source dicts."""
    flat, reqs = [], (SeamReq * max(1, len(hosts)))()
    for i, ((mn, size, colour), srcs) in enumerate(zip(hosts, sources_per_host)):
        reqs[i].host_min[:] = [int(v) for v in mn]
        reqs[i].host_size = int(size)
        reqs[i].colour[:] = [float(colour[0]), float(colour[1]), float(colour[2]), 0.0]
        reqs[i].first_source = len(flat)
        reqs[i].n_sources = len(srcs)
        flat.extend(srcs)
    return reqs, seam_sources(flat), flat


def _seam_meshes(fn, h, hosts, sources_per_host, v_caps=None, tri_caps=None):
    reqs, src, flat = _seam_request(hosts, sources_per_host)
    n = len(hosts)
    offered = [sum(len(s["nodes"]) if s["nodes"] is not None else 32768 for s in srcs) for srcs in sources_per_host]
    vc = list(v_caps) if v_caps is not None else offered
    tc = list(tri_caps) if tri_caps is not None else [4 * o + 64 for o in offered]
    for _ in range(2):
        bufs, outs = [], (SeamOut * max(1, n))()
        for i in range(n):
            v = np.zeros(max(1, vc[i]), dtype=VERTEX_DTYPE)
            t = np.zeros((max(1, tc[i]), 3), dtype=np.int32)
            bufs.append((v, t))
            outs[i].v, outs[i].tri = v.ctypes.data, t.ctypes.data
            outs[i].v_cap, outs[i].tri_cap = vc[i], tc[i]
        rc = fn(h, n, C.cast(reqs, C.c_void_p), C.cast(src, C.c_void_p), C.cast(outs, C.c_void_p))
        if rc != LVN_ERR_CAPACITY:
            break
        vc = [max(vc[i], outs[i].nv) for i in range(n)]
        tc = [max(tc[i], outs[i].nt) for i in range(n)]
    check(rc, "seam meshes")
    return [(bufs[i][0][:outs[i].nv].copy(), bufs[i][1][:outs[i].nt].copy()) for i in range(n)]


def _seam_meshes_on_device(fn, h, hosts, sources_per_host):
    reqs, src, flat = _seam_request(hosts, sources_per_host)
    n = len(hosts)
    outs = (SeamOut * max(1, n))()
    check(fn(h, n, C.cast(reqs, C.c_void_p), C.cast(src, C.c_void_p), C.cast(outs, C.c_void_p)), "seam meshes")
    return [(outs[i].nv, outs[i].nt) for i in range(n)]


def _seam_device_mesh(fn, h, i):
    v, t, nv, nt = C.c_void_p(), C.c_void_p(), C.c_int32(), C.c_int32()
    check(fn(h, int(i), C.byref(v), C.byref(t), C.byref(nv), C.byref(nt)), "seam device mesh")
    return v.value, t.value, nv.value, nt.value


class Context:
    """Compute_MeshGenContext over the C ABI."""

    def __init__(self, voxels_per_chunk=64, flags=0):
        h = C.c_void_p()
        check(lib().lvn_ctx_create(int(voxels_per_chunk), int(flags), C.byref(h)), "lvn_ctx_create")
        self._h = h
        self.n = voxels_per_chunk
        self.f = voxels_per_chunk + 2

    def close(self):
        if getattr(self, "_h", None):
            lib().lvn_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def voxels_per_chunk(self):
        return lib().lvn_ctx_voxels_per_chunk(self._h)

    def apply_csg(self, ops, min_xyz, size):
        o = np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
        check(lib().lvn_apply_csg(self._h, vptr(o), len(o), vptr(_i3(min_xyz)), int(size)), "lvn_apply_csg")

    def free_chunk(self, min_xyz, size):
        check(lib().lvn_free_chunk(self._h, vptr(_i3(min_xyz)), int(size)), "lvn_free_chunk")

    def has_surface(self, min_xyz, size):
        out = C.c_int32()
        check(lib().lvn_chunk_has_surface(self._h, vptr(_i3(min_xyz)), int(size), C.byref(out)),
              "lvn_chunk_has_surface")
        return bool(out.value)

    def generate_chunk_mesh(self, min_xyz, size, v_cap=None, tri_cap=None, s_cap=None):
        """Returns (vertices[VERTEX_DTYPE], triangles int32[T,3], seams[SEAM_DTYPE])."""
        n3 = self.n ** 3
        v_cap = n3 if v_cap is None else v_cap
        tri_cap = 6 * n3 if tri_cap is None else tri_cap
        s_cap = n3 if s_cap is None else s_cap
        v = np.zeros(v_cap, dtype=VERTEX_DTYPE)
        t = np.zeros((tri_cap, 3), dtype=np.int32)
        s = np.zeros(s_cap, dtype=SEAM_DTYPE)
        nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
        rc = lib().lvn_generate_chunk_mesh(self._h, vptr(_i3(min_xyz)), int(size), vptr(v), v_cap, C.byref(nv),
                                           vptr(t), tri_cap, C.byref(nt), vptr(s), s_cap, C.byref(ns))
        if rc == LVN_ERR_CAPACITY:
            raise LvnError(rc, f"lvn_generate_chunk_mesh (needs {nv.value} vertices, {nt.value} triangles, "
                               f"{ns.value} seam nodes)")
        check(rc, "lvn_generate_chunk_mesh")
        return v[:nv.value].copy(), t[:nt.value].copy(), s[:ns.value].copy()

    def seam_mesh(self, host_min, host_size, sources, colour=(1.0, 1.0, 1.0)):
        """The seam mesh of one clipmap node (GenerateClipmapSeamMesh, src/clipmap.cpp:573-611) from
        source dicts {min, size, neighbour, nodes}; nodes=None takes the seam nodes of this
        context's cached chunk at (min, size).  Returns (vertices, triangles int32[T, 3])."""
        return self.seam_meshes([(host_min, host_size, colour)], [sources])[0]

    def simplify_mesh(self, vertices, triangles, offset, options):
        """ngMeshSimplifier on one mesh (lvn_simplify_mesh).  Returns (vertices, triangles int32[T, 3],
        (iterations, collapses)); the inputs are not modified."""
        v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE).copy()
        t = np.ascontiguousarray(np.asarray(triangles, dtype=np.int32).reshape(-1, 3)).copy()
        nv, nt = C.c_int32(len(v)), C.c_int32(len(t))
        off = np.ascontiguousarray(offset, dtype=np.float32).reshape(4)
        stats = np.zeros(2, dtype=np.int32)
        check(lib().lvn_simplify_mesh(self._h, vptr(v), C.byref(nv), vptr(t), C.byref(nt), vptr(off),
                                      C.byref(options), vptr(stats)), "lvn_simplify_mesh")
        return v[:nv.value].copy(), t[:nt.value].copy(), (int(stats[0]), int(stats[1]))

    def seam_meshes(self, hosts, sources_per_host, v_caps=None, tri_caps=None):
        """Many seams in one call: hosts = [(min, size, colour)], sources_per_host = [[source dicts]].
        Returns [(vertices, triangles)]; raises LvnError on the first failing seam."""
        return _seam_meshes(lib().lvn_generate_seam_meshes, self._h, hosts, sources_per_host, v_caps, tri_caps)

    def seam_meshes_on_device(self, hosts, sources_per_host):
        """Stitch many seams and leave the meshes in device memory.  Returns [(nv, nt)];
        seam_device_mesh(i) gives their device pointers until the next seam call."""
        return _seam_meshes_on_device(lib().lvn_generate_seam_meshes, self._h, hosts, sources_per_host)

    def seam_device_mesh(self, i):
        """(vertex device pointer, triangle device pointer, nv, nt) of seam i of the last call."""
        return _seam_device_mesh(lib().lvn_seam_device_mesh, self._h, i)

    def debug_stages(self, min_xyz, size, ops=None, edge_cap=None, voxel_cap=None):
        """Every stage of one fresh chunk (density ... vertex positions)."""
        f3 = self.f ** 3
        h = self.n + 1
        edge_cap = 3 * h ** 3 if edge_cap is None else edge_cap
        voxel_cap = self.n ** 3 if voxel_cap is None else voxel_cap
        bufs = dict(
            density=np.zeros(f3, dtype=np.float32),
            materials=np.zeros(f3, dtype=np.uint8),
            edge_ids=np.zeros(edge_cap, dtype=np.int32),
            edge_data=np.zeros((edge_cap, 4), dtype=np.float32),
            voxel_codes=np.zeros(voxel_cap, dtype=np.int32),
            voxel_edge_masks=np.zeros(voxel_cap, dtype=np.int32),
            voxel_materials=np.zeros(voxel_cap, dtype=np.int32),
            voxel_positions=np.zeros((voxel_cap, 4), dtype=np.float32),
            voxel_normals=np.zeros((voxel_cap, 4), dtype=np.float32),
        )
        d = StageDump()
        d.density = ptr(bufs["density"], C.c_float)
        d.materials = ptr(bufs["materials"], C.c_uint8)
        d.edge_ids = ptr(bufs["edge_ids"], C.c_int32)
        d.edge_data = ptr(bufs["edge_data"], C.c_float)
        d.edge_cap = edge_cap
        d.voxel_codes = ptr(bufs["voxel_codes"], C.c_int32)
        d.voxel_edge_masks = ptr(bufs["voxel_edge_masks"], C.c_int32)
        d.voxel_materials = ptr(bufs["voxel_materials"], C.c_int32)
        d.voxel_positions = ptr(bufs["voxel_positions"], C.c_float)
        d.voxel_normals = ptr(bufs["voxel_normals"], C.c_float)
        d.voxel_cap = voxel_cap
        o = None if ops is None or len(ops) == 0 else np.ascontiguousarray(ops, dtype=CSG_OP_DTYPE)
        check(lib().lvn_debug_chunk_stages(self._h, vptr(_i3(min_xyz)), int(size), vptr(o),
                                           0 if o is None else len(o), C.byref(d)), "lvn_debug_chunk_stages")
        e, v = d.num_edges, d.num_voxels
        return dict(
            density=bufs["density"], materials=bufs["materials"],
            edge_ids=bufs["edge_ids"][:e].copy(), edge_data=bufs["edge_data"][:e].copy(),
            voxel_codes=bufs["voxel_codes"][:v].copy(), voxel_edge_masks=bufs["voxel_edge_masks"][:v].copy(),
            voxel_materials=bufs["voxel_materials"][:v].copy(),
            voxel_positions=bufs["voxel_positions"][:v].copy(), voxel_normals=bufs["voxel_normals"][:v].copy(),
        )


NUM_STAGES = 12  # LVN_NUM_STAGES


def chunk_reqs(mins, sizes):
    mins = np.ascontiguousarray(mins, dtype=np.int32).reshape(-1, 3)
    sizes = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1, 1)
    return np.ascontiguousarray(np.hstack([mins, sizes]), dtype=np.int32)


class Clipmap:
    """The clipmap request producer (lvn_clipmap_*): leven's Clipmap::update as chunk and seam
    requests.  Host-side only; runs without a GPU."""

    def __init__(self, world_min, world_max, voxels_per_chunk=64):
        h = C.c_void_p()
        check(lib().lvn_clipmap_create(int(voxels_per_chunk), vptr(_i3(world_min)), vptr(_i3(world_max)),
                                       C.byref(h)), "lvn_clipmap_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lvn_clipmap_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, camera, frustum=None):
        """Returns {to_mesh, released, selected} as (mins int32[n,3], sizes int32[n]) pairs."""
        cam = np.ascontiguousarray(camera, dtype=np.float32).reshape(3)
        fr = None if frustum is None else np.ascontiguousarray(frustum, dtype=np.float32).reshape(24)
        counts = np.zeros(3, dtype=np.int32)
        check(lib().lvn_clipmap_update(self._h, vptr(cam), vptr(fr) if fr is not None else None, vptr(counts)),
              "lvn_clipmap_update")
        return {name: self.list(k) for k, name in enumerate(("to_mesh", "released", "selected"))}

    def list(self, which):
        n = C.c_int32()
        lib().lvn_clipmap_list(self._h, int(which), None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), 4), dtype=np.int32)
        check(lib().lvn_clipmap_list(self._h, int(which), vptr(out), len(out), C.byref(n)), "lvn_clipmap_list")
        out = out[:n.value]
        return out[:, :3].copy(), out[:, 3].copy()

    def commit(self, results):
        """results: int32[n, 3] = (status, triangles, seam nodes) per chunk of the last update's
        to-mesh list.  Returns (SeamReq array, SeamSource array) of the seams to regenerate."""
        r = np.ascontiguousarray(results, dtype=np.int32).reshape(-1, 3)
        counts = np.zeros(2, dtype=np.int32)
        check(lib().lvn_clipmap_commit(self._h, len(r), vptr(r) if len(r) else None, vptr(counts)),
              "lvn_clipmap_commit")
        reqs = (SeamReq * max(1, int(counts[0])))()
        srcs = (SeamSource * max(1, int(counts[1])))()
        check(lib().lvn_clipmap_seams(self._h, C.cast(reqs, C.c_void_p), len(reqs), C.cast(srcs, C.c_void_p),
                                      len(srcs)), "lvn_clipmap_seams")
        return reqs, srcs, int(counts[0]), int(counts[1])

    def edit(self, aabb_min, aabb_max):
        n = C.c_int32()
        check(lib().lvn_clipmap_edit(self._h, vptr(_i3(aabb_min)), vptr(_i3(aabb_max)), C.byref(n)),
              "lvn_clipmap_edit")
        return self.list(4)


class Batch:
    """The batched form: many chunk slots in HBM, one launch per stage, graph replay."""

    def __init__(self, voxels_per_chunk=64, max_chunks=512, edge_cap=0, voxel_cap=0, flags=0):
        h = C.c_void_p()
        check(lib().lvn_batch_create(int(voxels_per_chunk), int(max_chunks), int(edge_cap), int(voxel_cap),
                                     int(flags), C.byref(h)), "lvn_batch_create")
        self._h = h
        self.n = 0
        self.vpc = voxels_per_chunk

    def close(self):
        if getattr(self, "_h", None):
            lib().lvn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_chunks(self, mins, sizes):
        reqs = chunk_reqs(mins, sizes)
        check(lib().lvn_batch_set_chunks(self._h, len(reqs), vptr(reqs)), "lvn_batch_set_chunks")
        self.n = len(reqs)

    def run(self):
        check(lib().lvn_batch_run(self._h), "lvn_batch_run")

    def wait(self):
        check(lib().lvn_batch_wait(self._h), "lvn_batch_wait")

    def run_timed(self):
        ms = np.zeros(NUM_STAGES, dtype=np.float32)
        check(lib().lvn_batch_run_timed(self._h, vptr(ms), NUM_STAGES), "lvn_batch_run_timed")
        names = [lib().lvn_batch_stage_name(i).decode() for i in range(NUM_STAGES)]
        return dict(zip(names, ms.astype(float).tolist()))

    TUNING = ["density_ysplit", "persistent_per_cu", "row_blocks", "voxel_blocks", "cross_lanes", "density_per_cu",
              "height_lines"]

    def tuning(self):
        t = Tuning()
        check(lib().lvn_batch_get_tuning(self._h, C.byref(t)), "lvn_batch_get_tuning")
        return {k: int(getattr(t, k)) for k in self.TUNING}

    def set_tuning(self, **kw):
        t = Tuning()
        check(lib().lvn_batch_get_tuning(self._h, C.byref(t)), "lvn_batch_get_tuning")
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().lvn_batch_set_tuning(self._h, C.byref(t)), "lvn_batch_set_tuning")

    def autotune(self, reps=3):
        """Time candidate launch shapes on the staged chunks and keep the fastest."""
        t = Tuning()
        check(lib().lvn_batch_autotune(self._h, int(reps), C.byref(t)), "lvn_batch_autotune")
        return {k: int(getattr(t, k)) for k in self.TUNING}

    def work(self):
        w = np.zeros(8, dtype=np.uint64)
        check(lib().lvn_batch_work(self._h, vptr(w)), "lvn_batch_work")
        keys = ["density_height", "density_cave", "cross_height", "cross_cave", "cross_brush", "samples", "edges"]
        return {k: int(v) for k, v in zip(keys, w)}

    def counts(self):
        """int32 [n, 6]: edges, voxels, triangles, seams, status, FP64 re-solved vertices."""
        out = np.zeros((max(self.n, 1), 6), dtype=np.int32)
        check(lib().lvn_batch_counts(self._h, vptr(out)), "lvn_batch_counts")
        return out[: self.n]

    def get_mesh(self, i):
        k = self.counts()[i]
        v = np.zeros(max(int(k[1]), 1), dtype=VERTEX_DTYPE)
        t = np.zeros((max(int(k[2]), 1), 3), dtype=np.int32)
        s = np.zeros(max(int(k[3]), 1), dtype=SEAM_DTYPE)
        nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().lvn_batch_get_mesh(self._h, int(i), vptr(v), len(v), C.byref(nv), vptr(t), len(t), C.byref(nt),
                                       vptr(s), len(s), C.byref(ns)), "lvn_batch_get_mesh")
        return v[:nv.value].copy(), t[:nt.value].copy(), s[:ns.value].copy()

    def get_seam_nodes(self, i, counts=None):
        """Chunk i's seam nodes only (lvn_batch_get_mesh without vertex or triangle buffers);
        counts = self.counts() when already at hand."""
        k = (self.counts() if counts is None else counts)[i]
        s = np.zeros(max(int(k[3]), 1), dtype=SEAM_DTYPE)
        nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
        big = 2 ** 31 - 1
        check(lib().lvn_batch_get_mesh(self._h, int(i), None, big, C.byref(nv), None, big, C.byref(nt), vptr(s),
                                       len(s), C.byref(ns)), "lvn_batch_get_mesh")
        return s[:ns.value].copy()

    def simplify(self, options):
        """Simplify every staged chunk's mesh in its slot (lvn_batch_simplify) with per-voxel
        limits (abi.simplify_options_per_voxel).  Returns int32 [n, 4]: vertices, triangles,
        iterations, collapses; get_mesh then returns the simplified meshes."""
        out = np.zeros((max(self.n, 1), 4), dtype=np.int32)
        check(lib().lvn_batch_simplify(self._h, C.byref(options), vptr(out)), "lvn_batch_simplify")
        return out[: self.n]

    def seam_meshes(self, hosts, sources_per_host):
        """Seams of this batch's chunks (sources with nodes=None are staged chunks, read from
        their slots).  Returns [(vertices, triangles)]."""
        return _seam_meshes(lib().lvn_batch_seam_meshes, self._h, hosts, sources_per_host)

    def seam_meshes_on_device(self, hosts, sources_per_host):
        """As seam_meshes, leaving the meshes in device memory; returns [(nv, nt)]."""
        return _seam_meshes_on_device(lib().lvn_batch_seam_meshes, self._h, hosts, sources_per_host)

    def seam_device_mesh(self, i):
        return _seam_device_mesh(lib().lvn_batch_seam_device_mesh, self._h, i)


class Server:
    """The serving form: continuous batching over device-resident chunk slots."""

    STATS = ["slots", "batches", "chunks_meshed", "cache_hits", "evictions", "remeshed_after_edit",
             "cached_chunks", "max_batch_seen"]

    def __init__(self, voxels_per_chunk=64, hbm_budget=64 << 30, max_batch=64, lanes=3, edge_cap=262144,
                 voxel_cap=131072, flags=0):
        h = C.c_void_p()
        check(lib().lvn_server_create(int(voxels_per_chunk), int(hbm_budget), int(max_batch), int(lanes),
                                      int(edge_cap), int(voxel_cap), int(flags), C.byref(h)), "lvn_server_create")
        self._h = h
        self.n = voxels_per_chunk

    def close(self):
        if getattr(self, "_h", None):
            lib().lvn_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_mesh(self, min_xyz, size):
        r = Request()
        r.kind = REQ_MESH
        for i in range(3):
            r.min[i] = int(min_xyz[i])
        r.size = int(size)
        t = C.c_int64()
        check(lib().lvn_server_submit(self._h, C.byref(r), C.byref(t)), "lvn_server_submit")
        return t.value

    def submit_meshes(self, mins, sizes):
        """Mesh requests for many chunks in one call; returns their tickets (int64 array)."""
        mins = np.asarray(mins, dtype=np.int64).reshape(-1, 3)
        n = len(mins)
        reqs = (Request * max(n, 1))()
        for k in range(n):
            reqs[k].kind = REQ_MESH
            for i in range(3):
                reqs[k].min[i] = int(mins[k][i])
            reqs[k].size = int(sizes[k])
        tickets = np.zeros(max(n, 1), dtype=np.int64)
        check(lib().lvn_server_submit_many(self._h, n, C.cast(reqs, C.c_void_p), vptr(tickets)),
              "lvn_server_submit_many")
        return tickets[:n]

    def submit_edit(self, op, aabb):
        r = Request()
        r.kind = REQ_EDIT
        o = np.ascontiguousarray(op, dtype=CSG_OP_DTYPE).reshape(1)
        C.memmove(C.byref(r.op), o.ctypes.data, 48)
        for i in range(3):
            r.aabb_min[i] = int(aabb[i])
            r.aabb_max[i] = int(aabb[3 + i])
        t = C.c_int64()
        check(lib().lvn_server_submit(self._h, C.byref(r), C.byref(t)), "lvn_server_submit")
        return t.value

    def poll(self, cap=65536):
        out = np.zeros(cap, dtype=RESULT_DTYPE)
        n = C.c_int32()
        check(lib().lvn_server_poll(self._h, vptr(out), cap, C.byref(n)), "lvn_server_poll")
        return out[: n.value].copy()

    def drain(self):
        check(lib().lvn_server_drain(self._h), "lvn_server_drain")

    def set_tuning(self, tuning):
        """Launch shapes for the batches launched from now on (a dict as Batch.tuning())."""
        t = Tuning()
        for k in Batch.TUNING:
            setattr(t, k, int(tuning.get(k, 0)))
        check(lib().lvn_server_set_tuning(self._h, C.byref(t)), "lvn_server_set_tuning")

    def stats(self):
        s = np.zeros(8, dtype=np.int64)
        check(lib().lvn_server_stats(self._h, vptr(s)), "lvn_server_stats")
        return dict(zip(self.STATS, (int(v) for v in s)))

    def get_mesh(self, ticket, v_cap=131072, tri_cap=393216, s_cap=131072):
        v = np.zeros(v_cap, dtype=VERTEX_DTYPE)
        t = np.zeros((tri_cap, 3), dtype=np.int32)
        s = np.zeros(s_cap, dtype=SEAM_DTYPE)
        nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().lvn_server_get_mesh(self._h, int(ticket), vptr(v), v_cap, C.byref(nv), vptr(t), tri_cap,
                                        C.byref(nt), vptr(s), s_cap, C.byref(ns)), "lvn_server_get_mesh")
        return v[:nv.value].copy(), t[:nt.value].copy(), s[:ns.value].copy()
