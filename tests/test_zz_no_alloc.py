"""The allocation-free entry point (lvn_generate_chunk_mesh_noalloc) against
lvn_generate_chunk_mesh on the same chunks, through the C ABI: vertices, triangles and
seam nodes bit-identical on two calls in a row, the second with a smaller chunk
(16^3 after 64^3), with host and with device output buffers; the limits are checked on
the host before any device work (needs a GPU)."""
import ctypes as C

import numpy as np
import pytest

from leven_amd import abi, mesher, workloads
from leven_amd.abi import VERTEX_DTYPE, vptr

pytestmark = pytest.mark.gpu

# (voxels per chunk, min in world units, size): a 64^3 chunk with two brushes and caves,
# then a 16^3 chunk with two brushes (427 active voxels)
CASES = [(64, (256, -256, 256), 256), (16, (64, -64, 128), 64)]


def _store_ops(w):
    mesher.clear_csg_ops()
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])


def _same(a, b):
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _reference_mesh(n, mn, size):
    ctx = mesher.Context(n)
    try:
        return ctx.generate_chunk_mesh(mn, size)
    finally:
        ctx.close()


def test_noalloc_matches_generate_chunk_mesh():
    w = workloads.terrain_512()
    mesher.init(w["desc"], w["seed"], 0)
    _store_ops(w)
    for n, mn, size in CASES:
        print(f"noalloc: N={n} min={mn} size={size}", flush=True)
        v0, t0, s0 = _reference_mesh(n, mn, size)
        v1, t1, s1 = mesher.generate_chunk_mesh_noalloc(n, mn, size)
        assert len(t0) > 0 and len(s0) > 0
        assert _same(v0, v1), "vertices"
        assert _same(t0, t1), "triangles"
        assert _same(s0, s1), "seam nodes"
    mesher.clear_csg_ops()


def _hip():
    """The HIP runtime the library itself links (a caller's own device buffers)."""
    mesher.lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def test_noalloc_device_outputs_and_empty_chunk():
    w = workloads.terrain_512()
    mesher.init(w["desc"], w["seed"], 0)
    _store_ops(w)
    n, mn, size = CASES[1]
    v0, t0, s0 = _reference_mesh(n, mn, size)
    hip = _hip()
    sizes = [v0.nbytes, t0.nbytes, s0.nbytes]
    bufs = [C.c_void_p() for _ in sizes]
    for b, nb in zip(bufs, sizes):
        assert hip.hipMalloc(C.byref(b), nb) == 0
    try:
        nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
        print(f"noalloc device outputs: N={n} min={mn} size={size}", flush=True)
        rc = mesher.lib().lvn_generate_chunk_mesh_noalloc(n, 0, vptr(np.array(mn, np.int32)), size, bufs[0], len(v0),
                                                          C.byref(nv), bufs[1], len(t0), C.byref(nt), bufs[2],
                                                          len(s0), C.byref(ns))
        assert rc == abi.LVN_OK
        assert (nv.value, nt.value, ns.value) == (len(v0), len(t0), len(s0))
        back = [np.zeros(nb, np.uint8) for nb in sizes]
        for h, d, nb in zip(back, bufs, sizes):
            assert hip.hipMemcpy(h.ctypes.data, d, nb, 2) == 0  # hipMemcpyDeviceToHost
        assert back[0].tobytes() == v0.tobytes()
        assert back[1].tobytes() == t0.tobytes()
        assert back[2].tobytes() == s0.tobytes()
    finally:
        for b in bufs:
            hip.hipFree(b)
    # all air: no edges, nothing exported
    v, t, s = mesher.generate_chunk_mesh_noalloc(16, (0, 4096, 0), 64)
    assert len(v) == len(t) == len(s) == 0
    mesher.clear_csg_ops()


def test_noalloc_limits_are_checked_before_device_work():
    w = workloads.terrain_512()
    mesher.init(w["desc"], w["seed"], 0)
    mesher.clear_csg_ops()
    L = mesher.lib()
    nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
    mn = np.array([0, 0, 0], np.int32)
    v = np.zeros(1, VERTEX_DTYPE)
    for n, size in [(128, 512), (48, 192), (16, 32), (16, 96)]:  # unsupported N, sizes off the N * 4 grid
        rc = L.lvn_generate_chunk_mesh_noalloc(n, 0, vptr(mn), size, vptr(v), 1, C.byref(nv), None, 0, C.byref(nt),
                                               None, 0, C.byref(ns))
        assert rc == abi.LVN_ERR_INVALID, (n, size)
    # more overlapping stored ops than the area holds
    op = w["ops"][:1]
    for _ in range(abi.NOALLOC_MAX_OPS + 1):
        mesher.store_csg_op(op[0], [-10, -10, -10], [10, 10, 10])
    rc = L.lvn_generate_chunk_mesh_noalloc(16, 0, vptr(mn), 64, vptr(v), 1, C.byref(nv), None, 0, C.byref(nt), None,
                                           0, C.byref(ns))
    assert rc == abi.LVN_ERR_OVERFLOW
    mesher.clear_csg_ops()
