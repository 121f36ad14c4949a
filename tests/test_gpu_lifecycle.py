"""Library lifetime, reseeding order and argument checks through the C ABI (needs a GPU).

* lvn_shutdown with a batch and a server still alive: every call that would launch work
  returns LVN_ERR_STATE, the tables stay allocated until the last object is destroyed,
  and lvn_init afterwards gives results equal to the oracle again;
* lvn_set_noise_seed right after an asynchronous lvn_batch_run: the batch in flight
  finishes with the old table (the reference's single in-order queue orders
  Compute_SetNoiseSeed after work already enqueued, src/compute.cpp:168-191);
* chunk sizes that are not a positive multiple of N * 4 are rejected by every
  single-chunk entry point.
"""
import ctypes as C

import numpy as np
import pytest

from leven_amd import abi, mesher, workloads
from leven_amd.abi import vptr
from test_gpu_parity import compare_stages, oracle_chunk

pytestmark = pytest.mark.gpu


def _mesh_all(b):
    return [b.get_mesh(i) for i in range(b.n)]


def _same_meshes(a, b):
    for (v0, t0, s0), (v1, t1, s1) in zip(a, b):
        assert (v0.view(np.uint32) == v1.view(np.uint32)).all() and (t0 == t1).all()
        assert (s0.view(np.uint32) == s1.view(np.uint32)).all()


def test_shutdown_with_live_objects_refuses_new_work(oracle):
    w = workloads.terrain_512()
    print("shutdown with a live batch and server", flush=True)
    mesher.init(w["desc"], w["seed"], 0)
    mesher.clear_csg_ops()
    b = mesher.Batch(16, 4)
    mins = [(0, -64, 0), (64, -64, 0), (0, -64, 64), (64, 0, 64)]
    b.set_chunks(mins, [64] * 4)
    b.run()
    b.wait()
    before = b.counts().copy()
    srv = mesher.Server(16, hbm_budget=64 << 20, max_batch=4, lanes=1, edge_cap=0, voxel_cap=0)
    mesher.shutdown()
    L = mesher.lib()
    state = abi.LVN_ERR_STATE
    assert L.lvn_batch_run(b._h) == state
    ms = np.zeros(mesher.NUM_STAGES, np.float32)
    assert L.lvn_batch_run_timed(b._h, vptr(ms), mesher.NUM_STAGES) == state
    assert L.lvn_batch_autotune(b._h, 1, None) == state
    work = np.zeros(8, np.uint64)
    assert L.lvn_batch_work(b._h, vptr(work)) == state
    assert (b.counts() == before).all()  # results of the last run stay readable
    srv.submit_mesh((0, -64, 0), 64)     # queued; the scheduler refuses to launch it
    with pytest.raises(mesher.LvnError) as e:
        srv.drain()
    assert e.value.code == state
    srv.close()
    b.close()
    h = C.c_void_p()
    assert L.lvn_ctx_create(16, 0, C.byref(h)) == state
    mesher.init(w["desc"], w["seed"], 0)
    ctx = mesher.Context(16)
    g = ctx.debug_stages((0, -64, 0), 64)
    compare_stages(g, oracle_chunk(oracle, w["desc"], w["seed"], 16, (0, -64, 0), 64), 16, 64)
    ctx.close()


def test_reseed_after_async_run_keeps_the_old_table_for_work_in_flight(oracle):
    w = workloads.terrain_512()
    print("reseed after an asynchronous batch run", flush=True)
    mesher.init(w["desc"], 2, 0)
    mesher.clear_csg_ops()
    mins = [tuple(m) for m in w["mins"][:64]]
    b = mesher.Batch(64, 64, 262144, 131072)
    b.set_chunks(mins, [256] * 64)
    b.run()
    b.wait()
    seed2 = _mesh_all(b)
    b.run()                     # asynchronous
    mesher.set_noise_seed(7)    # must not change the table under the batch in flight
    b.wait()
    _same_meshes(_mesh_all(b), seed2)
    b.run()
    b.wait()
    c = oracle_chunk(oracle, w["desc"], 7, 64, mins[20], 256)
    v, t, s = b.get_mesh(20)
    assert (t == c.triangles()).all() and (v.view(np.uint32) == c.vertices().view(np.uint32)).all()
    b.close()


def test_single_chunk_entry_points_reject_bad_sizes():
    print("chunk size validation", flush=True)
    mesher.init(workloads.heightfield_desc(), 1, 0)
    mesher.clear_csg_ops()
    ctx = mesher.Context(16)
    L = mesher.lib()
    mn = np.zeros(3, np.int32)
    v = np.zeros(16, abi.VERTEX_DTYPE)
    t = np.zeros((16, 3), np.int32)
    s = np.zeros(16, abi.SEAM_DTYPE)
    a, b_, c_ = C.c_int32(), C.c_int32(), C.c_int32()
    out = C.c_int32()
    for size in (0, -64, 4, 60, 100, 64 + 32):
        assert L.lvn_generate_chunk_mesh(ctx._h, vptr(mn), size, vptr(v), 16, C.byref(a), vptr(t), 16, C.byref(b_),
                                         vptr(s), 16, C.byref(c_)) == abi.LVN_ERR_INVALID, size
        assert L.lvn_chunk_has_surface(ctx._h, vptr(mn), size, C.byref(out)) == abi.LVN_ERR_INVALID, size
        assert L.lvn_apply_csg(ctx._h, None, 0, vptr(mn), size) == abi.LVN_ERR_INVALID, size
        dump = abi.StageDump()
        assert L.lvn_debug_chunk_stages(ctx._h, vptr(mn), size, None, 0, C.byref(dump)) == abi.LVN_ERR_INVALID, size
    assert L.lvn_chunk_has_surface(ctx._h, vptr(mn), 128, C.byref(out)) == abi.LVN_OK
    ctx.close()
