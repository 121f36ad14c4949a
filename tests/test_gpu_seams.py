"""Seam stitching on the GPU against the oracle: vertices and triangles bit for bit, in order.

Inputs are seam node lists of oracle-meshed chunks (equal-LOD block and a mixed LOD0/LOD1
layout), offered to each host node as GenerateClipmapSeamMesh gathers them
(src/clipmap.cpp:579-597).  The cached-source form reads the lists of chunks that the
context meshed itself.
"""
import contextlib
import itertools
import os

import numpy as np
import pytest

from leven_amd import abi, mesher, workloads
from leven_amd.abi import LVN_ERR_CAPACITY, LVN_ERR_INVALID, LVN_OK, SEAM_DTYPE, VERTEX_DTYPE
from oracle import py_oracle

from seam_scene import Scene, overlapping_offer

pytestmark = pytest.mark.gpu

SEED = 2


def _equal_lod_nodes():
    return [((x * 256, -256 + y * 256, z * 256), 256) for x, y, z in itertools.product(range(2), repeat=3)]


def _mixed_lod_nodes():
    """Two 512 regions split into LOD0 chunks among LOD1 chunks (sizes 256 and 512): LOD0
    hosts with coarser neighbours, LOD1 hosts with finer ones (at -x of the split regions)."""
    nodes = []
    for base in ((0, -512, 0), (0, 0, 0)):
        for x, y, z in itertools.product(range(2), repeat=3):
            nodes.append(((base[0] + 256 * x, base[1] + 256 * y, base[2] + 256 * z), 256))
    for x, y, z in ((512, -512, 0), (512, 0, 0), (0, -512, 512), (0, 0, 512), (512, -512, 512), (512, 0, 512),
                    (-512, -512, 0), (-512, 0, 0)):
        nodes.append(((x, y, z), 512))
    return nodes


@pytest.fixture(scope="module")
def ctx():
    mesher.init(workloads.terrain_desc(), SEED, 0)
    mesher.clear_csg_ops()
    c = mesher.Context(64)
    yield c
    c.close()


@pytest.fixture(scope="module")
def equal_lod():
    return Scene(workloads.terrain_desc(), SEED, _equal_lod_nodes())


@pytest.fixture(scope="module")
def mixed_lod():
    return Scene(workloads.terrain_desc(), SEED, _mixed_lod_nodes())


def _colour(i):
    return (0.25 * (i % 4), 0.5, 1.0 - 0.125 * (i % 8))


PATHS = ("grid", "workgroup")  # the two forms of the seam pass (LVN_SEAM_PATH)


@contextlib.contextmanager
def seam_path(path):
    """Run the seam calls inside with the given pass (the library reads LVN_SEAM_PATH per call)."""
    old = os.environ.get("LVN_SEAM_PATH")
    os.environ["LVN_SEAM_PATH"] = path
    try:
        yield
    finally:
        if old is None:
            del os.environ["LVN_SEAM_PATH"]
        else:
            os.environ["LVN_SEAM_PATH"] = old


def _check_scene(ctx, scene):
    srcs = scene.sources()
    hosts = [(mn, size, _colour(i)) for i, (mn, size, _) in enumerate(scene.nodes)]
    want = [py_oracle.seam_mesh(mn, size, s, colour=colour) for (mn, size, colour), s in zip(hosts, srcs)]
    tris = sum(len(t) for _, _, t in want)
    assert tris > 0
    for path in PATHS:
        with seam_path(path):
            got = ctx.seam_meshes(hosts, srcs)
        for i, ((mn, size, _), (rc, v, t)) in enumerate(zip(hosts, want)):
            assert rc == LVN_OK
            gv, gt = got[i]
            print(f"{path}: seam {i} host {mn} size {size}: {len(v)} vertices, {len(t)} triangles", flush=True)
            assert len(gv) == len(v) and len(gt) == len(t), f"{path}: seam {i}"
            assert gv.tobytes() == v.tobytes(), f"{path}: seam {i} vertices"
            assert (gt == t).all(), f"{path}: seam {i} triangles"
    return got


def test_equal_lod_seams_match_oracle(ctx, equal_lod):
    got = _check_scene(ctx, equal_lod)
    # one seam per call gives the same as the batch
    mn, size, _ = equal_lod.nodes[0]
    v, t = ctx.seam_mesh(mn, size, equal_lod.sources()[0], colour=_colour(0))
    assert v.tobytes() == got[0][0].tobytes() and (t == got[0][1]).all()


def test_mixed_lod_seams_match_oracle(ctx, mixed_lod):
    sizes = {s for _, s, _ in mixed_lod.nodes}
    assert sizes == {256, 512}
    _check_scene(ctx, mixed_lod)


def test_mixed_lod_plane_seams_match_oracle(ctx):
    """The coarse-fine sheet whose triangle count the CPU tests fix in closed form."""
    nodes = [((0, -512, 0), 512), ((0, 0, 0), 512)]
    nodes += [((512 + 256 * x, -512 + 256 * y, 256 * z), 256)
              for x, y, z in itertools.product(range(2), range(4), range(2))]
    _check_scene(ctx, Scene(abi.field_desc(height_octaves=0), SEED, nodes))


def test_cached_sources(ctx, equal_lod):
    """Seam lists of chunks the context meshed itself, read where they lie in device memory,
    give the seams the same lists give when passed from the host, and the oracle's."""
    nodes = []
    for mn, size, seams in equal_lod.nodes:
        v, t, s = ctx.generate_chunk_mesh(mn, size)
        assert s.tobytes() == seams.tobytes(), f"chunk {mn} seam nodes"
        nodes.append((mn, size, s))
    host_srcs = workloads.seam_sources(nodes)
    cached = [[dict(d, nodes=None) for d in srcs] for srcs in host_srcs]
    hosts = [(mn, size, _colour(i)) for i, (mn, size, _) in enumerate(nodes)]
    a = ctx.seam_meshes(hosts, host_srcs)
    b = ctx.seam_meshes(hosts, cached)
    for i in range(len(hosts)):
        assert a[i][0].tobytes() == b[i][0].tobytes() and (a[i][1] == b[i][1]).all(), f"seam {i}"
        rc, v, t = py_oracle.seam_mesh(hosts[i][0], hosts[i][1], host_srcs[i], colour=hosts[i][2])
        assert v.tobytes() == b[i][0].tobytes() and (t == b[i][1]).all(), f"seam {i} vs oracle"
    # device-resident results: nothing copied out, the meshes read back from device memory
    counts = ctx.seam_meshes_on_device(hosts, cached)
    for i in range(len(hosts)):
        pv, pt, nv, nt = ctx.seam_device_mesh(i)
        assert (nv, nt) == counts[i] == (len(b[i][0]), len(b[i][1])), f"seam {i} counts"
        assert _d2h(pv, 48 * nv) == b[i][0].tobytes(), f"seam {i} device vertices"
        assert _d2h(pt, 12 * nt) == b[i][1].tobytes(), f"seam {i} device triangles"
    for mn, size, _ in nodes:
        ctx.free_chunk(mn, size)


def _d2h(ptr, nbytes):
    """Copy device memory to the host with the HIP runtime (hipMemcpyDeviceToHost)."""
    import ctypes as C
    out = np.zeros(max(1, nbytes), dtype=np.uint8)
    if nbytes:
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        assert hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0
    return out[:nbytes].tobytes()


def test_batch_sources(ctx, equal_lod):
    """Seams stitched straight from a batch's slots equal the oracle's, on the host and in
    device memory."""
    mins = np.array([mn for mn, _, _ in equal_lod.nodes], dtype=np.int32)
    sizes = np.array([size for _, size, _ in equal_lod.nodes], dtype=np.int32)
    b = mesher.Batch(64, len(sizes))
    b.set_chunks(mins, sizes)
    b.run()
    b.wait()
    srcs = equal_lod.sources()
    slot_srcs = [[dict(d, nodes=None) for d in s] for s in srcs]
    hosts = [(mn, size, _colour(i)) for i, (mn, size, _) in enumerate(equal_lod.nodes)]
    got = b.seam_meshes(hosts, slot_srcs)
    counts = b.seam_meshes_on_device(hosts, slot_srcs)
    for i, ((mn, size, colour), s) in enumerate(zip(hosts, srcs)):
        rc, v, t = py_oracle.seam_mesh(mn, size, s, colour=colour)
        assert got[i][0].tobytes() == v.tobytes() and (got[i][1] == t).all(), f"seam {i}"
        pv, pt, nv, nt = b.seam_device_mesh(i)
        assert (nv, nt) == counts[i] == (len(v), len(t))
        assert _d2h(pv, 48 * nv) == v.tobytes() and _d2h(pt, 12 * nt) == t.tobytes(), f"seam {i} on device"
    # a source that is not a staged chunk
    with pytest.raises(mesher.LvnError) as e:
        b.seam_meshes([hosts[0]], [[dict(min=(4096, 0, 0), size=256, neighbour=4, nodes=None)]])
    assert e.value.code == -5
    b.close()


def test_clipmap_region_with_brushes():
    """A region of the 6-LOD clipmap selection (mixed LODs, caves, 256 brushes) meshed in one
    batch; its seams stitched from the slots equal the oracle's on the same lists."""
    wl = workloads.clipmap()
    mesher.init(wl["desc"], wl["seed"], wl["default_material"])
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    cam = wl["camera"]
    d = np.abs(wl["mins"] + wl["sizes"][:, None] // 2 - cam).max(axis=1)
    keep = np.argsort(d)[:160]  # the chunks nearest the camera: LOD 0-2
    mins, sizes = wl["mins"][keep], wl["sizes"][keep]
    assert len(set(sizes.tolist())) >= 3
    b = mesher.Batch(64, len(sizes))
    b.set_chunks(mins, sizes)
    b.run()
    b.wait()
    counts = b.counts()
    assert (counts[:, 4] == 0).all()
    nodes = [(tuple(int(v) for v in mins[i]), int(sizes[i]), b.get_mesh(i)[2]) for i in range(len(sizes))
             if counts[i, 2] > 0 or counts[i, 3] > 0]
    srcs = workloads.seam_sources(nodes)
    hosts = [(mn, size, _colour(i)) for i, (mn, size, _) in enumerate(nodes)]
    want = [py_oracle.seam_mesh(mn, size, s, colour=colour) for (mn, size, colour), s in zip(hosts, srcs)]
    for path in PATHS:
        with seam_path(path):
            got = b.seam_meshes(hosts, [[dict(x, nodes=None) for x in s] for s in srcs])
        tris = 0
        for i, ((mn, size, _), (rc, v, t)) in enumerate(zip(hosts, want)):
            assert rc == LVN_OK
            assert got[i][0].tobytes() == v.tobytes() and (got[i][1] == t).all(), \
                f"{path}: seam {i} host {mn} size {size}"
            tris += len(t)
        print(f"{path}: {len(nodes)} seams, {tris} triangles", flush=True)
        assert tris > 1000
    b.close()
    mesher.clear_csg_ops()


def test_clipmap_seams_all_match_oracle():
    """Every seam of the 6-LOD clipmap (1,798 seams; the largest offers ~1M nodes and a few
    outgrow their first scratch caps, so they are run again with larger ones), read from the
    clipmap batch's slots, in both forms of the pass: status, counts and the checksum of each
    mesh's bytes equal the oracle's on the same lists."""
    from leven_amd.mesher import _seam_request
    wl = workloads.clipmap()
    mesher.init(wl["desc"], wl["seed"], wl["default_material"])
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    n = len(wl["sizes"])
    b = mesher.Batch(64, n, edge_cap=0, voxel_cap=0, flags=wl["flags"])
    b.set_chunks(wl["mins"], wl["sizes"])
    b.run()
    b.wait()
    counts = b.counts()
    assert (counts[:, 4] == 0).all()
    nodes = [(tuple(int(v) for v in wl["mins"][i]), int(wl["sizes"][i]), b.get_seam_nodes(i, counts))
             for i in range(n) if counts[i, 2] > 0 or counts[i, 3] > 0]
    srcs = workloads.seam_sources(nodes)
    hosts = [(mn, size, _colour(i)) for i, (mn, size, _) in enumerate(nodes)]
    reqs, src, keep = _seam_request(hosts, srcs)
    _, totals, want = py_oracle.seam_meshes(reqs, src, threads=16, per_seam=True)
    assert totals[2] == 0 and (want[:, 0] == 0).all()
    assert len(hosts) > 1500 and totals[1] > 10_000_000  # the whole clipmap: 1,798 seams, ~33M triangles
    print(f"{len(hosts)} seams, {totals[0]} vertices, {totals[1]} triangles", flush=True)
    slot_srcs = [[dict(d, nodes=None) for d in s] for s in srcs]
    for path in PATHS:
        with seam_path(path):
            got = b.seam_meshes_on_device(hosts, slot_srcs)
        for i in range(len(hosts)):
            assert got[i] == (want[i, 1], want[i, 2]), f"{path}: seam {i} counts"
            pv, pt, nv, nt = b.seam_device_mesh(i)
            v = np.frombuffer(_d2h(pv, 48 * nv), dtype=VERTEX_DTYPE)
            t = np.frombuffer(_d2h(pt, 12 * nt), dtype=np.int32).reshape(-1, 3)
            assert py_oracle.mesh_checksum(v, t) == int(want[i, 3]) & (2 ** 64 - 1), f"{path}: seam {i} bytes"
        print(f"{path}: all seams equal the oracle", flush=True)
    b.close()
    mesher.clear_csg_ops()


def test_seam_errors(ctx, equal_lod):
    mn, size, _ = equal_lod.nodes[0]
    srcs = equal_lod.sources()[0]
    # empty
    v, t = ctx.seam_mesh(mn, size, [])
    assert len(v) == 0 and len(t) == 0
    bad = srcs + [overlapping_offer(mn, size, srcs)]
    rc, _, _ = py_oracle.seam_mesh(mn, size, bad)
    for path in PATHS:
        with seam_path(path):
            # a list offered twice: each node once
            v0, t0 = ctx.seam_mesh(mn, size, srcs)
            v1, t1 = ctx.seam_mesh(mn, size, srcs + srcs)
            assert v0.tobytes() == v1.tobytes() and (t0 == t1).all(), path
            # overlapping offers (a coarser node over one of the host's leaves): rejected by both
            with pytest.raises(mesher.LvnError) as e:
                ctx.seam_mesh(mn, size, bad)
            assert e.value.code == rc == LVN_ERR_INVALID, path
    # caller buffers too small: LVN_ERR_CAPACITY with the needed counts
    with pytest.raises(mesher.LvnError) as e:
        _seam_capacity(ctx, mn, size, srcs, len(v0), len(t0))
    assert e.value.code == LVN_ERR_CAPACITY


def _seam_capacity(ctx, mn, size, srcs, nv, nt):
    import ctypes as C
    from leven_amd.abi import seam_sources
    src = seam_sources(srcs)
    v = np.zeros(max(1, nv - 1), dtype=VERTEX_DTYPE)
    t = np.zeros((max(1, nt), 3), dtype=np.int32)
    gnv, gnt = C.c_int32(), C.c_int32()
    col = np.ones(3, dtype=np.float32)
    rc = mesher.lib().lvn_generate_seam_mesh(ctx._h, mesher.vptr(np.asarray(mn, dtype=np.int32)), int(size),
                                             C.cast(src, C.c_void_p), len(srcs), mesher.vptr(col), mesher.vptr(v),
                                             len(v), C.byref(gnv), mesher.vptr(t), len(t), C.byref(gnt))
    assert gnv.value == nv and gnt.value == nt
    mesher.check(rc, "lvn_generate_seam_mesh")
