"""The clipmap request producer feeding the server along a seeded camera path (needs a GPU).

Each update's chunks go to the server as mesh requests; the served results go back to the
producer (lvn_clipmap_commit), whose seam requests are stitched from the served chunks'
seam lists.  Checks: every served mesh equals the batched form's mesh of that chunk (bit
for bit); the requested seams equal the oracle's seam meshes; an edit made through both
the producer and the server re-meshes exactly the selected chunks it touches, and the
re-served meshes equal the batched form with the op stored."""
import numpy as np
import pytest

from leven_amd import abi, mesher, workloads

pytestmark = pytest.mark.gpu

R = workloads.LOD_MAX_NODE_SIZE
CAPS = dict(v_cap=262144, tri_cap=6 * 262144, s_cap=262144)


def _pairs(mins, sizes):
    return [(tuple(int(v) for v in m), int(s)) for m, s in zip(mins, sizes)]


def _serve(server, chunks):
    """Submit mesh requests for `chunks`, wait; returns {chunk: (ticket, status, nt, ns)}."""
    tickets = {server.submit_mesh(m, s): (m, s) for (m, s) in chunks}
    server.drain()
    res = server.poll(len(tickets) + 16)
    out = {}
    for r in res:
        if r["ticket"] in tickets:
            out[tickets[r["ticket"]]] = (int(r["ticket"]), int(r["status"]), int(r["nt"]), int(r["ns"]))
    assert len(out) == len(chunks)
    return out


def _batched_meshes(chunks):
    b = mesher.Batch(64, len(chunks), edge_cap=0, voxel_cap=0)
    b.set_chunks(np.array([m for m, _ in chunks], np.int32), np.array([s for _, s in chunks], np.int32))
    b.run()
    b.wait()
    out = {c: b.get_mesh(i) for i, c in enumerate(chunks)}
    b.close()
    return out


def _same(m1, m2):
    return all(x.tobytes() == y.tobytes() for x, y in zip(m1, m2))


def test_camera_path_through_server(oracle):
    w = workloads.terrain_512()
    mesher.init(w["desc"], w["seed"], 0)
    mesher.clear_csg_ops()
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    server = mesher.Server(64, hbm_budget=64 << 30, max_batch=64, lanes=3, edge_cap=0, voxel_cap=0)
    cm = mesher.Clipmap((-R, -R, -R), (2 * R, R, 2 * R))
    rng = np.random.default_rng(7)
    cam = np.array([700.0, 20.0, 900.0])
    served = {}  # chunk -> its latest (ticket, status, nt, ns)
    last_seams = None
    for step in range(5):
        print(f"producer step {step}: camera {cam.round(1).tolist()}", flush=True)
        u = cm.update(cam)
        chunks = _pairs(*u["to_mesh"])
        assert len(chunks) > 0
        res = _serve(server, chunks)
        served.update(res)
        results = np.array([res[c][1:] for c in chunks], np.int32)
        assert (results[:, 0] == 0).all()
        reqs, srcs, nreq, nsrc = cm.commit(results)
        assert nreq > 0
        last_seams = (reqs, srcs, nreq, nsrc)
        cam = cam + np.array([rng.uniform(300, 700), 0.0, rng.uniform(-400, 400)])
    # the served meshes of the active chunks equal the batched form's
    active = _pairs(*cm.list(abi.CLIPMAP_ACTIVE))
    ref = _batched_meshes(active)
    for c in active:
        assert _same(server.get_mesh(served[c][0], **CAPS), ref[c]), f"chunk {c}"
    # the last update's seams: host lists from the served chunks, against the oracle
    reqs, srcs, nreq, nsrc = last_seams
    hosts, per_host = [], []
    for i in range(nreq):
        q = reqs[i]
        hosts.append((tuple(q.host_min), q.host_size, (1.0, 1.0, 1.0)))
        per_host.append([dict(min=tuple(srcs[j].min), size=srcs[j].size, neighbour=srcs[j].neighbour,
                              nodes=ref[(tuple(srcs[j].min), srcs[j].size)][2])
                         for j in range(q.first_source, q.first_source + q.n_sources)])
    ctx = mesher.Context(64)
    meshes = ctx.seam_meshes(hosts, per_host)
    pick = np.linspace(0, nreq - 1, min(nreq, 12)).astype(int)
    for i in pick:
        st, v, t = oracle.seam_mesh(hosts[i][0], hosts[i][1], per_host[i])
        assert st == 0
        assert meshes[i][0].tobytes() == v.tobytes() and meshes[i][1].tobytes() == t.tobytes(), f"seam {hosts[i][:2]}"
    ctx.close()
    # an edit near the camera through both: the next update re-meshes the selected chunks it
    # touches, and the re-served meshes carry the op
    cam = cam - np.array([500.0, 0.0, 0.0])
    u = cm.update(cam)
    chunks = _pairs(*u["to_mesh"])
    res = _serve(server, chunks)
    cm.commit(np.array([res[c][1:] for c in chunks], np.int32).reshape(-1, 3))
    op = np.zeros(1, abi.CSG_OP_DTYPE)[0]
    op["type"], op["brush"], op["material"] = 0, 1, 2
    c = cam / 4
    op["origin"] = [c[0] + 0.5, 2.5, c[2] + 0.5, 0.0]
    op["dims"] = [9.0, 9.0, 9.0, 0.0]
    bb = workloads.op_aabb(op)
    touched = set(_pairs(*cm.edit(bb[:3], bb[3:])))
    server.submit_edit(op, bb)
    mesher.store_csg_op(op, bb[:3], bb[3:])
    u = cm.update(cam)
    again = _pairs(*u["to_mesh"])
    assert again and set(again) == touched & set(_pairs(*u["selected"]))
    res = _serve(server, again)
    cm.commit(np.array([res[c][1:] for c in again], np.int32))
    ref = _batched_meshes(again)
    for ch in again:
        assert _same(server.get_mesh(res[ch][0], **CAPS), ref[ch]), f"edited chunk {ch}"
    # the same chunks without the op differ: the served meshes carry it
    mesher.clear_csg_ops()
    for op0, bb0 in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op0, bb0[:3], bb0[3:])
    base = _batched_meshes(again)
    assert any(not _same(ref[ch], base[ch]) for ch in again)
    server.close()
    cm.close()
    mesher.clear_csg_ops()
