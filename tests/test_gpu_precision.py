"""Precision form (LVN_CTX_PRECISE_QEF): vertex error against the FP64
oracle, which solves each QEF in FP64 from the same hermite data (needs a GPU).

Target: <= 1e-4 voxel.  The plain FP32 form follows the reference's arithmetic
(ATb - ATA m, 10 FP32 Jacobi sweeps); its error against FP64 is measured and
printed, not bounded.  Vertex normals (the mean of the edges' central-difference
normals, h = 0.001) are compared with the oracle's FP64 normals at the same
crossings and printed: FP32 central differences at coordinates of O(100) carry
~1e-3 error, more where max() switches branch between the two samples.

End to end, both forms are also compared with the vertex computation done wholly in
FP64 on the same topology (crossings by the 17-sample search and normals of the FP64
field, QEF in FP64): there the reference's FP32 hermite data set the error (normals
~1e-3, an occasional crossing one sample away), which the precision form, whose
hermite data are the reference's, does not change.  Stated, loosely bounded.
"""
import numpy as np
import pytest

from leven_amd import mesher, workloads
from leven_amd.abi import CTX_PRECISE_QEF

pytestmark = pytest.mark.gpu

TARGET = 1e-4  # voxel


def _pick(name, wl):
    """Chunks compared per configuration: the heightfield chunk; 8 chunks of the 8x8x8 block
    (index = x + 8 * y + 64 * z; y = 3, 4 hold the surface and every brush, y = 1 is deep in
    the caves); two chunks of every clipmap LOD (at a third and two thirds of its list)."""
    if name == "plain_cpu":
        return [0]
    if name == "clipmap":
        sizes = np.asarray(wl["sizes"])
        out = []
        for s in np.unique(sizes):
            idx = np.flatnonzero(sizes == s)
            out += [int(idx[len(idx) // 3]), int(idx[2 * len(idx) // 3])]
        return out
    return [25, 100, 155, 230, 285, 357, 9, 333]


def _errors(oracle, wl, pick):
    """Both forms (plain FP32, precise) on the same chunks, each oracle chunk built once:
    per form the vertex error against the FP64 solve of the same hermite data and against
    the FP64 pipeline (per vertex, in voxels, with the chunk's size), the normal error, and
    the FP64 re-solves."""
    mesher.init(wl["desc"], wl["seed"], 0)
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    forms = {}
    for label, flags in (("plain", 0), ("precise", CTX_PRECISE_QEF)):
        b = mesher.Batch(64, len(pick), edge_cap=0, voxel_cap=0, flags=flags)
        b.set_chunks(wl["mins"][pick], wl["sizes"][pick])
        b.run()
        b.wait()
        forms[label] = (b, b.counts())
    perm = oracle.perm_table(wl["seed"])
    out = {label: dict(solve=[], e2e=[], nerr=[], size=[]) for label in forms}
    moved = edges = 0
    for i, j in enumerate(pick):
        mn, size = wl["mins"][j], int(wl["sizes"][j])
        keep = [k for k, bb in enumerate(wl["aabbs"])
                if not (mn[0] + size < bb[0] or mn[1] + size < bb[1] or mn[2] + size < bb[2]
                        or mn[0] > bb[3] or mn[1] > bb[4] or mn[2] > bb[5])]
        c = oracle.Chunk(64).generate(wl["desc"], perm, mn, size, ops=wl["ops"][keep]).build()
        tris = c.triangles()
        if c.counts()["vertices"] == 0 or len(tris) == 0:
            continue
        p64, n64 = c.positions_f64(), c.normals_f64()
        f64_pos, _, mv = c.fp64_pipeline()
        moved += mv
        edges += int(forms["plain"][1][i, 0])
        voxel = 4.0 * (size // 256)
        for label, (b, _) in forms.items():
            v, t, _ = b.get_mesh(i)
            assert (t == tris).all(), "topology must not depend on the precision form"
            pos = v["pos"][:, :3].astype(np.float64)
            o = out[label]
            o["solve"].append(np.abs(pos - p64[:, :3]).max(axis=1) / voxel)
            o["e2e"].append(np.abs(pos - f64_pos[:, :3]).max(axis=1) / voxel)
            o["nerr"].append(np.linalg.norm(v["normal"][:, :3].astype(np.float64) - n64[:, :3], axis=1))
            o["size"].append(np.full(len(v), size))
    for label, (b, counts) in forms.items():
        out[label] = {k: np.concatenate(v) for k, v in out[label].items()}
        out[label]["resolved"] = int(counts[:, 5].sum())
        b.close()
    mesher.clear_csg_ops()
    return out, moved, edges


@pytest.mark.parametrize("name", ["plain_cpu", "terrain_512", "precision", "clipmap"])
def test_precise_qef_vertex_error_vs_fp64(oracle, name):
    """Per configuration: the precision form's vertex error against the FP64 solve of the
    same hermite data is within TARGET; the end-to-end error against the FP64 pipeline is
    stated per chunk size (bounded on the LOD0 configurations; on the clipmap's coarse
    chunks the FP32 central-difference normals at world coordinates of 10^4 and more, the
    reference's method, dominate it), and the crossings that the FP64 field places one
    sample away stay under 1% of the edges."""
    wl = workloads.WORKLOADS[name]()
    res, moved, edges = _errors(oracle, wl, _pick(name, wl))
    plain, prec = res["plain"], res["precise"]
    verts = len(prec["solve"])
    print(f"\n{name}: {verts} vertices; against the FP64 solve of the same hermite data: FP32 plain max "
          f"{plain['solve'].max():.3e}, precise max {prec['solve'].max():.3e} p99.9 "
          f"{np.percentile(prec['solve'], 99.9):.3e} voxel; FP64 re-solves {prec['resolved']} "
          f"({100.0 * prec['resolved'] / max(verts, 1):.3f}%); normal |n - n64| p50 "
          f"{np.percentile(prec['nerr'], 50):.2e} p99 {np.percentile(prec['nerr'], 99):.2e}; crossings one "
          f"sample away {moved} of {edges} edges")
    for size in np.unique(prec["size"]):
        sel = prec["size"] == size
        e = prec["e2e"][sel]
        print(f"  end to end vs the FP64 pipeline, chunk size {size} ({sel.sum()} vertices): p50 "
              f"{np.percentile(e, 50):.2e} p99 {np.percentile(e, 99):.2e} max {e.max():.2e} voxel; "
              f"normal p50 {np.percentile(prec['nerr'][sel], 50):.2e}")
    assert verts > 0
    assert prec["solve"].max() <= TARGET
    assert moved < 0.01 * edges
    if name != "clipmap":
        assert np.percentile(prec["e2e"], 50) < 1e-3
        assert np.percentile(prec["nerr"], 50) < 1e-2  # FP32 central differences; stated, loosely bounded
