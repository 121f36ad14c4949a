"""Seam stitching: the oracle restatement (CPU) and the HIP path (GPU).

The oracle (oracle/lvn_oracle_seams.cpp) follows GenerateClipmapSeamMesh
(src/clipmap.cpp:573-611) and the octree contouring of src/octree.cpp.  It is pinned
on the CPU by what dual contouring fixes: for chunks of one LOD, the chunk meshes and
the seam meshes of every chunk together hold exactly the quads of one dual-contouring
grid over the whole block (one per sign-changing lattice edge whose four voxels lie in
the block), each wound by which end of its edge is solid.  No mesh holds a quad twice
and no seam repeats a chunk quad.  Seams may repeat each other's quads: a host's seam
octree holds the whole facing layers of, say, its +y and +xy neighbours
(FilterSeamNode, src/clipmap.cpp:508-542), so quads inside those layers that straddle
the two neighbours are also in the +y neighbour's own seam; the reference draws both.
The HIP path must then equal the oracle bit for bit (vertices and triangles, in order).
"""
import itertools

import numpy as np
import pytest

from leven_amd import workloads
from leven_amd.abi import LVN_ERR_INVALID, LVN_OK, SEAM_DTYPE
from oracle import py_oracle

from seam_scene import Scene, overlapping_offer


def _block_scene(lod_size=256, nx=2, ny=2, nz=2, y0=-256):
    nodes = [((x * lod_size, y0 + y * lod_size, z * lod_size), lod_size)
             for x, y, z in itertools.product(range(nx), range(ny), range(nz))]
    return Scene(workloads.terrain_desc(), 2, nodes)


@pytest.fixture(scope="module")
def block():
    return _block_scene()


def test_seam_oracle_closes_equal_lod_block(block):
    """Chunk quads + seam quads == the quads of one grid over the 128^3-voxel block."""
    chunk = block.chunk_quads()
    seam, repeats = {}, 0
    for (mn, size, _), srcs in zip(block.nodes, block.sources()):
        rc, v, t = py_oracle.seam_mesh(mn, size, srcs, colour=(0.5, 0.25, 1.0))
        assert rc == LVN_OK
        assert (v["pos"][:, 3] == 0).all() and (v["colour"][:, :3] == [0.5, 0.25, 1.0]).all()
        for k, sign in block.seam_quads(mn, size, srcs, t).items():
            if k in seam:
                assert seam[k] == sign, "a repeated seam quad with the other winding"
                repeats += 1
            seam[k] = sign
    assert not (set(chunk) & set(seam)), "a quad in a chunk mesh and a seam mesh"
    assert len(seam) > 1000 and repeats < len(seam) // 20
    lo, hi = (0, -64, 0), (128, 64, 128)
    S = block.solid(lo, hi)
    expect = {}
    for axis in range(3):
        b, c = [a for a in range(3) if a != axis]
        for p in itertools.product(*[range(0, 128) if a == axis else range(1, 128) for a in range(3)]):
            q = list(p)
            q[axis] += 1
            s0 = S[p]
            if s0 != S[tuple(q)]:
                world = tuple(int(p[a] + lo[a]) for a in range(3))
                expect[(axis, world)] = int(s0)
    got = dict(chunk)
    got.update(seam)
    assert set(got) == set(expect)
    # one winding convention: the sign along the axis is fixed by which end is solid
    conv = {(sign, expect[k]) for k, sign in got.items()}
    assert len(conv) == 2 and len({s for s, _ in conv}) == 2


def test_seam_oracle_edge_cases(block):
    mn, size, _ = block.nodes[0]
    srcs = block.sources()[0]
    rc, v, t = py_oracle.seam_mesh(mn, size, [])
    assert rc == LVN_OK and len(v) == 0 and len(t) == 0
    # a list offered twice selects each node once (the child slot keeps one)
    rc0, v0, t0 = py_oracle.seam_mesh(mn, size, srcs)
    rc1, v1, t1 = py_oracle.seam_mesh(mn, size, srcs + srcs)
    assert rc0 == rc1 == LVN_OK
    assert v0.tobytes() == v1.tobytes() and (t0 == t1).all()
    # overlapping nodes (a coarser node over one of the host's leaves) are rejected
    rc, _, _ = py_oracle.seam_mesh(mn, size, srcs + [overlapping_offer(mn, size, srcs)])
    assert rc == LVN_ERR_INVALID


def test_seam_sources_follow_find_active_nodes():
    """Neighbour lists of a mixed-LOD layout: a coarse node is offered for every neighbour
    region it covers, finer nodes for the region they lie in."""
    s = np.zeros(1, dtype=SEAM_DTYPE)
    nodes = [((256, 0, 0), 256, s), ((512, 0, 0), 512, s), ((256, 256, 0), 256, s), ((256, 0, 256), 256, s),
             ((256, 256, 256), 256, np.zeros(0, dtype=SEAM_DTYPE))]
    srcs = workloads.seam_sources(nodes)
    got = sorted((d["neighbour"], d["min"], d["size"]) for d in srcs[0])
    assert got == [(0, (256, 0, 0), 256), (1, (256, 0, 256), 256), (2, (256, 256, 0), 256),
                   (4, (512, 0, 0), 512), (5, (512, 0, 0), 512), (6, (512, 0, 0), 512), (7, (512, 0, 0), 512)]
    # the coarse node's own seam takes the finer nodes inside its +x... regions: none here
    assert sorted(d["neighbour"] for d in srcs[1]) == [0]
