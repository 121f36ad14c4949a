"""The C++ drop-in (include/lvn_compute.hpp) used the way leven's clipmap uses
src/compute.h, in a separate C++ program (tests/cpp/dropin_check.cpp, built by
__graft_entry__.build()), gives the same meshes as the same calls through the C ABI
(needs a GPU)."""
import os
import subprocess

import numpy as np
import pytest

from leven_amd import mesher, workloads
from leven_amd.abi import CSG_OP_DTYPE, MATERIAL_AIR

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_check")


def _fnv(*arrays):
    h = 1469598103934665603
    for a in arrays:
        for b in np.ascontiguousarray(a).view(np.uint8).tobytes():
            h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _line(tag, mesh):
    v, t, s = mesh
    return f"{tag} {len(v)} {len(t)} {len(s)} {_fnv(v, t, s):016x}"


def test_cpp_dropin_matches_c_abi():
    assert os.path.exists(BIN), "tests/cpp/dropin_check is not built (run __graft_entry__.build())"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = out.stdout.strip().splitlines()

    mesher.init(workloads.heightfield_desc(), 1, 0)
    mesher.clear_csg_ops()
    ctx = mesher.Context(64)
    mn, size = (0, -128, 0), 256
    want = ["voxels 64", _line("fresh", ctx.generate_chunk_mesh(mn, size))]
    want.append(f"isChunkEmpty {1 if ctx.has_surface(mn, size) else 0}")
    add = np.zeros(1, CSG_OP_DTYPE)
    add["type"], add["brush"], add["material"] = 0, 1, 2
    add["origin"] = [32.5, -4.5, 32.5, 0]
    add["dims"] = [7, 7, 7, 0]
    mesher.store_csg_op(add[0], (96, -52, 96), (164, 16, 164))
    ctx.free_chunk(mn, size)
    want.append(_line("stored", ctx.generate_chunk_mesh(mn, size)))
    box = np.zeros(1, CSG_OP_DTYPE)
    box["type"], box["brush"], box["material"] = 1, 0, MATERIAL_AIR
    box["origin"] = [20.5, -2.5, 44.5, 0]
    box["dims"] = [5, 6, 4, 0]
    ctx.apply_csg(box, mn, size)
    ctx.free_chunk(mn, size)
    applied = ctx.generate_chunk_mesh(mn, size)
    want.append(_line("applied", applied))
    nb_min = (mn[0] + size, mn[1], mn[2])
    neighbour = ctx.generate_chunk_mesh(nb_min, size)
    want.append(_line("neighbour", neighbour))
    v, t = ctx.seam_mesh(mn, size, [dict(min=mn, size=size, neighbour=0, nodes=applied[2]),
                                    dict(min=nb_min, size=size, neighbour=4, nodes=neighbour[2])],
                         colour=(1.0, 0.5, 0.25))
    assert len(t) > 0
    want.append(_line("seam", (v, t, np.zeros(0, np.uint8))))
    ctx.close()
    mesher.clear_csg_ops()
    assert got == want
    # the edits changed the mesh each time
    assert len({w.split()[-1] for w in want[1:5] if not w.startswith("isChunk")}) == 3
