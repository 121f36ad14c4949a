"""Parity of the HIP pipeline with the CPU oracle, through the C ABI (needs a gfx950 GPU).

Bar: bit-exact density, materials, edge list, edge data, active-voxel set, codes,
masks, materials and triangle index buffers; vertex positions within 1e-4 voxel
(they are bit-exact too when both sides follow the same FP32 operation order).
"""
import numpy as np
import pytest

from leven_amd import abi, mesher, workloads
from leven_amd.abi import CTX_FP16_DENSITY, MATERIAL_AIR

pytestmark = pytest.mark.gpu

VOXEL_TOL = 1e-4  # voxel units


def setup_field(desc, seed, default_material=0):
    mesher.init(desc, seed, default_material)
    mesher.clear_csg_ops()


def oracle_chunk(oracle, desc, seed, n, mn, size, ops=None, default_material=0):
    perm = oracle.perm_table(seed)
    return oracle.Chunk(n).generate(desc, perm, mn, size, default_material=default_material, ops=ops).build()


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and (a.view(np.uint32) == b.view(np.uint32)).all()


def compare_stages(g, c, n, size):
    assert same_bits(g["density"], c.density()), "density differs"
    assert (g["materials"] == c.materials()).all(), "materials differ"
    ids, data = c.edges()
    assert len(g["edge_ids"]) == len(ids) and (g["edge_ids"] == ids).all(), "edge list differs"
    assert same_bits(g["edge_data"], data), "edge crossing/normal data differs"
    codes, masks, mats = c.voxels()
    assert (g["voxel_codes"] == codes).all() and (g["voxel_edge_masks"] == masks).all()
    assert (g["voxel_materials"] == mats).all()
    pos, nrm = c.positions()
    voxel = 4.0 * (size // (n * 4))  # world units per voxel at this LOD
    err = np.abs(g["voxel_positions"][:, :3] - pos[:, :3]).max() / voxel if len(pos) else 0.0
    assert err <= VOXEL_TOL, f"vertex error {err} voxel"
    assert same_bits(g["voxel_positions"], pos), "vertex positions not bit-exact"
    assert same_bits(g["voxel_normals"], nrm)


CASES = [
    # name, desc, seed, n, min, size
    ("heightfield_n16", workloads.heightfield_desc(), 1, 16, (64, -32, 128), 64),
    ("heightfield_lod2_n16", workloads.heightfield_desc(), 1, 16, (0, -128, 0), 256),
    ("caves_n32", workloads.terrain_desc(), 2, 32, (128, -256, 0), 128),
    ("leven_terrain_n16", abi.field_desc(kind=abi.FIELD_LEVEN_TERRAIN), 93923590, 16, (0, 2320, 0), 64),
    ("plain_cpu_n64", workloads.heightfield_desc(), 1, 64, (0, -128, 0), 256),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_stages_bit_exact(oracle, case):
    name, desc, seed, n, mn, size = case
    setup_field(desc, seed)
    assert (mesher.perm_table() == oracle.perm_table(seed)).all()
    ctx = mesher.Context(n)
    g = ctx.debug_stages(mn, size)
    c = oracle_chunk(oracle, desc, seed, n, mn, size)
    assert c.counts()["edges"] > 0, "case must have a surface"
    compare_stages(g, c, n, size)


def test_caves_full_chunk_bit_exact(oracle):
    """A 64^3 underground chunk of the 3D terrain (cave term everywhere)."""
    desc = workloads.terrain_desc()
    setup_field(desc, 2)
    g = mesher.Context(64).debug_stages((256, -512, 0), 256)
    c = oracle_chunk(oracle, desc, 2, 64, (256, -512, 0), 256)
    assert c.counts()["voxels"] > 1000
    compare_stages(g, c, 64, 256)


def test_topology_given_same_density(oracle):
    """Feed the GPU's density field to the oracle: the edge list, active set and
    triangles must be identical."""
    desc = workloads.terrain_desc()
    setup_field(desc, 2)
    n, mn, size = 32, (0, -128, 128), 128
    ctx = mesher.Context(n)
    g = ctx.debug_stages(mn, size)
    c = oracle.Chunk(n).generate_from_density(desc, oracle.perm_table(2), mn, size, g["density"]).build()
    ids, _ = c.edges()
    assert (g["edge_ids"] == ids).all()
    codes, masks, mats = c.voxels()
    assert (g["voxel_codes"] == codes).all() and (g["voxel_materials"] == mats).all()
    v, t, s = ctx.generate_chunk_mesh(mn, size)
    assert (t == c.triangles()).all()


def _ops_for_chunk(w, mn, size):
    keep = [i for i, bb in enumerate(w["aabbs"])
            if not (mn[0] + size < bb[0] or mn[1] + size < bb[1] or mn[2] + size < bb[2]
                    or mn[0] > bb[3] or mn[1] > bb[4] or mn[2] > bb[5])]
    return w["ops"][keep]


def test_csg_fresh_chunk_bit_exact(oracle):
    """terrain_512 chunks with the brushes that overlap them (LoadDensityField's filter)."""
    w = workloads.terrain_512()
    setup_field(w["desc"], w["seed"])
    ctx = mesher.Context(64)
    tested = 0
    for mn in [(256, -256, 256), (768, 0, 512), (0, -256, 1024), (1280, 0, 1536)]:
        ops = _ops_for_chunk(w, mn, 256)
        if len(ops) == 0:
            continue
        g = ctx.debug_stages(mn, 256, ops)
        c = oracle_chunk(oracle, w["desc"], w["seed"], 64, mn, 256, ops=ops)
        compare_stages(g, c, 64, 256)
        tested += 1
    assert tested >= 2


def test_generate_chunk_mesh_matches_oracle(oracle):
    w = workloads.plain_cpu()
    setup_field(w["desc"], w["seed"])
    ctx = mesher.Context(64)
    mn, size = tuple(w["mins"][0]), int(w["sizes"][0])
    v, t, s = ctx.generate_chunk_mesh(mn, size)
    c = oracle_chunk(oracle, w["desc"], w["seed"], 64, mn, size)
    assert len(v) == c.counts()["vertices"] and len(t) == c.counts()["triangles"]
    assert (t == c.triangles()).all()
    assert (v.view(np.uint32) == c.vertices().view(np.uint32)).all()
    assert (s.view(np.uint32) == c.seams().view(np.uint32)).all()
    # cached second call returns the same
    v2, t2, s2 = ctx.generate_chunk_mesh(mn, size)
    assert (t2 == t).all() and (v2.view(np.uint32) == v.view(np.uint32)).all()


def test_capacity_error_reports_counts():
    w = workloads.plain_cpu()
    setup_field(w["desc"], w["seed"])
    ctx = mesher.Context(64)
    with pytest.raises(mesher.LvnError) as e:
        ctx.generate_chunk_mesh(tuple(w["mins"][0]), 256, v_cap=100, tri_cap=100, s_cap=10)
    assert e.value.code == abi.LVN_ERR_CAPACITY


def test_empty_chunks():
    setup_field(workloads.heightfield_desc(), 1)
    ctx = mesher.Context(16)
    for mn in [(0, 4096, 0), (0, -4096, 0)]:  # all air, all solid
        v, t, s = ctx.generate_chunk_mesh(mn, 64)
        assert len(v) == len(t) == len(s) == 0
        assert not ctx.has_surface(mn, 64)


def test_stored_ops_and_cache_semantics(oracle):
    """Stored ops are applied when a chunk is first generated; a cached chunk ignores
    ops stored later until freeChunkOctree; applyCSGOperations edits a cached field."""
    w = workloads.terrain_512()
    desc, seed = w["desc"], w["seed"]
    setup_field(desc, seed)
    perm = oracle.perm_table(seed)
    ctx = mesher.Context(64)
    mn, size = (256, -256, 256), 256
    ops = _ops_for_chunk(w, mn, size)
    assert len(ops) >= 2
    first, second = ops[:1], ops[1:2]
    bb = workloads.op_aabb(first[0])
    mesher.store_csg_op(first[0], bb[:3], bb[3:])
    _, t1, _ = ctx.generate_chunk_mesh(mn, size)
    c1 = oracle.Chunk(64).generate(desc, perm, mn, size, ops=first).build()
    assert (t1 == c1.triangles()).all()
    # a later stored op does not change the cached chunk ...
    bb2 = workloads.op_aabb(second[0])
    mesher.store_csg_op(second[0], bb2[:3], bb2[3:])
    _, t2, _ = ctx.generate_chunk_mesh(mn, size)
    assert (t2 == t1).all()
    # ... until it is freed: the cached field then takes the new op as its own batch
    ctx.free_chunk(mn, size)
    v3, t3, _ = ctx.generate_chunk_mesh(mn, size)
    c3 = oracle.Chunk(64).generate(desc, perm, mn, size, ops=first).apply_csg(second).build()
    assert (t3 == c3.triangles()).all()
    assert (v3.view(np.uint32) == c3.vertices().view(np.uint32)).all()
    # direct edit (Compute_ApplyCSGOperations) on the cached field
    extra = workloads.make_ops(3, seed=77, extent=64, y_range=(-60, -40))[0]
    extra["origin"][:, 0] += 64
    extra["origin"][:, 2] += 64
    ctx.apply_csg(extra, mn, size)
    ctx.free_chunk(mn, size)
    v4, t4, _ = ctx.generate_chunk_mesh(mn, size)
    c4 = oracle.Chunk(64).generate(desc, perm, mn, size, ops=first).apply_csg(second).apply_csg(extra).build()
    assert (t4 == c4.triangles()).all()
    assert (v4.view(np.uint32) == c4.vertices().view(np.uint32)).all()
    mesher.clear_csg_ops()


def test_fp16_density_keeps_topology(oracle):
    """FP16 storage: sign-preserving rounding keeps materials, edges and triangles equal
    to the FP32 form; density within half precision."""
    desc = workloads.terrain_desc()
    setup_field(desc, 2)
    n, mn, size = 32, (0, -128, 128), 128
    g32 = mesher.Context(n).debug_stages(mn, size)
    g16 = mesher.Context(n, flags=CTX_FP16_DENSITY).debug_stages(mn, size)
    assert (g16["materials"] == g32["materials"]).all()
    assert (g16["edge_ids"] == g32["edge_ids"]).all()
    rel = np.abs(g16["density"] - g32["density"]) / np.maximum(np.abs(g32["density"]), 1e-3)
    assert (rel[np.abs(g32["density"]) > 1e-3] <= 2.0 ** -11).all()
    neg = g32["density"] < 0
    assert (g16["density"][neg] < 0).all()
    # crossings are re-evaluated in FP32, so the hermite data stays bit-exact
    assert same_bits(g16["edge_data"], g32["edge_data"])


def test_batched_matches_oracle_and_replays(oracle):
    """The batched form (one launch per stage over all chunks, graph replay) meshes each
    chunk exactly as the oracle does; three replays, with other GPU work between them,
    give identical results."""
    w = workloads.terrain_512()
    setup_field(w["desc"], w["seed"])
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    pick = [0, 9, 27, 77, 100, 130, 200, 259, 300, 333, 450, 511]
    mins, sizes = w["mins"][pick], w["sizes"][pick]
    b = mesher.Batch(64, len(pick), edge_cap=262144, voxel_cap=131072)
    b.set_chunks(mins, sizes)
    first = None
    other = mesher.Context(16)
    for rep in range(3):
        b.run()
        b.wait()
        got = [b.get_mesh(i) for i in range(len(pick))]
        if first is None:
            first = got
            perm = oracle.perm_table(w["seed"])
            nonempty = 0
            for i, (mn, size) in enumerate(zip(mins, sizes)):
                c = oracle.Chunk(64).generate(w["desc"], perm, mn, int(size), ops=_ops_for_chunk(w, mn, int(size)))
                c.build()
                v, t, s = got[i]
                assert (t == c.triangles()).all(), f"chunk {i} triangles"
                assert (v.view(np.uint32) == c.vertices().view(np.uint32)).all(), f"chunk {i} vertices"
                assert (s.view(np.uint32) == c.seams().view(np.uint32)).all(), f"chunk {i} seams"
                nonempty += len(t) > 0
            assert nonempty >= 6
        else:
            for (v0, t0, s0), (v1, t1, s1) in zip(first, got):
                assert (t0 == t1).all() and (v0.view(np.uint32) == v1.view(np.uint32)).all()
                assert (s0.view(np.uint32) == s1.view(np.uint32)).all()
        other.debug_stages((0, -32, 0), 64)  # eager work between replays
    mesher.clear_csg_ops()


def test_launch_shapes_do_not_change_results():
    """Every launch shape (column split, persistent workgroups per CU, workgroups over the
    batch, lanes per edge in the crossing stage including the sample-parallel kernel, x/z-edge
    heights per sample or once per line) and the autotuner's pick give bit-identical meshes."""
    w = workloads.terrain_512()
    setup_field(w["desc"], w["seed"])
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    pick = list(range(150, 330, 9))
    b = mesher.Batch(64, len(pick), edge_cap=262144, voxel_cap=131072)
    b.set_chunks(w["mins"][pick], w["sizes"][pick])
    b.run()
    b.wait()
    ref = [b.get_mesh(i) for i in range(len(pick))]
    assert sum(len(m[1]) > 0 for m in ref) >= 5
    shapes = [dict(density_ysplit=3, persistent_per_cu=1, row_blocks=1, voxel_blocks=7, cross_lanes=2, density_per_cu=2,
                   height_lines=2),
              dict(density_ysplit=11, persistent_per_cu=2, row_blocks=37, voxel_blocks=1, cross_lanes=4, density_per_cu=1,
                   height_lines=1),
              dict(density_ysplit=66, persistent_per_cu=1, row_blocks=100000, voxel_blocks=100000, cross_lanes=8,
                   density_per_cu=1, height_lines=2),
              dict(density_ysplit=2, persistent_per_cu=2, row_blocks=4096, voxel_blocks=4096, cross_lanes=1,
                   density_per_cu=2, height_lines=2),
              dict(density_ysplit=1, persistent_per_cu=2, row_blocks=4096, voxel_blocks=4096, cross_lanes=1,
                   density_per_cu=1, height_lines=1),
              # sample-parallel crossing, with x/z-edge heights per line and per sample
              dict(density_ysplit=1, persistent_per_cu=2, row_blocks=4096, voxel_blocks=4096, cross_lanes=16,
                   density_per_cu=1, height_lines=2),
              dict(density_ysplit=6, persistent_per_cu=1, row_blocks=512, voxel_blocks=4096, cross_lanes=16,
                   density_per_cu=1, height_lines=1)]
    for shape in shapes:
        b.set_tuning(**shape)
        assert b.tuning() == shape
        b.run()
        b.wait()
        for i in range(len(pick)):
            assert all((x.view(np.uint32) == y.view(np.uint32)).all() for x, y in zip(b.get_mesh(i), ref[i])), \
                f"{shape}: chunk {pick[i]}"
    chosen = b.autotune(reps=1)
    assert b.tuning() == chosen
    b.run()
    b.wait()
    for i in range(len(pick)):
        assert all((x.view(np.uint32) == y.view(np.uint32)).all() for x, y in zip(b.get_mesh(i), ref[i]))
    with pytest.raises(RuntimeError):
        b.set_tuning(persistent_per_cu=3)
    # FP16 storage: the end samples are evaluated (by the row's last lane in the
    # sample-parallel kernel) instead of read from the field
    b16 = mesher.Batch(64, len(pick), edge_cap=262144, voxel_cap=131072, flags=CTX_FP16_DENSITY)
    b16.set_chunks(w["mins"][pick], w["sizes"][pick])
    meshes = []
    for lanes in (1, 16):
        b16.set_tuning(**dict(b16.tuning(), cross_lanes=lanes))
        b16.run()
        b16.wait()
        meshes.append([b16.get_mesh(i) for i in range(len(pick))])
    for i in range(len(pick)):
        assert all((x.view(np.uint32) == y.view(np.uint32)).all() for x, y in zip(meshes[0][i], meshes[1][i]))
    b16.close()
    with pytest.raises(RuntimeError):
        b.set_tuning(cross_lanes=3)
    with pytest.raises(RuntimeError):
        b.set_tuning(density_per_cu=3)
    with pytest.raises(RuntimeError):
        b.set_tuning(height_lines=3)
    mesher.clear_csg_ops()


def test_batched_mixed_lods_match_oracle(oracle):
    """A ragged batch: clipmap chunks of every LOD (voxel sizes 1..32) in one batch with
    full per-slot capacities, from empty chunks to dense LOD-5 cave chunks, each
    bit-identical to the oracle."""
    w = workloads.clipmap()
    setup_field(w["desc"], w["seed"])
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    # per LOD 0..3 and 5: a surface or cave chunk and an empty (or sparse) one; chunk 0 is a
    # LOD-5 cave chunk with ~237k of its 262,144 voxels active (the per-slot maximum case)
    pick = [930, 1110, 923, 1117, 755, 1137, 748, 1140, 0, 1888]
    mins, sizes = w["mins"][pick], w["sizes"][pick]
    assert sorted(set((sizes // 256).tolist())) == [1, 2, 4, 8, 32]
    b = mesher.Batch(64, len(pick), edge_cap=0, voxel_cap=0)
    b.set_chunks(mins, sizes)
    b.run()
    b.wait()
    counts = b.counts()
    assert (counts[:, 4] == 0).all()
    perm = oracle.perm_table(w["seed"])
    for i, (mn, size) in enumerate(zip(mins, sizes)):
        c = oracle.Chunk(64).generate(w["desc"], perm, mn, int(size), ops=_ops_for_chunk(w, mn, int(size))).build()
        v, t, s = b.get_mesh(i)
        assert (t == c.triangles()).all(), f"chunk {pick[i]} (size {size}) triangles"
        assert (v.view(np.uint32) == c.vertices().view(np.uint32)).all(), f"chunk {pick[i]} vertices"
        assert (s.view(np.uint32) == c.seams().view(np.uint32)).all(), f"chunk {pick[i]} seams"
    print(f"\nmixed LODs: {len(pick)} chunks, active voxels {counts[:, 1].min()}..{counts[:, 1].max()}, "
          f"empty {(counts[:, 2] == 0).sum()}")
    assert counts[:, 1].max() > 200000 and (counts[:, 2] == 0).any()
    mesher.clear_csg_ops()


def test_leven_clipmap_batch_matches_oracle(oracle):
    """The whole 6-LOD clipmap over the reference's own terrain field (DensityFunc) in one
    batch; per LOD the densest chunk, the two chunks with the most overlapping brushes and
    an empty chunk are bit-identical to the oracle, and every chunk stays within 0..60k
    active voxels."""
    w = workloads.clipmap_leven()
    setup_field(w["desc"], w["seed"])
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    n = len(w["sizes"])
    b = mesher.Batch(64, n, edge_cap=0, voxel_cap=0)
    b.set_chunks(w["mins"], w["sizes"])
    b.run()
    b.wait()
    counts = b.counts()
    assert (counts[:, 4] == 0).all()
    assert counts[:, 1].max() <= 60000
    lods = np.log2(w["sizes"] // 256).astype(int)
    nops = np.array([len(_ops_for_chunk(w, mn, int(size))) for mn, size in zip(w["mins"], w["sizes"])])
    pick = []
    for lod in range(6):
        sel = np.where(lods == lod)[0]
        assert len(sel), f"no chunk of LOD {lod}"
        pick.append(int(sel[np.argmax(counts[sel, 1])]))
        pick += [int(i) for i in sel[np.argsort(-nops[sel], kind="stable")[:2]] if nops[i] > 0]
        empty = sel[counts[sel, 1] == 0]
        if len(empty):
            pick.append(int(empty[len(empty) // 2]))
    pick = sorted(set(pick))
    perm = oracle.perm_table(w["seed"])
    for i in pick:
        mn, size = w["mins"][i], int(w["sizes"][i])
        c = oracle.Chunk(64).generate(w["desc"], perm, mn, size, ops=_ops_for_chunk(w, mn, size)).build()
        v, t, s = b.get_mesh(i)
        assert (t == c.triangles()).all(), f"chunk {i} (size {size}) triangles"
        assert (v.view(np.uint32) == c.vertices().view(np.uint32)).all(), f"chunk {i} vertices"
        assert (s.view(np.uint32) == c.seams().view(np.uint32)).all(), f"chunk {i} seams"
    print(f"\nleven clipmap: {n} chunks, {len(pick)} checked, active voxels up to {counts[:, 1].max()}, "
          f"{int((nops > 0).sum())} chunks under brushes")
    assert (nops[pick] > 0).sum() >= 4 and counts[pick, 1].max() > 1000
    b.close()
    mesher.clear_csg_ops()


@pytest.mark.parametrize("n", [16, 32, 64])
def test_height_lines_match_oracle(oracle, n):
    """x/z-edge heights evaluated once per line (k_height_lines, `height_lines` 2) and per
    sample (1) give the oracle's meshes bit for bit, for every chunk size: caves and brushes
    at several LODs, for 64^3 chunks the clipmap's cave chunks up to the dense LOD-5 one."""
    if n == 64:
        w = workloads.clipmap()
        pick = [930, 923, 755, 748, 0]
        mins, sizes = w["mins"][pick], w["sizes"][pick]
    else:
        w = workloads.terrain_512()
        leaf = 4 * n
        mins = np.array([(0, -leaf, 0), (leaf, -leaf, 2 * leaf), (2 * leaf, 0, leaf), (3 * leaf, -4 * leaf, 3 * leaf),
                         (0, -2 * leaf, 4 * leaf), (4 * leaf, -2 * leaf, 0)], np.int32)
        sizes = np.array([leaf, leaf, leaf, leaf, 2 * leaf, 2 * leaf], np.int32)
    setup_field(w["desc"], w["seed"])
    for op, bb in zip(w["ops"], w["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    perm = oracle.perm_table(w["seed"])
    want = [oracle.Chunk(n).generate(w["desc"], perm, mn, int(size), ops=_ops_for_chunk(w, mn, int(size))).build()
            for mn, size in zip(mins, sizes)]
    b = mesher.Batch(n, len(sizes), edge_cap=0, voxel_cap=0)
    b.set_chunks(mins, sizes)
    tris = 0
    for lines in (2, 1):
        b.set_tuning(height_lines=lines)
        b.run()
        b.wait()
        assert (b.counts()[:, 4] == 0).all()
        for i, c in enumerate(want):
            v, t, s = b.get_mesh(i)
            assert (t == c.triangles()).all(), f"N {n} height_lines {lines} chunk {i} triangles"
            assert (v.view(np.uint32) == c.vertices().view(np.uint32)).all(), f"N {n} chunk {i} vertices"
            assert (s.view(np.uint32) == c.seams().view(np.uint32)).all(), f"N {n} chunk {i} seams"
            tris += len(t)
    print(f"N {n}: {len(sizes)} chunks, {tris // 2} triangles per pass", flush=True)
    assert tris > 0
    b.close()
    mesher.clear_csg_ops()


def test_context_follows_reinitialisation(oracle):
    """A context (and its captured pipeline) used across lvn_init calls with different
    fields and seeds meshes each chunk with the current field."""
    ctx = mesher.Context(16)
    b = mesher.Batch(16, 1)
    cases = [(workloads.heightfield_desc(), 1, (0, -32, 0)), (workloads.terrain_desc(), 2, (64, -64, 0)),
             (workloads.heightfield_desc(), 7, (128, -32, 64))]
    for desc, seed, mn in cases:
        setup_field(desc, seed)
        v, t, s = ctx.generate_chunk_mesh(mn, 64)
        c = oracle_chunk(oracle, desc, seed, 16, mn, 64)
        assert len(t) > 0 and (t == c.triangles()).all() and (v.view(np.uint32) == c.vertices().view(np.uint32)).all()
        b.set_chunks(np.array([mn]), np.array([64]))
        b.run()
        b.wait()
        bv, bt, _ = b.get_mesh(0)
        assert (bt == t).all() and (bv.view(np.uint32) == v.view(np.uint32)).all()
    ctx.close()
    b.close()
