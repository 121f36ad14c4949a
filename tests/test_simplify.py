"""Pins of the simplifier oracle (CPU only): ngMeshSimplifier (src/ng_mesh_simplify.cpp)
with its two-endpoint SSE QEF (src/qef_simd.h:403-533).

* The random draws: an independent MT19937 pinned by the C++ standard's known answer, with
  libstdc++'s mapping, reproduces the oracle's iterations (which call std::mt19937 and
  std::uniform_int_distribution themselves).
* The QEF: closed forms where the solver performs no rotation (axis-aligned normals), the
  error consistent with the returned position, the error no lower than the least-squares
  minimum, symmetry in the two endpoints.
* The collapse rule: one iteration by brute force (sets and dictionaries) on small meshes.
* Invariants: edge collapses map the triangle chain by a simplicial map and drop exactly
  the degenerate triangles, so the oriented boundary of the mesh is kept (a closed mesh
  stays closed and consistently oriented; an open chunk mesh keeps its border bit for bit);
  meshes under 100 vertices or triangles are left alone; the iteration and target rules.
The reference's rule has no link-condition test, so a collapse can make an edge shared by
more than two triangles; that is its behaviour and is not checked against."""
import numpy as np
import pytest

from leven_amd import abi, workloads
from simplify_util import MT19937, boundary_chain, directed_edges, one_iteration, redraws, sphere_mesh, uniform_int

f32 = np.float32
OFF = np.array([64, 96, 64, 0], f32)


def test_mt19937_known_answer():
    """[rand.predef]: the 10000th output of a default-constructed std::mt19937 is 4123659995."""
    g = MT19937()
    for _ in range(9999):
        g()
    assert g() == 4123659995


def test_lemire_mapping_is_uniform_and_in_range():
    g = MT19937(42)
    n = 7
    counts = np.bincount([uniform_int(g, n) for _ in range(7000)], minlength=n)
    assert counts.sum() == 7000 and len(counts) == n
    assert counts.min() > 850 and counts.max() < 1150
    # the redraw rule: for n = 3 the low word must be >= 2^32 mod 3 = 1
    assert ((1 << 32) - 3) % 3 == 1


# ---- QEF ----------------------------------------------------------------------------

def _axis_normals(rng):
    axes = np.eye(3, dtype=f32)
    a, b = rng.integers(0, 3, 2)
    s0, s1 = rng.choice([-1.0, 1.0], 2)
    return (np.append(axes[a] * f32(s0), f32(0)), np.append(axes[b] * f32(s1), f32(0)))


def test_qef_axis_normals_closed_form(oracle):
    """With axis-aligned normals ATA is diagonal, every Givens coefficient is (1, 0) and the
    solve is x_c = (ATb_c - mp_c ATA_cc) / ATA_cc where min(|ATA_cc|, 1 / |ATA_cc|) >= 0.1,
    else 0; the error is |ATb - x ATA|^2 with x relative to the mass point."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        n0, n1 = _axis_normals(rng)
        p0 = np.append(rng.uniform(-20, 20, 3).astype(f32), f32(1))
        p1 = np.append(rng.uniform(-20, 20, 3).astype(f32), f32(1))
        pos, err = oracle.simplify_qef(p0, n0, p1, n1)
        a = (n0 * n0 + n1 * n1).astype(f32)[:3]                       # diagonal of ATA
        d0 = f32(np.dot(p0[:3].astype(np.float64), n0[:3]))            # one nonzero product: exact
        d1 = f32(np.dot(p1[:3].astype(np.float64), n1[:3]))
        atb = (d0 * n0 + d1 * n1).astype(f32)[:3]
        mp = ((p0 + p1) / f32(2)).astype(f32)
        rhs = (atb - mp[:3] * a).astype(f32)
        inv = np.where(np.minimum(np.abs(a), np.abs(f32(1) / np.where(a == 0, f32(1), a))) >= f32(0.1),
                       f32(1) / np.where(a == 0, f32(1), a), f32(0)).astype(f32)
        inv[a == 0] = 0
        x = (rhs * inv).astype(f32)
        t = (atb - x * a).astype(f32)
        assert pos[:3].tobytes() == (x + mp[:3]).astype(f32).tobytes()
        assert pos[3] == mp[3] == 1
        assert abs(float(err) - float(np.dot(t.astype(np.float64), t))) <= 1e-6 * max(1.0, float(err))


def test_qef_endpoint_symmetry_and_error_bounds(oracle):
    rng = np.random.default_rng(9)
    for _ in range(300):
        p0 = np.append(rng.uniform(-30, 30, 3), 1).astype(f32)
        p1 = (p0 + np.append(rng.uniform(-4, 4, 3), 0)).astype(f32)
        n0 = np.append(rng.normal(size=3), 0).astype(f32)
        n1 = np.append(rng.normal(size=3), 0).astype(f32)
        if rng.random() < 0.5:  # unit normals, as the mesher makes them
            n0[:3] /= np.linalg.norm(n0[:3])
            n1[:3] /= np.linalg.norm(n1[:3])
        pos, err = oracle.simplify_qef(p0, n0, p1, n1)
        pos2, err2 = oracle.simplify_qef(p1, n1, p0, n0)
        assert pos.tobytes() == pos2.tobytes() and err == err2  # sums of two terms commute
        # the error belongs to the returned position: |ATb - ATA (pos - mp)|^2
        A = np.outer(n0[:3], n0[:3]).astype(np.float64) + np.outer(n1[:3], n1[:3])
        atb = np.dot(p0[:3], n0[:3]) * n0[:3].astype(np.float64) + np.dot(p1[:3], n1[:3]) * n1[:3]
        mp = (p0[:3].astype(np.float64) + p1[:3]) / 2
        x = pos[:3].astype(np.float64) - mp
        r = atb - A @ x
        e64 = float(r @ r)
        assert np.isfinite(err) and err >= 0
        assert abs(float(err) - e64) <= 1e-3 * max(1.0, e64, float(np.abs(A).max() * np.abs(x).max()) ** 2)
        # no position beats the least-squares residual of ATA x = ATb
        sol = np.linalg.lstsq(A, atb, rcond=None)[0]
        r0 = atb - A @ sol
        assert float(err) >= float(r0 @ r0) * (1 - 1e-4) - 1e-4


# ---- the collapse rule ---------------------------------------------------------------

def _opts(**kw):
    return abi.simplify_options(256, **kw)


AGGRESSIVE = dict(max_error=1e9, max_edge_len=100.0, min_cos_angle=-1.0)


@pytest.mark.parametrize("case", range(4))
def test_one_iteration_brute_force(oracle, case):
    perm = oracle.perm_table(1)
    rng = np.random.default_rng(case)
    v, t = sphere_mesh(oracle, perm, rng.uniform(15, 17, 3), rng.uniform(4.5, 7))
    for kw in (dict(), AGGRESSIVE, dict(AGGRESSIVE, edge_fraction=0.6), dict(max_error=12.0, min_cos_angle=0.2)):
        opt = _opts(max_iterations=1, **kw)
        vo, to, st = oracle.simplify_mesh(v, t, OFF, opt)
        vb, tb = one_iteration(v, t, OFF, opt, oracle.simplify_qef)
        assert st[0] == 1
        assert vo.tobytes() == vb.tobytes() and to.tobytes() == tb.tobytes(), kw


def test_one_iteration_brute_force_open_mesh(oracle):
    """A piece of an open heightfield chunk (its border is boundary, never collapsed)."""
    w = workloads.plain_cpu()
    c = oracle.Chunk(16).generate(w["desc"], oracle.perm_table(w["seed"]), (0, -32, 0), 64).build()
    v, t = c.vertices(), c.triangles()
    assert len(t) >= 100 and len(v) >= 100
    off = np.array([32, 0, 32, 0], f32)
    for kw in (dict(), AGGRESSIVE):
        opt = abi.simplify_options(64, max_iterations=1, **kw)
        vo, to, _ = oracle.simplify_mesh(v, t, off, opt)
        vb, tb = one_iteration(v, t, off, opt, oracle.simplify_qef)
        assert vo.tobytes() == vb.tobytes() and to.tobytes() == tb.tobytes()


def test_redraw_seed_is_followed(oracle):
    """A seed whose first draws include a redraw (low word below 2^32 mod E): the oracle's
    iteration still equals the brute force, which redraws as libstdc++ does."""
    perm = oracle.perm_table(1)
    v, t = sphere_mesh(oracle, perm, (16.2, 15.9, 16.1), 5.0)
    opt0 = _opts(max_iterations=0, **AGGRESSIVE)
    _, _, st = oracle.simplify_mesh(v, t, OFF, opt0)
    E = int(st[2])  # candidate edges before any iteration
    seed = next(s for s in range(1, 20000) if redraws(s, E, E) > 0)
    opt = _opts(max_iterations=1, edge_fraction=1.0, seed=seed, **AGGRESSIVE)
    vo, to, _ = oracle.simplify_mesh(v, t, OFF, opt)
    vb, tb = one_iteration(v, t, OFF, opt, oracle.simplify_qef)
    assert vo.tobytes() == vb.tobytes() and to.tobytes() == tb.tobytes()


# ---- invariants ----------------------------------------------------------------------

@pytest.mark.parametrize("case", range(6))
def test_closed_mesh_stays_closed_and_oriented(oracle, case):
    perm = oracle.perm_table(1)
    rng = np.random.default_rng(40 + case)
    v, t = sphere_mesh(oracle, perm, rng.uniform(15, 17, 3), rng.uniform(5, 13))
    assert not boundary_chain(v, t)
    for kw in (dict(), AGGRESSIVE, dict(max_error=50.0, max_edge_len=10.0, min_cos_angle=0.0)):
        opt = _opts(max_iterations=40, **kw)
        vo, to, st = oracle.simplify_mesh(v, t, OFF, opt)
        assert len(to) < len(t) and st[1] > 0
        d = directed_edges(to)
        assert all(d.get((b, a), 0) == m for (a, b), m in d.items())  # every edge's two sides pair up
        assert all(len(set(tr)) == 3 for tr in to.tolist())
        assert to.min() >= 0 and to.max() == len(vo) - 1 and len(np.unique(to)) == len(vo)


def _round_trip(v, off):
    """Positions as the simplifier returns an unmoved vertex: (p - offset) + offset in FP32."""
    out = v.copy()
    out["pos"] = ((v["pos"] - off).astype(f32) + off).astype(f32)
    return out


def _border(v, t):
    """The directed sides (as position pairs) of the edges one triangle uses, leaving out the
    largest (max, min) index key, which the reference never examines (its loop ends before
    the last run, src/ng_mesh_simplify.cpp:144-166)."""
    count = {}
    for tr in t.tolist():
        for a, b in ((tr[0], tr[1]), (tr[1], tr[2]), (tr[2], tr[0])):
            k = (max(a, b), min(a, b))
            count.setdefault(k, []).append((a, b))
    last = max(count)
    return {(v["pos"][a].tobytes(), v["pos"][b].tobytes()) for k, sides in count.items()
            if len(sides) == 1 and k != last for (a, b) in sides}


def test_open_mesh_keeps_its_border(oracle):
    """An open chunk mesh: the ends of edges one triangle uses are never collapsed, and the
    triangle on such an edge can only lose its third vertex to a collapse elsewhere, so
    every border edge is in the output with its direction and its ends' positions."""
    w = workloads.plain_cpu()
    c = oracle.Chunk(64).generate(w["desc"], oracle.perm_table(w["seed"]), w["mins"][0], int(w["sizes"][0])).build()
    v, t = c.vertices(), c.triangles()
    off = np.append(w["mins"][0] + 128, 0).astype(f32)
    vr = _round_trip(v, off)
    border = _border(vr, t)
    assert len(border) > 100
    inputs = {vr["pos"][i].tobytes() for i in range(len(v))}
    for kw in (dict(), AGGRESSIVE):
        vo, to, st = oracle.simplify_mesh(v, t, off, _opts(**kw))
        assert 0 < st[0] <= 10 and len(to) < len(t)
        out = {(vo["pos"][a].tobytes(), vo["pos"][b].tobytes()) for (a, b) in directed_edges(to)}
        assert border <= out
        # every vertex either keeps its input position or took a collapse position, and
        # carries an input vertex's colour
        moved = sum(vo["pos"][i].tobytes() not in inputs for i in range(len(vo)))
        assert 0 < moved < len(vo)
        assert set(map(bytes, vo["colour"].view(np.uint8).reshape(len(vo), -1))) <= \
            set(map(bytes, v["colour"].view(np.uint8).reshape(len(v), -1)))


def test_small_meshes_unchanged_and_stopping_rules(oracle):
    perm = oracle.perm_table(1)
    v, t = sphere_mesh(oracle, perm, (16.3, 16.4, 15.7), 6.0)
    # under 100 triangles or 100 vertices: returned as given
    for vv, tt in ((v, t[:99]), (v[:99], t[np.all(t < 99, axis=1)])):
        vo, to, st = oracle.simplify_mesh(vv, tt, OFF, _opts(**AGGRESSIVE))
        assert vo.tobytes() == vv.tobytes() and to.tobytes() == tt.tobytes() and st[0] == 0
    # the triangle target stops the loop: at most one iteration past it is never run
    opt = _opts(max_iterations=100, target_percentage=0.5, **AGGRESSIVE)
    _, to, st = oracle.simplify_mesh(v, t, OFF, opt)
    assert len(to) <= int(f32(len(t)) * f32(0.5)) and st[0] < 100
    _, to_prev, _ = oracle.simplify_mesh(v, t, OFF, _opts(max_iterations=int(st[0]) - 1, target_percentage=0.5,
                                                          **AGGRESSIVE))
    assert len(to_prev) > int(f32(len(t)) * f32(0.5))
    # max_iterations = 0 leaves the mesh (compacted, offset round trip) as it was
    vo, to, st = oracle.simplify_mesh(v, t, OFF, _opts(max_iterations=0))
    assert st[0] == 0 and to.tobytes() == t.tobytes() and len(vo) == len(v)
