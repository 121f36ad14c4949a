"""Helpers of the simplifier tests: an independent MT19937, the libstdc++ draw mapping, a
brute-force restatement of one simplification iteration, and mesh invariants.

The brute-force iteration is written from ngMeshSimplifier's rules
(src/ng_mesh_simplify.cpp:118-381) with sets and dictionaries instead of the sorted
arrays, per-vertex minima and compaction passes the oracle and the kernel use, so a
mistake of order, index or tie-breaking in either shows up as a difference."""
import numpy as np


class MT19937:
    """The C++ standard's std::mt19937 ([rand.predef]: w 32, n 624, m 397, r 31)."""

    def __init__(self, seed=5489):
        mt = [0] * 624
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mt, self.i = mt, 624

    def _twist(self):
        mt = self.mt
        for k in range(624):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
            mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self.i = 0

    def __call__(self):
        if self.i >= 624:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def uniform_int(g, n):
    """std::uniform_int_distribution<int>(0, n - 1) as libstdc++ 11+ maps a 32-bit engine's
    output (Lemire's multiply-and-reject)."""
    prod = g() * n
    low = prod & 0xFFFFFFFF
    if low < n:
        thr = ((1 << 32) - n) % n
        while low < thr:
            prod = g() * n
            low = prod & 0xFFFFFFFF
    return prod >> 32


def redraws(seed, n, draws):
    """Number of redraws among the first `draws` values of uniform_int over [0, n)."""
    g = MT19937(seed)
    thr = ((1 << 32) - n) % n
    k, count = 0, 0
    while k < draws:
        if ((g() * n) & 0xFFFFFFFF) < thr:
            count += 1
        else:
            k += 1
    return count


f32 = np.float32


def glm_dot4(a, b):
    a, b = a.astype(f32), b.astype(f32)
    return f32(f32(a[0] * b[0] + a[1] * b[1]) + f32(a[2] * b[2] + a[3] * b[3]))


def one_iteration(vertices, triangles, offset, opt, qef):
    """One iteration of the simplifier, by brute force, followed by the final vertex
    compaction: returns (vertices, triangles) as the simplifier with max_iterations = 1
    does.  qef(p0, n0, p1, n1) -> (position, error) is the two-endpoint QEF."""
    pos = (vertices["pos"].astype(f32) - np.asarray(offset, f32)).astype(f32)
    nrm = vertices["normal"].astype(f32).copy()
    col = vertices["colour"].astype(f32)
    tris = [tuple(int(i) for i in t) for t in triangles]
    # candidate edges: each triangle side as (min, max); sides seen once mark their ends as
    # boundary; the edge with the largest (max, min) key is never examined
    count = {}
    for t in tris:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[0], t[2])):
            e = (min(a, b), max(a, b))
            count[e] = count.get(e, 0) + 1
    order = sorted(count, key=lambda e: (e[1], e[0]))
    examined = order[:-1]
    boundary = set()
    for e in examined:
        if count[e] == 1:
            boundary.update(e)
    edges = [e for e in examined if count[e] > 1 and e[0] not in boundary and e[1] not in boundary]
    degree = {}
    for t in tris:
        for i in t:
            degree[i] = degree.get(i, 0) + 1
    E = len(edges)
    draws = int(f32(f32(E) * f32(opt.edge_fraction)))
    g = MT19937(opt.seed)
    picks = {uniform_int(g, E) for _ in range(draws)}
    best = {}  # vertex -> (cost, edge index) of its cheapest valid candidate
    moved = {}
    for i in sorted(picks):
        a, b = edges[i]
        if glm_dot4(nrm[a], nrm[b]) < f32(opt.min_angle_cosine):
            continue
        d = (pos[b] - pos[a]).astype(f32)
        if glm_dot4(d, d) > f32(f32(opt.max_edge_size) * f32(opt.max_edge_size)):
            continue
        if abs(float(f32(col[a][3] - col[b][3]))) > 1e-3:
            continue
        deg = degree.get(a, 0) + degree.get(b, 0)
        if deg > 16:
            continue
        p, err = qef(pos[a], nrm[a], pos[b], nrm[b])
        err = f32(err)
        if err > 0:
            err = f32(f32(1) / err)
        err = f32(err + f32(max(0, deg - 10)) * f32(f32(opt.max_error) * f32(0.1)))
        if err > f32(opt.max_error):
            continue
        moved[i] = (np.array([p[0], p[1], p[2], 1.0], f32), ((nrm[a] - nrm[b]) * f32(0.5)).astype(f32))
        for v in (a, b):
            if (err, i) < best.get(v, (f32(np.inf), -1)) and err < np.finfo(f32).max:
                best[v] = (err, i)
    target = {}
    if moved:
        for i, (p, n) in moved.items():
            a, b = edges[i]
            if best.get(a, (None, -1))[1] == i and best.get(b, (None, -1))[1] == i:
                target[b] = a
                pos[a] = p
                nrm[a] = n
    out = []
    for t in tris:
        m = tuple(target.get(i, i) for i in t)
        if len(set(m)) == 3:
            out.append(m)
    used = sorted({i for t in out for i in t})
    remap = {v: k for k, v in enumerate(used)}
    vo = vertices[used].copy()
    vo["pos"] = (pos[used] + np.asarray(offset, f32)).astype(f32)
    vo["normal"] = nrm[used]
    return vo, np.array([[remap[i] for i in t] for t in out], np.int32).reshape(-1, 3)


def directed_edges(triangles):
    d = {}
    for t in triangles:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            d[(int(a), int(b))] = d.get((int(a), int(b)), 0) + 1
    return d


def boundary_chain(vertices, triangles):
    """The oriented boundary of the triangle chain as {(pos a, pos b): multiplicity}, with
    opposite directions cancelling (keys are position byte strings)."""
    key = [vertices["pos"][i].tobytes() for i in range(len(vertices))]
    chain = {}
    for (a, b), m in directed_edges(triangles).items():
        ka, kb = key[a], key[b]
        chain[(ka, kb)] = chain.get((ka, kb), 0) + m
        chain[(kb, ka)] = chain.get((kb, ka), 0) - m
    return {k: v for k, v in chain.items() if v > 0}


def sphere_mesh(oracle, perm, centre, radius, n=32):
    """An added sphere brush in air: a closed, consistently oriented mesh."""
    from leven_amd import abi
    op = np.zeros(1, abi.CSG_OP_DTYPE)
    op["type"], op["brush"], op["material"] = 0, 1, 1
    op["origin"] = [centre[0], centre[1] + 8.0, centre[2], 0]
    op["dims"] = [radius, radius, radius, 0]
    # the plane d = y lies below the chunk (voxel y >= 8): air everywhere but the brush
    c = oracle.Chunk(n).generate(abi.field_desc(height_octaves=0), perm, (0, 4 * 8, 0), n * 4, ops=op).build()
    return c.vertices(), c.triangles()
