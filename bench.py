"""Benchmark: 64^3 chunks meshed per second on one MI355X, batched vs plain.

    python bench.py [--steps K] [--warmup W] [--workload terrain_512]

A step is one batched pass of the whole pipeline over every chunk of the
workload (default: the 512-chunk 3D terrain with 256 CSG brushes), inputs staged
in HBM before the timed region.  The plain per-chunk form (one chunk per launch
sequence, synchronised per chunk) is timed over the same chunks; the CPU
baseline is the oracle on the host's cores.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from leven_amd import abi, mesher, workloads  # noqa: E402

METRIC = "64^3 chunks meshed/sec (batched vs plain, same GPU) + p99 chunk latency"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak (fused multiply-adds)
PEAK_HBM_GBS = 8000.0

# FP32 operations per evaluation, counted from the kernel source (add/sub/mul/div/sqrt = 1,
# fused multiply-add = 2; selects, max and floor not counted)
FLOPS_SNOISE2 = 49
FLOPS_SNOISE3 = 87


# leven's Terrain (cl/noise.cl:205-223) as height term: RidgedMultiFractal over 1 + 7 snoise2
# (9 operations per octave, 4 around them), fBm over 4 and over 2 snoise2 (5 per octave, 2 to
# scale the point), 13 for the blend; times 900 and y - H
FLOPS_LEVEN_TERRAIN = (8 * FLOPS_SNOISE2 + 7 * 9 + 4) + (4 * (FLOPS_SNOISE2 + 5) + 2) + (2 * (FLOPS_SNOISE2 + 5) + 2) + 13


def flops_height(desc):
    if desc.kind == 2:
        return FLOPS_LEVEN_TERRAIN + 2
    return 3 + (FLOPS_SNOISE2 + 5) * desc.height_octaves


def flops_cave(desc):
    return 5 + (FLOPS_SNOISE3 + 6) * desc.cave_octaves


def stage_flops(w, desc):
    """FLOPs the density and edge-crossing stages execute, from the kernels' work counters
    (after the cave-term cut, shared heights and stored end samples)."""
    fh, fc = flops_height(desc), flops_cave(desc)
    density = w["density_height"] * fh + w["density_cave"] * fc + w["samples"] * 2
    cross = w["cross_height"] * fh + w["cross_cave"] * fc + w["edges"] * 200 + w["cross_brush"] * 20
    return density, cross


def algorithmic_flops(chunks, edges, desc):
    """Algorithmic FLOPs per launch: 65^3 density evaluations per chunk, and 23 per edge
    (17 crossing samples + 6 normal samples), each a full evaluation of the field
    (height term + cave term) at the FLOP counts above."""
    fd = flops_height(desc) + (flops_cave(desc) if desc.kind == 1 else 0)
    return chunks * 65 ** 3 * fd, edges * 23 * fd


def hbm_traffic(kernel, workload):
    """HBM bytes per batched launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/hbm_traffic_<workload>.json, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    for name, entries in data["kernels"].items():
        if name.split("::")[-1].split("<")[0] == kernel and "<true>" not in name:
            e = max(entries, key=lambda x: x["grid"])  # the batched launch
            return e["read_bytes"] + e["write_bytes"]
    return None


# stage -> kernels whose HBM bytes it moves (batched launch)
STAGE_KERNELS = {"density": ["k_density"], "edge_rows": ["k_solid_rows", "k_edge_rows"],
                 "edge_compact": ["k_edge_compact"], "edge_cross": ["k_edge_cross"], "voxel_rows": ["k_voxel_rows"],
                 "voxel_compact": ["k_voxel_compact"], "qef": ["k_qef"], "emit": ["k_emit"]}


def hbm_rates(stage_ms, workload):
    """HBM GB/s of each stage against the 8 TB/s peak: the committed rocprofv3 byte counts
    per batched launch over the stage times measured in this run."""
    out = {}
    for stage, kernels in STAGE_KERNELS.items():
        total = [hbm_traffic(k, workload) for k in kernels]
        if any(t is None for t in total) or stage_ms.get(stage, 0) <= 0:
            continue
        gbs = sum(total) / (stage_ms[stage] * 1e-3) / 1e9
        out[stage] = {"gbs": round(gbs, 1), "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes": int(sum(total))}
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, threads):
    """The oracle (leven's semantics restated in C++, FP32) on the host: all chunks on
    `threads` threads (one chunk per thread at a time), plus a single-thread sample."""
    from oracle import py_oracle
    perm = py_oracle.perm_table(wl["seed"])
    wall, secs, totals = py_oracle.mesh_chunks(wl["desc"], perm, 64, wl["mins"], wl["sizes"], 0, wl["ops"],
                                               wl["aabbs"], threads=threads)
    n = len(wl["sizes"])
    one = np.linspace(0, n - 1, 16).astype(int)
    wall1, _, _ = py_oracle.mesh_chunks(wl["desc"], perm, 64, wl["mins"][one], wl["sizes"][one], 0, wl["ops"],
                                        wl["aabbs"], threads=1)
    return dict(value=n / wall, unit="chunks/s", cores=threads, kind="port",
                single_thread_chunks_per_s=round(len(one) / wall1, 2),
                sample=f"all {n} chunks of {wl['name']} on {threads} threads ({os.cpu_count()} logical CPUs "
                       f"visible, {_cpu_model()}), oracle (C++ restatement of leven's semantics, FP32, one chunk "
                       f"per thread; no CPU OpenCL device exists to run leven itself): {wall:.2f} s wall, p50 "
                       f"{np.percentile(secs, 50) * 1e3:.1f} ms, p99 {np.percentile(secs, 99) * 1e3:.1f} ms per chunk; "
                       f"single thread {len(one)} chunks in {wall1:.2f} s")


def precision_leg(args):
    """The precision form on the 512-chunk field with 1024 Y-rotated boxes, against the
    FP32 form on the same chunks: compensated FP32 QEF with FP64 re-solves near the
    truncation thresholds, with and without FP16 density storage; same batched pipeline."""
    from leven_amd.abi import CTX_FP16_DENSITY, CTX_PRECISE_QEF
    wl = workloads.precision()
    mesher.init(wl["desc"], wl["seed"], wl["default_material"])
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    n = len(wl["sizes"])
    out = {"chunks": n, "csg_ops": int(len(wl["ops"]))}
    for name, flags in (("fp32", 0), ("precise_qef", CTX_PRECISE_QEF),
                        ("precise_qef_fp16_density", CTX_PRECISE_QEF | CTX_FP16_DENSITY)):
        b = mesher.Batch(64, n, edge_cap=262144, voxel_cap=131072, flags=flags)
        b.set_chunks(wl["mins"], wl["sizes"])
        tuning = b.autotune(reps=1)
        for _ in range(max(1, args.warmup)):
            b.run()
        b.wait()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        b.wait()
        elapsed = time.perf_counter() - t0
        counts = b.counts()
        assert (counts[:, 4] == 0).all(), "a precision-form chunk overflowed"
        stage_ms = {k: round(v, 4) for k, v in b.run_timed().items()}
        b.close()
        out[name] = {"chunks_per_s": round(n * args.steps / elapsed, 1), "ms_per_step": round(elapsed / args.steps * 1e3, 3),
                     "vertices": int(counts[:, 1].sum()), "fp64_resolved": int(counts[:, 5].sum()),
                     "qef_ms": stage_ms["qef"], "density_ms": stage_ms["density"], "tuning": tuning}
    base = out["fp32"]["chunks_per_s"]
    for name in ("precise_qef", "precise_qef_fp16_density"):
        out[name]["over_fp32"] = round(out[name]["chunks_per_s"] / base, 3)
    out["accuracy"] = ("vertex error vs the FP64 solve of the same hermite data <= 1e-4 voxel; vs the FP64 "
                       "pipeline end to end see tests/test_gpu_precision.py")
    mesher.clear_csg_ops()
    return out


def seam_leg(batch, nodes, args):
    """Seam meshes of every active clipmap node (GenerateClipmapSeamMesh, src/clipmap.cpp:573-611):
    neighbour seam node lists gathered as the reference does, all seams in one call, in two
    forms: drop-in (host lists in, host meshes out, lvn_generate_seam_meshes) and
    device-resident (lists read from the clipmap batch's slots, meshes left in HBM,
    lvn_batch_seam_meshes); against the oracle on the host's cores."""
    import ctypes as C

    from leven_amd.abi import VERTEX_DTYPE, SeamOut, SeamReq, seam_sources, vptr
    per_host = workloads.seam_sources(nodes)
    flat, reqs = [], (SeamReq * len(nodes))()
    for i, ((mn, size, _), srcs) in enumerate(zip(nodes, per_host)):
        reqs[i].host_min[:] = list(mn)
        reqs[i].host_size = size
        reqs[i].colour[:] = [1.0, 1.0, 1.0, 0.0]
        reqs[i].first_source = len(flat)
        reqs[i].n_sources = len(srcs)
        flat.extend(srcs)
    src = seam_sources(flat)
    offered = [sum(len(d["nodes"]) for d in srcs) for srcs in per_host]
    outs = (SeamOut * len(nodes))()
    bufs = []
    for i, o in enumerate(offered):
        v = np.zeros(max(1, o), dtype=VERTEX_DTYPE)
        t = np.zeros((max(1, 6 * o + 64), 3), dtype=np.int32)
        bufs.append((v, t))
        outs[i].v, outs[i].tri, outs[i].v_cap, outs[i].tri_cap = v.ctypes.data, t.ctypes.data, len(v), len(t)
    ctx = mesher.Context(64)
    L = mesher.lib()
    # the drop-in form: host seam node lists in, host meshes out

    def run():
        mesher.check(L.lvn_generate_seam_meshes(ctx._h, len(nodes), C.cast(reqs, C.c_void_p), C.cast(src, C.c_void_p),
                                                C.cast(outs, C.c_void_p)), "lvn_generate_seam_meshes")

    for _ in range(max(1, args.warmup)):
        run()
    steps = max(3, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    elapsed = time.perf_counter() - t0
    nv = sum(o.nv for o in outs)
    nt = sum(o.nt for o in outs)
    # the device-resident form: the seam lists of the clipmap batch's slots, meshes left in HBM
    dev_src = seam_sources([dict(d, nodes=None) for d in flat])
    dev_outs = (SeamOut * len(nodes))()

    def run_dev():
        mesher.check(L.lvn_batch_seam_meshes(batch._h, len(nodes), C.cast(reqs, C.c_void_p),
                                             C.cast(dev_src, C.c_void_p), C.cast(dev_outs, C.c_void_p)),
                     "lvn_batch_seam_meshes")

    run_dev()
    assert sum(o.nv for o in dev_outs) == nv and sum(o.nt for o in dev_outs) == nt
    t0 = time.perf_counter()
    for _ in range(steps):
        run_dev()
    dev_elapsed = time.perf_counter() - t0
    # one seam per call (lvn_generate_seam_mesh, the drop-in for one GenerateClipmapSeamMesh)
    pick = np.linspace(0, len(nodes) - 1, 64).astype(int)
    lat = []
    snv, snt = C.c_int32(), C.c_int32()
    col = np.ones(3, dtype=np.float32)
    for i in pick:
        q = reqs[i]
        hmin = np.array(q.host_min[:], dtype=np.int32)
        first = C.cast(src, C.c_void_p).value + q.first_source * C.sizeof(src._type_)
        o = outs[i]
        t1 = time.perf_counter()
        mesher.check(L.lvn_generate_seam_mesh(ctx._h, vptr(hmin), q.host_size, C.c_void_p(first), q.n_sources,
                                              vptr(col), C.c_void_p(o.v), o.v_cap, C.byref(snv), C.c_void_p(o.tri),
                                              o.tri_cap, C.byref(snt)), "lvn_generate_seam_mesh")
        lat.append((time.perf_counter() - t1) * 1e3)
    ctx.close()
    out = {"seams": len(nodes), "sources": len(flat), "seam_nodes_offered": int(sum(offered)),
           "vertices": int(nv), "triangles": int(nt), "steps": steps,
           "seams_per_s": round(len(nodes) * steps / elapsed, 1), "ms_per_pass": round(elapsed / steps * 1e3, 3),
           "note": "drop-in form: host seam node lists in, host meshes out (one upload, one read-back per pass)",
           "device_resident": {"seams_per_s": round(len(nodes) * steps / dev_elapsed, 1),
                               "ms_per_pass": round(dev_elapsed / steps * 1e3, 3),
                               "note": "seam lists read from the clipmap batch's slots, meshes left in HBM"},
           "single_seam_ms": {"mean": round(float(np.mean(lat)), 3), "p50": round(float(np.percentile(lat, 50)), 3),
                              "p99": round(float(np.percentile(lat, 99)), 3),
                              "note": "64 seams spread over the clipmap, one lvn_generate_seam_mesh call each "
                                      "(host lists in, host mesh out)"}}
    if not args.no_cpu:
        from oracle import py_oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        wall, totals = py_oracle.seam_meshes(reqs, src, threads=threads)
        assert totals[2] == 0 and totals[0] == nv and totals[1] == nt, "oracle and GPU seam totals differ"
        one = (SeamReq * 64)()
        pick = np.linspace(0, len(nodes) - 1, 64).astype(int)
        for k, i in enumerate(pick):
            one[k] = reqs[i]
        wall1, _ = py_oracle.seam_meshes(one, src, threads=1)
        out["single_seam_ms"]["oracle_one_thread_mean"] = round(wall1 / 64 * 1e3, 3)
        out["cpu_baseline"] = {"value": round(len(nodes) / wall, 1), "unit": "seams/s", "cores": threads,
                               "kind": "port", "single_thread_seams_per_s": round(64 / wall1, 1),
                               "sample": f"all {len(nodes)} seams on {threads} threads, oracle restatement of the "
                                         f"reference's CPU seam path ({wall:.3f} s); single thread 64 seams in "
                                         f"{wall1:.3f} s"}
        out["gpu_over_cpu"] = round(out["seams_per_s"] / out["cpu_baseline"]["value"], 1)
        out["device_resident"]["gpu_over_cpu"] = round(out["device_resident"]["seams_per_s"] /
                                                       out["cpu_baseline"]["value"], 1)
    return out


def dropin_latency(wl, count):
    """The drop-in path a leven caller uses (Compute_MeshGenContext::generateChunkMesh =
    lvn_generate_chunk_mesh): one fresh chunk per call, mesh copied into host buffers."""
    import ctypes as C

    from leven_amd.abi import CSG_OP_DTYPE, MATERIAL_AIR, SEAM_DTYPE, VERTEX_DTYPE, vptr
    ctx = mesher.Context(64)
    n3 = 64 ** 3
    v = np.zeros(n3, dtype=VERTEX_DTYPE)
    t = np.zeros((6 * n3, 3), dtype=np.int32)
    s = np.zeros(n3, dtype=SEAM_DTYPE)
    nv, nt, ns = C.c_int32(), C.c_int32(), C.c_int32()
    lat = []
    idx = np.linspace(0, len(wl["sizes"]) - 1, count).astype(int)
    for i in idx:
        mn = np.ascontiguousarray(wl["mins"][i], dtype=np.int32)
        ts = time.perf_counter()
        rc = mesher.lib().lvn_generate_chunk_mesh(ctx._h, vptr(mn), int(wl["sizes"][i]), vptr(v), n3, C.byref(nv),
                                                  vptr(t), 6 * n3, C.byref(nt), vptr(s), n3, C.byref(ns))
        lat.append(time.perf_counter() - ts)
        assert rc == 0, f"lvn_generate_chunk_mesh failed: {rc}"
    # an edit the way the clipmap applies one (src/clipmap.cpp:1689-1710): applyCSGOperations
    # on the chunk's field, freeChunkOctree, generateChunkMesh.  The first edit of a chunk
    # loads its field into the context's cache; the second is timed: the incremental path
    # (cached materials, brush-touched edges re-crossed, the rest from the cached list).
    edit = []
    rng = np.random.default_rng(11)
    for i in idx[: min(len(idx), 64)]:
        mn = np.ascontiguousarray(wl["mins"][i], dtype=np.int32)
        size = int(wl["sizes"][i])
        for k in range(2):
            op = np.zeros(1, CSG_OP_DTYPE)
            op["type"], op["brush"] = k % 2, 1
            op["material"] = 2 if k % 2 == 0 else MATERIAL_AIR
            c = mn / 4 + rng.uniform(16, 48, 3)
            op["origin"] = [c[0] + 0.5, c[1] + 0.5, c[2] + 0.5, 0.0]
            op["dims"] = [6.0, 6.0, 6.0, 0.0]
            ts = time.perf_counter()
            mesher.check(mesher.lib().lvn_apply_csg(ctx._h, vptr(op), 1, vptr(mn), size), "lvn_apply_csg")
            mesher.check(mesher.lib().lvn_free_chunk(ctx._h, vptr(mn), size), "lvn_free_chunk")
            rc = mesher.lib().lvn_generate_chunk_mesh(ctx._h, vptr(mn), size, vptr(v), n3, C.byref(nv), vptr(t),
                                                      6 * n3, C.byref(nt), vptr(s), n3, C.byref(ns))
            assert rc == 0, f"lvn_generate_chunk_mesh after an edit failed: {rc}"
            if k == 1:
                edit.append(time.perf_counter() - ts)
    ctx.close()
    lat = np.array(lat) * 1e3
    edit = np.array(edit) * 1e3
    return {"chunks": int(count), "chunks_per_s": round(count / (lat.sum() * 1e-3), 1),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
            "edit_latency_ms": {"p50": round(float(np.percentile(edit, 50)), 3),
                                "p99": round(float(np.percentile(edit, 99)), 3),
                                "note": "applyCSGOperations + freeChunkOctree + generateChunkMesh on a chunk "
                                        "whose field is cached (lvn_apply_csg, incremental), mesh copied out"}}


def serve_burst():
    """Capacity of the served form: the 512 terrain_512 chunks submitted at once
    (lvn_server_submit_many) to a fresh server (all cache misses), meshed by continuous
    batching on 3 lanes of up to 176 chunks (so the burst is 3 concurrent batches), with
    launch shapes autotuned on a batch of the same chunks (lvn_server_set_tuning).  A
    burst's p99 is bounded below by the time to mesh ~99% of it."""
    wl = workloads.terrain_512()
    probe = mesher.Batch(64, len(wl["sizes"]), edge_cap=0, voxel_cap=0, flags=wl["flags"])
    probe.set_chunks(wl["mins"], wl["sizes"])
    tuning = probe.autotune(reps=2)
    probe.close()
    server = mesher.Server(64, hbm_budget=64 << 30, max_batch=176, lanes=3, edge_cap=0, voxel_cap=0)
    server.set_tuning(tuning)
    t0 = time.perf_counter()
    server.submit_meshes(wl["mins"], wl["sizes"])
    server.drain()
    wall = time.perf_counter() - t0
    res = server.poll(len(wl["sizes"]) + 16)
    lat = (res["complete_s"] - res["submit_s"]) * 1e3
    st = server.stats()
    server.close()
    return {"chunks": int(len(res)), "chunks_per_s": round(len(res) / wall, 1), "batches": st["batches"],
            "max_batch": 176, "lanes": 3, "tuning": tuning,
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)}}


def serve(args):
    """Serving form: replay the Poisson stream in real time against a server with a
    64 GB HBM budget; latency = completion - submission on the server's clock."""
    wl = workloads.serving(duration=args.serve_seconds)
    mesher.init(wl["desc"], wl["seed"], wl["default_material"])
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])
    # full per-slot capacities: a coarse clipmap chunk can have any number of active voxels
    server = mesher.Server(64, hbm_budget=64 << 30, max_batch=64, lanes=3, edge_cap=0, voxel_cap=0)
    events = wl["events"]
    t0 = time.perf_counter()
    for t, kind, payload in events:
        while time.perf_counter() - t0 < t:
            pass
        if kind == "mesh":
            server.submit_mesh(payload[0], payload[1])
        else:
            server.submit_edit(payload[0], payload[1])
    server.drain()
    wall = time.perf_counter() - t0
    res = server.poll(len(events) + 16)
    burst = serve_burst()
    mesh = res[res["kind"] == 0]
    bad = mesh[mesh["status"] != 0]
    assert len(bad) == 0, f"{len(bad)} served chunks failed, status {sorted(set(bad['status'].tolist()))}"
    lat = (mesh["complete_s"] - mesh["submit_s"]) * 1e3
    miss = mesh["cache_hit"] == 0
    st = server.stats()
    server.close()
    return {"requests": int(len(res)), "mesh_requests": int(len(mesh)), "edits": int((res["kind"] == 1).sum()),
            "offered_rate_per_s": wl["rate"], "seconds": round(wall, 3),
            "throughput_requests_per_s": round(len(res) / wall, 1),
            "chunks_meshed": st["chunks_meshed"], "batches": st["batches"], "max_batch_seen": st["max_batch_seen"],
            "cache_hits": st["cache_hits"], "remeshed_after_edit": st["remeshed_after_edit"],
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                           "max": round(float(lat.max()), 3)},
            "miss_latency_ms": {"p50": round(float(np.percentile(lat[miss], 50)), 3),
                                "p99": round(float(np.percentile(lat[miss], 99)), 3)} if miss.any() else None,
            "hbm_budget_gb": 64, "slots": st["slots"], "evictions": st["evictions"], "burst": burst}


def copy_probe(mib):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE read for it
    convert the counters of the mesher's kernels into bytes."""
    import torch

    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 5}), flush=True)


def stage_ops(wl):
    mesher.init(wl["desc"], wl["seed"], wl["default_material"])
    mesher.clear_csg_ops()
    for op, bb in zip(wl["ops"], wl["aabbs"]):
        mesher.store_csg_op(op, bb[:3], bb[3:])


def batched_leg(wl, args, steps, edge_cap, voxel_cap, plain=True, keep=False):
    """One workload through the batched form (one graph launch per pass over every chunk,
    inputs staged in HBM, launch shapes autotuned outside the timed region), per-stage HIP
    event times, work counters and the roofline of the dominant ALU stage; then the plain
    per-chunk form (one chunk per graph launch, synchronised per chunk) over the same
    chunks.  Returns (result dict, batch or None)."""
    stage_ops(wl)
    desc = wl["desc"]
    n = len(wl["sizes"])
    batch = mesher.Batch(64, n, edge_cap=edge_cap, voxel_cap=voxel_cap, flags=wl["flags"])
    batch.set_chunks(wl["mins"], wl["sizes"])
    tuning = batch.autotune(reps=1 if n > 1000 else 3) if not args.no_autotune else batch.tuning()
    for _ in range(max(1, args.warmup)):
        batch.run()
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    batch.wait()
    elapsed = time.perf_counter() - t0
    rate = n * steps / elapsed
    counts = batch.counts()
    assert (counts[:, 4] == 0).all(), f"a {wl['name']} chunk overflowed its slot capacity"
    # per-stage launch durations: HIP events around each stage's launches on the batch
    # stream, averaged over the same number of passes
    runs = [batch.run_timed() for _ in range(steps)]
    stage_ms = {k: sum(r[k] for r in runs) / len(runs) for k in runs[0]}
    work = batch.work()
    d_flops, c_flops = stage_flops(work, desc)
    dom = max(("density", "edge_cross"), key=lambda k: stage_ms[k])
    # the crossing stage's kernel for the tuned lanes per edge (1, 2/4/8, 16 = sample-parallel)
    lanes = tuning.get("cross_lanes", 1) if isinstance(tuning, dict) else 1
    dom_kernel = ("k_density" if dom == "density" else
                  "k_edge_cross_sp" if lanes == 16 else "k_edge_cross_group" if lanes in (2, 4, 8) else "k_edge_cross")
    a_d, a_c = algorithmic_flops(n, work["edges"], desc)
    dom_flops = a_d if dom == "density" else a_c
    exec_flops = d_flops if dom == "density" else c_flops
    achieved = dom_flops / (stage_ms[dom] * 1e-3) / 1e12
    lods = np.log2(wl["sizes"] // 256).astype(int)
    per_lod = {}
    for lod in sorted(set(lods.tolist())):
        sel = lods == lod
        per_lod[str(lod)] = {"chunks": int(sel.sum()), "max_active_voxels": int(counts[sel, 1].max()),
                             "mean_active_voxels": int(counts[sel, 1].mean())}
    out = {"chunks": n, "steps": steps, "chunks_per_s": round(rate, 2), "ms_per_pass": round(elapsed / steps * 1e3, 4),
           "totals": {"edges": int(counts[:, 0].sum()), "voxels": int(counts[:, 1].sum()),
                      "triangles": int(counts[:, 2].sum()), "nonempty_chunks": int((counts[:, 2] > 0).sum()),
                      "max_active_voxels": int(counts[:, 1].max()), "max_edges": int(counts[:, 0].max())},
           "per_lod": per_lod,
           "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()}, "work": work,
           "roofline": {"bound": "valu", "kernel": dom_kernel, "achieved": round(achieved, 3),
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                        "traffic": hbm_traffic(dom_kernel, wl["name"]), "flops_per_launch": int(dom_flops),
                        "flops_unit": "23 full field evaluations per edge" if dom == "edge_cross"
                        else "65^3 full field evaluations per chunk",
                        "launch_ms": round(stage_ms[dom], 4), "executed_flops_per_launch": int(exec_flops),
                        "executed_tflops": round(exec_flops / (stage_ms[dom] * 1e-3) / 1e12, 3)},
           "tuning": tuning}
    hbm = hbm_rates(stage_ms, wl["name"])
    if hbm:
        out["hbm_gbs"] = hbm
    if plain:
        # the plain per-chunk form: the same kernels, one chunk per graph launch + sync
        pb = mesher.Batch(64, 1, edge_cap=edge_cap, voxel_cap=voxel_cap, flags=wl["flags"])
        m = n if args.plain_chunks <= 0 else min(n, args.plain_chunks)
        idx = np.linspace(0, n - 1, m).astype(int)
        pb.set_chunks(wl["mins"][:1], wl["sizes"][:1])
        pb.run()
        pb.wait()
        lat = []
        t0 = time.perf_counter()
        for i in idx:
            ts = time.perf_counter()
            pb.set_chunks(wl["mins"][i:i + 1], wl["sizes"][i:i + 1])
            pb.run()
            pb.wait()
            lat.append(time.perf_counter() - ts)
        plain_rate = m / (time.perf_counter() - t0)
        pb.close()
        lat = np.array(lat) * 1e3
        out["plain"] = {"chunks": int(m), "chunks_per_s": round(plain_rate, 2),
                        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                       "p99": round(float(np.percentile(lat, 99)), 3)}}
        out["batched_over_plain"] = round(rate / plain_rate, 2)
    if keep:
        return out, batch
    batch.close()
    return out, None


def simplify_leg(batch, wl, args):
    """Post-mesh simplification (ngMeshSimplifier, as ConstructClipmapNodeData runs it on every
    chunk mesh with leven's options scaled by voxel size) over the meshed batch: one
    lvn_batch_simplify call per pass, each pass after a fresh meshing pass; then the
    pipeline and the simplifier together, and the oracle on a sample of the meshes."""
    opts = abi.simplify_options_per_voxel()
    n = len(wl["sizes"])
    batch.run()
    batch.wait()
    counts = batch.counts()
    idx = np.linspace(0, n - 1, min(n, args.cpu_sample)).astype(int)
    sample = [batch.get_mesh(i)[:2] for i in idx]
    res = batch.simplify(opts)  # warm-up (buffers grow here)
    times, both = [], []
    steps = max(2, args.steps // 2)
    for _ in range(steps):
        t0 = time.perf_counter()
        batch.run()
        batch.wait()
        t1 = time.perf_counter()
        res = batch.simplify(opts)
        t2 = time.perf_counter()
        times.append(t2 - t1)
        both.append(t2 - t0)
    ms = float(np.median(times)) * 1e3
    out = {"chunks": n, "steps": steps, "ms_per_pass": round(ms, 3), "chunks_per_s": round(n / (ms * 1e-3), 1),
           "mesh_and_simplify_chunks_per_s": round(n / float(np.median(both)), 1),
           "triangles_in": int(counts[:, 2].sum()), "triangles_out": int(res[:, 1].sum()),
           "vertices_out": int(res[:, 0].sum()), "collapses": int(res[:, 3].sum()),
           "iterations_max": int(res[:, 2].max()),
           "options": "leven's: max error 5 and max edge 2.5 voxel sizes, min cos 0.7, 1/8 of the edges per "
                      "iteration, 10 iterations, target 5 %, seed 42"}
    if not args.no_cpu:
        from oracle import py_oracle
        secs = []
        for i, (v, t) in zip(idx, sample):
            off = np.append(wl["mins"][i] + np.float32(wl["sizes"][i] / 2.0), 0).astype(np.float32)
            t0 = time.perf_counter()
            py_oracle.simplify_mesh(v, t, off, abi.simplify_options(int(wl["sizes"][i])))
            secs.append(time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": round(len(idx) / sum(secs), 2), "unit": "chunks/s", "cores": 1,
                               "kind": "port", "sample": f"the meshes of {len(idx)} of the {n} chunks, evenly "
                                                         f"spaced, simplified by the oracle on one thread"}
        out["gpu_over_cpu_single_thread"] = round(out["chunks_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def cpu_sample(wl, count, threads):
    """The oracle on `count` chunks of the workload spread evenly over its chunk list (in
    the clipmap: every LOD in proportion), `threads` chunks at a time, and the first 8 of
    them on one thread."""
    from oracle import py_oracle
    perm = py_oracle.perm_table(wl["seed"])
    n = len(wl["sizes"])
    idx = np.linspace(0, n - 1, min(n, count)).astype(int)
    wall, secs, _ = py_oracle.mesh_chunks(wl["desc"], perm, 64, wl["mins"][idx], wl["sizes"][idx],
                                          wl["default_material"], wl["ops"], wl["aabbs"], threads=threads)
    one = idx[:8]
    wall1, _, _ = py_oracle.mesh_chunks(wl["desc"], perm, 64, wl["mins"][one], wl["sizes"][one],
                                        wl["default_material"], wl["ops"], wl["aabbs"], threads=1)
    return dict(value=round(len(idx) / wall, 2), unit="chunks/s", cores=threads, kind="port",
                single_thread_chunks_per_s=round(len(one) / wall1, 3),
                sample=f"{len(idx)} of the {n} {wl['name']} chunks, evenly spaced over the chunk list, on "
                       f"{threads} threads ({os.cpu_count()} logical CPUs visible, {_cpu_model()}); oracle = "
                       f"C++ restatement of leven's semantics, FP32, one chunk per thread (no CPU OpenCL device "
                       f"exists to run leven itself): {wall:.2f} s wall, p50 {np.percentile(secs, 50) * 1e3:.1f} "
                       f"ms, p99 {np.percentile(secs, 99) * 1e3:.1f} ms per chunk; single thread "
                       f"{len(one)} chunks in {wall1:.2f} s")


def plain_cpu_leg(args, threads):
    """The one-chunk heightfield workload (64^3, 6-octave 2D fBm): the plain per-chunk GPU
    form's latency, and the oracle on one thread and on `threads` threads (that many copies
    of the chunk, one per thread)."""
    wl = workloads.plain_cpu()
    stage_ops(wl)
    pb = mesher.Batch(64, 1, edge_cap=0, voxel_cap=0, flags=wl["flags"])
    pb.set_chunks(wl["mins"], wl["sizes"])
    lat = []
    for k in range(60):
        ts = time.perf_counter()
        pb.run()
        pb.wait()
        if k >= 10:
            lat.append(time.perf_counter() - ts)
    counts = pb.counts()
    pb.close()
    lat = np.array(lat) * 1e3
    out = {"chunks": 1, "active_voxels": int(counts[0, 1]), "triangles": int(counts[0, 2]),
           "gpu_plain_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                    "p99": round(float(np.percentile(lat, 99)), 3)}}
    if not args.no_cpu:
        from oracle import py_oracle
        perm = py_oracle.perm_table(wl["seed"])
        reps = 8
        m1 = np.repeat(wl["mins"], reps, axis=0)
        s1 = np.repeat(wl["sizes"], reps)
        wall1, secs, _ = py_oracle.mesh_chunks(wl["desc"], perm, 64, m1, s1, 0, None, None, threads=1)
        mt = np.repeat(wl["mins"], threads, axis=0)
        st = np.repeat(wl["sizes"], threads)
        wallt, _, _ = py_oracle.mesh_chunks(wl["desc"], perm, 64, mt, st, 0, None, None, threads=threads)
        out["cpu_baseline"] = {"value": round(threads / wallt, 2), "unit": "chunks/s", "cores": threads,
                               "kind": "port", "single_thread_chunks_per_s": round(reps / wall1, 2),
                               "single_thread_latency_ms": round(float(np.median(secs)) * 1e3, 2),
                               "sample": f"the chunk meshed {reps} times on one thread and {threads} times on "
                                         f"{threads} threads (one copy per thread)"}
        out["gpu_over_cpu_single_thread"] = round(out["cpu_baseline"]["single_thread_latency_ms"] /
                                                  out["gpu_plain_latency_ms"]["p50"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clipmap_leven", choices=["clipmap_leven", "clipmap", "terrain_512"],
                    help="the headline workload (the others are reported beside it)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--cpu-sample", type=int, default=64, help="cave clipmap chunks in the CPU baseline sample")
    ap.add_argument("--cpu-sample-leven", type=int, default=4096,
                    help="chunks of the leven-terrain clipmap in its CPU baseline sample (capped at all of them)")
    ap.add_argument("--plain-chunks", type=int, default=0, help="chunks for the plain form (0 = all)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="length of the served request stream (0 = skip)")
    ap.add_argument("--only", default="", help="comma-separated legs to run (clipmap_leven, clipmap, terrain_512, "
                                                 "in_range, "
                                                 "precision, seams, simplify, serving, plain_cpu); empty = all")
    ap.add_argument("--no-seams", action="store_true", help="skip seam stitching of the clipmap")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-form batch")
    ap.add_argument("--no-autotune", action="store_true", help="time the default launch shapes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the build is single-GPU); got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)
    only = set(v for v in args.only.split(",") if v)

    def want(leg):
        return not only or leg in only

    threads = max(1, min(16, os.cpu_count() or 1))
    legs = {}
    # ---- the 6-LOD clipmap over the reference's own terrain field (DensityFunc, cl/noise.cl:225-266):
    # ~1,900 chunks at voxel sizes 1..32 with 0 to ~10k active voxels each, 256 brushes on the surface
    lev = workloads.clipmap_leven()
    if want("clipmap_leven"):
        legs["clipmap_leven"], _ = batched_leg(lev, args, args.steps, 0, 0, plain=True)
        if not args.no_cpu:
            legs["clipmap_leven"]["cpu_baseline"] = cpu_sample(lev, args.cpu_sample_leven, threads)
            legs["clipmap_leven"]["batched_over_cpu"] = round(legs["clipmap_leven"]["chunks_per_s"] /
                                                              legs["clipmap_leven"]["cpu_baseline"]["value"], 1)
    # ---- the 6-LOD clipmap: ~1,900 chunks at voxel sizes 1..32, full per-chunk capacities
    clip = workloads.clipmap()
    cb, nodes = None, None
    if want("clipmap") or want("seams") or want("in_range") or want("simplify"):
        legs["clipmap"], cb = batched_leg(clip, args, max(2, args.steps // 2), 0, 0, plain=want("clipmap"),
                                          keep=True)
        counts = cb.counts()
        if want("in_range"):
            # the clipmap's chunks within BASELINE's stated range of active voxels (0 to ~60k):
            # the coarse cave chunks above it come from the 1/64-frequency cave noise sampled
            # every 16-32 voxels, which is nearly white noise at those voxel sizes
            keep = counts[:, 1] <= 60000
            sub = dict(clip, name="clipmap_in_range", mins=clip["mins"][keep], sizes=clip["sizes"][keep])
            ecap = int(-(-int(counts[keep, 0].max()) // 4096) * 4096)  # slots sized for these chunks
            legs["clipmap"]["in_range"], _ = batched_leg(sub, args, max(2, args.steps // 2), ecap, 60000)
            legs["clipmap"]["in_range"]["active_voxels_max"] = 60000
        if want("simplify"):
            legs["simplify"] = simplify_leg(cb, clip, args)
        if want("seams") and not args.no_seams:
            n = len(clip["sizes"])
            nodes = [(tuple(int(v) for v in clip["mins"][i]), int(clip["sizes"][i]), cb.get_mesh(i)[2])
                     for i in range(n) if counts[i, 2] > 0 or counts[i, 3] > 0]
            stage_ops(clip)
            legs["seams"] = seam_leg(cb, nodes, args)
        cb.close()
        if want("clipmap") and not args.no_cpu:
            legs["clipmap"]["cpu_baseline"] = cpu_sample(clip, args.cpu_sample, threads)
            legs["clipmap"]["batched_over_cpu"] = round(legs["clipmap"]["chunks_per_s"] /
                                                        legs["clipmap"]["cpu_baseline"]["value"], 1)
    # ---- the 512-chunk terrain with 256 brushes (plain per-chunk configuration)
    if want("terrain_512"):
        t5 = workloads.terrain_512()
        legs["terrain_512"], _ = batched_leg(t5, args, args.steps, 262144, 131072)
        stage_ops(t5)
        legs["terrain_512"]["dropin_generate_chunk_mesh"] = dropin_latency(t5, 128)
        if not args.no_cpu:
            legs["terrain_512"]["cpu_baseline"] = cpu_baseline(t5, threads)
            legs["terrain_512"]["batched_over_cpu"] = round(legs["terrain_512"]["chunks_per_s"] /
                                                            legs["terrain_512"]["cpu_baseline"]["value"], 1)
    if want("plain_cpu"):
        legs["plain_cpu"] = plain_cpu_leg(args, threads)
    if want("precision") and not args.no_precision:
        legs["precision"] = precision_leg(args)
    if want("serving") and args.serve_seconds > 0:
        legs["serving"] = serve(args)

    head = legs.get(args.workload)
    if head is None:  # a partial run (--only) without the headline leg
        print(json.dumps({"metric": METRIC, "partial": True, **legs}), flush=True)
        return
    hw = {"clipmap_leven": lev, "clipmap": clip}.get(args.workload) or workloads.terrain_512()
    roof = dict(head["roofline"])
    roof["note"] = ("FP32 vector-ALU bound (neither HBM nor MFMA): achieved counts the method's field "
                    "evaluations (23 per edge, 65^3 per chunk), executed_tflops what the kernels evaluate after "
                    "the cave cut and shared heights; counters in profiles/kernels_" +
                    {"clipmap_leven": "clipmap_leven", "clipmap": "clipmap"}.get(args.workload, "terrain512") + ".md")
    fields = {"clipmap_leven": "y - 900 Terrain(x, z) (the reference's DensityFunc: ridged 7 + fBm 4 + fBm 2 "
                               "octaves of snoise2)",
              "clipmap": "max(y - 24 fbm2, fbm3 - 0.3), 5+5 octaves",
              "terrain_512": "max(y - 24 fbm2, fbm3 - 0.3), 5+5 octaves"}
    out = {
        "metric": METRIC,
        "value": head["chunks_per_s"],
        "unit": "chunks/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_pass"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": hw["name"], "chunks": len(hw["sizes"]), "voxels_per_chunk": 64,
                   "lods": "voxel sizes 1..32 (6-LOD clipmap selection around a seeded camera)"
                   if args.workload != "terrain_512" else "LOD0",
                   "field": fields[args.workload], "csg_ops": int(len(hw["ops"])),
                   "form": "batched (one launch per stage over all chunks, hipGraph)"},
        "plain_chunks_per_s": head.get("plain", {}).get("chunks_per_s"),
        "batched_over_plain": head.get("batched_over_plain"),
        "batched_over_cpu": head.get("batched_over_cpu"),
        "roofline": roof,
        "cpu_baseline": head.get("cpu_baseline"),
    }
    if "serving" in legs:
        out["p99_chunk_latency_ms"] = legs["serving"]["latency_ms"]["p99"]
    for k, v in legs.items():
        out[k] = v
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
