#!/usr/bin/env python3
"""Benchmark: bounce-rays/s of the stochastic energy path tracer on one MI355X.

Workload (the plain single-scene launch): a seeded ~50k-triangle city block
of 40 extruded buildings with door and window holes; 16 sources x 4096
rays x 3 bands (half the sources indoors), each ray scored against 8
listeners (1 ms x 2 s echograms, 64 bounces, relative floor 1e-7); plus
the 8 listeners' 512-ray Fibonacci dome fans against a 500 m sky sphere.
One step = one tracer launch + one dome launch over that batch, inputs and
outputs resident in HBM.  `value` counts tracer bounce-rays (iterations of
the reference's bounce loop, crates/omg-core/src/tracer.rs:103) plus dome
ray-events (escape events and straight-through segments).

Prints ONE JSON line.  --gpus above 1 is refused: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

# The serving leg keeps 8 batches in flight, each on a tracer stream and a
# dome stream, besides the context's and torch's streams; HIP maps streams
# onto GPU_MAX_HW_QUEUES hardware queues (4 by default) and streams that
# share a queue serialize.  It must be set before HIP starts (and the
# environment of the GPU hosts sets it to 4, so it is overridden here).
os.environ["GPU_MAX_HW_QUEUES"] = "24"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "omg-audio_amd"))

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
# HBM bytes per plain launch of the city-block workload (k_trace for the
# first 4 bounces + k_trace_resume16 for the rays still alive), measured
# with rocprofv3 PMC: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the scales a
# 1 GiB copy kernel (--copy-probe) reads in the same runs
# (profiles/k_trace_resume16_city_block_50k.md)
PLAIN_TRAFFIC_BYTES = 53.04e6
# the same for the k_trace_batch_regen_il launch of the mixed 256-query
# workload in the shape (1,12,6,2,4) (profiles/k_trace_batch_regen_il_mixed_256q.md)
BATCH_TRAFFIC_BYTES = 29.60e9
# the town: the routed plain launch (k_trace_batch_regen_il, one-query
# batch alike) and the single-job kernel (profiles/k_trace_town_2m.md)
STRESS_TRAFFIC_BYTES = 156.5e9
STRESS_SINGLE_TRAFFIC_BYTES = 32.6e9
# VALU issue of the same launches (SQ_INSTS_VALU per launch, and the mean
# active lanes per VALU instruction, SQ_THREAD_CYCLES_VALU /
# SQ_ACTIVE_INST_VALU / 64), from the same profiles
PLAIN_VALU_INSTS = 6.532e8
PLAIN_VALU_THREAD_UTIL = 0.463
BATCH_VALU_INSTS = 6.554e10
BATCH_VALU_THREAD_UTIL = 0.381
STRESS_VALU_INSTS = 5.345e11
STRESS_VALU_THREAD_UTIL = 0.388
STRESS_SINGLE_VALU_INSTS = 1.042e12
STRESS_SINGLE_VALU_THREAD_UTIL = 0.188
# Peak VALU issue: 128 FP32 lanes per CU per clock (the published 157.3
# TFLOP/s FP32 vector peak over 256 CUs at 2.4 GHz), i.e. two wave64 VALU
# instructions per CU per clock
VALU_PEAK_INSTS = 256 * 2 * 2.4e9
NODE_BYTES = 32           # bytes per AABB test (one child box, crates/omg-gpu/src/layout.rs:96-107)
PRIM_BYTES = 48           # bytes per triangle test (layout.rs:109-122)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-shoebox", action="store_true",
                    help="skip the shoebox room (trace and order-3 image sources vs one host core)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain workload")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-stress", action="store_true", help="skip the 2M-triangle town")
    ap.add_argument("--no-late", action="store_true", help="skip the late-field state workload")
    ap.add_argument("--no-paths", action="store_true", help="skip the knife-edge path search")
    ap.add_argument("--no-ism", action="store_true", help="skip the image-source discovery and solve")
    ap.add_argument("--build-only", action="store_true",
                    help="only the town's scene build, on the host and on the device")
    ap.add_argument("--batch-shape", default="",
                    help="lanes,stack_lds,waves_per_simd,regen,width for the batched launches "
                         "instead of the autotuner's choice (reproducible profiles)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run a plain elementwise copy of 1 GiB "
                         "(read 1 GiB, write 1 GiB): the calibration kernel of HBM "
                         "byte counters")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="minimum wall time of the CPU baseline sample")
    return ap.parse_args()


def fp32_error_report(mb, src, lis, n_rays, seed, gpu_bins):
    """The oracle's f64 re-accumulation of the same rays (counter RNG) and the
    f32 error of the GPU echograms against it, per (source, listener)
    echogram: max over bins of |gpu - f64| / f64 (bins with energy)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import numpy as np
    soup = mb.soup()
    pos, idx = O.dedup_tris(soup["tri_verts"])
    mats = [(list(m.absorption), m.scattering, list(m.transmission)) for m in soup["materials"]]
    om = O.Mesh(pos, idx, soup["tri_material"], mats, 3, soup["tri_surface"], soup["tri_max_edge"])
    g = O.mesh_geom(om)
    per = []
    for si in range(src.shape[0]):
        p = O.trace_params(n_rays, rng_mode=O.RNG_COUNTER, seed=seed, src_index=si, bin_dt=0.001,
                           max_time=2.0)
        o = O.trace(g, src[si], lis, p, want_f64=True)
        for li in range(lis.shape[0]):
            ref = o["bins64"][li]
            nz = ref > 0
            a = gpu_bins[si, li].astype(np.float64)
            bad = bool(np.any(a[~nz] != 0))
            per.append(float("inf") if bad else
                       float(np.max(np.abs(a[nz] - ref[nz]) / ref[nz])) if nz.any() else 0.0)
    per = np.array(per)
    return {"echograms": int(per.size), "max_rel": float(per.max()),
            "median_rel": float(np.median(per)), "tolerance": 1e-4,
            "per_echogram_max_rel": [float("%.3g" % v) for v in per]}


CPU_SHARE = 16   # host threads of one GPU's share of the box


def _time_oracle_jobs(O, g, src, lis, params, threads, min_seconds):
    """Repeats the jobs on a `threads`-thread pool for at least min_seconds."""
    reps, br, t0 = 0, 0, time.perf_counter()
    while True:
        _, stats = O.trace_jobs_threaded(g, src, lis, params, threads)
        br += int(stats[:, 2].sum())
        reps += 1
        if time.perf_counter() - t0 >= min_seconds:
            break
    return br, reps, time.perf_counter() - t0


def _hbm_use(alg_bytes, traffic_bytes, seconds):
    """A launch against the HBM roofline: its algorithmic bytes (the
    reference traversal's node and primitive bytes, most of which the
    caches serve) and its measured HBM traffic (PMC, from the kernel's
    profile), each per second as a fraction of the HBM peak."""
    return {"alg_bytes_per_launch": alg_bytes,
            "alg_bytes_frac_of_hbm_peak": alg_bytes / seconds / 1e9 / HBM_PEAK_GBS,
            "traffic_bytes_per_launch": traffic_bytes,
            "traffic_GBs": traffic_bytes / seconds / 1e9,
            "traffic_frac_of_hbm_peak": traffic_bytes / seconds / 1e9 / HBM_PEAK_GBS}


def _valu_roofline(insts, seconds, bounce_rays, thread_util, kernel):
    """VALU issue of a launch against the chip's peak: wave64 instructions
    per second (instruction count from the kernel's SQ profile, time
    measured here), instructions per bounce-ray at wave and lane level."""
    achieved = insts / seconds
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_INSTS,
            "unit": "wave64 VALU instructions/s", "frac": achieved / VALU_PEAK_INSTS,
            "insts_per_launch": insts, "wave_insts_per_bounce_ray": insts / bounce_rays,
            "lane_insts_per_bounce_ray": insts * 64 * thread_util / bounce_rays,
            "thread_utilization": thread_util, "kernel": kernel}


def cpu_baseline(mb, src, lis, n_rays, seed, max_seconds):
    """The oracle (C restatement of the reference tracer) on the host cores,
    same workload.  The headline figure: one job per source on a pool of
    CPU_SHARE threads (one GPU's share of the host), counter RNG.  Beside
    it: one thread, and the reference's own sequential xorshift RNG, at
    one thread and CPU_SHARE threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = O.mesh_geom(_oracle_mesh(mb))
    variants = []
    head = None
    for mode, mode_name in ((O.RNG_COUNTER, "counter"), (O.RNG_XORSHIFT, "xorshift")):
        params = [O.trace_params(n_rays, rng_mode=mode, seed=seed, src_index=i, bin_dt=0.001,
                                 max_time=2.0) for i in range(src.shape[0])]
        for threads in (CPU_SHARE, 1):
            secs = max_seconds if (mode == O.RNG_COUNTER and threads == CPU_SHARE) else 1.0
            # one thread: a bounded sample (the first 4 sources' jobs)
            jobs = params if threads > 1 else params[:4]
            br, reps, dt = _time_oracle_jobs(O, g, src, lis, jobs, threads, secs)
            v = dict(value=br / dt, unit="bounce-rays/s", cores=threads, kind="port", rng=mode_name,
                     sample=f"{len(jobs)} source jobs x {n_rays} rays x {lis.shape[0]} listeners, "
                            f"repeated {reps}x in {dt:.1f} s on a {threads}-thread pool")
            if head is None:
                head = dict(v)
            else:
                variants.append(v)
    head["variants"] = variants
    head["host_cpus"] = os.cpu_count()
    head["note"] = (f"{CPU_SHARE} threads = one GPU's share of the host's cores; "
                    "os.cpu_count() reports the machine's")
    return head


def shoebox_leg(ctx, dev, reps=50, cpu_seconds=3.0, want_cpu=True):
    """The shoebox config: a 12 x 8 x 4 m room as 12 triangles, per-surface
    absorption, one source and one listener, 4096 rays x 3 bands, 64
    bounces, floor 1e-6, 1 ms x 2 s, plus order-3 image sources on the BVH
    (4096-ray fan discovery over four rotations, then every chain solved).
    The tracer on the GPU as the drop-in is called (host buffers in and
    out) and device-resident, beside the oracle on one host thread."""
    import numpy as np
    import torch
    import omg_hip as H
    from omg_hip import scenes
    mb, src, lis, walls = scenes.shoebox(seed=1)
    scene = H.Scene.from_builder(ctx, mb)
    params = H.trace_params(4096, 3, max_bounces=64, seed=0xC1, bin_dt=0.001, max_time=2.0,
                            floor_rel=1e-6)
    counted = H.trace_world(scene, params, [src], [lis], counters=True)
    br = counted.counters["bounce_rays"]
    ts = []
    for k in range(reps + 2):
        t0 = time.perf_counter()
        H.trace_world(scene, params, [src], [lis])
        if k >= 2:
            ts.append(time.perf_counter() - t0)
    host_s = float(np.median(ts))
    d_src = torch.from_numpy(np.asarray([src], np.float32)).to(dev)
    d_lis = torch.from_numpy(np.asarray([lis], np.float32)).to(dev)
    nb = H.nbins(2.0, 0.001)
    d_bins = torch.empty((1, 1, nb, 3), dtype=torch.float32, device=dev)
    d_dirs = torch.empty((1, 1, nb, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    evs = []
    for k in range(reps + 2):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        H.trace_world_async(scene, params, d_src.data_ptr(), 1, d_lis.data_ptr(), 1,
                            d_bins.data_ptr(), d_dirs.data_ptr(), stream=stream.cuda_stream)
        e1.record(stream)
        if k >= 2:
            evs.append((e0, e1))
    torch.cuda.synchronize()
    dev_s = float(np.median([a.elapsed_time(b) for a, b in evs])) * 1e-3
    disc = H.GpuWorldDiscovery(scene, n_rays=4096)
    solver = H.GpuWorldSolve(scene)

    def ism():
        seen, chains, orders = set(), [], []
        for rot in range(4):
            ch, od = disc.discover(lis, rot)
            for c, o in zip(ch, od):
                key = tuple(int(x) for x in c[:o])
                if key not in seen:
                    seen.add(key)
                    chains.append(c)
                    orders.append(o)
        ch = np.vstack([np.full((1, 3), 0xFFFF, np.uint16), np.array(chains, np.uint16)])
        od = np.concatenate([[0], orders]).astype(np.uint8)
        return ch, od, solver.solve_batch([src], ch, od, lis)

    ism()
    ti = []
    for _ in range(10):
        t0 = time.perf_counter()
        ch, od, rec = ism()
        ti.append(time.perf_counter() - t0)
    ism_s = float(np.median(ti))
    leg = {"workload": "shoebox 12 x 8 x 4 m (12 tris), 1 source x 4096 rays x 3 bands, 64 bounces, "
                       "floor 1e-6, 1 ms x 2 s; order-3 image sources (4 x 4096-ray fan + solve)",
           "bounce_rays": br,
           "gpu_trace_ms": host_s * 1e3, "gpu_trace_bounce_rays_per_s": br / host_s,
           "gpu_trace_device_resident_ms": dev_s * 1e3,
           "gpu_trace_device_resident_bounce_rays_per_s": br / dev_s,
           "gpu_ism_ms": ism_s * 1e3, "ism_chains": int(len(od) - 1),
           "ism_valid_paths": int(rec["valid"].sum())}
    if want_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        om = _oracle_mesh(mb)
        g = O.mesh_geom(om)
        op = O.trace_params(4096, rng_mode=O.RNG_COUNTER, seed=0xC1, max_bounces=64,
                            bin_dt=0.001, max_time=2.0, floor_rel=1e-6)
        n, obr, t0 = 0, 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds or n == 0:
            o = O.trace(g, src, [lis], op)
            obr += int(o["stats"][2])
            n += 1
        cpu_s = (time.perf_counter() - t0) / n
        tab, n_surf = O.surface_table(om)
        t0 = time.perf_counter()
        ochains = set()
        for rot in range(4):
            c, o = O.mesh_chains(om, lis, 4096, rot)
            ochains |= {tuple(int(x) for x in cc[:oo]) for cc, oo in zip(c, o)}
        for key in [()] + sorted(ochains):
            O.mesh_record(om, tab, n_surf, list(key), len(key), src, lis)
        cpu_ism_s = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": obr / n / cpu_s, "unit": "bounce-rays/s", "cores": 1,
                               "kind": "port", "trace_ms": cpu_s * 1e3, "ism_ms": cpu_ism_s * 1e3,
                               "sample": f"the same trace {n}x and the image sources once, "
                                         "on one host thread"}
        leg["speedup_trace_vs_cpu_1core"] = cpu_s / host_s
        leg["speedup_ism_vs_cpu_1core"] = cpu_ism_s / ism_s
    scene.close()
    return leg


def _oracle_mesh(mb):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    soup = mb.soup()
    pos, idx = O.dedup_tris(soup["tri_verts"])
    mats = [(list(m.absorption), m.scattering, list(m.transmission)) for m in soup["materials"]]
    return O.Mesh(pos, idx, soup["tri_material"], mats, 3, soup["tri_surface"], soup["tri_max_edge"])


def path_search(ctx, dev, reps=5, cpu_seconds=4.0, want_cpu=True):
    """Knife-edge path search (AutoPaths::find, crates/omg-core/src/paths.rs:117-332)
    on two workloads: the reference's mesh benchmark (2.5k triangles, 64
    edges, default budget {24, 8, 4}, its 1000 queries;
    crates/omg-app/examples/mesh_bench.rs:81-92) and the city block with the
    world's graph and budget (512 edges, {48, 32, 4}; crates/omg-scene/src/
    world.rs:154, :638-642) over 4096 seeded (source, listener) pairs.  One
    step = one omg_paths_find_async call over every query (direct pass and
    bent pass), inputs and outputs resident in HBM, timed with HIP events on
    its stream.  The oracle runs the same queries on one host thread for a
    bounded sample and its paths are compared with the GPU's bit for bit."""
    import numpy as np
    import torch
    import omg_hip as H
    from omg_hip import scenes
    from omg_hip.paths import PATH_DTYPE, PathBudget
    mb2, bounds, boxes = scenes.city_block(seed=2)
    s2, l2 = scenes.path_queries(7, 4096, bounds, boxes)
    cases = [("mesh_bench", *scenes.mesh_bench(1000), 64, PathBudget()),
             ("city_block_world", mb2, s2, l2, 512, PathBudget(48, 32, 4))]
    legs = {}
    for name, mb, src, lis, ebud, budget in cases:
        scene = H.Scene.from_builder(ctx, mb)
        ap = H.AutoPaths(scene, ebud)
        n = src.shape[0]
        stride = 1 + budget.max_paths
        ap.set_counting(True)
        ref_raw, ref_cnt = ap.find_many(src, lis, budget, raw=True)
        cnt = ap.counters()
        ap.set_counting(False)
        d_src = torch.from_numpy(src).to(dev)
        d_lis = torch.from_numpy(lis).to(dev)
        d_out = torch.zeros(n * stride * PATH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        d_n = torch.zeros(n, dtype=torch.int32, device=dev)
        d_work = torch.zeros(n, dtype=torch.int32, device=dev)
        stream = torch.cuda.Stream(device=dev)

        def run():
            ap.find_async(d_src.data_ptr(), d_lis.data_ptr(), n, d_out.data_ptr(), d_n.data_ptr(),
                          d_work.data_ptr(), budget, stream.cuda_stream)

        for _ in range(2):
            run()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        for e in evs:
            e[0].record(stream)
            run()
            e[1].record(stream)
        torch.cuda.synchronize()
        ap.status()
        ms = float(np.median([e[0].elapsed_time(e[1]) for e in evs]))
        got = d_out.cpu().numpy().view(PATH_DTYPE).reshape(n, stride)
        got_n = d_n.cpu().numpy().astype(np.uint32)
        same = bool(np.array_equal(got_n, ref_cnt)) and all(
            np.array_equal(got[q, :got_n[q]], ref_raw[q, :got_n[q]]) for q in range(n))
        blocked = int(np.sum(ref_raw[:, 0]["gains"][:, 1] < 0.25))
        alg = NODE_BYTES * cnt["aabb"] + PRIM_BYTES * cnt["tri"]
        leg = {"queries": n, "tris": scene.info()["n_tris"], "edges": ap.edge_count(),
               "budget": [budget.edge_candidates, budget.pair_edges, budget.max_paths],
               "blocked": blocked, "paths_returned": int(ref_cnt.sum()),
               "ms_per_launch": ms, "queries_per_s": n / (ms * 1e-3),
               "us_per_query": ms * 1e3 / n,
               "aabb_tests": cnt["aabb"], "tri_tests": cnt["tri"],
               "alg_bytes_GBs": alg / (ms * 1e-3) / 1e9, "timed_equals_counted": same}
        if want_cpu:
            om = _oracle_mesh(mb)
            edges = om.path_edges(ebud)
            kw = dict(edge_candidates=budget.edge_candidates, pair_edges=budget.pair_edges,
                      max_paths=budget.max_paths, taps=budget.taps)
            k, bad, t0 = 0, 0, time.perf_counter()
            while k < n and (time.perf_counter() - t0 < cpu_seconds or k < 16):
                ref = om.find_paths(edges, src[k], lis[k], **kw)
                g = got[k, :got_n[k]]
                bad += int(len(ref) != len(g) or any(
                    r["key"] != int(x["key"]) or not np.array_equal(r["gains"], x["gains"]) or
                    r["length"] != x["length"] or
                    not np.array_equal(r["points"], x["points"][:int(x["n_points"])])
                    for r, x in zip(ref, g)))
                k += 1
            dt = time.perf_counter() - t0
            leg["cpu_baseline"] = {"value": k / dt, "unit": "queries/s", "cores": 1, "kind": "port",
                                   "us_per_query": dt * 1e6 / k,
                                   "sample": f"the first {k} queries on one host thread"}
            leg["oracle_mismatches"] = bad
            leg["oracle_compared"] = k
        legs[name] = leg
        ap.close()
        scene.close()
    return legs


def image_sources(ctx, scene, mb, src, lis, reps=3, cpu_seconds=3.0, want_cpu=True):
    """Image sources on the BVH (WorldDiscovery / WorldSolve,
    crates/omg-scene/src/early_world.rs:20-62) on the city block: per
    listener, chain discovery by the 4096-ray rotating fan (the reference's
    GPU budget, crates/omg-gpu/src/lib.rs:220) and the solve of every
    (source, chain) pair for the 16 sources (pt_mesh.rs:185-324).  Timed on
    the host around the synchronous C calls (inputs and records cross
    PCIe).  The oracle solves a bounded sample of the same pairs on one
    host thread and its chains and records are compared with the GPU's."""
    import numpy as np
    import omg_hip as H
    disc = H.GpuWorldDiscovery(scene, n_rays=4096)
    solver = H.GpuWorldSolve(scene)
    disc_s, solve_s, pairs, valid = [], [], 0, 0
    found = []
    for rep in range(reps + 1):
        for li in range(lis.shape[0]):
            t0 = time.perf_counter()
            ch, od = disc.discover(lis[li], li)
            t1 = time.perf_counter()
            rec = solver.solve_batch(src, ch, od, lis[li])
            t2 = time.perf_counter()
            if rep:
                disc_s.append(t1 - t0)
                solve_s.append(t2 - t1)
                pairs += src.shape[0] * len(od)
                valid += int(rec["valid"].sum())
            elif li == 0:
                found = (ch, od, rec)
    leg = {"workload": f"city block, {lis.shape[0]} listeners: 4096-ray chain discovery, "
                       f"then {src.shape[0]} sources x every chain solved",
           "discover_ms_per_listener": float(np.mean(disc_s)) * 1e3,
           "solve_ms_per_listener": float(np.mean(solve_s)) * 1e3,
           "chains_per_listener": pairs / src.shape[0] / len(solve_s),
           "pairs_per_s": pairs / float(np.sum(solve_s)),
           "valid_records_per_listener": valid / len(solve_s)}
    if want_cpu:
        om = _oracle_mesh(mb)
        import oracle as O
        ch, od, rec = found
        oc, oo = O.mesh_chains(om, lis[0], 4096, 0)
        leg["chains_identical"] = bool(np.array_equal(oc, ch) and np.array_equal(oo, od))
        tab, n_surf = O.surface_table(om)
        k, bad, t0 = 0, 0, time.perf_counter()
        n = len(od)
        while k < src.shape[0] * n and (time.perf_counter() - t0 < cpu_seconds or k < 64):
            si, ci = divmod(k, n)
            r = O.mesh_record(om, tab, n_surf, ch[ci][:od[ci]], int(od[ci]), src[si], lis[0])
            if (r is not None) != bool(rec["valid"][k]) or (
                    r is not None and (r["delay"] != rec["delay"][k] or
                                       not np.array_equal(r["dir"], rec["dir"][k]) or
                                       not np.allclose(r["gains"], rec["gains"][k], rtol=2e-6))):
                bad += 1
            k += 1
        dt = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": k / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"the first {k} (source, chain) pairs of listener 0 "
                                         "on one host thread"}
        leg["oracle_mismatches"] = bad
        leg["oracle_compared"] = k
    return leg


def batched_vs_plain(ctx, launches=3, fixed_shape=None):
    """Mixed-size workload: 256 queries over 32 procedural blocks (1k-200k
    tris, 3 or 6 bands, 1-32 sources x 1024-65536 rays, 1-4 listeners,
    1 ms x 1-4 s).  The batched persistent launch (autotuned, one hipGraph
    replay per pass) against the plain path: one omg_trace_world per query,
    host buffers in and out, as a caller of the single-job entry point runs
    it.  Same queries, same seeds; bounce-rays from a counted batch pass."""
    import numpy as np
    import omg_hip as H
    from omg_hip import scenes
    specs, qs = scenes.c3_workload()
    pool = [H.Scene.from_builder(ctx, mb, nbands=nb) for mb, nb in specs]
    Q = [H.QuerySpec(pool[q["scene"]],
                     H.trace_params(q["n_rays"], nbands=q["nbands"], seed=q["seed"], bin_dt=0.001,
                                    max_time=q["max_time"]),
                     q["sources"], q["listeners"]) for q in qs]
    bc = H.Batch(ctx, Q, counters=True)
    rays, items = bc.size()
    bc.launch()
    cnt = bc.counters()
    bc.close()
    b = H.Batch(ctx, Q)
    t0 = time.perf_counter()
    if fixed_shape:
        b.set_shape(*fixed_shape)
        shape = b.shape
    else:
        shape = b.tune(stride=16)
    tune_s = time.perf_counter() - t0
    b.launch()
    b.sync()
    ts = []
    for _ in range(launches):
        t0 = time.perf_counter()
        b.launch()
        b.sync()
        ts.append(time.perf_counter() - t0)
    batched_s = float(np.mean(ts))
    # the batched result of one query reproduces its plain trace
    ref = H.trace_world(Q[0].scene, Q[0].params, Q[0].sources, Q[0].listeners)
    got = b.result(0)
    rel = float(np.max(np.abs(got.bins - ref.bins)) / max(float(np.max(np.abs(ref.bins))), 1e-30))
    t0 = time.perf_counter()
    for q in Q:
        H.trace_world(q.scene, q.params, q.sources, q.listeners)
    plain_s = time.perf_counter() - t0
    b.close()
    # the plain path device-resident: inputs and outputs already in HBM, one
    # omg_trace_world_async per query back to back on one stream
    import torch
    dev = torch.device("cuda", 0)
    bufs = []
    for q in Q:
        ns, nl = len(q.sources), len(q.listeners)
        nbq = H.nbins(q.params.max_time, q.params.bin_dt)
        bufs.append((torch.from_numpy(np.ascontiguousarray(q.sources, np.float32)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(q.listeners, np.float32)).to(dev),
                     torch.empty((ns, nl, nbq, q.params.nbands), dtype=torch.float32, device=dev),
                     torch.empty((ns, nl, nbq, 3), dtype=torch.float32, device=dev)))
    stream = torch.cuda.Stream(device=dev)

    def device_pass():
        for q, (d_s, d_l, d_b, d_d) in zip(Q, bufs):
            H.trace_world_async(q.scene, q.params, d_s.data_ptr(), d_s.shape[0], d_l.data_ptr(),
                                d_l.shape[0], d_b.data_ptr(), d_d.data_ptr(),
                                stream=stream.cuda_stream)

    device_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    device_pass()
    torch.cuda.synchronize()
    plain_dev_s = time.perf_counter() - t0
    del bufs
    alg = NODE_BYTES * cnt["aabb_tests"] + PRIM_BYTES * cnt["tri_tests"]
    for s in pool:
        s.close()
    br = cnt["bounce_rays"]
    return {
        "workload": f"{len(Q)} queries over {len(pool)} city blocks "
                    f"({min(mb.n_tris for mb, _ in specs)}-{max(mb.n_tris for mb, _ in specs)} tris), "
                    "3/6 bands, 1-32 sources x 1024-65536 rays, 1-4 listeners, 1 ms x 1-4 s",
        "rays": rays, "work_items": items, "bounce_rays": br,
        "batched": {"s": batched_s, "bounce_rays_per_s": br / batched_s,
                    "queries_per_s": len(Q) / batched_s, "launches": launches,
                    "shape": shape, "tune_s": tune_s,
                    "hbm": _hbm_use(alg, BATCH_TRAFFIC_BYTES, batched_s),
                    "valu_roofline": _valu_roofline(BATCH_VALU_INSTS, batched_s, br,
                                                    BATCH_VALU_THREAD_UTIL,
                                                    "k_trace_batch_regen_il (1,12,6,2,4)")},
        "plain": {"s": plain_s, "bounce_rays_per_s": br / plain_s,
                  "queries_per_s": len(Q) / plain_s,
                  "path": "omg_trace_world per query, host buffers in and out"},
        "plain_device_resident": {"s": plain_dev_s, "bounce_rays_per_s": br / plain_dev_s,
                                  "queries_per_s": len(Q) / plain_dev_s,
                                  "path": "omg_trace_world_async per query back to back, "
                                          "buffers in HBM"},
        "speedup": plain_s / batched_s,
        "speedup_vs_device_resident_plain": plain_dev_s / batched_s,
        "query0_rel_diff_vs_plain": rel,
    }


def serving(ctx, qps_list=(100, 250, 500, 1000, 2000), seconds=1.0):
    """Serving loop on 8 resident city blocks (5k-500k tris): Poisson
    arrivals of requests from 1000 walking listeners, each 1-32 sources x
    512-16384 rays (10 ms x 3 s echograms, 3 bands) plus the listener's
    512-ray dome fan, 64 GB HBM budget.  Each rate replays the workload's
    first `seconds` of arrivals open-loop on a server created just before
    (cold: no request has run on it; omg_server_create sizes its arenas and
    staging from the budget and starts its streams); latency runs from the
    scheduled arrival to the results on the host."""
    import numpy as np
    import omg_hip as H
    from omg_hip import scenes
    specs, reqs = scenes.c5_workload(seconds=2.0, qps=2000.0)
    pool = [H.Scene.from_builder(ctx, mb, nbands=3) for mb, _, _ in specs]
    R = [dict(scene=pool[r["scene"]], params=H.trace_params(r["n_rays"], 3, seed=r["seed"]),
              sources=r["sources"], listeners=r["listener"], dome=True) for r in reqs]
    arr0 = np.array([r["t"] for r in reqs])
    shape = dict(lanes=1, stack_lds=12, waves_per_simd=6, regen=1, width=4)
    # every batch of this workload runs the quad-per-ray regenerating walk:
    # faster than the one-lane walk up to the largest (0.5 M-ray) requests
    slots, lat_rays = 8, 1 << 22
    rows = []
    st = {}
    for qps in qps_list:
        n = min(len(R), int(qps * seconds))
        arr = arr0[:n] * (2000.0 / qps)
        t0 = time.perf_counter()
        srv = H.Server(ctx, pool, hbm_budget=64 << 30, slots=slots, max_batch=64, shape=shape,
                       latency_rays=lat_rays)
        create_ms = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        lat, qs, _ = srv.replay(R[:n], arr, queue=True)
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        rays = sum(R[i]["params"].n_rays * len(R[i]["sources"]) for i in range(n))
        rows.append({"offered_qps": qps, "requests": n, "served_qps": n / wall,
                     "rays_per_s": rays / wall,
                     "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3),
                     "max_ms": float(np.max(lat) * 1e3),
                     "queue_p99_ms": float(np.percentile(qs, 99) * 1e3),
                     "server_create_ms": create_ms})
    for s in pool:
        s.close()
    held = [r["offered_qps"] for r in rows if r["p99_ms"] < 10.0]
    return {"workload": "8 resident city blocks "
                        f"({min(mb.n_tris for mb, _, _ in specs)}-{max(mb.n_tris for mb, _, _ in specs)} tris), "
                        "Poisson arrivals, 1000 walkers at 1.4 m/s, 1-32 sources x 512-16384 rays "
                        "+ 512-ray dome per request, 64 GB budget",
            "cold": "a new server per rate, measured from its first request (no warm-up replay)",
            "shape": shape, "slots": slots, "latency_rays": lat_rays,
            "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"), "rates": rows,
            "max_qps_with_p99_under_10ms": max(held) if held else None,
            "batches_at_top_rate": st.get("batches"), "mean_batch_at_top_rate": st.get("mean_batch"),
            "peak_committed_bytes_at_top_rate": st.get("peak_committed_bytes")}


def late_field(ctx, dev, n_slots=4096, reps=5):
    """Late-field state of 4096 listener slots (2 s x 1 ms x 3 bands): one
    omg_late_field_update call folds a fresh device echogram into every
    slot (EMA), reduces every slot (f64 decay integral, late energy and
    direction) and returns the reverb parameters to the host; the timed
    region is that C call (update table built once)."""
    import ctypes as C
    import numpy as np
    import torch
    import omg_hip as H
    from omg_hip import _lib
    lf = H.LateField(ctx, n_slots, nbands=3, bin_dt=0.001, max_time=2.0)
    n = lf.nbins
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    t = (torch.arange(n, device=dev, dtype=torch.float32) + 0.5) * 1e-3
    tail = torch.exp(-t[:, None] / torch.tensor([0.1, 0.15, 0.2], device=dev))
    cur_b = (tail[None] * torch.rand((n_slots, n, 3), generator=g, device=dev)).contiguous()
    cur_d = (torch.randn((n_slots, n, 3), generator=g, device=dev) * cur_b[..., 1:2]).contiguous()
    torch.cuda.synchronize()
    ups = (_lib.LateUpdate * n_slots)()
    for s in range(n_slots):
        ups[s].slot = s
        ups[s].traced = 1
        ups[s].cur_bins = cur_b[s].data_ptr()
        ups[s].cur_dirs = cur_d[s].data_ptr()
        ups[s].receiver[0] = 3.0
    out = (_lib.LateOut * n_slots)()
    L = _lib.lib()
    _lib.check(L.omg_late_field_update(lf.h, ups, n_slots, out, None))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(L.omg_late_field_update(lf.h, ups, n_slots, out, None))
        ts.append(time.perf_counter() - t0)
    lf.close()
    dt = float(np.median(ts))
    return {"workload": f"{n_slots} slots x {n} bins x 3 bands, every slot traced each call",
            "ms_per_call": dt * 1e3, "slot_updates_per_s": n_slots / dt,
            "rt60_slot0": [float(x) for x in out[0].rt60[:3]]}


def _same_tree(a, b):
    """Device build vs host build: primitives bit for bit, node references
    exact, node boxes equal as values."""
    import numpy as np
    (pa, na, n4a, sa), (pb, nb, n4b, sb) = a, b
    if sa != sb or not np.array_equal(pa.view(np.uint32), pb.view(np.uint32)):
        return False
    for x, y in ((na, nb), (n4a, n4b)):
        if x.shape != y.shape:
            return False
        refs = [3, 7, 11, 15] if x.shape[1] == 16 else [c for c in range(32) if c % 8 in (3, 7)]
        boxes = [c for c in range(x.shape[1]) if c not in refs]
        if not (np.array_equal(x.view(np.uint32)[:, refs], y.view(np.uint32)[:, refs]) and
                np.array_equal(x[:, boxes], y[:, boxes])):
            return False
    return True


def scene_build_leg(ctx, host_scene, mesh, reps=3):
    """The same scene built on the device (omg_scene_create_device: the
    tessellation, the SAH tree and the 4-wide collapse as kernels, inputs
    from host memory), timed on the C call after one warm-up build, checked
    against the host build; then refits of the unchanged vertices."""
    import omg_hip as H
    ms = []
    dev = None
    for k in range(reps + 1):
        if dev is not None:
            dev.close()
        t0 = time.perf_counter()
        dev = H.Scene.from_mesh(ctx, nbands=6, device=True, **mesh)
        if k:
            ms.append((time.perf_counter() - t0) * 1e3)
    same = _same_tree(host_scene.export(), dev.export())
    refit = []
    for _ in range(reps):
        t0 = time.perf_counter()
        dev.refit(mesh["positions"])
        refit.append((time.perf_counter() - t0) * 1e3)
    dev.close()
    ms.sort()
    refit.sort()
    return {"ms": ms[len(ms) // 2], "ms_all": ms, "refit_ms": refit[len(refit) // 2],
            "identical_to_host_build": same}


def build_only(ctx):
    """The town's scene (about 2.07M triangles) built on the host and on
    the device; see scene_build_leg."""
    import omg_hip as H
    from omg_hip import scenes
    mb, _, _ = scenes.c4_workload()
    soup = mb.soup()
    pos, idx = H.dedup_soup(soup["tri_verts"])
    mesh = dict(positions=pos, indices=idx, tri_material=soup["tri_material"],
                materials=soup["materials"], tri_surface=soup["tri_surface"],
                tri_max_edge=soup["tri_max_edge"])
    t0 = time.perf_counter()
    host = H.Scene.from_mesh(ctx, nbands=6, **mesh)
    host_s = time.perf_counter() - t0
    info = host.info()
    leg = scene_build_leg(ctx, host, mesh)
    host.close()
    return dict(leg, workload="town scene build", tris=info["n_tris"], prims=info["n_prims"],
                bvh_nodes=info["n_nodes"], host_build_s=host_s,
                speedup_vs_host=host_s / (leg["ms"] * 1e-3))


def stress(ctx, dev, launches=2, fixed_shape=None):
    """Large-scene stress: the ~2M-triangle town (400 buildings with
    storey slabs, partitions and apertures), 64 sources x 1M rays x 6
    bands scored at 4 listeners (1 ms x 2 s), 128 bounces, floor 1e-7.  The
    plain single launch on device buffers and the same job as a one-query
    batch; bounce-rays from a counted pass (binary walk: the reference's
    traversal counts)."""
    import numpy as np
    import torch
    import omg_hip as H
    from omg_hip import scenes
    t0 = time.perf_counter()
    mb, src, lis = scenes.c4_workload()
    gen_s = time.perf_counter() - t0
    soup = mb.soup()
    t0 = time.perf_counter()
    pos, idx = H.dedup_soup(soup["tri_verts"])
    dedup_s = time.perf_counter() - t0
    mesh = dict(positions=pos, indices=idx, tri_material=soup["tri_material"],
                materials=soup["materials"], tri_surface=soup["tri_surface"],
                tri_max_edge=soup["tri_max_edge"])
    t0 = time.perf_counter()
    scene = H.Scene.from_mesh(ctx, nbands=6, **mesh)
    build_s = time.perf_counter() - t0
    info = scene.info()
    build = scene_build_leg(ctx, scene, mesh)
    n_rays = 1 << 20
    params = H.trace_params(n_rays, 6, max_bounces=128, floor_rel=1e-7, seed=0xC4, bin_dt=0.001,
                            max_time=2.0)
    counted = H.trace_world(scene, params, src, lis, counters=True)
    tc = counted.counters
    n_src, n_lis, nb = src.shape[0], lis.shape[0], H.nbins(2.0, 0.001)
    d_src = torch.from_numpy(src).to(dev)
    d_lis = torch.from_numpy(lis).to(dev)
    d_bins = torch.empty((n_src, n_lis, nb, 6), dtype=torch.float32, device=dev)
    d_dirs = torch.empty((n_src, n_lis, nb, 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)

    def plain_launches():
        ms = []
        for k in range(launches + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            H.trace_world_async(scene, params, d_src.data_ptr(), n_src, d_lis.data_ptr(), n_lis,
                                d_bins.data_ptr(), d_dirs.data_ptr(), stream=stream.cuda_stream)
            e1.record(stream)
            torch.cuda.synchronize()
            if k:
                ms.append(e0.elapsed_time(e1))
        return ms

    # the drop-in as it is (64M rays: the persistent regenerating kernel),
    # then the single-job kernel alone (routing off)
    plain_ms = plain_launches()
    got = d_bins.cpu().numpy()
    rel = float(np.max(np.abs(got - counted.bins)) / max(float(np.max(np.abs(counted.bins))), 1e-30))
    ctx.set_route_rays(0)
    single_ms = plain_launches()
    got1 = d_bins.cpu().numpy()
    rel1 = float(np.max(np.abs(got1 - counted.bins)) / max(float(np.max(np.abs(counted.bins))), 1e-30))
    ctx.set_route_rays()
    b = H.Batch(ctx, [H.QuerySpec(scene, params, src, lis)])
    if fixed_shape:
        b.set_shape(*fixed_shape)
    else:
        b.tune(stride=64)
    b.launch()
    b.sync()
    batch_s = []
    for _ in range(launches):
        t0 = time.perf_counter()
        b.launch()
        b.sync()
        batch_s.append(time.perf_counter() - t0)
    shape = b.shape
    b.close()
    scene.close()
    br = tc["bounce_rays"]
    alg = NODE_BYTES * tc["aabb_tests"] + PRIM_BYTES * tc["tri_tests"]
    pm = float(np.mean(plain_ms)) * 1e-3
    sm = float(np.mean(single_ms)) * 1e-3
    bs = float(np.mean(batch_s))
    return {
        "workload": f"town of {info['n_tris']} tris (400 buildings with interiors), "
                    f"{n_src} sources x {n_rays} rays x 6 bands at {n_lis} listeners, "
                    "128 bounces, floor 1e-7",
        "tris": info["n_tris"], "prims": info["n_prims"], "bvh_nodes": info["n_nodes"],
        "scene_device_bytes": info["device_bytes"], "generate_s": gen_s, "dedup_s": dedup_s,
        "build_s": build_s,
        "device_build": dict(build, host_build_s=build_s,
                             speedup_vs_host=build_s / (build["ms"] * 1e-3)),
        "bounce_rays": br, "aabb_tests": tc["aabb_tests"], "tri_tests": tc["tri_tests"],
        "plain": {"ms": pm * 1e3, "bounce_rays_per_s": br / pm,
                  "hbm": _hbm_use(alg, STRESS_TRAFFIC_BYTES, pm),
                  "valu_roofline": _valu_roofline(STRESS_VALU_INSTS, pm, br, STRESS_VALU_THREAD_UTIL,
                                                  "k_trace_batch_regen_il (1,12,6,2,4)"),
                  "kernel": "persistent regenerating (omg_trace_world routes jobs of >= 1M rays)"},
        "plain_single_job_kernel": {"ms": sm * 1e3, "bounce_rays_per_s": br / sm,
                                    "hbm": _hbm_use(alg, STRESS_SINGLE_TRAFFIC_BYTES, sm),
                                    "valu_roofline": _valu_roofline(
                                        STRESS_SINGLE_VALU_INSTS, sm, br,
                                        STRESS_SINGLE_VALU_THREAD_UTIL,
                                        "k_trace one lane per ray, 4-wide (routing off)"),
                                    "timed_vs_counted_rel_err": rel1},
        "batched": {"ms": bs * 1e3, "bounce_rays_per_s": br / bs, "shape": shape,
                    "hbm": _hbm_use(alg, STRESS_TRAFFIC_BYTES, bs)},
        "timed_vs_counted_rel_err": rel,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by source and by query with no communication between GPUs)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import omg_hip as H
    from omg_hip import scenes

    dev = torch.device("cuda", 0)
    ctx = H.Context(0)
    if args.build_only:
        print(json.dumps(build_only(ctx)), flush=True)
        return
    mb, src, lis = scenes.c2_workload()
    scene = H.Scene.from_builder(ctx, mb)
    info = scene.info()
    n_rays, seed = 4096, 0xC2
    params = H.trace_params(n_rays, 3, seed=seed, bin_dt=0.001, max_time=2.0)
    dparams = H.dome_params(3, n_fan=512, horizon=500.0)
    n_src, n_lis = src.shape[0], lis.shape[0]
    nbins = H.nbins(2.0, 0.001)

    # one counted pass: traversal work of exactly the timed launch
    counted = H.trace_world(scene, params, src, lis, counters=True)
    tc = counted.counters
    dome = H.dome_sample(scene, lis, dparams, rays=True, counters=True)
    dome_events = int(dome["events"].sum() + dome["through"].sum())
    alg_bytes_trace = NODE_BYTES * tc["aabb_tests"] + PRIM_BYTES * tc["tri_tests"]

    d_src = torch.from_numpy(src).to(dev)
    d_lis = torch.from_numpy(lis).to(dev)
    d_bins = torch.empty((n_src, n_lis, nbins, 3), dtype=torch.float32, device=dev)
    d_dirs = torch.empty((n_src, n_lis, nbins, 3), dtype=torch.float32, device=dev)
    d_energy = torch.empty((n_lis, 9, 3), dtype=torch.float32, device=dev)
    d_dsum = torch.empty((n_lis, 9, 3), dtype=torch.float32, device=dev)
    d_rays = torch.empty((n_lis * dparams.n_rays, 16), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    sh = stream.cuda_stream

    # the dome fans run on a second stream, overlapped with the tracer launch
    stream2 = torch.cuda.Stream(device=dev)
    sh2 = stream2.cuda_stream

    def step(ev=None):
        fork = torch.cuda.Event()
        fork.record(stream)
        stream2.wait_event(fork)
        if ev is not None:
            ev[0].record(stream)
        H.trace_world_async(scene, params, d_src.data_ptr(), n_src, d_lis.data_ptr(), n_lis,
                            d_bins.data_ptr(), d_dirs.data_ptr(), stream=sh)
        if ev is not None:
            ev[1].record(stream)
            ev[2].record(stream2)
        H.dome_sample_async(scene, d_lis.data_ptr(), n_lis, d_energy.data_ptr(),
                            d_dsum.data_ptr(), d_rays.data_ptr(), dparams, stream=sh2)
        if ev is not None:
            ev[3].record(stream2)
        join = torch.cuda.Event()
        join.record(stream2)
        stream.wait_event(join)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for k in range(args.steps):
        step(evs[k])
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    trace_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    dome_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in evs]))
    H.dome_status(scene)

    # the timed launch reproduces the counted pass (same seed, same inputs)
    got = d_bins.cpu().numpy()
    rel = float(np.max(np.abs(got - counted.bins)) / max(float(np.max(np.abs(counted.bins))), 1e-30))

    ms_per_step = total_ms / args.steps
    work = tc["bounce_rays"] + dome_events
    value = work / (ms_per_step * 1e-3)
    achieved = alg_bytes_trace / (trace_ms * 1e-3) / 1e9
    out = {
        "metric": "bounce-rays/s + queries/s and p99 latency, batched vs plain launch, 1 MI355X",
        "value": value,
        "unit": "bounce-rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural city block)",
        "config": {
            "workload": "city block ~50k tris, 16 sources x 4096 rays x 3 bands scored at 8 "
                        "listeners (1 ms x 2 s), + 8 x 512-ray dome fans, 500 m sky",
            "tris": info["n_tris"], "prims": info["n_prims"], "bvh_nodes": info["n_nodes"],
            "sources": n_src, "rays_per_source": n_rays, "listeners": n_lis, "bands": 3,
            "parallelism": "single GPU",
        },
        "tracer": {"bounce_rays_per_step": tc["bounce_rays"], "ms": trace_ms,
                   "bounce_rays_per_s": tc["bounce_rays"] / (trace_ms * 1e-3),
                   "aabb_tests": tc["aabb_tests"], "tri_tests": tc["tri_tests"]},
        "dome": {"ray_events_per_step": dome_events, "ms": dome_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": PLAIN_TRAFFIC_BYTES,
                     "kernel": "k_trace (first 4 bounces) + k_trace_resume16 (the rest), "
                               "one plain launch",
                     "alg_bytes_per_launch": alg_bytes_trace},
        "valu_roofline": _valu_roofline(PLAIN_VALU_INSTS, trace_ms * 1e-3, tc["bounce_rays"],
                                        PLAIN_VALU_THREAD_UTIL,
                                        "k_trace + k_trace_resume16, one plain launch"),
        "wall_s": wall,
        "timed_vs_counted_rel_err": rel,
    }
    if args.copy_probe:
        # 1 GiB each way: four times the 256 MB Infinity Cache, so that the
        # copy's reads come from HBM whatever the cache holds
        x = torch.ones(256 << 20, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(2):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y
    fixed = [int(x) for x in args.batch_shape.split(",")] if args.batch_shape else None
    if not args.no_shoebox:
        out["shoebox"] = shoebox_leg(ctx, dev, want_cpu=not args.no_cpu_baseline)
    if not args.no_batch:
        out["batched_vs_plain"] = batched_vs_plain(ctx, fixed_shape=fixed)
    if not args.no_serve:
        out["serving"] = serving(ctx)
    if not args.no_stress:
        out["stress"] = stress(ctx, dev, fixed_shape=fixed)
    if not args.no_late:
        out["late_field"] = late_field(ctx, dev)
    if not args.no_paths:
        out["path_search"] = path_search(ctx, dev, want_cpu=not args.no_cpu_baseline)
    if not args.no_ism:
        out["image_sources"] = image_sources(ctx, scene, mb, src, lis,
                                             want_cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(mb, src, lis, n_rays, seed, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["fp32_error_vs_f64"] = fp32_error_report(mb, src, lis, n_rays, seed, got)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
