"""Benchmark: voxel-iterations per second of the gradient solve at 128^3.

Workload (one step): a complete gradient solve of the 128x128x128 volume,
spacing (1, 1, 1.5), target exp(smoothed Gaussian) clipped to [0.8, 1.25],
mask = union of random ellipsoids, zero initial field, it_max = 2000, the
reference's default parameters.  Inputs are resident in HBM before timing;
each step resets the field to zero and solves it to termination.

value = sum over steps of (voxels x iterations) / wall time of the steps.

roofline: the fused gradient pass (k_grad_stream).  Algorithmic bytes per
evaluated pass = 12 B per field volume read or written (the controller counts
them: 1 read + 1 speculative write on an accepted-path pass, 2 reads + 2
writes after a rejection) + 5 B of target and mask, per voxel; achieved =
those bytes / average launch duration, the device time of the launches (HIP
events on the solve's stream around each chunk of back-to-back launches) over
their count.

--algorithm greedy times greedy solves of the same volume instead (the
fused greedy pass, k_greedy_stream: 12 B per field volume read or written,
4 B per error map read or written, 5 B of target and mask, per voxel);
--algorithm matching times the matching algorithm (k_match_pass: 44 B/voxel
per cost pass, 28 B/voxel per gradient pass).

cpu_baseline: the reference C library built from its sources (oracle/_ref,
kind "reference"; the oracle restatement, kind "port", if that build is
absent) solving the same 128^3 inputs on the host cores, repeated for about
10 s of CPU work.

--mode copy runs a plain device-to-device copy of 1 GiB five times and prints
its byte count: the calibration of the FETCH_SIZE / WRITE_SIZE counters.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-iters/s at 128³ + p50/p99 solve latency @ fixed QPS vs plain GPU solve"
PEAK_HBM_GBS = 8000.0
# HBM bytes per k_grad_stream launch of the default solve come from the
# committed rocprofv3 summary of that kernel (its --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script, scaled by what the counters read for a
# 1 GiB copy): the line "bench-traffic-bytes-per-launch: N" there.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_grad_stream_128cube.md")


def measured_traffic():
    """(bytes per launch, source) from the kernel's committed profile, or (None, reason)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith("bench-traffic-bytes-per-launch:"):
                    return float(line.split(":", 1)[1]), os.path.relpath(TRAFFIC_PROFILE, ROOT)
    except OSError:
        pass
    return None, "no measured traffic in %s" % os.path.relpath(TRAFFIC_PROFILE, ROOT)


def physical_cores():
    """Physical cores among the CPUs this process may run on (affinity mask
    and `lscpu -p=CPU,CORE,SOCKET`); the affinity count if lscpu is missing."""
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = set(range(os.cpu_count() or 1))
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        cores = set()
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = (int(v) if v else 0 for v in line.split(",")[:3])
            if cpu in allowed:
                cores.add((sock, core))
        if cores:
            return len(cores), "lscpu: %d physical cores among %d allowed CPUs" % (len(cores), len(allowed))
    except Exception:
        pass
    return len(allowed), "%d allowed CPUs (lscpu unavailable)" % len(allowed)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a 128^3 gradient solve takes ~1.6 ms: 20 of them keep the timed region
    # well above the clock's and the host's jitter
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--it-max", type=int, default=2000)
    ap.add_argument("--cpu-iters", type=int, default=2000,
                    help="it_max of the CPU baseline solves (0 to skip the baseline)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--mode", choices=("solve", "batch", "serve", "copy", "pyramid", "jacobian", "sphere"), default="solve")
    ap.add_argument("--algorithm", choices=("gradient", "greedy", "matching"), default="gradient",
                    help="solver of the solve mode (the headline workload is gradient)")
    ap.add_argument("--batch", type=int, default=64,
                    help="volumes of the batched-vs-plain comparison (0 to skip it)")
    ap.add_argument("--serve-matching", type=float, default=0.0,
                    help="fraction of the serving requests turned into matching solves (--mode serve)")
    ap.add_argument("--batch-sides", default="48,160",
                    help="range of each axis of the batch volumes (--mode batch)")
    ap.add_argument("--serve", type=int, default=100,
                    help="requests of the serving stream (0 to skip it)")
    ap.add_argument("--qps", type=float, default=80.0)
    ap.add_argument("--graphs", type=int, default=0, help="engine replays chunks from hipGraphs")
    ap.add_argument("--slots", type=int, default=0, help="engine slots (instances solved at once; 0: 256)")
    ap.add_argument("--lane-variant", type=int, default=-1,
                    help="gradient column tile of the engine's lane (0-4; -1: the engine's choice)")
    ap.add_argument("--tz", type=int, default=0,
                    help="engine z-chunk policy (0: re-planned per set, -1: lone-solve, -2: <= 16, > 0: fixed)")
    ap.add_argument("--serve-dump", default="", help="write per-request records of the serving run (JSON)")
    return ap.parse_args()


def copy_calibration():
    """Five copies of a 1 GiB float32 buffer (known bytes read and written)."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "copy", "copies": 5, "bytes_read_per_copy": 4 * n,
                      "bytes_written_per_copy": 4 * n}))


def cpu_baseline(T, M, sp, iters, threads, min_seconds=10.0, max_repeats=50, algorithm="gradient"):
    """Solves of the same inputs on the host, repeated until about
    `min_seconds` of CPU work (bounded by `max_repeats`)."""
    from oracle.oracle import Oracle, Reference, reference_available
    orc = Oracle(32)
    orc.set_threads(threads)   # shared OpenMP runtime: applies to the reference build too
    n = T.size
    done = 0
    dt = 0.0
    reps = 0
    kind = "reference" if reference_available(32) else "port"
    ref = Reference(32) if kind == "reference" else None
    while reps < max_repeats and (reps == 0 or dt < min_seconds):
        t0 = time.perf_counter()
        if ref is not None:
            if algorithm == "matching":
                _, info = ref.matching(T, M, sp, it_max=iters)
                it = len(info["trace"])
            else:
                _, info = ref.solve(algorithm, T, M, sp, it_max=iters, capture=True)
                it = max(len(info["lines"]) - 1, 0)
        elif algorithm == "matching":
            _, info = orc.matching(T, sp, it_max=iters)
            it = info["iterations"]
        else:
            _, info = orc.solve(algorithm, T, M, sp, it_max=iters)
            it = max(len(info["trace"]) - 1, 0)
        dt += time.perf_counter() - t0
        done += it
        reps += 1
    return {"value": n * done / dt, "unit": "voxel-iters/s", "cores": threads, "kind": kind,
            "sample": "%d %s solve(s) of the same 128^3 inputs (it_max=%d), %d iterations "
                      "in %.1f s" % (reps, algorithm, iters, done, dt)}


def _iterations(info):
    """Iterations performed: the counter ends at it_max + 1 after a full run,
    and at the iteration whose trial triggered the stop otherwise."""
    return info["iterations"] - (1 if info["reason"] == 6 else 0)


def batch_bench(count, it_max, graphs, algorithm="gradient", sides=(48, 160), tz=0, lane_variant=-1):
    """64 volumes of uneven sizes (each axis U{48..160}, spacing 1): solved one
    dt_solve_device (matching: dt_matching_device) call after the other (the
    plain one-volume-per-launch solve) and all at once by the batch engine,
    inputs resident in HBM."""
    import numpy as np
    import torch
    from disptools_amd import native
    from disptools_amd.workloads import batch_uneven

    cases = batch_uneven(count, lo=sides[0], hi=sides[1], workers=8)
    dev = torch.device("cuda:0")
    vols = []
    for T, M, sp in cases:
        vols.append((torch.from_numpy(T).to(dev), torch.from_numpy(M.astype(np.uint8)).to(dev),
                     torch.zeros((3,) + T.shape, dtype=torch.float32, device=dev), sp))
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    def plain(i, its=it_max):
        T, M, F, sp = vols[i]
        nz, ny, nx = T.shape
        if algorithm == "matching":
            info, _ = native.matching_device(nx, ny, nz, sp, T.data_ptr(), F.data_ptr(), stream=sh, it_max=its)
            return info
        info, _ = native.solve_device(algorithm, nx, ny, nz, sp, T.data_ptr(), M.data_ptr(),
                                      F.data_ptr(), stream=sh, it_max=its)
        return info

    # matching counts accepted iterations
    count_its = (lambda i: i["iterations"]) if algorithm == "matching" else _iterations

    # warm-up of both paths on the first volume; the solver's device arena is
    # sized once for the largest volume (a 1-iteration solve), so the timed
    # plain solves do not pay its growth
    big = max(range(len(vols)), key=lambda i: vols[i][0].numel())
    plain(big, 1)
    vols[big][2].zero_()
    vols[0][2].zero_()
    plain(0)
    with native.Engine(use_graphs=bool(graphs), stream=sh, tz=tz, variant=lane_variant) as eng:
        vols[0][2].zero_()
        T, M, F, sp = vols[0]
        eng.submit(algorithm, T.shape[2], T.shape[1], T.shape[0], sp, T.data_ptr(), M.data_ptr(),
                   F.data_ptr(), it_max=it_max)
        eng.drain()
        eng.poll()
    torch.cuda.synchronize()

    for v in vols:
        v[2].zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_infos = [plain(i) for i in range(len(vols))]
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    plain_max = [i["final_max_error"] for i in plain_infos]
    plain_fields = [v[2].clone() for v in vols]

    for v in vols:
        v[2].zero_()
    torch.cuda.synchronize()
    with native.Engine(use_graphs=bool(graphs), stream=sh, tz=tz, variant=lane_variant) as eng:
        t0 = time.perf_counter()
        tickets = {}
        for i, (T, M, F, sp) in enumerate(vols):
            nz, ny, nx = T.shape
            t = eng.submit(algorithm, nx, ny, nz, sp, T.data_ptr(), M.data_ptr(), F.data_ptr(),
                           it_max=it_max)
            if t is None:
                raise RuntimeError("the engine's pool cannot hold the batch")
            tickets[t] = i
        eng.drain()
        res = {tickets[t]: (st, info) for t, st, info in eng.poll()}
        torch.cuda.synchronize()
        batch_wall = time.perf_counter() - t0
        stats = eng.stats()
    batch_infos = [res[i][1] for i in range(len(vols))]

    voxels = [int(np.prod(v[0].shape)) for v in vols]
    plain_vi = sum(n * count_its(i) for n, i in zip(voxels, plain_infos))
    batch_vi = sum(n * count_its(i) for n, i in zip(voxels, batch_infos))
    identical = sum(1 for a, v in zip(plain_fields, vols) if torch.equal(a, v[2]))
    diffs = [abs(b["final_max_error"] - p) for b, p in zip(batch_infos, plain_max)]
    same_reason = sum(1 for b, p in zip(batch_infos, plain_infos) if b["reason"] == p["reason"])
    return {
        "workload": "%d %s solves, each axis U{%d..%d}, spacing 1, smoothed log-normal target, "
                    "ellipsoid mask, zero init, it_max=%d" % (len(vols), algorithm, sides[0], sides[1], it_max),
        "voxels": sum(voxels),
        "plain": {"wall_s": plain_wall, "voxel_iters_per_s": plain_vi / plain_wall,
                  "iterations": sum(count_its(i) for i in plain_infos)},
        "batched": {"wall_s": batch_wall, "voxel_iters_per_s": batch_vi / batch_wall,
                    "iterations": sum(count_its(i) for i in batch_infos),
                    "chunks": stats["chunks"], "pass_launches": stats["passes"],
                    "graph_builds": stats["graph_builds"], "pool_peak_bytes": stats["pool_peak"],
                    "failed": sum(1 for i in range(len(vols)) if res[i][0] != 0)},
        "speedup": (batch_vi / batch_wall) / (plain_vi / plain_wall),
        "same_reason": "%d/%d" % (same_reason, len(vols)),
        "identical_fields": "%d/%d" % (identical, len(vols)),
        "max_abs_diff_final_max_error": max(diffs),
    }


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def serve_bench(count, qps, it_max, graphs, seed=7, slots=0, dump="", matching=0.0, tz=0, lane_variant=-1):
    """Open-loop serving stream: Poisson arrivals at `qps`, cube side
    log-normal (median 96, sigma_ln 0.35, clipped to [32, 256]), 80 % gradient
    and 20 % greedy requests, targets and masks of the 128^3 workload's
    distribution.  Requests arrive as pinned host buffers; latency runs from
    the scheduled arrival to the result being back in host memory.
    plain: one request at a time (upload, dt_solve_device, download), FIFO.
    batched: the batch engine with host staging, requests admitted as soon as
    they arrive and fit (continuous batching)."""
    import collections
    import numpy as np
    import torch
    from disptools_amd import native
    from disptools_amd.workloads import serving_stream, volume_case_torch

    dev = torch.device("cuda:0")
    reqs = serving_stream(count, qps, seed=seed)
    # optionally turn a fraction of the requests into matching solves
    # (dt_matching_device when plain, the engine's matching lane when batched)
    mrng = np.random.default_rng(seed + 1)
    for r in reqs:
        if mrng.random() < matching:
            r["algorithm"] = "matching"

    def its(info, alg):
        return info["iterations"] if alg == "matching" else _iterations(info)
    data = []
    for r in reqs:
        side = r["side"]
        T, M = volume_case_torch((side, side, side), r["seed"], dev)
        data.append((T.cpu().pin_memory(), M.cpu().pin_memory(),
                     torch.zeros((3, side, side, side), dtype=torch.float32).pin_memory()))
    torch.cuda.synchronize()
    voxels = [r["side"] ** 3 for r in reqs]
    nmax = max(voxels)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    # plain: FIFO, one solve at a time through device staging buffers
    dT = torch.empty(nmax, dtype=torch.float32, device=dev)
    dM = torch.empty(nmax, dtype=torch.uint8, device=dev)
    dF = torch.empty(3 * nmax, dtype=torch.float32, device=dev)
    # size the solver's device arena for the largest request before the
    # stream starts (a server allocates up front; a first-time growth after a
    # large free can take hundreds of milliseconds)
    side_max = max(r["side"] for r in reqs)
    dT[:nmax].fill_(1.0)
    dM[:nmax].zero_()
    dF[:3 * nmax].zero_()
    native.solve_device("gradient", side_max, side_max, side_max, (1.0, 1.0, 1.0), dT.data_ptr(), dM.data_ptr(),
                        dF.data_ptr(), stream=sh, it_max=1)
    torch.cuda.synchronize()
    lat_plain, vi_plain, info_plain, busy = [], 0, [], 0.0
    start = time.perf_counter()
    for i, r in enumerate(reqs):
        now = time.perf_counter() - start
        if now < r["arrival"]:
            time.sleep(r["arrival"] - now)
        T, M, F = data[i]
        n, side = voxels[i], r["side"]
        t_begin = time.perf_counter()
        dT[:n].copy_(T.view(-1), non_blocking=True)
        dM[:n].copy_(M.view(-1), non_blocking=True)
        dF[:3 * n].zero_()
        if r["algorithm"] == "matching":
            info, _ = native.matching_device(side, side, side, (1.0, 1.0, 1.0), dT.data_ptr(), dF.data_ptr(),
                                             stream=sh, it_max=it_max)
        else:
            info, _ = native.solve_device(r["algorithm"], side, side, side, (1.0, 1.0, 1.0), dT.data_ptr(),
                                          dM.data_ptr(), dF.data_ptr(), stream=sh, it_max=it_max)
        F.view(-1).copy_(dF[:3 * n], non_blocking=True)
        stream.synchronize()
        busy += time.perf_counter() - t_begin
        lat_plain.append(time.perf_counter() - start - r["arrival"])
        vi_plain += n * its(info, r["algorithm"])
        info_plain.append(info)
    plain_span = time.perf_counter() - start - reqs[0]["arrival"]
    del dT, dM, dF

    # batched: continuous batching
    lat_b = [None] * len(reqs)
    info_b = [None] * len(reqs)
    vi_b = 0
    for d in data:
        d[2].zero_()
    with native.Engine(use_graphs=bool(graphs), stream=sh, max_slots=slots, tz=tz, variant=lane_variant) as eng:
        queue = collections.deque()
        tickets = {}
        nxt = 0
        done = 0
        start = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - start
            while nxt < len(reqs) and reqs[nxt]["arrival"] <= now:
                queue.append(nxt)
                nxt += 1
            while queue:
                i = queue[0]
                side = reqs[i]["side"]
                T, M, F = data[i]
                t = eng.submit(reqs[i]["algorithm"], side, side, side, (1.0, 1.0, 1.0), T.data_ptr(),
                               M.data_ptr(), F.data_ptr(), host_io=True, it_max=it_max)
                if t is None:
                    break
                tickets[t] = queue.popleft()
            left = eng.progress(wait=True)
            for t, status, info in eng.poll():
                i = tickets.pop(t)
                lat_b[i] = time.perf_counter() - start - reqs[i]["arrival"]
                info_b[i] = info
                vi_b += voxels[i] * its(info, reqs[i]["algorithm"])
                done += 1
            if left == 0 and not queue and nxt < len(reqs):
                wait = reqs[nxt]["arrival"] - (time.perf_counter() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.001))
        batch_span = time.perf_counter() - start - reqs[0]["arrival"]
        stats = eng.stats()
    if dump:
        with open(dump, "w") as fh:
            json.dump([dict(side=r["side"], algorithm=r["algorithm"], arrival=r["arrival"],
                            plain_latency=lp, batched_latency=lb, iterations=ip["iterations"])
                       for r, lp, lb, ip in zip(reqs, lat_plain, lat_b, info_plain)], fh)
    same_reason = sum(1 for a, b in zip(info_plain, info_b) if a["reason"] == b["reason"])
    diff = max(abs(a["final_max_error"] - b["final_max_error"]) for a, b in zip(info_plain, info_b))
    return {
        "workload": "%d requests, Poisson %.1f req/s, cube side log-normal (median 96, sigma_ln 0.35, "
                    "clip [32, 256]), 80%% gradient / 20%% greedy%s, it_max=%d, host buffers (pinned)"
                    % (len(reqs), qps, ", then %.0f%% of all turned into matching" % (100 * matching)
                       if matching else "", it_max),
        "matching_requests": sum(1 for r in reqs if r["algorithm"] == "matching"),
        "engine_slots": slots or 256,
        "voxels": sum(voxels),
        "plain": {"p50_s": _pct(lat_plain, 50), "p99_s": _pct(lat_plain, 99),
                  "mean_s": float(np.mean(lat_plain)), "voxel_iters_per_s": vi_plain / plain_span,
                  "busy_fraction": busy / plain_span},
        "batched": {"p50_s": _pct(lat_b, 50), "p99_s": _pct(lat_b, 99), "mean_s": float(np.mean(lat_b)),
                    "voxel_iters_per_s": vi_b / batch_span, "chunks": stats["chunks"],
                    "graph_builds": stats["graph_builds"], "pool_peak_bytes": stats["pool_peak"]},
        "same_reason": "%d/%d" % (same_reason, len(reqs)),
        "max_abs_diff_final_max_error": diff,
    }


def pyramid_bench(it_max, reps=3):
    """displacement() on the 128^3 workload (numpy in and out) with 1 and with
    3 pyramid levels: the multi-resolution driver keeps the pyramid and the
    field on the device (one upload, one download per call)."""
    import numpy as np
    from disptools_amd import native
    from disptools_amd.displacements import displacement
    from disptools_amd.workloads import volume_128

    native.set_device(0)
    T, M, sp = volume_128()
    out = {"workload": "displacement() of the 128^3 gradient workload, numpy in/out, it_max=%d per level"
                       % it_max}
    for levels in (1, 3):
        displacement(T, mask=M, spacing=sp, levels=levels, it_max=it_max, gpu_id=0)   # warm-up
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f = displacement(T, mask=M, spacing=sp, levels=levels, it_max=it_max, gpu_id=0)
            walls.append(time.perf_counter() - t0)
        out["levels_%d" % levels] = {"wall_s": float(np.median(walls)), "finite": bool(np.isfinite(f).all())}
    return out


def jacobian_bench(sides=(128, 256), reps=50):
    """dt_jacobian_device (k_jacobian) on a random field of side^3 voxels,
    resident in HBM: average launch duration from HIP events on the launch
    stream, against 16 algorithmic bytes per voxel (3 field floats read, J
    written)."""
    import torch
    from disptools_amd import native

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    out = {"workload": "jacobian of a random field (N(0, 0.1) per component), spacing (1, 1, 1.5)",
           "bytes_per_voxel": 16}
    for side in sides:
        g = torch.Generator(device=dev).manual_seed(side)
        F = torch.randn((3, side, side, side), generator=g, device=dev, dtype=torch.float32) * 0.1
        J = torch.empty((side, side, side), device=dev, dtype=torch.float32)
        sp = (1.0, 1.0, 1.5)
        for _ in range(3):
            native.jacobian_device(side, side, side, sp, F.data_ptr(), J.data_ptr(), stream=stream.cuda_stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            native.jacobian_device(side, side, side, sp, F.data_ptr(), J.data_ptr(), stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        gbs = 16.0 * side ** 3 / (us * 1e-6) / 1e9
        out["%d^3" % side] = {"avg_launch_us": us, "achieved_GBps": gbs, "frac": gbs / PEAK_HBM_GBS}
    return out


def sphere_bench(threads, reps=20, min_cpu_seconds=5.0):
    """The reference's CPU-runnable configuration: a 40^3 sphere target
    (radius 10, 1.1 inside, 1.0 outside; mask = the sphere), spacing 1, zero
    init, gradient, it_max 500.  GPU: dt_solve_device on resident buffers,
    `reps` solves timed.  CPU: the reference C build (oracle/_ref; the
    restatement when it is missing) on `threads` threads and on one, same
    inputs, repeated for about `min_cpu_seconds` each.  Each side reports
    iterations, stop reason, time per solve, voxel-iters/s and the final
    masked max |J - T| of its field (FP32 jacobian of the oracle)."""
    import numpy as np
    import torch
    from disptools_amd import native
    from disptools_amd.workloads import sphere_case
    from oracle.oracle import Oracle, Reference, reference_available, REASON_NAMES

    T, M, sp = sphere_case(40)
    n = T.size
    o32 = Oracle(32)

    def masked_max(field):
        _, mx, _ = o32.compute_error(T, o32.jacobian(np.ascontiguousarray(field, np.float32), sp), M, 0.2)
        return mx

    dev = torch.device("cuda:0")
    dT = torch.from_numpy(T).to(dev)
    dM = torch.from_numpy(M.astype(np.uint8)).to(dev)
    dF = torch.zeros((3,) + T.shape, dtype=torch.float32, device=dev)
    sh = torch.cuda.current_stream(dev).cuda_stream

    def gpu_solve():
        dF.zero_()
        info, _ = native.solve_device("gradient", 40, 40, 40, sp, dT.data_ptr(), dM.data_ptr(), dF.data_ptr(),
                                      stream=sh, it_max=500)
        return info

    gpu_solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        info = gpu_solve()
    torch.cuda.synchronize()
    g_s = (time.perf_counter() - t0) / reps
    gpu = {"iterations": info["iterations"], "reason": native.REASONS[info["reason"]], "seconds_per_solve": g_s,
           "voxel_iters_per_s": n * info["iterations"] / g_s, "final_masked_max_error": masked_max(dF.cpu().numpy())}

    kind = "reference" if reference_available(32) else "port"
    ref = Reference(32) if kind == "reference" else None

    def cpu(nthreads):
        o32.set_threads(nthreads)   # shared OpenMP runtime: applies to the reference build too
        done = reps_c = 0
        dt = 0.0
        while reps_c == 0 or (dt < min_cpu_seconds and reps_c < 50):
            t0 = time.perf_counter()
            if ref is not None:
                f, inf = ref.solve("gradient", T, M, sp, it_max=500, capture=True)
                its, reason = inf["iterations"], inf["reason"]
            else:
                f, inf = o32.solve("gradient", T, M, sp, it_max=500)
                its, reason = inf["iterations"], inf["reason"]
            dt += time.perf_counter() - t0
            done += its
            reps_c += 1
        o32.set_threads(1)
        return {"threads": nthreads, "iterations": its, "reason": REASON_NAMES.get(reason, reason),
                "seconds_per_solve": dt / reps_c, "voxel_iters_per_s": n * done / dt,
                "final_masked_max_error": masked_max(f), "solves": reps_c}

    out = {"workload": "40^3 sphere (radius 10, 1.1 inside / 1.0 outside, mask = sphere), spacing 1, zero init, "
                       "gradient, it_max 500", "gpu": gpu, "cpu_kind": kind, "cpu": cpu(threads)}
    if threads > 1:
        out["cpu_one_thread"] = cpu(1)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (each GPU box has one device); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        copy_calibration()
        return
    if args.mode == "pyramid":
        print(json.dumps({"mode": "pyramid", **pyramid_bench(args.it_max)}))
        return
    if args.mode == "sphere":
        from disptools_amd import native
        if not native.set_device(0):
            raise RuntimeError("no HIP device")
        cores, cores_source = physical_cores()
        out = sphere_bench(args.cpu_threads or cores)
        out["cores_source"] = cores_source if not args.cpu_threads else "--cpu-threads"
        print(json.dumps({"mode": "sphere", **out}))
        return
    if args.mode == "jacobian":
        from disptools_amd import native
        if not native.set_device(0):
            raise RuntimeError("no HIP device")
        print(json.dumps({"mode": "jacobian", **jacobian_bench()}))
        return
    if args.mode in ("batch", "serve"):
        from disptools_amd import native
        if not native.set_device(0):
            raise RuntimeError("no HIP device")
        if args.mode == "batch":
            sides = tuple(int(v) for v in args.batch_sides.split(","))
            print(json.dumps(batch_bench(args.batch or 64, args.it_max, args.graphs, args.algorithm, sides, args.tz,
                                          args.lane_variant)))
        else:
            print(json.dumps(serve_bench(args.serve or 100, args.qps, args.it_max, args.graphs,
                                         matching=args.serve_matching,
                                         slots=args.slots, dump=args.serve_dump, tz=args.tz,
                                         lane_variant=args.lane_variant)))
        return

    import numpy as np
    import torch
    from disptools_amd import native
    from disptools_amd.workloads import volume_128

    if not native.set_device(0):
        raise RuntimeError("no HIP device")
    T, M, sp = volume_128()
    nz, ny, nx = T.shape
    n = T.size
    dev = torch.device("cuda:0")
    d_T = torch.from_numpy(T).to(dev)
    d_M = torch.from_numpy(M.astype(np.uint8)).to(dev)
    d_F = torch.zeros((3, nz, ny, nx), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    def step():
        d_F.zero_()
        if args.algorithm == "matching":
            info, _ = native.matching_device(nx, ny, nz, sp, d_T.data_ptr(), d_F.data_ptr(), stream=sh,
                                             it_max=args.it_max)
            return info
        info, _ = native.solve_device(args.algorithm, nx, ny, nz, sp, d_T.data_ptr(), d_M.data_ptr(),
                                      d_F.data_ptr(), stream=sh, it_max=args.it_max)
        return info

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    infos = []
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        infos.append(step())
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)

    # iterations performed: the counter ends at it_max + 1 after a full run, and
    # at the iteration whose trial triggered the stop otherwise
    launches = sum(i["launches"] for i in infos)
    if args.algorithm == "matching":
        # accepted iterations; a cost pass reads the accepted positions, the
        # gradient and the target and writes the trial and its cell weights
        # (44 B/voxel; the initial evaluation 20 B), a gradient pass reads the
        # positions and weights and writes the gradient (28 B/voxel)
        iters = sum(i["iterations"] for i in infos)
        costs = sum(i["cost_passes"] for i in infos)
        grads = sum(i["gradient_passes"] for i in infos)
        alg_bytes = n * (44.0 * (costs - len(infos)) + 20.0 * len(infos) + 28.0 * grads)
    else:
        iters = sum(i["iterations"] - (1 if i["reason"] == 6 else 0) for i in infos)
        evals = sum(i["evaluations"] for i in infos)
        field_io = sum(i["field_io"] for i in infos)
        alg_bytes = n * (12.0 * field_io + 5.0 * evals)
        if args.algorithm == "greedy":
            # error maps: every pass writes one, every pass but the initial
            # evaluation also reads the previous one
            alg_bytes += n * (4.0 * evals + 4.0 * (evals - len(infos)))
    voxel_iters = n * iters
    value = voxel_iters / wall
    # device time of the pass launches (HIP events recorded by the library on
    # the solve's stream around each chunk of back-to-back launches)
    avg_launch_s = sum(i["kernel_seconds"] for i in infos) / launches
    achieved = alg_bytes / launches / avg_launch_s / 1e9

    batch = (batch_bench(args.batch, args.it_max, args.graphs, tz=args.tz, lane_variant=args.lane_variant)
             if args.batch > 0 else None)
    serving = (serve_bench(args.serve, args.qps, args.it_max, args.graphs, slots=args.slots,
                           dump=args.serve_dump, tz=args.tz, lane_variant=args.lane_variant)
               if args.serve > 0 else None)

    cores, cores_source = physical_cores()
    threads = args.cpu_threads or cores
    # the reference's matching is single-threaded and takes seconds per
    # iteration at 128^3: its sample is a few iterations
    cpu_iters = min(args.cpu_iters, 5) if args.algorithm == "matching" else args.cpu_iters
    cpu = None
    if args.cpu_iters > 0:
        cpu = cpu_baseline(T, M, sp, cpu_iters, threads, algorithm=args.algorithm)
        cpu["cores_source"] = cores_source if not args.cpu_threads else "--cpu-threads"
        if threads > 1 and args.algorithm != "matching":
            one = cpu_baseline(T, M, sp, cpu_iters, 1, algorithm=args.algorithm, max_repeats=2)
            cpu["one_thread"] = {"value": one["value"], "unit": one["unit"], "sample": one["sample"]}

    default_workload = (args.it_max == 2000 and args.algorithm == "gradient" and
                        "DISPTOOLS_GRAD_VARIANT" not in os.environ)
    traffic, traffic_source = measured_traffic() if default_workload else (None, "not the profiled workload")
    last = infos[-1]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s solve, 128x128x128, spacing (1,1,1.5), smoothed log-normal "
                               "target, ellipsoid mask, zero init, it_max=%d" % (args.algorithm, args.it_max),
                   "voxels": n, "it_max": args.it_max},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": {"gradient": "k_grad_stream", "greedy": "k_greedy_stream",
                                "matching": "k_match_pass"}[args.algorithm],
                     "avg_launch_us": avg_launch_s * 1e6,
                     "alg_bytes_per_launch": alg_bytes / launches},
        "cpu_baseline": cpu,
        "batch": batch,
        "serving": serving,
        "solve": dict({k: v for k, v in last.items() if k not in ("seconds", "trace_len", "kernel_seconds")},
                      reason=native.REASONS[last["reason"]], gpu_ms_per_step=gpu_ms / args.steps),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
