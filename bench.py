"""Benchmark: voxel-iterations per second of the gradient solve at 128^3.

Workload (one step): a complete gradient solve of the 128x128x128 volume,
spacing (1, 1, 1.5), target exp(smoothed Gaussian) clipped to [0.8, 1.25],
mask = union of random ellipsoids, zero initial field, it_max = 2000, the
reference's default parameters.  Inputs are resident in HBM before timing;
each step resets the field to zero and solves it to termination.

value = sum over steps of (voxels x iterations) / wall time of the steps.

roofline: the fused gradient pass (k_grad_pass).  Algorithmic bytes per
evaluated pass = 12 B per field volume read or written (the controller counts
them: 1 read + 1 speculative write on an accepted-path pass, 2 reads + 2
writes after a rejection) + 5 B of target and mask, per voxel; achieved =
those bytes / (HIP-event time of the solves on their stream / pass launches).

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
# HBM bytes per k_grad_pass launch of the default solve (128^3, it_max 2000,
# default pass variant), from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# runs of this script, scaled by what the counters read for a 1 GiB copy
# (--mode copy): read bytes = 2.0 x FETCH_SIZE, written = 1.0 x WRITE_SIZE.
MEASURED_TRAFFIC_128 = 68.85e6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--it-max", type=int, default=2000)
    ap.add_argument("--cpu-iters", type=int, default=2000,
                    help="it_max of the CPU baseline solves (0 to skip the baseline)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--mode", choices=("solve", "copy"), default="solve")
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


def cpu_baseline(T, M, sp, iters, threads, min_seconds=10.0, max_repeats=50):
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
            _, info = ref.solve("gradient", T, M, sp, it_max=iters, capture=True)
            it = max(len(info["lines"]) - 1, 0)
        else:
            _, info = orc.solve("gradient", T, M, sp, it_max=iters)
            it = max(len(info["trace"]) - 1, 0)
        dt += time.perf_counter() - t0
        done += it
        reps += 1
    return {"value": n * done / dt, "unit": "voxel-iters/s", "cores": threads, "kind": kind,
            "sample": "%d gradient solve(s) of the same 128^3 inputs (it_max=%d), %d iterations "
                      "in %.1f s" % (reps, iters, done, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (each GPU box has one device); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        copy_calibration()
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
        info, _ = native.solve_device("gradient", nx, ny, nz, sp, d_T.data_ptr(), d_M.data_ptr(),
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
    iters = sum(i["iterations"] - (1 if i["reason"] == 6 else 0) for i in infos)
    voxel_iters = n * iters
    value = voxel_iters / wall
    launches = sum(i["launches"] for i in infos)
    evals = sum(i["evaluations"] for i in infos)
    field_io = sum(i["field_io"] for i in infos)
    alg_bytes = n * (12.0 * field_io + 5.0 * evals)
    avg_launch_s = (gpu_ms / 1e3) / launches
    achieved = alg_bytes / launches / avg_launch_s / 1e9

    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    cpu = cpu_baseline(T, M, sp, args.cpu_iters, threads) if args.cpu_iters > 0 else None

    default_workload = args.it_max == 2000 and "DISPTOOLS_GRAD_VARIANT" not in os.environ
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
        "config": {"workload": "gradient solve, 128x128x128, spacing (1,1,1.5), smoothed log-normal "
                               "target, ellipsoid mask, zero init, it_max=%d" % args.it_max,
                   "voxels": n, "it_max": args.it_max},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": MEASURED_TRAFFIC_128 if default_workload else None,
                     "kernel": "k_grad_pass", "avg_launch_us": avg_launch_s * 1e6,
                     "alg_bytes_per_launch": alg_bytes / launches},
        "cpu_baseline": cpu,
        "solve": {"iterations": last["iterations"], "reason": native.REASONS[last["reason"]],
                  "evaluations": last["evaluations"], "launches": last["launches"],
                  "rejections": last["rejections"], "final_error": last["final_error"],
                  "final_max_error": last["final_max_error"], "gpu_ms_per_step": gpu_ms / args.steps},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
