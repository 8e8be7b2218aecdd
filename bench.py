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
absent) on the same 128^3 inputs for a bounded number of iterations.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--it-max", type=int, default=2000)
    ap.add_argument("--cpu-iters", type=int, default=200,
                    help="iterations of the CPU baseline sample (0 to skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(T, M, sp, iters, threads):
    from oracle.oracle import Oracle, Reference, reference_available
    orc = Oracle(32)
    orc.set_threads(threads)   # shared OpenMP runtime: applies to the reference build too
    n = T.size
    if reference_available(32):
        ref = Reference(32)
        t0 = time.perf_counter()
        _, info = ref.solve("gradient", T, M, sp, it_max=iters, capture=True)
        dt = time.perf_counter() - t0
        done = max(len(info["lines"]) - 1, 0)
        kind = "reference"
    else:
        t0 = time.perf_counter()
        _, info = orc.solve("gradient", T, M, sp, it_max=iters)
        dt = time.perf_counter() - t0
        done = max(len(info["trace"]) - 1, 0)
        kind = "port"
    return {"value": n * done / dt, "unit": "voxel-iters/s", "cores": threads, "kind": kind,
            "sample": "gradient solve of the same 128^3 inputs, first %d iterations (%.1f s)" % (done, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (each GPU box has one device); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

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
                     "frac": achieved / PEAK_HBM_GBS, "traffic": None,
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
