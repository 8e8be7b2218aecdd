"""Per-shape timings of the column tiles of the solver passes on this GPU.

    python3 -m disptools_amd.tile_sweep --algorithm gradient   # dt_tune_gradient per shape
    python3 -m disptools_amd.tile_sweep --algorithm greedy     # 32x16 vs 32x32 fused greedy pass

One JSON line per shape.  The gradient sweep times every compiled variant of
the gradient pass with the autotuner (stop tests disabled, so every timed
launch evaluates a trial); the greedy sweep runs 100-iteration greedy solves
on device buffers with each fused tile (DISPTOOLS_GREEDY_VARIANT) and reports
the device time per launch, best of three.  These are the data behind the
single-solve tile rules (the gradient cost model and the greedy plane-size
threshold).
"""

import argparse
import json
import os

GRADIENT_SHAPES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (80, 80, 80), (96, 96, 96), (112, 112, 112),
                   (128, 128, 128), (144, 144, 144), (160, 160, 160), (192, 192, 192), (224, 224, 224),
                   (256, 256, 256), (140, 60, 100), (96, 128, 64), (128, 128, 64), (64, 200, 120),
                   (57, 91, 133), (150, 110, 70)]                 # (nx, ny, nz)
GREEDY_SHAPES = [(48, 48, 48), (64, 64, 64), (80, 80, 80), (96, 96, 96), (112, 112, 112), (128, 128, 128),
                 (160, 160, 160), (200, 200, 200), (60, 100, 140), (140, 100, 60), (70, 60, 100),
                 (64, 128, 128), (128, 64, 96), (90, 90, 40)]     # (nz, ny, nx)


def sweep_gradient(native, passes=12):
    for nx, ny, nz in GRADIENT_SHAPES:
        best, us = native.tune_gradient(nx, ny, nz, spacing=(1.0, 1.0, 1.0), passes=passes)
        print(json.dumps({"shape_xyz": [nx, ny, nz], "best": best,
                          "us_per_pass": {k: round(v, 2) for k, v in us.items()}}), flush=True)


def sweep_greedy(native, it_max=100, reps=3):
    import torch
    from .workloads import volume_case_torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    old = os.environ.get("DISPTOOLS_GREEDY_VARIANT")
    try:
        for shape in GREEDY_SHAPES:
            T, M = volume_case_torch(shape, 3, dev)
            nz, ny, nx = shape
            F = torch.zeros((3,) + shape, dtype=torch.float32, device=dev)
            out = {}
            for v, name in (("1", "32x16"), ("2", "32x32")):
                os.environ["DISPTOOLS_GREEDY_VARIANT"] = v
                best = None
                for _ in range(reps):
                    F.zero_()
                    torch.cuda.synchronize()
                    info, _ = native.solve_device("greedy", nx, ny, nz, (1.0, 1.0, 1.0), T.data_ptr(),
                                                  M.data_ptr(), F.data_ptr(), stream=stream, it_max=it_max)
                    us = 1e6 * info["kernel_seconds"] / info["launches"]
                    best = us if best is None else min(best, us)
                out[name] = round(best, 2)
            print(json.dumps({"shape_zyx": list(shape), "nx_ny": nx * ny, "us_per_launch": out}), flush=True)
    finally:
        if old is None:
            os.environ.pop("DISPTOOLS_GREEDY_VARIANT", None)
        else:
            os.environ["DISPTOOLS_GREEDY_VARIANT"] = old


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--algorithm", choices=("gradient", "greedy"), default="gradient")
    args = ap.parse_args()
    from . import native
    if not native.set_device(0):
        raise SystemExit("no HIP device visible")
    if args.algorithm == "gradient":
        sweep_gradient(native)
    else:
        sweep_greedy(native)


if __name__ == "__main__":
    main()
