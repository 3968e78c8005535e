#!/usr/bin/env python
"""End-to-end timing of HDoG3D_main on a synthetic stack written as slice files.

Renders the 2160 x 2560 stack model (background mean 150 -> 400 along z, sd 20, Gaussian
blobs with sigma_xy in [1.5, 3.5], sigma_z = 0.8 sigma_xy, one blob per 2e4 voxels;
512 slices by default) on the device, writes it as {Y}/{Y}_{X}/{Z:06d}.bin slices under
OUTDIR, writes a parameter file (radius 10, DoG sigma1 = (z 1.6, xy 2.0), sigma2 = 1.6
sigma1, truncation 4, margins 100) and runs the host program on it, timing the whole run
from the process start (slice reads, H2D, GPU work, D2H, record writes).

With --stacks N the parameter file lists the same slices N times (N result files), as a
parameter file of N stacks of one brain would: the program then runs one pipeline over
all of them (the next stack's slices are read while the previous one runs).

Prints one JSON line: wall seconds, stack voxels per second, the number of substacks and
the host program's per-substack timings.

    python e2e_stack.py OUTDIR [--depth 512] [--slices 512] [--stacks 1] [--exe PATH]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("outdir")
    ap.add_argument("--slices", type=int, default=512, help="stack length (z)")
    ap.add_argument("--depth", type=int, default=512, help="HDoG_param.depth (substack depth)")
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--width", type=int, default=2560)
    ap.add_argument("--stacks", type=int, default=1, help="times the stack is listed")
    ap.add_argument("--exe", default=os.path.join(HERE, "HDoG3D_main"))
    ap.add_argument("--keep", action="store_true", help="reuse slices already in OUTDIR")
    args = ap.parse_args()

    shape = (args.slices, args.height, args.width)
    srcdir = os.path.join(args.outdir, "src")
    sdir = os.path.join(srcdir, "000000", "000000_000000")
    paths = [os.path.join(sdir, f"{z:06d}.bin") for z in range(shape[0])]
    if not (args.keep and all(os.path.exists(p) for p in paths)):
        import numpy as np
        import torch
        from hdog import HDoG, synth
        cfg = synth.WORKLOADS["c1"]
        nb = int(shape[0] * shape[1] * shape[2] / 2e4)
        blobs = synth.blob_table(shape, nb, cfg["seed"], cfg["sigma"], cfg["zfactor"], cfg["amp"])
        hd = HDoG(0)
        img = hd.synth_volume(shape, blobs, cfg["seed"], cfg["mean0"], cfg["mean1"], cfg["sd"])
        vol = img.cpu().numpy()
        del img
        hd.close()
        torch.cuda.empty_cache()
        os.makedirs(sdir, exist_ok=True)
        for z, p in enumerate(paths):
            np.ascontiguousarray(vol[z]).astype("<u2").tofile(p)
        del vol
    param = {
        "devID": 0, "verbose": False,
        "input_image_info": {"width": shape[2], "height": shape[1], "left_margin": 100,
                             "right_margin": 100, "top_margin": 100, "bottom_margin": 100},
        "coordinate_info": {"scale_x": 1.0, "scale_y": 1.0, "scale_z": 1.0},
        "HDoG_param": {"depth": args.depth, "extra_depth_margin": 3, "min_intensity_skip": 0,
                       "radius_norm": {"xy_large": 10, "z_large": 10}, "min_intensity": 0,
                       "gamma_n": 1.0, "min_size": 0, "truncate": 4.0,
                       "sigma_dog": {"xy_small": 2.0, "z_small": 1.6, "xy_large": 3.2,
                                     "z_large": 2.56}},
        "stacks": [{"src_paths": paths, "dst_path": os.path.join(args.outdir, f"cells_{k}.bin")}
                   for k in range(args.stacks)],
    }
    pfile = os.path.join(args.outdir, "param.json")
    with open(pfile, "w") as f:
        json.dump(param, f)
    t0 = time.perf_counter()
    r = subprocess.run([args.exe, pfile], capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        print(r.stdout[-2000:], r.stderr[-2000:], file=sys.stderr)
        sys.exit(r.returncode)
    out = r.stdout

    def col(pat):
        return [float(x) for x in re.findall(pat, out)]
    n_vox = shape[0] * shape[1] * shape[2] * args.stacks
    res = {
        "stack_zyx": list(shape), "stacks": args.stacks, "depth": args.depth, "wall_s": wall,
        "init_ms": col(r"Initialization took ([0-9.]+)msec"),
        "gvoxel_per_s": n_vox / wall / 1e9,
        "substacks": len(col(r"processSubstack\(z0=(\d+)")),
        "load_ms": col(r"loadImage took ([0-9.]+)msec"),
        "gpu_ms": col(r"GPU memcpy & kernel execution took ([0-9.]+)msec"),
        "save_ms": col(r"saveImage took ([0-9.]+)msec"),
        "records_bytes": os.path.getsize(param["stacks"][0]["dst_path"]),
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
