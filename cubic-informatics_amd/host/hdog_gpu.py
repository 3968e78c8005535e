#!/usr/bin/env python
"""Assign image stacks to GPUs and run HDoG3D_main on each GPU in parallel.

The drop-in for script/HDoG_gpu.py (script/HDoG_gpu.py:1-126): the same parameter file
(`gpus`, `src_basedir`, `dst_basedir` plus the HDoG3D_main keys), the same stack layout
{src_basedir}/{Y}/{Y}_{X}/{Z}.bin, the same per-GPU parameter files param_gpu{NN}.json and
logs log_gpu{NN}.txt beside PARAM_FILE, and one result file {dst_basedir}/{Y}_{X}.bin per
stack.  Differences:

* stacks are balanced over the GPUs longest first onto the least loaded GPU (``--balance
  lpt``, the default), which evens out uneven stacks better than the reference's
  round-robin over the length-sorted list (``--balance rr`` reproduces that assignment);
* stacks of equal length are ordered by path, so the assignment does not depend on the
  file system's listing order;
* the exit status is 1 when any HDoG3D_main run fails (the reference ignores it).

Usage:
  hdog_gpu.py PARAM_FILE [--zmin Z] [--zmax Z] [--exec PATH] [--balance lpt|rr] [--dry-run]
  hdog_gpu.py PARAM_FILE single XNAME YNAME [same options]
"""
import argparse
import copy
import glob
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def default_exec():
    local = os.path.join(HERE, "HDoG3D_main")
    return local if os.path.exists(local) else "/usr/local/bin/HDoG3D_main"


def list_stacks(params, zmin=None, zmax=None, single=None):
    """[(stack_dir, n_images)], longest first, ties by path (script/HDoG_gpu.py:61-72)."""
    base = params["src_basedir"]
    if single:
        xname, yname = single
        dirs = [os.path.join(base, yname, "{}_{}".format(yname, xname))]
    else:
        dirs = [d for d in glob.glob(os.path.join(base, "*", "*")) if os.path.isdir(d)]
    stacks = [(d, len(sorted(glob.glob(os.path.join(d, "*.bin")))[zmin:zmax])) for d in dirs]
    stacks.sort(key=lambda s: (-s[1], s[0]))
    return stacks


def assign(stacks, gpus, balance="lpt"):
    """{gpu: [stack_dir, ...]} for stacks sorted longest first.

    rr: stack i goes to gpus[i % len(gpus)] (script/HDoG_gpu.py:73-81).
    lpt: each stack goes to the GPU with the fewest images so far (first such GPU in the
    `gpus` order on ties)."""
    per = {g: [] for g in gpus}
    load = {g: 0 for g in gpus}
    for i, (path, n) in enumerate(stacks):
        if balance == "rr":
            g = gpus[i % len(gpus)]
        elif balance == "lpt":
            g = min(gpus, key=lambda k: (load[k], gpus.index(k)))
        else:
            raise ValueError("balance must be 'lpt' or 'rr', not %r" % balance)
        per[g].append(path)
        load[g] += n
    return per, load


def gpu_params(params, gpu, stack_dirs, zmin=None, zmax=None):
    """The HDoG3D_main parameter file of one GPU (script/HDoG_gpu.py:83-101)."""
    p = copy.deepcopy(params)
    p["devID"] = gpu
    p["stacks"] = []
    for d in stack_dirs:
        yxname = os.path.basename(os.path.normpath(d))
        znames = sorted(glob.glob(os.path.join(d, "*.bin")))[zmin:zmax]
        p["stacks"].append({
            "dst_path": os.path.join(params["dst_basedir"], "{}.bin".format(yxname)),
            "src_paths": [os.path.join(d, os.path.basename(z)) for z in znames],
        })
    return p


def parse(argv):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("param_file")
    ap.add_argument("single", nargs="*", help="single XNAME YNAME")
    ap.add_argument("--zmin", type=int, default=None)
    ap.add_argument("--zmax", type=int, default=None)
    ap.add_argument("--exec", dest="exe", default=None)
    ap.add_argument("--balance", default="lpt", choices=["lpt", "rr"])
    ap.add_argument("--dry-run", action="store_true",
                    help="write the per-GPU parameter files and print the plan only")
    a = ap.parse_args(argv)
    if a.single and (len(a.single) != 3 or a.single[0] != "single"):
        ap.error("expected: PARAM_FILE single XNAME YNAME")
    a.single = tuple(a.single[1:]) if a.single else None
    return a


def main(argv=None):
    a = parse(sys.argv[1:] if argv is None else argv)
    with open(a.param_file) as f:
        params = json.load(f)
    gpus = list(params["gpus"])
    os.makedirs(params["dst_basedir"], exist_ok=True)
    stacks = list_stacks(params, a.zmin, a.zmax, a.single)
    per, load = assign(stacks, gpus, a.balance)
    exe = a.exe or default_exec()
    pdir = os.path.dirname(os.path.abspath(a.param_file))
    jobs = []
    for g in gpus:
        print("GPU:{}\timages:{}".format(g, load[g]))
        pfile = os.path.join(pdir, "param_gpu{:02d}.json".format(g))
        lfile = os.path.join(pdir, "log_gpu{:02d}.txt".format(g))
        with open(pfile, "w") as f:
            json.dump(gpu_params(params, g, per[g], a.zmin, a.zmax), f, indent=4)
        print("GPU:{}\tparamfile: {}".format(g, pfile))
        jobs.append((g, pfile, lfile))
    if a.dry_run:
        return 0
    procs = []
    for g, pfile, lfile in jobs:
        print("Executing : {} {}".format(exe, pfile))
        log = open(lfile, "wb")
        procs.append((g, subprocess.Popen([exe, pfile], stdout=log, stderr=subprocess.STDOUT),
                      log, pfile))
    rc = 0
    for g, p, log, pfile in procs:
        code = p.wait()
        log.close()
        print("[*] finished GPU {} (exit {})".format(g, code))
        if code != 0:
            rc = 1
    for _, _, _, pfile in procs:
        os.remove(pfile)
    return rc


if __name__ == "__main__":
    sys.exit(main())
