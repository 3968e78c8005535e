"""The multi-GPU fan-out (host/hdog_gpu.py), the drop-in for script/HDoG_gpu.py.

CPU tests build a stack tree of empty slice files and run the fan-out with a stand-in
executable that copies the parameter file it was given, then check the per-GPU parameter
files (the reference's round-robin assignment, the balanced one, z slicing, result paths,
devID), the exit status and the clean-up.  The GPU test runs the real HDoG3D_main through
the fan-out and compares its result files with a direct run."""
import importlib.util
import json
import os
import stat
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cubic-informatics_amd", "host")
SCRIPT = os.path.join(HOST, "hdog_gpu.py")

spec = importlib.util.spec_from_file_location("hdog_gpu", SCRIPT)
hdog_gpu = importlib.util.module_from_spec(spec)
spec.loader.exec_module(hdog_gpu)

# (Y, X, number of slices)
TREE = [("100000", "200000", 10), ("100000", "210000", 9), ("110000", "200000", 8),
        ("110000", "210000", 7), ("120000", "200000", 6), ("120000", "210000", 5),
        ("130000", "200000", 4)]


def make_tree(tmp_path, tree=TREE, gpus=(0, 1, 2)):
    src = os.path.join(str(tmp_path), "src")
    for y, x, n in tree:
        d = os.path.join(src, y, "{}_{}".format(y, x))
        os.makedirs(d)
        for z in range(n):
            open(os.path.join(d, "{:06d}.bin".format(z * 10)), "wb").close()
    params = {"gpus": list(gpus), "src_basedir": src,
              "dst_basedir": os.path.join(str(tmp_path), "dst"), "devID": 0,
              "HDoG_param": {"depth": 32}, "stacks": []}
    pfile = os.path.join(str(tmp_path), "param.json")
    json.dump(params, open(pfile, "w"))
    return params, pfile


def fake_exec(tmp_path, fail_dev=None):
    """An executable that copies its parameter file into tmp/seen/ (and fails for one GPU)."""
    seen = os.path.join(str(tmp_path), "seen")
    os.makedirs(seen, exist_ok=True)
    exe = os.path.join(str(tmp_path), "fake_hdog")
    body = "#!/bin/sh\ncp \"$1\" \"%s/$(basename \"$1\")\"\necho ran \"$1\"\n" % seen
    if fail_dev is not None:
        body += "case \"$1\" in *gpu%02d.json) exit 3;; esac\n" % fail_dev
    open(exe, "w").write(body)
    os.chmod(exe, os.stat(exe).st_mode | stat.S_IXUSR)
    return exe, seen


def loads(per, tree=TREE):
    n = {"{}_{}".format(y, x): k for y, x, k in tree}
    return {g: sum(n[os.path.basename(d)] for d in ds) for g, ds in per.items()}


def test_round_robin_matches_reference_assignment(tmp_path):
    params, _ = make_tree(tmp_path)
    stacks = hdog_gpu.list_stacks(params)
    assert [n for _, n in stacks] == [10, 9, 8, 7, 6, 5, 4]
    per, load = hdog_gpu.assign(stacks, [0, 1, 2], "rr")
    # script/HDoG_gpu.py:73-81: stack i of the length-sorted list goes to gpus[i % 3]
    for i, (d, _) in enumerate(stacks):
        assert d in per[[0, 1, 2][i % 3]]
    assert load == {0: 21, 1: 15, 2: 13} == loads(per)


def test_longest_first_balances_uneven_stacks(tmp_path):
    params, _ = make_tree(tmp_path)
    stacks = hdog_gpu.list_stacks(params)
    per, load = hdog_gpu.assign(stacks, [0, 1, 2], "lpt")
    assert load == loads(per) and sum(load.values()) == 49
    assert max(load.values()) == 19 < 21  # round-robin's maximum
    assert sorted(sum(per.values(), [])) == sorted(d for d, _ in stacks)
    with pytest.raises(ValueError):
        hdog_gpu.assign(stacks, [0], "random")


def test_fan_out_writes_per_gpu_params_and_runs_each_gpu(tmp_path):
    params, pfile = make_tree(tmp_path)
    exe, seen = fake_exec(tmp_path)
    rc = hdog_gpu.main([pfile, "--exec", exe, "--zmin", "1", "--zmax", "6"])
    assert rc == 0
    got = {}
    for g in params["gpus"]:
        p = json.load(open(os.path.join(seen, "param_gpu{:02d}.json".format(g))))
        assert p["devID"] == g and p["HDoG_param"] == params["HDoG_param"]
        for st in p["stacks"]:
            yx = os.path.basename(os.path.dirname(st["src_paths"][0]))
            assert st["dst_path"] == os.path.join(params["dst_basedir"], yx + ".bin")
            names = [os.path.basename(s) for s in st["src_paths"]]
            n = [k for y, x, k in TREE if "{}_{}".format(y, x) == yx][0]
            assert names == ["{:06d}.bin".format(z * 10) for z in range(n)][1:6]
            got[yx] = g
        assert os.path.exists(os.path.join(str(tmp_path), "log_gpu{:02d}.txt".format(g)))
        # the temporary parameter files are removed after the runs
        assert not os.path.exists(os.path.join(str(tmp_path), "param_gpu{:02d}.json".format(g)))
    assert len(got) == len(TREE) and os.path.isdir(params["dst_basedir"])


def test_single_stack_and_dry_run(tmp_path):
    params, pfile = make_tree(tmp_path, gpus=(3, 5))
    rc = hdog_gpu.main([pfile, "single", "210000", "110000", "--dry-run"])
    assert rc == 0
    p3 = json.load(open(os.path.join(str(tmp_path), "param_gpu03.json")))
    p5 = json.load(open(os.path.join(str(tmp_path), "param_gpu05.json")))
    assert [len(s["src_paths"]) for s in p3["stacks"]] == [7] and p5["stacks"] == []
    assert p3["stacks"][0]["dst_path"].endswith("110000_210000.bin")


def test_failed_run_gives_exit_1(tmp_path):
    params, pfile = make_tree(tmp_path, gpus=(0, 1))
    exe, _ = fake_exec(tmp_path, fail_dev=1)
    r = subprocess.run([sys.executable, SCRIPT, pfile, "--exec", exe], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "GPU:0\timages:" in r.stdout and "exit 3" in r.stdout


@pytest.mark.gpu
def test_fan_out_results_equal_oracle_records(tmp_path):
    # every stack's result file holds the records the oracle gives for it: the substack
    # loop, margins and record filter of src/3D/HDoG3D_main.cpp:294-315 and
    # src/3D/utils.h:204-250 applied to the oracle's regions of each substack
    import numpy as np
    from hdog import synth
    from test_host_program import DT_LOCAL, expected_records, make_param
    import oracle
    exe = os.path.join(HOST, "HDoG3D_main")
    if not os.path.exists(exe):
        pytest.skip("host program not built")
    src = os.path.join(str(tmp_path), "src")
    vols = {}
    for k, (y, x, L) in enumerate([("000000", "000000", 30), ("000000", "000010", 20)]):
        vol = synth.make_volume("c0", shape=(L, 40, 56), n_blobs=20, seed=7 + k)
        d = os.path.join(src, y, "{}_{}".format(y, x))
        os.makedirs(d)
        for z in range(L):
            vol[z].astype("<u2").tofile(os.path.join(d, "{:06d}.bin".format(z)))
        vols[(y, x)] = vol
    base = json.load(open(make_param(tmp_path, [], depth=24, W=56, H=40)))
    base.update(src_basedir=src, dst_basedir=os.path.join(str(tmp_path), "fan"), gpus=[0])
    pfile = os.path.join(str(tmp_path), "fan.json")
    json.dump(base, open(pfile, "w"))
    assert hdog_gpu.main([pfile, "--exec", exe]) == 0
    hp = base["HDoG_param"]
    sd = hp["sigma_dog"]
    p = dict(min_intensity=hp["min_intensity"], rxy=hp["radius_norm"]["xy_large"],
             rz=hp["radius_norm"]["z_large"], sigma_xy1=sd["xy_small"], sigma_z1=sd["z_small"],
             sigma_xy2=sd["xy_large"], sigma_z2=sd["z_large"], truncate=2.0, gamma=hp["gamma_n"])
    ii = base["input_image_info"]
    P = dict(depth=hp["depth"],
             margin=oracle.gaussian_radius(sd["z_large"], 2.0) + hp["extra_depth_margin"],
             left=ii["left_margin"], right=ii["right_margin"], top=ii["top_margin"],
             bottom=ii["bottom_margin"], min_size=hp["min_size"], p=p)
    for (y, x), vol in vols.items():
        got = np.fromfile(os.path.join(str(tmp_path), "fan", "{}_{}.bin".format(y, x)),
                          dtype=DT_LOCAL)
        exp = expected_records(vol, P)
        assert len(got) == len(exp) and len(exp) > 0, (y, x, len(got), len(exp))
        for k in ("local_x", "local_y", "local_z", "intensity", "size", "padding"):
            np.testing.assert_array_equal(got[k], exp[k], err_msg=k)
        for k in ("structureness", "blobness"):
            np.testing.assert_allclose(got[k], exp[k], rtol=1e-5, err_msg=k)
