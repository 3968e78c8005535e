"""The HDoG3D_main host program: CLI, parameter schema, substack loop and record format.

CPU tests cover what runs before the device is touched (usage, parameter and file
errors) and the substack / record logic restated from src/3D/HDoG3D_main.cpp:294-315
and src/3D/utils.h:204-250.  The GPU test runs the program end to end and compares its
result file with records built from the oracle's regions.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from hdog import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cubic-informatics_amd", "host", "HDoG3D_main")
# record layout of script/HalfBrainCells.py (dt_local)
DT_LOCAL = np.dtype([("local_x", "f4"), ("local_y", "f4"), ("local_z", "f4"),
                     ("structureness", "f4"), ("blobness", "f4"), ("intensity", "f4"),
                     ("size", "u2"), ("padding", "u2")])


def make_param(tmp_path, stacks, depth=32, W=64, H=48, margins=(2, 2, 2, 2), extra=1,
               rxy=3, rz=2, min_intensity=0, sigmas=(1.5, 1.5, 2.4, 2.4), truncate=None,
               min_size=0, precision=None):
    j = {
        "src_basedir": "/unused", "dst_basedir": "/unused", "gpus": [0],
        "devID": 0,
        "input_image_info": {"width": W, "height": H, "#": "comment",
                             "left_margin": margins[0], "right_margin": margins[1],
                             "top_margin": margins[2], "bottom_margin": margins[3]},
        "coordinate_info": {"scale_x": -0.65, "scale_y": -0.65, "scale_z": 5.0,
                            "affine_global": [[0, 0.1, 0, 0], [0.1, 0, 0, 0], [0, 0, 5, -1]]},
        "HDoG_param": {"depth": depth, "extra_depth_margin": extra, "cub_tmp_size_factor": 8.001,
                       "min_intensity_skip": 1000, "radius_norm": {"xy_large": rxy, "z_large": rz},
                       "min_intensity": min_intensity, "gamma_n": 1.0,
                       "sigma_dog": {"xy_small": sigmas[0], "z_small": sigmas[1],
                                     "xy_large": sigmas[2], "z_large": sigmas[3]},
                       "min_size": min_size},
        "stacks": stacks,
    }
    if truncate is not None:
        j["HDoG_param"]["truncate"] = truncate
    if precision is not None:
        j["HDoG_param"]["precision"] = precision
    path = os.path.join(str(tmp_path), "param.json")
    with open(path, "w") as f:
        json.dump(j, f)
    return path


def run(args):
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module", autouse=True)
def need_exe():
    if not os.path.exists(EXE):
        pytest.skip("HDoG3D_main not built")


def test_usage_exits_2():
    r = run([])
    assert r.returncode == 2 and "Usage" in r.stdout
    assert run(["a", "b"]).returncode == 2


def test_missing_param_file_exits_2(tmp_path):
    r = run([os.path.join(str(tmp_path), "nope.json")])
    assert r.returncode == 2 and "Unable to open file" in r.stdout


def test_bad_json_exits_2(tmp_path):
    p = os.path.join(str(tmp_path), "bad.json")
    open(p, "w").write('{"devID": 0, ')
    assert run([p]).returncode == 2


@pytest.mark.parametrize("section,key", [("HDoG_param", "sigma_dog"), ("input_image_info", "width"),
                                         (None, "stacks"), ("HDoG_param", "min_size")])
def test_missing_required_key_exits_2(tmp_path, section, key):
    path = make_param(tmp_path, [])
    j = json.load(open(path))
    if section:
        del j[section][key]
    else:
        del j[key]
    json.dump(j, open(path, "w"))
    r = run([path])
    assert r.returncode == 2 and key in r.stdout


def test_duplicate_comment_keys_are_accepted(tmp_path):
    # the reference's example files repeat the "#" key; the last one wins
    path = make_param(tmp_path, [])
    text = open(path).read().replace('"#": "comment"', '"#": "a", "#": "b"')
    open(path, "w").write(text)
    r = run([path])
    assert r.returncode in (0, 1), r.stdout   # 1 only when no device is present


# ---------------------------------------------------------------- substack / record logic
def substack_starts(L, depth, margin):
    """src/3D/HDoG3D_main.cpp:303-314."""
    return list(range(0, max(L - margin, 0), depth - 2 * margin)) if L - margin > 0 else []


def expected_records(vol, P, regions_of=None):
    """Records the reference writes for one stack, from the oracle's regions (or from the
    regions regions_of(substack) returns for each zero-filled substack)."""
    depth, margin = P["depth"], P["margin"]
    D, H, W = vol.shape
    out = []
    for z0 in substack_starts(D, depth, margin):
        d = min(depth, D - z0)
        sub = np.zeros((depth, H, W), np.uint16)
        sub[:d] = vol[z0:z0 + d]
        reg = regions_of(sub) if regions_of else oracle.substack(sub, P["p"])["regions"]
        c = reg["centroid"]
        for i in range(1, len(reg["label"])):
            cx, cy, cz = c[i]
            if cz < margin or cz >= d - margin:
                continue
            if cx < P["left"] or cx >= W - P["right"] or cy < P["top"] or cy >= H - P["bottom"]:
                continue
            if reg["size16"][i] < P["min_size"] or reg["norm"][i] <= 0:
                continue
            out.append((cx, cy, np.float32(cz + np.float32(z0)), reg["eig"][i, 0], reg["eig"][i, 1],
                        reg["norm"][i], reg["size16"][i], 0))
    return np.array(out, dtype=DT_LOCAL)


def test_substack_starts_match_reference_loop():
    assert substack_starts(100, 32, 8) == [0, 16, 32, 48, 64, 80]
    assert substack_starts(32, 32, 8) == [0, 16]
    assert substack_starts(8, 32, 8) == []
    assert substack_starts(9, 32, 8) == [0]


def test_record_dtype_is_28_bytes():
    assert DT_LOCAL.itemsize == 28


def test_unknown_precision_exits_2(tmp_path):
    param = make_param(tmp_path, [{"src_paths": [], "dst_path": "/unused/x.bin"}],
                       precision="fp8")
    r = run([param])
    assert r.returncode == 2
    assert "precision" in r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("L,depth,truncate", [(40, 24, 2.0), (24, 24, 4.0), (61, 32, 2.0)])
def test_host_program_end_to_end(tmp_path, L, depth, truncate):
    vol = synth.make_volume("c0", shape=(L, 48, 64), n_blobs=25, seed=L)
    paths = synth.write_stack(vol, os.path.join(str(tmp_path), "src"))
    dst = os.path.join(str(tmp_path), "out", "000000_000000.bin")
    os.makedirs(os.path.dirname(dst))
    open(dst, "wb").write(b"stale")  # removed first by the program
    sig = (1.5, 1.5, 2.4, 2.4)
    param = make_param(tmp_path, [{"src_paths": paths, "dst_path": dst}], depth=depth,
                       truncate=truncate, sigmas=sig, min_size=2)
    r = run([param])
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(dst, dtype=DT_LOCAL)
    rz2 = oracle.gaussian_radius(sig[3], truncate)
    P = dict(depth=depth, margin=rz2 + 1, left=2, right=2, top=2, bottom=2, min_size=2,
             p=dict(min_intensity=0, rxy=3, rz=2, sigma_xy1=sig[0], sigma_z1=sig[1],
                    sigma_xy2=sig[2], sigma_z2=sig[3], truncate=truncate, gamma=1.0))
    exp = expected_records(vol, P)
    assert len(got) == len(exp) and len(exp) > 0
    for k in ("local_x", "local_y", "local_z", "intensity", "size", "padding"):
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)
    for k in ("structureness", "blobness"):
        np.testing.assert_allclose(got[k], exp[k], rtol=1e-5, err_msg=k)


@pytest.mark.gpu
def test_host_program_several_stacks(tmp_path):
    # the substacks of all stacks run as one pipeline (the next stack is read while the
    # previous one runs, records are written by a writer thread): every stack's file holds
    # exactly its own records, and a stack too short for any substack has its stale result
    # removed and none written, as in the reference's loop
    sig, truncate = (1.5, 1.5, 2.4, 2.4), 2.0
    lengths = (40, 6, 29)
    vols, stacks = [], []
    for k, L in enumerate(lengths):
        if L > 10:
            vol = synth.make_volume("c0", shape=(L, 48, 64), n_blobs=25, seed=100 + k)
        else:  # too short for blobs away from the edges, and for any substack
            vol = np.random.default_rng(k).integers(150, 250, (L, 48, 64)).astype(np.uint16)
        paths = synth.write_stack(vol, os.path.join(str(tmp_path), f"src{k}"))
        dst = os.path.join(str(tmp_path), "out", f"stack{k}.bin")
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        open(dst, "wb").write(b"stale")
        vols.append(vol)
        stacks.append({"src_paths": paths, "dst_path": dst})
    param = make_param(tmp_path, stacks, depth=24, truncate=truncate, sigmas=sig, min_size=2)
    r = run([param])
    assert r.returncode == 0, r.stdout + r.stderr
    rz2 = oracle.gaussian_radius(sig[3], truncate)
    P = dict(depth=24, margin=rz2 + 1, left=2, right=2, top=2, bottom=2, min_size=2,
             p=dict(min_intensity=0, rxy=3, rz=2, sigma_xy1=sig[0], sigma_z1=sig[1],
                    sigma_xy2=sig[2], sigma_z2=sig[3], truncate=truncate, gamma=1.0))
    for k, vol in enumerate(vols):
        dst = stacks[k]["dst_path"]
        exp = expected_records(vol, P)
        if len(substack_starts(len(vol), 24, P["margin"])) == 0:
            assert not os.path.exists(dst), f"stack {k}: stale result kept"
            continue
        got = np.fromfile(dst, dtype=DT_LOCAL)
        assert len(got) == len(exp) and len(exp) > 0, (k, len(got), len(exp))
        for key in ("local_x", "local_y", "local_z", "intensity", "size", "padding"):
            np.testing.assert_array_equal(got[key], exp[key], err_msg=f"stack {k}: {key}")
        for key in ("structureness", "blobness"):
            np.testing.assert_allclose(got[key], exp[key], rtol=1e-5, err_msg=f"stack {k}: {key}")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp16", "bf16_dog32"])
def test_host_program_precision_key(tmp_path, precision):
    # HDoG_param.precision selects the library's reduced-precision form for every substack:
    # the program's records equal those the same form gives for each zero-filled substack
    # through the library's Python binding (same kernels, so equal bit for bit)
    from hdog import HDoG
    sig, truncate, depth, L = (1.5, 1.5, 2.4, 2.4), 2.0, 24, 40
    vol = synth.make_volume("c0", shape=(L, 48, 64), n_blobs=25, seed=7)
    paths = synth.write_stack(vol, os.path.join(str(tmp_path), "src"))
    dst = os.path.join(str(tmp_path), "out", "000000_000000.bin")
    os.makedirs(os.path.dirname(dst))
    param = make_param(tmp_path, [{"src_paths": paths, "dst_path": dst}], depth=depth,
                       truncate=truncate, sigmas=sig, min_size=2, precision=precision)
    r = run([param])
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(dst, dtype=DT_LOCAL)
    h = HDoG(0)
    prm = h.make_params((depth, 48, 64), 3, 2, 0, *sig, truncate=truncate, gamma=1.0,
                        precision=precision)

    def regions_of(sub):
        h.run_device(sub, prm)
        return h.download_regions()

    rz2 = oracle.gaussian_radius(sig[3], truncate)
    P = dict(depth=depth, margin=rz2 + 1, left=2, right=2, top=2, bottom=2, min_size=2)
    exp = expected_records(vol, P, regions_of)
    assert len(got) == len(exp) and len(exp) > 0
    for k in DT_LOCAL.names:
        np.testing.assert_array_equal(got[k], exp[k], err_msg=k)
    # and the reduced-precision form still finds nearly the FP32 cells
    P["p"] = dict(min_intensity=0, rxy=3, rz=2, sigma_xy1=sig[0], sigma_z1=sig[1],
                  sigma_xy2=sig[2], sigma_z2=sig[3], truncate=truncate, gamma=1.0)
    ref = expected_records(vol, P)
    near = sum(np.any((np.abs(ref["local_x"] - g["local_x"]) < 1)
                      & (np.abs(ref["local_y"] - g["local_y"]) < 1)
                      & (np.abs(ref["local_z"] - g["local_z"]) < 1)) for g in got)
    assert near >= 0.9 * len(ref), (near, len(ref), len(got))
