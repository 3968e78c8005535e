"""The step-runner programs (Normalize_test ... AverageNormalized_test): the reference's
test/*_test CLIs and output files, on libhdog.so.

CPU tests cover name dispatch and the parameter / stack checks that run before the device
is touched.  GPU tests run every runner on a small stack in a scratch directory and compare
each dumped file with the oracle, bit for bit where the parity bar is bit-exact."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from hdog import synth
from test_host_program import make_param

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cubic-informatics_amd", "host")
STEP = os.path.join(HOST, "hdog_step_runner")
NAMES = ["Normalize_test", "DoG_test", "Hessian_test", "CCL_test", "Regional_test",
         "MaxNormalized_test", "AverageNormalized_test"]


@pytest.fixture(scope="module", autouse=True)
def need_exe():
    if not os.path.exists(STEP):
        pytest.skip("step runners not built")


def runner(name, args, cwd):
    exe = os.path.join(HOST, "test", name)
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=300, cwd=cwd)


def stack(tmp_path, shape=(24, 40, 56), seed=5, n_blobs=30):
    vol = synth.make_volume("c0", shape=shape, n_blobs=n_blobs, seed=seed)
    paths = synth.write_stack(vol, os.path.join(str(tmp_path), "src"))
    return vol, paths


def test_every_runner_is_installed():
    for n in NAMES:
        assert os.access(os.path.join(HOST, "test", n), os.X_OK), n


def test_unknown_name_and_usage_exit_2(tmp_path):
    r = subprocess.run([STEP, "Nope_test", "x"], capture_output=True, text=True)
    assert r.returncode == 2 and "Normalize_test" in r.stdout
    r = subprocess.run([os.path.join(HOST, "test", "DoG_test")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stdout


def test_single_stack_of_depth_slices_is_required(tmp_path):
    vol, paths = stack(tmp_path, shape=(10, 16, 16))
    two = [{"src_paths": paths, "dst_path": ""}] * 2
    p = make_param(tmp_path, two, depth=10, W=16, H=16)
    r = runner("Normalize_test", [p], str(tmp_path))
    assert r.returncode == 2 and "A single stack should be provided!" in r.stdout
    p = make_param(tmp_path, [{"src_paths": paths[:9], "dst_path": ""}], depth=10, W=16, H=16)
    r = runner("Regional_test", [p], str(tmp_path))
    assert r.returncode == 2 and "depth(=9) should be 10" in r.stdout


def test_missing_slice_exits_2(tmp_path):
    vol, paths = stack(tmp_path, shape=(10, 16, 16))
    paths[3] = os.path.join(str(tmp_path), "missing.bin")
    p = make_param(tmp_path, [{"src_paths": paths, "dst_path": ""}], depth=10, W=16, H=16)
    r = runner("CCL_test", [p], str(tmp_path))
    assert r.returncode == 2 and "Unable to open" in r.stdout


# ------------------------------------------------------------------------- GPU
P = dict(min_intensity=0, rxy=3, rz=2, sigma_xy1=1.5, sigma_z1=1.5, sigma_xy2=2.4,
         sigma_z2=2.4, truncate=2.0, gamma=1.0)


def gpu_case(tmp_path, verbose, is_ave=False):
    import json
    vol, paths = stack(tmp_path)
    D, H, W = vol.shape
    p = make_param(tmp_path, [{"src_paths": paths, "dst_path": ""}], depth=D, W=W, H=H,
                   rxy=P["rxy"], rz=P["rz"], min_intensity=P["min_intensity"],
                   sigmas=(P["sigma_xy1"], P["sigma_z1"], P["sigma_xy2"], P["sigma_z2"]))
    j = json.load(open(p))
    j["verbose"] = verbose
    if is_ave:
        j["is_ave_mode"] = True
    json.dump(j, open(p, "w"))
    o = oracle.substack(vol, dict(P, is_ave_mode=is_ave))
    return vol, p, o


def load(path, dtype, n=None):
    a = np.fromfile(path, dtype=dtype)
    if n is not None:
        assert a.size == n, (path, a.size, n)
    return a


@pytest.mark.gpu
def test_normalize_dog_hessian_runners_match_oracle(tmp_path):
    vol, p, o = gpu_case(tmp_path, verbose=True)
    N = vol.size
    d = str(tmp_path)
    r = runner("Normalize_test", [p, "e.bin", "d.bin", "n.bin"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    norm, E, Dl = oracle.normalize3d(vol, P["min_intensity"], P["rxy"], P["rz"], return_ed=True)
    np.testing.assert_array_equal(load(os.path.join(d, "e.bin"), "<f4", N), E.ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "d.bin"), "<f4", N), Dl.ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "n.bin"), "<f4", N), o["norm"].ravel())
    r = runner("DoG_test", [p], d)  # verbose: dog.bin and normalized.bin
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sigma_xy1: 1.5" in r.stdout
    np.testing.assert_array_equal(load(os.path.join(d, "dog.bin"), "<f4", N), o["dog"].ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "normalized.bin"), "<f4", N),
                                  o["norm"].ravel())
    r = runner("Hessian_test", [p, "g.bin", "pd.bin", "h.bin"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    np.testing.assert_array_equal(load(os.path.join(d, "g.bin"), "<f4", N), o["dog"].ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "pd.bin"), "u1", N), o["pd"].ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "h.bin"), "<f4", 6 * N), o["hess"].ravel())


@pytest.mark.gpu
def test_ccl_and_regional_runners_match_oracle(tmp_path):
    vol, p, o = gpu_case(tmp_path, verbose=False)
    N = vol.size
    d = str(tmp_path)
    r = runner("CCL_test", [p], d)
    assert r.returncode == 0, r.stdout + r.stderr
    np.testing.assert_array_equal(load(os.path.join(d, "labels.bin"), "<u4", N),
                                  o["labels"].ravel())
    np.testing.assert_array_equal(load(os.path.join(d, "hessian_pd.bin"), "u1", N), o["pd"].ravel())
    r = runner("Regional_test", [p], d)
    assert r.returncode == 0, r.stdout + r.stderr
    g = o["regions"]
    R = len(g["label"])
    assert f"num_regions: {R}" in r.stdout
    n = R - 1
    np.testing.assert_array_equal(load(os.path.join(d, "labels_region.bin"), "<i4", n),
                                  g["label"][1:].astype(np.int32))
    np.testing.assert_array_equal(load(os.path.join(d, "size_region.bin"), "<u2", n),
                                  g["size16"][1:])
    np.testing.assert_array_equal(load(os.path.join(d, "max_normalized_region.bin"), "<f4", n),
                                  g["norm"][1:])
    for k, name in enumerate(("centroid_x", "centroid_y", "centroid_z")):
        np.testing.assert_array_equal(load(os.path.join(d, name + ".bin"), "<f4", n),
                                      g["centroid"][1:, k])
    for k, name in enumerate(("structureness_region", "blobness_region")):
        np.testing.assert_allclose(load(os.path.join(d, name + ".bin"), "<f4", n),
                                   g["eig"][1:, k], rtol=1e-5, atol=1e-30)
    np.testing.assert_array_equal(load(os.path.join(d, "labels.bin"), "<u4", N),
                                  o["labels"].ravel())
    r = runner("MaxNormalized_test", [p, "mx.bin"], d)
    assert r.returncode == 0, r.stdout + r.stderr
    np.testing.assert_array_equal(load(os.path.join(d, "mx.bin"), "<f4", n), g["norm"][1:])


@pytest.mark.gpu
def test_average_normalized_runner_matches_oracle(tmp_path):
    vol, p, o = gpu_case(tmp_path, verbose=False, is_ave=True)
    d = str(tmp_path)
    r = runner("AverageNormalized_test", [p], d)
    assert r.returncode == 0, r.stdout + r.stderr
    g = o["regions"]
    n = len(g["label"]) - 1
    got = load(os.path.join(d, "average_normalized_region.bin"), "<f4", n)
    # sum of normalized intensity (FP64 sum rounded to FP32) / u16 size, in FP32
    want = (g["norm"][1:] / g["size16"][1:].astype(np.float32)).astype(np.float32)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=0)
