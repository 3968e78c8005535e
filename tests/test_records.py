"""Downstream consumers of the candidate records (hdog.records): the record layout, the
per-stack placement in the half-brain frame, the scale-merge coordinates and the
classifier's log-features.

CPU tests pin each transform with known answers.  The GPU test runs HDoG3D_main on a
stack and checks that the classifier features of its records equal those of the records
derived from the oracle (exact for intensity, within the FP32 feature tolerance for
structureness and blobness)."""
import os

import numpy as np
import pytest

from hdog import records


def recs(rows):
    return np.array(rows, dtype=records.DT_LOCAL)


def test_layout_matches_the_host_program_record():
    # float x, y, z, structureness, blobness, intensity; u16 size, padding: 28 bytes
    dt = records.DT_LOCAL
    assert dt.itemsize == 28
    assert [dt.fields[k][1] for k in dt.names] == [0, 4, 8, 12, 16, 20, 24, 26]
    raw = np.zeros(2, dtype=dt)
    raw["size"] = [3, 65535]
    raw["local_z"] = [1.5, -2.0]
    b = raw.tobytes()
    assert len(b) == 56 and b[24:26] == (3).to_bytes(2, "little")
    assert np.frombuffer(b[8:12], "<f4")[0] == 1.5


def test_read_records_round_trip_and_missing_file(tmp_path):
    r = recs([(1, 2, 3, 4, 5, 6, 7, 0), (8, 9, 10, 11, 12, 13, 14, 0)])
    p = os.path.join(str(tmp_path), "s.bin")
    r.tofile(p)
    np.testing.assert_array_equal(records.read_records(p), r)
    assert records.read_records(p + ".none").shape == (0,)


def test_to_global_offsets_scales_and_flips():
    r = recs([(10, 20, 3, 0, 0, 0, 1, 0), (0, 0, 0, 0, 0, 0, 1, 0)])
    g = records.to_global(r, (100.0, 200.0, 50.0), (0.5, 2.0, 5.0), (64, 48), True, (2, 3))
    np.testing.assert_array_equal(g["merged_x"], [105.0, 100.0])
    np.testing.assert_array_equal(g["merged_y"], [240.0, 200.0])
    np.testing.assert_array_equal(g["merged_z"], [65.0, 50.0])
    assert list(g["i_x"]) == [2, 2] and list(g["i_y"]) == [3, 3] and list(g["FWRV"]) == [0, 0]
    # a negative scale mirrors the axis across the image: x -> offset + (W - x) * |s|
    f = records.to_global(r, (100.0, 200.0, 50.0), (-0.5, -2.0, 5.0), (64, 48), False)
    np.testing.assert_array_equal(f["merged_x"], [100 + (64 - 10) * 0.5, 100 + 64 * 0.5])
    np.testing.assert_array_equal(f["merged_y"], [200 + (48 - 20) * 2.0, 200 + 48 * 2.0])
    assert list(f["FWRV"]) == [1, 1]


def test_scale_merge_bounds_shift_and_flip():
    g = np.zeros(4, dtype=records.DT_GLOBAL)
    g["merged_x"] = [10, 20, 30, 40]
    g["merged_y"] = [1, 2, 3, 4]
    g["merged_z"] = [5, 10, 15, 20]
    eye = np.eye(3)
    out = records.scale_merge(g, (10, 15), 1.0, 0.5, (10, 0, 5), (0.5, 2.0, 0.25), eye, np.zeros(3))
    assert list(out["is_valid"]) == [False, True, True, False]  # inclusive bounds
    np.testing.assert_array_equal(out["scaled_x"], [-1.0, 4.0, 9.0, 14.0])
    np.testing.assert_array_equal(out["scaled_y"], [1.5, 3.5, 5.5, 7.5])
    np.testing.assert_array_equal(out["scaled_z"], [0.0, 1.25, 2.5, 3.75])
    # swapped bounds mean the same range; a flip in x about a merged width of 100
    A = np.diag([-1.0, 1.0, 1.0])
    out2 = records.scale_merge(g, (15, 10), 1.0, 0.5, (10, 0, 5), (0.5, 2.0, 0.25), A,
                               np.array([100.0, 0, 0]))
    assert list(out2["is_valid"]) == list(out["is_valid"])
    np.testing.assert_array_equal(out2["scaled_x"], 100.0 - out["scaled_x"])


def test_get_X_3d_log_features():
    r = recs([(0, 0, 0, 1000.0, 0.01, 100.0, 4, 0), (0, 0, 0, 1.0, 0.0, 10.0, 5, 0)])
    X = records.get_X_3d(r)
    assert X.shape == (2, 3)
    np.testing.assert_allclose(X[0], [2.0, 3.0, -2.0], rtol=0, atol=1e-6)
    # log10(0) = -inf is mapped to the most negative finite value of the dtype
    assert X[1, 2] == np.finfo(X.dtype).min and X[1, 0] == pytest.approx(1.0)
    Xa = records.get_X_3d(r, is_ave=True)
    np.testing.assert_allclose(Xa[:, 0], np.log10([25.0, 2.0]), rtol=1e-6)
    Xb = records.get_X_3d(r, bias=np.array([1.0, 0, 0]), scale=np.array([2.0, 1, 1]),
                          bias_before_log=np.array([0.0, 0, 1.0]))
    np.testing.assert_allclose(Xb[0], [6.0, 3.0, np.log10(1.01)], rtol=1e-6)


@pytest.mark.gpu
def test_classifier_features_of_pipeline_records_match_the_oracle(tmp_path):
    import subprocess
    import oracle
    from hdog import synth
    from test_host_program import EXE, make_param, expected_records
    if not os.path.exists(EXE):
        pytest.skip("host program not built")
    vol = synth.make_volume("c0", shape=(28, 48, 64), n_blobs=30, seed=11)
    paths = synth.write_stack(vol, os.path.join(str(tmp_path), "src"))
    dst = os.path.join(str(tmp_path), "cells.bin")
    sig = (1.5, 1.5, 2.4, 2.4)
    param = make_param(tmp_path, [{"src_paths": paths, "dst_path": dst}], depth=28,
                       truncate=2.0, sigmas=sig)
    r = subprocess.run([EXE, param], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = records.read_records(dst)
    rz2 = oracle.gaussian_radius(sig[3], 2.0)
    P = dict(depth=28, margin=rz2 + 1, left=2, right=2, top=2, bottom=2, min_size=0,
             p=dict(min_intensity=0, rxy=3, rz=2, sigma_xy1=sig[0], sigma_z1=sig[1],
                    sigma_xy2=sig[2], sigma_z2=sig[3], truncate=2.0, gamma=1.0))
    exp = expected_records(vol, P).view(records.DT_LOCAL)
    assert len(got) == len(exp) > 0
    Xg, Xe = records.get_X_3d(got), records.get_X_3d(exp)
    np.testing.assert_array_equal(Xg[:, 0], Xe[:, 0])
    # structureness / blobness within 1e-5 relative -> log10 within ~5e-6 absolute
    np.testing.assert_allclose(Xg[:, 1:], Xe[:, 1:], rtol=0, atol=1e-5)
    g = records.to_global(got, (0.0, 0.0, 0.0), (-0.65, -0.65, 5.0), (64, 48))
    e = records.to_global(exp, (0.0, 0.0, 0.0), (-0.65, -0.65, 5.0), (64, 48))
    for k in ("merged_x", "merged_y", "merged_z"):
        np.testing.assert_array_equal(g[k], e[k])
