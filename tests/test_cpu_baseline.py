"""The CPU baseline: the scipy counterpart of the reference's CPU detector
(oracle/hdog_cpu_scipy.py, script/HDoG_cpu.py:57-146).  Its slab form (filter stages in
worker processes, each slab with a halo of 2 r_norm + r_dog + 2 slices) must give exactly
the single-process result along every axis, so the all-cores timing measures the same
work; a halo one slice short must not."""
import numpy as np
import pytest

import hdog_cpu_scipy as hc
from hdog import synth


@pytest.fixture(scope="module")
def case():
    vol = synth.make_volume("c1", shape=(40, 72, 90), n_blobs=60, seed=5)
    det = hc.HDoGCPU((1.6, 2.0, 2.0), (2.56, 3.2, 3.2), 4, 3, 0, 1.0, 4.0)
    return vol, det, det.run(vol)


def same(a, b):
    return a["n"] == b["n"] and all(np.array_equal(a[k], b[k], equal_nan=True) for k in a if k != "n")


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_slabs_equal_single_process(case, axis):
    vol, det, ref = case
    assert ref["n"] > 20
    assert same(det.run_slabs(vol, 3, axis), ref)


def test_halo_is_needed(case, monkeypatch):
    vol, det, ref = case
    h = det.halo(0)
    monkeypatch.setattr(det, "halo", lambda axis=0: h - 1)
    assert not same(det.run_slabs(vol, 3, 0), ref)


def test_stage_semantics_on_known_answers():
    det = hc.HDoGCPU((1.0, 1.0, 1.0), (1.5, 1.5, 1.5), 1, 1, 0, 1.0, 2.0)
    # norm = I / (D - E) with D, E box means (zero padding) of the max / min filter
    img = np.zeros((5, 5, 5), np.uint16)
    img[2, 2, 2] = 270
    norm = det.normalize(img)
    # around the centre the max filter is 270 on the 3^3 cube, so D(centre) = 270 and E = 0
    assert norm[2, 2, 2] == np.float32(270) / np.float32(270)
    # negative definite Hessian of a centred bump
    z, y, x = np.mgrid[-4:5, -4:5, -4:5].astype(np.float32)
    pd, hess = det.hessian_pd(-(x * x + 2 * y * y + 3 * z * z))
    assert pd[4, 4, 4] and hess[0][4, 4, 4] == -2 and hess[3][4, 4, 4] == -4
    # 26-connected labels: diagonal neighbours join
    m = np.zeros((3, 3, 3), bool)
    m[0, 0, 0] = m[1, 1, 1] = m[2, 2, 2] = True
    assert det.ccl(m)[1] == 1
