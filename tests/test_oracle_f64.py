"""Pins the FP64 restatement of the oracle (the reference of the reduced-precision storage
forms): each function against a float64 library routine or a closed form, and the FP32
oracle against it within FP32 rounding."""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle
from hdog import synth

RNG = np.random.default_rng(64)


@pytest.mark.parametrize("sigma,trunc", [(2.0, 4.0), (3.2, 4.0), (1.6, 2.0), (2.56, 4.0)])
def test_weights_are_scipys_gaussian(sigma, trunc):
    s = float(np.float32(sigma))
    w = oracle.gaussian_weights_f64(s, trunc)
    r = len(w) // 2
    assert r == int(trunc * s + 0.5)
    imp = np.zeros(4 * r + 1)
    imp[2 * r] = 1.0
    ref = ndi.gaussian_filter1d(imp, s, truncate=trunc, mode="constant")[r:3 * r + 1]
    np.testing.assert_allclose(w, ref, rtol=1e-14, atol=1e-17)
    assert abs(w.sum() - 1.0) < 1e-14


def test_gaussian3d_is_scipys_gaussian_filter():
    vol = RNG.random((17, 23, 29))
    wxy, wz = oracle.gaussian_weights_f64(2.0, 4.0), oracle.gaussian_weights_f64(1.6, 4.0)
    got = oracle.gaussian3d_f64(vol, wxy, wz)
    ref = ndi.gaussian_filter(vol, (float(np.float32(1.6)), 2.0, 2.0), truncate=4.0,
                              mode="constant", cval=0.0)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)


def test_normalize_is_scipys_erosion_dilation_and_box():
    img = RNG.integers(0, 3000, (14, 30, 33), dtype=np.uint16)
    img[:, 5, :] = 65535
    size = (2 * 2 + 1, 2 * 4 + 1, 2 * 4 + 1)
    E = ndi.uniform_filter(ndi.grey_erosion(img, size=size, mode="nearest").astype(np.float64),
                           size=size, mode="nearest")
    Dl = ndi.uniform_filter(ndi.grey_dilation(img, size=size, mode="nearest").astype(np.float64),
                            size=size, mode="nearest")
    I = img.astype(np.float64)
    want = np.where(I >= 100, (I - E) / (Dl - E), 0.0)
    got = oracle.normalize3d_f64(img, 100, 4, 2)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # radius 0: the thresholded copy
    np.testing.assert_array_equal(oracle.normalize3d_f64(img, 100, 0, 2), np.where(I >= 100, I, 0))


def test_hessian_gradients_and_minors():
    dog = RNG.standard_normal((11, 13, 15))
    hess, pd = oracle.hessian3d_f64(dog)
    # interior: numpy's central differences (np.gradient) applied twice
    gz, gy, gx = np.gradient(dog)
    gxz, gxy, gxx = np.gradient(gx)
    gyz, gyy, _ = np.gradient(gy)
    gzz = np.gradient(gz)[0]
    inner = (slice(2, -2),) * 3
    for k, ref in enumerate((gxx, gxy, gxz, gyy, gyz, gzz)):
        np.testing.assert_allclose(hess[k][inner], ref[inner], rtol=0, atol=1e-14)
    # zero padding at the border: the first difference of a field of ones is 0.5, 0, 0, 0,
    # -0.5 along a line of 5, so the second is 0, -0.25, 0, -0.25, 0
    h1, _ = oracle.hessian3d_f64(np.ones((5, 5, 5)))
    np.testing.assert_array_equal(h1[0][2, 2], [0.0, -0.25, 0.0, -0.25, 0.0])
    # a quadratic has constant second derivatives away from the border
    z, y, x = np.mgrid[0:9, 0:9, 0:9].astype(np.float64)
    hq, pq = oracle.hessian3d_f64(-(x * x) - 2 * y * y - 3 * z * z + x * y)
    c = (slice(2, -2),) * 3
    for k, v in enumerate((-2.0, 1.0, 0.0, -4.0, 0.0, -6.0)):
        assert np.all(hq[k][c] == v)
    assert np.all(pq[c] == 1)
    assert oracle.negdef_f64(-1, 0, 0, -1, 0, -1) == 1
    assert oracle.negdef_f64(-1, 0, 0, -1, 0, 1) == 0
    assert oracle.negdef_f64(-1, 0, 0, 1, 0, -1) == 0
    assert oracle.negdef_f64(1, 0, 0, -1, 0, -1) == 0


def test_fp32_oracle_is_fp64_within_fp32_rounding():
    vol = synth.make_volume("c1", shape=(40, 96, 110), n_blobs=40, seed=6)
    p = dict(min_intensity=0, rxy=10, rz=10, sigma_xy1=2.0, sigma_z1=1.6, sigma_xy2=3.2,
             sigma_z2=2.56, truncate=4.0, gamma=1.0)
    o32, o64 = oracle.substack(vol, p), oracle.substack_f64(vol, p)
    rel = lambda a, b: np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
    assert rel(o32["norm"], o64["norm"]) < 4e-6
    assert rel(o32["dog"], o64["dog"]) < 2e-5
    assert rel(o32["hess"], o64["hess"]) < 1e-4
    a, b = o32["pd"].astype(bool), o64["pd"].astype(bool)
    assert (a & b).sum() / (a | b).sum() > 0.9999
    # the FP64 region features of the FP32 oracle's own labels and Hessian match its FP32
    # features to FP32 precision
    r64 = oracle.regions_f64(o32["labels"], o32["hess"].astype(np.float64), o32["norm"])
    r32 = o32["regions"]
    np.testing.assert_array_equal(r64["label"], r32["label"][1:])
    np.testing.assert_array_equal(r64["count"], r32["count"][1:].astype(np.float64))
    np.testing.assert_allclose(r64["structureness"], r32["eig"][1:, 0], rtol=1e-5)
    np.testing.assert_array_equal(r64["norm"], r32["norm"][1:])
