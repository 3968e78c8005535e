"""Pin the CPU oracle against routes that do not share its code.

Each stage is checked against a library routine (scipy.ndimage / numpy.linalg),
a closed form, or brute force on tiny inputs, chosen so that a dropped term, a
wrong sign, a wrong border rule or a transposed axis fails at least one check.
The reference holds no golden vectors for this path (its step runners dump
files without comparing them), so these are the pins.
"""
import collections

import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle
from hdog import synth

RNG = np.random.default_rng(1234)


def rand_u16(shape, lo=0, hi=4000, rng=RNG):
    return rng.integers(lo, hi, size=shape, dtype=np.uint16)


# ---------------------------------------------------------------- a1 min/max
@pytest.mark.parametrize("shape,rxy,rz", [((7, 9, 11), 1, 1), ((12, 17, 23), 3, 2),
                                          ((5, 6, 40), 10, 10), ((9, 8, 7), 2, 0)])
@pytest.mark.parametrize("is_min", [True, False])
def test_minmax_matches_scipy_grey_morphology(shape, rxy, rz, is_min):
    img = rand_u16(shape)
    size = (2 * rz + 1, 2 * rxy + 1, 2 * rxy + 1)
    fn = ndi.grey_erosion if is_min else ndi.grey_dilation
    ref = fn(img, size=size, mode="nearest")
    np.testing.assert_array_equal(oracle.minmax3d(img, rxy, rz, is_min), ref)


def test_minmax_brute_force_tiny():
    img = rand_u16((4, 5, 6))
    r, rz = 2, 1
    out = oracle.minmax3d(img, r, rz, True)
    D, H, W = img.shape
    for z in range(D):
        for y in range(H):
            for x in range(W):
                zs = np.clip(np.arange(z - rz, z + rz + 1), 0, D - 1)
                ys = np.clip(np.arange(y - r, y + r + 1), 0, H - 1)
                xs = np.clip(np.arange(x - r, x + r + 1), 0, W - 1)
                assert out[z, y, x] == img[np.ix_(zs, ys, xs)].min()


# ---------------------------------------------------------------- a2 box mean
@pytest.mark.parametrize("shape,rxy,rz", [((6, 10, 12), 2, 1), ((11, 13, 17), 10, 10),
                                          ((8, 30, 25), 5, 3)])
def test_box_matches_scipy_uniform_filter(shape, rxy, rz):
    img = rand_u16(shape, 0, 65535)
    size = (2 * rz + 1, 2 * rxy + 1, 2 * rxy + 1)
    ref = ndi.uniform_filter(img.astype(np.float64), size=size, mode="nearest")
    out = oracle.box3d(img, rxy, rz).astype(np.float64)
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=0)


def test_box_x_pass_is_exact_integer_window_sum():
    # H = D = 1: the x pass sums integers (exact in any order): a = fl32(S) * fl32(1/(2r+1))
    # with S the exact clamped window sum.  The y pass (same radius, H = 1) then adds 2r+1
    # clamped copies of a; the z pass with r = 0 multiplies by fl(1/1) = 1.
    img = rand_u16((1, 1, 300), 0, 65535)
    r = 10
    out = oracle.box3d(img, r, 0)
    x = np.arange(300)
    idx = np.clip(x[:, None] + np.arange(-r, r + 1)[None, :], 0, 299)
    S = img[0, 0][idx].astype(np.int64).sum(axis=1)
    k = np.float32(1.0) / np.float32(2 * r + 1)
    a = S.astype(np.float32) * k
    acc = np.zeros_like(a)
    for _ in range(2 * r + 1):
        acc = (acc + a).astype(np.float32)
    np.testing.assert_array_equal(out[0, 0], acc * k)


def test_box_constant_volume_is_uniform():
    img = np.full((9, 14, 15), 777, np.uint16)
    out = oracle.box3d(img, 3, 2)
    assert np.all(out == out.flat[0]) and abs(out.flat[0] - 777) < 1e-3


# ---------------------------------------------------------------- a3 normalize
def test_normalize_closed_form_and_threshold():
    img = rand_u16((10, 24, 26), 100, 3000)
    rxy, rz, m = 4, 2, 1500
    norm, E, Dl = oracle.normalize3d(img, m, rxy, rz, return_ed=True)
    size = (2 * rz + 1, 2 * rxy + 1, 2 * rxy + 1)
    Er = ndi.uniform_filter(ndi.grey_erosion(img, size=size, mode="nearest").astype(np.float64),
                            size=size, mode="nearest")
    Dr = ndi.uniform_filter(ndi.grey_dilation(img, size=size, mode="nearest").astype(np.float64),
                            size=size, mode="nearest")
    np.testing.assert_allclose(E, Er, rtol=2e-6)
    np.testing.assert_allclose(Dl, Dr, rtol=2e-6)
    I = img.astype(np.float32)
    below = I < m
    assert np.all(norm[below] == 0)
    # exact per-voxel formula on the oracle's own E and D (IEEE division)
    np.testing.assert_array_equal(norm[~below], ((I - E) / (Dl - E))[~below])
    ref = (img.astype(np.float64) - Er) / (Dr - Er)
    np.testing.assert_allclose(norm[~below], ref[~below], rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("rxy,rz", [(0, 3), (3, 0)])
def test_normalize_zero_radius_is_thresholded_copy(rxy, rz):
    img = rand_u16((4, 8, 8), 0, 3000)
    out = oracle.normalize3d(img, 1000, rxy, rz)
    ref = np.where(img >= 1000, img.astype(np.float32), 0).astype(np.float32)
    np.testing.assert_array_equal(out, ref)


# ---------------------------------------------------------------- a4 weights
@pytest.mark.parametrize("sigma,trunc", [(2.0, 2.0), (3.2, 4.0), (1.6, 4.0), (2.56, 4.0),
                                         (4.0, 2.0), (6.0, 2.0), (0.3, 2.0)])
def test_gaussian_weights_match_scipy_kernel(sigma, trunc):
    from scipy.ndimage._filters import _gaussian_kernel1d
    r = oracle.gaussian_radius(sigma, trunc)
    assert r == int(trunc * float(np.float32(sigma)) + 0.5)
    w = oracle.gaussian_weights(sigma, trunc)
    ref = _gaussian_kernel1d(float(np.float32(sigma)), 0, r)
    assert len(w) == 2 * r + 1
    np.testing.assert_allclose(w, ref, rtol=3e-7, atol=0)
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    np.testing.assert_array_equal(w, w[::-1])


def test_dog_normalizer_values():
    assert oracle.dog_normalizer(2.0, 2.0, 1.0) == np.float32(8.0)
    ref = np.float32(np.float32(2.0) ** np.float32(2.0)) * np.float32(1.6)
    assert abs(oracle.dog_normalizer(2.0, 1.6, 1.0) - ref) <= 1e-6


# ---------------------------------------------------------------- a5/a6 Gaussian + DoG
def test_gaussian_impulse_response_is_rounded_outer_product():
    D, H, W = 15, 17, 19
    vol = np.zeros((D, H, W), np.float32)
    vol[7, 8, 9] = 1.0
    wxy = oracle.gaussian_weights(2.0, 2.0)
    wz = oracle.gaussian_weights(1.5, 2.0)
    out = oracle.gaussian3d(vol, wxy, wz)
    rxy, rz = len(wxy) // 2, len(wz) // 2
    ref = np.zeros_like(vol)
    yx = (wxy[:, None] * wxy[None, :]).astype(np.float32)        # column pass: fl(wy*wx)
    zyx = (wz[:, None, None] * yx[None]).astype(np.float32)       # layer pass: fl(wz*.)
    ref[7 - rz:7 + rz + 1, 8 - rxy:8 + rxy + 1, 9 - rxy:9 + rxy + 1] = zyx
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("trunc", [2.0, 4.0])
def test_dog_matches_scipy_gaussian_filter(trunc):
    vol = synth.make_volume("c0", shape=(24, 40, 36), n_blobs=30, seed=5)
    norm = oracle.normalize3d(vol, 0, 4, 3)
    s = (2.0, 2.0, 3.2, 3.2)
    dog, rz2 = oracle.dog3d(norm, *s, truncate=trunc, gamma=1.0)
    assert rz2 == int(trunc * 3.2 + 0.5)
    n64 = norm.astype(np.float64)
    g1 = ndi.gaussian_filter(n64, sigma=(s[1], s[0], s[0]), truncate=trunc, mode="constant", cval=0)
    g2 = ndi.gaussian_filter(n64, sigma=(s[3], s[2], s[2]), truncate=trunc, mode="constant", cval=0)
    ref = (g1 - g2) * 8.0
    scale = np.abs(n64).max() * 8.0
    assert np.max(np.abs(dog - ref)) < 2e-6 * scale


def test_gaussian_of_constant_interior_is_constant():
    vol = np.full((30, 30, 30), 3.0, np.float32)
    w = oracle.gaussian_weights(2.0, 2.0)
    out = oracle.gaussian3d(vol, w, w)
    inner = out[5:-5, 5:-5, 5:-5]
    assert np.all(inner == inner.flat[0]) and abs(inner.flat[0] - 3.0) < 1e-5


# ---------------------------------------------------------------- a7/a8 Hessian + PD
def test_hessian_interior_equals_numpy_gradient_bitwise():
    dog = RNG.standard_normal((10, 12, 14)).astype(np.float32)
    hess, _ = oracle.hessian3d(dog)
    # reference pass order (Hessian3DFilter_element.cu:373-383): Hxy = g_y(g_x f),
    # Hxz = g_z(g_x f), Hyz = g_z(g_y f); the other operand order rounds differently
    Hz, Hy, Hx = np.gradient(dog)
    Hxz, Hxy, Hxx = np.gradient(Hx)
    Hyz, Hyy = np.gradient(Hy, axis=(0, 1))
    Hzz = np.gradient(Hz, axis=0)
    inner = (slice(2, -2),) * 3
    for k, ref in enumerate([Hxx, Hxy, Hxz, Hyy, Hyz, Hzz]):
        np.testing.assert_array_equal(hess[k][inner], ref[inner].astype(np.float32))


def test_hessian_zero_padding_at_borders():
    dog = RNG.standard_normal((3, 4, 5)).astype(np.float32)
    hess, _ = oracle.hessian3d(dog)
    p = np.pad(dog, 1)                                   # zero padding
    Hx = ((p[1:-1, 1:-1, 2:] - p[1:-1, 1:-1, :-2]) * np.float32(0.5)).astype(np.float32)
    q = np.pad(Hx, 1)
    Hxx = (q[1:-1, 1:-1, 2:] - q[1:-1, 1:-1, :-2]) * np.float32(0.5)
    Hxy = (q[1:-1, 2:, 1:-1] - q[1:-1, :-2, 1:-1]) * np.float32(0.5)
    np.testing.assert_array_equal(hess[0], Hxx)
    np.testing.assert_array_equal(hess[1], Hxy)


def test_hessian_of_quadratic_is_constant():
    z, y, x = np.mgrid[0:12, 0:12, 0:12].astype(np.float64)
    f = (-1.0 * x * x - 2.0 * y * y - 0.5 * z * z + 0.25 * x * y + 0.125 * y * z).astype(np.float32)
    hess, pd = oracle.hessian3d(f)
    inner = (slice(2, -2),) * 3
    # 5-point wide second difference of x^2 is exactly 2; of x*y exactly 1
    expect = {0: -2.0, 1: 0.25, 2: 0.0, 3: -4.0, 4: 0.125, 5: -1.0}
    for k, v in expect.items():
        np.testing.assert_allclose(hess[k][inner], v, atol=1e-4)
    assert np.all(pd[inner] == 1)


@pytest.mark.parametrize("h,expect", [((-1, 0, 0, -1, 0, -1), 1), ((-1, 0, 0, -1, 0, 1), 0),
                                      ((-1, 0, 0, 1, 0, -1), 0), ((1, 0, 0, -1, 0, -1), 0),
                                      ((-2, 1, 0, -2, 1, -2), 1), ((-1, 2, 0, -1, 0, -1), 0),
                                      ((0, 0, 0, 0, 0, 0), 0)])
def test_negdef_known_answers(h, expect):
    assert oracle.negdef(*map(float, h)) == expect


def test_negdef_agrees_with_eigvalsh_away_from_singular():
    A = RNG.standard_normal((4000, 3, 3))
    S = (A + A.transpose(0, 2, 1)) / 2
    ev = np.linalg.eigvalsh(S)
    ok = np.abs(ev).min(axis=1) > 0.05
    for M, e in zip(S[ok][:1500], ev[ok][:1500]):
        got = oracle.negdef(M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2])
        assert got == int(np.all(e < 0))


# ---------------------------------------------------------------- a9 CCL
def _canonical_min_index(lab):
    """Relabel a scipy label image so each component carries its minimum linear index."""
    flat = lab.ravel()
    out = np.zeros(flat.shape, np.uint32)
    n = flat.max()
    mins = np.full(n + 1, np.iinfo(np.int64).max, np.int64)
    idx = np.arange(flat.size, dtype=np.int64)
    np.minimum.at(mins, flat, idx)
    fg = flat > 0
    out[fg] = mins[flat[fg]]
    return out.reshape(lab.shape)


def test_ccl_matches_scipy_label_without_border_contact():
    mask = (RNG.random((12, 20, 22)) < 0.3).astype(np.uint8)
    mask[:, :, 0] = mask[:, :, -1] = 0
    mask[:, 0, :] = mask[:, -1, :] = 0
    mask.flat[0] = 0
    lab, _ = ndi.label(mask, structure=np.ones((3, 3, 3)))
    np.testing.assert_array_equal(oracle.ccl3d(mask), _canonical_min_index(lab))


def _bfs_wrapped(mask):
    D, H, W = mask.shape
    N = mask.size
    flat = mask.ravel().astype(bool).copy()
    flat[0] = False
    offs = sorted({dz * W * H + dy * W + dx for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                   for dx in (-1, 0, 1)} - {0})
    lab = np.zeros(N, np.uint32)
    for s in range(N):
        if not flat[s] or lab[s]:
            continue
        comp, q = [s], collections.deque([s])
        lab[s] = 1
        while q:
            v = q.popleft()
            for o in offs:
                u = v + o
                if 0 <= u < N and flat[u] and not lab[u]:
                    lab[u] = 1
                    comp.append(u)
                    q.append(u)
        lab[comp] = min(comp)
    return lab.reshape(mask.shape)


@pytest.mark.parametrize("shape,p", [((3, 4, 5), 0.4), ((4, 6, 3), 0.3), ((2, 3, 7), 0.5),
                                     ((5, 1, 9), 0.4), ((1, 1, 30), 0.5), ((3, 5, 1), 0.5)])
def test_ccl_brute_force_wrapped_adjacency(shape, p):
    for seed in range(4):
        mask = (np.random.default_rng(seed).random(shape) < p).astype(np.uint8)
        np.testing.assert_array_equal(oracle.ccl3d(mask), _bfs_wrapped(mask))


def test_ccl_wrap_joins_row_end_to_next_row_start():
    mask = np.zeros((1, 3, 4), np.uint8)
    mask[0, 0, 3] = 1      # x = W-1 of row 0
    mask[0, 1, 0] = 1      # x = 0 of row 1: linear neighbour (+1)
    lab = oracle.ccl3d(mask)
    assert lab[0, 0, 3] == lab[0, 1, 0] == 3


def test_ccl_voxel_zero_is_background():
    mask = np.ones((2, 2, 2), np.uint8)
    lab = oracle.ccl3d(mask)
    assert lab.flat[0] == 0 and np.all(lab.ravel()[1:] == 1)


# ---------------------------------------------------------------- a10-a13 regions
def _regions_reference(labels, hess, norm, img):
    flat = labels.ravel()
    fg = flat > 0
    roots = np.unique(flat[fg])
    inv = np.searchsorted(roots, flat[fg])
    D, H, W = labels.shape
    z, y, x = np.unravel_index(np.nonzero(fg)[0], labels.shape)
    cnt = np.bincount(inv, minlength=len(roots))
    sums = [np.bincount(inv, weights=c.astype(np.float64), minlength=len(roots)) for c in (x, y, z)]
    hs = [np.bincount(inv, weights=hess[k].ravel()[fg].astype(np.float64), minlength=len(roots))
          for k in range(6)]
    mx = ndi.maximum(norm.ravel()[fg], labels=inv + 1, index=np.arange(1, len(roots) + 1))
    isum = np.bincount(inv, weights=img.ravel()[fg].astype(np.float64), minlength=len(roots))
    return roots, cnt, sums, hs, np.asarray(mx, np.float32), isum


def test_regions_match_numpy_reductions_and_eigvalsh():
    vol = synth.make_volume("c0", shape=(24, 48, 40), n_blobs=40, seed=7)
    o = oracle.substack(vol, dict(min_intensity=0, rxy=4, rz=3, sigma_xy1=1.5, sigma_z1=1.5,
                                  sigma_xy2=2.4, sigma_z2=2.4, truncate=2.0, gamma=1.0))
    reg = o["regions"]
    roots, cnt, sums, hs, mx, isum = _regions_reference(o["labels"], o["hess"], o["norm"], vol)
    assert len(reg["label"]) == len(roots) + 1 and len(roots) > 5
    np.testing.assert_array_equal(reg["label"][1:], roots)
    np.testing.assert_array_equal(reg["count"][1:], cnt)
    np.testing.assert_array_equal(reg["size16"][1:], cnt.astype(np.uint16))
    for a in range(3):
        np.testing.assert_array_equal(reg["sumxyz"][1:, a], sums[a].astype(np.uint64))
        np.testing.assert_allclose(reg["centroid"][1:, a], sums[a] / cnt, rtol=1e-6)
    for k in range(6):
        tol = 1e-6 * np.abs(o["hess"][k]).max() * cnt
        assert np.all(np.abs(reg["hsum"][1:, k] - hs[k]) <= tol + 1e-7 * np.abs(hs[k]))
    np.testing.assert_array_equal(reg["norm"][1:], mx)
    np.testing.assert_array_equal(reg["isum"][1:], isum.astype(np.uint64))
    # eigen features vs LAPACK on the float sums
    h = reg["hsum"][1:].astype(np.float64)
    M = np.stack([np.stack([h[:, 0], h[:, 1], h[:, 2]], -1), np.stack([h[:, 1], h[:, 3], h[:, 4]], -1),
                  np.stack([h[:, 2], h[:, 4], h[:, 5]], -1)], -2)
    ev = np.linalg.eigvalsh(M)
    st = (ev ** 2).sum(-1)
    bl = ev[:, 2] ** 2 / (ev[:, 0] * ev[:, 1])
    offd = h[:, 1] ** 2 + h[:, 2] ** 2 + h[:, 4] ** 2
    sel = offd > 1e-30
    np.testing.assert_allclose(reg["eig"][1:, 0][sel], st[sel], rtol=1e-5)
    ok = sel & (np.abs(ev[:, 0] * ev[:, 1]) > 1e-6 * st)
    np.testing.assert_allclose(reg["eig"][1:, 1][ok], bl[ok], rtol=1e-4)


@pytest.mark.parametrize("h,st,bl", [((-1, 0, 0, -2, 0, -3), 14.0, 4.5),
                                     ((0, 0, 0, 0, 0, 0), 0.0, 0.0),
                                     ((-3, 0, 0, -3, 0, -3), 27.0, 1.0 / 9.0 * 9.0)])
def test_eigen_known_answers(h, st, bl):
    out = oracle.eigen_features(np.array(h, np.float32))
    assert abs(out[0] - st) < 1e-5 * max(1, st)
    assert abs(out[1] - bl) < 1e-5 * max(1, abs(bl))


def test_eigen_trig_branch_matches_eigvalsh():
    for _ in range(200):
        a = RNG.standard_normal(6).astype(np.float32)
        M = np.array([[a[0], a[1], a[2]], [a[1], a[3], a[4]], [a[2], a[4], a[5]]], np.float64)
        ev = np.linalg.eigvalsh(M)
        out = oracle.eigen_features(a)
        assert abs(out[0] - (ev ** 2).sum()) < 1e-5 * (ev ** 2).sum()
        b = ev[2] ** 2 / (ev[0] * ev[1])
        if abs(ev[0] * ev[1]) > 1e-3:
            assert abs(out[1] - b) < 1e-4 * max(1.0, abs(b))


def test_region_size_wraps_modulo_65536():
    D, H, W = 4, 160, 120           # one block of 70 000 voxels
    labels = np.zeros((D, H, W), np.uint32)
    blk = np.zeros((D, H, W), bool)
    blk[:, 2:147, 1:121] = True     # 4*145*120 = 69600
    first = np.flatnonzero(blk.ravel())[0]
    labels[blk] = first
    hess = np.zeros((6, D, H, W), np.float32)
    norm = np.ones((D, H, W), np.float32)
    img = np.ones((D, H, W), np.uint16)
    reg = oracle.regions(labels, hess, norm, img)
    n = int(blk.sum())
    assert reg["count"][1] == n and reg["size16"][1] == n % 65536
    x = np.nonzero(blk)[2]
    assert reg["centroid"][1, 0] == np.float32(np.float32(x.sum()) / np.float32(n % 65536))
