"""ctypes front-end of the CPU oracle (``libhdog_oracle.so``).

TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the CPU-baseline
leg of ``bench.py`` use it as the checker.  No product path imports it.

Each wrapper restates one stage of the reference HDoG path; the C source cites
the reference file:line for every function (hdog_oracle.cpp).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libhdog_oracle.so")
_lib = None

_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_l = ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_minmax3d.argtypes = [_u16p, _u16p, _i, _i, _i, _i, _i, _i]
        L.oracle_box3d.argtypes = [_u16p, _f32p, _i, _i, _i, _i, _i]
        L.oracle_normalize3d.argtypes = [_u16p, _f32p, _f, _i, _i, _i, _i, _i,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_gaussian_radius.argtypes = [_f, _d]
        L.oracle_gaussian_radius.restype = _i
        L.oracle_gaussian_weights.argtypes = [_f, _d, _f32p]
        L.oracle_gaussian_weights.restype = _i
        L.oracle_dog_normalizer.argtypes = [_f, _f, _f]
        L.oracle_dog_normalizer.restype = _f
        L.oracle_gaussian3d.argtypes = [_f32p, _f32p, _i, _i, _i, _f32p, _i, _f32p, _i]
        L.oracle_dog3d.argtypes = [_f32p, _f32p, _i, _i, _i, _f, _f, _f, _f, _d, _f]
        L.oracle_dog3d.restype = _i
        L.oracle_negdef.argtypes = [_f] * 6
        L.oracle_negdef.restype = ctypes.c_uint8
        L.oracle_hessian3d.argtypes = [_f32p, _f32p, _u8p, _i, _i, _i]
        L.oracle_positive_offsets.argtypes = [_i, _i, _i64p]
        L.oracle_positive_offsets.restype = _i
        L.oracle_ccl3d.argtypes = [_u8p, _u32p, _i, _i, _i]
        L.oracle_eigen_features.argtypes = [_f32p, _f32p]
        L.oracle_regions.argtypes = [_u32p, _f32p, _f32p, _u16p, _i, _i, _i, _i, _l,
                                     _u32p, _u16p, _u64p, _u64p, _f32p, _f32p, _f32p,
                                     _f32p, _u64p, _u16p]
        L.oracle_regions.restype = _l
        L.oracle64_normalize3d.argtypes = [_u16p, _f64p, _d, _i, _i, _i, _i, _i]
        L.oracle64_gaussian_weights.argtypes = [_f, _d, _f64p]
        L.oracle64_gaussian_weights.restype = _i
        L.oracle64_gaussian3d.argtypes = [_f64p, _f64p, _i, _i, _i, _f64p, _i, _f64p, _i]
        L.oracle64_dog3d.argtypes = [_f64p, _f64p, _i, _i, _i, _f, _f, _f, _f, _d, _f]
        L.oracle64_dog3d.restype = _i
        L.oracle64_negdef.argtypes = [_d] * 6
        L.oracle64_negdef.restype = ctypes.c_uint8
        L.oracle64_hessian3d.argtypes = [_f64p, _f64p, _u8p, _i, _i, _i]
        _lib = L
    return _lib


def _shape(a):
    D, H, W = a.shape
    return W, H, D


def minmax3d(img, rxy, rz, is_min):
    img = np.ascontiguousarray(img, np.uint16)
    out = np.empty_like(img)
    lib().oracle_minmax3d(img, out, *_shape(img), rxy, rz, int(bool(is_min)))
    return out


def box3d(img_u16, rxy, rz):
    img = np.ascontiguousarray(img_u16, np.uint16)
    out = np.empty(img.shape, np.float32)
    lib().oracle_box3d(img, out, *_shape(img), rxy, rz)
    return out


def normalize3d(img, min_intensity, rxy, rz, return_ed=False):
    img = np.ascontiguousarray(img, np.uint16)
    norm = np.empty(img.shape, np.float32)
    if return_ed:
        E = np.empty(img.shape, np.float32)
        Dl = np.empty(img.shape, np.float32)
        lib().oracle_normalize3d(img, norm, float(min_intensity), *_shape(img), rxy, rz,
                                 E.ctypes.data, Dl.ctypes.data)
        return norm, E, Dl
    lib().oracle_normalize3d(img, norm, float(min_intensity), *_shape(img), rxy, rz,
                             None, None)
    return norm


def gaussian_radius(sigma, truncate=2.0):
    return lib().oracle_gaussian_radius(float(sigma), float(truncate))


def gaussian_weights(sigma, truncate=2.0):
    r = gaussian_radius(sigma, truncate)
    w = np.zeros(2 * r + 1, np.float32)
    lib().oracle_gaussian_weights(float(sigma), float(truncate), w)
    return w


def dog_normalizer(sxy1, sz1, gamma):
    return lib().oracle_dog_normalizer(float(sxy1), float(sz1), float(gamma))


def gaussian3d(vol, wxy, wz):
    vol = np.ascontiguousarray(vol, np.float32)
    out = np.empty_like(vol)
    wxy = np.ascontiguousarray(wxy, np.float32)
    wz = np.ascontiguousarray(wz, np.float32)
    lib().oracle_gaussian3d(vol, out, *_shape(vol), wxy, len(wxy) // 2, wz, len(wz) // 2)
    return out


def dog3d(norm, sigma_xy1, sigma_z1, sigma_xy2, sigma_z2, truncate=2.0, gamma=1.0):
    norm = np.ascontiguousarray(norm, np.float32)
    out = np.empty_like(norm)
    rz2 = lib().oracle_dog3d(norm, out, *_shape(norm), sigma_xy1, sigma_z1, sigma_xy2,
                             sigma_z2, float(truncate), float(gamma))
    return out, rz2


def negdef(xx, xy, xz, yy, yz, zz):
    return lib().oracle_negdef(xx, xy, xz, yy, yz, zz)


def hessian3d(dog):
    dog = np.ascontiguousarray(dog, np.float32)
    hess = np.empty((6,) + dog.shape, np.float32)
    pd = np.empty(dog.shape, np.uint8)
    lib().oracle_hessian3d(dog, hess, pd, *_shape(dog))
    return hess, pd


def positive_offsets(W, H):
    offs = np.zeros(13, np.int64)
    n = lib().oracle_positive_offsets(W, H, offs)
    return offs[:n]


def ccl3d(pd):
    pd = np.ascontiguousarray(pd, np.uint8)
    labels = np.empty(pd.shape, np.uint32)
    lib().oracle_ccl3d(pd, labels, *_shape(pd))
    return labels


def eigen_features(h6):
    h6 = np.ascontiguousarray(h6, np.float32)
    out = np.zeros(2, np.float32)
    lib().oracle_eigen_features(h6, out)
    return out


def regions(labels, hess, norm, img, is_ave_mode=False):
    """Per-region features; dict of arrays indexed by region id (0 = background)."""
    labels = np.ascontiguousarray(labels, np.uint32)
    hess = np.ascontiguousarray(hess, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    img = np.ascontiguousarray(img, np.uint16)
    W, H, D = _shape(labels)
    flat = labels.ravel()
    cap = int(np.count_nonzero(flat[1:] == np.arange(1, flat.size, dtype=np.uint32))) + 1
    o = dict(label=np.zeros(cap, np.uint32), size16=np.zeros(cap, np.uint16),
             count=np.zeros(cap, np.uint64), sumxyz=np.zeros((cap, 3), np.uint64),
             centroid=np.zeros((cap, 3), np.float32), hsum=np.zeros((cap, 6), np.float32),
             eig=np.zeros((cap, 2), np.float32), norm=np.zeros(cap, np.float32),
             isum=np.zeros(cap, np.uint64), imax=np.zeros(cap, np.uint16))
    R = lib().oracle_regions(labels, hess, norm, img, W, H, D, int(bool(is_ave_mode)), cap,
                             o["label"], o["size16"], o["count"], o["sumxyz"], o["centroid"],
                             o["hsum"], o["eig"], o["norm"], o["isum"], o["imax"])
    assert R == cap, (R, cap)
    return o


def substack(img, p):
    """Run the whole oracle path on one substack; p is a dict of HDoG parameters."""
    norm = normalize3d(img, p["min_intensity"], p["rxy"], p["rz"])
    dog, rz2 = dog3d(norm, p["sigma_xy1"], p["sigma_z1"], p["sigma_xy2"], p["sigma_z2"],
                     p.get("truncate", 2.0), p.get("gamma", 1.0))
    hess, pd = hessian3d(dog)
    labels = ccl3d(pd)
    reg = regions(labels, hess, norm, img, p.get("is_ave_mode", False))
    return dict(norm=norm, dog=dog, hess=hess, pd=pd, labels=labels, regions=reg, rz2=rz2)


# ---- FP64 restatement (the reference of the reduced-precision storage forms) ----------
def normalize3d_f64(img, min_intensity, rxy, rz):
    img = np.ascontiguousarray(img, np.uint16)
    norm = np.empty(img.shape, np.float64)
    lib().oracle64_normalize3d(img, norm, float(min_intensity), *_shape(img), rxy, rz)
    return norm


def gaussian_weights_f64(sigma, truncate=2.0):
    r = gaussian_radius(sigma, truncate)
    w = np.zeros(2 * r + 1, np.float64)
    lib().oracle64_gaussian_weights(float(sigma), float(truncate), w)
    return w


def gaussian3d_f64(vol, wxy, wz):
    vol = np.ascontiguousarray(vol, np.float64)
    out = np.empty_like(vol)
    wxy = np.ascontiguousarray(wxy, np.float64)
    wz = np.ascontiguousarray(wz, np.float64)
    lib().oracle64_gaussian3d(vol, out, *_shape(vol), wxy, len(wxy) // 2, wz, len(wz) // 2)
    return out


def dog3d_f64(norm, sigma_xy1, sigma_z1, sigma_xy2, sigma_z2, truncate=2.0, gamma=1.0):
    norm = np.ascontiguousarray(norm, np.float64)
    out = np.empty_like(norm)
    rz2 = lib().oracle64_dog3d(norm, out, *_shape(norm), sigma_xy1, sigma_z1, sigma_xy2,
                               sigma_z2, float(truncate), float(gamma))
    return out, rz2


def negdef_f64(xx, xy, xz, yy, yz, zz):
    return lib().oracle64_negdef(xx, xy, xz, yy, yz, zz)


def hessian3d_f64(dog):
    dog = np.ascontiguousarray(dog, np.float64)
    hess = np.empty((6,) + dog.shape, np.float64)
    pd = np.empty(dog.shape, np.uint8)
    lib().oracle64_hessian3d(dog, hess, pd, *_shape(dog))
    return hess, pd


def regions_f64(labels, hess, norm):
    """Per-region features of an FP64 result, in float64 (numpy): label, voxel count,
    centroid (sum / count), Hessian sums, structureness and blobness from the sorted
    eigenvalues of the summed Hessian, max of norm.  Region 0 (background) excluded."""
    flat = np.ascontiguousarray(labels).ravel()
    on = np.nonzero(flat)[0]
    lab = flat[on]
    uniq, inv = np.unique(lab, return_inverse=True)
    R = len(uniq)
    cnt = np.bincount(inv, minlength=R).astype(np.float64)
    D, H, W = labels.shape
    x, y, z = on % W, (on // W) % H, on // (W * H)
    cen = np.stack([np.bincount(inv, weights=c.astype(np.float64), minlength=R) / cnt
                    for c in (x, y, z)], axis=1)
    hs = np.stack([np.bincount(inv, weights=hess[k].ravel()[on], minlength=R) for k in range(6)],
                  axis=1)
    M = np.stack([hs[:, [0, 1, 2]], hs[:, [1, 3, 4]], hs[:, [2, 4, 5]]], axis=1)
    ev = np.linalg.eigvalsh(M)
    st = (ev ** 2).sum(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        bl = ev[:, 2] ** 2 / (ev[:, 0] * ev[:, 1])
    nmax = np.full(R, -np.inf)
    np.maximum.at(nmax, inv, np.ascontiguousarray(norm).ravel()[on])
    return dict(label=uniq.astype(np.uint32), count=cnt, centroid=cen, hsum=hs,
                structureness=st, blobness=bl, norm=nmax)


def substack_f64(img, p):
    """Filter stages in FP64, then labels and region features of the FP64 candidates."""
    norm = normalize3d_f64(img, p["min_intensity"], p["rxy"], p["rz"])
    dog, _ = dog3d_f64(norm, p["sigma_xy1"], p["sigma_z1"], p["sigma_xy2"], p["sigma_z2"],
                       p.get("truncate", 2.0), p.get("gamma", 1.0))
    hess, pd = hessian3d_f64(dog)
    labels = ccl3d(pd)
    return dict(norm=norm, dog=dog, hess=hess, pd=pd, labels=labels,
                regions=regions_f64(labels, hess, norm))
