"""From-scratch CPU counterpart of the reference's CPU detector, script/HDoG_cpu.py:57-146,
used as the CPU baseline that bench.py times beside the GPU path.

TEST AND BENCHMARK INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and the tests use it;
no product path imports it.

It keeps the script's stage semantics (which differ from the CUDA path the GPU build
reproduces, see oracle/hdog_oracle.cpp):
  Normalize  grey dilation / erosion (mode "nearest") of the box (2 rz + 1, 2 rxy + 1,
             2 rxy + 1), box means of both (uniform_filter, mode "constant"), then
             norm = (I >= min_intensity) * I / (D - E)          (HDoG_cpu.py:57-67)
  DoG        gaussian_filter of sigma1 and sigma2 (z, y, x), mode "constant", then
             (G1 - G2) * sigma_xy1^(2 gamma) * sigma_z1^gamma   (:69-79, :46)
             the truncation is a parameter (the script fixes 2.0)
  Hessian    np.gradient twice, negative definite by leading minors   (:81-90)
  CCL        26-connected labels: scipy.ndimage.label with a 3 x 3 x 3 structure, which
             is what the script's skimage.morphology.label does with full connectivity
             (skimage is not installed here)                     (:96-98)
  Regions    max of norm, sizes, Hessian sums by bincount (float64), eigvalsh,
             blobness l3^2 / (l1 l2), structureness l1^2 + l2^2 + l3^2, centroids
                                                                   (:100-146)
The slab form runs Normalize, DoG and Hessian on slabs of the stack (along its longest
axis by default) in worker processes, each slab with a halo of 2 r_norm + r_dog + 2 slices
along that axis (the min / max filter and its box mean reach r_norm each, the Gaussian
r_dog, the two gradients one each: the slab's own slices then equal the whole-stack
result), then labels and reduces the whole stack in the parent process.
"""
import os
import time

import numpy as np
import scipy.ndimage as ndi


def gauss_radius(sigma, truncate):
    return int(truncate * float(sigma) + 0.5)


class HDoGCPU:
    def __init__(self, sigma1_zyx, sigma2_zyx, radius_xy, radius_z, min_intensity=0,
                 gamma=1.0, truncate=2.0):
        self.s1 = tuple(float(s) for s in sigma1_zyx)
        self.s2 = tuple(float(s) for s in sigma2_zyx)
        self.size = (2 * radius_z + 1, 2 * radius_xy + 1, 2 * radius_xy + 1)
        self.rz = radius_z
        self.min_intensity = min_intensity
        self.truncate = truncate
        self.normalizer = self.s1[1] ** (gamma * 2) * self.s1[0] ** gamma

    # -- stages ------------------------------------------------------------------
    def normalize(self, img):
        dil = ndi.uniform_filter(ndi.grey_dilation(img, size=self.size, mode="nearest")
                                 .astype(np.float32), size=self.size, mode="constant", cval=0)
        ero = ndi.uniform_filter(ndi.grey_erosion(img, size=self.size, mode="nearest")
                                 .astype(np.float32), size=self.size, mode="constant", cval=0)
        inten = img.astype(np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            return (inten >= self.min_intensity) * inten / (dil - ero)

    def dog(self, norm):
        g1 = ndi.gaussian_filter(norm, sigma=self.s1, truncate=self.truncate, mode="constant",
                                 cval=0)
        g2 = ndi.gaussian_filter(norm, sigma=self.s2, truncate=self.truncate, mode="constant",
                                 cval=0)
        return (g1 - g2) * np.float32(self.normalizer)

    @staticmethod
    def hessian_pd(dog):
        hz, hy, hx = np.gradient(dog)
        hzz, hyz, hxz = np.gradient(hz)
        _, hyy, hxy = np.gradient(hy)
        _, _, hxx = np.gradient(hx)
        det = hxx * hyy * hzz + 2 * hxy * hyz * hxz - hxx * hyz * hyz - hyy * hxz * hxz - hzz * hxy * hxy
        pd = (hxx < 0) & (hxx * hyy - hxy * hxy > 0) & (det < 0)
        return pd, np.array([hxx, hxy, hxz, hyy, hyz, hzz])

    @staticmethod
    def ccl(pd):
        labels, n = ndi.label(pd, structure=np.ones((3, 3, 3), bool))
        return labels, n

    @staticmethod
    def regional_features(norm, hess, labels, n):
        on = np.nonzero(labels)
        lab = labels[on] - 1
        out = dict(n=n)
        if n == 0:
            return out
        out["max_normalized"] = ndi.maximum(norm, labels=labels, index=np.arange(1, n + 1))
        size = np.bincount(lab, minlength=n).astype(np.float64)
        HT = np.stack([np.bincount(lab, weights=hess[k][on], minlength=n) for k in range(6)])
        M = np.array([[HT[0], HT[1], HT[2]], [HT[1], HT[3], HT[4]], [HT[2], HT[4], HT[5]]]).T
        ev = np.linalg.eigvalsh(M)
        l1, l2, l3 = ev[:, 0], ev[:, 1], ev[:, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            out["blobness"] = l3 * l3 / (l1 * l2)
        out["structureness"] = l1 * l1 + l2 * l2 + l3 * l3
        out["size"] = size
        for ax, c in zip("zyx", on):
            out["centroid_" + ax] = np.bincount(lab, weights=c.astype(np.float64), minlength=n) / size
        return out

    # -- whole path --------------------------------------------------------------
    def filters(self, img):
        norm = self.normalize(img)
        pd, hess = self.hessian_pd(self.dog(norm))
        return norm, pd, hess

    def run(self, img, times=None):
        t = time.perf_counter()
        norm = self.normalize(img)
        t1 = time.perf_counter()
        dog = self.dog(norm)
        t2 = time.perf_counter()
        pd, hess = self.hessian_pd(dog)
        t3 = time.perf_counter()
        labels, n = self.ccl(pd)
        t4 = time.perf_counter()
        reg = self.regional_features(norm, hess, labels, n)
        t5 = time.perf_counter()
        if times is not None:
            times.update(normalize=t1 - t, dog=t2 - t1, hessian=t3 - t2, ccl=t4 - t3,
                         regions=t5 - t4)
        return reg

    def halo(self, axis=0):
        """Slices of context a slab needs along `axis` (0 z, 1 y, 2 x)."""
        r_norm = self.size[axis] // 2
        sig = (self.s1[axis], self.s2[axis])
        return 2 * r_norm + max(gauss_radius(s, self.truncate) for s in sig) + 2

    def run_slabs(self, img, procs, axis=None, times=None):
        """Normalize / DoG / Hessian on slabs along `axis` (default: the longest) in `procs`
        worker processes (fork), then labels and regions of the whole stack here."""
        import multiprocessing as mp
        if axis is None:
            axis = int(np.argmax(img.shape))
        n = img.shape[axis]
        h = self.halo(axis)
        edges = np.linspace(0, n, procs + 1).astype(int)
        jobs = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
        t = time.perf_counter()
        global _SLAB_CTX
        _SLAB_CTX = (self, img, h, axis)
        with mp.get_context("fork").Pool(len(jobs)) as pool:
            parts = pool.map(_slab_filters, jobs)
        _SLAB_CTX = None
        norm = np.concatenate([p[0] for p in parts], axis=axis)
        pd = np.concatenate([p[1] for p in parts], axis=axis)
        hess = np.concatenate([p[2] for p in parts], axis=axis + 1)
        t1 = time.perf_counter()
        labels, n = self.ccl(pd)
        t2 = time.perf_counter()
        reg = self.regional_features(norm, hess, labels, n)
        t3 = time.perf_counter()
        if times is not None:
            times.update(filters=t1 - t, ccl=t2 - t1, regions=t3 - t2)
        return reg


_SLAB_CTX = None


def _take(a, axis, lo, hi):
    idx = [slice(None)] * a.ndim
    idx[axis] = slice(lo, hi)
    return a[tuple(idx)]


def _slab_filters(job):
    self, img, h, axis = _SLAB_CTX
    a, b = job
    lo, hi = max(0, a - h), min(img.shape[axis], b + h)
    norm, pd, hess = self.filters(_take(img, axis, lo, hi))
    return (_take(norm, axis, a - lo, b - lo), _take(pd, axis, a - lo, b - lo),
            _take(hess, axis + 1, a - lo, b - lo))


def cores():
    return len(os.sched_getaffinity(0))
