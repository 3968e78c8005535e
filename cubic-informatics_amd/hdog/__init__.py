"""MI355X HDoG cell-candidate detection: Python front-end of libhdog.so.

The stage methods keep the names and argument meaning of the reference's launch
wrappers (src/3D/*.cuh of lsb-riken/CUBIC-informatics); buffers are torch tensors on
the HIP device (torch is only the allocator and stream owner here), and every call
runs the hand-written gfx950 kernels through the C-ABI of include/hdog.h.  Errors
raise HDoGError with the library's message instead of the reference's print-and-exit.

    hd = HDoG(device=0)
    norm = hd.Normalize3DFilter(img_u16, min_intensity, radius_xy, radius_z)
    rz2 = hd.initGaussian3DKernel(sxy1, sz1, sxy2, sz2, truncate=2.0)
    dog = hd.DoG3DFilter(norm, gamma)
    hess, pd = hd.HessianPositiveDefiniteWithElement(dog)
    labels = hd.CCL(pd)
    regions = hd.RegionalFeatures(labels, hess, norm, img_u16, is_ave_mode=False)
"""
import ctypes

import numpy as np
import torch

from . import _capi as C
from . import synth  # noqa: F401

__all__ = ["HDoG", "HDoGError", "PRECISIONS"]

PRECISIONS = {"fp32": C.HDOG_FP32, "fp16": C.HDOG_FP16, "bf16": C.HDOG_BF16,
              "fp16_dog32": C.HDOG_FP16_DOG32, "bf16_dog32": C.HDOG_BF16_DOG32}
# storage dtypes (norm and hessian, dog) of each precision form
_TORCH_DTYPE = {C.HDOG_FP32: (torch.float32, torch.float32),
                C.HDOG_FP16: (torch.float16, torch.float16),
                C.HDOG_BF16: (torch.bfloat16, torch.bfloat16),
                C.HDOG_FP16_DOG32: (torch.float16, torch.float32),
                C.HDOG_BF16_DOG32: (torch.bfloat16, torch.float32)}


class HDoGError(RuntimeError):
    def __init__(self, message, status=None):
        super().__init__(message)
        self.status = status


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _shape(t):
    D, H, W = t.shape[-3:]
    return int(W), int(H), int(D)


class HDoG:
    """One library context bound to one HIP device."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise HDoGError("no HIP device visible: the HDoG path runs only on the GPU")
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        rc = C.hdog_create(ctypes.byref(h), device)
        if rc != C.HDOG_OK:
            raise HDoGError(f"hdog_create({device}) failed with status {rc}")
        self._h = h
        self.r_z2 = None
        self._last_R = 0

    def release_scratch(self):
        """Free the context's device scratch (not a batch plan, not the caller's buffers)."""
        self._check(C.hdog_release_scratch(self._h))

    def close(self):
        if getattr(self, "_h", None):
            C.hdog_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -----------------------------------------------------------------
    def _check(self, rc):
        if rc != C.HDOG_OK:
            raise HDoGError(C.hdog_last_error(self._h).decode() or f"status {rc}", rc)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, a, dtype):
        if isinstance(a, np.ndarray):
            a = torch.from_numpy(np.ascontiguousarray(a))
        a = a.to(self.device)
        if a.dtype != dtype:
            raise HDoGError(f"expected {dtype}, got {a.dtype}")
        return a.contiguous()

    # -- reference-named stages --------------------------------------------------
    def Normalize3DFilter(self, img, min_intensity, radius_xy, radius_z, out=None):
        """src/3D/NormalizeFilter.cu:603-654: uint16 (D,H,W) -> float32 normalized."""
        img = self._dev(img, torch.uint16)
        W, H, D = _shape(img)
        out = torch.empty(img.shape, dtype=torch.float32, device=self.device) if out is None else out
        self._check(C.hdog_normalize3d(self._h, self._stream(), _ptr(img), _ptr(out),
                                       float(min_intensity), W, H, D, int(radius_xy), int(radius_z)))
        return out

    def Normalize3DFilterWithBounds(self, img, min_intensity, radius_xy, radius_z):
        """Normalize3DFilter plus its erosion_l / dilation_l volumes (box means of the min and
        max filters): returns (norm, E, D), float32 (D,H,W) each."""
        img = self._dev(img, torch.uint16)
        W, H, D = _shape(img)
        norm, ero, dil = (torch.empty(img.shape, dtype=torch.float32, device=self.device)
                          for _ in range(3))
        self._check(C.hdog_normalize3d_ed(self._h, self._stream(), _ptr(img), _ptr(norm), _ptr(ero),
                                          _ptr(dil), float(min_intensity), W, H, D, int(radius_xy),
                                          int(radius_z)))
        return norm, ero, dil

    def initGaussian3DKernel(self, sigma_xy1, sigma_z1, sigma_xy2, sigma_z2, truncate=2.0):
        """src/3D/DoG3DFilter.cu:394-415; returns the z radius of sigma2."""
        r = ctypes.c_int()
        self._check(C.hdog_init_gaussian3d(self._h, sigma_xy1, sigma_z1, sigma_xy2, sigma_z2,
                                           float(truncate), ctypes.byref(r)))
        self.r_z2 = r.value
        return r.value

    def DoG3DFilter(self, norm, gamma=1.0, out=None):
        """src/3D/DoG3DFilter.cu:431-437."""
        norm = self._dev(norm, torch.float32)
        W, H, D = _shape(norm)
        out = torch.empty_like(norm) if out is None else out
        self._check(C.hdog_dog3d(self._h, self._stream(), _ptr(norm), _ptr(out), W, H, D,
                                 float(gamma)))
        return out

    def HessianPositiveDefiniteWithElement(self, dog):
        """src/3D/Hessian3DFilter_element.cu:355-388 -> (hess[6,D,H,W], pd uint8)."""
        dog = self._dev(dog, torch.float32)
        W, H, D = _shape(dog)
        hess = torch.empty((6,) + tuple(dog.shape), dtype=torch.float32, device=self.device)
        pd = torch.empty(dog.shape, dtype=torch.uint8, device=self.device)
        self._check(C.hdog_hessian_pd(self._h, self._stream(), _ptr(dog), _ptr(hess), _ptr(pd),
                                      W, H, D))
        return hess, pd

    def CCL(self, pd):
        """src/3D/CCL3D.cu:81-117 -> uint32 labels (component minimum linear index)."""
        pd = self._dev(pd, torch.uint8)
        W, H, D = _shape(pd)
        labels = torch.empty(pd.shape, dtype=torch.uint32, device=self.device)
        self._check(C.hdog_ccl3d(self._h, self._stream(), _ptr(pd), _ptr(labels), W, H, D))
        return labels

    def RegionalFeatures(self, labels, hess, norm, img=None, is_ave_mode=False):
        """HessianFeatures + Eigenvalues + MaxNormalized/SumNormalized +
        RegionalSizeAndCentroid (src/3D/HDoG3D_main.cpp:184-214) -> dict of numpy arrays."""
        labels = self._dev(labels, torch.uint32)
        hess = self._dev(hess, torch.float32)
        norm = self._dev(norm, torch.float32)
        img = self._dev(img, torch.uint16) if img is not None else None
        W, H, D = _shape(labels)
        R = ctypes.c_int64()
        self._check(C.hdog_regional_features(self._h, self._stream(), _ptr(labels), _ptr(hess),
                                             _ptr(norm), _ptr(img), W, H, D, int(bool(is_ave_mode)),
                                             ctypes.byref(R)))
        self._last_R = R.value
        return self.download_regions()

    @staticmethod
    def _region_arrays(R):
        o = dict(label=np.zeros(R, np.uint32), size16=np.zeros(R, np.uint16),
                 count=np.zeros(R, np.uint64), cx=np.zeros(R, np.float32),
                 cy=np.zeros(R, np.float32), cz=np.zeros(R, np.float32),
                 structureness=np.zeros(R, np.float32), blobness=np.zeros(R, np.float32),
                 norm=np.zeros(R, np.float32), hsum=np.zeros((6, R), np.float32),
                 isum=np.zeros(R, np.uint64), imax=np.zeros(R, np.uint16))
        reg = C.Regions()
        reg.capacity = R
        for k, f in (("label", "label"), ("size16", "size"), ("count", "voxels"),
                     ("cx", "centroid_x"), ("cy", "centroid_y"), ("cz", "centroid_z"),
                     ("structureness", "structureness"), ("blobness", "blobness"),
                     ("norm", "max_normalized"), ("hsum", "hessian"), ("isum", "intensity_sum"),
                     ("imax", "intensity_max")):
            setattr(reg, f, o[k].ctypes.data)
        return o, reg

    @staticmethod
    def _finish_regions(o):
        o["centroid"] = np.stack([o["cx"], o["cy"], o["cz"]], axis=1)
        o["eig"] = np.stack([o["structureness"], o["blobness"]], axis=1)
        o["hsum"] = np.ascontiguousarray(o["hsum"].T)
        return o

    def download_regions(self):
        """Copy the last region result to host numpy arrays (index 0 = background)."""
        o, reg = self._region_arrays(self._last_R)
        self._check(C.hdog_regions_download(self._h, self._stream(), ctypes.byref(reg)))
        return self._finish_regions(o)

    # -- autotuned launch shapes ------------------------------------------------
    def autotune(self, img, params, cache_path=None):
        """Measured choice of the row walkers' segment counts for the stack shape of
        `params` (timed on `img`, a device uint16 (D, H, W) tensor, unless `cache_path`
        already holds the shape); kept in the context for later runs of that shape.
        Returns (normalization segments, DoG segments)."""
        ns, gs = ctypes.c_int(), ctypes.c_int()
        ptr = _ptr(self._dev(img, torch.uint16)) if img is not None else None
        path = cache_path.encode() if cache_path else None
        self._check(C.hdog_autotune(self._h, self._stream(), ctypes.byref(params), ptr, path,
                                    ctypes.byref(ns), ctypes.byref(gs)))
        return ns.value, gs.value

    # -- batched stacks (one captured graph) -------------------------------------
    def batch_plan(self, images, params, max_regions_per_stack=0, arena_capacity=0,
                   autotune_cache=None):
        """Capture the pipeline of every stack in `images` (uint16/int16 device tensors of
        shape (D, H, W), kept alive by this object) into one graph.  `params` from
        make_params; its sizes are ignored.  With `autotune_cache` (a file path) every
        stack shape's launch shapes are first chosen by the autotuner (measured, or read
        from the file)."""
        self._batch_images = [im if im.is_contiguous() else im.contiguous() for im in images]
        if autotune_cache:
            done = set()
            for im in self._batch_images:
                shape = tuple(im.shape)
                if shape in done:
                    continue
                done.add(shape)
                p = type(params).from_buffer_copy(params)
                p.depth, p.height, p.width = shape
                self.autotune(im, p, autotune_cache)
        tiles = (C.Tile * len(images))()
        for t, im in enumerate(self._batch_images):
            D, H, W = im.shape
            tiles[t].image = im.data_ptr()
            tiles[t].width, tiles[t].height, tiles[t].depth = W, H, D
        self._check(C.hdog_batch_plan(self._h, ctypes.byref(params), tiles, len(images),
                                      int(max_regions_per_stack), int(arena_capacity)))

    def batch_run(self):
        self._check(C.hdog_batch_run(self._h, self._stream()))

    def batch_tile_info(self, tile):
        n = ctypes.c_int64()
        over = ctypes.c_int()
        self._check(C.hdog_batch_tile_info(self._h, int(tile), ctypes.byref(n), ctypes.byref(over)))
        return n.value, bool(over.value)

    def batch_download(self, tile):
        n, _ = self.batch_tile_info(tile)
        o, reg = self._region_arrays(n + 1)
        self._check(C.hdog_batch_download(self._h, self._stream(), int(tile), ctypes.byref(reg)))
        return self._finish_regions(o)

    def batch_release(self):
        self._check(C.hdog_batch_release(self._h))
        self._batch_images = None

    # -- whole path ---------------------------------------------------------------
    @staticmethod
    def make_params(shape, radius_xy, radius_z, min_intensity, sigma_xy1, sigma_z1, sigma_xy2,
                    sigma_z2, truncate=2.0, gamma=1.0, is_ave_mode=False, precision="fp32",
                    minor_band=0.0):
        D, H, W = shape
        p = C.Params()
        p.width, p.height, p.depth = W, H, D
        p.radius_xy, p.radius_z = radius_xy, radius_z
        p.min_intensity = float(min_intensity)
        p.sigma_xy1, p.sigma_z1, p.sigma_xy2, p.sigma_z2 = sigma_xy1, sigma_z1, sigma_xy2, sigma_z2
        p.truncate = float(truncate)
        p.gamma = float(gamma)
        p.is_ave_mode = int(bool(is_ave_mode))
        p.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        p.minor_band = float(minor_band)
        return p

    def alloc_intermediates(self, shape, precision="fp32"):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        dt, dd = _TORCH_DTYPE[prec]
        D, H, W = shape
        kw = dict(device=self.device)
        return dict(norm=torch.empty((D, H, W), dtype=dt, **kw),
                    dog=torch.empty((D, H, W), dtype=dd, **kw),
                    hess=torch.empty((6, D, H, W), dtype=dt, **kw),
                    pd=torch.empty((D, H, W), dtype=torch.uint8, **kw),
                    labels=torch.empty((D, H, W), dtype=torch.uint32, **kw))

    def run_device(self, img, params, bufs=None):
        """Every stage on a device-resident uint16 volume; returns the region count.
        bufs (from alloc_intermediates) receives the intermediates when given."""
        img = self._dev(img, torch.uint16)
        b = bufs or {}
        R = ctypes.c_int64()
        self._check(C.hdog_run_device(self._h, self._stream(), ctypes.byref(params), _ptr(img),
                                      _ptr(b.get("norm")), _ptr(b.get("dog")), _ptr(b.get("hess")),
                                      _ptr(b.get("pd")), _ptr(b.get("labels")), ctypes.byref(R)))
        self._last_R = R.value
        return R.value

    def set_profiling(self, on=True):
        """Record a HIP event after every kernel of run_device (see stage_times)."""
        self._check(C.hdog_set_profiling(self._h, int(bool(on))))

    def stage_times(self):
        """[("stage/kernel", device ms)] of the last profiled run_device, in launch order."""
        ms = (ctypes.c_double * 64)()
        names = (ctypes.c_char_p * 64)()
        n = C.hdog_stage_times(self._h, ms, names, 64)
        return [(names[k].decode(), ms[k]) for k in range(n)]

    def synth_volume(self, shape, blobs, seed, mean0, mean1, sd, out=None):
        """Render a synthetic stack on the device (see hdog.synth for the model)."""
        D, H, W = shape
        img = torch.empty((D, H, W), dtype=torch.uint16, device=self.device) if out is None else out
        scratch = torch.empty((D, H, W), dtype=torch.float32, device=self.device)
        blobs = np.ascontiguousarray(blobs, np.float64)
        self._check(C.hdog_synth_volume(self._h, self._stream(), _ptr(img), _ptr(scratch), W, H, D,
                                        int(seed), float(mean0), float(mean1), float(sd),
                                        ctypes.c_void_p(blobs.ctypes.data), len(blobs)))
        del scratch
        return img
