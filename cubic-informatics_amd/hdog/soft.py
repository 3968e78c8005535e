"""Differentiable DoG -> Hessian soft blob response (torch.autograd.Function).

    dog  = (sx1 sy1 sz1)^gamma * (G_sigma1 v - G_sigma2 v)       zero-padded, fixed radii
    H    = Hessian(dog)  (the detector's central differences),   M = candidate mask of H
    loss = mean over M of BCE(sigmoid(-det(H) / tau), target)

optionally behind the detector's local-contrast normalization of the input (normalize):

    n = v >= m ? (v - E) / (D - E) : 0,  E = box(min3(v)), D = box(max3(v))

whose backward pass takes the min / max filters by their subgradient (the gradient of a
window's extreme goes to the voxel that attains it) and the clamped box means by their
transposes (hdog_soft_normalize_forward / _backward).

Forward and backward run the hand-written HIP kernels of libhdog.so
(hdog_soft_forward / hdog_soft_backward): the DoG volume is checkpointed in the forward
pass and H is recomputed from it in the backward pass.  The mask M is treated as a
constant.  Sigmas are given in (z, y, x) order like the volume's shape; each axis has its
own sigma and its own radius int(truncate * sigma_init + 0.5), fixed at construction, so
the loss is smooth in the sigmas.
"""
import ctypes

import torch

from . import _capi as C


def radii(sigma_zyx, truncate=4.0):
    return tuple(int(truncate * float(s) + 0.5) for s in sigma_zyx)


def run_forward(hd, v, target, sigma1, sigma2, tau, rad1, rad2, gamma=1.0):
    """Forward pass only: (loss, |M|, dog, mask).  v: float32 (D, H, W) on the device,
    target: uint8 of the same shape, sigma1 / sigma2 (z, y, x)."""
    D, H, W = v.shape
    s1 = [float(x) for x in torch.as_tensor(sigma1).detach().cpu().tolist()][::-1]  # -> x, y, z
    s2 = [float(x) for x in torch.as_tensor(sigma2).detach().cpu().tolist()][::-1]
    r1, r2 = list(rad1)[::-1], list(rad2)[::-1]
    v = v.contiguous()
    target = target.contiguous()
    dog = torch.empty_like(v)
    mask = torch.empty(v.shape, dtype=torch.uint8, device=v.device)
    loss = ctypes.c_double()
    count = ctypes.c_int64()
    hd._check(C.hdog_soft_forward(hd._h, hd._stream(), ctypes.c_void_p(v.data_ptr()),
                                  ctypes.c_void_p(target.data_ptr()), W, H, D,
                                  (ctypes.c_double * 3)(*s1), (ctypes.c_double * 3)(*s2),
                                  (ctypes.c_int * 3)(*r1), (ctypes.c_int * 3)(*r2),
                                  float(gamma), float(tau), ctypes.c_void_p(dog.data_ptr()),
                                  ctypes.c_void_p(mask.data_ptr()), ctypes.byref(loss),
                                  ctypes.byref(count)))
    return loss.value, count.value, dog, mask


class _SoftFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, sigma1, sigma2, tau, target, hd, rad1, rad2, gamma):
        v = v.contiguous()
        target = target.contiguous()
        loss, count, dog, mask = run_forward(hd, v, target, sigma1, sigma2, tau, rad1, rad2, gamma)
        ctx.save_for_backward(v, dog, mask, target)
        ctx.hd = hd
        ctx.s1 = [float(x) for x in sigma1.detach().cpu().tolist()][::-1]
        ctx.s2 = [float(x) for x in sigma2.detach().cpu().tolist()][::-1]
        ctx.r1, ctx.r2 = list(rad1)[::-1], list(rad2)[::-1]
        ctx.gamma, ctx.tau, ctx.count = float(gamma), float(tau), count
        ctx.dev_sigma = sigma1.device
        return torch.tensor(loss, dtype=torch.float64, device=sigma1.device)

    @staticmethod
    def backward(ctx, grad_out):
        v, dog, mask, target = ctx.saved_tensors
        hd = ctx.hd
        D, H, W = v.shape
        gv = torch.empty_like(v)
        g1 = (ctypes.c_double * 3)()
        g2 = (ctypes.c_double * 3)()
        gt = ctypes.c_double()
        hd._check(C.hdog_soft_backward(hd._h, hd._stream(), ctypes.c_void_p(v.data_ptr()),
                                       ctypes.c_void_p(dog.data_ptr()),
                                       ctypes.c_void_p(mask.data_ptr()),
                                       ctypes.c_void_p(target.data_ptr()), W, H, D,
                                       (ctypes.c_double * 3)(*ctx.s1),
                                       (ctypes.c_double * 3)(*ctx.s2),
                                       (ctypes.c_int * 3)(*ctx.r1), (ctypes.c_int * 3)(*ctx.r2),
                                       ctx.gamma, ctx.tau, ctx.count, float(grad_out), 
                                       ctypes.c_void_p(gv.data_ptr()), g1, g2, ctypes.byref(gt)))
        dev = ctx.dev_sigma
        gs1 = torch.tensor(list(g1)[::-1], dtype=torch.float64, device=dev)  # -> z, y, x
        gs2 = torch.tensor(list(g2)[::-1], dtype=torch.float64, device=dev)
        return gv, gs1, gs2, torch.tensor(gt.value, dtype=torch.float64, device=dev), \
            None, None, None, None, None


class _NormFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, hd, min_intensity, rxy, rz):
        v = v.contiguous()
        D, H, W = v.shape
        norm = torch.empty_like(v)
        ed = torch.empty((D, H, W, 2), dtype=torch.float32, device=v.device)
        arg = torch.empty((2, D, H, W), dtype=torch.int32, device=v.device)  # uint32 indices
        hd._check(C.hdog_soft_normalize_forward(hd._h, hd._stream(), ctypes.c_void_p(v.data_ptr()),
                                                ctypes.c_void_p(norm.data_ptr()),
                                                ctypes.c_void_p(ed.data_ptr()),
                                                ctypes.c_void_p(arg.data_ptr()),
                                                float(min_intensity), W, H, D, int(rxy), int(rz)))
        ctx.save_for_backward(v, ed, arg)
        ctx.hd, ctx.args = hd, (float(min_intensity), int(rxy), int(rz))
        return norm

    @staticmethod
    def backward(ctx, grad_norm):
        v, ed, arg = ctx.saved_tensors
        hd = ctx.hd
        m, rxy, rz = ctx.args
        D, H, W = v.shape
        gn = grad_norm.contiguous().float()
        gv = torch.empty_like(v)
        hd._check(C.hdog_soft_normalize_backward(hd._h, hd._stream(), ctypes.c_void_p(v.data_ptr()),
                                                 ctypes.c_void_p(ed.data_ptr()),
                                                 ctypes.c_void_p(arg.data_ptr()),
                                                 ctypes.c_void_p(gn.data_ptr()),
                                                 ctypes.c_void_p(gv.data_ptr()), m, W, H, D,
                                                 rxy, rz))
        return gv, None, None, None, None


def normalize(hd, v, min_intensity, radius_xy, radius_z):
    """Differentiable local-contrast normalization of a float32 (D, H, W) device volume."""
    return _NormFn.apply(v, hd, min_intensity, radius_xy, radius_z)


class BlobResponseLoss(torch.nn.Module):
    """Loss module with learnable DoG sigmas (z, y, x) and threshold tau; with
    `normalization=(min_intensity, radius_xy, radius_z)` the input is normalized first."""

    def __init__(self, hd, sigma1_zyx, sigma2_zyx, tau, truncate=4.0, gamma=1.0,
                 normalization=None):
        super().__init__()
        self.hd = hd
        self.sigma1 = torch.nn.Parameter(torch.tensor(sigma1_zyx, dtype=torch.float64))
        self.sigma2 = torch.nn.Parameter(torch.tensor(sigma2_zyx, dtype=torch.float64))
        self.tau = torch.nn.Parameter(torch.tensor(float(tau), dtype=torch.float64))
        self.rad1 = radii(sigma1_zyx, truncate)
        self.rad2 = radii(sigma2_zyx, truncate)
        self.gamma = gamma
        self.normalization = normalization

    def forward(self, v, target):
        if self.normalization is not None:
            v = normalize(self.hd, v, *self.normalization)
        return _SoftFn.apply(v, self.sigma1, self.sigma2, self.tau, target, self.hd, self.rad1,
                             self.rad2, self.gamma)


def soft_loss(hd, v, target, sigma1, sigma2, tau, rad1, rad2, gamma=1.0):
    """Functional form; sigma1 / sigma2 / tau are float64 tensors (z, y, x)."""
    return _SoftFn.apply(v, sigma1, sigma2, tau, target, hd, rad1, rad2, gamma)
