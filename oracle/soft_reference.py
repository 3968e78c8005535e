"""FP64 reference of the differentiable blob response (torch float64 on the CPU).

TEST INFRASTRUCTURE ONLY: the gradient checks of the differentiable form compare the HIP
forward / backward against this restatement, and it is itself checked against central
finite differences and against the CPU oracle's float32 stages (tests/test_soft_reference.py).

Same definition as the library (hdog.soft): anisotropic normalised Gaussians with fixed
radii and zero padding, dog = (sx1 sy1 sz1)^gamma (G1 v - G2 v), the Hessian by the
detector's nine zero-padded central differences in the order of
src/3D/Hessian3DFilter_element.cu:373-383, the candidate mask of
src/3D/Hessian3DFilter_element.cu:317-321, and the mean BCE of sigmoid(-det / tau) over the
mask.  Sigmas are (z, y, x).
"""
import numpy as np
import torch
import torch.nn.functional as F


def weights(sigma, r):
    i = torch.arange(-r, r + 1, dtype=torch.float64)
    e = torch.exp(-i * i / (2.0 * sigma * sigma))
    return e / e.sum()


def sep_gauss(v, sig_zyx, rad_zyx):
    """Separable zero-padded Gaussian of a (D, H, W) float64 tensor."""
    x = v[None, None]
    for axis, (s, r) in enumerate(zip(sig_zyx, rad_zyx)):
        w = weights(s, r)
        shape = [1, 1, 1, 1, 1]
        shape[2 + axis] = 2 * r + 1
        pad = [0, 0, 0]
        pad[axis] = r
        x = F.conv3d(x, w.reshape(shape), padding=tuple(pad))
    return x[0, 0]


def cdiff(f, axis):
    """Zero-padded central difference along axis (0 = z, 1 = y, 2 = x)."""
    pad = [0, 0, 0, 0, 0, 0]
    pad[2 * (2 - axis)] = 1
    pad[2 * (2 - axis) + 1] = 1
    g = F.pad(f[None, None], pad)[0, 0]
    hi = g.narrow(axis, 2, f.shape[axis])
    lo = g.narrow(axis, 0, f.shape[axis])
    return (hi - lo) * 0.5


def hessian(dog):
    hx, hy, hz = cdiff(dog, 2), cdiff(dog, 1), cdiff(dog, 0)
    return (cdiff(hx, 2), cdiff(hx, 1), cdiff(hx, 0), cdiff(hy, 1), cdiff(hy, 0), cdiff(hz, 0))


def forward(v, sigma1, sigma2, tau, target, rad1, rad2, gamma=1.0, mask=None):
    """Returns (loss, mask, dog).  sigma1/sigma2: float64 tensors (z, y, x)."""
    s = (sigma1[0] * sigma1[1] * sigma1[2]) ** gamma
    dog = s * (sep_gauss(v, sigma1, rad1) - sep_gauss(v, sigma2, rad2))
    xx, xy, xz, yy, yz, zz = hessian(dog)
    det = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy
    if mask is None:
        mask = ((xx < 0) & (xx * yy - xy * xy > 0) & (det < 0)).detach()
    u = -det / tau
    t = target.to(torch.float64)
    terms = F.softplus(u) - t * u
    m = mask.bool()
    loss = terms[m].mean() if m.any() else terms.sum() * 0.0
    return loss, mask, dog


def blob_target(shape, blobs):
    """Blob centres (rounded) dilated by one voxel (a 3x3x3 cube)."""
    t = np.zeros(shape, np.uint8)
    D, H, W = shape
    for cz, cy, cx in np.rint(blobs[:, :3]).astype(int):
        t[max(cz - 1, 0):min(cz + 2, D), max(cy - 1, 0):min(cy + 2, H), max(cx - 1, 0):min(cx + 2, W)] = 1
    return t
