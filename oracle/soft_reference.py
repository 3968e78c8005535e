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
    """Separable zero-padded Gaussian of a (D, H, W) float64 tensor: one 1-D convolution per
    axis over the lines along it (the weights are symmetric, so correlation = convolution)."""
    x = v
    for axis, (s, r) in enumerate(zip(sig_zyx, rad_zyx)):
        w = weights(s, r).to(v.device)
        xl = x.movedim(axis, -1)
        shp = xl.shape
        y = F.conv1d(xl.reshape(-1, 1, shp[-1]), w.reshape(1, 1, -1), padding=r)
        x = y.reshape(shp).movedim(-1, axis)
    return x


def _clamp_pad(x, axis, r):
    pad = [0, 0, 0, 0, 0, 0]
    pad[2 * (2 - axis)] = r
    pad[2 * (2 - axis) + 1] = r
    return F.pad(x[None, None], pad, mode="replicate")


def _window_max(x, axis, r):
    k = [1, 1, 1]
    k[axis] = 2 * r + 1
    return F.max_pool3d(_clamp_pad(x, axis, r), kernel_size=k, stride=1)[0, 0]


def _box(x, axis, r):
    k = [1, 1, 1]
    k[axis] = 2 * r + 1
    return F.avg_pool3d(_clamp_pad(x, axis, r), kernel_size=k, stride=1)[0, 0]


def normalize(v, min_intensity, rxy, rz):
    """Local-contrast normalization of a float64 volume, src/3D/NormalizeFilter.cu:53-654:
    clamp-to-edge min / max filters and box means of radius (rz, rxy, rxy) (z, y, x),
    n = v >= m ? (v - E) / (D - E) : 0.  Autograd takes each window extreme's gradient to
    the voxel that attains it."""
    rad = (rz, rxy, rxy)
    mn, mx = v, v
    for axis in (2, 1, 0):
        mn = -_window_max(-mn, axis, rad[axis])
        mx = _window_max(mx, axis, rad[axis])
    E, Dl = mn, mx
    for axis in (2, 1, 0):
        E = _box(E, axis, rad[axis])
        Dl = _box(Dl, axis, rad[axis])
    return torch.where(v >= min_intensity, (v - E) / (Dl - E), torch.zeros_like(v))


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
    t = target.to(torch.float64).to(det.device)
    terms = F.softplus(u) - t * u
    m = mask.bool().to(det.device)
    loss = terms[m].mean() if m.any() else terms.sum() * 0.0
    return loss, mask, dog


def blob_target(shape, blobs):
    """Blob centres (rounded) dilated by one voxel (a 3x3x3 cube)."""
    t = np.zeros(shape, np.uint8)
    D, H, W = shape
    for cz, cy, cx in np.rint(blobs[:, :3]).astype(int):
        t[max(cz - 1, 0):min(cz + 2, D), max(cy - 1, 0):min(cy + 2, H), max(cx - 1, 0):min(cx + 2, W)] = 1
    return t
