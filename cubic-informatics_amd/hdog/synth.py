"""Seeded synthetic light-sheet stacks: Gaussian noise background plus Gaussian blobs.

    v   = noise + sum_b A_b * exp(-sum_a (x_a - c_a)^2 / (2 sigma_a^2))   (within 4 sigma_b)
    u16 = clip(rint(v), 0, 65535)

Blob parameters come from NumPy ``Generator(PCG64(seed))``.  Noise is counter
based (splitmix64 of the voxel index) so the CPU generator here and the HIP
generator in the library draw the same Gaussian variates for a voxel; the two
renderings can still differ in the last bit of the float blob sum.

Stacks are written in the reference's on-disk layout
``{Y}/{Y}_{X}/{Z:06d}.bin`` of little-endian uint16 slices (README.md:57-64 of
the reference).
"""
import os

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _unit(h):
    # (0, 1) open interval from the top 53 bits
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def noise_normal(seed, start, count):
    """Standard normal variates for voxel indices [start, start+count) (Box-Muller)."""
    idx = np.arange(start, start + count, dtype=np.uint64)
    key = np.uint64((int(seed) * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        h1 = _splitmix64((idx * np.uint64(2)) ^ key)
        h2 = _splitmix64((idx * np.uint64(2) + np.uint64(1)) ^ key)
    u1, u2 = _unit(h1), _unit(h2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# Synthetic workloads: shape (z, y, x), blob count, noise model and blob model.
WORKLOADS = {
    # 128x256x256, noise N(200, 20^2), 2000 isotropic blobs sigma U[1.5,3.5], A U[500,3000]
    "c0": dict(shape=(128, 256, 256), n_blobs=2000, seed=0, mean0=200.0, mean1=200.0,
               sd=20.0, sigma=(1.5, 3.5), zfactor=1.0, amp=(500.0, 3000.0)),
    # 512x2160x2560, background mean 150 -> 400 along z, 150000 blobs with sigma_z = 0.8 sigma_xy
    "c1": dict(shape=(512, 2160, 2560), n_blobs=150000, seed=1, mean0=150.0, mean1=400.0,
               sd=20.0, sigma=(1.5, 3.5), zfactor=0.8, amp=(500.0, 3000.0)),
    # differentiable form: 128x512x512, config-0 noise, 12000 blobs
    "c4": dict(shape=(128, 512, 512), n_blobs=12000, seed=4, mean0=200.0, mean1=200.0,
               sd=20.0, sigma=(1.5, 3.5), zfactor=1.0, amp=(500.0, 3000.0)),
}


def blob_table(shape, n_blobs, seed, sigma=(1.5, 3.5), zfactor=1.0, amp=(500.0, 3000.0),
               margin=4):
    """(n, 7) float64 table: cz, cy, cx, sz, sy, sx, A."""
    rng = np.random.Generator(np.random.PCG64(seed))
    D, H, W = shape
    cz = rng.uniform(margin, D - 1 - margin, n_blobs)
    cy = rng.uniform(margin, H - 1 - margin, n_blobs)
    cx = rng.uniform(margin, W - 1 - margin, n_blobs)
    sxy = rng.uniform(sigma[0], sigma[1], n_blobs)
    A = rng.uniform(amp[0], amp[1], n_blobs)
    return np.stack([cz, cy, cx, sxy * zfactor, sxy, sxy, A], axis=1)


def render(shape, blobs, seed, mean0, mean1, sd):
    """Render noise + blobs into a uint16 volume (CPU, float64 accumulation)."""
    D, H, W = shape
    N = D * H * W
    v = noise_normal(seed, 0, N).reshape(D, H, W) * sd
    zmean = mean0 + (mean1 - mean0) * np.arange(D, dtype=np.float64) / max(D - 1, 1)
    v += zmean[:, None, None]
    np.maximum(v, 0.0, out=v)  # noise clipped at 0
    for cz, cy, cx, sz, sy, sx, A in blobs:
        lo = [max(0, int(np.floor(c - 4 * s))) for c, s in ((cz, sz), (cy, sy), (cx, sx))]
        hi = [min(n, int(np.ceil(c + 4 * s)) + 1) for c, s, n in ((cz, sz, D), (cy, sy, H), (cx, sx, W))]
        z = np.arange(lo[0], hi[0], dtype=np.float64)
        y = np.arange(lo[1], hi[1], dtype=np.float64)
        x = np.arange(lo[2], hi[2], dtype=np.float64)
        gz = np.exp(-(z - cz) ** 2 / (2 * sz * sz))
        gy = np.exp(-(y - cy) ** 2 / (2 * sy * sy))
        gx = np.exp(-(x - cx) ** 2 / (2 * sx * sx))
        v[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += A * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return np.clip(np.rint(v), 0, 65535).astype(np.uint16)


def make_volume(name=None, shape=None, n_blobs=None, seed=None, **over):
    """Build a named configuration, optionally overriding its shape/blob count/seed."""
    cfg = dict(WORKLOADS[name or "c0"])
    cfg.update(over)
    if shape is not None:
        cfg["shape"] = tuple(shape)
    if n_blobs is not None:
        cfg["n_blobs"] = n_blobs
    if seed is not None:
        cfg["seed"] = seed
    blobs = blob_table(cfg["shape"], cfg["n_blobs"], cfg["seed"], cfg["sigma"],
                       cfg["zfactor"], cfg["amp"])
    return render(cfg["shape"], blobs, cfg["seed"], cfg["mean0"], cfg["mean1"], cfg["sd"])


def write_stack(vol, basedir, yname="000000", xname="000000"):
    """Write vol (z, y, x) as {basedir}/{Y}/{Y}_{X}/{Z:06d}.bin slices; returns the paths."""
    d = os.path.join(basedir, yname, f"{yname}_{xname}")
    os.makedirs(d, exist_ok=True)
    paths = []
    for z in range(vol.shape[0]):
        p = os.path.join(d, f"{z:06d}.bin")
        np.ascontiguousarray(vol[z], dtype="<u2").tofile(p)
        paths.append(p)
    return paths
