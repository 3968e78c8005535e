"""Pins of the FP64 reference of the differentiable blob response (CPU).

The reference (oracle/soft_reference.py) is what the GPU gradients are checked against,
so it is itself checked: its stages against the float32 detector oracle and scipy, and
its autograd gradient against central finite differences (FP64, step 1e-3, mask frozen).
"""
import numpy as np
import pytest
import torch

import oracle
import soft_reference as ref
from hdog import synth


def small_case(seed=3, shape=(12, 20, 22), n_blobs=6):
    vol = synth.make_volume("c0", shape=shape, n_blobs=n_blobs, seed=seed).astype(np.float64)
    v = torch.from_numpy((vol - vol.mean()) / vol.std())
    blobs = synth.blob_table(shape, n_blobs, seed)
    target = torch.from_numpy(ref.blob_target(shape, blobs))
    return v, target


def test_weights_match_scipy_kernel():
    from scipy.ndimage._filters import _gaussian_kernel1d
    for s, r in ((2.0, 8), (1.6, 6), (3.2, 13)):
        np.testing.assert_allclose(ref.weights(torch.tensor(s, dtype=torch.float64), r).numpy(),
                                   _gaussian_kernel1d(s, 0, r), rtol=1e-14)


def test_gaussian_and_hessian_match_detector_oracle():
    v, _ = small_case()
    s1, r1 = (2.0, 2.0, 2.0), (8, 8, 8)
    g = ref.sep_gauss(v, torch.tensor(s1, dtype=torch.float64), r1).numpy()
    w = ref.weights(torch.tensor(2.0, dtype=torch.float64), 8).numpy().astype(np.float32)
    go = oracle.gaussian3d(v.numpy().astype(np.float32), w, w)
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-5 * np.abs(go).max())
    dog = (torch.randn(9, 10, 11, dtype=torch.float64, generator=torch.Generator().manual_seed(0)))
    h = ref.hessian(dog)
    ho, _ = oracle.hessian3d(dog.numpy().astype(np.float32))
    for k in range(6):
        np.testing.assert_allclose(h[k].numpy(), ho[k], rtol=0, atol=1e-6)


def test_adjoint_of_zero_padded_central_difference_is_its_negative():
    a = torch.randn(7, 8, 9, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    b = torch.randn(7, 8, 9, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    for axis in range(3):
        lhs = (ref.cdiff(a, axis) * b).sum()
        rhs = (a * -ref.cdiff(b, axis)).sum()
        assert abs(float(lhs - rhs)) < 1e-12


@pytest.mark.parametrize("which", ["sigma1", "sigma2", "tau"])
def test_autograd_matches_central_finite_differences(which):
    v, target = small_case()
    s1 = torch.tensor([1.6, 2.0, 2.0], dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor([2.56, 3.2, 3.2], dtype=torch.float64, requires_grad=True)
    tau = torch.tensor(0.05, dtype=torch.float64, requires_grad=True)
    r1, r2 = (6, 8, 8), (10, 13, 13)
    loss, mask, _ = ref.forward(v, s1, s2, tau, target, r1, r2)
    loss.backward()
    h = 1e-3
    params = {"sigma1": s1, "sigma2": s2, "tau": tau}
    p = params[which]
    grad = p.grad.reshape(-1)
    for k in range(grad.numel()):
        with torch.no_grad():
            def at(delta):
                q = {n: t.detach().clone() for n, t in params.items()}
                q[which].reshape(-1)[k] += delta
                return float(ref.forward(v, q["sigma1"], q["sigma2"], q["tau"], target, r1, r2,
                                         mask=mask)[0])
            fd = (at(h) - at(-h)) / (2 * h)
            fd2 = (at(h / 2) - at(-h / 2)) / h
        g = float(grad[k])
        # the central difference at step 1e-3 has an O(h^2) truncation error (visible for tau,
        # where the loss is most curved); its Richardson extrapolation is O(h^4)
        assert abs(fd - g) <= 1e-3 * max(1e-3, abs(g)), (which, k, fd, g)
        rich = (4 * fd2 - fd) / 3
        assert abs(rich - g) <= 1e-6 * max(1.0, abs(g)), (which, k, rich, g)


def test_reference_normalization_is_the_detectors():
    # on integer inputs the float64 restatement equals the FP32 oracle's normalization to
    # FP32 rounding (same clamp-to-edge min / max and box windows, same formula)
    import oracle
    rng = np.random.default_rng(12)
    img = rng.integers(0, 4000, (12, 30, 26), dtype=np.uint16)
    want = oracle.normalize3d(img, 150, 4, 2).astype(np.float64)
    got = ref.normalize(torch.from_numpy(img.astype(np.float64)), 150.0, 4, 2).numpy()
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-7)


def test_reference_normalization_gradient_is_the_subgradient():
    # continuous values (no ties in any window): the normalization is differentiable and
    # autograd's gradient, which routes each window extreme to its voxel, matches central
    # finite differences
    rng = np.random.default_rng(13)
    v = torch.from_numpy(rng.random((5, 7, 6)) * 100).requires_grad_()
    assert torch.autograd.gradcheck(lambda x: ref.normalize(x, 10.0, 2, 1), (v,), eps=1e-6,
                                    atol=1e-7, rtol=1e-5)
