"""Differentiable form on the GPU: hand-written adjoint kernels against the FP64 reference
(torch float64 autograd, oracle/soft_reference.py) and against central finite differences
(FP64, step 1e-3, Richardson-extrapolated, mask held fixed); at the workload's own size
(128 x 512 x 512, 12000 blobs) with the local-contrast normalization in front, against
the FP64 reference gradient and finite differences at single input voxels."""
import numpy as np
import pytest
import torch

import soft_reference as ref
from hdog import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hd():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hdog import HDoG
    h = HDoG(0)
    yield h
    h.close()


def case(shape=(32, 64, 64), n_blobs=40, seed=4):
    vol = synth.make_volume("c4", shape=shape, n_blobs=n_blobs, seed=seed).astype(np.float64)
    v = ((vol - vol.mean()) / vol.std()).astype(np.float32)
    blobs = synth.blob_table(shape, n_blobs, seed)
    return v, ref.blob_target(shape, blobs)


S1, S2, TAU = [1.6, 2.0, 2.0], [2.56, 3.2, 3.2], 0.05


def test_soft_forward_and_gradients_match_fp64_reference(hd):
    from hdog import soft
    v, t = case()
    rad1, rad2 = soft.radii(S1), soft.radii(S2)
    vg = torch.from_numpy(v).to(hd.device).requires_grad_()
    tg = torch.from_numpy(t).to(hd.device)
    s1 = torch.tensor(S1, dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor(S2, dtype=torch.float64, requires_grad=True)
    tau = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    loss = soft.soft_loss(hd, vg, tg, s1, s2, tau, rad1, rad2)
    loss.backward()
    _, count, _, mask = soft.run_forward(hd, vg.detach(), tg, S1, S2, TAU, rad1, rad2)
    mask = mask.cpu()
    # reference on the same float32 input values, with the GPU mask held fixed
    v64 = torch.from_numpy(v.astype(np.float64)).requires_grad_()
    r1 = torch.tensor(S1, dtype=torch.float64, requires_grad=True)
    r2 = torch.tensor(S2, dtype=torch.float64, requires_grad=True)
    rt = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    rl, rmask, _ = ref.forward(v64, r1, r2, rt, torch.from_numpy(t), rad1, rad2, mask=mask.bool())
    rl.backward()
    _, own_mask, _ = ref.forward(v64.detach(), r1.detach(), r2.detach(), rt.detach(),
                                 torch.from_numpy(t), rad1, rad2)
    jac = (own_mask & mask.bool()).sum().item() / max(1, (own_mask | mask.bool()).sum().item())
    print(f"mask Jaccard GPU FP32 vs FP64: {jac:.6f}, |M| = {count}")
    assert count > 1000 and jac > 0.995
    assert abs(float(loss) - float(rl)) <= 1e-4 * abs(float(rl))
    for g, r, name in ((s1.grad, r1.grad, "sigma1"), (s2.grad, r2.grad, "sigma2"),
                       (tau.grad, rt.grad, "tau")):
        np.testing.assert_allclose(g.numpy(), r.numpy(), rtol=2e-3, atol=1e-6, err_msg=name)
    gv, rv = vg.grad.cpu().double(), v64.grad
    rel = float((gv - rv).norm() / rv.norm())
    print(f"input gradient relative L2 error: {rel:.2e}")
    assert rel < 2e-3


def test_soft_gradients_match_finite_differences(hd):
    from hdog import soft
    v, t = case(shape=(24, 48, 48), n_blobs=20, seed=6)
    rad1, rad2 = soft.radii(S1), soft.radii(S2)
    vg = torch.from_numpy(v).to(hd.device)
    tg = torch.from_numpy(t).to(hd.device)
    s1 = torch.tensor(S1, dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor(S2, dtype=torch.float64, requires_grad=True)
    tau = torch.tensor(TAU, dtype=torch.float64, requires_grad=True)
    soft.soft_loss(hd, vg, tg, s1, s2, tau, rad1, rad2).backward()
    _, _, _, mask = soft.run_forward(hd, vg, tg, S1, S2, TAU, rad1, rad2)
    mask = mask.cpu().bool()
    v64 = torch.from_numpy(v.astype(np.float64))
    t64 = torch.from_numpy(t)
    base = {"s1": S1, "s2": S2, "tau": [TAU]}
    grads = {"s1": s1.grad.numpy(), "s2": s2.grad.numpy(), "tau": np.array([float(tau.grad)])}
    h = 1e-3

    def loss_at(name, k, delta):
        q = {n: list(vals) for n, vals in base.items()}
        q[name][k] += delta
        with torch.no_grad():
            return float(ref.forward(v64, torch.tensor(q["s1"], dtype=torch.float64),
                                     torch.tensor(q["s2"], dtype=torch.float64),
                                     torch.tensor(q["tau"][0], dtype=torch.float64), t64, rad1,
                                     rad2, mask=mask)[0])
    for name, vals in base.items():
        for k in range(len(vals)):
            fd = (loss_at(name, k, h) - loss_at(name, k, -h)) / (2 * h)
            fd2 = (loss_at(name, k, h / 2) - loss_at(name, k, -h / 2)) / h
            rich = (4 * fd2 - fd) / 3
            g = grads[name][k]
            assert abs(rich - g) <= 2e-3 * max(1e-2, abs(rich)), (name, k, rich, g)


def test_fitting_sigmas_and_tau_reduces_the_loss(hd):
    from hdog import soft
    v, t = case(shape=(32, 64, 64), n_blobs=40, seed=9)
    vg = torch.from_numpy(v).to(hd.device)
    tg = torch.from_numpy(t).to(hd.device)
    model = soft.BlobResponseLoss(hd, [1.4, 1.8, 1.8], [2.4, 3.0, 3.0], 0.05)
    opt = torch.optim.Adam(model.parameters(), lr=0.02)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        loss = model(vg, tg)
        loss.backward()
        opt.step()
        losses.append(float(loss))
    print("fit losses:", [round(x, 5) for x in losses])
    assert losses[-1] < losses[0]


@pytest.fixture(scope="module")
def c4_case(hd):
    # the differentiable form's workload: 128 x 512 x 512 with the config-0 noise model and
    # 12000 blobs, as float32 with a uniform [0, 1) jitter so that no window has two equal
    # extremes (the min / max filters are then differentiable where checked)
    cfg = synth.WORKLOADS["c4"]
    shape = cfg["shape"]
    blobs = synth.blob_table(shape, cfg["n_blobs"], cfg["seed"], cfg["sigma"], cfg["zfactor"],
                             cfg["amp"])
    img = hd.synth_volume(shape, blobs, cfg["seed"], cfg["mean0"], cfg["mean1"], cfg["sd"])
    gen = torch.Generator(device=hd.device).manual_seed(44)
    v = img.float() + torch.rand(shape, device=hd.device, generator=gen)
    t = torch.from_numpy(ref.blob_target(shape, blobs)).to(hd.device)
    return v, t


S1C, S2C, NORM = [2.0, 2.0, 2.0], [3.2, 3.2, 3.2], (0.0, 10, 10)


def test_normalized_soft_loss_at_c4_size_matches_fp64(hd, c4_case):
    from hdog import soft
    v, t = c4_case
    rad1, rad2 = soft.radii(S1C), soft.radii(S2C)
    # tau: the median |det H| over the candidates of the FP64 reference (fixed from here on)
    with torch.no_grad():
        n64 = ref.normalize(v.double(), *NORM)
        _, m0, dog0 = ref.forward(n64, torch.tensor(S1C, dtype=torch.float64),
                                  torch.tensor(S2C, dtype=torch.float64),
                                  torch.tensor(1.0, dtype=torch.float64), t, rad1, rad2)
        xx, xy, xz, yy, yz, zz = ref.hessian(dog0)
        det = xx * yy * zz + 2 * xy * yz * xz - xx * yz * yz - yy * xz * xz - zz * xy * xy
        TAUC = float(det[m0.bool()].abs().median())
        del n64, dog0, xx, xy, xz, yy, yz, zz, det
    vg = v.clone().requires_grad_()
    s1 = torch.tensor(S1C, dtype=torch.float64, requires_grad=True)
    s2 = torch.tensor(S2C, dtype=torch.float64, requires_grad=True)
    tau = torch.tensor(TAUC, dtype=torch.float64, requires_grad=True)
    n = soft.normalize(hd, vg, *NORM)
    loss = soft.soft_loss(hd, n, t, s1, s2, tau, rad1, rad2)
    loss.backward()
    _, count, _, mask = soft.run_forward(hd, n.detach(), t, S1C, S2C, TAUC, rad1, rad2)
    v64 = v.double().requires_grad_()
    r1 = torch.tensor(S1C, dtype=torch.float64, requires_grad=True)
    r2 = torch.tensor(S2C, dtype=torch.float64, requires_grad=True)
    rt = torch.tensor(TAUC, dtype=torch.float64, requires_grad=True)
    rl, _, _ = ref.forward(ref.normalize(v64, *NORM), r1, r2, rt, t, rad1, rad2, mask=mask.bool())
    rl.backward()
    jac = float((m0.bool() & mask.bool()).sum()) / max(1, float((m0.bool() | mask.bool()).sum()))
    print(f"C4 {tuple(v.shape)}: |M| = {count}, mask Jaccard FP32 vs FP64 {jac:.6f}, "
          f"tau {TAUC:.3e}, loss {float(loss):.6f} vs {float(rl):.6f}", flush=True)
    assert count > 100000 and jac > 0.995
    assert abs(float(loss) - float(rl)) <= 1e-4 * abs(float(rl))
    for g, r, name in ((s1.grad, r1.grad, "sigma1"), (s2.grad, r2.grad, "sigma2"),
                       (tau.grad, rt.grad, "tau")):
        np.testing.assert_allclose(g.numpy(), r.cpu().numpy(), rtol=2e-3, atol=1e-6, err_msg=name)
    gv, rv = vg.grad.double(), v64.grad
    rel = float((gv - rv).norm() / rv.norm())
    print(f"C4 input gradient relative L2 error vs FP64: {rel:.2e}", flush=True)
    assert rel < 2e-3
    # central finite differences of the FP64 loss (mask held fixed) at the input voxels
    # with the largest gradients, which include window extremes of the min / max filters
    idx = torch.topk(rv.abs().flatten(), 4).indices
    h = 1e-4
    for i in idx.tolist():
        with torch.no_grad():
            def lf(delta):
                w = v.double().flatten().clone()
                w[i] += delta
                return float(ref.forward(ref.normalize(w.view(v.shape), *NORM),
                                         r1.detach(), r2.detach(), rt.detach(), t, rad1, rad2,
                                         mask=mask.bool())[0])
            fd = (lf(h) - lf(-h)) / (2 * h)
        g = float(vg.grad.flatten()[i])
        print(f"voxel {i}: finite difference {fd:.6e}, GPU gradient {g:.6e}", flush=True)
        assert abs(fd - g) <= 5e-3 * abs(fd) + 1e-12, (i, fd, g)
