"""Differentiable form on the GPU: hand-written adjoint kernels against the FP64 reference
(torch float64 autograd, oracle/soft_reference.py) and against central finite differences
(FP64, step 1e-3, Richardson-extrapolated, mask held fixed)."""
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
