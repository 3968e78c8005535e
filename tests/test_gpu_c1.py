"""The headline workload itself: the 512 x 2160 x 2560 tile (2.83 Gvoxel, beyond the
reference's 32-bit int indexing, src/3D/utils.h:122) with the config's noise and blob
model, rendered on the device.

- A full-plane 32 x 2160 x 2560 substack of it (z 240..271) against the CPU oracle:
  normalized volume, DoG, the six Hessian elements, candidate mask, labels and regions
  bit-exact (Hessian sums and eigen features to the usual bar).
- The full tile through hdog_run_device with the mask and the label volume materialized,
  checked on the device with torch, slab by slab (the oracle cannot hold the tile):
  background labels 0; every candidate's label is at most its own index and is the
  index of a candidate carrying that same label (so each label is the minimum index of
  the voxels that carry it), labels at or above 2^31 included; every pair of candidates
  related by one of the 13 positive linear offsets (rows and slices wrapping like the
  reference) carries one label; the regions' voxel counts add up to the candidates; and
  a second run gives identical labels and integer region arrays."""
import numpy as np
import pytest

import oracle
from hdog import synth

pytestmark = pytest.mark.gpu

C1P = dict(min_intensity=0, rxy=10, rz=10, sigma_xy1=2.0, sigma_z1=1.6, sigma_xy2=3.2,
           sigma_z2=2.56, truncate=4.0, gamma=1.0)


@pytest.fixture(scope="module")
def hd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hdog import HDoG
    h = HDoG(0)
    yield h
    h.close()


@pytest.fixture(scope="module")
def c1(hd):
    import torch
    cfg = synth.WORKLOADS["c1"]
    shape = cfg["shape"]
    blobs = synth.blob_table(shape, cfg["n_blobs"], cfg["seed"], cfg["sigma"], cfg["zfactor"],
                             cfg["amp"])
    img = hd.synth_volume(shape, blobs, cfg["seed"], cfg["mean0"], cfg["mean1"], cfg["sd"])
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    yield img
    del img
    torch.cuda.empty_cache()


def params_for(hd, shape):
    p = C1P
    return hd.make_params(shape, p["rxy"], p["rz"], p["min_intensity"], p["sigma_xy1"],
                          p["sigma_z1"], p["sigma_xy2"], p["sigma_z2"], p["truncate"], p["gamma"])


def test_full_plane_substack_bitexact(hd, c1):
    import torch
    from test_gpu_parity import assert_labels_equal, assert_regions_match
    sub = c1[240:272].contiguous()
    vol = sub.cpu().numpy()
    print(f"full-plane substack {vol.shape}", flush=True)
    o = oracle.substack(vol, C1P)
    bufs = hd.alloc_intermediates(vol.shape)
    R = hd.run_device(sub, params_for(hd, vol.shape), bufs)
    g = hd.download_regions()
    np.testing.assert_array_equal(bufs["norm"].cpu().numpy(), o["norm"])
    np.testing.assert_array_equal(bufs["dog"].cpu().numpy(), o["dog"])
    for k in range(6):
        np.testing.assert_array_equal(bufs["hess"][k].cpu().numpy(), o["hess"][k], err_msg=str(k))
    np.testing.assert_array_equal(bufs["pd"].cpu().numpy(), o["pd"])
    assert_labels_equal(bufs["labels"].cpu().numpy(), o["labels"])
    assert R == len(o["regions"]["label"]) and R > 100000
    assert_regions_match(g, o["regions"], o["hess"])
    del bufs
    torch.cuda.empty_cache()


def forward_offsets(W, H):
    WH = W * H
    return sorted({dz * WH + dy * W + dx for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                   for dx in (-1, 0, 1) if dz * WH + dy * W + dx > 0})


def test_full_tile_label_invariants_and_determinism(hd, c1):
    import torch
    D, H, W = c1.shape
    N = D * H * W
    assert N > 2**31
    dev = c1.device
    params = params_for(hd, (D, H, W))
    labels = torch.empty((D, H, W), dtype=torch.uint32, device=dev)
    pd = torch.empty((D, H, W), dtype=torch.uint8, device=dev)
    bufs = dict(norm=torch.empty((D, H, W), dtype=torch.float32, device=dev),
                dog=torch.empty((D, H, W), dtype=torch.float32, device=dev), pd=pd, labels=labels)
    print("full tile, run 1", flush=True)
    R = hd.run_device(c1, params, bufs)
    g1 = hd.download_regions()
    del bufs["norm"], bufs["dog"]
    torch.cuda.empty_cache()
    lab = labels.view(-1).view(torch.int32)  # uint32 bits
    m = pd.view(-1)
    S = 1 << 28  # voxels per slab
    cand = 0
    high = 0
    for s in range(0, N, S):
        e = min(N, s + S)
        L = lab[s:e].to(torch.int64) & 0xFFFFFFFF
        f = m[s:e] != 0
        if s == 0:
            f[0] = False  # voxel 0 is background whatever its mask
        assert not bool((L[~f] != 0).any()), f"background labels in [{s}, {e})"
        idx = torch.arange(s, e, device=dev, dtype=torch.int64)
        Lf = L[f]
        assert bool((Lf > 0).all()) and bool((Lf <= idx[f]).all()), f"label above index in [{s}, {e})"
        owner = lab[Lf].to(torch.int64) & 0xFFFFFFFF
        assert bool((owner == Lf).all()), f"a label's own voxel carries another label in [{s}, {e})"
        cand += int(f.sum().item())
        high += int((Lf >= 2**31).sum().item())
        del L, f, idx, Lf, owner
    print(f"{cand} candidates, {high} with labels >= 2^31", flush=True)
    assert high > 0
    for o in forward_offsets(W, H):
        for s in range(0, N - o, S):
            e = min(N - o, s + S)
            a, b = lab[s:e], lab[s + o:e + o]
            both = (m[s:e] != 0) & (m[s + o:e + o] != 0)
            if s == 0:
                both[0] = False
            assert not bool(((a != b) & both).any()), f"offset {o}: linked candidates differ in [{s}, {e})"
            del a, b, both
    assert R == len(g1["label"])
    assert int(g1["count"][1:].sum()) == cand
    # determinism: a second run, the same labels and integer region arrays
    first = labels.clone()
    print("full tile, run 2", flush=True)
    bufs2 = dict(pd=pd, labels=labels)
    R2 = hd.run_device(c1, params, bufs2)
    g2 = hd.download_regions()
    assert R2 == R
    assert torch.equal(first, labels)
    for k in ("label", "count", "size16", "isum", "imax", "centroid", "norm"):
        np.testing.assert_array_equal(g1[k], g2[k], err_msg=k)
    del first, labels, pd, bufs, bufs2
    torch.cuda.empty_cache()
