"""Batched stacks: one captured graph over several resident stacks of different shapes.

Each stack's regions from the graph must equal, bit for bit, what the per-stack path
(hdog_run_device, itself pinned to the oracle in test_gpu_parity.py) returns for it; one
stack is also checked against the oracle directly.  Replays must follow new image
contents, and region overflow must be reported per stack without touching memory out
of bounds."""
import numpy as np
import pytest

import oracle
from hdog import synth

pytestmark = pytest.mark.gpu

P = dict(min_intensity=0, rxy=10, rz=10, sigma_xy1=2.0, sigma_z1=1.6, sigma_xy2=3.2,
         sigma_z2=2.56, truncate=4.0, gamma=1.0)
SHAPES = [(24, 64, 72), (40, 48, 64), (16, 80, 40), (32, 64, 64), (12, 33, 47)]
EXACT = ("label", "count", "size16", "isum", "imax", "centroid", "norm")


@pytest.fixture(scope="module")
def hd():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hdog import HDoG
    h = HDoG(0)
    yield h
    h.close()


def params(hd, shape, precision="fp32"):
    return hd.make_params(shape, P["rxy"], P["rz"], P["min_intensity"], P["sigma_xy1"],
                          P["sigma_z1"], P["sigma_xy2"], P["sigma_z2"], P["truncate"],
                          P["gamma"], False, precision)


def volumes(seed0):
    return [synth.make_volume("c1", shape=s, n_blobs=max(4, int(np.prod(s) / 4000)), seed=seed0 + k)
            for k, s in enumerate(SHAPES)]


def per_stack(hd, vol, precision):
    import torch
    img = torch.from_numpy(vol).to(hd.device)
    hd.run_device(img, params(hd, vol.shape, precision))
    return hd.download_regions()


def assert_same(a, b, what):
    """Integer sums, centroids and max-normalized intensity bit-exact.  The Hessian sums are
    FP64 atomics whose order varies from run to run: equal to FP32 rounding noise, and the
    eigen features bit-exact wherever the six sums came out identical."""
    assert len(a["label"]) == len(b["label"]), (what, len(a["label"]), len(b["label"]))
    for k in EXACT:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    ha, hb = a["hsum"].astype(np.float64), b["hsum"].astype(np.float64)
    scale = np.abs(hb).max(axis=0, keepdims=True) + 1e-30
    assert np.all(np.abs(ha - hb) <= 2.4e-7 * np.abs(hb) + 1e-9 * scale), what
    same = np.all(a["hsum"] == b["hsum"], axis=1)
    for k in ("structureness", "blobness"):
        np.testing.assert_array_equal(a[k][same], b[k][same], err_msg=f"{what}: {k}")


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_batch_matches_per_stack_path(hd, precision):
    import torch
    vols = volumes(20)
    imgs = [torch.from_numpy(v).to(hd.device) for v in vols]
    hd.batch_plan(imgs, params(hd, SHAPES[0], precision))
    try:
        hd.batch_run()
        got = [hd.batch_download(t) for t in range(len(vols))]
        for t, v in enumerate(vols):
            n, over = hd.batch_tile_info(t)
            assert not over and n == len(got[t]["label"]) - 1
            assert_same(got[t], per_stack(hd, v, precision), f"stack {t} {SHAPES[t]}")
        # per-stack calls on the same context leave the plan intact: replay, same result
        hd.batch_run()
        for t in range(len(vols)):
            assert_same(hd.batch_download(t), got[t], f"replay stack {t}")
    finally:
        hd.batch_release()


def test_batch_stack_matches_oracle(hd):
    import torch
    vols = volumes(40)
    imgs = [torch.from_numpy(v).to(hd.device) for v in vols]
    hd.batch_plan(imgs, params(hd, SHAPES[0]))
    try:
        hd.batch_run()
        t = 3
        g = hd.batch_download(t)
        o = oracle.substack(vols[t], P)["regions"]
        assert len(g["label"]) == len(o["label"])
        for k in ("label", "count", "size16", "isum", "imax", "centroid", "norm"):
            np.testing.assert_array_equal(g[k][1:], o[k][1:], err_msg=k)
    finally:
        hd.batch_release()


def test_batch_replay_follows_new_contents(hd):
    import torch
    vols = volumes(60)
    imgs = [torch.from_numpy(v).to(hd.device) for v in vols]
    hd.batch_plan(imgs, params(hd, SHAPES[0]))
    try:
        hd.batch_run()
        first = [hd.batch_download(t) for t in range(len(vols))]
        new = synth.make_volume("c1", shape=SHAPES[1], n_blobs=30, seed=999)
        imgs[1].copy_(torch.from_numpy(new).to(hd.device))
        hd.batch_run()
        assert_same(hd.batch_download(1), per_stack(hd, new, "fp32"), "replaced stack 1")
        for t in (0, 2, 3, 4):
            assert_same(hd.batch_download(t), first[t], f"untouched stack {t}")
    finally:
        hd.batch_release()


def test_batch_reports_region_overflow(hd):
    import torch
    from hdog import HDoGError, _capi as C
    vols = volumes(80)
    imgs = [torch.from_numpy(v).to(hd.device) for v in vols]
    ref = [per_stack(hd, v, "fp32") for v in vols]
    counts = [len(r["label"]) - 1 for r in ref]
    # per-stack accumulator too small for the largest stacks
    cap = sorted(counts)[1]
    hd.batch_plan(imgs, params(hd, SHAPES[0]), max_regions_per_stack=cap)
    try:
        with pytest.raises(HDoGError) as e:
            hd.batch_run()
        assert e.value.status == C.HDOG_ERR_CAPACITY
        for t in range(len(vols)):
            n, over = hd.batch_tile_info(t)
            assert n == counts[t] and over == (counts[t] > cap)
            if not over:
                assert_same(hd.batch_download(t), ref[t], f"stack {t} within capacity")
    finally:
        hd.batch_release()
    # arena too small: the stacks that reserved their range before it filled are complete,
    # the others report overflow (the graph's branches reserve in the order they get there)
    arena = counts[0] + counts[1]
    hd.batch_plan(imgs, params(hd, SHAPES[0]), arena_capacity=arena)
    try:
        with pytest.raises(HDoGError):
            hd.batch_run()
        fitted, overflowed = 0, 0
        for t in range(len(vols)):
            n, over = hd.batch_tile_info(t)
            assert n == counts[t]
            if over:
                overflowed += 1
            else:
                fitted += n
                assert_same(hd.batch_download(t), ref[t], f"stack {t} in arena")
        assert fitted <= arena and overflowed >= 1
    finally:
        hd.batch_release()


def test_batch_rejects_bad_plans(hd):
    import torch
    from hdog import HDoGError
    img = torch.zeros((4, 8, 8), dtype=torch.uint16, device=hd.device)
    with pytest.raises(HDoGError):
        hd.batch_plan([], params(hd, (4, 8, 8)))
    with pytest.raises(HDoGError):
        hd.batch_run()
    bad = torch.zeros((0, 8, 8), dtype=torch.uint16, device=hd.device)
    with pytest.raises(HDoGError):
        hd.batch_plan([img, bad], params(hd, (4, 8, 8)))


def c3_shapes(n=40, seed=3):
    """Stack shapes (z, y, x) of the batched workload: z ~ U{64..640}, y and x each from
    {512, 1024, 1536, 2160} (the bench's generator)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rng.integers(64, 641, n)
    y = rng.choice([512, 1024, 1536, 2160], n)
    x = rng.choice([512, 1024, 1536, 2160], n)
    return [(int(a), int(b), int(c)) for a, b, c in zip(z, y, x)]


def test_autotune_cache_and_unchanged_results(hd, tmp_path):
    # the tuner's choice is read back from its cache file by a fresh context without a
    # stack to time, and tuned launch shapes give the untuned regions bit for bit
    import torch
    from hdog import HDoG
    shape = (24, 700, 260)
    vol = synth.make_volume("c1", shape=shape, n_blobs=120, seed=51)
    img = torch.from_numpy(vol).to(hd.device)
    p = params(hd, shape)
    ref = per_stack(hd, vol, "fp32")
    cache = str(tmp_path / "tune.txt")
    ns, gs = hd.autotune(img, p, cache)
    assert 1 <= ns <= 4 and 1 <= gs <= 4
    line = open(cache).read().split()
    assert [int(v) for v in line] == [shape[2], shape[1], shape[0], 0, ns, gs]
    fresh = HDoG(0)
    try:
        assert fresh.autotune(None, p, cache) == (ns, gs)
        with pytest.raises(Exception):
            fresh.autotune(None, params(hd, (24, 700, 261)), cache)  # not cached, no stack
    finally:
        fresh.close()
    for want in range(1, 5):  # every launch shape the tuner may pick gives the same regions
        with open(cache, "w") as f:
            f.write(f"{shape[2]} {shape[1]} {shape[0]} 0 {want} {want}\n")
        assert hd.autotune(None, p, cache) == (want, want)
        assert_same(per_stack(hd, vol, "fp32"), ref, f"segments {want}")


def test_batch_of_c3_stacks_tuned_matches_per_stack_and_oracle(hd, tmp_path):
    # the three smallest stacks of the batched workload itself, rendered by its generator,
    # through the autotuned graph, the per-stack path and the oracle
    import torch
    cfg = synth.WORKLOADS["c1"]
    shapes = c3_shapes()
    pick = sorted(range(len(shapes)), key=lambda t: np.prod(shapes[t]))[:3]
    imgs, vols = [], []
    for t in pick:
        sh = shapes[t]
        nb = int(sh[0] * sh[1] * sh[2] / 2e4)
        blobs = synth.blob_table(sh, nb, 1000 + t, cfg["sigma"], cfg["zfactor"], cfg["amp"])
        im = hd.synth_volume(sh, blobs, 1000 + t, cfg["mean0"], cfg["mean1"], cfg["sd"])
        imgs.append(im)
        vols.append(im.cpu().numpy())
    print("C3 stacks", [shapes[t] for t in pick], flush=True)
    hd.batch_plan(imgs, params(hd, shapes[pick[0]]), autotune_cache=str(tmp_path / "c3.txt"))
    try:
        hd.batch_run()
        got = [hd.batch_download(t) for t in range(len(imgs))]
    finally:
        hd.batch_release()
    for k, v in enumerate(vols):
        assert_same(got[k], per_stack(hd, v, "fp32"), f"C3 stack {v.shape}")
        o = oracle.substack(v, P)["regions"]
        assert len(got[k]["label"]) == len(o["label"])
        for key in ("label", "count", "size16", "isum", "imax"):
            np.testing.assert_array_equal(got[k][key][1:], o[key][1:], err_msg=key)
        np.testing.assert_array_equal(got[k]["centroid"][1:], o["centroid"][1:])
    del imgs
    torch.cuda.empty_cache()
