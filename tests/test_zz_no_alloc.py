"""hdog_ccl3d_noalloc (labeling with a fixed per-GPU scratch, no allocation) against
hdog_ccl3d and the oracle, through the C-ABI: two calls in a row, the second on a smaller
volume, so that the second reuses the scratch the first left behind.  Volumes that pad to
more than the scratch holds are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def ccl_both(hd, C, mask):
    import torch
    D, H, W = mask.shape
    pd = torch.from_numpy(mask).to(hd.device)
    want = torch.empty((D, H, W), dtype=torch.int32, device=hd.device)  # uint32 labels
    got = torch.full((D, H, W), 0xFFFF, dtype=torch.int32, device=hd.device)
    s = hd._stream()
    assert C.hdog_ccl3d(hd._h, s, ctypes.c_void_p(pd.data_ptr()),
                        ctypes.c_void_p(want.data_ptr()), W, H, D) == C.HDOG_OK
    rc = C.hdog_ccl3d_noalloc(hd._h, s, ctypes.c_void_p(pd.data_ptr()),
                              ctypes.c_void_p(got.data_ptr()), W, H, D)
    assert rc == C.HDOG_OK, C.hdog_last_error(hd._h)
    return got.cpu().numpy().view(np.uint32), want.cpu().numpy().view(np.uint32)


def test_noalloc_labels_equal_the_allocating_entry_point():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hdog import HDoG
    from hdog import _capi as C
    hd = HDoG(0)
    rng = np.random.default_rng(5)
    try:
        # the largest volume the scratch holds (128 x 256 x 256 pads to exactly 2^23 voxels),
        # then a smaller, ragged one
        for shape, dens in (((128, 256, 256), 0.3), ((37, 90, 130), 0.35)):
            mask = (rng.random(shape) < dens).astype(np.uint8)
            print(f"hdog_ccl3d_noalloc {shape}", flush=True)
            got, want = ccl_both(hd, C, mask)
            np.testing.assert_array_equal(got, want)
            np.testing.assert_array_equal(got, oracle.ccl3d(mask))
        # one more tile row than the scratch holds: refused, labels untouched
        big = torch.zeros((129, 256, 256), dtype=torch.uint8, device=hd.device)
        out = torch.full(big.shape, 7, dtype=torch.int32, device=hd.device)
        rc = C.hdog_ccl3d_noalloc(hd._h, hd._stream(), ctypes.c_void_p(big.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), 256, 256, 129)
        assert rc == C.HDOG_ERR_ARG and b"scratch" in C.hdog_last_error(hd._h)
        assert bool((out == 7).all())
    finally:
        hd.close()
