"""Precision forms (FP16 / BF16 storage of the normalized volume, DoG and Hessian elements,
FP32 accumulation) against the FP64 restatement of the oracle, on a crop of the headline
workload (its noise and blob model and parameters): per-stage error, candidate-set
Jaccard and region-feature error (hdog.precision), printed as one JSON line per form.

Bars:
- FP32 agrees with FP64 to FP32 rounding (stage errors, Jaccard > 0.9999);
- a stored 2-byte normalized volume is exactly the FP32 one rounded to nearest;
- the FP64 re-evaluation band of the leading minors: the quoted 1e-3 band decides the
  same candidates as the default 1e-5 band;
- candidate-set agreement with FP64 at the 99.9 % goal.  Forms measured below it are
  expected failures with the measured value in the reason (strict: the mark has to go
  when a form reaches the goal)."""
import json

import numpy as np
import pytest

import oracle
from hdog import precision, synth

pytestmark = pytest.mark.gpu

P = dict(min_intensity=0, rxy=10, rz=10, sigma_xy1=2.0, sigma_z1=1.6, sigma_xy2=3.2,
         sigma_z2=2.56, truncate=4.0, gamma=1.0)
SHAPE = (48, 360, 400)
GOAL = 0.999
# measured candidate-set Jaccard against FP64 on this crop, below the goal
SHORT = {"fp16": 0.9944, "bf16": 0.9576, "fp16_dog32": 0.9985, "bf16_dog32": 0.9892}


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
def case():
    n_blobs = int(round(150000 * np.prod(SHAPE) / (512 * 2160 * 2560)))
    vol = synth.make_volume("c1", shape=SHAPE, n_blobs=n_blobs, seed=31)
    return vol, oracle.substack_f64(vol, P)


def run(hd, vol, form, band=0.0):
    import torch
    params = hd.make_params(vol.shape, P["rxy"], P["rz"], P["min_intensity"], P["sigma_xy1"],
                            P["sigma_z1"], P["sigma_xy2"], P["sigma_z2"], P["truncate"],
                            P["gamma"], False, form, minor_band=band)
    bufs = hd.alloc_intermediates(vol.shape, form)
    hd.run_device(torch.from_numpy(vol).to(hd.device), params, bufs)
    return bufs, hd.download_regions()


def report(hd, case, form):
    vol, ref = case
    bufs, g = run(hd, vol, form)
    err = precision.stage_errors(bufs["norm"], bufs["dog"], bufs["hess"], ref)
    jac = precision.jaccard(bufs["pd"].cpu(), ref["pd"])
    reg = precision.region_errors(
        dict(label=g["label"], count=g["count"], structureness=g["structureness"],
             blobness=g["blobness"], norm=g["norm"]), ref["regions"])
    print(json.dumps(dict(form=form, shape=SHAPE, jaccard_vs_fp64=jac, stages=err, regions=reg)),
          flush=True)
    return bufs, err, jac, reg


def test_fp32_is_fp64_to_fp32_rounding(hd, case):
    _, err, jac, reg = report(hd, case, "fp32")
    assert err["norm"]["max_rel"] < 4e-6 and err["dog"]["max_rel"] < 2e-5
    assert max(err[k]["max_rel"] for k in precision.STAGES[2:]) < 1e-4
    assert jac > 0.9999
    # a region whose candidates sit at the FP32 / FP64 decision boundary can gain or lose
    # a voxel (measured: 99.84 % of the FP64 regions reproduced with label and size)
    assert reg["reproduced"] > 0.997 and reg["structureness"]["median_rel"] < 1e-5


@pytest.mark.parametrize("form,dt", [("fp16", "float16"), ("bf16", "bfloat16"),
                                     ("fp16_dog32", "float16"), ("bf16_dog32", "bfloat16")])
def test_stored_norm_is_the_fp32_one_rounded(hd, case, form, dt):
    import torch
    vol, _ = case
    b32, _ = run(hd, vol, "fp32")
    b, _ = run(hd, vol, form)
    assert torch.equal(b["norm"], b32["norm"].to(getattr(torch, dt)))


@pytest.mark.parametrize("form", ["fp16", "bf16"])
def test_quoted_band_decides_the_same_candidates(hd, case, form):
    import torch
    vol, _ = case
    b5, _ = run(hd, vol, form, 1e-5)
    b3, _ = run(hd, vol, form, 1e-3)
    assert torch.equal(b5["pd"], b3["pd"])


@pytest.mark.parametrize("form", [
    pytest.param(f, marks=pytest.mark.xfail(
        strict=True, reason=f"measured candidate Jaccard vs FP64 {SHORT[f]} < {GOAL}"))
    if f in SHORT else f for f in ("fp16", "bf16", "fp16_dog32", "bf16_dog32")])
def test_candidate_agreement_with_fp64_reaches_the_goal(hd, case, form):
    _, err, jac, reg = report(hd, case, form)
    assert jac >= GOAL, jac
