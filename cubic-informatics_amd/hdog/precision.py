"""Agreement of a reduced-precision storage form with a reference result of the same stack
(the FP32 path, or the FP64 restatement): per-stage error, candidate-set agreement and
region-feature error.  Torch tensors on any device, or numpy arrays.

Per stage (normalized volume, DoG, each Hessian element), with a the form's values and b
the reference's:
    max_rel  = max |a - b| / max |b|       (error against the stage's range)
    mean_rel = mean |a - b| / mean |b|
Candidates: Jaccard |A & B| / |A | B| of the two masks.
Regions: a reference region is reproduced when the form has a region with the same label
(component minimum) and voxel count; over those, the largest and the median relative error
of structureness, blobness and max-normalized intensity."""
import numpy as np

STAGES = ("norm", "dog", "hxx", "hxy", "hxz", "hyy", "hyz", "hzz")


def _t(x):
    import torch
    return x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))


def stage_error(a, b):
    a, b = _t(a), _t(b).to(_t(a).device)
    a64, b64 = a.double(), b.double()
    d = (a64 - b64).abs()
    bmax, bmean = float(b64.abs().max()), float(b64.abs().mean())
    return dict(max_rel=float(d.max()) / max(bmax, 1e-300),
                mean_rel=float(d.mean()) / max(bmean, 1e-300))


def jaccard(a, b):
    a, b = _t(a).bool(), _t(b).to(_t(a).device).bool()
    inter = int((a & b).sum())
    union = int((a | b).sum())
    return inter / max(union, 1)


def stage_errors(norm, dog, hess, ref):
    """ref: dict with norm, dog, hess (6 leading) of the reference run."""
    out = {"norm": stage_error(norm, ref["norm"]), "dog": stage_error(dog, ref["dog"])}
    for k, name in enumerate(STAGES[2:]):
        out[name] = stage_error(hess[k], ref["hess"][k])
    return out


def region_errors(g, ref):
    """g: the form's regions (label, count or size, structureness, blobness, norm; region 0
    excluded or not); ref: the reference's.  Both dicts of numpy arrays."""
    def table(r):
        lab = np.asarray(r["label"]).astype(np.int64)
        keep = lab > 0
        cnt = np.asarray(r["count"]).astype(np.float64)[keep]
        st = np.asarray(r["structureness"], np.float64)[keep]
        bl = np.asarray(r["blobness"], np.float64)[keep]
        nm = np.asarray(r["norm"], np.float64)[keep]
        return lab[keep], cnt, st, bl, nm
    la, ca, sa, ba, na = table(g)
    lb, cb, sb, bb, nb = table(ref)
    common, ia, ib = np.intersect1d(la, lb, assume_unique=True, return_indices=True)
    same = ca[ia] == cb[ib]
    ia, ib = ia[same], ib[same]
    out = dict(reference_regions=int(len(lb)), regions=int(len(la)),
               reproduced=float(len(ib)) / max(len(lb), 1))
    for name, x, y in (("structureness", sa, sb), ("blobness", ba, bb), ("max_normalized", na, nb)):
        xs, ys = x[ia], y[ib]
        fin = np.isfinite(xs) & np.isfinite(ys) & (ys != 0)
        rel = np.abs(xs[fin] - ys[fin]) / np.abs(ys[fin])
        out[name] = dict(max_rel=float(rel.max()) if rel.size else 0.0,
                         median_rel=float(np.median(rel)) if rel.size else 0.0)
    return out
