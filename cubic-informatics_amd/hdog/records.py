"""Readers of the candidate records and the downstream transforms that consume them.

HDoG3D_main appends, per substack, packed 28-byte records {x, y, z, structureness,
blobness, intensity (max or sum of the normalized intensity), size (u16), padding}
(src/3D/utils.h:188-250).  The reference's downstream scripts read them as `dt_local`
(script/HalfBrainCells.py:11-15), place each stack's cells in the half-brain frame
(script/HalfBrainCells.py:22-58), scale and flip them into the merged frame
(script/MergeBrain.py:546-580) and classify them on log-features
(script/HDoG_classifier.py:158-176).  This module restates those transforms on NumPy
arrays so that results of this build can be checked through them.
"""
import numpy as np

# script/HalfBrainCells.py:11-15
DT_LOCAL = np.dtype([
    ("local_x", "f4"), ("local_y", "f4"), ("local_z", "f4"),
    ("structureness", "f4"), ("blobness", "f4"), ("intensity", "f4"),
    ("size", "u2"), ("padding", "u2"),
])
# script/HalfBrainCells.py:16-20
DT_GLOBAL = np.dtype([
    ("merged_x", "f4"), ("merged_y", "f4"), ("merged_z", "f4"),
    ("i_x", "u1"), ("i_y", "u1"), ("FWRV", "u1"),
])
# script/MergeBrain.py:32-35: scaled coordinates and the z-bound flag
DT_SCALEMERGED = np.dtype([
    ("scaled_x", "f4"), ("scaled_y", "f4"), ("scaled_z", "f4"), ("is_valid", "bool"),
])
assert DT_LOCAL.itemsize == 28


def read_records(path):
    """All records of one result file (empty when the file is absent)."""
    import os
    if not os.path.exists(path):
        return np.zeros(0, dtype=DT_LOCAL)
    return np.fromfile(path, dtype=DT_LOCAL)


def to_global(data_local, offset_xyz, scales, image_size, is_fw=True, i_xy=(0, 0)):
    """Cells of one stack in the half-brain frame (script/HalfBrainCells.py:44-56):
    merged = offset + local * scale, and a negative scale flips the axis across the image
    (merged -= extent * scale)."""
    width, height = image_size
    sx, sy, sz = scales
    ox, oy, oz = offset_xyz
    g = np.zeros(data_local.shape[0], dtype=DT_GLOBAL)
    g["FWRV"] = 0 if is_fw else 1
    g["i_x"], g["i_y"] = i_xy
    g["merged_x"] = ox + data_local["local_x"] * sx
    g["merged_y"] = oy + data_local["local_y"] * sy
    g["merged_z"] = oz + data_local["local_z"] * sz
    if sx < 0:
        g["merged_x"] -= width * sx
    if sy < 0:
        g["merged_y"] -= height * sy
    return g


def scale_merge(data_global, bound_z, margin_left, margin_top, offset_xyz, coeff_xyz, A, b):
    """Merged-frame coordinates of one stack's cells (script/MergeBrain.py:546-580): the z
    range test, the shift/scale into the downscaled frame (FP32), then the flip/rotation
    x' = A x + b (FP64, as NumPy's dot of the float32 array with a float64 matrix)."""
    lo, hi = (bound_z[1], bound_z[0]) if bound_z[0] > bound_z[1] else tuple(bound_z)
    out = np.zeros(data_global.shape[0], dtype=DT_SCALEMERGED)
    out["is_valid"] = np.bitwise_and(lo <= data_global["merged_z"], data_global["merged_z"] <= hi)
    c = np.zeros((data_global.shape[0], 3), dtype=np.float32)
    ox, oy, oz = offset_xyz
    cx, cy, cz = coeff_xyz
    c[:, 0] = (data_global["merged_x"] - ox) * cx - margin_left
    c[:, 1] = (data_global["merged_y"] - oy) * cy - margin_top
    c[:, 2] = (data_global["merged_z"] - oz) * cz
    f = np.asarray(A).dot(c.T).T + np.asarray(b)
    out["scaled_x"], out["scaled_y"], out["scaled_z"] = f[:, 0], f[:, 1], f[:, 2]
    return out


def get_X_3d(data, bias=np.array([0, 0, 0]), scale=np.array([1.0, 1.0, 1.0]),
             bias_before_log=np.array([0.0, 0.0, 0.0]), is_ave=False):
    """Classifier features (script/HDoG_classifier.py:158-176): log10 of (intensity,
    structureness, blobness), intensity divided by size in average mode, then bias and
    scale; NaN and infinities mapped by np.nan_to_num."""
    inten = data["intensity"] / data["size"] if is_ave else data["intensity"]
    x = np.array([inten, data["structureness"], data["blobness"]]).T
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.nan_to_num((np.log10(x + bias_before_log) + bias) * scale)
