"""Comparison helpers shared by the tests."""
import numpy as np

VEC = {"Fb": slice(0, 3), "Tb": slice(3, 6), "Fdq": slice(6, 9), "Tdq": slice(9, 12),
       "Fdf": slice(12, 15), "Tdf": slice(15, 18)}


def norm_rel(x, ref, floor):
    """|x - ref|_2 / max(|ref|_2, floor): the norm-relative error of one F/T vector."""
    x, ref = np.asarray(x, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), floor))


def compare_out(gpu, ref, scale_F, scale_T, rtol=1e-6, fields=VEC, mask=None):
    """Worst norm-relative error per F/T vector over a batch of sfhydro
    outputs (n, 24).  Vectors whose reference norm is below 1e-6 of the
    batch's force/torque scale are compared against that floor."""
    gpu, ref = np.atleast_2d(gpu), np.atleast_2d(ref)
    worst = {}
    for name, sl in fields.items():
        floor = 1e-6 * (scale_T if name.startswith("T") else scale_F)
        errs = [norm_rel(gpu[i, sl], ref[i, sl], floor) for i in range(gpu.shape[0])
                if mask is None or mask[i]]
        worst[name] = max(errs) if errs else 0.0
    return worst
