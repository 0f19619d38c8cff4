"""Comparison helpers shared by the tests.

Error statement of the GPU path against the FP64 oracle, per F/T vector x
(a sum over faces of per-face contributions x_i):

    |x - x64|  <=  1e-6 |x64|  +  2e-7 S,      S = sum_i |x_i|

The first term is the 1e-6 relative bound; the second is what FP32 rounding
of each contribution allows when the contributions cancel, e.g. the torque
of a symmetric hull: a face term is a composition of several FP32 operations
(centroid, relative velocity, normal velocity, product, cross product), whose
rounding errors add to a few units of the unit roundoff u = 6e-8; 2e-7 is
3.4 u.  For sums that do not cancel (S ~ |x|) the bound is a 1.2e-6
relative bound.  For the quadratic drag S is the size
of each face term's operands, |vc|^2 A (|r| |vc|^2 A for its moment): the
normal velocity vc.n is a difference of products of size |vc|, so a face the
flow passes nearly tangentially carries an FP32 rounding error of order
u |vc|^2 A however small its own term.  For the pressure buoyancy with waves
S is A (|d| + e) per wet face, e the wet polygon's extent from its centroid:
a clipped polygon's points are FP32, so its centroid is known to u e, and
near the waterline its depth d (a small difference of surface-scale values)
is of that order (the oracle's orc_result.S).  After CorrectHydrodynamicForces
(x_out = k(x) x with k = kappa sum_i |d_i| C_i) the bound is carried through
the correction: |d x_out| <= (k + 2 kappa |C|) |d x_raw|.
"""
import numpy as np

VEC = {"Fb": slice(0, 3), "Tb": slice(3, 6), "Fdq": slice(6, 9), "Tdq": slice(9, 12),
       "Fdf": slice(12, 15), "Tdf": slice(15, 18)}
NAMES = ("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf")
RTOL = 1e-6
STOL = 2e-7


def norm_rel(x, ref, floor):
    """|x - ref|_2 / max(|ref|_2, floor): the norm-relative error of one F/T vector."""
    x, ref = np.asarray(x, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), floor))


def compare_out(gpu, ref, scale_F, scale_T, rtol=1e-6, fields=VEC, mask=None):
    """Worst norm-relative error per F/T vector over a batch of sfhydro
    outputs (n, 24), with a floor of 1e-6 of the batch's force/torque scale."""
    gpu, ref = np.atleast_2d(gpu), np.atleast_2d(ref)
    worst = {}
    for name, sl in fields.items():
        floor = 1e-6 * (scale_T if name.startswith("T") else scale_F)
        errs = [norm_rel(gpu[i, sl], ref[i, sl], floor) for i in range(gpu.shape[0])
                if mask is None or mask[i]]
        worst[name] = max(errs) if errs else 0.0
    return worst


def bound_ratio(gpu, ref, extras, names=NAMES, Cd=None, Cf=None, rho=1000.0):
    """Worst |x - x64| / bound per vector over a batch (ratios <= 1 pass) and
    the worst plain norm-relative error |x - x64| / |x64| (for reporting).
    gpu, ref: (n, 24); extras: (n, 22) from the oracle (S[6], raw[12], k[4]).
    Cd, Cf: (n, 3) coefficients when the outputs went through Correct."""
    gpu, ref, extras = np.atleast_2d(gpu), np.atleast_2d(ref), np.atleast_2d(extras)
    worst, plain = {}, {}
    for j, name in enumerate(NAMES):
        if name not in names:
            continue
        sl = VEC[name]
        r_worst, p_worst = 0.0, 0.0
        for i in range(gpu.shape[0]):
            if not (int(ref[i, 20]) >> j) & 1:
                continue
            d = np.linalg.norm(gpu[i, sl] - ref[i, sl])
            S = extras[i, j]
            if j >= 2:
                raw = extras[i, 6 + 3 * (j - 2):9 + 3 * (j - 2)]
                k = extras[i, 18 + (j - 2)]
                bound = RTOL * np.linalg.norm(raw) + STOL * S
                if Cd is not None and k != 1.0:
                    C = (Cd if j < 4 else Cf)[i]
                    kappa = (0.5 if j < 4 else 1.0) * rho
                    bound *= k + 2.0 * kappa * np.linalg.norm(C)
            else:
                bound = RTOL * np.linalg.norm(ref[i, sl]) + STOL * S
            bound = max(bound, 1e-300)
            r_worst = max(r_worst, d / bound)
            nr = np.linalg.norm(ref[i, sl])
            if nr > 0:
                p_worst = max(p_worst, d / nr)
        worst[name] = r_worst
        plain[name] = p_worst
    return worst, plain


TOTALS = {"F": (slice(0, 3), slice(6, 9), slice(12, 15)), "T": (slice(3, 6), slice(9, 12), slice(15, 18))}


def component_errors(gpu, ref, names=NAMES, floor_frac=1e-3):
    """Per-component error table of the error statement, worst over a batch.

    For every F/T vector x (the six outputs, and the totals F = Fb + Fdq + Fdf,
    T = Tb + Tdq + Tdf over the slots the reference wrote) and each component
    k: |dx_k| / |x64_k| where |x64_k| >= floor_frac |x64|_2; a component below
    that floor gets |dx_k| / |x64|_2 instead (marked "abs").  Returns
    {name: {"rel": [x, y, z], "abs": [x, y, z], "norm_rel": worst |dx|/|x64|}}
    with None where no body had a component of that kind."""
    gpu, ref = np.atleast_2d(gpu), np.atleast_2d(ref)
    table = {}

    def vecs(i, name):
        if name in VEC:
            j = NAMES.index(name)
            if not (int(ref[i, 20]) >> j) & 1:
                return None
            return gpu[i, VEC[name]], ref[i, VEC[name]]
        g = np.zeros(3)
        r = np.zeros(3)
        for j, sl in zip((0, 2, 4) if name == "F" else (1, 3, 5), TOTALS[name]):
            if (int(ref[i, 20]) >> j) & 1:
                g = g + gpu[i, sl]
                r = r + ref[i, sl]
        return g, r

    for name in list(names) + ["F", "T"]:
        rel = [None, None, None]
        ab = [None, None, None]
        nr = None
        for i in range(gpu.shape[0]):
            v = vecs(i, name)
            if v is None:
                continue
            g, r = v
            n = float(np.linalg.norm(r))
            if n == 0.0:
                continue
            d = np.abs(g - r)
            nr = max(nr or 0.0, float(np.linalg.norm(g - r)) / n)
            for k in range(3):
                if abs(r[k]) >= floor_frac * n:
                    rel[k] = max(rel[k] or 0.0, float(d[k] / abs(r[k])))
                else:
                    ab[k] = max(ab[k] or 0.0, float(d[k] / n))
        table[name] = {"rel": [None if x is None else float(f"{x:.2e}") for x in rel],
                       "abs": [None if x is None else float(f"{x:.2e}") for x in ab],
                       "norm_rel": None if nr is None else float(f"{nr:.2e}")}
    return table


def format_component_table(table) -> str:
    """The component table as text lines (for test logs)."""
    lines = ["vector  norm_rel   |dx|/|x_k| (x, y, z)            |dx_k|/|x| below the 1e-3 floor"]
    for name, t in table.items():
        f = lambda v: "    -    " if v is None else f"{v:9.2e}"
        lines.append(f"{name:6s} {f(t['norm_rel'])}  {' '.join(f(v) for v in t['rel'])}   "
                     f"{' '.join(f(v) for v in t['abs'])}")
    return "\n".join(lines)
