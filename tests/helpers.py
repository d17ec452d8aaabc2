"""Comparison helpers shared by the parity tests."""
import numpy as np

# Fields compared after a step, per phase material (1 water, 2 soil).
STATE_FIELDS = {
    1: ("x", "vx", "rho", "p", "c", "vof", "drho", "dvx", "dvof", "av", "txx", "txy", "tyy", "vof2"),
    2: ("x", "vx", "rho", "p", "c", "vof", "sxx", "sxy", "syy", "epsilon_p", "drho", "dvx",
        "dsxx", "dsxy", "dsyy", "dp", "av", "txx", "txy", "tyy", "vcc", "wxy"),
}


def normwise(a, b):
    """max|a-b| / max|b|: the norm-wise relative error of field a against b."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    if scale == 0.0:
        return d
    return d / scale


def field_errors(gpu_phase, ref_phase, fields=None):
    n = ref_phase.n
    fields = fields or STATE_FIELDS[ref_phase.material]
    out = {}
    for f in fields:
        out[f] = normwise(getattr(gpu_phase, f)[:n], getattr(ref_phase, f)[:n])
    return out


def assert_fields(gpu_phase, ref_phase, tol, fields=None, label=""):
    errs = field_errors(gpu_phase, ref_phase, fields)
    bad = {k: v for k, v in errs.items() if not (v <= tol)}
    assert not bad, f"{label} fields over {tol:g}: {bad}  (all: {errs})"
    return errs


def canonical_pairs(pi, pj):
    a = np.stack([np.minimum(pi, pj), np.maximum(pi, pj)], 1)
    order = np.lexsort((a[:, 1], a[:, 0]))
    return a[order]
