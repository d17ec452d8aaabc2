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


def crop_phase(ph, idx):
    """The particles idx (sorted; real ones stay before virtual ones) of a
    phase as a phase of their own: every per-particle array is taken at idx,
    the pair arrays sized for the crop."""
    from sources_amd.state import Phase
    idx = np.asarray(idx, np.int64)
    q = Phase.__new__(Phase)
    m = ph.maxn
    for k, v in ph.__dict__.items():
        if k in ("pair_i", "pair_j", "w", "dwdx", "_view"):
            continue
        if isinstance(v, np.ndarray) and v.shape[:1] == (m,):
            setattr(q, k, np.ascontiguousarray(v[idx]))
        elif isinstance(v, np.ndarray):
            setattr(q, k, v.copy())
        else:
            setattr(q, k, v)
    q._view = None
    q.ntotal = int(np.count_nonzero(idx < ph.ntotal))
    q.nvirt = len(idx) - q.ntotal
    q.maxn = max(len(idx), 1)
    q.niac = 0
    q.set_max_interaction(64 * max(len(idx), 1))
    return q


def window_indices(ph, lo, hi):
    """Indices of the particles of ph inside the box [lo, hi) (sorted)."""
    x = ph.x[:ph.n]
    inside = (x[:, 0] >= lo[0]) & (x[:, 0] < hi[0]) & (x[:, 1] >= lo[1]) & (x[:, 1] < hi[1])
    return np.nonzero(inside)[0]
