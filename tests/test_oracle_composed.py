"""The oracle's composed two-phase step (single_step,
declarations_sph.f90:1131-1435) pinned by symmetries of the equations,
beyond the per-routine checks:

- mirror equivariance, bit for bit: x -> -x is exact in floating point, so
  the pair sets, the pair order and every |term| are unchanged; x components
  of vectors and the shear components of tensors change sign, everything
  else is identical.  A transposed operand (an x gradient where a y gradient
  belongs, a swapped tensor component) breaks it;
- Galilean invariance: a uniform velocity added to every particle leaves all
  rates unchanged up to rounding (every velocity enters through
  differences)."""
import numpy as np
import pytest

import oracle
from helpers import normwise
from sources_amd import scenarios

RATES = ("drho", "dvx", "txx", "txy", "tyy", "vcc", "wxy", "dsxx", "dsxy", "dsyy", "dp", "av", "dvof", "vof2")
ODD = ("txy", "wxy", "dsxy")   # shear components: change sign under x -> -x


MAKERS = {
    "two_phase": lambda: scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                                    jitter=0.05, seed=7),
    "water_only": lambda: scenarios.dam_break(dx=0.01, jitter=0.05, seed=5),
    "soil_only": lambda: scenarios.soil_collapse(jitter=0.05, seed=4),
}


def _phases(sc):
    return [ph for ph in (sc.water, sc.soil) if ph is not None]


def _case(seed=3, kind="two_phase"):
    sc = MAKERS[kind]()
    rng = np.random.default_rng(seed)
    for ph in _phases(sc):
        ph.vx[:ph.n] = rng.normal(scale=0.05, size=(ph.n, 2))
    if sc.soil is not None:
        sc.soil.sxy[:sc.soil.n] = rng.normal(scale=50.0, size=sc.soil.n)
    return sc


def _mirror(sc, flip_velocity=True):
    b = sc.copy()
    for ph in _phases(b):
        ph.x[:, 0] = -ph.x[:, 0]
        if flip_velocity:
            ph.vx[:, 0] = -ph.vx[:, 0]
        ph.sxy[:] = -ph.sxy
    b.params.x_mingeom, b.params.x_maxgeom = -sc.params.x_maxgeom, -sc.params.x_mingeom
    return b


def _mirror_mismatch(a, b):
    bad = []
    for name in ("water", "soil"):
        pa, pb = getattr(a, name), getattr(b, name)
        if pa is None:
            continue
        n = pa.n
        for f in RATES:
            A = getattr(pa, f)[:n]
            B = getattr(pb, f)[:n].copy()
            if B.ndim == 2:
                B[:, 0] = -B[:, 0]
            elif f in ODD:
                B = -B
            if not np.array_equal(A, B):
                bad.append((name, f))
    return bad


@pytest.mark.parametrize("kind", list(MAKERS))
def test_single_step_is_mirror_equivariant_bit_for_bit(kind):
    # two-phase single_step, single_step_for_water and single_step_for_soil
    a = _case(kind=kind)
    b = _mirror(a)
    oracle.single_step(a.params, a.water, a.soil, 3)
    oracle.single_step(b.params, b.water, b.soil, 3)
    assert _mirror_mismatch(a, b) == []
    # the check is not vacuous: the rates are far from zero
    for ph in _phases(a):
        assert np.max(np.abs(ph.dvx[:ph.n])) > 1.0
    if a.soil is not None:
        assert np.max(np.abs(a.soil.dsxy[:a.soil.n])) > 1e5


def test_mirror_check_fails_for_an_unmirrored_input():
    # mirrored positions with the velocities left as they were: not a mirror
    # image, and the check says so
    a = _case()
    b = _mirror(a, flip_velocity=False)
    oracle.single_step(a.params, a.water, a.soil, 3)
    oracle.single_step(b.params, b.water, b.soil, 3)
    assert len(_mirror_mismatch(a, b)) > 5


def test_single_step_is_galilean_invariant():
    a = _case()
    b = a.copy()
    for ph in (b.water, b.soil):
        ph.vx[:ph.n] += [0.3, -0.2]
    oracle.single_step(a.params, a.water, a.soil, 3)
    oracle.single_step(b.params, b.water, b.soil, 3)
    for name in ("water", "soil"):
        pa, pb = getattr(a, name), getattr(b, name)
        for f in RATES:
            e = normwise(getattr(pb, f)[:pb.n], getattr(pa, f)[:pa.n])
            assert e <= 1e-12, (name, f, e)
