"""Neighbour lists with a skin in the FP64 two-phase step: the lists hold
the candidates within R + skin at a rebuild, and each step filters them with
the exact pair predicate.  Checked against the oracle (which searches every
step) and against the row-scan form of the same kernels (SPHG_NLIST=0);
the displacement trigger and the past-the-skin error are forced from the
host through the skin length (SPHG_NL_SKIN) -- every kernel exits normally."""
import numpy as np
import pytest

import oracle
from helpers import assert_fields, field_errors
from sources_amd import abi, scenarios
from sources_amd.sph import Solver

pytestmark = pytest.mark.gpu


def _column():
    return scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12, jitter=0.05,
                                      seed=7)


def _slope():
    return scenarios.slope(dx=0.01, scale=0.25, jitter=0.1, seed=3)


def _run(monkeypatch, make, nsteps, **env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sc = make()
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(nsteps, sync=True)
    stats = s.nlist_stats()
    s.close()
    for k in env:
        monkeypatch.delenv(k)
    return sc, stats


@pytest.mark.parametrize("make", [_column, _slope], ids=["column", "slope"])
def test_lists_agree_with_row_scans(monkeypatch, make):
    # same kernels, candidates from the lists or from the rows of the cell
    # list rebuilt every step: the pair sets are identical, only the order of
    # the partners (and so the rounding of the sums) differs
    print(f"case={make.__name__}", flush=True)
    a, (na, _, skin) = _run(monkeypatch, make, 40, SPHG_NL_EVERY="16")
    b, (nb, _, _) = _run(monkeypatch, make, 40, SPHG_NLIST="0")
    assert na >= 3 and nb == 0 and skin > 0     # steps 1, 17 and 33 rebuild
    for ph in ("water", "soil"):
        e = field_errors(getattr(a, ph), getattr(b, ph))
        assert max(e.values()) <= 1e-11, (ph, e)
        assert np.array_equal(getattr(a, ph).fail, getattr(b, ph).fail)


def test_displacement_trigger_rebuilds(monkeypatch):
    # a skin of 1% of R and no periodic rebuild: the water particles of the
    # column pass a quarter of the skin (0.25% of R) after about 60 steps
    sc, ref = _column(), _column()
    monkeypatch.setenv("SPHG_NL_SKIN", "0.01")
    monkeypatch.setenv("SPHG_NL_EVERY", "100000")
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(200, sync=True)
    rebuilds, spells, skin = s.nlist_stats()
    s.close()
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 200)
    print("rebuilds", rebuilds, "skin", skin, flush=True)
    assert rebuilds >= 2 and spells == 0
    hmax = max(float(np.max(ref.water.hsml)), float(np.max(ref.soil.hsml)))
    assert skin == pytest.approx(0.01 * 2.0 * hmax, rel=1e-6)     # cubic spline: R = 2h
    for ph in ("water", "soil"):
        assert_fields(getattr(sc, ph), getattr(ref, ph), 1e-11, fields=("x", "vx", "rho", "p"), label=ph)


def test_past_half_the_skin_is_an_error(monkeypatch):
    # a skin of 1e-9 R: the first step after the rebuild moves particles
    # farther than half of it, and the lagged check reports it
    monkeypatch.setenv("SPHG_NL_SKIN", "1e-9")
    sc = _column()
    s = Solver(sc.params, sc.water, sc.soil)
    with pytest.raises(abi.SphgError) as e:
        s.time_integration(5, sync=True)
    assert e.value.code == abi.SPHG_ERR_MOVED_TOO_FAR
    s.close()


def test_fast_motion_turns_the_lists_off(monkeypatch):
    # summation density with the water renormalisation: the column collapses
    # (about 1% of R per step by step 40), triggered rebuilds come every few
    # steps and the lists are turned off; results stay at rounding level
    def make():
        sc = _column()
        sc.params.summation_density = 1
        sc.params.volume_fraction_renorm = 1
        return sc
    sc, ref = make(), make()
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(45, sync=True)
    rebuilds, spells, _ = s.nlist_stats()
    s.close()
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 45)
    print("rebuilds", rebuilds, "off spells", spells, flush=True)
    assert spells >= 1
    for ph in ("water", "soil"):
        assert_fields(getattr(sc, ph), getattr(ref, ph), 1e-10, label=ph)


def _stress2():
    sc = _column()
    sc.params.stress_integration = 2
    return sc


def _renorm():
    sc = _column()
    sc.params.summation_density = 1
    sc.params.volume_fraction_renorm = 1
    return sc


@pytest.mark.parametrize("make,nsteps", [(_column, 20), (_slope, 20), (_stress2, 20), (_renorm, 45)],
                         ids=["column", "slope", "stress_integration_2", "summation_renorm"])
def test_fused_final_and_pack_are_bit_identical(make, nsteps):
    # Inside one time_integration call the final update of each step also
    # packs the next step (one pass); a call's last step does not.  N calls
    # of one step take the unfused path throughout: the states agree bit
    # for bit.
    print(f"case={make.__name__} steps={nsteps}", flush=True)
    a, b = make(), make()
    sa = Solver(a.params, a.water, a.soil)
    sb = Solver(b.params, b.water, b.soil)
    sa.time_integration(nsteps, sync=True)
    for _ in range(nsteps):
        sb.time_integration(1, sync=False)
    sb.time_integration(0, sync=True)
    for ph in ("water", "soil"):
        pa, pb = getattr(a, ph), getattr(b, ph)
        n = pa.n
        for f in ("x", "vx", "rho", "p", "c", "vof", "sxx", "sxy", "syy", "drho", "dvx", "fail"):
            if hasattr(pa, f):
                assert np.array_equal(getattr(pa, f)[:n], getattr(pb, f)[:n]), (ph, f)
    sa.close()
    sb.close()
