"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Pair lists bit-exact; FP64 fields within 1e-12
norm-wise (max|gpu-ref| / max|ref|) after one step."""
import numpy as np
import pytest

import oracle
from helpers import assert_fields, canonical_pairs, field_errors
from sources_amd import abi, scenarios
from sources_amd.sph import Solver

pytestmark = pytest.mark.gpu

FP64_TOL_1STEP = 1e-12


def _column():
    return scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12, jitter=0.05,
                                      seed=7)


def _column_varied_h(seed=5):
    """Smoothing length varied by up to 4% per particle: the kernels' general
    form (pair h = mean of the two, h read from the partner record)."""
    sc = _column()
    rng = np.random.default_rng(seed)
    for ph in (sc.water, sc.soil):
        n = ph.ntotal + ph.nvirt
        ph.hsml[:n] *= 1.0 + 0.04 * rng.uniform(-1.0, 1.0, n)
    return sc


def _column_fractional_repulsion():
    """Non-integer repulsion exponents: the kernels' general pow() form
    (integer exponents take binary powering)."""
    sc = _column()
    sc.params.p1, sc.params.p2 = 12.5, 3.5
    return sc


def _column_large_h():
    """h = 2.5 dx: about 25 particles per cell and 75 candidates per cell row,
    so the cell sort takes its path for cells over 16 entries and the partner
    loops walk rows over 64 candidates directly."""
    sc = _column()
    for ph in (sc.water, sc.soil):
        ph.hsml[:] *= 2.5 / 1.2
        ph.set_max_interaction(120 * ph.n)
    return sc


def small_cases():
    return {
        "column_large_h": _column_large_h,
        "savage_skf4": lambda: scenarios.savage_column(),
        "savage_skf1": lambda: scenarios.savage_column(skf=1),
        "column_jitter": _column,
        "column_varied_h": _column_varied_h,
        "column_frac_rep": _column_fractional_repulsion,
        "slope_small": lambda: scenarios.slope(dx=0.01, scale=0.25, jitter=0.1, seed=3),
    }


def _both(make):
    sc = make()
    return sc, sc.copy()


@pytest.mark.parametrize("case", list(small_cases()))
@pytest.mark.parametrize("phase", ["water", "soil"])
def test_direct_find_bit_exact(case, phase):
    print(f"case={case} phase={phase}", flush=True)
    sc, ref = _both(small_cases()[case])
    s = Solver(sc.params, sc.water, sc.soil)
    s.direct_find(phase)
    g = getattr(sc, phase)
    r = getattr(ref, phase)
    oracle.call_phase("direct_find", ref.params, r)
    assert g.niac == r.niac > 0
    k = r.niac
    np.testing.assert_array_equal(g.pair_i[:k], r.pair_i[:k])
    np.testing.assert_array_equal(g.pair_j[:k], r.pair_j[:k])
    np.testing.assert_array_equal(g.countiac[:r.n], r.countiac[:r.n])
    assert np.array_equal(g.w[:k], r.w[:k]), np.max(np.abs(g.w[:k] - r.w[:k]))
    assert np.array_equal(g.dwdx[:k], r.dwdx[:k]), np.max(np.abs(g.dwdx[:k] - r.dwdx[:k]))
    s.close()


@pytest.mark.parametrize("case", list(small_cases()))
def test_direct_find_2_bit_exact(case):
    print(f"case={case}", flush=True)
    sc, ref = _both(small_cases()[case])
    s = Solver(sc.params, sc.water, sc.soil)
    s.direct_find_2()
    oracle.call_cross("direct_find_2", ref.params, ref.water, ref.soil)
    k = ref.water.niac
    assert sc.water.niac == k > 0
    np.testing.assert_array_equal(sc.water.pair_i[:k], ref.water.pair_i[:k])
    np.testing.assert_array_equal(sc.water.pair_j[:k], ref.water.pair_j[:k])
    np.testing.assert_array_equal(sc.water.countiac[:ref.water.n], ref.water.countiac[:ref.water.n])
    np.testing.assert_array_equal(sc.soil.countiac[:ref.soil.n], ref.soil.countiac[:ref.soil.n])
    assert np.array_equal(sc.water.w[:k], ref.water.w[:k])
    assert np.array_equal(sc.water.dwdx[:k], ref.water.dwdx[:k])
    s.close()


def _tight_box_column(**kw):
    # Box so tight that the reference sorting grid's cells are smaller than
    # the 2h support: link_list then misses pairs (m_particles.f90:1107-1109).
    return scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                      jitter=0.05, seed=11, x_mingeom=-0.02, x_maxgeom=0.32,
                                      y_mingeom=-0.02, y_maxgeom=0.15, **kw)


@pytest.mark.parametrize("make", [scenarios.savage_column, _tight_box_column], ids=["savage", "tight_box"])
def test_link_list_same_set(make):
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    s.link_list("water")
    oracle.call_phase("link_list", ref.params, ref.water)
    k = ref.water.niac
    assert sc.water.niac == k
    a = canonical_pairs(sc.water.pair_i[:k], sc.water.pair_j[:k])
    b = canonical_pairs(ref.water.pair_i[:k], ref.water.pair_j[:k])
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sc.water.countiac[:ref.water.n], ref.water.countiac[:ref.water.n])
    s.close()


def test_link_list_reproduces_reference_misses():
    sc, ref = _both(_tight_box_column)
    oracle.call_phase("link_list", ref.params, ref.water)
    full = sc.copy()
    oracle.call_phase("direct_find", full.params, full.water)
    assert ref.water.niac < full.water.niac  # the reference's reach really is short here
    s = Solver(sc.params, sc.water, sc.soil)
    s.link_list("water")
    assert sc.water.niac == ref.water.niac
    s.close()


@pytest.mark.parametrize("case", list(small_cases()))
def test_one_step_two_phase(case):
    print(f"case={case}", flush=True)
    sc, ref = _both(small_cases()[case])
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(1, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 1)
    ew = assert_fields(sc.water, ref.water, FP64_TOL_1STEP, label=f"{case} water")
    es = assert_fields(sc.soil, ref.soil, FP64_TOL_1STEP, label=f"{case} soil")
    print("water", max(ew.values()), "soil", max(es.values()))
    np.testing.assert_array_equal(sc.soil.fail[:ref.soil.n], ref.soil.fail[:ref.soil.n])
    s.close()


@pytest.mark.parametrize("nsteps,tol", [(2, 1e-12), (10, 1e-10)])
def test_steps_two_phase(nsteps, tol):
    sc, ref = _both(_column)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(nsteps, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, nsteps)
    assert s.itimestep == nsteps
    ew = assert_fields(sc.water, ref.water, tol, label="water")
    es = assert_fields(sc.soil, ref.soil, tol, label="soil")
    print(nsteps, "water", max(ew.values()), "soil", max(es.values()))
    s.close()


def test_long_run_stays_at_rounding_level():
    # 300 FP64 steps of the jittered column: the GPU stays as close to the
    # oracle as the oracle stays to itself after a 1-ulp perturbation of its
    # input density (about 1e-13 norm-wise for velocity and pressure).
    sc, ref = _both(_column)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(300, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 300)
    for ph in ("water", "soil"):
        e = assert_fields(getattr(sc, ph), getattr(ref, ph), 1e-11, fields=("x", "vx", "rho", "p"), label=ph)
        print(ph, {k: f"{v:.1e}" for k, v in e.items()})
    s.close()


@pytest.mark.parametrize("every", ["1", "5"])
def test_state_reorder_is_transparent(monkeypatch, every):
    # The device keeps particles in cell order, permuting the state every N
    # steps; host views, pair lists and results stay in reference order.
    monkeypatch.setenv("SPHG_REORDER_EVERY", every)
    monkeypatch.setenv("SPHG_NL_EVERY", every)     # the list rebuild reorders too
    sc, ref = _both(small_cases()["slope_small"])
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(7, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 7)
    assert_fields(sc.water, ref.water, 1e-11, label="water")
    assert_fields(sc.soil, ref.soil, 1e-11, label="soil")
    s.direct_find("soil")
    oracle.call_phase("direct_find", ref.params, ref.soil)
    k = ref.soil.niac
    assert sc.soil.niac == k
    np.testing.assert_array_equal(sc.soil.pair_i[:k], ref.soil.pair_i[:k])
    np.testing.assert_array_equal(sc.soil.pair_j[:k], ref.soil.pair_j[:k])
    s.close()


def _sensitivity(make, nsteps, fields, seed=123):
    """Norm-wise change of the oracle's result when the input density is
    perturbed by one ulp (relative 2**-52, random sign): the conditioning of
    each field.  Summation-density and delta-SPH terms difference nearly equal
    densities, so a field can legitimately differ between two correct
    summation orders by far more than 1e-12 of its own magnitude."""
    a, b = make(), make()
    u = np.random.default_rng(seed).choice([-1.0, 1.0], size=b.water.rho.shape)
    b.water.rho[:] = b.water.rho * (1.0 + u * 2.0 ** -52)
    oracle.time_integration(a.params, a.water, None, 0, nsteps)
    oracle.time_integration(b.params, b.water, None, 0, nsteps)
    return field_errors(b.water, a.water, fields)


@pytest.mark.parametrize("skf", [1, 4])
def test_single_phase_dam_break(skf):
    make = lambda: scenarios.dam_break(dx=0.01, skf=skf, jitter=0.05, seed=5)  # noqa: E731
    sc, ref = _both(make)
    fields = ("x", "vx", "rho", "p", "drho", "dvx", "av", "txx", "txy", "tyy")
    s = Solver(sc.params, sc.water)
    for nsteps, done in ((1, 0), (5, 1)):
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, None, done, nsteps - done)
        sens = _sensitivity(make, nsteps, fields)
        errs = field_errors(sc.water, ref.water, fields)
        # within 1e-12 norm-wise, or within 16x the oracle's own 1-ulp sensitivity
        bad = {f: (errs[f], sens[f]) for f in fields if not (errs[f] <= max(1e-12, 16.0 * sens[f]))}
        print(nsteps, "errs", errs, "sens", sens)
        assert not bad, f"{nsteps} steps: {bad}"
    s.close()


def test_verlet_dam_break():
    """integrate_scheme = 2 (time_integration_for_water_by_verlet): one call of
    one step, then one call of 54 steps, across the 50th step's restart."""
    def make():
        sc = scenarios.dam_break(dx=0.01, skf=1, jitter=0.05, seed=5)
        sc.params.integrate_scheme = 2
        return sc
    sched = ((1, 0), (55, 1))
    sc, ref = _both(make)
    fields = ("x", "vx", "rho", "p", "drho", "dvx", "av")
    s = Solver(sc.params, sc.water)
    for nsteps, done in sched:
        print(f"verlet steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, None, done, nsteps - done)
    assert s.itimestep == 55
    # conditioning: the same schedule from a one-ulp density perturbation
    a, b = make(), make()
    u = np.random.default_rng(123).choice([-1.0, 1.0], size=b.water.rho.shape)
    b.water.rho[:] = b.water.rho * (1.0 + u * 2.0 ** -52)
    for nsteps, done in sched:
        oracle.time_integration(a.params, a.water, None, done, nsteps - done)
        oracle.time_integration(b.params, b.water, None, done, nsteps - done)
    sens = field_errors(b.water, a.water, fields)
    errs = field_errors(sc.water, ref.water, fields)
    print("errs", errs, "sens", sens)
    bad = {f: (errs[f], sens[f]) for f in fields if not (errs[f] <= max(1e-12, 16.0 * sens[f]))}
    assert not bad, bad
    s.close()


@pytest.mark.parametrize("skf,nnps", [(1, 1), (4, 1), (1, 2)])
def test_soil_only_collapse(skf, nnps):
    """Single-phase soil: time_integration_for_soil / single_step_for_soil."""
    make = lambda: scenarios.soil_collapse(dx=0.005, skf=skf, nnps=nnps, jitter=0.05, seed=9)  # noqa: E731
    sc, ref = _both(make)
    s = Solver(sc.params, None, sc.soil)
    for nsteps, done, tol in ((1, 0, FP64_TOL_1STEP), (10, 1, 1e-10)):
        print(f"soil only skf={skf} nnps={nnps} steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, None, ref.soil, done, nsteps - done)
        es = assert_fields(sc.soil, ref.soil, tol, label=f"soil only {nsteps} steps")
        print(nsteps, "soil", max(es.values()))
        _assert_fail_flags(sc.soil, ref.soil, ref.params)
    s.close()


def test_soil_only_summation_density():
    """Single-phase soil under summation_density: single_step_for_soil
    replaces rho by sum_density (declarations_sph.f90:1558-1562) and the
    leapfrog leaves it alone; 10 steps against the oracle.  The initial
    density is perturbed by +-5 % (seeded): from a uniform density the
    summation reproduces it to rounding and drho (the delta-SPH term of the
    summed density) would be rounding noise."""
    def make():
        sc = scenarios.soil_collapse(dx=0.005, skf=1, nnps=1, jitter=0.05, seed=9)
        sc.params.summation_density = 1
        rng = np.random.default_rng(13)
        sc.soil.rho[:sc.soil.n] *= 1.0 + 0.05 * rng.uniform(-1.0, 1.0, sc.soil.n)
        return sc
    sc, ref = _both(make)
    s = Solver(sc.params, None, sc.soil)
    for nsteps, done, tol in ((1, 0, FP64_TOL_1STEP), (10, 1, 1e-10)):
        print(f"soil only summation density steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, None, ref.soil, done, nsteps - done)
        es = assert_fields(sc.soil, ref.soil, tol, label=f"soil only summation {nsteps} steps")
        print(nsteps, "soil", max(es.values()))
        _assert_fail_flags(sc.soil, ref.soil, ref.params)
    s.close()


def _assert_fail_flags(g, r, P):
    """The Drucker-Prager flags agree, except where the oracle's state sits on
    the yield surface to rounding (sqrt(J2) = alpha1 I1, as a projection
    leaves it), where 1e-15 differences decide the comparison either way."""
    n = r.n
    bad = np.nonzero(g.fail[:n] != r.fail[:n])[0]
    if len(bad) == 0:
        return
    t = np.tan(P.soil_phi)
    alpha1 = t / np.sqrt(9.0 + 12.0 * t * t)
    sxx, sxy, syy, p = r.sxx[bad], r.sxy[bad], r.syy[bad], r.p[bad]
    j2 = np.sqrt((sxx * sxx + 2.0 * sxy * sxy + syy * syy + (sxx + syy) ** 2) / 2.0) + 1.0e-6
    lim = alpha1 * 3.0 * p
    print("fail flags differ at", bad.tolist(), "sqrt(J2)", j2.tolist(), "alpha1 I1", lim.tolist())
    assert np.all(np.abs(j2 - lim) <= 1e-9 * np.maximum(np.abs(j2), np.abs(lim))), (j2, lim)


@pytest.mark.parametrize("cohesion,ptop", [(0.0, -40.0), (100.0, -600.0)])
def test_two_phase_mohr_coulomb(cohesion, ptop):
    """yield_criterion = 1: the soil criterion of single_step is Mohr-Coulomb.
    phi = 10 degrees and a shear-stress perturbation make it scale the
    deviator of most soil particles; a top layer in tension (p below
    -c cot(phi)) takes the cut-off, which resets rho between the terms
    evaluated before the criterion and after it."""
    def make():
        sc = _column()
        P = sc.params
        P.yield_criterion = 1
        P.soil_cohesion = cohesion
        P.soil_phi = np.deg2rad(10.0)
        so = sc.soil
        n = so.ntotal
        so.sxy[:n] += np.random.default_rng(3).normal(0.0, 60.0, n)
        top = so.x[:n, 1] > so.x[:n, 1].max() - 0.006
        so.p[:n][top] = ptop
        return sc
    sc, ref = _both(make)
    probe = make()
    oracle.single_step(probe.params, probe.water, probe.soil, 1)
    assert (probe.soil.rho == probe.params.soil_rho0).sum() > 0   # the cut-off is exercised
    s = Solver(sc.params, sc.water, sc.soil)
    for nsteps, done, tol in ((1, 0, FP64_TOL_1STEP), (6, 1, 1e-10)):
        print(f"mohr-coulomb c={cohesion} steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, ref.soil, done, nsteps - done)
        assert_fields(sc.water, ref.water, tol, label="water")
        es = assert_fields(sc.soil, ref.soil, tol, label="soil")
        print(nsteps, "soil", max(es.values()))
    s.close()


def test_soil_only_mohr_coulomb():
    make = lambda: scenarios.soil_collapse(dx=0.005, jitter=0.05, seed=9, yield_criterion=1,  # noqa: E731
                                           soil_cohesion=50.0, soil_phi=np.deg2rad(10.0))
    sc, ref = _both(make)
    s = Solver(sc.params, None, sc.soil)
    s.time_integration(5, sync=True)
    oracle.time_integration(ref.params, None, ref.soil, 0, 5)
    es = assert_fields(sc.soil, ref.soil, 1e-10, label="soil only, Mohr-Coulomb")
    print("soil", max(es.values()))
    s.close()


def test_plastic_flow_rule_1():
    """plasticity = 1: plastic_flow_rule at every point, two-phase and soil only."""
    def make2():
        sc = _column()
        sc.params.plasticity = 1
        return sc
    sc, ref = _both(make2)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(4, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 4)
    assert_fields(sc.water, ref.water, 1e-11, label="water")
    es = assert_fields(sc.soil, ref.soil, 1e-11, label="soil")
    print("two-phase, rule 1:", max(es.values()))
    s.close()
    sc, ref = _both(lambda: scenarios.soil_collapse(dx=0.005, jitter=0.05, seed=9, plasticity=1))
    s = Solver(sc.params, None, sc.soil)
    s.time_integration(4, sync=True)
    oracle.time_integration(ref.params, None, ref.soil, 0, 4)
    es = assert_fields(sc.soil, ref.soil, 1e-11, label="soil only, rule 1")
    print("soil only, rule 1:", max(es.values()))
    s.close()


@pytest.mark.parametrize("si,plasticity", [(2, 3), (1, 2), (2, 2)])
def test_stress_integration_and_return_mapping(si, plasticity):
    """stress_integration = 2 (leapfrog stress levels) and plasticity = 2
    (plastic_flow_rule2 + plastic_or_not + return_mapping after the update),
    two-phase; and the same switches on the soil-only path (whose integrator
    has no return mapping)."""
    def make():
        sc = _column()
        sc.params.stress_integration = si
        sc.params.plasticity = plasticity
        return sc
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    for nsteps, done, tol in ((1, 0, FP64_TOL_1STEP), (5, 1, 1e-10)):
        print(f"si={si} plasticity={plasticity} steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, ref.soil, done, nsteps - done)
        assert_fields(sc.water, ref.water, tol, label="water")
        es = assert_fields(sc.soil, ref.soil, tol, label="soil")
        es2 = field_errors(sc.soil, ref.soil, ("sxx_min", "p_min", "dsxx2", "dp2"))
        assert max(es2.values()) <= tol, es2
        print(nsteps, "soil", max(es.values()), es2)
    s.close()
    sc, ref = _both(lambda: scenarios.soil_collapse(dx=0.005, jitter=0.05, seed=9, stress_integration=si,
                                                    plasticity=plasticity))
    s = Solver(sc.params, None, sc.soil)
    s.time_integration(4, sync=True)
    oracle.time_integration(ref.params, None, ref.soil, 0, 4)
    es = assert_fields(sc.soil, ref.soil, 1e-11, label="soil only")
    print("soil only", max(es.values()))
    s.close()


def test_summation_density_switch():
    """summation_density = 1: two-phase single_step keeps con_density and the
    leapfrog leaves rho alone; the single-phase water integrator likewise
    (its right-hand side always uses summation density)."""
    def make():
        sc = _column()
        sc.params.summation_density = 1
        return sc
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(4, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 4)
    assert_fields(sc.water, ref.water, 1e-11, label="water")
    assert_fields(sc.soil, ref.soil, 1e-11, label="soil")
    s.close()
    make1 = lambda: scenarios.dam_break(dx=0.01, skf=1, jitter=0.05, seed=5, summation_density=1)  # noqa: E731
    sc, ref = _both(make1)
    s = Solver(sc.params, sc.water)
    s.time_integration(5, sync=True)
    oracle.time_integration(ref.params, ref.water, None, 0, 5)
    fields = ("x", "vx", "rho", "p", "dvx", "av")
    errs = field_errors(sc.water, ref.water, fields)
    sens = _sensitivity(make1, 5, fields)
    bad = {f: (errs[f], sens[f]) for f in fields if not (errs[f] <= max(1e-12, 16.0 * sens[f]))}
    assert not bad, bad
    s.close()


def test_summation_density_with_vof_renormalisation_over_45_steps():
    """summation_density = 1 with volume_fraction_renorm = 1: the leapfrog
    leaves rho alone, so the water renormalisation of step 40
    (declarations_sph.f90:1414-1420) persists; compared before, at and after
    it."""
    def make():
        sc = _column()
        sc.params.summation_density = 1
        sc.params.volume_fraction_renorm = 1
        return sc
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    done = 0
    for upto in (39, 40, 45):
        print(f"steps {done}->{upto}", flush=True)
        s.time_integration(upto - done, sync=True)
        oracle.time_integration(ref.params, ref.water, ref.soil, done, upto - done)
        done = upto
        ew = assert_fields(sc.water, ref.water, 1e-10, label=f"water after {upto}")
        es = assert_fields(sc.soil, ref.soil, 1e-10, label=f"soil after {upto}")
        print(upto, "water", max(ew.values()), "soil", max(es.values()))
    n = ref.water.n
    assert np.max(np.abs(ref.water.vof2[:n] - ref.water.vof[:n])) > 1e-3   # the renormalisation acted
    s.close()


def test_single_step_renormalises_on_the_40th_step():
    """The bare right-hand side at itimestep 40: single_step itself rescales
    water rho by vof2 / vof and sets vof = vof2."""
    sc, ref = _both(_column)
    s = Solver(sc.params, sc.water, sc.soil)
    s.itimestep = 40
    s.single_step(sync=True)
    oracle.single_step(ref.params, ref.water, ref.soil, 40)
    assert_fields(sc.water, ref.water, FP64_TOL_1STEP, label="water")
    assert_fields(sc.soil, ref.soil, FP64_TOL_1STEP, label="soil")
    n = ref.water.n
    assert np.array_equal(ref.water.vof[:n], ref.water.vof2[:n])
    s.close()


@pytest.mark.parametrize("skf", [2, 3])
def test_gaussian_and_quintic_kernels(skf):
    """skf 2 (Gaussian, m_particles.f90:1524-1538) and 3 (quintic,
    :1539-1574) have a 3h support (get_scale_k, :916-934): wider cells and
    candidate rows.  Pair lists bit-exact, one step within 1e-12.  The
    Gaussian's W and grad W hold exp(-q^2): the device exp and the host libm
    exp are both faithfully rounded but not always to the same double, so
    for skf 2 W and grad W agree to a few ulp (held to 1e-14 relative to the
    largest value) instead of bit for bit; the pair lists stay bit-exact."""
    def same_values(a, b):
        if skf == 2:
            np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14 * float(np.max(np.abs(b))))
        else:
            np.testing.assert_array_equal(a, b)

    def make():
        sc = _column()
        sc.params.skf = skf
        for ph in (sc.water, sc.soil):
            ph.set_max_interaction(40 * ph.n)
        return sc
    for phase in ("water", "soil"):
        sc, ref = _both(make)
        s = Solver(sc.params, sc.water, sc.soil)
        s.direct_find(phase)
        g, r = getattr(sc, phase), getattr(ref, phase)
        oracle.call_phase("direct_find", ref.params, r)
        k = r.niac
        assert g.niac == k > 0
        np.testing.assert_array_equal(g.pair_i[:k], r.pair_i[:k])
        np.testing.assert_array_equal(g.pair_j[:k], r.pair_j[:k])
        same_values(g.w[:k], r.w[:k])
        same_values(g.dwdx[:k], r.dwdx[:k])
        s.close()
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    s.direct_find_2()
    oracle.call_cross("direct_find_2", ref.params, ref.water, ref.soil)
    k = ref.water.niac
    assert sc.water.niac == k > 0
    np.testing.assert_array_equal(sc.water.pair_i[:k], ref.water.pair_i[:k])
    np.testing.assert_array_equal(sc.water.pair_j[:k], ref.water.pair_j[:k])
    same_values(sc.water.w[:k], ref.water.w[:k])
    same_values(sc.water.dwdx[:k], ref.water.dwdx[:k])
    s.close()
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    for nsteps, done, tol in ((1, 0, FP64_TOL_1STEP), (3, 1, 1e-11)):
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, ref.soil, done, nsteps - done)
        ew = assert_fields(sc.water, ref.water, tol, label=f"skf {skf} water")
        es = assert_fields(sc.soil, ref.soil, tol, label=f"skf {skf} soil")
        print(skf, nsteps, "water", max(ew.values()), "soil", max(es.values()))
    s.close()


def test_dam_break_config_as_stated():
    """The single-phase dam break exactly as configured: 50 x 100 water
    lattice at dx = 0.005 (no jitter), 3 wall layers in a 1.0 x 0.6 tank,
    cubic spline, h = 1.2 dx, c0 = 10 sqrt(2 g H), dt = 0.2 h / c0; 100 steps
    against the oracle, each field within 1e-12 norm-wise or within 16x the
    oracle's own change under a 1-ulp perturbation of its input density."""
    make = lambda: scenarios.dam_break(dx=0.005, skf=1, jitter=0.0)  # noqa: E731
    sc, ref = _both(make)
    assert sc.water.ntotal == 5000
    fields = ("x", "vx", "rho", "p", "drho", "dvx", "av", "txx", "txy", "tyy")
    s = Solver(sc.params, sc.water)
    done = 0
    for nsteps in (1, 100):
        print(f"steps {done}->{nsteps}", flush=True)
        s.time_integration(nsteps - done, sync=True)
        oracle.time_integration(ref.params, ref.water, None, done, nsteps - done)
        done = nsteps
        sens = _sensitivity(make, nsteps, fields)
        errs = field_errors(sc.water, ref.water, fields)
        bad = {f: (errs[f], sens[f]) for f in fields if not (errs[f] <= max(1e-12, 16.0 * sens[f]))}
        print(nsteps, "errs", {k: f"{v:.1e}" for k, v in errs.items()},
              "sens", {k: f"{v:.1e}" for k, v in sens.items()})
        assert not bad, f"{nsteps} steps: {bad}"
    s.close()


def test_two_phase_without_delta_and_xsph():
    """artificial_density and average_velocity off: no delta-SPH on rho, and
    av keeps its value (the reference does not touch it)."""
    def make():
        sc = _column()
        sc.params.artificial_density = 0
        sc.params.average_velocity = 0
        for ph in (sc.water, sc.soil):
            ph.av[:] = np.random.default_rng(1).normal(0.0, 1e-4, ph.av.shape)
        return sc
    sc, ref = _both(make)
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(3, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 3)
    assert_fields(sc.water, ref.water, 1e-11, label="water")
    assert_fields(sc.soil, ref.soil, 1e-11, label="soil")
    s.close()


def test_nnps2_fused_step_matches_link_list_oracle():
    sc, ref = _both(lambda: _tight_box_column(nnps=2))
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(1, sync=True)
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 1)
    assert_fields(sc.water, ref.water, FP64_TOL_1STEP, label="water")
    assert_fields(sc.soil, ref.soil, FP64_TOL_1STEP, label="soil")
    s.close()


def _advanced_state():
    """A two-phase state a few steps in, so velocities and rates are non-zero."""
    sc = small_cases()["column_jitter"]()
    oracle.time_integration(sc.params, sc.water, sc.soil, 0, 3)
    return sc


@pytest.mark.parametrize("phase", ["water", "soil"])
def test_granular_con_density_and_delta(phase):
    sc = _advanced_state()
    ref = sc.copy()
    s = Solver(sc.params, sc.water, sc.soil)
    s.con_density(phase)
    s.sync(phase, abi.F_DRHO)
    r = getattr(ref, phase)
    oracle.call_phase("find_pairs", ref.params, r)
    oracle.call_phase("con_density", ref.params, r)
    g = getattr(sc, phase)
    assert field_errors(g, r, ("drho",))["drho"] <= 1e-12
    s.delta_sph_density(phase)
    s.sync(phase, abi.F_DRHO)
    oracle.delta_sph(ref.params, r, r.rho, r.drho)
    assert field_errors(g, r, ("drho",))["drho"] <= 1e-12
    s.close()


@pytest.mark.parametrize("phase", ["water", "soil"])
def test_granular_internal_force_and_art_visc(phase):
    sc = _advanced_state()
    ref = sc.copy()
    s = Solver(sc.params, sc.water, sc.soil)
    g = getattr(sc, phase)
    r = getattr(ref, phase)
    # reference: zero rates, pairs, then the internal-force group (declarations_sph.f90:1222-1308)
    for ph in (g, r):
        ph.dvx[:] = 0.0
        if ph.material == abi.SPHG_SOIL:
            for f in ("dsxx", "dsxy", "dsyy", "dp"):
                getattr(ph, f)[:] = 0.0
    s.internal_force(phase)
    s.sync(phase, abi.F_ALL)
    oracle.call_phase("find_pairs", ref.params, r)
    P = ref.params
    oracle.call_phase("shear_strain_rate", P, r)
    if phase == "soil":
        oracle.call_phase("velocity_divergence", P, r)
    oracle.call_phase("pressure", P, r)
    n = r.n
    if phase == "water":
        oracle.call_phase("newtonian_fluid", P, r)
        t1 = oracle.df(r, r.p, 0); t2 = oracle.df(r, r.vof * r.sxx, 0); t3 = oracle.df(r, r.vof * r.sxy, 1)
        ax = -(r.vof * t1) + t2 + t3
        t1 = oracle.df(r, r.p, 1); t2 = oracle.df(r, r.vof * r.sxy, 0); t3 = oracle.df(r, r.vof * r.syy, 1)
        ay = -(r.vof * t1) + t2 + t3
    else:
        oracle.call_phase("drucker_prager", P, r)
        fp = r.vof * r.p
        ax = -oracle.df(r, fp, 0, "df3") + oracle.df(r, r.vof * r.sxx, 0, "df3") + oracle.df(r, r.vof * r.sxy, 1, "df3")
        ay = -oracle.df(r, fp, 1, "df3") + oracle.df(r, r.vof * r.sxy, 0, "df3") + oracle.df(r, r.vof * r.syy, 1, "df3")
    pos = r.rho > 0
    r.dvx[:, 0] = np.where(pos, ax / np.where(pos, r.rho, 1.0), ax)
    r.dvx[:, 1] = np.where(pos, ay / np.where(pos, r.rho, 1.0), ay)
    fields = ["dvx", "p", "c", "txx", "txy", "tyy"]
    if phase == "soil":
        oracle.call_phase("plastic_or_not", P, r)
        oracle.call_phase("plastic_flow_rule3", P, r)
        oracle.call_phase("jaumann_rate", P, r)
        fields += ["dsxx", "dsxy", "dsyy", "dp", "epsilon_p", "sxx", "sxy", "syy", "vcc", "wxy"]
    assert_fields(g, r, 1e-12, fields, label=f"internal_force {phase}")
    s.art_visc(phase)
    s.sync(phase, abi.F_DVX)
    oracle.call_phase("art_visc", P, r)
    assert_fields(g, r, 1e-12, ["dvx"], label=f"art_visc {phase}")
    assert n == g.n
    s.close()


def test_out_of_range_is_reported():
    sc = scenarios.savage_column()
    sc.water.x[0, 0] = 10.0  # outside the geometry box
    s = Solver(sc.params, sc.water, sc.soil)
    with pytest.raises(abi.SphgError) as e:
        s.time_integration(1, sync=True)
    assert e.value.code == abi.SPHG_ERR_OUT_OF_RANGE
    s.close()


def test_too_many_interactions_is_reported():
    sc = scenarios.savage_column(max_interaction=100)
    s = Solver(sc.params, sc.water, sc.soil)
    with pytest.raises(abi.SphgError) as e:
        s.direct_find("water")
    assert e.value.code == abi.SPHG_ERR_TOO_MANY_INTERACTIONS
    s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_counts_match_the_pair_searches(precision):
    # sphg_pair_counts after one fused step: the same-phase and cross pairs of
    # the state the step started from, each pair once -- direct_find's and
    # direct_find_2's niac
    sc = _column()
    refs = [sc.copy() for _ in range(3)]
    if precision == "mixed":
        sc.params.precision = abi.SPHG_MIXED
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(1, sync=True)
    nw, ns, nx = s.pair_counts()
    s.close()
    oracle.call_phase("direct_find", refs[0].params, refs[0].water)
    oracle.call_phase("direct_find", refs[1].params, refs[1].soil)
    oracle.call_cross("direct_find_2", refs[2].params, refs[2].water, refs[2].soil)
    got, want = (nw, ns, nx), (refs[0].water.niac, refs[1].soil.niac, refs[2].water.niac)
    if precision == "fp64":
        assert got == want
    else:
        # FP32 separations may flip pairs at r = 2h, where W and grad W vanish
        assert all(abs(g - w) <= 1e-4 * w for g, w in zip(got, want)), (got, want)


@pytest.mark.parametrize("kind", ["water_only", "soil_only"])
def test_pair_counts_single_phase(kind):
    # the single-phase water step and the single-phase soil step count their
    # same-phase pairs too (each pair once; no cross pairs)
    if kind == "water_only":
        sc = scenarios.dam_break(dx=0.01, jitter=0.05, seed=5)
    else:
        sc = scenarios.soil_collapse(jitter=0.05, seed=4)
    ref = sc.copy()
    s = Solver(sc.params, sc.water, sc.soil)
    s.time_integration(1, sync=True)
    nw, ns, nx = s.pair_counts()
    s.close()
    ph = ref.water if kind == "water_only" else ref.soil
    oracle.call_phase("direct_find", ref.params, ph)
    got = nw if kind == "water_only" else ns
    assert got == ph.niac > 0
    assert nx == 0 and (ns if kind == "water_only" else nw) == 0
