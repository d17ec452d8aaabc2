"""Batched scenarios: contexts advanced by replays of one captured hipGraph
(per period: the neighbour-list rebuild period of FP64 contexts; staged: the
period's first step context by context over several streams, every later
step one launch per kernel for all contexts; or every step context by
context) give bit-identical states to the same contexts stepped alone."""
import numpy as np
import pytest

from sources_amd import abi, scenarios
from sources_amd.batch import Batch
from sources_amd.sph import Solver

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vx", "rho", "p", "c", "vof", "drho", "dvx", "sxx", "sxy", "syy", "epsilon_p", "dsxx", "dp")


def cases():
    return scenarios.sweep(4, seed=5, nmin=2000, nmax=8000)


@pytest.mark.parametrize("nsteps,staged,prec,vary_h", [(40, "1", "fp64", False), (140, "1", "fp64", False),
                                                       (140, "0", "fp64", False), (140, "1", "mixed", False),
                                                       (140, "1", "fp64", True)])
def test_batch_is_bit_identical_to_solo(nsteps, staged, prec, vary_h, monkeypatch):
    # 40 steps: eager only; 140: eager to step 64, one replayed period of 64
    # steps (65-128, starting with the list rebuild), eager remainder; the
    # captured period staged or context by context; FP64 or mixed precision;
    # vary_h: one case's water smoothing lengths not uniform, so its water
    # kernels are other instantiations than the other cases' and the staged
    # capture replays those steps case by case over the streams
    monkeypatch.setenv("SPHG_BATCH_STAGED", staged)
    solo = cases()
    bat = cases()
    for c in solo + bat:
        c.params.precision = abi.SPHG_MIXED if prec == "mixed" else abi.SPHG_FP64
    if vary_h:
        for c in (solo[1], bat[1]):
            c.water.hsml[0] *= 1.001
    print("particles", [c.water.n + c.soil.n for c in solo], flush=True)
    ss = [Solver(c.params, c.water, c.soil) for c in solo]
    for s in ss:
        s.time_integration(nsteps, sync=True)
    sb = [Solver(c.params, c.water, c.soil) for c in bat]
    b = Batch(sb)
    b.run(nsteps)
    assert b.staged == (staged == "1")
    for s in sb:
        s.sync_all()
    for i, (a, c) in enumerate(zip(solo, bat)):
        for ph in ("water", "soil"):
            pa, pc = getattr(a, ph), getattr(c, ph)
            n = pa.n
            for f in FIELDS:
                if not hasattr(pa, f):
                    continue
                x, y = getattr(pa, f)[:n], getattr(pc, f)[:n]
                assert np.array_equal(x, y), (i, ph, f, float(np.max(np.abs(x - y))))
    assert all(s.itimestep == nsteps for s in sb)
    b.close()
    for s in ss + sb:
        s.close()


def test_full_sweep_is_bit_identical_to_solo(monkeypatch):
    """The whole 64-case sweep (N ~ logU[5k, 120k], 2.41M particles, seed 11):
    33 steps batched (eager to the period boundary, two captured periods of
    16 steps replayed in the staged form, eager remainder) equal the same 33
    steps of each case alone (both with the neighbour lists rebuilt every 16
    steps)."""
    monkeypatch.setenv("SPHG_NL_EVERY", "16")
    nsteps = 33
    solo = scenarios.sweep(64, seed=11)
    bat = scenarios.sweep(64, seed=11)
    print("cases", len(solo), "particles", sum(c.water.n + c.soil.n for c in solo), flush=True)
    for c in solo:
        s = Solver(c.params, c.water, c.soil)
        s.time_integration(nsteps, sync=True)
        s.close()
    sb = [Solver(c.params, c.water, c.soil) for c in bat]
    b = Batch(sb)
    b.run(nsteps)
    assert b.staged
    for s in sb:
        s.sync_all()
    for i, (a, c) in enumerate(zip(solo, bat)):
        for ph in ("water", "soil"):
            pa, pc = getattr(a, ph), getattr(c, ph)
            n = pa.n
            for f in FIELDS:
                x, y = getattr(pa, f)[:n], getattr(pc, f)[:n]
                assert np.array_equal(x, y), (i, ph, f, float(np.max(np.abs(x - y))))
    assert all(s.itimestep == nsteps for s in sb)
    b.close()
    for s in sb:
        s.close()


def test_renormalising_contexts_are_refused():
    """A captured period replays fixed step numbers, so a context whose
    every-40-step renormalisation persists (summation_density) is refused."""
    c = cases()[0]
    c.params.summation_density = 1
    s = Solver(c.params, c.water, c.soil)
    with pytest.raises(abi.SphgError) as e:
        Batch([s])
    assert e.value.code == abi.SPHG_ERR_UNSUPPORTED
    s.close()


def test_autotuner_picks_a_stream_count():
    bat = cases()
    sb = [Solver(c.params, c.water, c.soil) for c in bat]
    b = Batch(sb)
    ns, ms = b.tune(trial_periods=1)
    print("best streams", ns, "ms per step", ms)
    assert ns in (0, 1, 2, 4) and ms > 0.0
    b.run(20)
    for s in sb:
        s.sync_all()
    for c in bat:
        assert np.isfinite(c.water.rho[:c.water.n]).all() and np.isfinite(c.soil.x[:c.soil.n]).all()
    b.close()
    for s in sb:
        s.close()
