"""Closed-form and inequality pins of the oracle routines whose sign,
exponent or denominator the invariant tests (conservation, mirror and
Galilean symmetry) cannot see:

  art_visc          m_particles.f90:2504-2564  dissipation; two-particle Monaghan form
  pressure          m_particles.f90:2133-2175  c^2 = dp/drho of the Tait EOS
  sum_density       m_particles.f90:2243-2298  Shepard reproduction of a constant
  repulsive_force   m_particles.f90:1605-1642  real-virtual pair, Lennard-Jones form
  darcy_law         m_particles.f90:3832-3874  drag opposes the relative velocity
  av_vel            m_particles.f90:2761-2793  XSPH pulls v towards the partner
  plastic_or_not    m_particles.f90:3447-3505  the 1e-5 band around the yield surface

Each check is stated from the mathematics of the method (the textbook form of
the term, or an inequality it must satisfy), not from the oracle's code.
tests/test_oracle_mutants.py builds mutated oracles (a flipped sign, a wrong
exponent, a wrong denominator, a wider tolerance) and shows each one fails
the check here that is meant to catch it."""
import math

import numpy as np

import oracle
from sources_amd import abi
from sources_amd.state import Phase, default_params


def _pair_phase(material, xs, itypes, h, rho, mass, c=30.0, dspp=0.0):
    """A phase of the given particles (real ones first), pairs by direct_find."""
    itypes = np.asarray(itypes)
    nreal = int(np.count_nonzero(itypes > 0))
    ph = Phase(material, nreal, len(xs) - nreal, max_interaction=64, dspp=dspp)
    ph.x[:] = np.asarray(xs, float)
    ph.itype[:] = itypes
    ph.hsml[:] = h
    ph.rho[:] = rho
    ph.mass[:] = mass
    ph.c[:] = c
    ph.vof[:] = 1.0
    return ph


def _lattice(n, dx, jitter, seed, material=abi.SPHG_WATER, rho=1000.0):
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], 1)
    x = x + rng.uniform(-jitter * dx, jitter * dx, x.shape)
    ph = Phase(material, len(x), 0, max_interaction=40 * len(x), dspp=dx)
    ph.x[:] = x
    ph.hsml[:] = 1.2 * dx
    ph.itype[:] = 2 if material == abi.SPHG_WATER else 3
    ph.rho[:] = rho
    ph.mass[:] = rho * dx * dx
    ph.c[:] = 30.0
    ph.vof[:] = 1.0
    return ph


def _params(**kw):
    base = dict(skf=1, single_phase=1, x_mingeom=-1.0, x_maxgeom=1.0, y_mingeom=-1.0, y_maxgeom=1.0)
    base.update(kw)
    return default_params(**base)


# ---------------------------------------------------------------------------
# art_visc: Monaghan's viscosity Pi_ij = (-alpha cbar mu + beta mu^2) / rhobar,
# mu = hbar v_ij.x_ij / (r^2 + eta^2 hbar^2), acting only on approaching pairs,
# a_i = -sum_j m_j Pi_ij grad_i W.  It can only remove kinetic energy.
# ---------------------------------------------------------------------------

def test_art_visc_dissipates_kinetic_energy():
    P = _params()
    ph = _lattice(20, 0.005, 0.25, seed=11)
    rng = np.random.default_rng(12)
    ph.vx[:] = rng.normal(scale=0.3, size=ph.vx.shape)
    ph.rho[:] = 1000.0 * (1.0 + 0.02 * rng.normal(size=ph.maxn))
    ph.c[:] = 30.0 * (1.0 + 0.1 * rng.uniform(size=ph.maxn))
    oracle.call_phase("direct_find", P, ph)
    ph.dvx[:] = 0.0
    oracle.call_phase("art_visc", P, ph)
    power = float(np.sum(ph.mass[:, None] * ph.vx * ph.dvx))   # dE_kin/dt
    scale = float(np.sum(np.abs(ph.mass[:, None] * ph.vx * ph.dvx)))
    assert scale > 0.0
    assert power < -1e-3 * scale, (power, scale)


def test_art_visc_two_particle_closed_form():
    P = _params()
    h, r, u = 0.006, 0.004, 0.7
    m_i, m_j, rho_i, rho_j, c_i, c_j = 0.025, 0.03, 1000.0, 1010.0, 30.0, 32.0
    ph = _pair_phase(abi.SPHG_WATER, [[0.0, 0.0], [r, 0.0]], [2, 2], h, 0.0, 0.0)
    ph.mass[:2] = [m_i, m_j]
    ph.rho[:2] = [rho_i, rho_j]
    ph.c[:2] = [c_i, c_j]
    ph.vx[0] = [u, 0.0]          # particle 0 runs into particle 1
    oracle.call_phase("direct_find", P, ph)
    assert ph.niac == 1
    ph.dvx[:] = 0.0
    oracle.call_phase("art_visc", P, ph)
    # Monaghan's form, alpha = 0.1, beta = 0, eta = 0.1
    _, gradW_i = oracle.kernel(1, r, np.array([-r, 0.0]), h)   # grad_i W, x_i - x_j = (-r, 0)
    mu = h * (u * -r) / (r * r + (0.1 * h) ** 2)
    Pi = (-0.1 * 0.5 * (c_i + c_j) * mu) / (0.5 * (rho_i + rho_j))
    assert Pi > 0.0
    np.testing.assert_allclose(ph.dvx[0], -m_j * Pi * gradW_i, rtol=1e-13)
    np.testing.assert_allclose(ph.dvx[1], m_i * Pi * gradW_i, rtol=1e-13)
    assert ph.dvx[0, 0] < 0.0 < ph.dvx[1, 0]            # the approach is braked
    # receding particles feel nothing
    ph.vx[0] = [-u, 0.0]
    ph.dvx[:] = 0.0
    oracle.call_phase("art_visc", P, ph)
    assert np.all(ph.dvx == 0.0)


# ---------------------------------------------------------------------------
# pressure: Tait EOS p = B((rho/(rho0 vof))^gamma - 1).  With B = rho0 c0^2 /
# gamma the sound speed of the EOS is c^2 = vof dp/drho = c0^2 (rho/rho0 vof)^(gamma-1),
# i.e. c = c0 r^3 for gamma = 7: the exponents of both lines are fixed.
# ---------------------------------------------------------------------------

def test_pressure_sound_speed_is_the_eos_slope():
    c0, rho0, gamma = 31.32, 1000.0, 7.0
    P = _params(water_rho0=rho0, water_c=c0, water_gamma=gamma, water_b=rho0 * c0 * c0 / gamma)
    for vof in (1.0, 0.4):
        ph = _lattice(6, 0.005, 0.0, seed=0)
        n = ph.maxn
        rho = rho0 * vof * np.linspace(0.98, 1.02, n)
        d = 1e-4 * rho0 * vof

        def eos(r):
            ph.rho[:] = r
            ph.vof[:] = vof
            oracle.call_phase("pressure", P, ph)
            return ph.p.copy(), ph.c.copy()
        pp, _ = eos(rho + d)
        pm, _ = eos(rho - d)
        _, c = eos(rho)
        slope = (pp - pm) / (2 * d)
        np.testing.assert_allclose(c * c, vof * slope, rtol=1e-7)
        # and the pressure vanishes at the rest density
        p0, _ = eos(np.full(n, rho0 * vof))
        assert np.all(np.abs(p0) < 1e-9 * P.water_b)


# ---------------------------------------------------------------------------
# sum_density: rho_i = sum_j m_j W_ij / sum_j (m_j / rho_j) W_ij (j = i
# included, with W(0)), Shepard-normalised by the old densities.  A constant
# density is reproduced exactly whatever the particle volumes, and the
# result is homogeneous of degree one in the old densities.
# ---------------------------------------------------------------------------

def test_sum_density_reproduces_a_constant_density():
    P = _params(water_rho0=1000.0)
    ph = _lattice(16, 0.005, 0.3, seed=21, rho=1137.0)
    rng = np.random.default_rng(22)
    vol = 0.005 ** 2 * rng.uniform(0.8, 1.2, ph.maxn)
    ph.mass[:] = 1137.0 * vol
    oracle.call_phase("direct_find", P, ph)
    oracle.call_phase("sum_density", P, ph)
    np.testing.assert_allclose(ph.rho, 1137.0, rtol=1e-13)


def test_sum_density_scales_with_the_old_density():
    P = _params(water_rho0=1000.0)
    base = _lattice(12, 0.005, 0.3, seed=23)
    rng = np.random.default_rng(24)
    base.rho[:] = 1000.0 * (1.0 + 0.05 * rng.normal(size=base.maxn))
    oracle.call_phase("direct_find", P, base)
    a, b = base.copy(), base.copy()
    b.rho[:] = 1.7 * base.rho
    oracle.call_phase("sum_density", P, a)
    oracle.call_phase("sum_density", P, b)
    np.testing.assert_allclose(b.rho, 1.7 * a.rho, rtol=1e-13)


# ---------------------------------------------------------------------------
# repulsive_force: the wall force of Monaghan (1994) in Lennard-Jones form,
# a_real += D [(r0/r)^p1 - (r0/r)^p2] (x_real - x_wall) / r^2 for r <= r0,
# on the real particle only, pushing it away from the virtual one.
# ---------------------------------------------------------------------------

def test_repulsive_force_real_virtual_pair():
    P = _params()
    r0, h = 0.005, 0.006
    xr, xv = np.array([0.01, 0.02]), np.array([0.01 + 0.0024, 0.02 + 0.0032])   # r = 0.004 < r0
    ph = _pair_phase(abi.SPHG_WATER, [xr, xv], [2, -2], h, 1000.0, 0.025, dspp=r0)
    oracle.call_phase("direct_find", P, ph)
    assert ph.niac == 1
    ph.dvx[:] = 0.0
    oracle.call_phase("repulsive_force", P, ph)
    dx = xr - xv
    r = math.hypot(*dx)
    expect = P.dd * ((r0 / r) ** P.p1 - (r0 / r) ** P.p2) / r ** 2 * dx
    np.testing.assert_allclose(ph.dvx[0], expect, rtol=1e-13)
    assert np.dot(ph.dvx[0], dx) > 0.0          # away from the wall particle
    assert np.all(ph.dvx[1] == 0.0)             # the wall particle does not move
    # beyond r0, and between two real particles, there is no force
    far = _pair_phase(abi.SPHG_WATER, [xr, xr + [0.0055, 0.0]], [2, -2], h, 1000.0, 0.025, dspp=r0)
    oracle.call_phase("direct_find", P, far)
    assert far.niac == 1
    oracle.call_phase("repulsive_force", P, far)
    assert np.all(far.dvx == 0.0)
    both = _pair_phase(abi.SPHG_WATER, [xr, xv], [2, 2], h, 1000.0, 0.025, dspp=r0)
    oracle.call_phase("direct_find", P, both)
    oracle.call_phase("repulsive_force", P, both)
    assert np.all(both.dvx == 0.0)


# ---------------------------------------------------------------------------
# darcy_law: the drag of Darcy's law between co-located water and soil,
# f = n_w n_s rho_w |g| / k (v_w - v_s) per unit volume, smoothed with W:
# a_w = -m_s cf (v_w - v_s) W / (rho_w rho_s), a_s = +m_w cf (v_w - v_s) W / (rho_w rho_s).
# ---------------------------------------------------------------------------

def test_darcy_drag_two_particle_closed_form():
    P = _params(single_phase=0)
    h, k = 0.006, P.soil_permeability
    vof_w, vof_s, rho_w, rho_s, m_w, m_s = 0.4, 0.6, 400.0, 1590.0, 0.01, 0.04
    wa = _pair_phase(abi.SPHG_WATER, [[0.0, 0.0]], [2], h, rho_w, m_w)
    so = _pair_phase(abi.SPHG_SOIL, [[0.003, 0.001]], [3], h, rho_s, m_s)
    wa.vof[:] = vof_w
    so.vof[:] = vof_s
    wa.vx[0] = [0.2, -0.1]
    so.vx[0] = [-0.05, 0.04]
    oracle.call_cross("direct_find_2", P, wa, so)
    assert wa.niac == 1
    wa.dvx[:] = 0.0
    so.dvx[:] = 0.0
    oracle.call_cross("darcy_law", P, wa, so)
    r = math.hypot(0.003, 0.001)
    W, _ = oracle.kernel(1, r, np.array([-0.003, -0.001]), h)
    cf = vof_w * vof_s * rho_w * abs(P.numeric_gravity) / k
    dv = wa.vx[0] - so.vx[0]
    np.testing.assert_allclose(wa.dvx[0], -m_s * cf * dv * W / (rho_w * rho_s), rtol=1e-13)
    np.testing.assert_allclose(so.dvx[0], m_w * cf * dv * W / (rho_w * rho_s), rtol=1e-13)
    assert np.dot(wa.dvx[0], dv) < 0.0 < np.dot(so.dvx[0], dv)   # opposes the slip


# ---------------------------------------------------------------------------
# av_vel: XSPH, v_i <- v_i + eps sum_j m_j (v_j - v_i) / rhobar W_ij for the
# real particles: the velocity is pulled towards the neighbours' mean.
# ---------------------------------------------------------------------------

def test_xsph_two_particle_closed_form():
    P = _params()
    h, r = 0.006, 0.005
    ph = _pair_phase(abi.SPHG_WATER, [[0.0, 0.0], [0.0, r]], [2, 2], h, 1000.0, 0.025)
    ph.rho[:2] = [990.0, 1020.0]
    ph.mass[:2] = [0.024, 0.026]
    ph.vx[0] = [0.3, 0.1]
    ph.vx[1] = [-0.2, 0.4]
    oracle.call_phase("direct_find", P, ph)
    oracle.call_phase("av_vel", P, ph)
    W, _ = oracle.kernel(1, r, np.array([0.0, -r]), h)
    rbar = 0.5 * (990.0 + 1020.0)
    np.testing.assert_allclose(ph.av[0], P.epsilon * 0.026 * (ph.vx[1] - ph.vx[0]) / rbar * W, rtol=1e-13)
    np.testing.assert_allclose(ph.av[1], P.epsilon * 0.024 * (ph.vx[0] - ph.vx[1]) / rbar * W, rtol=1e-13)
    assert np.dot(ph.av[0], ph.vx[1] - ph.vx[0]) > 0.0
    assert np.dot(ph.av[1], ph.vx[0] - ph.vx[1]) > 0.0


# ---------------------------------------------------------------------------
# plastic_or_not: a point is flagged when |sqrt(J2) - alpha1 I1| < 1e-5 (FP32
# literal), sqrt(J2) = sqrt(s:s / 2) + 1e-6 over the plane-strain deviator.
# States are built at fixed distances from the yield surface.
# ---------------------------------------------------------------------------

def test_plastic_or_not_tolerance_band():
    phi = math.radians(30.0)
    P = default_params(soil_phi=phi)
    alpha1 = math.tan(phi) / math.sqrt(9.0 + 12.0 * math.tan(phi) ** 2)
    offsets = np.array([0.0, 0.5e-5, -0.5e-5, 0.9e-5, -0.9e-5, 2e-5, -2e-5, 1e-3, -1e-3])
    n = len(offsets)
    ph = Phase(abi.SPHG_SOIL, n, 0, max_interaction=1)
    p = np.full(n, 100.0)
    d = np.array([1.0, 0.5, -0.3])                       # direction of (sxx, sxy, syy)
    norm = math.sqrt((d[0] ** 2 + 2 * d[1] ** 2 + d[2] ** 2 + (d[0] + d[2]) ** 2) / 2.0)
    s = (alpha1 * 3.0 * p + offsets - 1e-6) / norm
    ph.p[:] = p
    ph.sxx[:], ph.sxy[:], ph.syy[:] = s * d[0], s * d[1], s * d[2]
    oracle.call_phase("plastic_or_not", P, ph)
    expect = (np.abs(offsets) < 1e-5).astype(np.int32)
    assert list(ph.fail[:n]) == list(expect), (list(ph.fail[:n]), list(expect))
    assert ph.nfail == int(expect.sum())
