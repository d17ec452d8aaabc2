"""Seeded scenario generators.

They follow the conventions of the reference's input routines
(input_mix_Savage.f90:1-145, input_tank.f90:1-75): particles on the cell
centres of a lattice block (get_cell_center_sub, m_particles.f90:461-479),
real particles first and fixed virtual boundary particles after them
(take_real_points1 / take_virtual_points1, m_particles.f90:808-893), itype
+-2 / +-3 (setup_itype, :896-913), hydrostatic water pressure with density
from the inverse Tait law (initial_density, :2105-2130), K0 soil stresses
sxx = 0.1905 p, syy = -0.381 p, p = 0.619 p (input_mix_Savage.f90:122-127)
and co-located water/soil particles for saturated soil.

Jitter comes from a SplitMix64 counter generator, so any host computes the
same positions bit for bit from a seed.
"""
import math

import numpy as np

from . import abi
from .state import Phase, default_params, f32

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n, stream=0):
    """n uint64 outputs of SplitMix64 started at seed + stream * 2**32."""
    with np.errstate(over="ignore"):
        gamma = np.uint64(0x9E3779B97F4A7C15)
        base = np.uint64((int(seed) + (int(stream) << 32)) & 0xFFFFFFFFFFFFFFFF)
        z = base + gamma * (np.arange(1, n + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, stream=0):
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def block_centers(xl, yl, m, n):
    """Cell centres of an m x n block, x-major / y-minor (m_particles.f90:461-479)."""
    dx = xl / m
    dy = yl / n
    i = np.repeat(np.arange(m, dtype=np.float64), n)
    j = np.tile(np.arange(n, dtype=np.float64), m)
    return i * dx + dx / 2.0, j * dy + dy / 2.0, dx, dy


def _make_phase(material, xr, yr, xv, yv, dx, dy, hfac, max_interaction):
    nr, nv = len(xr), len(xv)
    ph = Phase(material, nr, nv, maxn=nr + nv, max_interaction=max_interaction, dspp=dx)
    ph.x[:nr, 0], ph.x[:nr, 1] = xr, yr
    ph.x[nr:, 0], ph.x[nr:, 1] = xv, yv
    ph.hsml[:] = dx * hfac
    tag = 2 if material == abi.SPHG_WATER else 3
    ph.itype[:nr] = tag
    ph.itype[nr:] = -tag
    ph.vol = np.full(nr + nv, dx * dy)
    return ph


def _water_fields(ph, P, p, vof):
    """p, c, vof, rho = rho0 (p/b+1)^(1/gamma) * vof, mass = vol*rho (input_mix_Savage.f90:93-112)."""
    ph.p[:] = p
    ph.c[:] = P.water_c
    ph.vof[:] = vof
    ph.rho[:] = P.water_rho0 * (ph.p / P.water_b + 1.0) ** (1.0 / P.water_gamma)
    ph.rho[:] = ph.rho * ph.vof
    ph.mass[:] = ph.vol * ph.rho


def _soil_fields(ph, P, p0, vof):
    """K0 stresses and density (input_mix_Savage.f90:118-141)."""
    ph.sxy[:] = 0.0
    ph.sxx[:] = f32(0.1905) * p0
    ph.syy[:] = -f32(0.381) * p0
    ph.p[:] = f32(0.619) * p0
    ph.c[:] = P.soil_c
    ph.vof[:] = vof
    ph.rho[:] = P.soil_rho0 * ph.vof
    ph.mass[:] = ph.vol * ph.rho


class Scenario:
    """A ready-to-run state: params, water phase and (two-phase) soil phase;
    a single-phase soil run has no water phase."""

    def __init__(self, name, params, water, soil=None, nsteps=1, notes=""):
        self.name = name
        self.params = params
        self.water = water
        self.soil = soil
        self.nsteps = nsteps
        self.notes = notes

    @property
    def n_particles(self):
        return (self.water.n if self.water is not None else 0) + (self.soil.n if self.soil is not None else 0)

    def copy(self):
        from .state import params_copy
        return Scenario(self.name, params_copy(self.params), self.water.copy() if self.water is not None else None,
                        self.soil.copy() if self.soil is not None else None,
                        self.nsteps, self.notes)


def savage_column(max_interaction=None, **pover):
    """The reference's own two-phase case, input_mix_Savage.f90: a 52 x 26
    lattice (dx = 0.005) shifted by -0.01, water zones 2/4 real, soil zone 2
    real, zones 1/3/5/6 virtual for both phases, water vof 0.45 in zones 1-2."""
    P = default_params(**pover)
    g = P.gravity
    x, y, dx, dy = block_centers(0.26, 0.13, 52, 26)
    x = x - f32(0.01)
    y = y - f32(0.01)
    zone = np.full(x.shape, 2, np.int32)
    zone[(x < 0.0) | (x > f32(0.24)) | (y < 0.0)] = 1
    zone[(zone == 1) & (x > f32(0.06))] = 3
    zone[(zone == 1) & (y > f32(0.048))] = 5
    zone[(zone == 5) & (y > f32(0.1))] = 6
    zone[(zone == 3) & (y > f32(0.1))] = 6
    zone[(zone == 2) & (x > f32(0.06))] = 4
    zone[(zone == 2) & (y > f32(0.048))] = 4
    zone[(zone == 4) & (y > f32(0.1))] = 7

    def take(zs):
        sel = np.concatenate([np.nonzero(zone == z)[0] for z in zs])
        return x[sel], y[sel], zone[sel]

    xr, yr, zr = take([2, 4])
    xv, yv, zv = take([1, 3, 5, 6])
    hfac = f32(1.2)
    mi = max_interaction if max_interaction is not None else 40 * (len(xr) + len(xv))
    wa = _make_phase(abi.SPHG_WATER, xr, yr, xv, yv, dx, dy, hfac, mi)
    zw = np.concatenate([zr, zv])
    water_surface = f32(0.1)
    p = P.water_rho0 * g * (wa.x[:, 1] - water_surface)
    p[zw == 6] = 0.0
    vof = np.where((zw == 1) | (zw == 2), f32(0.45), 1.0)
    _water_fields(wa, P, p, vof)

    xr, yr, zr = take([2])
    xv, yv, zv = take([1, 3, 5, 6])
    so = _make_phase(abi.SPHG_SOIL, xr, yr, xv, yv, dx, dy, hfac, mi)
    zs = np.concatenate([zr, zv])
    soil_surface = f32(0.048)
    p0 = (P.soil_rho0 - P.water_rho0) * g * (so.x[:, 1] - soil_surface)
    p0[(zs == 3) | (zs == 5) | (zs == 6)] = 0.0
    _soil_fields(so, P, p0, f32(0.55))
    return Scenario("savage_column", P, wa, so, nsteps=1,
                    notes="input_mix_Savage.f90 lattice, 52x26 at dx=0.005")


def dam_break(dx=0.005, width=0.25, height=0.5, tank_w=1.0, tank_h=0.6, layers=3,
              skf=1, max_interaction=None, jitter=0.0, seed=0, **pover):
    """Single-phase dam break: water column width x height at the left wall of
    a tank_w x tank_h tank, `layers` rows of fixed virtual particles on the
    bottom and both sides.  Cubic spline, Tait EOS with c0 = 10 sqrt(2 g H),
    B = rho0 c0^2 / gamma, dt = 0.2 h / c0."""
    g = -9.81
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * height)
    rho0, gam = 1000.0, 7.0
    b = rho0 * c0 * c0 / gam
    m = int(round((tank_w + 2 * layers * dx) / dx))
    n = int(round((tank_h + layers * dx) / dx))
    x, y, ddx, ddy = block_centers(m * dx, n * dx, m, n)
    x = x - layers * dx
    y = y - layers * dx
    real = (x > 0.0) & (x < width) & (y > 0.0) & (y < height)
    virt = (x < 0.0) | (x > tank_w) | (y < 0.0)
    xr, yr = x[real], y[real]
    if jitter > 0.0:
        u = uniform(seed, 2 * len(xr))
        xr = xr + (2.0 * u[0::2] - 1.0) * jitter * dx
        yr = yr + (2.0 * u[1::2] - 1.0) * jitter * dx
    xv, yv = x[virt], y[virt]
    h = ddx * f32(1.2)
    P = default_params(single_phase=1, skf=skf, gravity=g, numeric_gravity=g,
                       water_rho0=rho0, water_b=b, water_gamma=gam, water_c=c0,
                       water_viscosity=1.0e-3, dt=0.2 * h / c0,
                       x_mingeom=-0.05, x_maxgeom=tank_w + 0.05,
                       y_mingeom=-0.05, y_maxgeom=2.0 * tank_h)
    for k, v in pover.items():
        setattr(P, k, v)
    mi = max_interaction if max_interaction is not None else 16 * (len(xr) + len(xv))
    wa = _make_phase(abi.SPHG_WATER, xr, yr, xv, yv, ddx, ddy, f32(1.2), mi)
    p = np.maximum(P.water_rho0 * g * (wa.x[:, 1] - height), 0.0)
    _water_fields(wa, P, p, 1.0)
    return Scenario("dam_break", P, wa, None, nsteps=2000,
                    notes=f"{len(xr)} water + {len(xv)} virtual, dx={dx}")


def soil_collapse(dx=0.005, width=0.1, height=0.05, tank_w=0.3, tank_h=0.12, layers=3, skf=1,
                  max_interaction=None, jitter=0.0, seed=0, **pover):
    """Single-phase soil (the reference's time_integration_for_soil /
    single_step_for_soil, declarations_sph.f90:978-1127,1530-1655): a dry
    granular column width x height at the left wall of a tank_w x tank_h
    tank, `layers` rows of fixed virtual particles on the bottom and both
    sides, vof = 1, K0 stresses of the dry column with the
    input_mix_Savage.f90 factors."""
    P = default_params(single_phase=1, skf=skf, x_mingeom=-0.05, x_maxgeom=tank_w + 0.05,
                       y_mingeom=-0.05, y_maxgeom=2.0 * tank_h)
    for k, v in pover.items():
        setattr(P, k, v)
    m = int(round((tank_w + 2 * layers * dx) / dx))
    n = int(round((tank_h + layers * dx) / dx))
    x, y, ddx, ddy = block_centers(m * dx, n * dx, m, n)
    x = x - layers * dx
    y = y - layers * dx
    real = (x > 0.0) & (x < width) & (y > 0.0) & (y < height)
    virt = (x < 0.0) | (x > tank_w) | (y < 0.0)
    xr, yr = x[real], y[real]
    if jitter > 0.0:
        u = uniform(seed, 2 * len(xr))
        xr = xr + (2.0 * u[0::2] - 1.0) * jitter * dx
        yr = yr + (2.0 * u[1::2] - 1.0) * jitter * dx
    xv, yv = x[virt], y[virt]
    mi = max_interaction if max_interaction is not None else 16 * (len(xr) + len(xv))
    so = _make_phase(abi.SPHG_SOIL, xr, yr, xv, yv, ddx, ddy, f32(1.2), mi)
    p0 = np.maximum(P.soil_rho0 * P.gravity * (so.x[:, 1] - height), 0.0)
    _soil_fields(so, P, p0, 1.0)
    return Scenario("soil_collapse", P, None, so, nsteps=1000,
                    notes=f"{len(xr)} soil + {len(xv)} virtual, dx={dx}")


def submerged_column(dx=0.0005, soil_w=0.2, soil_h=0.1, tank_w=0.8, tank_h=0.3, layers=3,
                     jitter=0.05, seed=1, skf=1, porosity=0.4, phi_deg=30.0,
                     permeability=1.0e-3, max_interaction=None, **pover):
    """Two-phase submerged granular column collapse: a soil column soil_w x
    soil_h at the left wall of a water-filled tank_w x tank_h tank, pore water
    co-located inside the column (water vof = porosity, soil vof = 1-porosity).
    Drucker-Prager soil, Darcy coupling, seeded +-jitter*dx lattice jitter."""
    g = -9.81
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * tank_h)
    E, nu, rho_g = 1.8e6, 0.3, 2650.0
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    cs = math.sqrt((K + 4.0 * G / 3.0) / (rho_g * (1.0 - porosity)))
    m = int(round((tank_w + 2 * layers * dx) / dx))
    n = int(round((tank_h + layers * dx) / dx))
    x, y, ddx, ddy = block_centers(m * dx, n * dx, m, n)
    x = x - layers * dx
    y = y - layers * dx
    inside = (x > 0.0) & (x < tank_w) & (y > 0.0) & (y < tank_h)
    virt = (x < 0.0) | (x > tank_w) | (y < 0.0)
    in_soil = inside & (x < soil_w) & (y < soil_h)
    h = ddx * f32(1.2)
    P = default_params(single_phase=0, skf=skf, gravity=g, numeric_gravity=g,
                       water_rho0=1000.0, water_b=1000.0 * c0 * c0 / 7.0, water_gamma=7.0,
                       water_c=c0, water_viscosity=1.0e-3,
                       soil_rho0=rho_g, soil_k=K, soil_E=E, soil_niu=nu, soil_c=cs,
                       soil_porosity=porosity, soil_permeability=permeability,
                       soil_cohesion=0.0, soil_phi=math.radians(phi_deg),
                       dt=0.2 * h / max(c0, cs),
                       x_mingeom=-0.05 * tank_w - 4 * layers * dx, x_maxgeom=1.05 * tank_w + 4 * layers * dx,
                       y_mingeom=-0.05 * tank_h - 4 * layers * dx, y_maxgeom=2.0 * tank_h)
    for k, v in pover.items():
        setattr(P, k, v)

    def jit(xs, ys, stream):
        if jitter <= 0.0:
            return xs, ys
        u = uniform(seed, 2 * len(xs), stream)
        return xs + (2.0 * u[0::2] - 1.0) * jitter * ddx, ys + (2.0 * u[1::2] - 1.0) * jitter * ddx

    xw, yw = jit(x[inside], y[inside], 0)
    xs, ys = jit(x[in_soil], y[in_soil], 1)
    xv, yv = x[virt], y[virt]
    nw, ns, nv = len(xw), len(xs), len(xv)
    mi_w = max_interaction if max_interaction is not None else 16 * (nw + nv) + 32 * ns
    mi_s = max_interaction if max_interaction is not None else 16 * (ns + nv)
    wa = _make_phase(abi.SPHG_WATER, xw, yw, xv, yv, ddx, ddy, f32(1.2), mi_w)
    so = _make_phase(abi.SPHG_SOIL, xs, ys, xv, yv, ddx, ddy, f32(1.2), mi_s)
    wsoil = np.concatenate([in_soil[inside], np.zeros(nv, bool)])
    wvirt_soil = np.concatenate([np.zeros(nw, bool), (xv < soil_w) & (yv < soil_h)])
    vof_w = np.where(wsoil | wvirt_soil, porosity, 1.0)
    pw = P.water_rho0 * g * (wa.x[:, 1] - tank_h)
    _water_fields(wa, P, pw, vof_w)
    p0 = (P.soil_rho0 - P.water_rho0) * g * (so.x[:, 1] - soil_h)
    p0[(so.x[:, 0] > soil_w) | (so.x[:, 1] > soil_h)] = 0.0
    p0 = np.maximum(p0, 0.0)
    _soil_fields(so, P, p0, 1.0 - porosity)
    return Scenario("submerged_column", P, wa, so, nsteps=5000,
                    notes=f"water {nw}+{nv} virt, soil {ns}+{nv} virt, dx={dx}")


def slope(dx=0.001, scale=1.0, layers=3, jitter=0.1, seed=2, skf=1, porosity=0.4,
          phi_deg=35.0, cohesion=2.0e3, slope_deg=30.0, permeability=1.0e-3,
          max_interaction=None, **pover):
    """Two-phase slope failure: a soil body whose surface rises at slope_deg
    from 0.15*scale at the left wall to 0.65*scale, in a tank 1.2*scale wide,
    partially submerged: water stands at 0.35*scale, as free water above the
    slope toe and as pore water co-located in the soil below that level.
    At dx = 0.001, scale = 1 this is ~563k soil and ~420k water real
    particles (~1.0M with the walls)."""
    g = -9.81
    W, h0, hmax, hw = 1.2 * scale, 0.15 * scale, 0.65 * scale, 0.35 * scale
    tan_s = math.tan(math.radians(slope_deg))
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * hw)
    E, nu, rho_g = 1.8e6, 0.3, 2650.0
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    cs = math.sqrt((K + 4.0 * G / 3.0) / (rho_g * (1.0 - porosity)))
    top = hmax + 2 * layers * dx
    m = int(round((W + 2 * layers * dx) / dx))
    n = int(round((top + layers * dx) / dx))
    x, y, ddx, ddy = block_centers(m * dx, n * dx, m, n)
    x = x - layers * dx
    y = y - layers * dx
    ys = np.minimum(hmax, h0 + tan_s * x)
    inside = (x > 0.0) & (x < W) & (y > 0.0)
    soil_r = inside & (y < ys)
    water_r = inside & (y < hw)
    virt = ((x < 0.0) | (x > W) | (y < 0.0)) & (y < top)
    h = ddx * f32(1.2)
    P = default_params(single_phase=0, skf=skf, gravity=g, numeric_gravity=g,
                       water_rho0=1000.0, water_b=1000.0 * c0 * c0 / 7.0, water_gamma=7.0,
                       water_c=c0, water_viscosity=1.0e-3,
                       soil_rho0=rho_g, soil_k=K, soil_E=E, soil_niu=nu, soil_c=cs,
                       soil_porosity=porosity, soil_permeability=permeability,
                       soil_cohesion=cohesion, soil_phi=math.radians(phi_deg),
                       dt=0.2 * h / max(c0, cs),
                       x_mingeom=-0.1 * W, x_maxgeom=1.1 * W,
                       y_mingeom=-0.1 * W, y_maxgeom=1.5 * top)
    for k, v in pover.items():
        setattr(P, k, v)

    def jit(xs, yv, stream):
        if jitter <= 0.0:
            return xs, yv
        u = uniform(seed, 2 * len(xs), stream)
        return xs + (2.0 * u[0::2] - 1.0) * jitter * ddx, yv + (2.0 * u[1::2] - 1.0) * jitter * ddx

    xw, yw = jit(x[water_r], y[water_r], 0)
    xs_, ys_ = jit(x[soil_r], y[soil_r], 1)
    xv, yv = x[virt], y[virt]
    nw, ns, nv = len(xw), len(xs_), len(xv)
    mi_w = max_interaction if max_interaction is not None else 16 * (nw + nv) + 32 * ns
    mi_s = max_interaction if max_interaction is not None else 16 * (ns + nv)
    wa = _make_phase(abi.SPHG_WATER, xw, yw, xv, yv, ddx, ddy, f32(1.2), mi_w)
    so = _make_phase(abi.SPHG_SOIL, xs_, ys_, xv, yv, ddx, ddy, f32(1.2), mi_s)
    w_in_soil = np.concatenate([(y[water_r] < ys[water_r]), (yv < np.minimum(hmax, h0 + tan_s * np.clip(xv, 0, W)))])
    vof_w = np.where(w_in_soil, porosity, 1.0)
    pw = P.water_rho0 * g * (wa.x[:, 1] - hw)
    pw = np.maximum(pw, 0.0)
    _water_fields(wa, P, pw, vof_w)
    surf = np.minimum(hmax, h0 + tan_s * np.clip(so.x[:, 0], 0.0, W))
    p0 = np.maximum((P.soil_rho0 - P.water_rho0) * g * (so.x[:, 1] - surf), 0.0)
    _soil_fields(so, P, p0, 1.0 - porosity)
    return Scenario("slope", P, wa, so, nsteps=1000,
                    notes=f"water {nw}+{nv} virt, soil {ns}+{nv} virt, dx={dx}, scale={scale}")


def sweep(ncases=64, seed=11, nmin=5000, nmax=120000, jitter=0.05):
    """A scenario sweep of submerged granular column collapses with uneven
    sizes and materials, all drawn from one seed (SplitMix64):
    total particles ~ logU[nmin, nmax], column aspect (width / height)
    ~ U[0.5, 2], friction angle ~ U[25, 40] degrees, permeability
    ~ logU[1e-4, 1e-2] m/s.  The tank is three column widths wide and 1.5
    column heights tall; dx follows from the particle target."""
    u = uniform(seed, 4 * ncases, 7)
    cases = []
    for i in range(ncases):
        a, b, c, d = u[4 * i:4 * i + 4]
        n_target = math.exp(math.log(nmin) + a * (math.log(nmax) - math.log(nmin)))
        aspect = 0.5 + 1.5 * b
        phi = 25.0 + 15.0 * c
        perm = math.exp(math.log(1e-4) + d * (math.log(1e-2) - math.log(1e-4)))
        soil_h = 0.1
        soil_w = aspect * soil_h
        tank_w, tank_h = 3.0 * soil_w, 1.5 * soil_h
        dx = math.sqrt((tank_w * tank_h + soil_w * soil_h) / n_target)
        cases.append(submerged_column(dx=dx, soil_w=soil_w, soil_h=soil_h, tank_w=tank_w, tank_h=tank_h,
                                      jitter=jitter, seed=seed * 1000 + i, phi_deg=phi, permeability=perm,
                                      max_interaction=0))
    return cases
