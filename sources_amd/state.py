"""Host-side particle state and parameters.

`Phase` is the Python counterpart of the reference's ``type(particles)``
(m_particles.f90:62-384) restricted to the arrays the per-step loop touches.
Arrays keep the reference layout: vectors are Fortran ``x(2,maxn)``, i.e. a
C-contiguous ``(maxn, 2)`` numpy array; scalars have length ``maxn``.
"""
import ctypes as C

import numpy as np

from . import abi


def f32(v):
    """A Fortran default-real literal widened to double (e.g. -9.8 -> -9.800000190734863)."""
    return float(np.float32(v))


def default_params(**overrides):
    """Reference defaults: `type numerical` (m_particles.f90:31-55), module
    param switches (param.f90:78-103), grid box (m_particles.f90:89-102) and
    the module gravity (declarations_sph.f90:13)."""
    P = abi.Params()
    P.water_rho0, P.water_b, P.water_gamma, P.water_c, P.water_viscosity = 1000.0, 1.0e5, 7.0, 30.0, 1.0e-3
    P.soil_rho0, P.soil_k, P.soil_E, P.soil_niu, P.soil_c = 2650.0, 1.5e6, 1.8e6, 0.3, 30.0
    P.soil_porosity, P.soil_permeability = 0.4, 1.0e-3
    P.soil_cohesion, P.soil_phi = 0.0, np.deg2rad(30.0)
    P.nnps, P.skf = 1, 4
    P.numeric_gravity = f32(-9.8)
    P.alpha, P.beta, P.etq = 0.1, 0.0, 0.1
    P.dd, P.p1, P.p2 = 0.1, 12.0, 4.0
    P.delta = 0.1
    P.epsilon = f32(0.001)
    P.gravity = f32(-9.8)
    P.single_phase = 0
    P.summation_density = 0
    P.average_velocity = 1
    P.visc_artificial = 1
    P.artificial_density = 1
    P.soil_artificial_stress = 1
    P.volume_fraction = 1
    P.water_artificial_volume = 1
    P.volume_fraction_renorm = 1
    P.stress_integration = 1
    P.yield_criterion = 2
    P.plasticity = 3
    P.x_mingeom, P.x_maxgeom = f32(-0.05), f32(0.3)
    P.y_mingeom, P.y_maxgeom = f32(-0.05), f32(0.15)
    P.maxngx, P.maxngy = 100, 100
    P.dt = 1.0e-5
    P.precision = abi.SPHG_FP64
    for k, v in overrides.items():
        setattr(P, k, v)
    return P


def params_copy(P):
    Q = abi.Params()
    C.memmove(C.addressof(Q), C.addressof(P), C.sizeof(abi.Params))
    return Q


_VEC = {"x", "vx", "v_min", "dvx", "av"}


class Phase:
    """All arrays of one phase (water or soil), owned by the host."""

    def __init__(self, material, ntotal, nvirt, maxn=None, max_interaction=0, dspp=0.0):
        self.material = material
        self.ntotal = int(ntotal)
        self.nvirt = int(nvirt)
        n = self.ntotal + self.nvirt
        self.maxn = int(maxn if maxn is not None else max(n, 1))
        self.max_interaction = int(max_interaction)
        self.niac = 0
        self.nfail = 0
        self.dspp = float(dspp)
        m = self.maxn
        self.itype = np.zeros(m, np.int32)
        self.fail = np.zeros(m, np.int32)
        self.countiac = np.zeros(m, np.int32)
        for name in abi.VIEW_DBL_ARRAYS:
            if name in ("w", "dwdx"):
                continue
            shape = (m, 2) if name in _VEC else (m,)
            setattr(self, name, np.zeros(shape, np.float64))
        mi = max(self.max_interaction, 1)
        self.pair_i = np.zeros(mi, np.int32)
        self.pair_j = np.zeros(mi, np.int32)
        self.w = np.zeros(mi, np.float64)
        self.dwdx = np.zeros((mi, 2), np.float64)
        self._view = None

    @property
    def n(self):
        return self.ntotal + self.nvirt

    def set_max_interaction(self, mi):
        self.max_interaction = int(mi)
        mi = max(self.max_interaction, 1)
        self.pair_i = np.zeros(mi, np.int32)
        self.pair_j = np.zeros(mi, np.int32)
        self.w = np.zeros(mi, np.float64)
        self.dwdx = np.zeros((mi, 2), np.float64)

    def copy(self):
        q = Phase.__new__(Phase)
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                setattr(q, k, v.copy())
            elif k != "_view":
                setattr(q, k, v)
        q._view = None
        return q

    def view(self):
        """Build a PhaseView pointing at this object's arrays (kept alive by self)."""
        v = abi.PhaseView()
        v.ntotal, v.nvirt, v.maxn = self.ntotal, self.nvirt, self.maxn
        v.max_interaction = self.max_interaction
        v.material = self.material
        v.niac, v.nfail = self.niac, self.nfail
        v.dspp = self.dspp
        for name in abi.VIEW_INT_ARRAYS:
            a = getattr(self, name)
            assert a.dtype == np.int32 and a.flags.c_contiguous
            setattr(v, name, a.ctypes.data_as(C.POINTER(C.c_int32)))
        for name in abi.VIEW_DBL_ARRAYS:
            a = getattr(self, name)
            assert a.dtype == np.float64 and a.flags.c_contiguous
            setattr(v, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        self._view = v
        return v

    def pull_scalars(self, v):
        """Copy the scalar outputs (niac, nfail) of a view back."""
        self.niac = int(v.niac)
        self.nfail = int(v.nfail)

    def pairs(self):
        """Pair list of the last pair search, as (pair_i, pair_j, w, dwdx)."""
        k = self.niac
        return self.pair_i[:k].copy(), self.pair_j[:k].copy(), self.w[:k].copy(), self.dwdx[:k].copy()
