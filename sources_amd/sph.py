"""Python host of the per-step particle loop, mirroring the reference's call
structure (sph.f90 -> time_integration -> single_step -> link_list /
direct_find / density / internal force / art_visc) on top of the C ABI in
include/sphg.h.  Every call goes to libsphg.so; there is no CPU fallback.
"""
import ctypes as C

from . import abi
from .state import Phase  # noqa: F401  (re-exported for callers)


def _phase_id(phase):
    if phase in (abi.SPHG_WATER, "water", "parts"):
        return abi.SPHG_WATER
    if phase in (abi.SPHG_SOIL, "soil"):
        return abi.SPHG_SOIL
    raise ValueError(f"unknown phase {phase!r}")


class Solver:
    """One device context holding the water phase (`parts`) and, for
    two-phase runs, the soil phase (`soil`); a single-phase soil run binds
    the soil phase alone (water=None)."""

    def __init__(self, params, water, soil=None, lib=None):
        self.lib = lib or abi.load_library()
        self.params = params
        self.water = water
        self.soil = soil
        self._ctx = C.c_void_p()
        self._check(self.lib.sphg_create(C.byref(params), C.byref(self._ctx)), "sphg_create")
        self._views = {}
        if water is not None:
            self.bind(abi.SPHG_WATER, water)
        if soil is not None:
            self.bind(abi.SPHG_SOIL, soil)

    # -- plumbing --------------------------------------------------------
    def _check(self, rc, where):
        if rc != abi.SPHG_OK:
            if rc == abi.SPHG_ERR_HIP and self._ctx:
                where = f"{where} (hipError {self.lib.sphg_last_hip_error(self._ctx)})"
            raise abi.SphgError(rc, where)

    def _ph(self, phase):
        return self.water if _phase_id(phase) == abi.SPHG_WATER else self.soil

    def _pull(self, phase):
        pid = _phase_id(phase)
        v = self._views[pid]
        self._ph(pid).pull_scalars(v)

    def bind(self, phase, ph):
        pid = _phase_id(phase)
        v = ph.view()
        self._views[pid] = v
        self._check(self.lib.sphg_bind_phase(self._ctx, pid, C.byref(v)), "sphg_bind_phase")

    def set_params(self, params):
        self.params = params
        self._check(self.lib.sphg_set_params(self._ctx, C.byref(params)), "sphg_set_params")

    def sync(self, phase="water", mask=abi.F_ALL, to_host=True):
        pid = _phase_id(phase)
        self._check(self.lib.sphg_sync(self._ctx, pid, C.c_uint64(mask), 1 if to_host else 0), "sphg_sync")
        self._pull(pid)

    def sync_all(self, mask=abi.F_ALL):
        if self.water is not None:
            self.sync("water", mask)
        if self.soil is not None:
            self.sync("soil", mask)

    def upload_all(self, mask=abi.F_ALL):
        """Host arrays -> device for every bound phase."""
        if self.water is not None:
            self.sync("water", mask, to_host=False)
        if self.soil is not None:
            self.sync("soil", mask, to_host=False)

    @property
    def itimestep(self):
        return self.lib.sphg_get_itimestep(self._ctx)

    @itimestep.setter
    def itimestep(self, v):
        self._check(self.lib.sphg_set_itimestep(self._ctx, int(v)), "sphg_set_itimestep")

    @property
    def stream(self):
        return self.lib.sphg_stream(self._ctx)

    def close(self):
        if self._ctx:
            self.lib.sphg_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference procedures ------------------------------------------
    def direct_find(self, phase="water"):
        """direct_find(parts), m_particles.f90:938-1001."""
        pid = _phase_id(phase)
        self._check(self.lib.sphg_direct_find(self._ctx, pid), "direct_find")
        self._pull(pid)

    def link_list(self, phase="water"):
        """link_list(parts), m_particles.f90:1146-1265."""
        pid = _phase_id(phase)
        self._check(self.lib.sphg_link_list(self._ctx, pid), "link_list")
        self._pull(pid)

    def find_pairs(self, phase="water"):
        """find_pairs(parts), m_particles.f90:1446-1466."""
        if self.params.nnps == 1:
            self.direct_find(phase)
        elif self.params.nnps == 2:
            self.link_list(phase)
        else:
            raise abi.SphgError(abi.SPHG_ERR_UNSUPPORTED, "find_pairs: tree_search method not implemented")

    def direct_find_2(self):
        """direct_find_2(parts, soil), m_particles.f90:1004-1062."""
        self._check(self.lib.sphg_direct_find_2(self._ctx), "direct_find_2")
        self._pull("water")
        self._pull("soil")

    def con_density(self, phase="water"):
        """con_density(parts), m_particles.f90:2402-2431."""
        self._check(self.lib.sphg_con_density(self._ctx, _phase_id(phase)), "con_density")

    def sum_density(self, phase="water"):
        """sum_density(parts), m_particles.f90:2243-2298."""
        self._check(self.lib.sphg_sum_density(self._ctx, _phase_id(phase)), "sum_density")

    def delta_sph_density(self, phase="water"):
        """delta_sph(parts, parts%rho, parts%drho), m_particles.f90:2658-2695."""
        self._check(self.lib.sphg_delta_sph_density(self._ctx, _phase_id(phase)), "delta_sph")

    def internal_force(self, phase="water"):
        """Internal-force group of single_step, declarations_sph.f90:1222-1308."""
        pid = _phase_id(phase)
        self._check(self.lib.sphg_internal_force(self._ctx, pid), "internal_force")
        self._pull(pid)

    def art_visc(self, phase="water"):
        """art_visc(parts), m_particles.f90:2504-2564."""
        self._check(self.lib.sphg_art_visc(self._ctx, _phase_id(phase)), "art_visc")

    def single_step(self, sync=True):
        """single_step / single_step_for_water (declarations_sph.f90:1131, :1439)."""
        self._check(self.lib.sphg_single_step(self._ctx, 1 if sync else 0), "single_step")
        if sync:
            self._pull_bound()

    def time_integration(self, nsteps, sync=True):
        """time_integration / time_integration_for_water / _for_soil /
        _for_water_by_verlet (declarations_sph.f90:525, :794, :978, :889)."""
        self._check(self.lib.sphg_time_integration(self._ctx, int(nsteps), 1 if sync else 0),
                    "time_integration")
        if sync:
            self._pull_bound()

    def _pull_bound(self):
        if self.water is not None:
            self._pull("water")
        if self.soil is not None:
            self._pull("soil")

    # -- measurement ------------------------------------------------------
    def profile(self, enable=True):
        self._check(self.lib.sphg_profile_enable(self._ctx, 1 if enable else 0), "profile_enable")

    def profile_reset(self):
        self._check(self.lib.sphg_profile_reset(self._ctx), "profile_reset")

    def profile_get(self, name):
        n = C.c_int64()
        ms = C.c_double()
        rc = self.lib.sphg_profile_get(self._ctx, name.encode(), C.byref(n), C.byref(ms))
        if rc != abi.SPHG_OK:
            return 0, 0.0
        return n.value, ms.value

    def pair_counts(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.sphg_pair_counts(self._ctx, C.byref(a), C.byref(b), C.byref(c)), "pair_counts")
        return a.value, b.value, c.value

    def nlist_stats(self):
        """(rebuilds, off_spells, skin) of the neighbour lists of the FP64
        two-phase step."""
        n, off, skin = C.c_int64(), C.c_int64(), C.c_double()
        self._check(self.lib.sphg_nlist_stats(self._ctx, C.byref(n), C.byref(off), C.byref(skin)), "nlist_stats")
        return n.value, off.value, skin.value
