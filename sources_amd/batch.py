"""Batched scenarios (the sphg_batch_* entry points): many independent
two-phase solvers advanced together by replays of one captured hipGraph per
period (in the staged form, one launch per kernel for all solvers in every
step), bit-identical to stepping each solver alone."""
import ctypes as C

from . import abi


class Batch:
    def __init__(self, solvers):
        self.solvers = list(solvers)
        self.lib = self.solvers[0].lib
        arr = (C.c_void_p * len(self.solvers))(*[s._ctx.value for s in self.solvers])
        self._b = C.c_void_p()
        self._check(self.lib.sphg_batch_create(arr, len(self.solvers), C.byref(self._b)), "sphg_batch_create")

    def _check(self, rc, where):
        if rc != abi.SPHG_OK:
            raise abi.SphgError(rc, where)

    def run(self, nsteps):
        self._check(self.lib.sphg_batch_run(self._b, nsteps), "sphg_batch_run")

    def set_streams(self, n):
        self._check(self.lib.sphg_batch_set_streams(self._b, n), "sphg_batch_set_streams")

    def tune(self, trial_periods=2):
        """Autotune the launch mode; returns (streams, ms per step), streams = 0
        meaning eager launches on the solvers' own streams."""
        ns = C.c_int()
        ms = C.c_double()
        self._check(self.lib.sphg_batch_tune(self._b, trial_periods, C.byref(ns), C.byref(ms)), "sphg_batch_tune")
        return ns.value, ms.value

    @property
    def staged(self):
        """True when the captured steps are one launch per kernel for all
        solvers."""
        return self.lib.sphg_batch_staged(self._b) == 1

    def close(self):
        if self._b:
            self.lib.sphg_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
