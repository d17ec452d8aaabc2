"""Learned multi-step surrogate over a Solver (the sphg_gnn_* entry points).

A learned advance of K steps is one leapfrog step of size K dt plus the
velocity and density correction that a message-passing network on the
same-phase neighbour graph (BF16 MFMA, FP32 accumulation, hand-written
forward and backward kernels) predicts towards K steps of size dt.  It
trains on samples that the plain FP64 step generates; the plain step judges
it and takes over whenever the learned density deviates by more than the
tolerance.  No counterpart in the reference.
"""
import ctypes as C

import numpy as np

from . import abi

FN, RAW, DEG, NO = 32, 12, 48, 4


class Surrogate:
    def __init__(self, solver, ksteps=2, max_samples=64, seed=1):
        self.solver = solver
        self.lib = solver.lib
        self.ksteps = ksteps
        self._g = C.c_void_p()
        self._check(self.lib.sphg_gnn_create(solver._ctx, ksteps, max_samples, C.c_uint64(seed),
                                             C.byref(self._g)), "sphg_gnn_create")
        self.nparams = self.lib.sphg_gnn_nparams()

    def _check(self, rc, where):
        self.solver._check(rc, where)

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None

    @staticmethod
    def _ip(a):
        return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None

    def weights(self):
        w = np.zeros(self.nparams, np.float32)
        self._check(self.lib.sphg_gnn_weights(self._g, self._fp(w), 1), "sphg_gnn_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparams
        self._check(self.lib.sphg_gnn_weights(self._g, self._fp(w), 0), "sphg_gnn_weights")

    @property
    def num_samples(self):
        return self.lib.sphg_gnn_num_samples(self._g)

    def record(self, n=1):
        for _ in range(n):
            self._check(self.lib.sphg_gnn_record(self._g), "sphg_gnn_record")

    def train(self, iters, lr=1e-3):
        loss = C.c_float()
        self._check(self.lib.sphg_gnn_train(self._g, iters, C.c_float(lr), C.byref(loss)), "sphg_gnn_train")
        return loss.value

    def gains(self):
        """Fitted gains of the correction channels (vx, vy, rho, spare)."""
        g = np.zeros(4, np.float32)
        self._check(self.lib.sphg_gnn_gains(self._g, self._fp(g)), "sphg_gnn_gains")
        return g

    def step(self, use_net=True):
        self._check(self.lib.sphg_gnn_step(self._g, 1 if use_net else 0), "sphg_gnn_step")

    def rollout(self, nlearned, judge_every=4, tol=0.01, use_net=True):
        """Returns (fallbacks, max relative density error, rms relative
        density error) of the judged advances."""
        nfb = C.c_int()
        err = C.c_double()
        rms = C.c_double()
        self._check(self.lib.sphg_gnn_rollout(self._g, nlearned, judge_every, C.c_double(tol),
                                              1 if use_net else 0, C.byref(nfb), C.byref(err), C.byref(rms)),
                    "sphg_gnn_rollout")
        return nfb.value, err.value, rms.value

    def debug(self, sample):
        n = self.lib.sphg_gnn_sample_size(self._g, sample)
        if n < 0:
            raise IndexError(sample)
        out = {
            "nf": np.zeros((n, FN), np.float32), "raw": np.zeros((n, RAW), np.float32),
            "nb": np.zeros((n, DEG), np.int32), "nd": np.zeros(n, np.int32),
            "tgt": np.zeros((n, NO), np.float32), "tw": np.zeros(n, np.float32),
            "y": np.zeros((n, NO), np.float32), "grad": np.zeros(self.nparams, np.float32),
            "loss": np.zeros(1, np.float32), "out_scale": np.zeros(NO, np.float32),
        }
        self._check(self.lib.sphg_gnn_debug(
            self._g, sample, self._fp(out["nf"]), self._fp(out["raw"]), self._ip(out["nb"]), self._ip(out["nd"]),
            self._fp(out["tgt"]), self._fp(out["tw"]), self._fp(out["y"]), self._fp(out["grad"]),
            self._fp(out["loss"]), self._fp(out["out_scale"])), "sphg_gnn_debug")
        out["loss"] = float(out["loss"][0])
        return out

    def close(self):
        if self._g:
            self.lib.sphg_gnn_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
