"""Binary state snapshot: the fixture, restart and rollout format.

Layout (little endian, no padding between records):
  int32  magic 0x53504847 ('SPHG'), version 1, nphases, itimestep
  int32  sizeof(sphg_params); the raw sphg_params bytes (include/sphg.h layout)
  per phase:
    int32  material, ntotal, nvirt
    float64 dspp
    int32  itype[n]
    float64 x[2n], vx[2n]                 (Fortran x(2,n) order)
    float64 mass, hsml, rho, p, c, vof, sxx, sxy, syy, epsilon_p, drho  [n each]
    float64 dvx[2n]
    float64 dvof, dsxx, dsxy, dsyy, dp    [n each]
The Fortran host (sources_amd/fortran/sph_main.f90) reads and writes the same
records with stream access.
"""
import ctypes as C

import numpy as np

from . import abi
from .state import Phase

MAGIC = 0x53504847
VERSION = 1
SCALARS_A = ("mass", "hsml", "rho", "p", "c", "vof", "sxx", "sxy", "syy", "epsilon_p", "drho")
SCALARS_B = ("dvof", "dsxx", "dsxy", "dsyy", "dp")


def write(path, params, phases, itimestep=0):
    with open(path, "wb") as f:
        np.array([MAGIC, VERSION, len(phases), itimestep], np.int32).tofile(f)
        np.array([C.sizeof(abi.Params)], np.int32).tofile(f)
        f.write(bytes(params))
        for ph in phases:
            n = ph.n
            np.array([ph.material, ph.ntotal, ph.nvirt], np.int32).tofile(f)
            np.array([ph.dspp], np.float64).tofile(f)
            ph.itype[:n].astype(np.int32).tofile(f)
            ph.x[:n].astype(np.float64).tofile(f)
            ph.vx[:n].astype(np.float64).tofile(f)
            for name in SCALARS_A:
                getattr(ph, name)[:n].astype(np.float64).tofile(f)
            ph.dvx[:n].astype(np.float64).tofile(f)
            for name in SCALARS_B:
                getattr(ph, name)[:n].astype(np.float64).tofile(f)


def read(path):
    """Returns (params, [phases], itimestep)."""
    with open(path, "rb") as f:
        magic, version, nph, it = np.fromfile(f, np.int32, 4)
        if magic != MAGIC or version != VERSION:
            raise ValueError(f"{path}: not a version-{VERSION} state snapshot")
        (psize,) = np.fromfile(f, np.int32, 1)
        if psize != C.sizeof(abi.Params):
            raise ValueError(f"{path}: params size {psize} != {C.sizeof(abi.Params)}")
        P = abi.Params.from_buffer_copy(f.read(psize))
        phases = []
        for _ in range(nph):
            mat, nt, nv = np.fromfile(f, np.int32, 3)
            (dspp,) = np.fromfile(f, np.float64, 1)
            n = int(nt + nv)
            ph = Phase(int(mat), int(nt), int(nv), maxn=n, dspp=float(dspp))
            ph.itype[:] = np.fromfile(f, np.int32, n)
            ph.x[:] = np.fromfile(f, np.float64, 2 * n).reshape(n, 2)
            ph.vx[:] = np.fromfile(f, np.float64, 2 * n).reshape(n, 2)
            for name in SCALARS_A:
                getattr(ph, name)[:] = np.fromfile(f, np.float64, n)
            ph.dvx[:] = np.fromfile(f, np.float64, 2 * n).reshape(n, 2)
            for name in SCALARS_B:
                getattr(ph, name)[:] = np.fromfile(f, np.float64, n)
            phases.append(ph)
    return P, phases, int(it)
