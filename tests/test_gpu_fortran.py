"""The Fortran host (sources_amd/fortran/sph_main.f90, the structure of the
reference's program SPH) drives the GPU path through the ISO_C_BINDING
module: its result after a few steps matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import assert_fields
from sources_amd import scenarios, snapshot

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sources_amd", "fortran", "sph_gpu")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (flang absent at build time)")
@pytest.mark.parametrize("nsteps", [1, 3])
def test_fortran_host_matches_oracle(tmp_path, nsteps):
    sc = scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                    jitter=0.05, seed=7)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    snapshot.write(fin, sc.params, [sc.water, sc.soil])
    print(f"sph_gpu {nsteps} steps", flush=True)
    r = subprocess.run([EXE, fin, fout, str(nsteps)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    _, (w, s), it = snapshot.read(fout)
    assert it == nsteps
    ref = sc.copy()
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, nsteps)
    fields_w = ("x", "vx", "rho", "p", "c", "vof", "drho", "dvx", "dvof")
    fields_s = ("x", "vx", "rho", "p", "c", "sxx", "sxy", "syy", "epsilon_p", "drho", "dvx", "dsxx", "dp")
    assert_fields(w, ref.water, 1e-12 if nsteps == 1 else 1e-11, fields_w, "fortran water")
    assert_fields(s, ref.soil, 1e-12 if nsteps == 1 else 1e-11, fields_s, "fortran soil")
