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


EXE_CAS = os.path.join(ROOT, "sources_amd", "fortran", "sph_cas")


def tecplot_zones(path):
    """Zones of a Tecplot BLOCK file written by `output`: a list of
    {variable: array}."""
    import re
    text = open(path).read()
    zones = []
    for chunk in text.split("VARIABLES=")[1:]:
        head, body = chunk.split("F=BLOCK", 1)
        names = re.findall(r'"([^"]+)"', head)
        n = int(re.search(r"I=\s*(\d+)", head).group(1))
        vals = np.array([float(t) for t in body.split()])
        assert vals.size == n * len(names), (vals.size, n, names)
        zones.append({nm: vals[i * n:(i + 1) * n] for i, nm in enumerate(names)})
    return zones


@pytest.mark.skipif(not os.path.exists(EXE_CAS), reason="Fortran host not built (flang absent at build time)")
def test_steering_driver_writes_reference_tecplot_zones(tmp_path):
    sc = scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                    jitter=0.05, seed=7)
    fin = str(tmp_path / "in.bin")
    snapshot.write(fin, sc.params, [sc.water, sc.soil])
    cas = tmp_path / "run.cas"
    # list-directed input ends an unquoted string at '/': paths are quoted
    cas.write_text(f"Print step = 1\nSave step = 2\n"
                   f"Results file = '{tmp_path}/xv.dat', '{tmp_path}/state.dat', '{tmp_path}/other.dat'\n")
    print("sph_cas 3 steps", flush=True)
    r = subprocess.run([EXE_CAS, str(cas), fin], input="3\n0\n", capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-500:])
    assert r.returncode == 0
    assert "current number of time step =" in r.stdout and "Minimum time step is" in r.stdout
    xv = tecplot_zones(str(tmp_path / "xv.dat"))
    st = tecplot_zones(str(tmp_path / "state.dat"))
    assert len(xv) == 2 and len(st) == 2          # initial state and step 2
    ref = sc.copy()
    oracle.time_integration(ref.params, ref.water, ref.soil, 0, 2)
    w, s = ref.water, ref.soil
    for name, want in (("X", w.x[:, 0]), ("Y", w.x[:, 1]), ("Pressure", w.p), ("rho", w.rho),
                       ("vx", w.vx[:, 0]), ("vy", w.vx[:, 1]), ("VoF", w.vof), ("mass", w.mass)):
        got = xv[1][name]
        assert np.max(np.abs(got - want[:w.n])) <= 1e-12 * max(np.max(np.abs(want[:w.n])), 1e-300), name
    for name, want in (("X", s.x[:, 0]), ("Pressure", s.p), ("sxx", s.sxx), ("sxy", s.sxy), ("syy", s.syy),
                       ("rho", s.rho), ("ep", s.epsilon_p)):
        got = st[1][name]
        assert np.max(np.abs(got - want[:s.n])) <= 1e-12 * max(np.max(np.abs(want[:s.n])), 1e-300), name
    np.testing.assert_allclose(st[1]["sigma_yy"], -s.p[:s.n] + s.syy[:s.n], rtol=0,
                               atol=1e-12 * np.max(np.abs(s.syy[:s.n])))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (flang absent at build time)")
def test_fortran_density_procedures_match_oracle(tmp_path):
    """The granular wrappers con_density, delta_sph_density and sum_density
    of module sphg, called by the Fortran host at the reference's call sites
    (declarations_sph.f90:1195-1207, :1465), against the oracle's
    con_density, delta_sph(rho, drho) and sum_density on the same state."""
    sc = scenarios.submerged_column(dx=0.005, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                    jitter=0.05, seed=7)
    # a state with motion: two oracle steps from the column
    oracle.time_integration(sc.params, sc.water, sc.soil, 0, 2)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    snapshot.write(fin, sc.params, [sc.water, sc.soil], itimestep=2)
    print("sph_gpu density", flush=True)
    r = subprocess.run([EXE, fin, fout, "density"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    _, (w, s), _ = snapshot.read(fout)
    ref = sc.copy()
    for ph in (ref.water, ref.soil):
        oracle.call_phase("find_pairs", ref.params, ph)
        oracle.call_phase("con_density", ref.params, ph)
        oracle.delta_sph(ref.params, ph, ph.rho, ph.drho)
    oracle.call_phase("sum_density", ref.params, ref.water)
    assert_fields(w, ref.water, 1e-12, ("drho", "rho"), "fortran water density")
    assert_fields(s, ref.soil, 1e-12, ("drho", "rho"), "fortran soil density")
