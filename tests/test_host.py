"""Host-side logic: scenario generators, the RNG, the snapshot format and the
Fortran host build (no GPU calls)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from sources_amd import abi, scenarios, snapshot
from sources_amd.state import default_params, f32


def test_splitmix64_reference_vector():
    # SplitMix64 from state 0: first outputs of the published generator
    out = scenarios.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_reference_defaults_are_fp32_literals():
    P = default_params()
    assert P.gravity == float(np.float32(-9.8)) == -9.800000190734863
    assert P.epsilon == float(np.float32(0.001))
    assert P.skf == 4 and P.nnps == 1 and P.plasticity == 3 and P.yield_criterion == 2


def test_savage_column_matches_reference_input():
    sc = scenarios.savage_column()
    # input_mix_Savage.f90: 52x26 lattice, dx = 0.26/52, h = dx*1.2 (FP32 literal)
    assert sc.water.ntotal + sc.water.nvirt + 0 <= 52 * 26
    assert sc.soil.ntotal == 120 and sc.soil.nvirt == sc.water.nvirt
    assert np.all(sc.water.hsml == 0.26 / 52 * f32(1.2))
    assert np.all(sc.water.itype[:sc.water.ntotal] == 2) and np.all(sc.water.itype[sc.water.ntotal:] == -2)
    assert np.all(sc.soil.itype[:sc.soil.ntotal] == 3) and np.all(sc.soil.itype[sc.soil.ntotal:] == -3)
    # K0 stresses and mass = vol * rho
    p0 = sc.soil.p / f32(0.619)
    np.testing.assert_allclose(sc.soil.sxx, f32(0.1905) * p0, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sc.water.mass, (0.26 / 52) * (0.13 / 26) * sc.water.rho, rtol=1e-15)


def test_slope_full_size_counts():
    sc = scenarios.slope(dx=0.001, max_interaction=0)
    assert 0.95e6 < sc.n_particles < 1.05e6
    assert sc.soil.ntotal > sc.water.ntotal > 0


def test_snapshot_round_trip(tmp_path):
    sc = scenarios.submerged_column(dx=0.01, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12)
    path = tmp_path / "s.bin"
    snapshot.write(str(path), sc.params, [sc.water, sc.soil], itimestep=7)
    P, (w, s), it = snapshot.read(str(path))
    assert it == 7 and bytes(P) == bytes(sc.params)
    for a, b in ((w, sc.water), (s, sc.soil)):
        for f in ("x", "vx", "rho", "p", "sxx", "mass", "itype"):
            assert np.array_equal(getattr(a, f)[:a.n], getattr(b, f)[:b.n])


@pytest.mark.skipif(not (shutil.which("flang") or os.path.exists("/opt/rocm/lib/llvm/bin/flang")),
                    reason="flang not available")
def test_fortran_host_builds_and_links():
    from sources_amd import fortran_build
    exe = fortran_build.build()
    assert os.path.exists(exe)


CAS = """# a steering file in the reference's syntax (declarations_sph.f90:61-335)
Print step = 10
SAVE STEP = 20
particle pairs find method = 2
Results file = xv.dat, state.dat,
   other.dat
Single phase = .false.
First phase = water
Second phase = soil
Water property = 1000.d0, 1.4d5, 7.0, 31.3, 1.0d-3
Soil property = 2650., 1.5d6, 1.8d6, 0.3, 40.
Darcy parameter = 0.4, 1.0d-3
Soil failure criteria = 0.0, 0.5235987755982988
Time step size = 2.0d-5
&
Print step = 99
"""


@pytest.mark.skipif(not (shutil.which("flang") or os.path.exists("/opt/rocm/lib/llvm/bin/flang")),
                    reason="flang not available")
def test_steering_file_is_parsed_like_the_reference(tmp_path):
    from sources_amd import fortran_build
    fortran_build.build()
    f = tmp_path / "run.cas"
    f.write_text(CAS)
    r = subprocess.run([fortran_build.EXE_CAS, str(f), "--parse-only"], capture_output=True, text=True)
    kv = dict(line.split(" = ", 1) for line in r.stdout.splitlines() if " = " in line)
    assert kv["print_step"] == "10"          # the record after '&' is not read
    assert kv["save_step"] == "20"
    assert kv["nnps"] == "2"
    assert (kv["results"], kv["res_soil"], kv["res_other"]) == ("xv.dat", "state.dat", "other.dat")
    assert kv["single_phase"] == "F"
    assert [float(v) for v in kv["water"].split()] == [1000.0, 1.4e5, 7.0, 31.3, 1.0e-3]
    assert [float(v) for v in kv["soil"].split()] == [2650.0, 1.5e6, 1.8e6, 0.3, 40.0]
    assert [float(v) for v in kv["darcy"].split()] == [0.4, 1.0e-3]
    assert [float(v) for v in kv["failure"].split()] == [0.0, 0.5235987755982988]
    assert float(kv["dt"]) == 2.0e-5


@pytest.mark.skipif(not (shutil.which("flang") or os.path.exists("/opt/rocm/lib/llvm/bin/flang")),
                    reason="flang not available")
def test_unknown_steering_keyword_stops(tmp_path):
    from sources_amd import fortran_build
    fortran_build.build()
    f = tmp_path / "bad.cas"
    f.write_text("Print step = 10\nNo such thing = 3\n")
    r = subprocess.run([fortran_build.EXE_CAS, str(f), "--parse-only"], capture_output=True, text=True)
    assert "No such keyword" in r.stdout
    assert "print_step" not in r.stdout
