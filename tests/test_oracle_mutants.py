"""The pins of the oracle can fail.  Each case below plants one plausible
mistake in a copy of oracle/sph_oracle.c (a flipped sign, a wrong exponent
or denominator, a wider tolerance, a force on the wrong particle), builds it
with the oracle's own flags, and runs the pin meant to catch that mistake
against the mutant: the pin must fail.  The planted mistakes live only in
the temporary copies."""
import os
import subprocess

import pytest

import oracle
import test_oracle_physics as phys
import test_oracle_pins as pins
import test_oracle_soil_only as soil

ORACLE_DIR = os.path.dirname(os.path.abspath(oracle.__file__))
SRC = os.path.join(ORACLE_DIR, "sph_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c99", "-shared"]

# name: (exact text in sph_oracle.c, its replacement, pins that must fail)
MUTANTS = {
    "art_visc_sign": (
        "piv = (beta * muv - alpha * mc) * muv / mrho;",
        "piv = -((beta * muv - alpha * mc) * muv / mrho);",
        [pins.test_art_visc_dissipates_kinetic_energy, pins.test_art_visc_two_particle_closed_form]),
    "eos_c_exponent_gamma": (
        "pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), 3.0)",
        "pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), P->water_gamma)",
        [pins.test_pressure_sound_speed_is_the_eos_slope]),
    "eos_p_exponent_3": (
        "pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), P->water_gamma) - 1.0",
        "pow(ph->rho[i] / (P->water_rho0 * ph->vof[i]), 3.0) - 1.0",
        [pins.test_pressure_sound_speed_is_the_eos_slope]),
    "sum_density_self_rho0": (
        "wi[i] = selfdens * ph->mass[i] / ph->rho[i];",
        "wi[i] = selfdens * ph->mass[i] / P->water_rho0;",
        [pins.test_sum_density_reproduces_a_constant_density, pins.test_sum_density_scales_with_the_old_density]),
    "sum_density_partner_rho0": (
        "if (I_END) wi[i] = wi[i] + ph->mass[j] / ph->rho[j] * ph->w[k];",
        "if (I_END) wi[i] = wi[i] + ph->mass[j] / P->water_rho0 * ph->w[k];",
        [pins.test_sum_density_reproduces_a_constant_density]),
    "repulsion_on_virtual": (
        "if (ph->itype[i] < 0) ii = j;",
        "if (ph->itype[i] > 0) ii = j;",
        [pins.test_repulsive_force_real_virtual_pair]),
    "darcy_sign": (
        "if (I_END) wa->dvx[2 * i + d] = wa->dvx[2 * i + d] - so->mass[j] * sp;\n"
        "            if (J_END) so->dvx[2 * j + d] = so->dvx[2 * j + d] + wa->mass[i] * sp;",
        "if (I_END) wa->dvx[2 * i + d] = wa->dvx[2 * i + d] + so->mass[j] * sp;\n"
        "            if (J_END) so->dvx[2 * j + d] = so->dvx[2 * j + d] - wa->mass[i] * sp;",
        [pins.test_darcy_drag_two_particle_closed_form]),
    "xsph_sign": (
        "if (I_END) ph->av[2 * i + e] = ph->av[2 * i + e] - ph->mass[j] * dvx / mrho * ph->w[k];\n"
        "            if (J_END) ph->av[2 * j + e] = ph->av[2 * j + e] + ph->mass[i] * dvx / mrho * ph->w[k];",
        "if (I_END) ph->av[2 * i + e] = ph->av[2 * i + e] + ph->mass[j] * dvx / mrho * ph->w[k];\n"
        "            if (J_END) ph->av[2 * j + e] = ph->av[2 * j + e] - ph->mass[i] * dvx / mrho * ph->w[k];",
        [pins.test_xsph_two_particle_closed_form]),
    "plastic_or_not_tolerance": (
        "if (fabs(J2 - alpha1 * I1) < EPS_YIELD_F) { k++; ph->fail[i] = 1; }",
        "if (fabs(J2 - alpha1 * I1) < 1.0e-3) { k++; ph->fail[i] = 1; }",
        [pins.test_plastic_or_not_tolerance_band]),
    "soil_only_con_density": (
        "if (P->summation_density) oracle_sum_density(P, pl);\n    else oracle_con_density(P, pl);",
        "oracle_con_density(P, pl);",
        [soil.test_soil_only_step_uses_summation_density]),
    "renorm_skipped": (
        "if (P->volume_fraction_renorm && itimestep % 40 == 0) {",
        "if (P->volume_fraction_renorm && itimestep % 40 == 0 && 0) {",
        [phys.test_vof_renormalisation_keeps_the_intrinsic_density]),
}


@pytest.fixture(scope="module")
def source():
    with open(SRC) as f:
        return f.read()


@pytest.mark.parametrize("name", list(MUTANTS))
def test_pin_catches_mutant(name, source, tmp_path):
    old, new, checks = MUTANTS[name]
    assert source.count(old) == 1, f"mutation site of {name} not found exactly once"
    src = tmp_path / "sph_oracle_mut.c"
    src.write_text(source.replace(old, new))
    lib = tmp_path / f"liboracle_{name}.so"
    subprocess.run(["gcc", *CFLAGS, "-I", ORACLE_DIR, str(src), "-o", str(lib), "-lm"], check=True)
    with oracle.using(str(lib)):
        for check in checks:
            with pytest.raises(AssertionError):
                check()
    # and the unmutated oracle passes the same pins
    for check in checks:
        check()
