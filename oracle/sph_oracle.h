/*
 * sph_oracle.h -- CPU restatement of the reference's per-step loop.
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it as the checker.  The product library never links it.
 * It operates on the same sphg_params / sphg_phase_view structures as the
 * public C ABI (include/sphg.h), so one host-side state feeds both.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include "../include/sphg.h"
#ifdef __cplusplus
extern "C" {
#endif
int  oracle_scale_k(int skf);
void oracle_kernel(int skf, double r, const double dx[2], double hsml, double *w, double dwdx[2]);
int  oracle_direct_find(const sphg_params *P, sphg_phase_view *ph);
int  oracle_direct_find_2(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so);
void oracle_init_grid(const sphg_params *P, int ntotal, double hsml, int ngridx[2], int ghsmlx[2],
                      double mingridx[2], double dgeomx[2]);
int  oracle_grid_geom(const sphg_params *P, const int ngridx[2], const double mingridx[2],
                      const double dgeomx[2], const double x[2], int gcell[2]);
int  oracle_link_list(const sphg_params *P, sphg_phase_view *ph);
int  oracle_find_pairs(const sphg_params *P, sphg_phase_view *ph);
void oracle_con_density(const sphg_params *P, sphg_phase_view *ph);
void oracle_sum_density(const sphg_params *P, sphg_phase_view *ph);
void oracle_delta_sph(const sphg_params *P, sphg_phase_view *ph, const double *f, double *df);
void oracle_shear_strain_rate(const sphg_params *P, sphg_phase_view *ph);
void oracle_velocity_divergence(const sphg_params *P, sphg_phase_view *ph);
void oracle_pressure(const sphg_params *P, sphg_phase_view *ph);
void oracle_newtonian_fluid(const sphg_params *P, sphg_phase_view *ph);
void oracle_drucker_prager(const sphg_params *P, sphg_phase_view *ph);
void oracle_mohr_coulomb(const sphg_params *P, sphg_phase_view *ph);
void oracle_plastic_or_not(const sphg_params *P, sphg_phase_view *ph);
void oracle_plastic_flow_rule3(const sphg_params *P, sphg_phase_view *ph);
void oracle_plastic_flow_rule(const sphg_params *P, sphg_phase_view *ph);
void oracle_plastic_flow_rule2(const sphg_params *P, sphg_phase_view *ph);
void oracle_first_half(const sphg_params *P, sphg_phase_view *pl);
void oracle_first_step(const sphg_params *P, sphg_phase_view *pl);
void oracle_second_half(const sphg_params *P, sphg_phase_view *pl);
void oracle_return_mapping(const sphg_params *P, sphg_phase_view *pl);
void oracle_jaumann_rate(const sphg_params *P, sphg_phase_view *ph);
void oracle_art_visc(const sphg_params *P, sphg_phase_view *ph);
void oracle_repulsive_force(const sphg_params *P, sphg_phase_view *ph);
void oracle_av_vel(const sphg_params *P, sphg_phase_view *ph);
void oracle_df(sphg_phase_view *ph, const double *f, int d, double *out);
void oracle_df2(sphg_phase_view *ph, const double *f, int d, double *out);
void oracle_df3(sphg_phase_view *ph, const double *f, int d, double *out);
void oracle_darcy_law(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so);
void oracle_pore_water_pressure(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so);
void oracle_volume_fraction_soil(const sphg_params *P, sphg_phase_view *so);
void oracle_volume_fraction_water2(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so);
void oracle_volume_fraction_water(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so);
int  oracle_single_step(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so, int itimestep);
int  oracle_single_step_for_water(const sphg_params *P, sphg_phase_view *pl, int itimestep);
int  oracle_time_integration(const sphg_params *P, sphg_phase_view *wa, sphg_phase_view *so,
                             int *itimestep, int nsteps);
int  oracle_time_integration_for_water(const sphg_params *P, sphg_phase_view *wa, int *itimestep, int nsteps);
int  oracle_single_step_for_soil(const sphg_params *P, sphg_phase_view *pl, int itimestep);
int  oracle_time_integration_for_soil(const sphg_params *P, sphg_phase_view *so, int *itimestep, int nsteps);
void oracle_verlet_update(const sphg_params *P, sphg_phase_view *wa, double *lastvx, double *lastrho, int step);
int  oracle_time_integration_for_water_by_verlet(const sphg_params *P, sphg_phase_view *wa, int *itimestep,
                                                 int nsteps);
#ifdef __cplusplus
}
#endif
#endif
