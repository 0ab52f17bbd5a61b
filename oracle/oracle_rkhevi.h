/*
 * oracle_rkhevi.h -- CPU oracle of the RKHEVI step (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, as the checker.  The product path never links it.
 * Every function restates one reference function with run-time sizes and the
 * reference's operation order; the names carry an `or_` prefix so that the
 * oracle and the reference's own objects (oracle/_ref) can share a process.
 */
#ifndef ORACLE_RKHEVI_H
#define ORACLE_RKHEVI_H

#include "../include/game_rkhevi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx {
    game_dims_t d;
    game_grid_view_t g;
    game_dualgrid_view_t dg;
} or_ctx_t;

/* spatial operators (src/spatial_operators/) */
void or_grad_hor_cov(const or_ctx_t *c, const double *in, double *out);
void or_grad_vert_cov(const or_ctx_t *c, const double *in, double *out);
void or_vector_field_hor_cov_to_con(const or_ctx_t *c, double *field);
void or_grad(const or_ctx_t *c, const double *in, double *out);
double or_remap_verpri2horpri_vector(const or_ctx_t *c, const double *field, int layer, int h);
double or_vertical_contravariant_corr(const or_ctx_t *c, const double *field, int layer, int h);
double or_horizontal_covariant(const or_ctx_t *c, const double *field, int layer, int h);
void or_divv_h(const or_ctx_t *c, const double *in, double *out);
void or_divv_h_tracer(const or_ctx_t *c, const double *in, const double *density, const double *wind, double *out);
void or_scalar_times_vector_h(const or_ctx_t *c, const double *s, const double *v, double *out);
void or_scalar_times_vector_h_upstream(const or_ctx_t *c, const double *s, const double *v, double *out);
void or_scalar_times_vector_v(const or_ctx_t *c, const double *s, const double *v, double *out);
void or_scalar_times_vector(const or_ctx_t *c, const double *s, const double *v, double *out);
void or_inner_product(const or_ctx_t *c, const double *a, const double *b, double *out);
void or_calc_rel_vort_on_triangles(const or_ctx_t *c, const double *wind, double *out);
void or_calc_rel_vort(const or_ctx_t *c, const double *wind, game_diag_view_t *diag);
void or_add_f_to_rel_vort(const or_ctx_t *c, const double *rel_vort, double *out);
void or_calc_pot_vort(const or_ctx_t *c, const double *wind, const double *density, game_diag_view_t *diag);
/* the density calc_pot_vort divides the rhombus PV of edge h in layer by
   (vorticities.c:33-43) */
double or_density_to_rhombus(const or_ctx_t *c, const double *density, int layer, int h);
void or_vorticity_flux(const or_ctx_t *c, const double *flux, const double *pv, double *out);
void or_linear_combine_two_states(const or_ctx_t *c, const game_state_view_t *s0, const game_state_view_t *s1,
                                  game_state_view_t *out, double c0, double c1);

/* constituents (src/constituents/) */
void or_temperature_diagnostics(const or_ctx_t *c, const game_state_view_t *s, double *temperature);
double or_c_p_cond(int const_id, double temperature);
double or_phase_trans_heat(int direction, double temperature);
double or_saturation_pressure_over_water(double t);
double or_saturation_pressure_over_ice(double t);
void or_calc_h2otracers_source_rates(const or_ctx_t *c, const game_state_view_t *s, const game_diag_view_t *diag,
                                     const game_config_t *cfg, game_irrev_view_t *irrev, double dt);

/* time stepping (src/time_stepping/) */
void or_thomas(const double *c, const double *d, const double *e, const double *r, double *x, int n);
int or_manage_rkhevi(const or_ctx_t *c, game_state_view_t *s_old, game_state_view_t *s_new,
                     game_state_view_t *tend, game_diag_view_t *diag, game_forcings_view_t *f,
                     game_irrev_view_t *irrev, game_config_t *cfg, double dt);
void or_manage_pressure_gradient(const or_ctx_t *c, const game_state_view_t *s, game_diag_view_t *diag,
                                 game_forcings_view_t *f, game_irrev_view_t *irrev, const game_config_t *cfg);
void or_calc_pressure_grad_condensates_v(const or_ctx_t *c, const game_state_view_t *s, game_forcings_view_t *f,
                                         game_irrev_view_t *irrev);
void or_vector_tendencies_expl(const or_ctx_t *c, const game_state_view_t *s, game_state_view_t *tend,
                               game_diag_view_t *diag, game_forcings_view_t *f, game_irrev_view_t *irrev,
                               const game_config_t *cfg, int rk_step);
void or_scalar_tendencies_expl(const or_ctx_t *c, const game_state_view_t *s_old, const game_state_view_t *s,
                               game_state_view_t *tend, game_diag_view_t *diag, game_forcings_view_t *f,
                               game_irrev_view_t *irrev, const game_config_t *cfg, int rk_step);
void or_three_band_solver_ver_waves(const or_ctx_t *c, const game_state_view_t *s_old, game_state_view_t *s_new,
                                    const game_state_view_t *tend, const game_config_t *cfg, double dt, int rk_step);
void or_three_band_solver_gen_densities(const or_ctx_t *c, const game_state_view_t *s_old, game_state_view_t *s_new,
                                        const game_state_view_t *tend, const game_diag_view_t *diag,
                                        game_irrev_view_t *irrev, const game_config_t *cfg, double dt, int rk_step);

/* global integrals (src/io/write_output.c:1515-1596, 1659-1668).
   naive = the reference's serial left-to-right sum; otherwise Neumaier. */
double or_global_scalar_integrator(const or_ctx_t *c, const double *field, int naive);
void or_integrals(const or_ctx_t *c, const game_state_view_t *s, const double *temperature, double *out, int naive);

/* the time loop of the driver (src/coordinator.c:292-324), for timing */
int or_run_steps(const or_ctx_t *c, game_state_view_t *s_old, game_state_view_t *s_new, game_state_view_t *tend,
                 game_diag_view_t *diag, game_forcings_view_t *f, game_irrev_view_t *irrev,
                 game_config_t *cfg, double dt, int n_steps);

#ifdef __cplusplus
}
#endif
#endif
