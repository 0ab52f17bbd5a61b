/*
 * init_state.h -- analytic balanced initial states (set_ideal_init restated,
 * src/io/set_initial_state.c:23-242).
 */
#ifndef GAME_INIT_STATE_H
#define GAME_INIT_STATE_H

#include <stdint.h>
#include "../../include/game_rkhevi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct game_init_params {
    double t_sfc;          /* surface temperature (K) */
    double lapse_rate;     /* dT/dz below the tropopause (K/m) */
    double z_tropopause;   /* isothermal above (m) */
    double p_sfc;          /* surface pressure (Pa) */
    double u_max;          /* jet maximum (m/s): u = u_max cos^2(lat) sin(pi min(z, z_jet)/z_jet) */
    double z_jet_top;
    double sigma_t;        /* std of the seeded temperature noise (K) */
    uint64_t seed;
    int moist;             /* water vapour below z_moist_top */
    double rh_fraction;    /* q = rh_fraction*q_sat*(1 + q_noise*U(-1, 1)) */
    double q_noise;
    double z_moist_top;
} game_init_params_t;

typedef struct game_init_work {
    double *temperature;        /* S */
    double *temperature_v;      /* S */
    double *pressure;           /* S */
    double *specific_humidity;  /* S */
    double *density;            /* S: moist-air density (before, then after the balance) */
} game_init_work_t;

void game_init_default_params(game_init_params_t *p);
/* T, p, q, the unbalanced density and the wind.  The caller then evaluates
   v_squared = inner_product(wind, wind) and pot_vort_tend =
   vorticity_flux(density*wind, pv(wind, density)) (set_initial_state.c:155-166). */
int game_init_phase1(const game_init_params_t *p, const game_dims_t *d, const game_grid_view_t *g,
                     const double *latitude_vector, game_state_view_t *state, game_init_work_t *w);
/* the hydrostatic balance and the prognostic thermodynamic fields */
int game_init_phase2(const game_init_params_t *p, const game_dims_t *d, const game_grid_view_t *g,
                     const double *v_squared, const double *pot_vort_tend,
                     game_state_view_t *state, game_init_work_t *w);
/* The re-derivation of a state read from a file (read_init_data,
   src/io/set_initial_state.c:244-336): given the densities (n_const x S) and
   the temperature (S), the water vapour is limited to saturation in place
   (rel_humidity <= 1, :281-293), then rhotheta_v, theta_v_pert and
   exner_pert follow from the equation of state (:296-314).  rho and wind of
   `state` are not touched; returns 1 if a density is negative (the
   reference aborts there, :318-327). */
int game_state_from_temperature(const game_dims_t *d, const game_grid_view_t *g, double *rho,
                                const double *temperature, game_state_view_t *state);
/* rel_humidity (src/constituents/derived_quantities.c:63-82) */
double game_rel_humidity(double abs_humidity, double temperature);
double game_gauss(uint64_t seed, long long cell, int layer);
double game_uniform01(uint64_t seed, long long cell, int layer, int stream);

#ifdef __cplusplus
}
#endif
#endif
