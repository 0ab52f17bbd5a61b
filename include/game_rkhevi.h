/*
 * game_rkhevi.h -- runtime-sized C ABI of the MI355X RKHEVI dynamical-core step.
 *
 * The reference (GAME, src/game_types.h:12-258) fixes every array size at
 * compile time through `enum grid_integers`.  This header describes the same
 * fields with run-time sizes: every view struct below carries plain pointers
 * whose names, meaning, element order and lengths are those of the reference
 * struct member of the same name.  The index conventions are the reference's:
 *   scalar   (h, layer)      -> layer*NSH + h, layer 0 on top
 *   vector   w at level l    -> l*VPL + h          (VPL = NSH + NVH)
 *            u at edge e     -> NSH + layer*VPL + e
 *   curl     tangential      -> layer*2*NVH + e,   vertical (rhombus) -> layer*2*NVH + NVH + e
 *   densities                -> k*S + i  (k: snow, rain, cloud ice, cloud water, gas, vapour)
 *
 * The drop-in entry point with the reference's exact struct signature
 * (`manage_rkhevi`, src/time_stepping/time_stepping.h:6) lives in
 * include/game_abi.h; the functions here are what it forwards to.
 */
#ifndef GAME_RKHEVI_H
#define GAME_RKHEVI_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* sizes (src/game_types.h:12-60, evaluated at run time)               */
/* ------------------------------------------------------------------ */
typedef struct game_dims {
    int res_id;            /* RES_ID */
    int n_layers;          /* NO_OF_LAYERS */
    int moisture_on;       /* MOISTURE_ON (0 or 1) */
    int n_oro_layers;      /* grid->no_of_oro_layers */
    /* derived by game_dims_derive() */
    int n_const;           /* NO_OF_CONSTITUENTS (1 or 6) */
    int n_cond;            /* NO_OF_CONDENSED_CONSTITUENTS (0 or 4) */
    int n_levels;          /* NO_OF_LEVELS */
    int nsh;               /* NO_OF_SCALARS_H */
    int nvh;               /* NO_OF_VECTORS_H */
    int nt;                /* NO_OF_DUAL_SCALARS_H = NO_OF_TRIANGLES */
    int vpl;               /* NO_OF_VECTORS_PER_LAYER */
    int dvpl;              /* NO_OF_DUAL_VECTORS_PER_LAYER */
    long long s;           /* NO_OF_SCALARS */
    long long v;           /* NO_OF_VECTORS */
    long long curl;        /* length of Curl_field */
    long long dual_v;      /* NO_OF_DUAL_VECTORS */
    long long dual_v_vectors; /* NO_OF_DUAL_V_VECTORS (triangle vorticity) */
} game_dims_t;

static inline void game_dims_derive(game_dims_t *d)
{
    long long four_r = 1LL << (2 * d->res_id);
    long long hexagons = 10 * (four_r - 1);
    d->n_cond = 4 * d->moisture_on;
    d->n_const = 1 + d->moisture_on + d->n_cond;
    d->n_levels = d->n_layers + 1;
    d->nsh = (int)(12 + hexagons);
    d->nvh = (int)(30 + 3 * hexagons);
    d->nt = (int)(20 * four_r);
    d->vpl = d->nsh + d->nvh;
    d->dvpl = d->nvh + d->nt;
    d->s = (long long)d->nsh * d->n_layers;
    d->v = (long long)d->n_layers * d->nvh + (long long)d->n_levels * d->nsh;
    d->curl = (long long)d->n_layers * 2 * d->nvh + d->nvh;
    d->dual_v = (long long)d->n_levels * d->nvh + (long long)d->n_layers * d->nt;
    d->dual_v_vectors = (long long)d->n_layers * d->nt;
}

/* ------------------------------------------------------------------ */
/* views of the reference structs (pointers into caller memory)        */
/* ------------------------------------------------------------------ */
/* struct grid, src/game_types.h:70-114 (fields the step reads) */
typedef struct game_grid_view {
    int no_of_oro_layers;
    double *normal_distance;            /* V   */
    double *volume;                     /* S   */
    double *area;                       /* V   */
    double *z_scalar;                   /* S   */
    double *z_vector;                   /* V   */
    double *gravity_potential;          /* S   */
    double *gravity_m;                  /* V   */
    double *slope;                      /* V   */
    double *theta_v_bg;                 /* S   */
    double *exner_bg;                   /* S   */
    double *exner_bg_grad;              /* V   */
    double *layer_thickness;            /* S   */
    int *trsk_indices;                  /* 10*NVH */
    int *trsk_modified_curl_indices;    /* 10*NVH */
    int *from_index;                    /* NVH */
    int *to_index;                      /* NVH */
    int *adjacent_vector_indices_h;     /* 6*NSH */
    int *adjacent_signs_h;              /* 6*NSH */
    int *density_to_rhombi_indices;     /* 4*NVH */
    double *latitude_scalar;            /* NSH */
    double *longitude_scalar;           /* NSH */
    double *inner_product_weights;      /* 8*S */
    double *direction;                  /* NVH */
    double *density_to_rhombi_weights;  /* 4*NVH */
    double *trsk_weights;               /* 10*NVH */
    int *is_land;                       /* NSH */
    double toa;
    double radius;
    double stretching_parameter;
    int oro_id;
} game_grid_view_t;

/* struct dualgrid, src/game_types.h:117-126 */
typedef struct game_dualgrid_view {
    double *area;                       /* curl */
    double *z_vector;                   /* dual_v */
    double *normal_distance;            /* dual_v */
    int *from_index;                    /* NVH */
    int *to_index;                      /* NVH */
    int *vorticity_indices_triangles;   /* 3*NT */
    int *vorticity_signs_triangles;     /* 3*NT */
    double *f_vec;                      /* 2*NVH */
} game_dualgrid_view_t;

/* struct state, src/game_types.h:128-135 */
typedef struct game_state_view {
    double *rho;                        /* n_const*S */
    double *rhotheta_v;                 /* S */
    double *theta_v_pert;               /* S */
    double *exner_pert;                 /* S */
    double *wind;                       /* V */
} game_state_view_t;

/* the members of struct diagnostics (src/game_types.h:138-162) the step touches */
typedef struct game_diag_view {
    double *flux_density;               /* V */
    double *flux_density_divv;          /* S */
    double *rel_vort_on_triangles;      /* L*NT */
    double *rel_vort;                   /* curl */
    double *pot_vort;                   /* curl */
    double *temperature;                /* S */
    double *scalar_field_placeholder;   /* S */
    double *v_squared;                  /* S */
} game_diag_view_t;

/* struct forcings, src/game_types.h:180-190 */
typedef struct game_forcings_view {
    double *pgrad_acc_old;                  /* V */
    double *pressure_gradient_acc_neg_nl;   /* V */
    double *pressure_gradient_acc_neg_l;    /* V */
    double *pressure_grad_condensates_v;    /* V */
    double *v_squared_grad;                 /* V */
    double *pot_vort_tend;                  /* V */
    double *radiation_tendency;             /* S */
} game_forcings_view_t;

/* the members of struct irreversible_quantities (src/game_types.h:218-238) the step touches */
typedef struct game_irrev_view {
    double *temperature_diffusion_heating;  /* S */
    double *friction_acc;                   /* V */
    double *heating_diss;                   /* S */
    double *pressure_gradient_decel_factor; /* S */
    double *condensates_sediment_heat;      /* S */
    double *mass_diff_tendency;             /* n_const*S */
    double *phase_trans_rates;              /* (n_cond+1)*S */
    double *phase_trans_heating_rate;       /* S */
} game_irrev_view_t;

/* struct config, src/game_types.h:193-216: scalars only, so its layout does
   not depend on the grid size and the reference struct is used as is. */
typedef struct game_config {
    int totally_first_step_bool;
    int temperature_diff_h;
    int temperature_diff_v;
    int momentum_diff_h;
    int momentum_diff_v;
    int mass_diff_h;
    int mass_diff_v;
    int rad_on;
    int prog_soil_temp;
    int sfc_phase_trans;
    int sfc_sensible_heat_flux;
    int rad_update;
    int time_to_next_analysis;
    int pbl_scheme;
    int total_run_span;
    double damping_start_height_over_toa;
    double damping_coeff_max;
    double impl_thermo_weight;
    double cloud_droplets_velocity;
    double rain_velocity;
    double snow_velocity;
    double radiation_delta_t;
} game_config_t;

/* ------------------------------------------------------------------ */
/* GPU path (libgame_rkhevi.so)                                         */
/* ------------------------------------------------------------------ */
/* precision forms of the device step.
   GAME_FP64: every array and every operation in FP64, the reference's
     operation order (bit-identical to the reference's C on its inputs).
   GAME_MIXED: the work arrays of the horizontal operators (pressure gradient,
     mass flux, vorticities, potential vorticity, TRSK flux, kinetic energy
     and its gradient), the momentum tendency, temperature, phase-transition
     rates and heating and the sediment heat are stored in FP32; the
     prognostic state, the density and rhotheta_v tendencies, both vertical
     implicit solvers, the grid metrics and the integrals stay FP64, so every
     constituent's mass is conserved as in the FP64 form.  Forcings and
     diagnostics read back through the getters are widened to double. */
enum game_precision { GAME_FP64 = 0, GAME_MIXED = 1 };

/* error codes (0 = success) */
enum game_status {
    GAME_OK = 0,
    GAME_ERR_HIP = 1,        /* a HIP runtime call failed */
    GAME_ERR_ARG = 2,        /* invalid argument or unsupported configuration */
    GAME_ERR_NOT_INIT = 3,   /* game_gpu_init was not called */
    GAME_ERR_NONFINITE = 4,  /* a NaN/Inf was found in the state */
    GAME_ERR_UNSUPPORTED = 5 /* a physics switch outside the implemented path is on */
};

/* Uploads the grid and allocates the device-resident state.  Replaces the
   one-time set-up the reference does in set_grid_properties
   (src/io/set_grid_properties.c:196-253) plus the struct callocs of
   src/coordinator.c:36-46.  Only one grid is resident per process. */
int game_gpu_init(const game_dims_t *dims, const game_grid_view_t *grid,
                  const game_dualgrid_view_t *dualgrid, const game_config_t *config,
                  int precision);
int game_gpu_finalize(void);
const char *game_gpu_last_error(void);

/* which device-resident state */
enum game_state_slot { GAME_STATE_OLD = 0, GAME_STATE_NEW = 1, GAME_STATE_TENDENCY = 2 };
int game_gpu_set_state(int slot, const game_state_view_t *host);
int game_gpu_get_state(int slot, game_state_view_t *host);
/* host copies of the carried fields (forcings, diagnostics, irrev) */
int game_gpu_get_forcings(game_forcings_view_t *host);
int game_gpu_get_diagnostics(game_diag_view_t *host);
int game_gpu_get_irrev(game_irrev_view_t *host);
int game_gpu_set_config(const game_config_t *config);

/* One RKHEVI step on the device-resident states: replaces manage_rkhevi
   (src/time_stepping/manage_rkhevi.c:20-100).  Reads state_old, updates
   state_new, state_tendency, diagnostics, forcings and irrev in place.
   `totally_first_step_bool` mirrors config->totally_first_step_bool. */
int game_gpu_manage_rkhevi(double delta_t, int totally_first_step_bool);

/* state_out = c0*state_0 + c1*state_1 on the device, theta_v_pert re-derived
   (src/spatial_operators/linear_combine_two_states.c:14-43). */
int game_gpu_linear_combine_two_states(int slot_0, int slot_1, int slot_out,
                                       double coeff_0, double coeff_1);

/* The reference time loop body (src/coordinator.c:302-390 minus output):
   for n steps: copy new->old, manage_rkhevi, clear the first-step flag.
   The copy is an exchange of the two state buffers (the step's inputs are
   the same).  Every step but a first one replays a HIP graph of the step's
   launches, captured on first use for each assignment of the two buffers
   and rebuilt when delta_t or the configuration changes; with kernel timing
   on, or GAME_GPU_NO_GRAPH set at init, the steps launch eagerly (same bits). */
int game_gpu_run_steps(int n_steps, double delta_t, int first_step_bool);

/* inner_product (src/spatial_operators/inner_product.c:13-43) of two host
   vector fields (V each) into a host scalar field (S); uses the device state
   slots as staging, so it is meant for the driver's set-up calls */
int game_gpu_inner_product(const double *in_field_0, const double *in_field_1, double *out_field);

/* temperature_diagnostics (src/constituents/derived_quantities.c:15-40) of a slot */
int game_gpu_temperature_diagnostics(int slot);

/* Global integrals of a slot (src/io/write_output.c:1515-1596, 1659-1668),
   summed with compensated (Neumaier) arithmetic:
   out[0..n_const) masses, out[n_const] rhotheta_v, out[n_const+1] kinetic,
   out[n_const+2] potential, out[n_const+3] internal energy. */
int game_gpu_integrals(int slot, double *out);

/* The three records write_out_integral appends to the files "masses",
   "potential_temperature_density" and "energy" in `directory`
   (src/io/write_output.c:1515-1596; same columns and %lf format), from
   game_gpu_integrals of a slot at time_since_init. */
int game_gpu_write_integrals(int slot, double time_since_init, const char *directory);
/* n steps of the reference time loop with its integral output
   (src/coordinator.c:302-336): each step as game_gpu_run_steps, then the
   records of the new state at time t_start + (k + 1)*delta_t. */
int game_gpu_run_steps_write_integrals(int n_steps, double delta_t, int first_step_bool, double t_start,
                                       const char *directory);

/* v_squared (S) and pot_vort_tend (V) of a host (wind, density) pair: the
   kinetic-energy and vorticity-flux terms of the initial hydrostatic balance
   (src/io/set_initial_state.c:155-166).  Uses the tendency slot as scratch,
   so call it before game_gpu_set_state. */
int game_gpu_init_balance_terms(const double *wind, const double *density, double *v_squared,
                                double *pot_vort_tend);

/* Per-kernel timing over the next calls; see game_gpu_kernel_stats. */
int game_gpu_enable_timing(int on);
/* name, launches and accumulated milliseconds of kernel k (k < count). */
int game_gpu_kernel_stats(int k, const char **name, long long *launches, double *ms);
int game_gpu_kernel_count(void);
int game_gpu_reset_timing(void);
int game_gpu_synchronize(void);
/* bytes of device memory held by the library */
long long game_gpu_bytes_allocated(void);

/* ------------------------------------------------------------------ */
/* Single operators and their adjoints (FP64 form only).                */
/* All array arguments are DEVICE pointers in the layouts above (S, V,   */
/* curl, L*NT triangle or L*NVH edge fields); each call runs on the      */
/* library's stream and returns after it has finished.  The adjoints     */
/* are exact transposes of the forward operators, terrain terms          */
/* included, computed as gathers over reverse tables (no atomics).       */
/* A null buffer argument is rejected with GAME_ERR_ARG.                 */
/* The reference has no adjoints; the forward operators replace:         */
/* ------------------------------------------------------------------ */
/* grad, gradient_operators.c:56-74: S -> V (top and surface w are 0) */
int game_gpu_op_grad(const double *f, double *out);
int game_gpu_op_grad_adjoint(const double *out_bar, double *f_bar);
/* divv_h, divergences.c:14-64: V (horizontal part read) -> S */
int game_gpu_op_divv_h(const double *in, double *out);
int game_gpu_op_divv_h_adjoint(const double *out_bar, double *in_bar);
/* calc_rel_vort_on_triangles, vorticities.c:167-222: V -> L*NT */
int game_gpu_op_curl_triangles(const double *wind, double *out);
int game_gpu_op_curl_triangles_adjoint(const double *out_bar, double *wind_bar);
/* rhombus density interpolation of calc_pot_vort, vorticities.c:74-80: S -> L*NVH */
int game_gpu_op_rhombus_density(const double *rho, double *out);
int game_gpu_op_rhombus_density_adjoint(const double *out_bar, double *rho_bar);
/* vorticity_flux, vorticity_flux.c:16-187 and :373-417, linear in the mass
   flux for a fixed potential vorticity (curl layout): V -> V */
int game_gpu_op_vorticity_flux(const double *flux, const double *pot_vort, double *out);
int game_gpu_op_vorticity_flux_adjoint(const double *pot_vort, const double *out_bar, double *flux_bar);
/* inner_product, inner_product.c:13-43, linear in a for a fixed b: V x V -> S */
int game_gpu_op_inner_product(const double *a, const double *b, double *out);
int game_gpu_op_inner_product_adjoint(const double *b, const double *out_bar, double *a_bar);
/* three-band solve, column_solvers.c:641-667, of ncol independent systems of
   n rows: row j reads c[j-1] x[j-1] + d[j] x[j] + e[j] x[j+1] = r[j];
   element j of system k at j*ncol + k (c, e: n - 1 rows used);
   work: n*ncol doubles */
int game_gpu_op_thomas(int n, int ncol, const double *c, const double *d, const double *e, const double *r,
                       double *x, double *work);
/* The same solve without a caller-supplied work buffer, for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
   It calls no allocating routine: the elimination quotients go to a
   file-scope device array of 2^23 doubles (64 MiB) that each GPU the library
   runs on holds until the process ends; n*ncol above 2^23 is rejected with
   GAME_ERR_ARG.  Every element of it that a call reads, the same call wrote
   first, so nothing needs clearing; calls that overlap on one GPU share the
   array and race. */
int game_gpu_op_thomas_noalloc(int n, int ncol, const double *c, const double *d, const double *e, const double *r,
                               double *x);
/* its adjoint: r_bar = A^-T x_bar, c_bar/d_bar/e_bar = -(r_bar x^T) on the bands */
int game_gpu_op_thomas_adjoint(int n, int ncol, const double *c, const double *d, const double *e, const double *x,
                               const double *x_bar, double *r_bar, double *c_bar, double *d_bar, double *e_bar,
                               double *work);

/* ------------------------------------------------------------------ */
/* Window of steps and its gradient (FP64 form of a dry model only).     */
/* The reference has no adjoint.  Control variables: rho (S), rhotheta_v */
/* (S) and wind (V), DEVICE pointers; theta_v_pert and exner_pert are    */
/* diagnosed from them by the equation of state, exner = (R_d rhotheta_v */
/* / p_0)^(R_d/c_v) (set_initial_state.c:302-314 written for rhotheta_v). */
/* The first step takes the first-step branches (coordinator.c:299).     */
/* The runs use the model's state slots and carried fields, which hold   */
/* the window's last recomputed state afterwards.  The initial state,    */
/* the target and the gradient buffers must be given (GAME_ERR_ARG).     */
/* ------------------------------------------------------------------ */
/* n_steps steps from the control variables; outputs (device, may be null):
   rho, rhotheta_v, exner_pert and wind of the final state */
int game_gpu_window_forward(int n_steps, double dt, const double *rho0, const double *rt0, const double *wind0,
                            double *rho_out, double *rt_out, double *exner_out, double *wind_out);
/* the vector-Jacobian product of the window map (rho, rhotheta_v, wind)_0 ->
   (rho, rhotheta_v, exner_pert, wind)_N: given cotangents of the final
   state (device, null for zero), the cotangents of the control variables.
   Forward sweep with checkpoints every checkpoint_every steps (state and
   carried forcings), then per segment a recomputation that copies every
   step's inputs and substep results, reversed step by step with the
   hand-written adjoint kernels. */
int game_gpu_window_vjp(int n_steps, double dt, int checkpoint_every, const double *rho0, const double *rt0,
                        const double *wind0, const double *rho_bar_n, const double *rt_bar_n,
                        const double *exner_bar_n, const double *wind_bar_n, double *rho_bar, double *rt_bar,
                        double *wind_bar);
/* J = 1/2 sum (rho_N - rho_target)^2 (weights 1) into *cost (host) and its
   gradient with respect to the control variables (device) */
int game_gpu_window_gradient(int n_steps, double dt, int checkpoint_every, const double *rho0, const double *rt0,
                             const double *wind0, const double *rho_target, double *cost, double *rho_bar,
                             double *rt_bar, double *wind_bar);
/* device time (ms, from HIP events on the library's stream) of the phases of
   the last game_gpu_window_vjp/_gradient: [0] the forward sweep with its
   checkpoints and the taped last segment, [1] the recomputation of the other
   segments with their tapes, [2] the reverse of every step from its tape,
   [3] the equation-of-state cotangent and the copies out */
int game_gpu_window_last_timing(double *ms);

#ifdef __cplusplus
}
#endif
#endif /* GAME_RKHEVI_H */
