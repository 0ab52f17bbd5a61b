/*
 * game_abi.h -- the reference's struct layout and its drop-in entry points.
 *
 * GAME's driver passes fixed-size structs whose layout depends on the
 * compile-time sizes RES_ID, NO_OF_LAYERS and MOISTURE_ON (src/game_types.h:12-60,
 * 62-258).  A drop-in library must use the very same layout, so this header
 * declares layout-identical structs (same members, order and types) with the
 * sizes selectable by -DGAME_RES_ID=.. -DGAME_NO_OF_LAYERS=.. -DGAME_MOISTURE_ON=..
 * (defaults: the reference's 5, 26, 1).  Each compiled layout becomes one
 * library, lib/libgame_dropin_r{R}l{L}m{M}.so; game_abi_fingerprint() lets a
 * driver check that it was compiled with the same sizes.
 *
 * The entry points below are exactly the reference's (same names, argument
 * order and meaning); they forward to the device step of include/game_rkhevi.h.
 */
#ifndef GAME_ABI_H
#define GAME_ABI_H

#ifndef GAME_RES_ID
#define GAME_RES_ID 5
#endif
#ifndef GAME_NO_OF_LAYERS
#define GAME_NO_OF_LAYERS 26
#endif
#ifndef GAME_MOISTURE_ON
#define GAME_MOISTURE_ON 1
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* src/game_types.h:12-60 */
#define GA_NO_OF_SOIL_LAYERS 5
#define GA_NO_OF_RAD_BLOCKS 18
#define GA_NO_OF_CONDENSED (4*GAME_MOISTURE_ON)
#define GA_NO_OF_CONSTITUENTS (1 + GAME_MOISTURE_ON + GA_NO_OF_CONDENSED)
#define GA_NO_OF_HEXAGONS (10*((1 << (2*GAME_RES_ID)) - 1))
#define GA_NO_OF_LEVELS (GAME_NO_OF_LAYERS + 1)
#define GA_NO_OF_SCALARS_H (12 + GA_NO_OF_HEXAGONS)
#define GA_NO_OF_VECTORS_H (30 + 3*GA_NO_OF_HEXAGONS)
#define GA_NO_OF_H_VECTORS (GAME_NO_OF_LAYERS*GA_NO_OF_VECTORS_H)
#define GA_NO_OF_V_VECTORS (GA_NO_OF_LEVELS*GA_NO_OF_SCALARS_H)
#define GA_NO_OF_VECTORS_PER_LAYER (GA_NO_OF_VECTORS_H + GA_NO_OF_SCALARS_H)
#define GA_NO_OF_TRIANGLES (20*(1 << (2*GAME_RES_ID)))
#define GA_NO_OF_SCALARS (GA_NO_OF_SCALARS_H*GAME_NO_OF_LAYERS)
#define GA_NO_OF_VECTORS (GA_NO_OF_H_VECTORS + GA_NO_OF_V_VECTORS)
#define GA_NO_OF_SCALARS_RAD (GA_NO_OF_SCALARS/GA_NO_OF_RAD_BLOCKS)
#define GA_NO_OF_SCALARS_RAD_PER_LAYER (GA_NO_OF_SCALARS_RAD/GAME_NO_OF_LAYERS)
#define GA_NO_OF_DUAL_SCALARS_H GA_NO_OF_TRIANGLES
#define GA_NO_OF_DUAL_H_VECTORS (GA_NO_OF_LEVELS*GA_NO_OF_VECTORS_H)
#define GA_NO_OF_DUAL_V_VECTORS (GAME_NO_OF_LAYERS*GA_NO_OF_DUAL_SCALARS_H)
#define GA_NO_OF_DUAL_VECTORS (GA_NO_OF_DUAL_H_VECTORS + GA_NO_OF_DUAL_V_VECTORS)
#define GA_NO_OF_LATLON_IO_POINTS ((4*(1 << GAME_RES_ID))*(2*(1 << GAME_RES_ID)))

typedef double Scalar_field[GA_NO_OF_SCALARS];
typedef double Vector_field[GA_NO_OF_VECTORS];
typedef double Dual_vector_field[GA_NO_OF_DUAL_VECTORS];
typedef double Curl_field[GAME_NO_OF_LAYERS*2*GA_NO_OF_VECTORS_H + GA_NO_OF_VECTORS_H];
typedef double Mass_densities[GA_NO_OF_CONSTITUENTS*GA_NO_OF_SCALARS];

/* struct grid, src/game_types.h:70-114 */
typedef struct grid {
    int no_of_oro_layers;
    Vector_field normal_distance;
    Scalar_field volume;
    Vector_field area;
    Scalar_field z_scalar;
    Vector_field z_vector;
    Scalar_field gravity_potential;
    Vector_field gravity_m;
    Vector_field slope;
    Scalar_field theta_v_bg;
    Scalar_field exner_bg;
    Vector_field exner_bg_grad;
    Scalar_field layer_thickness;
    int trsk_indices[10*GA_NO_OF_VECTORS_H];
    int trsk_modified_curl_indices[10*GA_NO_OF_VECTORS_H];
    int from_index[GA_NO_OF_VECTORS_H];
    int to_index[GA_NO_OF_VECTORS_H];
    int adjacent_vector_indices_h[6*GA_NO_OF_SCALARS_H];
    int adjacent_signs_h[6*GA_NO_OF_SCALARS_H];
    int density_to_rhombi_indices[4*GA_NO_OF_VECTORS_H];
    double latitude_scalar[GA_NO_OF_SCALARS_H];
    double longitude_scalar[GA_NO_OF_SCALARS_H];
    double inner_product_weights[8*GA_NO_OF_SCALARS];
    double direction[GA_NO_OF_VECTORS_H];
    double density_to_rhombi_weights[4*GA_NO_OF_VECTORS_H];
    double trsk_weights[10*GA_NO_OF_VECTORS_H];
    double sfc_albedo[GA_NO_OF_SCALARS_H];
    double sfc_rho_c[GA_NO_OF_SCALARS_H];
    double t_conduc_soil[GA_NO_OF_SCALARS_H];
    double roughness_length[GA_NO_OF_SCALARS_H];
    int is_land[GA_NO_OF_SCALARS_H];
    int latlon_interpol_indices[5*GA_NO_OF_LATLON_IO_POINTS];
    double latlon_interpol_weights[5*GA_NO_OF_LATLON_IO_POINTS];
    double z_soil_interface[GA_NO_OF_SOIL_LAYERS + 1];
    double z_soil_center[GA_NO_OF_SOIL_LAYERS];
    double mean_velocity_area;
    double t_const_soil[GA_NO_OF_SCALARS_H];
    double z_t_const;
    double toa;
    int oro_id;
    double stretching_parameter;
    double radius;
    double eff_hor_res;
} Grid;

/* struct dualgrid, src/game_types.h:117-126 */
typedef struct dualgrid {
    Curl_field area;
    Dual_vector_field z_vector;
    Dual_vector_field normal_distance;
    int from_index[GA_NO_OF_VECTORS_H];
    int to_index[GA_NO_OF_VECTORS_H];
    int vorticity_indices_triangles[3*GA_NO_OF_DUAL_SCALARS_H];
    int vorticity_signs_triangles[3*GA_NO_OF_DUAL_SCALARS_H];
    double f_vec[2*GA_NO_OF_VECTORS_H];
} Dualgrid;

/* struct state, src/game_types.h:128-135 */
typedef struct state {
    Mass_densities rho;
    Scalar_field rhotheta_v;
    Scalar_field theta_v_pert;
    Scalar_field exner_pert;
    Vector_field wind;
    double temperature_soil[GA_NO_OF_SOIL_LAYERS*GA_NO_OF_SCALARS_H];
} State;

/* struct diagnostics, src/game_types.h:138-162 */
typedef struct diagnostics {
    Vector_field flux_density;
    Scalar_field flux_density_divv;
    double rel_vort_on_triangles[GA_NO_OF_DUAL_V_VECTORS];
    Curl_field rel_vort;
    Curl_field pot_vort;
    Scalar_field temperature;
    Scalar_field c_g_p_field;
    Scalar_field v_squared;
    Scalar_field wind_divv;
    Vector_field curl_of_vorticity;
    Scalar_field scalar_field_placeholder;
    Vector_field vector_field_placeholder;
    Vector_field u_at_edge;
    Vector_field v_at_edge;
    Scalar_field u_at_cell;
    Scalar_field v_at_cell;
    Scalar_field n_squared;
    double dv_hdz[GA_NO_OF_H_VECTORS + GA_NO_OF_VECTORS_H];
    double scalar_flux_resistance[GA_NO_OF_SCALARS_H];
    double power_flux_density_sensible[GA_NO_OF_SCALARS_H];
    double power_flux_density_latent[GA_NO_OF_SCALARS_H];
    double roughness_velocity[GA_NO_OF_SCALARS_H];
    double monin_obukhov_length[GA_NO_OF_SCALARS_H];
} Diagnostics;

/* struct forcings, src/game_types.h:180-190 */
typedef struct forcings {
    Vector_field pgrad_acc_old;
    Vector_field pressure_gradient_acc_neg_nl;
    Vector_field pressure_gradient_acc_neg_l;
    Vector_field pressure_grad_condensates_v;
    Vector_field v_squared_grad;
    Vector_field pot_vort_tend;
    double sfc_sw_in[GA_NO_OF_SCALARS_H];
    double sfc_lw_out[GA_NO_OF_SCALARS_H];
    Scalar_field radiation_tendency;
} Forcings;

/* struct config, src/game_types.h:193-216 */
typedef struct config {
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
} Config;

/* struct irreversible_quantities, src/game_types.h:218-238 */
typedef struct irreversible_quantities {
    Scalar_field temperature_diffusion_heating;
    Vector_field friction_acc;
    Scalar_field heating_diss;
    Scalar_field molecular_diffusion_coeff;
    Scalar_field mass_diffusion_coeff_numerical_h;
    Scalar_field mass_diffusion_coeff_numerical_v;
    Scalar_field temp_diffusion_coeff_numerical_h;
    Scalar_field temp_diffusion_coeff_numerical_v;
    Scalar_field pressure_gradient_decel_factor;
    Scalar_field condensates_sediment_heat;
    double mass_diff_tendency[GA_NO_OF_CONSTITUENTS*GA_NO_OF_SCALARS];
    double phase_trans_rates[(GA_NO_OF_CONDENSED + 1)*GA_NO_OF_SCALARS];
    double phase_trans_heating_rate[GA_NO_OF_SCALARS];
    Scalar_field viscosity;
    Vector_field viscosity_rhombi;
    double viscosity_triangles[GA_NO_OF_DUAL_V_VECTORS];
    double vert_hor_viscosity[GA_NO_OF_H_VECTORS + GA_NO_OF_VECTORS_H];
    Scalar_field tke;
} Irreversible_quantities;

/* ---------------- drop-in entry points ---------------- */

/* One RKHEVI step; replaces src/time_stepping/manage_rkhevi.c:20-100
   (declared at src/time_stepping/time_stepping.h:6, called at src/coordinator.c:322).
   Reads state_old, updates state_new in place; forcings, the tendency state,
   diagnostics and irreversible quantities are carried on the device between
   calls and copied into the caller's structs after every call, as the
   reference writes them in place (with GAME_DROPIN_DEFER_AUX set they are
   copied only by game_dropin_sync_aux).  The device copies are the ones the
   next call reads.  time_coordinate is unused (radiation is outside this
   path).  Returns 0, or a game_status code. */
int manage_rkhevi(State *state_old, State *state_new, Grid *grid, Dualgrid *dualgrid, State *state_tendency,
                  Diagnostics *diagnostics, Forcings *forcings, Irreversible_quantities *irrev, Config *config,
                  double delta_t, double time_coordinate);

/* replaces src/spatial_operators/linear_combine_two_states.c:14-43
   (src/spatial_operators/spatial_operators.h:25; called at src/coordinator.c:292, 305) */
int linear_combine_two_states(State *state_0, State *state_1, State *state_out, double coeff_0, double coeff_1,
                              Grid *grid);

/* replaces src/constituents/derived_quantities.c:15-40
   (src/constituents/constituents.h:24; called at src/coordinator.c:250) */
int temperature_diagnostics(State *state, Grid *grid, Diagnostics *diagnostics);

/* replaces src/spatial_operators/inner_product.c:13-43
   (src/spatial_operators/spatial_operators.h:26; called at src/coordinator.c:251) */
int inner_product(Vector_field in_field_0, Vector_field in_field_1, Scalar_field out_field, Grid *grid);

/* replaces src/io/set_grid_properties.c:21-253 (src/io/io.h:6; called at
   src/coordinator.c:~240): reads the grid file written by the reference's
   grid generator or by game_grid_file_write, without the netCDF library,
   derives layer thickness, slope, gravity_m and exner_bg_grad and sets up
   the soil grid.  Exits with status 2 on a read error, as the reference
   does, unless GAME_GPU_NOABORT is set. */
int set_grid_properties(Grid *grid, Dualgrid *dualgrid, char grid_file_name[]);

/* the integrals of write_out_integral (src/io/write_output.c:1515-1596, io.h:12)
   without the text files: out[0..NO_OF_CONSTITUENTS) masses, then rhotheta_v,
   kinetic, potential and internal energy; compensated sums on the device */
int game_dropin_integrals(State *state, Grid *grid, double *out);

/* explicit set-up and tear-down (set-up otherwise happens on the first call) */
int game_dropin_init(Grid *grid, Dualgrid *dualgrid, Config *config);
int game_dropin_finalize(void);
/* copies the device-carried forcings, tendency state, diagnostics and
   irreversible quantities into the caller's structs */
int game_dropin_sync_aux(State *state_tendency, Diagnostics *diagnostics, Forcings *forcings,
                         Irreversible_quantities *irrev);
/* layout fingerprint of the compiled sizes (a driver compares it with its own) */
unsigned long long game_abi_fingerprint(void);

#ifdef __cplusplus
}
#endif
#endif /* GAME_ABI_H */
