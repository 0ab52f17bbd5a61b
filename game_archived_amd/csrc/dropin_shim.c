/*
 * dropin_shim.c -- the reference's entry points over the device step.
 *
 * Compiled once per layout (-DGAME_RES_ID, -DGAME_NO_OF_LAYERS,
 * -DGAME_MOISTURE_ON) into lib/libgame_dropin_r{R}l{L}m{M}.so.  A GAME driver
 * linked against it instead of src/time_stepping/ and src/spatial_operators/
 * calls manage_rkhevi exactly as src/coordinator.c:322 does.
 *
 * Coherence: the driver owns host structs and mutates them between calls
 * (src/coordinator.c:305 copies state_new into state_old), so every call
 * uploads the states it reads and downloads the state it writes.  The fields
 * the step carries from call to call (forcings, the tendency state,
 * irreversible quantities, diagnostics) stay on the device;
 * game_dropin_sync_aux copies them out.  Not thread-safe, like the reference.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "../../include/game_abi.h"
#include "../../include/game_rkhevi.h"
#include "../../include/game_files.h"

static const Grid *g_grid = NULL;
static game_grid_view_t g_gv;
static game_dualgrid_view_t g_dgv;

static game_state_view_t state_view(State *s)
{
    game_state_view_t v;
    v.rho = s->rho;
    v.rhotheta_v = s->rhotheta_v;
    v.theta_v_pert = s->theta_v_pert;
    v.exner_pert = s->exner_pert;
    v.wind = s->wind;
    return v;
}

static game_config_t config_copy(const Config *c)
{
    game_config_t o;
    memcpy(&o, c, sizeof o);
    return o;
}

static void fail(int rc, const char *what)
{
    fprintf(stderr, "game drop-in: %s failed (%d): %s\n", what, rc, game_gpu_last_error());
    if (!getenv("GAME_GPU_NOABORT")) exit(1);   /* the reference aborts on errors too */
}

unsigned long long game_abi_fingerprint(void)
{
    unsigned long long h = 1469598103934665603ULL;
    unsigned long long v[] = {GAME_RES_ID, GAME_NO_OF_LAYERS, GAME_MOISTURE_ON, sizeof(Grid), sizeof(Dualgrid),
                              sizeof(State), sizeof(Diagnostics), sizeof(Forcings), sizeof(Config),
                              sizeof(Irreversible_quantities)};
    for (size_t i = 0; i < sizeof v/sizeof v[0]; ++i) { h ^= v[i]; h *= 1099511628211ULL; }
    return h;
}

static game_dims_t layout_dims(int n_oro_layers)
{
    game_dims_t d;
    memset(&d, 0, sizeof d);
    d.res_id = GAME_RES_ID;
    d.n_layers = GAME_NO_OF_LAYERS;
    d.moisture_on = GAME_MOISTURE_ON;
    d.n_oro_layers = n_oro_layers;
    game_dims_derive(&d);
    return d;
}

/* views of the members of struct grid / dualgrid the library reads or fills */
static void struct_views(Grid *grid, Dualgrid *dualgrid, game_grid_view_t *g, game_dualgrid_view_t *dg)
{
    memset(g, 0, sizeof *g);
    g->no_of_oro_layers = grid->no_of_oro_layers;
    g->normal_distance = grid->normal_distance; g->volume = grid->volume; g->area = grid->area;
    g->z_scalar = grid->z_scalar; g->z_vector = grid->z_vector; g->gravity_potential = grid->gravity_potential;
    g->gravity_m = grid->gravity_m; g->slope = grid->slope; g->theta_v_bg = grid->theta_v_bg;
    g->exner_bg = grid->exner_bg; g->exner_bg_grad = grid->exner_bg_grad; g->layer_thickness = grid->layer_thickness;
    g->trsk_indices = grid->trsk_indices; g->trsk_modified_curl_indices = grid->trsk_modified_curl_indices;
    g->from_index = grid->from_index; g->to_index = grid->to_index;
    g->adjacent_vector_indices_h = grid->adjacent_vector_indices_h; g->adjacent_signs_h = grid->adjacent_signs_h;
    g->density_to_rhombi_indices = grid->density_to_rhombi_indices;
    g->latitude_scalar = grid->latitude_scalar; g->longitude_scalar = grid->longitude_scalar;
    g->inner_product_weights = grid->inner_product_weights; g->direction = grid->direction;
    g->density_to_rhombi_weights = grid->density_to_rhombi_weights; g->trsk_weights = grid->trsk_weights;
    g->is_land = grid->is_land; g->toa = grid->toa; g->radius = grid->radius;
    g->stretching_parameter = grid->stretching_parameter; g->oro_id = grid->oro_id;
    dg->area = dualgrid->area; dg->z_vector = dualgrid->z_vector; dg->normal_distance = dualgrid->normal_distance;
    dg->from_index = dualgrid->from_index; dg->to_index = dualgrid->to_index;
    dg->vorticity_indices_triangles = dualgrid->vorticity_indices_triangles;
    dg->vorticity_signs_triangles = dualgrid->vorticity_signs_triangles; dg->f_vec = dualgrid->f_vec;
}

int game_dropin_init(Grid *grid, Dualgrid *dualgrid, Config *config)
{
    game_dims_t d = layout_dims(grid->no_of_oro_layers);
    struct_views(grid, dualgrid, &g_gv, &g_dgv);
    game_config_t c = config_copy(config);
    int rc = game_gpu_init(&d, &g_gv, &g_dgv, &c, GAME_FP64);
    if (rc) { fail(rc, "game_gpu_init"); return rc; }
    g_grid = grid;
    return 0;
}

/* src/io/set_grid_properties.c:21-253 without the netCDF library: the grid
   file's variables into the structs, the derived metrics, and the soil
   grid the reference sets up at the end */
int set_grid_properties(Grid *grid, Dualgrid *dualgrid, char grid_file_name[])
{
    int n_oro = -1;
    game_nc_file_t *f = game_nc_open(grid_file_name);
    if (!f || game_nc_get(f, "no_of_oro_layers", GAME_NC_INT, &n_oro, 1) != 0) {
        fprintf(stderr, "game drop-in: reading %s failed: %s\n", grid_file_name, game_nc_last_error());
        if (f) game_nc_close(f);
        if (!getenv("GAME_GPU_NOABORT")) exit(2);
        return GAME_NC_EIO;
    }
    game_nc_close(f);
    game_dims_t d = layout_dims(n_oro);
    game_grid_view_t g;
    game_dualgrid_view_t dg;
    struct_views(grid, dualgrid, &g, &dg);
    game_grid_file_extras_t x;
    memset(&x, 0, sizeof x);
    x.interpol_indices = grid->latlon_interpol_indices;
    x.interpol_weights = grid->latlon_interpol_weights;
    x.sfc_albedo = grid->sfc_albedo;
    x.sfc_rho_c = grid->sfc_rho_c;
    x.t_conductivity = grid->t_conduc_soil;
    x.roughness_length = grid->roughness_length;
    int rc = game_grid_file_read(grid_file_name, &d, &g, &dg, &x);
    if (rc) {
        fprintf(stderr, "game drop-in: reading %s failed: %s\n", grid_file_name, game_nc_last_error());
        if (!getenv("GAME_GPU_NOABORT")) exit(2);
        return rc;
    }
    grid->no_of_oro_layers = g.no_of_oro_layers;
    grid->toa = g.toa;
    grid->radius = g.radius;
    grid->stretching_parameter = g.stretching_parameter;
    /* set_grid_properties.c:223-253: the soil's constant-temperature depth
       and temperature, and the soil layers (sigma 0.352, rescaled to the depth) */
    const double t_0 = 273.15, sigma_soil = 0.352;
    grid->z_t_const = -10.0;
    for (int i = 0; i < GA_NO_OF_SCALARS_H; ++i) grid->t_const_soil[i] = t_0 + 25.0*cos(2.0*grid->latitude_scalar[i]);
    grid->z_soil_interface[0] = 0;
    for (int i = 1; i < GA_NO_OF_SOIL_LAYERS + 1; ++i)
        grid->z_soil_interface[i] = grid->z_soil_interface[i - 1] + pow(sigma_soil, GA_NO_OF_SOIL_LAYERS - i);
    const double rescale_factor = grid->z_t_const/grid->z_soil_interface[GA_NO_OF_SOIL_LAYERS];
    for (int i = 1; i < GA_NO_OF_SOIL_LAYERS + 1; ++i) grid->z_soil_interface[i] = rescale_factor*grid->z_soil_interface[i];
    for (int i = 0; i < GA_NO_OF_SOIL_LAYERS; ++i)
        grid->z_soil_center[i] = 0.5*(grid->z_soil_interface[i] + grid->z_soil_interface[i + 1]);
    return 0;
}

int game_dropin_finalize(void)
{
    g_grid = NULL;
    return game_gpu_finalize();
}

static int ensure_init(Grid *grid, Dualgrid *dualgrid, Config *config)
{
    if (g_grid == grid) return 0;
    if (!dualgrid || !config) { fprintf(stderr, "game drop-in: call manage_rkhevi or game_dropin_init first\n"); return GAME_ERR_NOT_INIT; }
    return game_dropin_init(grid, dualgrid, config);
}

int manage_rkhevi(State *state_old, State *state_new, Grid *grid, Dualgrid *dualgrid, State *state_tendency,
                  Diagnostics *diagnostics, Forcings *forcings, Irreversible_quantities *irrev, Config *config,
                  double delta_t, double time_coordinate)
{
    (void)time_coordinate;
    int rc = ensure_init(grid, dualgrid, config);
    if (rc) return rc;
    game_config_t c = config_copy(config);
    if ((rc = game_gpu_set_config(&c))) { fail(rc, "config"); return rc; }
    game_state_view_t vo = state_view(state_old), vn = state_view(state_new);
    if ((rc = game_gpu_set_state(GAME_STATE_OLD, &vo))) { fail(rc, "upload state_old"); return rc; }
    if ((rc = game_gpu_set_state(GAME_STATE_NEW, &vn))) { fail(rc, "upload state_new"); return rc; }
    if ((rc = game_gpu_manage_rkhevi(delta_t, config->totally_first_step_bool))) { fail(rc, "manage_rkhevi"); return rc; }
    if ((rc = game_gpu_get_state(GAME_STATE_NEW, &vn))) { fail(rc, "download state_new"); return rc; }
    /* the reference's step writes the tendency state, diagnostics, forcings
       and irreversible quantities in place, and a driver may read them at
       output time: they are copied back after every step unless the driver
       asks to fetch them itself (GAME_DROPIN_DEFER_AUX, then
       game_dropin_sync_aux) */
    if (!getenv("GAME_DROPIN_DEFER_AUX")
        && (rc = game_dropin_sync_aux(state_tendency, diagnostics, forcings, irrev))) {
        fail(rc, "download of the carried fields");
        return rc;
    }
    return 0;
}

int linear_combine_two_states(State *s0, State *s1, State *out, double c0, double c1, Grid *grid)
{
    int rc = ensure_init(grid, NULL, NULL);
    if (rc) return rc;
    game_state_view_t v0 = state_view(s0), v1 = state_view(s1), vo = state_view(out);
    if ((rc = game_gpu_set_state(GAME_STATE_OLD, &v0))) { fail(rc, "upload"); return rc; }
    if ((rc = game_gpu_set_state(GAME_STATE_NEW, &v1))) { fail(rc, "upload"); return rc; }
    if ((rc = game_gpu_linear_combine_two_states(GAME_STATE_OLD, GAME_STATE_NEW, GAME_STATE_OLD, c0, c1))) {
        fail(rc, "linear_combine_two_states"); return rc; }
    if ((rc = game_gpu_get_state(GAME_STATE_OLD, &vo))) { fail(rc, "download"); return rc; }
    /* soil temperature is not carried by the device step */
    for (size_t i = 0; i < sizeof out->temperature_soil/sizeof(double); ++i)
        out->temperature_soil[i] = c0*s0->temperature_soil[i] + c1*s1->temperature_soil[i];
    return 0;
}

int temperature_diagnostics(State *state, Grid *grid, Diagnostics *diagnostics)
{
    int rc = ensure_init(grid, NULL, NULL);
    if (rc) return rc;
    game_state_view_t v = state_view(state);
    if ((rc = game_gpu_set_state(GAME_STATE_OLD, &v))) { fail(rc, "upload"); return rc; }
    if ((rc = game_gpu_temperature_diagnostics(GAME_STATE_OLD))) { fail(rc, "temperature"); return rc; }
    game_diag_view_t dv;
    memset(&dv, 0, sizeof dv);
    dv.temperature = diagnostics->temperature;
    if ((rc = game_gpu_get_diagnostics(&dv))) { fail(rc, "download"); return rc; }
    return 0;
}

int inner_product(Vector_field a, Vector_field b, Scalar_field out, Grid *grid)
{
    int rc = ensure_init(grid, NULL, NULL);
    if (rc) return rc;
    if ((rc = game_gpu_inner_product(a, b, out))) { fail(rc, "inner_product"); return rc; }
    return 0;
}

int game_dropin_integrals(State *state, Grid *grid, double *out)
{
    int rc = ensure_init(grid, NULL, NULL);
    if (rc) return rc;
    game_state_view_t v = state_view(state);
    if ((rc = game_gpu_set_state(GAME_STATE_OLD, &v))) { fail(rc, "upload"); return rc; }
    if ((rc = game_gpu_integrals(GAME_STATE_OLD, out))) { fail(rc, "integrals"); return rc; }
    return 0;
}

int game_dropin_sync_aux(State *tend, Diagnostics *diag, Forcings *forc, Irreversible_quantities *irrev)
{
    int rc;
    if (tend) {
        game_state_view_t v = state_view(tend);
        if ((rc = game_gpu_get_state(GAME_STATE_TENDENCY, &v))) return rc;
    }
    if (forc) {
        game_forcings_view_t f = {forc->pgrad_acc_old, forc->pressure_gradient_acc_neg_nl,
                                  forc->pressure_gradient_acc_neg_l, forc->pressure_grad_condensates_v,
                                  forc->v_squared_grad, forc->pot_vort_tend, forc->radiation_tendency};
        if ((rc = game_gpu_get_forcings(&f))) return rc;
    }
    if (diag) {
        game_diag_view_t d = {diag->flux_density, diag->flux_density_divv, diag->rel_vort_on_triangles, NULL,
                              diag->pot_vort, diag->temperature, NULL, diag->v_squared};
        if ((rc = game_gpu_get_diagnostics(&d))) return rc;
    }
    if (irrev) {
        game_irrev_view_t i;
        memset(&i, 0, sizeof i);
        i.condensates_sediment_heat = irrev->condensates_sediment_heat;
        i.phase_trans_rates = irrev->phase_trans_rates;
        i.phase_trans_heating_rate = irrev->phase_trans_heating_rate;
        if ((rc = game_gpu_get_irrev(&i))) return rc;
    }
    return 0;
}
