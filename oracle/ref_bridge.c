/*
 * ref_bridge.c -- link-time companion of the reference build in oracle/_ref
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Compiled against the reference's own src/game_types.h (RES_ID 5,
 * 26 layers, MOISTURE_ON 1 as the reference ships it).  It provides:
 *   - the symbols the reference's buildable objects need from the four
 *     reference files that cannot compile here (they include the absent
 *     geos95.h): averaging.c and vorticities.c / vorticity_flux.c functions are
 *     forwarded to the oracle's restatement; the flag-guarded momentum
 *     diffusion hooks and set_scalar_h_dual_coords abort if ever called
 *     (every measured configuration keeps them off / unused);
 *   - the struct layout (sizeof/offsetof) of the reference's structs, so the
 *     tests can lay out reference structs in numpy buffers.
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "game_types.h"
#include "oracle_rkhevi.h"

static or_ctx_t ctx_of(Grid *g, Dualgrid *dg)
{
    or_ctx_t c;
    memset(&c, 0, sizeof c);
    c.d.res_id = RES_ID;
    c.d.n_layers = NO_OF_LAYERS;
    c.d.moisture_on = MOISTURE_ON;
    c.d.n_oro_layers = g->no_of_oro_layers;
    game_dims_derive(&c.d);
    c.g.no_of_oro_layers = g->no_of_oro_layers;
    c.g.normal_distance = g->normal_distance;
    c.g.volume = g->volume;
    c.g.area = g->area;
    c.g.z_scalar = g->z_scalar;
    c.g.z_vector = g->z_vector;
    c.g.gravity_potential = g->gravity_potential;
    c.g.gravity_m = g->gravity_m;
    c.g.slope = g->slope;
    c.g.theta_v_bg = g->theta_v_bg;
    c.g.exner_bg = g->exner_bg;
    c.g.exner_bg_grad = g->exner_bg_grad;
    c.g.layer_thickness = g->layer_thickness;
    c.g.trsk_indices = g->trsk_indices;
    c.g.trsk_modified_curl_indices = g->trsk_modified_curl_indices;
    c.g.from_index = g->from_index;
    c.g.to_index = g->to_index;
    c.g.adjacent_vector_indices_h = g->adjacent_vector_indices_h;
    c.g.adjacent_signs_h = g->adjacent_signs_h;
    c.g.density_to_rhombi_indices = g->density_to_rhombi_indices;
    c.g.latitude_scalar = g->latitude_scalar;
    c.g.longitude_scalar = g->longitude_scalar;
    c.g.inner_product_weights = g->inner_product_weights;
    c.g.direction = g->direction;
    c.g.density_to_rhombi_weights = g->density_to_rhombi_weights;
    c.g.trsk_weights = g->trsk_weights;
    c.g.is_land = g->is_land;
    c.g.toa = g->toa;
    c.g.radius = g->radius;
    if (dg) {
        c.dg.area = dg->area;
        c.dg.z_vector = dg->z_vector;
        c.dg.normal_distance = dg->normal_distance;
        c.dg.from_index = dg->from_index;
        c.dg.to_index = dg->to_index;
        c.dg.vorticity_indices_triangles = dg->vorticity_indices_triangles;
        c.dg.vorticity_signs_triangles = dg->vorticity_signs_triangles;
        c.dg.f_vec = dg->f_vec;
    }
    return c;
}

/* ---- forwarded to the oracle (reference files that need geos95.h) ---- */
int vector_field_hor_cov_to_con(Vector_field f, Grid *g)
{
    or_ctx_t c = ctx_of(g, NULL);
    or_vector_field_hor_cov_to_con(&c, f);
    return 0;
}

int vertical_contravariant_corr(Vector_field f, int layer, int h, Grid *g, double *result)
{
    or_ctx_t c = ctx_of(g, NULL);
    *result = or_vertical_contravariant_corr(&c, f, layer, h);
    return 0;
}

int calc_pot_vort(Vector_field wind, Scalar_field density, Diagnostics *diag, Grid *g, Dualgrid *dg)
{
    or_ctx_t c = ctx_of(g, dg);
    game_diag_view_t d;
    memset(&d, 0, sizeof d);
    d.rel_vort_on_triangles = diag->rel_vort_on_triangles;
    d.rel_vort = diag->rel_vort;
    d.pot_vort = diag->pot_vort;
    or_calc_pot_vort(&c, wind, density, &d);
    return 0;
}

int vorticity_flux(Vector_field flux, Dual_vector_field pv, Vector_field out, Grid *g, Dualgrid *dg)
{
    or_ctx_t c = ctx_of(g, dg);
    or_vorticity_flux(&c, flux, pv, out);
    return 0;
}

/* ---- never reached with the measured switches ---- */
static void unavailable(const char *name)
{
    fprintf(stderr, "ref_bridge: %s is not part of this reference build\n", name);
    abort();
}

int hor_momentum_diffusion(State *a, Diagnostics *b, Irreversible_quantities *c, Config *d, Grid *e, Dualgrid *f)
{ (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; unavailable("hor_momentum_diffusion"); return 1; }
int vert_momentum_diffusion(State *a, Diagnostics *b, Irreversible_quantities *c, Grid *d, Config *e, double f)
{ (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; unavailable("vert_momentum_diffusion"); return 1; }
int simple_dissipation_rate(State *a, Irreversible_quantities *b, Grid *c)
{ (void)a; (void)b; (void)c; unavailable("simple_dissipation_rate"); return 1; }
int set_scalar_h_dual_coords(double a[], double b[], double c[], double d[], int e[][3], int f[][3], int g[][3])
{ (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; (void)g; unavailable("set_scalar_h_dual_coords"); return 1; }

/* ---- layout of the reference structs ---- */
#define MEMBER(S, M) {#S, #M, (long)offsetof(S, M)}
static const struct { const char *s, *m; long off; } members[] = {
    MEMBER(Grid, no_of_oro_layers), MEMBER(Grid, normal_distance), MEMBER(Grid, volume), MEMBER(Grid, area),
    MEMBER(Grid, z_scalar), MEMBER(Grid, z_vector), MEMBER(Grid, gravity_potential), MEMBER(Grid, gravity_m),
    MEMBER(Grid, slope), MEMBER(Grid, theta_v_bg), MEMBER(Grid, exner_bg), MEMBER(Grid, exner_bg_grad),
    MEMBER(Grid, layer_thickness), MEMBER(Grid, trsk_indices), MEMBER(Grid, trsk_modified_curl_indices),
    MEMBER(Grid, from_index), MEMBER(Grid, to_index), MEMBER(Grid, adjacent_vector_indices_h),
    MEMBER(Grid, adjacent_signs_h), MEMBER(Grid, density_to_rhombi_indices), MEMBER(Grid, latitude_scalar),
    MEMBER(Grid, longitude_scalar), MEMBER(Grid, inner_product_weights), MEMBER(Grid, direction),
    MEMBER(Grid, density_to_rhombi_weights), MEMBER(Grid, trsk_weights), MEMBER(Grid, is_land),
    MEMBER(Grid, toa), MEMBER(Grid, radius), MEMBER(Grid, oro_id), MEMBER(Grid, stretching_parameter),
    MEMBER(Grid, sfc_albedo), MEMBER(Grid, sfc_rho_c), MEMBER(Grid, t_conduc_soil), MEMBER(Grid, roughness_length),
    MEMBER(Grid, latlon_interpol_indices), MEMBER(Grid, latlon_interpol_weights), MEMBER(Grid, z_soil_interface),
    MEMBER(Grid, z_soil_center), MEMBER(Grid, mean_velocity_area), MEMBER(Grid, t_const_soil),
    MEMBER(Grid, z_t_const), MEMBER(Grid, eff_hor_res),
    MEMBER(Dualgrid, area), MEMBER(Dualgrid, z_vector), MEMBER(Dualgrid, normal_distance),
    MEMBER(Dualgrid, from_index), MEMBER(Dualgrid, to_index), MEMBER(Dualgrid, vorticity_indices_triangles),
    MEMBER(Dualgrid, vorticity_signs_triangles), MEMBER(Dualgrid, f_vec),
    MEMBER(State, rho), MEMBER(State, rhotheta_v), MEMBER(State, theta_v_pert), MEMBER(State, exner_pert),
    MEMBER(State, wind), MEMBER(State, temperature_soil),
    MEMBER(Diagnostics, flux_density), MEMBER(Diagnostics, flux_density_divv),
    MEMBER(Diagnostics, rel_vort_on_triangles), MEMBER(Diagnostics, rel_vort), MEMBER(Diagnostics, pot_vort),
    MEMBER(Diagnostics, temperature), MEMBER(Diagnostics, scalar_field_placeholder), MEMBER(Diagnostics, v_squared),
    MEMBER(Forcings, pgrad_acc_old), MEMBER(Forcings, pressure_gradient_acc_neg_nl),
    MEMBER(Forcings, pressure_gradient_acc_neg_l), MEMBER(Forcings, pressure_grad_condensates_v),
    MEMBER(Forcings, v_squared_grad), MEMBER(Forcings, pot_vort_tend), MEMBER(Forcings, radiation_tendency),
    MEMBER(Irreversible_quantities, temperature_diffusion_heating), MEMBER(Irreversible_quantities, friction_acc),
    MEMBER(Irreversible_quantities, heating_diss), MEMBER(Irreversible_quantities, pressure_gradient_decel_factor),
    MEMBER(Irreversible_quantities, condensates_sediment_heat), MEMBER(Irreversible_quantities, mass_diff_tendency),
    MEMBER(Irreversible_quantities, phase_trans_rates), MEMBER(Irreversible_quantities, phase_trans_heating_rate),
    MEMBER(Config, totally_first_step_bool), MEMBER(Config, impl_thermo_weight),
};

long ref_offsetof(const char *s, const char *m)
{
    for (size_t k = 0; k < sizeof members/sizeof members[0]; ++k)
        if (!strcmp(members[k].s, s) && !strcmp(members[k].m, m)) return members[k].off;
    return -1;
}

long ref_sizeof(const char *s)
{
    if (!strcmp(s, "Grid")) return (long)sizeof(Grid);
    if (!strcmp(s, "Dualgrid")) return (long)sizeof(Dualgrid);
    if (!strcmp(s, "State")) return (long)sizeof(State);
    if (!strcmp(s, "Diagnostics")) return (long)sizeof(Diagnostics);
    if (!strcmp(s, "Forcings")) return (long)sizeof(Forcings);
    if (!strcmp(s, "Irreversible_quantities")) return (long)sizeof(Irreversible_quantities);
    if (!strcmp(s, "Config")) return (long)sizeof(Config);
    return -1;
}

/* the compiled-in sizes (src/game_types.h:12-60) */
void ref_sizes(long *out)
{
    out[0] = RES_ID; out[1] = NO_OF_LAYERS; out[2] = MOISTURE_ON; out[3] = NO_OF_SCALARS_H;
    out[4] = NO_OF_VECTORS_H; out[5] = NO_OF_SCALARS; out[6] = NO_OF_VECTORS; out[7] = NO_OF_CONSTITUENTS;
}
