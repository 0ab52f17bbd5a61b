/*
 * gridfile.c -- the reference's grid file and state file on top of ncio.c.
 *
 * The variable names, dimension names, types, units and order are the
 * reference's (grid_generator/src/coordinator.c:354-611 for the grid file,
 * src/io/write_output.c:1428-1503 for the state file): they are the file
 * format, so a file written here is read by the reference and the other
 * way round.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/game_files.h"
#include "gridgen.h"

#define NV_MAX 64

typedef struct {
    game_nc_var_t v[NV_MAX];
    int n;
} vlist_t;

static void add(vlist_t *L, const char *name, const char *dim, long long len, int type, const void *data,
                const char *units)
{
    game_nc_var_t *v = &L->v[L->n++];
    v->name = name; v->dim = dim; v->len = len; v->type = type; v->data = data; v->units = units;
}

#define D GAME_NC_DOUBLE
#define I GAME_NC_INT

int game_grid_file_write(const char *path, const game_dims_t *d, const game_grid_view_t *g,
                         const game_dualgrid_view_t *dg, const game_grid_file_extras_t *x)
{
    if (!path || !d || !g || !dg || !x || !x->latitude_vector || !x->longitude_vector || !x->latitude_scalar_dual
        || !x->longitude_scalar_dual || !x->interpol_indices || !x->interpol_weights || !x->sfc_albedo
        || !x->sfc_rho_c || !x->t_conductivity || !x->roughness_length)
        return GAME_NC_EARG;
    const long long S = d->s, V = d->v, NSH = d->nsh, NVH = d->nvh, NT = d->nt, DV = d->dual_v;
    const long long NLL = game_latlon_points(d->res_id);
    /* the generator leaves the pentagons' sixth adjacency slot at -1 (index.c:66-70) */
    int *adj = malloc(sizeof(int)*6*(size_t)NSH);
    if (!adj) return GAME_NC_EIO;
    memcpy(adj, g->adjacent_vector_indices_h, sizeof(int)*6*(size_t)NSH);
    for (int i = 0; i < 12; ++i) adj[6*i + 5] = -1;
    const int n_lloyd = x->n_lloyd_iterations, n_oro = g->no_of_oro_layers;
    const double stretching = g->stretching_parameter, toa = g->toa, radius = g->radius;
    vlist_t L = {.n = 0};
    add(&L, "no_of_lloyd_iterations", "single_int_dimid_index", 1, I, &n_lloyd, NULL);
    add(&L, "no_of_oro_layers", "single_int_dimid_index", 1, I, &n_oro, NULL);
    add(&L, "stretching_parameter", "single_double_dimid_index", 1, D, &stretching, NULL);
    add(&L, "toa", "single_double_dimid_index", 1, D, &toa, "m");
    add(&L, "radius", "single_double_dimid_index", 1, D, &radius, "m");
    add(&L, "latitude_scalar", "scalar_h_index", NSH, D, g->latitude_scalar, NULL);
    add(&L, "longitude_scalar", "scalar_h_index", NSH, D, g->longitude_scalar, NULL);
    add(&L, "latitude_scalar_dual", "scalar_dual_h_index", NT, D, x->latitude_scalar_dual, NULL);
    add(&L, "longitude_scalar_dual", "scalar_dual_h_index", NT, D, x->longitude_scalar_dual, NULL);
    add(&L, "z_scalar", "scalar_index", S, D, g->z_scalar, "m");
    add(&L, "theta_v_bg", "scalar_index", S, D, g->theta_v_bg, "K");
    add(&L, "exner_bg", "scalar_index", S, D, g->exner_bg, NULL);
    add(&L, "gravity_potential", "scalar_index", S, D, g->gravity_potential, "m^2/s^2");
    add(&L, "z_vector", "vector_index", V, D, g->z_vector, "m");
    add(&L, "normal_distance", "vector_index", V, D, g->normal_distance, "m");
    add(&L, "volume", "scalar_index", S, D, g->volume, "m^3");
    add(&L, "area", "vector_index", V, D, g->area, "m^2");
    add(&L, "trsk_weights", "vector_h_10_index", 10*NVH, D, g->trsk_weights, NULL);
    add(&L, "z_vector_dual", "vector_index_dual", DV, D, dg->z_vector, "m");
    add(&L, "normal_distance_dual", "vector_index_dual", DV, D, dg->normal_distance, "m");
    add(&L, "area_dual", "vector_index_dual", DV, D, dg->area, "m^2");
    add(&L, "f_vec", "f_vec_index", 2*NVH, D, dg->f_vec, "1/s");
    add(&L, "direction", "vector_h_index", NVH, D, g->direction, NULL);
    add(&L, "latitude_vector", "vector_h_index", NVH, D, x->latitude_vector, NULL);
    add(&L, "longitude_vector", "vector_h_index", NVH, D, x->longitude_vector, NULL);
    add(&L, "inner_product_weights", "scalar_8_index", 8*S, D, g->inner_product_weights, NULL);
    add(&L, "density_to_rhombi_weights", "vector_h_4_index", 4*NVH, D, g->density_to_rhombi_weights, NULL);
    add(&L, "interpol_weights", "latlon_3_index", 5*NLL, D, x->interpol_weights, NULL);
    add(&L, "from_index", "vector_h_index", NVH, I, g->from_index, NULL);
    add(&L, "to_index", "vector_h_index", NVH, I, g->to_index, NULL);
    add(&L, "from_index_dual", "vector_h_index", NVH, I, dg->from_index, NULL);
    add(&L, "to_index_dual", "vector_h_index", NVH, I, dg->to_index, NULL);
    add(&L, "adjacent_vector_indices_h", "scalar_h_6_index", 6*NSH, I, adj, NULL);
    add(&L, "interpol_indices", "latlon_3_index", 5*NLL, I, x->interpol_indices, NULL);
    add(&L, "trsk_indices", "vector_h_10_index", 10*NVH, I, g->trsk_indices, NULL);
    add(&L, "trsk_modified_curl_indices", "vector_h_10_index", 10*NVH, I, g->trsk_modified_curl_indices, NULL);
    add(&L, "adjacent_signs_h", "scalar_h_6_index", 6*NSH, I, g->adjacent_signs_h, NULL);
    add(&L, "vorticity_signs_triangles", "scalar_dual_h_3_index", 3*NT, I, dg->vorticity_signs_triangles, NULL);
    add(&L, "vorticity_indices_triangles", "scalar_dual_h_3_index", 3*NT, I, dg->vorticity_indices_triangles, NULL);
    add(&L, "density_to_rhombi_indices", "vector_h_4_index", 4*NVH, I, g->density_to_rhombi_indices, NULL);
    add(&L, "sfc_albedo", "scalar_h_index", NSH, D, x->sfc_albedo, NULL);
    add(&L, "sfc_rho_c", "scalar_h_index", NSH, D, x->sfc_rho_c, "J/(K*m**3)");
    add(&L, "is_land", "scalar_h_index", NSH, I, g->is_land, NULL);
    add(&L, "t_conductivity", "scalar_h_index", NSH, D, x->t_conductivity, "m^2/2");
    add(&L, "roughness_length", "scalar_h_index", NSH, D, x->roughness_length, "m");
    const int rc = game_nc_write(path, L.v, L.n);
    free(adj);
    return rc;
}

/* reads one variable when the destination is non-NULL */
static int get(game_nc_file_t *f, const char *name, int type, void *out, long long len)
{
    return out ? game_nc_get(f, name, type, out, len) : 0;
}

int game_grid_file_read(const char *path, const game_dims_t *d, game_grid_view_t *g, game_dualgrid_view_t *dg,
                        game_grid_file_extras_t *x)
{
    if (!path || !d || !g || !dg) return GAME_NC_EARG;
    game_nc_file_t *f = game_nc_open(path);
    if (!f) return GAME_NC_EIO;
    const long long S = d->s, V = d->v, NSH = d->nsh, NVH = d->nvh, NT = d->nt, DV = d->dual_v;
    const long long NLL = game_latlon_points(d->res_id);
    int n_oro = 0, n_lloyd = 0;
    int rc = get(f, "no_of_oro_layers", I, &n_oro, 1)
             || get(f, "toa", D, &g->toa, 1)
             || get(f, "radius", D, &g->radius, 1)
             || get(f, "stretching_parameter", D, &g->stretching_parameter, 1)
             || get(f, "normal_distance", D, g->normal_distance, V)
             || get(f, "volume", D, g->volume, S)
             || get(f, "area", D, g->area, V)
             || get(f, "z_scalar", D, g->z_scalar, S)
             || get(f, "theta_v_bg", D, g->theta_v_bg, S)
             || get(f, "exner_bg", D, g->exner_bg, S)
             || get(f, "gravity_potential", D, g->gravity_potential, S)
             || get(f, "z_vector", D, g->z_vector, V)
             || get(f, "trsk_weights", D, g->trsk_weights, 10*NVH)
             || get(f, "area_dual", D, dg->area, DV)
             || get(f, "z_vector_dual", D, dg->z_vector, DV)
             || get(f, "f_vec", D, dg->f_vec, 2*NVH)
             || get(f, "to_index", I, g->to_index, NVH)
             || get(f, "to_index_dual", I, dg->to_index, NVH)
             || get(f, "direction", D, g->direction, NVH)
             || get(f, "normal_distance_dual", D, dg->normal_distance, DV)
             || get(f, "from_index", I, g->from_index, NVH)
             || get(f, "from_index_dual", I, dg->from_index, NVH)
             || get(f, "adjacent_vector_indices_h", I, g->adjacent_vector_indices_h, 6*NSH)
             || get(f, "vorticity_indices_triangles", I, dg->vorticity_indices_triangles, 3*NT)
             || get(f, "vorticity_signs_triangles", I, dg->vorticity_signs_triangles, 3*NT)
             || get(f, "trsk_indices", I, g->trsk_indices, 10*NVH)
             || get(f, "trsk_modified_curl_indices", I, g->trsk_modified_curl_indices, 10*NVH)
             || get(f, "adjacent_signs_h", I, g->adjacent_signs_h, 6*NSH)
             || get(f, "inner_product_weights", D, g->inner_product_weights, 8*S)
             || get(f, "density_to_rhombi_weights", D, g->density_to_rhombi_weights, 4*NVH)
             || get(f, "density_to_rhombi_indices", I, g->density_to_rhombi_indices, 4*NVH)
             || get(f, "latitude_scalar", D, g->latitude_scalar, NSH)
             || get(f, "longitude_scalar", D, g->longitude_scalar, NSH)
             || get(f, "is_land", I, g->is_land, NSH);
    if (!rc && x)
        rc = get(f, "no_of_lloyd_iterations", I, &n_lloyd, 1)
             || get(f, "interpol_indices", I, x->interpol_indices, 5*NLL)
             || get(f, "interpol_weights", D, x->interpol_weights, 5*NLL)
             || get(f, "sfc_rho_c", D, x->sfc_rho_c, NSH)
             || get(f, "sfc_albedo", D, x->sfc_albedo, NSH)
             || get(f, "roughness_length", D, x->roughness_length, NSH)
             || get(f, "t_conductivity", D, x->t_conductivity, NSH)
             || get(f, "latitude_vector", D, x->latitude_vector, NVH)
             || get(f, "longitude_vector", D, x->longitude_vector, NVH)
             || get(f, "latitude_scalar_dual", D, x->latitude_scalar_dual, NT)
             || get(f, "longitude_scalar_dual", D, x->longitude_scalar_dual, NT);
    game_nc_close(f);
    if (rc) return GAME_NC_EARG;
    if (n_oro != d->n_oro_layers) return GAME_NC_EARG;
    g->no_of_oro_layers = n_oro;
    if (x) x->n_lloyd_iterations = n_lloyd;
    game_grid_derived_metrics(d, g);
    return 0;
}

int game_state_file_write(const char *path, const game_dims_t *d, const double *rho, const double *temperature,
                          const double *wind, const double *rh, const double *rel_vort, const double *divv,
                          const double *tke, const double *t_soil, int n_soil)
{
    if (!path || !d || !rho || !temperature || !wind || n_soil < 1) return GAME_NC_EARG;
    const long long S = d->s, V = d->v, NSOIL = (long long)n_soil*d->nsh;
    double *zeros = calloc((size_t)(S > NSOIL ? S : NSOIL), sizeof(double));
    if (!zeros) return GAME_NC_EIO;
    vlist_t L = {.n = 0};
    add(&L, "densities", "densities_index", (long long)d->n_const*S, D, rho, "kg/m^3");
    add(&L, "temperature", "scalar_index", S, D, temperature, "K");
    add(&L, "wind", "vector_index", V, D, wind, "m/s");
    add(&L, "rh", "scalar_index", S, D, rh ? rh : zeros, "%");
    add(&L, "rel_vort", "scalar_index", S, D, rel_vort ? rel_vort : zeros, "1/s");
    add(&L, "divv_h_all_layers", "scalar_index", S, D, divv ? divv : zeros, "1/s");
    add(&L, "tke", "scalar_index", S, D, tke ? tke : zeros, "J/kg");
    add(&L, "t_soil", "soil_index", NSOIL, D, t_soil ? t_soil : zeros, "K");
    const int rc = game_nc_write(path, L.v, L.n);
    free(zeros);
    return rc;
}

int game_state_file_read(const char *path, const game_dims_t *d, double *rho, double *temperature, double *wind,
                         double *tke, int *tke_found)
{
    if (!path || !d || !rho || !temperature || !wind) return GAME_NC_EARG;
    game_nc_file_t *f = game_nc_open(path);
    if (!f) return GAME_NC_EIO;
    const long long S = d->s, V = d->v;
    int rc = game_nc_get(f, "densities", D, rho, (long long)d->n_const*S)
             || game_nc_get(f, "temperature", D, temperature, S)
             || game_nc_get(f, "wind", D, wind, V);
    int found = 0;
    if (!rc && game_nc_inq(f, "tke", NULL, NULL) == 0) {
        found = 1;
        if (tke) rc = game_nc_get(f, "tke", D, tke, S);
    }
    game_nc_close(f);
    if (tke_found) *tke_found = found;
    return rc ? GAME_NC_EARG : 0;
}

/* ------------------------------------------------------------------ */
/* the generator's statistics file                                      */
/* ------------------------------------------------------------------ */
/* extremes with the reference's scan (index_helpers.c:13-47: the first
   strict improvement from element 0) */
static void extremes(const double *x, int n, double *lo, double *hi)
{
    double a = x[0], b = x[0];
    for (int i = 1; i < n; ++i) {
        if (x[i] < a) a = x[i];
        if (x[i] > b) b = x[i];
    }
    *lo = a;
    *hi = b;
}

int game_grid_statistics_write(const char *path, const char *grid_name, int n_lloyd, int nsh, int nvh,
                               const double *face_area_unity, const double *normal_distance,
                               const double *normal_distance_dual)
{
    if (!path || !grid_name || nsh < 1 || nvh < 1 || !face_area_unity || !normal_distance || !normal_distance_dual)
        return GAME_NC_EARG;
    double amin, amax, dmin, dmax, ddmin, ddmax;
    extremes(face_area_unity, nsh, &amin, &amax);
    extremes(normal_distance + nsh, nvh, &dmin, &dmax);     /* the highest layer's edges */
    extremes(normal_distance_dual, nvh, &ddmin, &ddmax);    /* the highest level's dual edges */
    FILE *f = fopen(path, "w");
    if (!f) return GAME_NC_EIO;
    fprintf(f, "Statistical properties of grid %s:\n\n", grid_name);
    fprintf(f, "Number of Lloyd iterations: %d\n", n_lloyd);
    fprintf(f, "Ratio of minimum to maximum area: %lf\n", amin/amax);
    fprintf(f, "Shortest horizontal normal distance (highest layer): %lf m\n", dmin);
    fprintf(f, "Longest horizontal normal distance (highest layer): %lf m\n", dmax);
    fprintf(f, "Ratio of shortest to longest horizontal normal distance: %lf\n", dmin/dmax);
    fprintf(f, "Shortest horizontal normal distance dual (highest level): %lf m\n", ddmin);
    fprintf(f, "Longest horizontal normal distance dual (highest level): %lf m\n", ddmax);
    fprintf(f, "Ratio of shortest to longest dual horizontal normal distance: %lf\n", ddmin/ddmax);
    return fclose(f) ? GAME_NC_EIO : 0;
}
