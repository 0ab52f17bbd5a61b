/*
 * gridgen.h -- host-side generator of GAME's icosahedral hexagonal C-grid.
 *
 * Restates the reference's grid_generator executable
 * (grid_generator/src/coordinator.c:40-352) plus the derived metrics the
 * model computes when it loads a grid file (src/io/set_grid_properties.c:196-221),
 * with run-time sizes and linear-time neighbour searches.  The numbering of
 * cells, edges and triangles and the slot order of every neighbour table are
 * the reference's; the netCDF writer, the land-sea mask, real orography and
 * the lat-lon output interpolation are not part of this module.
 */
#ifndef GAME_GRIDGEN_H
#define GAME_GRIDGEN_H

#include "../../include/game_rkhevi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct game_grid_params {
    int res_id;
    int n_layers;
    int n_oro_layers;          /* grid_generator/run_script.sh:18 (23 for 26 layers) */
    int n_lloyd_iterations;    /* grid_generator/run_script.sh:13 (2000) */
    double stretching_parameter; /* 1.3 */
    double toa;                /* 41152 m */
    double radius_rescale;     /* 1.0 */
    int oro_id;                /* 0: flat; 2: analytic Gaussian mountain (new) */
    double oro_h0;             /* mountain height (m) */
    double oro_half_width;     /* e-folding distance (m): h = h0*exp(-(d/w)^2) */
    double oro_lat, oro_lon;   /* mountain centre (rad) */
    /* optional: 2*NSH generator coordinates (lat, lon interleaved) that replace
       the generation + Lloyd iterations (the reference's use_scalar_h_coords_file) */
    const double *generator_latlon;
    /* bound on |sum of a rhombus's density weights - 1| (grid_generator/src/
       rhombus_averaging.c:262-265); 0 selects the reference's 1e-10 */
    double rhombus_tolerance;
    /* bound of the TRSK checks 30 (kite areas sum to the cell area) and 39
       (Eq. (39) antisymmetry), grid_generator/src/coriolis.c:188-194, 393-427;
       0 selects the reference's 1e-10 */
    double trsk_tolerance;
} game_grid_params_t;

/* arrays of the grid file that are not members of struct grid / dualgrid */
typedef struct game_grid_extras {
    double *latitude_vector;            /* NVH */
    double *longitude_vector;           /* NVH */
    double *latitude_scalar_dual;       /* NT */
    double *longitude_scalar_dual;      /* NT */
    double *direction_dual;             /* NVH */
    double *pent_hex_face_unity_sphere; /* NSH */
    double *triangle_face_unit_sphere;  /* NT */
    double *oro;                        /* NSH */
    int *vorticity_indices_rhombi;      /* 4*NVH */
} game_grid_extras_t;

/* Fills every array of the three views (caller-allocated, zero-initialised).
   Returns 0 on success, a positive code when a self-check of the reference
   fails (the reference aborts there). */
int game_grid_generate(const game_grid_params_t *params, const game_dims_t *dims,
                       game_grid_view_t *grid, game_dualgrid_view_t *dualgrid,
                       game_grid_extras_t *extras);

/* Only the horizontal generators after `n_lloyd_iterations` Lloyd steps
   (2*NSH doubles, lat/lon interleaved); lets callers cache the expensive part. */
int game_grid_generators(int res_id, int n_lloyd_iterations, double *latlon_out);

/* The state one Lloyd iteration (grid_generator/src/optimize.c:17-31) starts
   from after n_lloyd_iterations: generators (NSH x 2 lat/lon), their Voronoi
   vertices (NT x 2, may be NULL), the adjacency with -1 in the pentagons'
   sixth slot (6 x NSH, may be NULL) and the dual from/to indices before their
   reorientation (NVH each, may be NULL). */
int game_grid_lloyd_state(int res_id, int n_lloyd_iterations, double *latlon_out, double *latlon_dual_out,
                          int *adj_out, int *from_dual_out, int *to_dual_out);

/* src/io/set_grid_properties.c:196-221 on a grid read from a file or just
   generated: -1 adjacency slots to 0, layer thickness, slope, gravity_m,
   exner_bg_grad */
void game_grid_derived_metrics(const game_dims_t *dims, game_grid_view_t *grid);

const char *game_grid_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
