/*
 * game_files.h -- the reference's on-disk formats, in lib/libgame_host.so:
 * netCDF classic files written and read without the netCDF library, the
 * grid file of the grid generator and the state file of the model.
 *
 * Replaces:
 *   grid_generator/src/coordinator.c:354-611   (writes the grid file)
 *   src/io/set_grid_properties.c:21-253        (reads it, derives the metrics)
 *   src/io/write_output.c:1428-1503            (writes the state file)
 *   src/io/set_initial_state.c:244-336         (read_init_data reads it)
 *   grid_generator/src/interpolate_ll.c:16-74  (lat-lon output interpolation)
 * All functions return 0 or a positive error code with a message in
 * game_nc_last_error().
 */
#ifndef GAME_FILES_H
#define GAME_FILES_H

#include "game_rkhevi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* netCDF classic (CDF-1, and CDF-2 when offsets pass 2^31)              */
/* ------------------------------------------------------------------ */
#define GAME_NC_INT 4
#define GAME_NC_DOUBLE 6
#define NC_NAME_MAX 256
enum { GAME_NC_EARG = 1, GAME_NC_EIO = 2, GAME_NC_ENOTVAR = 3 };

/* one one-dimensional variable: its dimension (shared by name), type, data
   and optional "units" text attribute */
typedef struct game_nc_var {
    const char *name;
    const char *dim;
    long long len;
    int type;               /* GAME_NC_INT or GAME_NC_DOUBLE */
    const void *data;
    const char *units;      /* or NULL */
} game_nc_var_t;

/* writes the variables in the given order (nc_create(NC_CLOBBER), nc_def_dim,
   nc_def_var, nc_put_att_text, nc_enddef, nc_put_var_*) */
int game_nc_write(const char *path, const game_nc_var_t *vars, int n_vars);

typedef struct game_nc_file game_nc_file_t;
/* nc_open(NC_NOWRITE); NULL on failure */
game_nc_file_t *game_nc_open(const char *path);
/* nc_inq_varid + the variable's type and number of elements; GAME_NC_ENOTVAR if absent */
int game_nc_inq(const game_nc_file_t *f, const char *name, int *type, long long *len);
/* nc_get_var_int / nc_get_var_double: type and length must match the file's */
int game_nc_get(game_nc_file_t *f, const char *name, int type, void *out, long long len);
void game_nc_close(game_nc_file_t *f);
const char *game_nc_last_error(void);

/* ------------------------------------------------------------------ */
/* the grid file                                                         */
/* ------------------------------------------------------------------ */
/* lat-lon output points of the reference (src/game_types.h:54-56) */
static inline long long game_latlon_points(int res_id) { return 8LL << (2*res_id); }

/* the grid file's arrays that are not members of struct grid / dualgrid
   (reading: NULL members are skipped) */
typedef struct game_grid_file_extras {
    double *latitude_vector;            /* NVH */
    double *longitude_vector;           /* NVH */
    double *latitude_scalar_dual;       /* NT */
    double *longitude_scalar_dual;      /* NT */
    int *interpol_indices;              /* 5*game_latlon_points(res) */
    double *interpol_weights;           /* 5*game_latlon_points(res) */
    double *sfc_albedo;                 /* NSH */
    double *sfc_rho_c;                  /* NSH */
    double *t_conductivity;             /* NSH */
    double *roughness_length;           /* NSH */
    int n_lloyd_iterations;
} game_grid_file_extras_t;

/* The reference's grid file: every variable, dimension and unit of
   grid_generator/src/coordinator.c:354-611, in its order; the pentagons'
   sixth adjacency slot is written as -1 as the generator leaves it. */
int game_grid_file_write(const char *path, const game_dims_t *dims, const game_grid_view_t *grid,
                         const game_dualgrid_view_t *dualgrid, const game_grid_file_extras_t *extras);
/* src/io/set_grid_properties.c:21-221: the variables, -1 adjacency slots to
   0, then layer thickness, slope (grad_hor_cov of z_scalar), gravity_m and
   exner_bg_grad (grad).  dims must match the file's sizes. */
int game_grid_file_read(const char *path, const game_dims_t *dims, game_grid_view_t *grid,
                        game_dualgrid_view_t *dualgrid, game_grid_file_extras_t *extras);

/* grid_generator/src/interpolate_ll.c:16-74: the five nearest cells of every
   lat-lon output point and their inverse-distance weights, found by a walk
   over the cell adjacency instead of a scan of every cell, with the
   reference's selection rule (index_helpers.c:49-67) reproduced exactly. */
int game_latlon_interpolation(int res_id, const double *latitude_scalar, const double *longitude_scalar,
                              const int *adjacent_vector_indices_h, const int *from_index, const int *to_index,
                              int *interpol_indices, double *interpol_weights);

/* the ocean surface properties the generator gives a cell that is not land
   (grid_generator/src/phys_sfc_properties.c:150-198); land cells are
   rejected (no land mask is generated here) */
int game_sfc_properties(int nsh, const int *is_land, double *sfc_albedo, double *sfc_rho_c, double *t_conductivity,
                        double *roughness_length);

/* The generator's statistics file (grid_generator/src/miscellaneous.c:92-135,
   written at grid_generator/src/coordinator.c:337-338 as
   statistics/RES<r>_L<l>_ORO<o>.txt, grid_name RES<r>_L<l>_ORO<o>): the Lloyd
   iterations and the extremes of the cells' unit-sphere areas (NSH), of the
   highest layer's horizontal normal distances (normal_distance[NSH + e]) and
   of the highest level's dual ones (normal_distance_dual[e]), in the
   reference's words and %lf format. */
int game_grid_statistics_write(const char *path, const char *grid_name, int n_lloyd, int nsh, int nvh,
                               const double *face_area_unity, const double *normal_distance,
                               const double *normal_distance_dual);

/* ------------------------------------------------------------------ */
/* the state file                                                        */
/* ------------------------------------------------------------------ */
/* src/io/write_output.c:1428-1503: densities (n_const*S), temperature (S),
   wind (V), and rh, rel_vort, divv_h_all_layers, tke (S) and t_soil
   (soil layers*NSH), each NULL for zeros */
int game_state_file_write(const char *path, const game_dims_t *dims, const double *densities,
                          const double *temperature, const double *wind, const double *rh, const double *rel_vort,
                          const double *divv_h_all_layers, const double *tke, const double *t_soil, int n_soil_layers);
/* the reads of read_init_data (src/io/set_initial_state.c:252-282):
   densities, temperature and wind; tke when the file has it (*tke_found 1)
   and tke is non-NULL.  game_state_from_temperature then completes the
   state as read_init_data does. */
int game_state_file_read(const char *path, const game_dims_t *dims, double *densities, double *temperature,
                         double *wind, double *tke, int *tke_found);

#ifdef __cplusplus
}
#endif
#endif
