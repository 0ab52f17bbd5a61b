/*
 * gridgen.c -- GAME's icosahedral hexagonal C-grid, generated on the host.
 *
 * Every function cites the reference function whose result it reproduces.
 * The arithmetic of the geometric helpers keeps the reference's operation
 * order (including pow(x, 2)) so that the generated floating-point weights
 * can be compared bit for bit with the reference's own routines.  The
 * combinatorics differ only where the reference scans all N elements per
 * element: those searches are replaced by single passes that produce the
 * same slot order (ascending scan order), which is checked in the tests.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gridgen.h"

#define N_PENT 12
#define N_ICO_EDGES 30
#define N_FACES 20
#define ORTH_CRITERION_DEG 89.99
#define EPS_SECURITY 1e-10
#define RADIUS_EARTH 6371000.789927
#define OMEGA_EARTH (7.292115e-5)
#define G_MEAN 9.80616
#define P_0_REF 100000.0
#define C_D_P_REF 1005.0
#define R_D_REF 287.057811

static char g_err[256];

/* the error flag of a parallel construction loop: threads record the first
   failure they meet and skip their remaining work; atomic accesses keep the
   flag free of data races (any code a thread records makes the loop fail) */
#define SET_ERR(e, v) do { _Pragma("omp atomic write") (e) = (v); } while (0)
static inline int get_err_(const int *e) { int v; _Pragma("omp atomic read") v = *e; return v; }
#define GET_ERR(e) get_err_(&(e))
const char *game_grid_last_error(void) { return g_err; }
#define FAIL(code, ...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return (code); } while (0)

/* ------------------------------------------------------------------ */
/* geometry helpers: grid_generator/src/geodesy.c                       */
/* ------------------------------------------------------------------ */
static void find_global_normal(double lat, double lon, double *x, double *y, double *z)
{   /* geodesy.c:152-161 */
    *x = cos(lat)*cos(lon);
    *y = cos(lat)*sin(lon);
    *z = sin(lat);
}

static void find_geos(double x, double y, double z, double *lat, double *lon)
{   /* geodesy.c:142-150 */
    *lat = asin(z/sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2)));
    *lon = atan2(y, x);
}

static double dist_cart(double lat1, double lon1, double lat2, double lon2, double r1, double r2)
{   /* geodesy.c:51-72 */
    if (lat1 == lat2 && lon1 == lon2) return 0;
    double x1, y1, z1, x2, y2, z2;
    find_global_normal(lat1, lon1, &x1, &y1, &z1);
    find_global_normal(lat2, lon2, &x2, &y2, &z2);
    x1 = r1*x1; y1 = r1*y1; z1 = r1*z1;
    x2 = r2*x2; y2 = r2*y2; z2 = r2*z2;
    return sqrt(pow(x2 - x1, 2) + pow(y2 - y1, 2) + pow(z2 - z1, 2));
}

static double dist_h(double lat_a, double lon_a, double lat_b, double lon_b, double radius)
{   /* geodesy.c:42-49 */
    return 2.0*radius*asin(sqrt(0.5 - 0.5*(cos(lat_a)*cos(lat_b)*cos(lon_b - lon_a) + sin(lat_a)*sin(lat_b))));
}

static void find_geodetic(double lat1, double lon1, double lat2, double lon2, double parameter,
                          double *lat_out, double *lon_out)
{   /* geodesy.c:15-29 */
    double d = dist_cart(lat1, lon1, lat2, lon2, 1.0, 1.0);
    double theta = 2.0*asin(d/2.0);
    double tau_dash = 0.5 + sqrt(1.0/pow(d, 2) - 0.25)*tan(theta*(parameter - 0.5));
    double z = tau_dash*sin(lat2) + (1.0 - tau_dash)*sin(lat1);
    double x = tau_dash*cos(lat2)*cos(lon2) + (1.0 - tau_dash)*cos(lat1)*cos(lon1);
    double y = tau_dash*cos(lat2)*sin(lon2) + (1.0 - tau_dash)*cos(lat1)*sin(lon1);
    *lat_out = asin(z/sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2)));
    *lon_out = atan2(y, x);
}

static double dot3(const double a[3], const double b[3])
{   /* geodesy.c:94-105 (left-to-right accumulation from 0) */
    double s = 0;
    for (int i = 0; i < 3; ++i) s = s + a[i]*b[i];
    return s;
}

static double dot2(const double a[2], const double b[2])
{   /* geodesy.c:107-118 */
    double s = 0;
    for (int i = 0; i < 2; ++i) s = s + a[i]*b[i];
    return s;
}

static double geodetic_direction(double lat1, double lon1, double lat2, double lon2, double parameter)
{   /* geodesy.c:74-92 */
    double rel[3], li[3], lj[3];
    rel[0] = cos(lat2)*cos(lon2) - cos(lat1)*cos(lon1);
    rel[1] = cos(lat2)*sin(lon2) - cos(lat1)*sin(lon1);
    rel[2] = sin(lat2) - sin(lat1);
    double lat, lon = 0;
    find_geodetic(lat1, lon1, lat2, lon2, parameter, &lat, &lon);
    li[0] = -sin(lon); li[1] = cos(lon); li[2] = 0;                          /* geodesy.c:120-129 */
    lj[0] = -sin(lat)*cos(lon); lj[1] = -sin(lat)*sin(lon); lj[2] = cos(lat); /* geodesy.c:131-140 */
    double xc = dot3(li, rel);
    double yc = dot3(lj, rel);
    return atan2(yc, xc);
}

static double turn_angle(double a0, double a1)
{   /* geodesy.c:496-511 */
    double r = a1 - a0;
    if (r > M_PI) r -= 2.0*M_PI;
    if (r < -M_PI) r += 2.0*M_PI;
    return r;
}

static double rad2deg(double x) { return x*360.0/(2.0*M_PI); }   /* geodesy.c:522-529 */

static void active_turn_x(double angle, const double in[3], double out[3])
{   /* geodesy.c:248-257 */
    out[0] = in[0];
    out[1] = cos(angle)*in[1] + -sin(angle)*in[2];
    out[2] = sin(angle)*in[1] + cos(angle)*in[2];
}

static double triangle_area(double lat0, double lon0, double lat1, double lon1, double lat2, double lon2)
{   /* geodesy.c:259-322 */
    double avg_lat = (lat0 + lat1 + lat2)/3.0;
    if (fabs(avg_lat) > 0.9*M_PI/2) {
        double v_in[3], v_out[3];
        find_global_normal(lat0, lon0, &v_in[0], &v_in[1], &v_in[2]);
        active_turn_x(avg_lat, v_in, v_out);
        find_geos(v_out[0], v_out[1], v_out[2], &lat0, &lon0);
        find_global_normal(lat1, lon1, &v_in[0], &v_in[1], &v_in[2]);
        active_turn_x(avg_lat, v_in, v_out);
        find_geos(v_out[0], v_out[1], v_out[2], &lat1, &lon1);
        find_global_normal(lat2, lon2, &v_in[0], &v_in[1], &v_in[2]);
        active_turn_x(avg_lat, v_in, v_out);
        find_geos(v_out[0], v_out[1], v_out[2], &lat2, &lon2);
    }
    double d01 = geodetic_direction(lat0, lon0, lat1, lon1, 0);
    double d02 = geodetic_direction(lat0, lon0, lat2, lon2, 0);
    double d10 = geodetic_direction(lat1, lon1, lat0, lon0, 0);
    double d12 = geodetic_direction(lat1, lon1, lat2, lon2, 0);
    double d20 = geodetic_direction(lat2, lon2, lat0, lon0, 0);
    double d21 = geodetic_direction(lat2, lon2, lat1, lon1, 0);
    double v01[2] = {cos(d01), sin(d01)}, v02[2] = {cos(d02), sin(d02)};
    double v10[2] = {cos(d10), sin(d10)}, v12[2] = {cos(d12), sin(d12)};
    double v20[2] = {cos(d20), sin(d20)}, v21[2] = {cos(d21), sin(d21)};
    double a0 = acos(dot2(v01, v02));
    double a1 = acos(dot2(v10, v12));
    double a2 = acos(dot2(v20, v21));
    return a0 + a1 + a2 - M_PI;
}

static int min_index(const double *v, int n)
{   /* index_helpers.c:13-29 */
    int r = 0;
    double m = v[0];
    for (int i = 1; i < n; ++i) if (v[i] < m) { m = v[i]; r = i; }
    return r;
}

static int max_index(const double *v, int n)
{   /* index_helpers.c:31-47 */
    int r = 0;
    double m = v[0];
    for (int i = 1; i < n; ++i) if (v[i] > m) { m = v[i]; r = i; }
    return r;
}

static int in_list(int value, const int *arr, int n)
{   /* index_helpers.c:125-140 */
    for (int i = 0; i < n; ++i) if (arr[i] == value) return 1;
    return 0;
}

/* Orders the corners of a polygon counter-clockwise (geodesy.c:389-494):
   nearest-neighbour chaining from corner 0, reversed when the summed turn
   angle is negative. */
static void sort_edge_indices(const double *lat, const double *lon, int n, int *sorted)
{
    double dist_arr[8];
    int idx_arr[8], nb[16];
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j) {
            double dc = dist_cart(lat[i], lon[i], lat[j], lon[j], 1, 1);
            if (dc != 0) { idx_arr[c] = j; dist_arr[c] = dc; ++c; }
        }
        nb[2*i + 0] = idx_arr[min_index(dist_arr, n - 1)];
        dist_arr[min_index(dist_arr, n - 1)] = 2.1;
        nb[2*i + 1] = idx_arr[min_index(dist_arr, n - 1)];
    }
    for (int i = 1; i < n; ++i) sorted[i] = -1;
    sorted[0] = 0;
    int cand[2] = {0, 0};
    for (int i = 1; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j)
            if (nb[2*j + 0] == sorted[i - 1] || nb[2*j + 1] == sorted[i - 1]) { if (c < 2) cand[c] = j; ++c; }
        sorted[i] = in_list(cand[0], sorted, n) ? cand[1] : cand[0];
    }
    double angle_sum = 0.0;
    for (int i = 0; i < n; ++i) {
        int a = sorted[i], b = sorted[(i + 1) % n], c = sorted[(i + 2) % n];
        double dir0 = geodetic_direction(lat[a], lon[a], lat[b], lon[b], 1.0);
        double dir1 = geodetic_direction(lat[b], lon[b], lat[c], lon[c], 0.0);
        angle_sum += turn_angle(dir0, dir1);
    }
    if (angle_sum < -0.9*2.0*M_PI) {
        int tmp[8];
        for (int i = 0; i < n; ++i) tmp[i] = sorted[n - 1 - i];
        for (int i = 0; i < n; ++i) sorted[i] = tmp[i];
    }
}

static double spherical_polygon_area(const double *lat, const double *lon, int n)
{   /* geodesy.c:324-367 */
    double x[8], y[8], z[8];
    for (int i = 0; i < n; ++i) find_global_normal(lat[i], lon[i], &x[i], &y[i], &z[i]);
    double xc = 0, yc = 0, zc = 0;
    for (int i = 0; i < n; ++i) {
        xc += 1.0/n*x[i];
        yc += 1.0/n*y[i];
        zc += 1.0/n*z[i];
    }
    double latc, lonc;
    find_geos(xc, yc, zc, &latc, &lonc);
    int srt[8];
    sort_edge_indices(lat, lon, n, srt);
    double ls[8], os[8], tri[8];
    for (int i = 0; i < n; ++i) { ls[i] = lat[srt[i]]; os[i] = lon[srt[i]]; }
    for (int i = 0; i < n; ++i)
        tri[i] = triangle_area(latc, lonc, ls[i], os[i], ls[(i + 1) % n], os[(i + 1) % n]);
    double area = 0;
    for (int i = 0; i < n; ++i) area += tri[i];
    return area;
}

static void voronoi_center(double lat0, double lon0, double lat1, double lon1, double lat2, double lon2,
                           double *lat_out, double *lon_out)
{   /* geodesy.c:173-194 */
    double x0, y0, z0, x1, y1, z1, x2, y2, z2;
    find_global_normal(lat0, lon0, &x0, &y0, &z0);
    find_global_normal(lat1, lon1, &x1, &y1, &z1);
    find_global_normal(lat2, lon2, &x2, &y2, &z2);
    double a[3] = {x1 - x0, y1 - y0, z1 - z0};
    double b[3] = {x2 - x0, y2 - y0, z2 - z0};
    double c[3];
    c[0] = a[1]*b[2] - a[2]*b[1];
    c[1] = a[2]*b[0] - a[0]*b[2];
    c[2] = a[0]*b[1] - a[1]*b[0];
    find_geos(c[0], c[1], c[2], lat_out, lon_out);
}

static double vertical_area(double base, double r1, double r2)
{   /* geodesy.c:163-171 */
    return base*(0.5*pow(r2, 2)/r1 - 0.5*r1);
}

static double box_volume(double area1, double r1, double r2)
{   /* geodesy.c:196-204 */
    return area1/(3.0*pow(r1, 2))*(pow(r2, 3) - pow(r1, 3));
}

static double min_dist_rel_on_line(double lat0, double lon0, double lat1, double lon1, double latp, double lonp)
{   /* geodesy.c:369-387: 1001-point search along the geodesic */
    const int np = 1000 + 1;
    int best = 0;
    double best_d = 0;
    for (int i = 0; i < np; ++i) {
        double par = (i + 0.0)/(np + 1.0), lat, lon;
        find_geodetic(lat0, lon0, lat1, lon1, par, &lat, &lon);
        double d = dist_cart(latp, lonp, lat, lon, 1, 1);
        if (i == 0 || d < best_d) { best_d = d; best = i; }
    }
    return (best + 0.0)/(np + 1.0);
}

/* ------------------------------------------------------------------ */
/* discrete coordinates on the icosahedron:                             */
/* grid_generator/src/discrete_coordinate_trafos.c                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int res;                       /* target resolution (RES_ID) */
    int face_vertices[N_FACES][3];
    int face_edges[N_FACES][3];
    int face_edges_reverse[N_FACES][3];
    int edge_vertices[N_ICO_EDGES][2];
    double lat_ico[N_PENT], lon_ico[N_PENT];
} ico_t;

static int ppe_of(int res) { return (1 << res) - 1; }                               /* :479-486 */
static int sppif_of(int res) { return ((1 << res) - 2)*((1 << res) - 1)/2; }        /* :488-496 */
static int tpf_of(int res) { return 1 << (2*res); }                                 /* :498-505 */
static int vppif_of(int res) { return 3*((1 << res) - 1)*(1 << res)/2; }            /* game_types.h:59 */

static void coords_from_tofi(int tofi, int res, int *c0, int *c1, int *c0pa)
{   /* discrete_coordinate_trafos.c:15-40 */
    int ppe = ppe_of(res), c1p = -1, maxi = -1, mini;
    for (;;) {
        ++c1p;
        *c0pa = ppe - c1p;
        maxi = maxi + *c0pa;
        mini = maxi - (*c0pa - 1);
        if (tofi <= maxi && tofi >= mini) { *c0 = tofi - mini; break; }
    }
    *c1 = c1p;
}

static int tofi_from_coords(int c0, int c1, int res)
{   /* discrete_coordinate_trafos.c:42-61 */
    int t = 0, amount = ppe_of(res);
    for (int i = 0; i < c1; ++i) { t += amount; amount -= 1; }
    return t + c0;
}

static void triangle_edge_points(const ico_t *ico, int tofi, int face, int res, int *p0, int *p1, int *p2,
                                 int *p3, int *p4, int *p5, int *dsofi)
{   /* discrete_coordinate_trafos.c:79-190 */
    int c0, c1, c0pa;
    coords_from_tofi(tofi, res, &c0, &c1, &c0pa);
    *dsofi = 1 + 2*tofi + c1;
    int ppe = ppe_of(res), sppif = sppif_of(res);
    const int *fe = ico->face_edges[face], *fr = ico->face_edges_reverse[face], *fv = ico->face_vertices[face];
    if (c1 == 0) {
        if (fr[0] == 0) *p0 = N_PENT + fe[0]*ppe + c0;
        else *p0 = N_PENT + (fe[0] + 1)*ppe - 1 - c0;
    } else
        *p0 = N_PENT + ppe*N_ICO_EDGES + face*sppif + tofi - ppe;
    if (c0 == ppe - 1 - c1) {
        if (fr[1] == 0) *p1 = N_PENT + fe[1]*ppe + c1;
        else *p1 = N_PENT + (fe[1] + 1)*ppe - 1 - c1;
    } else
        *p1 = N_PENT + ppe*N_ICO_EDGES + face*sppif + tofi - c1;
    if (c0 == 0) {
        if (fr[2] == 0) *p2 = N_PENT + (fe[2] + 1)*ppe - 1 - c1;
        else *p2 = N_PENT + fe[2]*ppe + c1;
    } else
        *p2 = N_PENT + ppe*N_ICO_EDGES + face*sppif + tofi - 1 - c1;
    if (c1 == 0) {
        if (c0 == 0) *p3 = fv[0];
        else *p3 = (fr[0] == 0) ? *p0 - 1 : *p0 + 1;
    } else if (c0 == 0)
        *p3 = (fr[2] == 0) ? *p2 + 1 : *p2 - 1;
    else
        *p3 = *p0 - 1;
    *p4 = -1;
    *p5 = -1;
    if (c0 == c0pa - 1) {
        if (c1 == 0) *p4 = fv[1];
        else *p4 = (fr[1] == 0) ? *p1 - 1 : *p1 + 1;
        if (c1 == ppe - 1) *p5 = fv[2];
    }
}

/* Inverse of the dual-scalar-on-face numbering at one resolution level:
   for every dual scalar index d, the triangle-on-face index and the flags
   computed by discrete_coordinate_trafos.c:192-252 (which scans t upwards;
   the sets {d0,d1,d2,d3}(t) partition [0, 4^res), so a table is exact). */
typedef struct { int tofi, down, special, last; } dual_entry_t;

static dual_entry_t *dual_table(int res)
{
    int n = tpf_of(res), ppe = ppe_of(res);
    dual_entry_t *tab = malloc(sizeof(dual_entry_t)*(size_t)n);
    for (int i = 0; i < n; ++i) tab[i].tofi = -1;
    int ntri = ppe*(ppe + 1)/2;
    for (int t = 0; t < ntri; ++t) {
        int c0, c1, c0pa;
        coords_from_tofi(t, res, &c0, &c1, &c0pa);
        int d0 = 2*t + 1 + c1, d1 = d0 - 1, d2 = -1, d3 = -1;
        if (c0 == c0pa - 1) { d2 = d0 + 1; if (c1 == ppe - 1) d3 = d2 + 1; }
        dual_entry_t e0 = {t, 1, 0, 0}, e1 = {t, 0, 0, 0}, e2 = {t, 0, 1, 0}, e3 = {t, 0, 0, 1};
        if (tab[d0].tofi < 0) tab[d0] = e0;
        if (tab[d1].tofi < 0) tab[d1] = e1;
        if (d2 >= 0 && tab[d2].tofi < 0) tab[d2] = e2;
        if (d3 >= 0 && tab[d3].tofi < 0) tab[d3] = e3;
    }
    return tab;
}

static void tri_points_from_dsofi(const ico_t *ico, const dual_entry_t *tab, int dsofi, int face, int res,
                                  int *p0, int *p1, int *p2)
{   /* discrete_coordinate_trafos.c:254-322 */
    dual_entry_t e = tab[dsofi];
    int c0, c1, c0pa, r0, r1, r2, r3, a0, a1, dump;
    coords_from_tofi(e.tofi, res, &c0, &c1, &c0pa);
    int ppe = ppe_of(res);
    triangle_edge_points(ico, e.tofi, face, res, &r0, &r1, &r2, &r3, &a0, &a1, &dump);
    if (e.down == 1) { *p0 = r0; *p1 = r1; *p2 = r2; return; }
    if (c0 == c0pa - 1) {
        if (c1 == ppe - 1) {
            if (e.last == 1) { *p0 = r2; *p1 = r1; *p2 = a1; }
            else if (e.special == 1) { *p0 = r0; *p1 = a0; *p2 = r1; }
            else { *p0 = r3; *p1 = r0; *p2 = r2; }
        } else {
            if (e.special == 1) { *p0 = r0; *p1 = a0; *p2 = r1; }
            else { *p0 = r3; *p1 = r0; *p2 = r2; }
        }
    } else { *p0 = r3; *p1 = r0; *p2 = r2; }
}

static int upscale_point(int res, int old, int R)
{   /* discrete_coordinate_trafos.c:324-355 */
    int ppe = ppe_of(res), sppif = sppif_of(res);
    int ppeR = ppe_of(R), sppifR = sppif_of(R);
    int scale = 1 << (R - res);
    if (old < N_PENT) return old;
    if (old < N_PENT + N_ICO_EDGES*ppe) {
        int edge = (old - N_PENT)/ppe;
        int oe = old - (N_PENT + edge*ppe);
        return N_PENT + edge*ppeR + scale*(oe + 1) - 1;
    }
    int face = (old - (N_PENT + N_ICO_EDGES*ppe))/sppif;
    int ofi = old - (N_PENT + N_ICO_EDGES*ppe + face*sppif);
    int c0, c1, c0pa;
    coords_from_tofi(ofi + ppe, res, &c0, &c1, &c0pa);
    c0 = (c0 + 1)*scale - 1;
    c1 = c1*scale;
    int t = tofi_from_coords(c0, c1, R);
    return N_PENT + N_ICO_EDGES*ppeR + face*sppifR + t - ppeR;
}

static void write_scalar_coordinates(int e0, int e1, int e2, int p0, int p1, int p2, int up,
                                     double *x, double *y, double *z, double *lat, double *lon)
{   /* discrete_coordinate_trafos.c:357-443 */
    int ends[3][2] = {{e0, e1}, {e1, e2}, {e2, e0}};
    int dst[3];
    if (up == 1) { dst[0] = p0; dst[1] = p1; dst[2] = p2; }
    else { dst[0] = p1; dst[1] = p2; dst[2] = p0; }
    for (int m = 0; m < 3; ++m) {
        int a = ends[m][0], b = ends[m][1];
        double xr = x[a] + 0.5*(x[b] - x[a]);   /* geodesy.c:31-40 */
        double yr = y[a] + 0.5*(y[b] - y[a]);
        double zr = z[a] + 0.5*(z[b] - z[a]);
        double len = sqrt(pow(xr, 2) + pow(yr, 2) + pow(zr, 2));  /* geodesy.c:225-235 */
        xr = xr/len; yr = yr/len; zr = zr/len;
        x[dst[m]] = xr; y[dst[m]] = yr; z[dst[m]] = zr;
        find_geos(xr, yr, zr, &lat[dst[m]], &lon[dst[m]]);
    }
}

static int build_icosahedron(ico_t *ico)
{   /* discrete_coordinate_trafos.c:507-759 */
    ico->lat_ico[0] = M_PI/2;
    for (int i = 1; i <= 5; ++i) ico->lat_ico[i] = atan(0.5);
    for (int i = 6; i <= 10; ++i) ico->lat_ico[i] = -atan(0.5);
    ico->lat_ico[11] = -M_PI/2;
    ico->lon_ico[0] = 0;
    ico->lon_ico[1] = 0;
    ico->lon_ico[2] = 1*2*M_PI/5;
    ico->lon_ico[3] = 2*2*M_PI/5;
    ico->lon_ico[4] = 3*2*M_PI/5;
    ico->lon_ico[5] = 4*2*M_PI/5;
    ico->lon_ico[6] = 2*M_PI/10;
    ico->lon_ico[7] = 2*M_PI/10 + 1*2*M_PI/5;
    ico->lon_ico[8] = 2*M_PI/10 + 2*2*M_PI/5;
    ico->lon_ico[9] = 2*M_PI/10 + 3*2*M_PI/5;
    ico->lon_ico[10] = 2*M_PI/10 + 4*2*M_PI/5;
    ico->lon_ico[11] = 0;
    static const int ev[N_ICO_EDGES][2] = {
        {0,1},{0,2},{0,3},{0,4},{0,5},{1,2},{2,3},{3,4},{4,5},{5,1},
        {1,6},{2,6},{2,7},{3,7},{3,8},{4,8},{4,9},{5,9},{5,10},{1,10},
        {10,6},{6,7},{7,8},{8,9},{9,10},{6,11},{7,11},{8,11},{9,11},{10,11}};
    static const int fv[N_FACES][3] = {
        {0,1,2},{0,2,3},{0,3,4},{0,4,5},{0,5,1},{1,10,6},{6,2,1},{2,6,7},{7,3,2},{3,7,8},
        {8,4,3},{4,8,9},{9,5,4},{5,9,10},{10,1,5},{11,6,10},{11,7,6},{11,8,7},{11,9,8},{11,10,9}};
    memcpy(ico->edge_vertices, ev, sizeof ev);
    memcpy(ico->face_vertices, fv, sizeof fv);
    int edge_count[N_ICO_EDGES] = {0};
    for (int i = 0; i < N_FACES; ++i)
        for (int j = 0; j < 3; ++j) {
            int other_face_vertex = fv[i][(j + 1) % 3];
            for (int k = 0; k < N_ICO_EDGES; ++k) {
                int other;
                if (ev[k][0] == fv[i][j]) other = 1;
                else if (ev[k][1] == fv[i][j]) other = 0;
                else continue;
                if (ev[k][other] == other_face_vertex) {
                    ico->face_edges[i][j] = k;
                    ico->face_edges_reverse[i][j] = (other == 1) ? 0 : 1;
                    ++edge_count[k];
                }
            }
        }
    for (int k = 0; k < N_ICO_EDGES; ++k)
        if (edge_count[k] != 2) FAIL(10, "icosahedron edge %d belongs to %d faces", k, edge_count[k]);
    return 0;
}

/* ------------------------------------------------------------------ */
/* horizontal structure                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
    int R, nsh, nvh, nt;
    ico_t ico;
    int *tri_pts;   /* 3*NT: ordered corners of every triangle as used for its
                       Voronoi centre and area (horizontal_generation.c:316-362, 176-239) */
} hgrid_t;

static void generate_generators(const hgrid_t *hg, double *lat, double *lon, double *x, double *y, double *z)
{   /* horizontal_generation.c:19-110 */
    const ico_t *ico = &hg->ico;
    int R = hg->R;
    for (int i = 0; i < N_PENT; ++i) {
        lat[i] = ico->lat_ico[i];
        lon[i] = ico->lon_ico[i];
        find_global_normal(ico->lat_ico[i], ico->lon_ico[i], &x[i], &y[i], &z[i]);
    }
    dual_entry_t **tabs = calloc((size_t)R + 2, sizeof(dual_entry_t *));
    for (int j = 0; j <= R; ++j) tabs[j] = dual_table(j);
    for (int i = 0; i < N_FACES; ++i) {
        for (int j = 0; j < R; ++j) {
            int ntpf = tpf_of(j);
            for (int k = 0; k < ntpf; ++k) {
                int p0, p1, p2;
                if (j == 0) {
                    tri_points_from_dsofi(ico, tabs[1], 1, i, 1, &p0, &p1, &p2);
                    p0 = upscale_point(1, p0, R);
                    p1 = upscale_point(1, p1, R);
                    p2 = upscale_point(1, p2, R);
                    write_scalar_coordinates(ico->face_vertices[i][0], ico->face_vertices[i][1], ico->face_vertices[i][2],
                                             p0, p1, p2, 1, x, y, z, lat, lon);
                } else {
                    int e0, e1, e2;
                    tri_points_from_dsofi(ico, tabs[j], k, i, j, &e0, &e1, &e2);
                    dual_entry_t de = tabs[j][k];
                    int c0, c1, c0pa;
                    coords_from_tofi(de.tofi, j, &c0, &c1, &c0pa);
                    int ppe = ppe_of(j);
                    int base_old = 0, base_down = 0, base_up = base_down + 4*ppe + 3;
                    for (int l = 0; l < c1; ++l) {
                        int amount = ppe - l;
                        base_old += 2*amount + 1;
                        base_down += 4*(2*amount + 1);
                        base_up = base_down + 4*(ppe - l) + 3;
                    }
                    if (de.last == 1) { base_old += 3; base_down += 12; base_up = base_down + 3; }
                    int old_on_line = k - base_old;
                    int dsofi = (de.down == 0) ? base_down + 1 + 2*old_on_line : base_up + 2*old_on_line;
                    tri_points_from_dsofi(ico, tabs[j + 1], dsofi, i, j + 1, &p0, &p1, &p2);
                    e0 = upscale_point(j, e0, R);
                    e1 = upscale_point(j, e1, R);
                    e2 = upscale_point(j, e2, R);
                    p0 = upscale_point(j + 1, p0, R);
                    p1 = upscale_point(j + 1, p1, R);
                    p2 = upscale_point(j + 1, p2, R);
                    write_scalar_coordinates(e0, e1, e2, p0, p1, p2, de.down == 1 ? 0 : 1, x, y, z, lat, lon);
                }
            }
        }
    }
    for (int j = 0; j <= R; ++j) free(tabs[j]);
    free(tabs);
}

static void inner_edge_points(const hgrid_t *hg, int i, int *p, int *dsofi, int *small, int *face,
                              int *c0, int *c1, int *c0pa)
{   /* discrete_coordinate_trafos.c:63-77 */
    int R = hg->R;
    int first = N_ICO_EDGES*(ppe_of(R) + 1), vppif = vppif_of(R);
    *face = (i - first)/vppif;
    int ofi = i - (first + *face*vppif);
    int t = ofi/3;
    *small = ofi - 3*t;
    triangle_edge_points(&hg->ico, t, *face, R, &p[0], &p[1], &p[2], &p[3], &p[4], &p[5], dsofi);
    coords_from_tofi(t, R, c0, c1, c0pa);
}

static void set_from_to_index(const hgrid_t *hg, int *from, int *to)
{   /* horizontal_generation.c:262-314 */
    int R = hg->R, ppe = ppe_of(R);
    #pragma omp parallel for
    for (int i = 0; i < hg->nvh; ++i) {
        if (i < N_ICO_EDGES*(ppe + 1)) {
            int edge = i/(ppe + 1), oe = i - edge*(ppe + 1);
            if (oe == 0) { from[i] = hg->ico.edge_vertices[edge][0]; to[i] = N_PENT + edge*ppe; }
            else if (oe == ppe) { from[i] = N_PENT + (edge + 1)*ppe - 1; to[i] = hg->ico.edge_vertices[edge][1]; }
            else { from[i] = N_PENT + edge*ppe + oe - 1; to[i] = N_PENT + edge*ppe + oe; }
        } else {
            int p[6], ds, sm, fc, c0, c1, c0pa;
            inner_edge_points(hg, i, p, &ds, &sm, &fc, &c0, &c1, &c0pa);
            if (sm == 0) { from[i] = p[0]; to[i] = p[2]; }
            if (sm == 1) { from[i] = p[0]; to[i] = p[1]; }
            if (sm == 2) { from[i] = p[2]; to[i] = p[1]; }
        }
    }
}

static void set_from_to_index_dual(const hgrid_t *hg, int *from_d, int *to_d)
{   /* horizontal_generation.c:364-517 */
    int R = hg->R, ppe = ppe_of(R), tpf = tpf_of(R), vppif = vppif_of(R);
    const ico_t *ico = &hg->ico;
    #pragma omp parallel for
    for (int i = 0; i < hg->nvh; ++i) {
        if (i < N_ICO_EDGES*(ppe + 1)) {
            int edge = i/(ppe + 1), oe = i - edge*(ppe + 1);
            int f0 = 0, f1 = 0, found = 0;
            for (int j = 0; j < N_FACES; ++j)
                if (ico->face_edges[j][0] == edge || ico->face_edges[j][1] == edge || ico->face_edges[j][2] == edge) {
                    if (!found) { f0 = j; found = 1; } else f1 = j;
                }
            int faces[2] = {f0, f1}, tofi[2] = {0, 0};
            for (int m = 0; m < 2; ++m) {
                int f = faces[m], rel = 0;
                if (ico->face_edges[f][0] == edge) rel = 0;
                if (ico->face_edges[f][1] == edge) rel = 1;
                if (ico->face_edges[f][2] == edge) rel = 2;
                int rev = ico->face_edges_reverse[f][rel];
                if (rel == 0) tofi[m] = (rev == 0) ? 2*oe : 2*ppe - 2*oe;
                if (rel == 1) tofi[m] = (rev == 0) ? -1 + (oe + 1)*(2*ppe - oe + 1) : tpf - oe*oe - 1;
                if (rel == 2) tofi[m] = (rev == 0) ? tpf - 1 - oe*(oe + 2) : oe*(2*ppe + 2 - oe);
            }
            to_d[i] = f0*tpf + tofi[0];
            from_d[i] = f1*tpf + tofi[1];
        } else {
            int first = N_ICO_EDGES*(ppe + 1);
            int face = (i - first)/vppif;
            int ofi = i - (first + face*vppif);
            int t = ofi/3, sm = ofi - 3*t, c0, c1, c0pa;
            coords_from_tofi(t, R, &c0, &c1, &c0pa);
            if (sm == 0) { from_d[i] = face*tpf + 2*t + c1; to_d[i] = from_d[i] + 1; }
            if (sm == 1) { from_d[i] = face*tpf + 2*t + 1 + c1; to_d[i] = from_d[i] + 1; }
            if (sm == 2) { from_d[i] = face*tpf + 2*t + 1 + c1; to_d[i] = from_d[i] + 2*c0pa; }
        }
    }
}

/* ordered corners of every triangle (horizontal_generation.c:327-360) */
static int set_triangle_points(hgrid_t *hg)
{
    int R = hg->R, ppe = ppe_of(R), tpf = tpf_of(R);
    int *tp = hg->tri_pts;
    for (int i = 0; i < 3*hg->nt; ++i) tp[i] = -1;
    for (int i = N_ICO_EDGES*(ppe + 1); i < hg->nvh; ++i) {
        int p[6], ds, sm, fc, c0, c1, c0pa;
        inner_edge_points(hg, i, p, &ds, &sm, &fc, &c0, &c1, &c0pa);
        int d = ds + fc*tpf;
        int *t0 = &tp[3*d], *tm = &tp[3*(d - 1)];
        t0[0] = p[0]; t0[1] = p[1]; t0[2] = p[2];
        tm[0] = p[3]; tm[1] = p[0]; tm[2] = p[2];
        if (c0 == c0pa - 1) {
            int *tp1 = &tp[3*(d + 1)];
            tp1[0] = p[0]; tp1[1] = p[4]; tp1[2] = p[1];
            if (c1 == ppe - 1) {
                int *tp2 = &tp[3*(d + 2)];
                tp2[0] = p[2]; tp2[1] = p[1]; tp2[2] = p[5];
            }
        }
    }
    for (int i = 0; i < 3*hg->nt; ++i)
        if (tp[i] < 0) FAIL(11, "triangle %d has no corner list", i/3);
    return 0;
}

static void set_dual_coords(const hgrid_t *hg, const double *lat, const double *lon, double *lat_d, double *lon_d)
{   /* horizontal_generation.c:316-362 */
    #pragma omp parallel for
    for (int d = 0; d < hg->nt; ++d) {
        const int *p = &hg->tri_pts[3*d];
        voronoi_center(lat[p[0]], lon[p[0]], lat[p[1]], lon[p[1]], lat[p[2]], lon[p[2]], &lat_d[d], &lon_d[d]);
    }
}

static int find_adjacent_vector_indices(const hgrid_t *hg, const int *from, const int *to, int *signs, int *adj)
{   /* index.c:20-110; one pass over the edges in ascending order fills each
       cell's slots in the order the reference's per-cell scan finds them */
    int *count = calloc((size_t)hg->nsh, sizeof(int));
    for (int j = 0; j < hg->nvh; ++j) {
        if (from[j] == to[j]) { free(count); FAIL(12, "from_index == to_index at edge %d", j); }
        int c[2] = {from[j], to[j]};
        for (int m = 0; m < 2; ++m) {
            int i = c[m];
            if (count[i] >= 6) { free(count); FAIL(13, "cell %d has more than 6 edges", i); }
            adj[6*i + count[i]] = j;
            signs[6*i + count[i]] = (m == 0) ? 1 : -1;
            ++count[i];
        }
    }
    for (int i = 0; i < hg->nsh; ++i) {
        int expect = (i < N_PENT) ? 5 : 6;
        if (count[i] != expect) { free(count); FAIL(14, "cell %d has %d edges", i, count[i]); }
        if (i < N_PENT) { adj[6*i + 5] = -1; signs[6*i + 5] = 0; }
    }
    free(count);
    return 0;
}

static void cell_vertex_lists(const hgrid_t *hg, const int *adj, const int *from_d, const int *to_d, int *verts)
{   /* vertex gathering of optimize.c:54-79 and coriolis.c:109-134 */
    for (int i = 0; i < hg->nsh; ++i) {
        int ne = (i < N_PENT) ? 5 : 6;
        int *vi = &verts[6*i];
        for (int l = 0; l < 6; ++l) vi[l] = -1;
        int c = 0;
        for (int j = 0; j < ne; ++j) {
            int a = from_d[adj[6*i + j]], b = to_d[adj[6*i + j]];
            if (!in_list(a, vi, ne)) vi[c++] = a;
            if (!in_list(b, vi, ne)) vi[c++] = b;
        }
    }
}

static void find_cell_cgs(const hgrid_t *hg, const int *verts, double *lat, double *lon,
                          const double *lat_d, const double *lon_d)
{   /* optimize.c:33-113 */
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < hg->nsh; ++i) {
        int ne = (i < N_PENT) ? 5 : 6;
        double lv[6], ov[6];
        int srt[6], vr[6];
        for (int j = 0; j < ne; ++j) { lv[j] = lat_d[verts[6*i + j]]; ov[j] = lon_d[verts[6*i + j]]; }
        sort_edge_indices(lv, ov, ne, srt);
        for (int j = 0; j < ne; ++j) vr[j] = verts[6*i + srt[j]];
        double xr = 0, yr = 0, zr = 0;
        for (int j = 0; j < ne; ++j) {
            double la0 = lat[i], lo0 = lon[i];
            double la1 = lat_d[vr[j]], lo1 = lon_d[vr[j]];
            double la2 = lat_d[vr[(j + 1) % ne]], lo2 = lon_d[vr[(j + 1) % ne]];
            double x0, y0, z0, x1, y1, z1, x2, y2, z2;
            find_global_normal(la0, lo0, &x0, &y0, &z0);
            find_global_normal(la1, lo1, &x1, &y1, &z1);
            find_global_normal(la2, lo2, &x2, &y2, &z2);
            double a = triangle_area(la0, lo0, la1, lo1, la2, lo2);
            xr += a*(x0 + x1 + x2);
            yr += a*(y0 + y1 + y2);
            zr += a*(z0 + z1 + z2);
        }
        find_geos(xr, yr, zr, &lat[i], &lon[i]);
    }
}

static int hgrid_setup(hgrid_t *hg, int R)
{
    memset(hg, 0, sizeof *hg);
    hg->R = R;
    long long four_r = 1LL << (2*R);
    hg->nsh = (int)(12 + 10*(four_r - 1));
    hg->nvh = (int)(30 + 30*(four_r - 1));
    hg->nt = (int)(20*four_r);
    int rc = build_icosahedron(&hg->ico);
    if (rc) return rc;
    hg->tri_pts = malloc(sizeof(int)*3*(size_t)hg->nt);
    return set_triangle_points(hg);
}

/* generator coordinates (+ Lloyd iterations), horizontal_generation.c:19-110 and optimize.c:17-31 */
static int horizontal_generators(hgrid_t *hg, int n_iter, const int *adj, const int *from_d, const int *to_d,
                                 double *lat, double *lon, double *lat_d, double *lon_d)
{
    double *x = malloc(sizeof(double)*(size_t)hg->nsh);
    double *y = malloc(sizeof(double)*(size_t)hg->nsh);
    double *z = malloc(sizeof(double)*(size_t)hg->nsh);
    generate_generators(hg, lat, lon, x, y, z);
    free(x); free(y); free(z);
    if (n_iter > 0) {
        int *verts = malloc(sizeof(int)*6*(size_t)hg->nsh);
        cell_vertex_lists(hg, adj, from_d, to_d, verts);
        for (int it = 0; it < n_iter; ++it) {
            set_dual_coords(hg, lat, lon, lat_d, lon_d);
            find_cell_cgs(hg, verts, lat, lon, lat_d, lon_d);
        }
        free(verts);
    }
    return 0;
}

int game_grid_lloyd_state(int res_id, int n_lloyd_iterations, double *latlon_out, double *latlon_dual_out,
                          int *adj_out, int *from_dual_out, int *to_dual_out)
{
    hgrid_t hg;
    int rc = hgrid_setup(&hg, res_id);
    if (rc) return rc;
    int *from = malloc(sizeof(int)*(size_t)hg.nvh), *to = malloc(sizeof(int)*(size_t)hg.nvh);
    int *from_d = malloc(sizeof(int)*(size_t)hg.nvh), *to_d = malloc(sizeof(int)*(size_t)hg.nvh);
    int *adj = malloc(sizeof(int)*6*(size_t)hg.nsh), *sg = malloc(sizeof(int)*6*(size_t)hg.nsh);
    double *lat = malloc(sizeof(double)*(size_t)hg.nsh), *lon = malloc(sizeof(double)*(size_t)hg.nsh);
    double *lat_d = malloc(sizeof(double)*(size_t)hg.nt), *lon_d = malloc(sizeof(double)*(size_t)hg.nt);
    set_from_to_index(&hg, from, to);
    set_from_to_index_dual(&hg, from_d, to_d);
    rc = find_adjacent_vector_indices(&hg, from, to, sg, adj);
    if (!rc) rc = horizontal_generators(&hg, n_lloyd_iterations, adj, from_d, to_d, lat, lon, lat_d, lon_d);
    if (!rc) {
        for (int i = 0; i < hg.nsh; ++i) { latlon_out[2*i] = lat[i]; latlon_out[2*i + 1] = lon[i]; }
        if (latlon_dual_out) {
            /* the Voronoi vertices the next iteration starts from (optimize.c:26) */
            set_dual_coords(&hg, lat, lon, lat_d, lon_d);
            for (int i = 0; i < hg.nt; ++i) { latlon_dual_out[2*i] = lat_d[i]; latlon_dual_out[2*i + 1] = lon_d[i]; }
        }
        if (adj_out) memcpy(adj_out, adj, sizeof(int)*6*(size_t)hg.nsh);
        if (from_dual_out) memcpy(from_dual_out, from_d, sizeof(int)*(size_t)hg.nvh);
        if (to_dual_out) memcpy(to_dual_out, to_d, sizeof(int)*(size_t)hg.nvh);
    }
    free(from); free(to); free(from_d); free(to_d); free(adj); free(sg);
    free(lat); free(lon); free(lat_d); free(lon_d); free(hg.tri_pts);
    return rc;
}

int game_grid_generators(int res_id, int n_lloyd_iterations, double *latlon_out)
{
    return game_grid_lloyd_state(res_id, n_lloyd_iterations, latlon_out, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------ */
/* vertical grid and metric: grid_generator/src/vertical_grid.c          */
/* ------------------------------------------------------------------ */
static const double T_SFC = 273.15 + 15;
static const double TEMP_GRADIENT = -0.65/100;
static const double P_0_STANDARD = 101325;
static const double TROPO_HEIGHT_STANDARD = 11e3;
static const double INVERSE_HEIGHT_STANDARD = 20e3;
static const double TEMP_GRADIENT_INV_STANDARD = 0.1/100;

static double standard_temp(double z)
{   /* vertical_grid.c:350-369 */
    const double tropo = T_SFC + TROPO_HEIGHT_STANDARD*TEMP_GRADIENT;
    if (z < TROPO_HEIGHT_STANDARD) return T_SFC + z*TEMP_GRADIENT;
    if (z < INVERSE_HEIGHT_STANDARD) return tropo;
    return tropo + TEMP_GRADIENT_INV_STANDARD*(z - INVERSE_HEIGHT_STANDARD);
}

static double standard_pres(double z)
{   /* vertical_grid.c:371-394 */
    const double tropo = T_SFC + TROPO_HEIGHT_STANDARD*TEMP_GRADIENT;
    if (z < TROPO_HEIGHT_STANDARD)
        return P_0_STANDARD*pow(1 + TEMP_GRADIENT*z/T_SFC, -G_MEAN/(R_D_REF*TEMP_GRADIENT));
    if (z < INVERSE_HEIGHT_STANDARD)
        return P_0_STANDARD*pow(1 + TEMP_GRADIENT*TROPO_HEIGHT_STANDARD/T_SFC, -G_MEAN/(R_D_REF*TEMP_GRADIENT))
               *exp(-G_MEAN*(z - TROPO_HEIGHT_STANDARD)/(R_D_REF*tropo));
    double p_inv = P_0_STANDARD*pow(1 + TEMP_GRADIENT*TROPO_HEIGHT_STANDARD/T_SFC, -G_MEAN/(R_D_REF*TEMP_GRADIENT))
                   *exp(-G_MEAN*(INVERSE_HEIGHT_STANDARD - TROPO_HEIGHT_STANDARD)/(R_D_REF*tropo));
    return p_inv*pow(1 + TEMP_GRADIENT*(z - INVERSE_HEIGHT_STANDARD)/T_SFC, -G_MEAN/(R_D_REF*TEMP_GRADIENT));
}

double game_standard_temp(double z) { return standard_temp(z); }
double game_standard_pres(double z) { return standard_pres(z); }

/* ------------------------------------------------------------------ */
/* host gradient for the derived metrics (src/spatial_operators/        */
/* gradient_operators.c:14-74, averaging.c:109-134, 208-228)            */
/* ------------------------------------------------------------------ */
static void host_grad_hor_cov(const game_dims_t *d, const game_grid_view_t *g, const double *in, double *out)
{
    int nsh = d->nsh, vpl = d->vpl;
    #pragma omp parallel for
    for (int e = 0; e < d->nvh; ++e)
        for (int l = 0; l < d->n_layers; ++l) {
            long long vi = nsh + (long long)l*vpl + e;
            out[vi] = (in[g->to_index[e] + (long long)l*nsh] - in[g->from_index[e] + (long long)l*nsh])/g->normal_distance[vi];
        }
}

static void host_grad(const game_dims_t *d, const game_grid_view_t *g, const double *in, double *out)
{
    int nsh = d->nsh, vpl = d->vpl, L = d->n_layers;
    host_grad_hor_cov(d, g, in, out);
    #pragma omp parallel for
    for (long long i = nsh; i < (long long)d->n_levels*nsh - nsh; ++i) {
        int l = (int)(i/nsh), h = (int)(i - (long long)l*nsh);
        long long vi = h + (long long)l*vpl;
        out[vi] = (in[h + (long long)(l - 1)*nsh] - in[h + (long long)l*nsh])/g->normal_distance[vi];
    }
    int n_oro = g->no_of_oro_layers;
    #pragma omp parallel for
    for (long long i = 0; i < (long long)n_oro*d->nvh; ++i) {
        int lo = (int)(i/d->nvh), e = (int)(i - (long long)lo*d->nvh);
        int l = lo + (L - n_oro);
        int f = g->from_index[e], t = g->to_index[e];
        const double *ipw = g->inner_product_weights;
        double c = ipw[8*((long long)l*nsh + f) + 6]*out[(long long)l*vpl + f];
        c += ipw[8*((long long)l*nsh + t) + 6]*out[(long long)l*vpl + t];
        if (l < L - 1) {
            c += ipw[8*((long long)l*nsh + f) + 7]*out[(long long)(l + 1)*vpl + f];
            c += ipw[8*((long long)l*nsh + t) + 7]*out[(long long)(l + 1)*vpl + t];
        }
        c = 0.5*c;
        long long vi = nsh + (long long)l*vpl + e;
        out[vi] += -g->slope[vi]*c;
    }
}

/* ------------------------------------------------------------------ */
/* TRSK weights and Gassmann (2018) resorting, coriolis.c:16-438          */
/* ------------------------------------------------------------------ */
static int coriolis(const game_dims_t *dm, game_grid_view_t *g, game_dualgrid_view_t *dg,
                    const double *lat_v, const double *lon_v, const double *lat_d, const double *lon_d, double tol)
{
    int nsh = dm->nsh, nvh = dm->nvh;
    const int *from = g->from_index, *to = g->to_index, *adj = g->adjacent_vector_indices_h;
    const int *from_d = dg->from_index, *to_d = dg->to_index;
    const double *nd = g->normal_distance, *nd_dual = dg->normal_distance, *area = g->area;
    const double *zv = g->z_vector;
    double radius = g->radius;
    double rescale_1d = (radius + g->z_scalar[0])/(radius + zv[0]);
    double rescale_2d = pow(rescale_1d, 2);
    int *ti = g->trsk_indices, *tmc = g->trsk_modified_curl_indices;
    double *tw = g->trsk_weights;
    int err = 0;
    double bad_ratio = 0.0;   /* the failing sum ratio of check 30, for the message */
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < nvh; ++i) {
        if (GET_ERR(err)) continue;
        int offset = 0;
        for (int k = 0; k < 10; ++k) {
            if (k == 0 || k == 5) offset = 0;
            int index_offset, sign_0, cell;
            if (k < 5) { index_offset = 0; sign_0 = -1; cell = from[i]; }
            else { index_offset = 5; sign_0 = 1; cell = to[i]; }
            if (adj[6*cell + k - index_offset] == i) offset += 1;
            if (offset > 1) { SET_ERR(err, 21); break; }
            ti[10*i + k] = adj[6*cell + k - index_offset + offset];
            if (ti[10*i + k] == -1) { tw[10*i + k] = 0; continue; }
            int sign_1 = (from[ti[10*i + k]] == cell) ? 1 : -1;
            int ne = (cell < N_PENT) ? 5 : 6;
            int vi[6], ei[6], srt[6], vr[6];
            double lv[6], ov[6], le[6], oe[6], areas[6];
            for (int l = 0; l < ne; ++l) vi[l] = -1;
            int counter = 0;
            for (int l = 0; l < ne; ++l) {
                int a = from_d[adj[6*cell + l]], b = to_d[adj[6*cell + l]];
                if (!in_list(a, vi, ne)) { vi[counter] = a; lv[counter] = lat_d[a]; ov[counter] = lon_d[a]; ++counter; }
                if (!in_list(b, vi, ne)) { vi[counter] = b; lv[counter] = lat_d[b]; ov[counter] = lon_d[b]; ++counter; }
            }
            if (counter != ne) { SET_ERR(err, 22); break; }
            sort_edge_indices(lv, ov, ne, srt);
            for (int l = 0; l < ne; ++l) vr[l] = vi[srt[l]];
            for (int l = 0; l < ne; ++l)
                for (int m = 0; m < ne; ++m) {
                    int em = adj[6*cell + m];
                    if ((from_d[em] == vr[l] && to_d[em] == vr[(l + 1) % ne])
                        || (to_d[em] == vr[l] && from_d[em] == vr[(l + 1) % ne]))
                        ei[l] = em;
                }
            for (int l = 0; l < ne; ++l) { le[l] = lat_v[ei[l]]; oe[l] = lon_v[ei[l]]; }
            double check_sum = 0;
            for (int l = 0; l < ne; ++l) {
                double t0;
                if (l == 0)
                    t0 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lv[srt[l]], ov[srt[l]], le[ne - 1], oe[ne - 1]);
                else
                    t0 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lv[srt[l]], ov[srt[l]], le[l - 1], oe[l - 1]);
                double t1 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lv[srt[l]], ov[srt[l]], le[l], oe[l]);
                areas[l] = pow(radius + zv[nsh + i], 2)*(t0 + t1);
                check_sum += areas[l];
            }
            if (fabs(check_sum/(rescale_2d*area[cell]) - 1) > tol) {
                #pragma omp critical(gridgen_trsk_ratio)
                bad_ratio = check_sum/(rescale_2d*area[cell]);
                SET_ERR(err, 30);
                break;
            }
            int first_index = -1, last_index = -1;
            for (int l = 0; l < ne; ++l) {
                if (ei[l] == i) last_index = l;
                if (ei[l] == ti[10*i + k]) first_index = (l + 1) % ne;
            }
            double sum = 0.0;   /* index_helpers.c:98-123 */
            if (first_index <= last_index) {
                for (int l = first_index; l <= last_index; ++l) sum += areas[l];
            } else {
                for (int l = first_index; l < ne; ++l) sum += areas[l];
                for (int l = 0; l <= last_index; ++l) sum += areas[l];
            }
            sum = sum/(rescale_2d*area[cell]);
            if (sum < 0 || sum > 1) { SET_ERR(err, 34); break; }
            tw[10*i + k] = sign_0*(sum - 0.5)*sign_1;
            tw[10*i + k] = -rescale_1d*nd_dual[ti[10*i + k]]/nd[nsh + i]*tw[10*i + k];
        }
        if (GET_ERR(err)) continue;
        /* Gassmann (2018) resorting, coriolis.c:223-378 */
        int pre_i[10];
        double pre_w[10];
        for (int j = 0; j < 10; ++j) { pre_i[j] = ti[10*i + j]; pre_w[j] = tw[10*i + j]; }
        for (int side = 0; side < 2; ++side) {
            int cell = side == 0 ? from[i] : to[i];
            int ne = (cell < N_PENT) ? 5 : 6;
            int base = side == 0 ? 0 : 5;
            int next_v = side == 0 ? to_d[i] : from_d[i];
            int used[5] = {-1, -1, -1, -1, -1}, used_n = 0;
            for (int j = 0; j < ne - 1; ++j) {
                int written = 0;
                for (int k = 0; k < ne - 1; ++k) {
                    int e = pre_i[base + k];
                    if ((from_d[e] == next_v || to_d[e] == next_v) && !in_list(k, used, ne - 1) && !written) {
                        ti[10*i + base + j] = e;
                        tw[10*i + base + j] = pre_w[base + k];
                        used[used_n++] = k;
                        written = 1;
                    }
                }
                int cand = to_d[ti[10*i + base + j]];
                next_v = (cand == next_v) ? from_d[ti[10*i + base + j]] : cand;
            }
            if (used_n != ne - 1) { err = 42 + side; break; }
        }
        if (GET_ERR(err)) continue;
        int *m = &tmc[10*i];
        int *t = &ti[10*i];
        m[0] = (to[i] < N_PENT) ? t[8] : t[9];
        m[1] = t[0];
        if (from[i] < N_PENT) {
            m[2] = t[3]; m[3] = t[5]; m[4] = 0;
            if (tw[10*i + 4] != 0) { SET_ERR(err, 40); continue; }
        } else { m[2] = t[2]; m[3] = t[4]; m[4] = t[5]; }
        m[5] = (from[i] < N_PENT) ? t[3] : t[4];
        m[6] = t[5];
        if (to[i] < N_PENT) {
            m[7] = t[8]; m[8] = t[0]; m[9] = 0;
            if (tw[10*i + 9] != 0) { SET_ERR(err, 41); continue; }
        } else { m[7] = t[7]; m[8] = t[9]; m[9] = t[0]; }
        for (int j = 0; j < 10; ++j)
            for (int k = j + 1; k < 10; ++k)
                if (t[j] == t[k] && tw[10*i + j] != 0 && tw[10*i + k] != 0) SET_ERR(err, 29);
    }
    if (err == 30)
        FAIL(err, "TRSK construction failed (reference problem code 30: kite areas / cell area = %.17g, "
             "bound %g)", bad_ratio, tol);
    if (err) FAIL(err, "TRSK construction failed (reference problem code %d)", err);
    /* Eq. (39) of Thuburn et al. (2009), coriolis.c:393-427 */
    #pragma omp parallel for
    for (int i = 0; i < nvh; ++i) {
        for (int j = 0; j < 10; ++j) {
            int fi = ti[10*i + j];
            if (fi == -1) continue;
            double v0 = nd[nsh + i]/(rescale_1d*nd_dual[fi])*tw[10*i + j];
            int second = -1;
            for (int k = 0; k < 10; ++k) if (ti[10*fi + k] == i) second = 10*fi + k;
            if (second == -1) { SET_ERR(err, 38); continue; }
            double v1 = nd[nsh + fi]/(rescale_1d*nd_dual[i])*tw[second];
            if (fabs(v0 + v1) > tol) SET_ERR(err, 39);
        }
    }
    if (err) FAIL(err, "TRSK antisymmetry check failed (reference problem code %d)", err);
    for (int i = 0; i < 10*nvh; ++i) if (ti[i] == -1) ti[i] = 0;
    return 0;
}

/* rhombus_averaging.c:16-269 */
static int rhombus_averaging(const game_dims_t *dm, game_grid_view_t *g, game_dualgrid_view_t *dg,
                             const double *lat_v, const double *lon_v, const double *lat_d, const double *lon_d,
                             int *vir, double rh_tol)
{
    int nvh = dm->nvh, nsh = dm->nsh;
    const int *vit = dg->vorticity_indices_triangles, *from_d = dg->from_index, *to_d = dg->to_index;
    const int *from = g->from_index, *to = g->to_index;
    int *dri = g->density_to_rhombi_indices;
    double *drw = g->density_to_rhombi_weights;
    double radius = g->radius;
    int err = 0;
    #pragma omp parallel for
    for (int i = 0; i < nvh; ++i) {
        if (GET_ERR(err)) continue;
        int pre[6], list[4], dbl[2] = {-1, -1};
        for (int j = 0; j < 3; ++j) pre[j] = vit[3*to_d[i] + j];
        for (int j = 0; j < 3; ++j) pre[3 + j] = vit[3*from_d[i] + j];
        for (int j = 0; j < 6; ++j)
            for (int k = j + 1; k < 6; ++k)
                if (pre[j] == pre[k]) { dbl[0] = j; dbl[1] = k; }
        int c = 0;
        for (int j = 0; j < 6; ++j)
            if (j != dbl[0] && j != dbl[1]) { if (c < 4) list[c] = pre[j]; c++; }
        if (c != 4) { SET_ERR(err, 51); continue; }
        for (int j = 0; j < 4; ++j) vir[4*i + j] = list[j];
        int dpre[4] = {-1, -1, -1, -1}, cc = 0;
        for (int j = 0; j < 4; ++j) {
            int cand = from[vir[4*i + j]];
            if (!in_list(cand, dpre, 4)) { if (cc < 4) dpre[cc] = cand; ++cc; }
            cand = to[vir[4*i + j]];
            if (!in_list(cand, dpre, 4)) { if (cc < 4) dpre[cc] = cand; ++cc; }
        }
        if (cc != 4) { SET_ERR(err, 52); continue; }
        for (int j = 0; j < 4; ++j) dri[4*i + j] = dpre[j];
        double rh_area = dg->area[nvh + from_d[i]] + dg->area[nvh + to_d[i]];
        int first_case = 0, second_case = 0;
        for (int j = 0; j < 4; ++j) {
            int cell = dri[4*i + j];
            int k = 0, v0, v1;
            while (!(from[vir[4*i + k]] == cell || to[vir[4*i + k]] == cell)) ++k;
            v0 = vir[4*i + k];
            int ds0 = from_d[v0], which = 1;
            for (int l = 0; l < 4; ++l)
                if (l != k && (from_d[vir[4*i + l]] == ds0 || to_d[vir[4*i + l]] == ds0)) which = 0;
            if (which == 1) ds0 = to_d[v0];
            double t0 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lat_d[ds0], lon_d[ds0], lat_v[v0], lon_v[v0]);
            k = 0;
            while (!((from[vir[4*i + k]] == cell || to[vir[4*i + k]] == cell) && vir[4*i + k] != v0)) ++k;
            v1 = vir[4*i + k];
            if (cell == from[i] || cell == to[i]) {
                first_case++;
                double t1 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lat_d[ds0], lon_d[ds0], lat_v[i], lon_v[i]);
                int ds1 = from_d[v1];
                which = 1;
                for (int l = 0; l < 4; ++l)
                    if (l != k && (from_d[vir[4*i + l]] == ds1 || to_d[vir[4*i + l]] == ds1)) which = 0;
                if (which == 1) ds1 = to_d[v1];
                double t2 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lat_d[ds1], lon_d[ds1], lat_v[i], lon_v[i]);
                double t3 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lat_d[ds1], lon_d[ds1], lat_v[v1], lon_v[v1]);
                drw[4*i + j] = pow(radius + g->z_vector[nsh], 2)*(t0 + t1 + t2 + t3)/rh_area;
            } else {
                second_case++;
                double t1 = triangle_area(g->latitude_scalar[cell], g->longitude_scalar[cell], lat_d[ds0], lon_d[ds0], lat_v[v1], lon_v[v1]);
                drw[4*i + j] = pow(radius + g->z_vector[nsh], 2)*(t0 + t1)/rh_area;
            }
        }
        if (first_case != 2 || second_case != 2) { SET_ERR(err, 53); continue; }
        double check = 0;
        for (int j = 0; j < 4; ++j) {
            check += drw[4*i + j];
            if (drw[4*i + j] <= 0 || drw[4*i + j] >= 1) SET_ERR(err, 54);
        }
        if (fabs(check - 1) > rh_tol) { SET_ERR(err, 55); fprintf(stderr, "rhombus %d sum %.17g w %g %g %g %g\n", i, check, drw[4*i], drw[4*i+1], drw[4*i+2], drw[4*i+3]); }
    }
    if (err) FAIL(err, "rhombus weights failed (code %d)", err);
    return 0;
}

/* the metrics the model derives from a grid file, src/io/set_grid_properties.c:196-221 */
void game_grid_derived_metrics(const game_dims_t *dm, game_grid_view_t *g)
{
    const int nsh = dm->nsh, vpl = dm->vpl;
    for (int i = 0; i < 6*nsh; ++i) if (g->adjacent_vector_indices_h[i] == -1) g->adjacent_vector_indices_h[i] = 0;
    for (long long i = 0; i < dm->s; ++i) {
        int l = (int)(i/nsh), h = (int)(i - (long long)l*nsh);
        g->layer_thickness[i] = g->z_vector[h + (long long)l*vpl] - g->z_vector[h + (long long)(l + 1)*vpl];
    }
    host_grad_hor_cov(dm, g, g->z_scalar, g->slope);
    host_grad(dm, g, g->gravity_potential, g->gravity_m);
    host_grad(dm, g, g->exner_bg, g->exner_bg_grad);
}

/* ------------------------------------------------------------------ */
/* lat-lon output interpolation, grid_generator/src/interpolate_ll.c     */
/* ------------------------------------------------------------------ */
/* The reference scans every cell for each point and picks five times
   (index_helpers.c:49-67): the pick starts from cell 0 with its distance as
   the running minimum, without testing whether cell 0 is excluded, and
   replaces it by each later non-excluded cell with a strictly smaller
   distance.  The nearest non-excluded cells lie within two rings of the
   point's nearest cell, which a greedy walk over the cell adjacency finds
   (the adjacency is the Delaunay graph of the generators), so the pick is
   made over those candidates and cell 0. */
int game_latlon_interpolation(int res_id, const double *lat, const double *lon, const int *adj, const int *from,
                              const int *to, int *out_idx, double *out_w)
{
    if (res_id < 0 || res_id > 12 || !lat || !lon || !adj || !from || !to || !out_idx || !out_w) {
        snprintf(g_err, sizeof g_err, "bad arguments");
        return 1;
    }
    const int nsh = 12 + 10*((1 << (2*res_id)) - 1);
    const int n_lon = 4 << res_id, n_lat = 2 << res_id;
    const double delta_latitude = M_PI/n_lat, delta_longitude = 2*M_PI/n_lon;
    int err = 0;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int la = 0; la < n_lat; ++la) {
        int cur = 0;
        for (int lo = 0; lo < n_lon; ++lo) {
            const long long i = (long long)la*n_lon + lo;
            const double lat_value = M_PI/2 - 0.5*delta_latitude - la*delta_latitude;
            const double lon_value = lo*delta_longitude;
            /* greedy walk to the nearest cell, from the previous point's */
            double dc = dist_h(lat_value, lon_value, lat[cur], lon[cur], 1);
            for (int moved = 1; moved;) {
                moved = 0;
                const int ne = cur < N_PENT ? 5 : 6;
                for (int j = 0; j < ne; ++j) {
                    const int e = adj[6*cur + j], nb = from[e] == cur ? to[e] : from[e];
                    const double d = dist_h(lat_value, lon_value, lat[nb], lon[nb], 1);
                    if (d < dc) { dc = d; cur = nb; moved = 1; }
                }
            }
            /* candidates: the nearest cell and two rings around it */
            int cand[1 + 6 + 36];
            int nc = 0;
            cand[nc++] = cur;
            for (int ring = 0, lo_i = 0; ring < 2; ++ring) {
                const int hi_i = nc;
                for (int k = lo_i; k < hi_i; ++k) {
                    const int c = cand[k], ne = c < N_PENT ? 5 : 6;
                    for (int j = 0; j < ne; ++j) {
                        const int e = adj[6*c + j], nb = from[e] == c ? to[e] : from[e];
                        int seen = 0;
                        for (int q = 0; q < nc; ++q) if (cand[q] == nb) { seen = 1; break; }
                        if (!seen && nc < 43) cand[nc++] = nb;
                    }
                }
                lo_i = hi_i;
            }
            double cd[43];
            for (int q = 0; q < nc; ++q) cd[q] = dist_h(lat_value, lon_value, lat[cand[q]], lon[cand[q]], 1);
            const double d0 = dist_h(lat_value, lon_value, lat[0], lon[0], 1);
            int picked[5] = {-1, -1, -1, -1, -1};
            double w[5], wsum = 0;
            for (int p = 0; p < 5; ++p) {
                int best = 0;
                double bmin = d0;
                /* the smallest index among equal distances: the reference's
                   scan keeps the first of equal strict minima */
                for (int q = 0; q < nc; ++q) {
                    const int c = cand[q];
                    if (c == 0) continue;
                    int excl = 0;
                    for (int r = 0; r < 5; ++r) if (picked[r] == c) excl = 1;
                    if (excl) continue;
                    if (cd[q] < bmin || (cd[q] == bmin && best != 0 && c < best)) { bmin = cd[q]; best = c; }
                }
                picked[p] = best;
                const double d = best == 0 ? d0 : bmin;
                w[p] = 1/(pow(d, 2 + EPS_SECURITY) + EPS_SECURITY);
                wsum += w[p];
            }
            for (int p = 0; p < 5; ++p) {
                if (picked[p] < 0 || picked[p] >= nsh) err = 1;
                out_idx[5*i + p] = picked[p];
                out_w[5*i + p] = w[p]/wsum;
            }
        }
    }
    if (err) { snprintf(g_err, sizeof g_err, "interpolation index out of range"); return 2; }
    return 0;
}

/* grid_generator/src/phys_sfc_properties.c:150-198 for cells that are not land */
int game_sfc_properties(int nsh, const int *is_land, double *albedo, double *rho_c, double *t_cond, double *rough)
{
    const double c_p_water = 4184.0, albedo_water = 0.06, t_conductivity_water = 1.4e-7, density_water = 1024.0;
    for (int i = 0; i < nsh; ++i) {
        if (is_land && is_land[i]) {
            snprintf(g_err, sizeof g_err, "cell %d is land: no land surface properties are generated here", i);
            return 1;
        }
        albedo[i] = albedo_water;
        rho_c[i] = density_water*c_p_water;
        rough[i] = fmax(0.0001, 0.08);
        t_cond[i] = t_conductivity_water;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* the full generator, grid_generator/src/coordinator.c:178-352          */
/* ------------------------------------------------------------------ */
int game_grid_generate(const game_grid_params_t *P, const game_dims_t *dm,
                       game_grid_view_t *g, game_dualgrid_view_t *dg, game_grid_extras_t *X)
{
    g_err[0] = 0;
    if (dm->res_id != P->res_id || dm->n_layers != P->n_layers || dm->n_oro_layers != P->n_oro_layers)
        FAIL(1, "dims do not match the parameters");
    if (P->n_oro_layers < 0 || P->n_oro_layers >= P->n_layers) FAIL(2, "it must be 0 <= orography_layers < NO_OF_LAYERS");
    if (P->stretching_parameter < 1) FAIL(3, "stretching_parameter must be >= 1");
    hgrid_t hg;
    int rc = hgrid_setup(&hg, P->res_id);
    if (rc) return rc;
    int nsh = dm->nsh, nvh = dm->nvh, nt = dm->nt, L = dm->n_layers, vpl = dm->vpl, dvpl = dm->dvpl;
    double radius = P->radius_rescale*RADIUS_EARTH;
    g->radius = radius;
    g->toa = P->toa;
    g->no_of_oro_layers = P->n_oro_layers;
    g->stretching_parameter = P->stretching_parameter;
    g->oro_id = P->oro_id;
    double *lat = g->latitude_scalar, *lon = g->longitude_scalar;
    double *lat_d = X->latitude_scalar_dual, *lon_d = X->longitude_scalar_dual;

    set_from_to_index(&hg, g->from_index, g->to_index);
    set_from_to_index_dual(&hg, dg->from_index, dg->to_index);
    rc = find_adjacent_vector_indices(&hg, g->from_index, g->to_index, g->adjacent_signs_h, g->adjacent_vector_indices_h);
    if (rc) goto done;
    if (P->generator_latlon) {
        for (int i = 0; i < nsh; ++i) { lat[i] = P->generator_latlon[2*i]; lon[i] = P->generator_latlon[2*i + 1]; }
    } else {
        rc = horizontal_generators(&hg, P->n_lloyd_iterations, g->adjacent_vector_indices_h, dg->from_index,
                                   dg->to_index, lat, lon, lat_d, lon_d);
        if (rc) goto done;
    }
    set_dual_coords(&hg, lat, lon, lat_d, lon_d);

    /* set_vector_h_doubles, horizontal_generation.c:241-260 */
    #pragma omp parallel for
    for (int i = 0; i < nvh; ++i) {
        int f = g->from_index[i], t = g->to_index[i];
        double x0, y0, z0, x1, y1, z1;
        find_global_normal(lat[f], lon[f], &x0, &y0, &z0);
        find_global_normal(lat[t], lon[t], &x1, &y1, &z1);
        double xr = x0 + 0.5*(x1 - x0), yr = y0 + 0.5*(y1 - y0), zr = z0 + 0.5*(z1 - z0);
        find_geos(xr, yr, zr, &X->latitude_vector[i], &X->longitude_vector[i]);
        g->direction[i] = geodetic_direction(lat[f], lon[f], lat[t], lon[t], 0.5);
    }
    /* set_dual_vector_h_doubles + direct_tangential_unity, horizontal_generation.c:519-576 */
    double *rel_on_line = malloc(sizeof(double)*(size_t)nvh);
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < nvh; ++i) {
        int fd = dg->from_index[i], td = dg->to_index[i];
        rel_on_line[i] = min_dist_rel_on_line(lat_d[fd], lon_d[fd], lat_d[td], lon_d[td],
                                              X->latitude_vector[i], X->longitude_vector[i]);
        X->direction_dual[i] = geodetic_direction(lat_d[fd], lon_d[fd], lat_d[td], lon_d[td], rel_on_line[i]);
    }
    #pragma omp parallel for
    for (int i = 0; i < nvh; ++i) {
        double dc = turn_angle(g->direction[i], X->direction_dual[i]);
        if (rad2deg(dc) < -ORTH_CRITERION_DEG) {
            int tmp = dg->from_index[i];
            dg->from_index[i] = dg->to_index[i];
            dg->to_index[i] = tmp;
            rel_on_line[i] = 1 - rel_on_line[i];
            int fd = dg->from_index[i], td = dg->to_index[i];
            X->direction_dual[i] = geodetic_direction(lat_d[fd], lon_d[fd], lat_d[td], lon_d[td], rel_on_line[i]);
        }
    }
    free(rel_on_line);
    /* the reference only reports non-orthogonal intersections (horizontal_generation.c:565-574) */
    for (int i = 0; i < nvh; ++i) {
        double dc = fabs(rad2deg(turn_angle(g->direction[i], X->direction_dual[i])));
        if (dc < ORTH_CRITERION_DEG || dc > 90 + (90 - ORTH_CRITERION_DEG))
            fprintf(stderr, "Grid non-orthogonal: intersection angle of %lf degrees at edge %d.\n", dc, i);
    }
    /* set_f_vec, miscellaneous.c:21-43 */
    for (int i = 0; i < 2*nvh; ++i) {
        if (i < nvh) dg->f_vec[i] = 2*OMEGA_EARTH/P->radius_rescale*cos(X->latitude_vector[i])*sin(X->direction_dual[i]);
        else dg->f_vec[i] = 2*OMEGA_EARTH/P->radius_rescale*sin(X->latitude_vector[i - nvh]);
    }
    /* calc_triangle_area_unity, horizontal_generation.c:176-239 */
    #pragma omp parallel for
    for (int d = 0; d < nt; ++d) {
        const int *p = &hg.tri_pts[3*d];
        X->triangle_face_unit_sphere[d] = triangle_area(lat[p[0]], lon[p[0]], lat[p[1]], lon[p[1]], lat[p[2]], lon[p[2]]);
    }
    {
        double sum = 0, ideal = 4*M_PI/nt;
        for (int d = 0; d < nt; ++d) {
            sum += X->triangle_face_unit_sphere[d];
            if (X->triangle_face_unit_sphere[d] <= 0 || fabs(X->triangle_face_unit_sphere[d]/ideal - 1) > 0.4) { rc = 61; snprintf(g_err, sizeof g_err, "bad triangle area %d", d); goto done; }
        }
        if (fabs(sum/(4*M_PI) - 1) > EPS_SECURITY) { rc = 62; snprintf(g_err, sizeof g_err, "triangle areas do not sum to 4 pi"); goto done; }
    }
    /* calc_vorticity_indices_triangles, miscellaneous.c:45-90: edges in ascending order */
    {
        int *cnt = calloc((size_t)nt, sizeof(int));
        int *vit = dg->vorticity_indices_triangles, *vst = dg->vorticity_signs_triangles;
        for (int j = 0; j < nvh; ++j) {
            int tri[2] = {dg->from_index[j], dg->to_index[j]};
            for (int m = 0; m < 2; ++m) {
                int i = tri[m];
                if (cnt[i] >= 3) { rc = 63; break; }
                int sign = 1;
                double dc = turn_angle(X->direction_dual[j], g->direction[j]);
                if (m == 0 && rad2deg(dc) < -ORTH_CRITERION_DEG) sign = -1;
                if (m == 1 && rad2deg(dc) > ORTH_CRITERION_DEG) sign = -1;
                vit[3*i + cnt[i]] = j;
                vst[3*i + cnt[i]] = sign;
                cnt[i]++;
            }
            if (rc) break;
        }
        for (int i = 0; i < nt && !rc; ++i) if (cnt[i] != 3) rc = 63;
        free(cnt);
        if (rc) { snprintf(g_err, sizeof g_err, "triangle edge lists broken"); goto done; }
    }
    /* calc_cell_area_unity, horizontal_generation.c:112-174: triangles in ascending order */
    {
        int bad = 0;
        #pragma omp parallel for
        for (int i = 0; i < nsh; ++i) {
            int ne = (i < N_PENT) ? 5 : 6, tri[12], n = 0;
            for (int j = 0; j < ne; ++j) {
                int e = g->adjacent_vector_indices_h[6*i + j];
                int a = dg->from_index[e], b = dg->to_index[e];
                if (!in_list(a, tri, n)) tri[n++] = a;
                if (!in_list(b, tri, n)) tri[n++] = b;
            }
            if (n != ne) { bad = 1; continue; }
            for (int a = 1; a < n; ++a)
                for (int b = a; b > 0 && tri[b - 1] > tri[b]; --b) { int t = tri[b]; tri[b] = tri[b - 1]; tri[b - 1] = t; }
            double la[6], lo[6];
            for (int j = 0; j < n; ++j) { la[j] = lat_d[tri[j]]; lo[j] = lon_d[tri[j]]; }
            X->pent_hex_face_unity_sphere[i] = spherical_polygon_area(la, lo, ne);
        }
        if (bad) { rc = 64; snprintf(g_err, sizeof g_err, "cell corner search failed"); goto done; }
        double sum = 0, ideal = 4*M_PI/nsh;
        for (int i = 0; i < nsh; ++i) {
            sum += X->pent_hex_face_unity_sphere[i];
            if (X->pent_hex_face_unity_sphere[i] <= 0 || fabs(X->pent_hex_face_unity_sphere[i]/ideal - 1) > 0.4) { rc = 65; snprintf(g_err, sizeof g_err, "bad cell area %d", i); goto done; }
        }
        if (fabs(sum/(4*M_PI) - 1) > EPS_SECURITY) { rc = 66; snprintf(g_err, sizeof g_err, "cell areas do not sum to 4 pi"); goto done; }
    }
    /* orography: flat, or the analytic Gaussian mountain */
    for (int i = 0; i < nsh; ++i) {
        X->oro[i] = 0.0;
        if (P->oro_id == 2) {
            double dist = dist_h(lat[i], lon[i], P->oro_lat, P->oro_lon, radius);
            X->oro[i] = P->oro_h0*exp(-pow(dist/P->oro_half_width, 2));
        }
        g->is_land[i] = 0;
    }
    /* set_z_scalar, vertical_grid.c:27-76 */
    {
        double *zpre = malloc(sizeof(double)*(size_t)(L + 1));
        double max_oro = X->oro[max_index(X->oro, nsh)];
        int n_oro = P->n_oro_layers;
        for (int h = 0; h < nsh; ++h) {
            for (int j = 0; j < L + 1; ++j) {
                double z_rel = 1 - (j + 0.0)/L;
                double sigma_z = pow(z_rel, P->stretching_parameter);
                double A = sigma_z*P->toa, B;
                if (j >= L - n_oro) B = (j - (L - n_oro) + 0.0)/n_oro;
                else B = 0;
                zpre[j] = A + B*X->oro[h];
            }
            if (h == 0 && max_oro >= zpre[L - n_oro]) { free(zpre); rc = 70; snprintf(g_err, sizeof g_err, "maximum of orography larger or equal to the height of the lowest flat level"); goto done; }
            for (int l = 0; l < L; ++l) g->z_scalar[(long long)l*nsh + h] = 0.5*(zpre[l] + zpre[l + 1]);
        }
        free(zpre);
    }
    /* set_z_vector_and_normal_distance, vertical_grid.c:78-142 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->v; ++i) {
        int l = (int)(i/vpl), h = (int)(i - (long long)l*vpl);
        if (h >= nsh) {
            int e = h - nsh;
            g->z_vector[i] = 0.5*(g->z_scalar[(long long)l*nsh + g->from_index[e]] + g->z_scalar[(long long)l*nsh + g->to_index[e]]);
            g->normal_distance[i] = dist_h(lat[g->from_index[e]], lon[g->from_index[e]], lat[g->to_index[e]], lon[g->to_index[e]], radius + g->z_vector[i]);
        } else {
            long long up = h + (long long)(l - 1)*nsh, lo = h + (long long)l*nsh;
            if (l == 0) { g->z_vector[i] = P->toa; g->normal_distance[i] = P->toa - g->z_scalar[lo]; }
            else if (l == L) { g->z_vector[i] = X->oro[h]; g->normal_distance[i] = g->z_scalar[up] - g->z_vector[i]; }
            else { g->normal_distance[i] = g->z_scalar[up] - g->z_scalar[lo]; g->z_vector[i] = g->z_scalar[lo] + 0.5*g->normal_distance[i]; }
        }
    }
    /* set_z_scalar_dual (vertical_grid.c:144-166) feeds only the dual normal distances */
    double *z_scalar_dual = malloc(sizeof(double)*(size_t)dm->n_levels*(size_t)nt);
    #pragma omp parallel for
    for (long long i = 0; i < (long long)dm->n_levels*nt; ++i) {
        int l = (int)(i/nt), h = (int)(i - (long long)l*nt);
        const int *vt = &dg->vorticity_indices_triangles[3*h];
        const int *fr = g->from_index, *to = g->to_index;
        long long b = (long long)l*vpl;
        z_scalar_dual[i] = 1.0/6*(g->z_vector[b + fr[vt[0]]] + g->z_vector[b + fr[vt[1]]] + g->z_vector[b + fr[vt[2]]]
                                  + g->z_vector[b + to[vt[0]]] + g->z_vector[b + to[vt[1]]] + g->z_vector[b + to[vt[2]]]);
    }
    /* calc_z_vector_dual_and_normal_distance_dual, vertical_grid.c:266-309 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->dual_v; ++i) {
        int l = (int)(i/dvpl), h = (int)(i - (long long)l*dvpl);
        if (h >= nvh) {
            int t = h - nvh;
            dg->normal_distance[i] = z_scalar_dual[t + (long long)l*nt] - z_scalar_dual[t + (long long)(l + 1)*nt];
            const int *vt = &dg->vorticity_indices_triangles[3*t];
            long long b = nsh + (long long)l*vpl;
            dg->z_vector[i] = 1.0/3*(g->z_vector[b + vt[0]] + g->z_vector[b + vt[1]] + g->z_vector[b + vt[2]]);
        } else {
            if (l == 0) dg->z_vector[i] = P->toa;
            else if (l == L) dg->z_vector[i] = 0.5*(g->z_vector[(long long)L*vpl + g->from_index[h]] + g->z_vector[(long long)L*vpl + g->to_index[h]]);
            else dg->z_vector[i] = 0.5*(g->z_vector[nsh + h + (long long)(l - 1)*vpl] + g->z_vector[nsh + h + (long long)l*vpl]);
            int fd = dg->from_index[h], td = dg->to_index[h];
            dg->normal_distance[i] = dist_h(lat_d[fd], lon_d[fd], lat_d[td], lon_d[td], radius + dg->z_vector[i]);
        }
    }
    free(z_scalar_dual);
    /* set_area, vertical_grid.c:236-264 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->v; ++i) {
        int l = (int)(i/vpl), h = (int)(i - (long long)l*vpl);
        if (h < nsh) g->area[i] = X->pent_hex_face_unity_sphere[h]*pow(radius + g->z_vector[i], 2);
        else {
            long long dvi = (long long)(l + 1)*dvpl + h - nsh;
            double r0 = radius + dg->z_vector[dvi], r1 = radius + dg->z_vector[dvi - dvpl];
            g->area[i] = vertical_area(dg->normal_distance[dvi], r0, r1);
        }
    }
    /* set_area_dual, vertical_grid.c:189-234 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->dual_v; ++i) {
        int l = (int)(i/dvpl), h = (int)(i - (long long)l*dvpl);
        if (h >= nvh) dg->area[i] = pow(radius + dg->z_vector[i], 2)*X->triangle_face_unit_sphere[h - nvh];
        else {
            long long pvi;
            double r0, r1, base;
            if (l == 0) { pvi = nsh + h; r0 = radius + g->z_vector[pvi]; r1 = radius + P->toa; base = g->normal_distance[pvi]; }
            else if (l == L) {
                pvi = nsh + (long long)(L - 1)*vpl + h;
                r0 = radius + 0.5*(g->z_vector[(long long)L*vpl + g->from_index[h]] + g->z_vector[(long long)L*vpl + g->to_index[h]]);
                r1 = radius + g->z_vector[pvi];
                base = g->normal_distance[pvi]*r0/r1;
            } else { pvi = nsh + (long long)l*vpl + h; r0 = radius + g->z_vector[pvi]; r1 = radius + g->z_vector[pvi - vpl]; base = g->normal_distance[pvi]; }
            dg->area[i] = vertical_area(base, r0, r1);
        }
    }
    /* set_volume, vertical_grid.c:168-187 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->s; ++i) {
        int l = (int)(i/nsh), h = (int)(i - (long long)l*nsh);
        double base = g->area[h + (long long)(l + 1)*vpl];
        double r0 = radius + g->z_vector[h + (long long)(l + 1)*vpl];
        double r1 = radius + g->z_vector[h + (long long)l*vpl];
        g->volume[i] = box_volume(base, r0, r1);
    }
    /* set_gravity_potential, gravity_potential.c:13-25 */
    for (long long i = 0; i < dm->s; ++i)
        g->gravity_potential[i] = -G_MEAN*(radius*radius/(radius + g->z_scalar[i]) - radius);
    /* set_background_state, vertical_grid.c:311-348 */
    for (int h = 0; h < nsh; ++h)
        for (int l = L - 1; l >= 0; --l) {
            long long si = (long long)l*nsh + h;
            double temperature = standard_temp(g->z_scalar[si]);
            if (l == L - 1) {
                double pressure = standard_pres(g->z_scalar[si]);
                g->exner_bg[si] = pow(pressure/P_0_REF, R_D_REF/C_D_P_REF);
            } else {
                double b = -0.5*g->exner_bg[si + nsh]/standard_temp(g->z_scalar[si + nsh])
                           *(temperature - standard_temp(g->z_scalar[si + nsh])
                             + 2.0/C_D_P_REF*(g->gravity_potential[si] - g->gravity_potential[si + nsh]));
                double c = pow(g->exner_bg[si + nsh], 2)*temperature/standard_temp(g->z_scalar[si + nsh]);
                g->exner_bg[si] = b + pow((pow(b, 2) + c), 0.5);
            }
            g->theta_v_bg[si] = temperature/g->exner_bg[si];
        }
    /* calc_inner_product, inner_product.c:15-63 */
    #pragma omp parallel for
    for (long long i = 0; i < dm->s; ++i) {
        int l = (int)(i/nsh), h = (int)(i - (long long)l*nsh);
        double *w = &g->inner_product_weights[8*i];
        for (int j = 0; j < 6; ++j) {
            if (j < 5 || h >= N_PENT) {
                long long vi = nsh + (long long)l*vpl + g->adjacent_vector_indices_h[6*h + j];
                w[j] = g->area[vi];
                w[j] = w[j]*g->normal_distance[vi];
                w[j] = w[j]/(2*g->volume[i]);
            } else w[j] = 0;
        }
        double dz = (l == 0) ? 2*(g->z_vector[h] - g->z_scalar[i]) : g->z_scalar[i - nsh] - g->z_scalar[i];
        w[6] = g->area[h + (long long)l*vpl]*dz/(2*g->volume[i]);
        dz = (l == L - 1) ? 2*(g->z_scalar[i] - g->z_vector[(long long)L*vpl + h]) : g->z_scalar[i] - g->z_scalar[i + nsh];
        w[7] = g->area[h + (long long)(l + 1)*vpl]*dz/(2*g->volume[i]);
    }
    rc = rhombus_averaging(dm, g, dg, X->latitude_vector, X->longitude_vector, lat_d, lon_d, X->vorticity_indices_rhombi,
                           P->rhombus_tolerance > 0 ? P->rhombus_tolerance : EPS_SECURITY);
    if (rc) goto done;
    rc = coriolis(dm, g, dg, X->latitude_vector, X->longitude_vector, lat_d, lon_d,
                  P->trsk_tolerance > 0 ? P->trsk_tolerance : EPS_SECURITY);
    if (rc) goto done;
    game_grid_derived_metrics(dm, g);
done:
    free(hg.tri_pts);
    return rc;
}
