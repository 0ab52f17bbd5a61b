/*
 * oracle_rkhevi.c -- CPU oracle of GAME's RKHEVI step.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never linked by the product.
 *
 * Each function restates the reference function cited in its comment, loop
 * by loop, with run-time sizes (src/game_types.h:12-60 evaluated at run
 * time) and the same floating-point operation order, so that with
 * `gcc -O2 -ffp-contract=off` (the reference's own flags, CMakeLists.txt:70-71)
 * it reproduces the reference bit for bit; tests/test_oracle_vs_reference.py
 * checks that against the reference's objects built in oracle/_ref.
 * Only the switches of the measured workloads are supported: all diffusion,
 * radiation, PBL and surface fluxes off (the reference's flag-guarded hooks
 * are rejected by or_manage_rkhevi when on).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle_rkhevi.h"

/* src/game_constants.h:10-27 */
#define T_0 273.15
#define K_B (1.380649e-23)
#define N_A (6.02214076e23)
#define P_0 100000.0
#define C_D_V 717.942189
#define C_V_V 1396.475121
#define C_D_P 1005.0
#define C_V_P 1858.0
#define M_D (N_A*0.004810e-23)
#define M_V (N_A*0.002991e-23)
#define R_D 287.057811
#define R_V 461.524879
#define N_PENT 12

#define NSH (c->d.nsh)
#define NVH (c->d.nvh)
#define NT (c->d.nt)
#define VPL (c->d.vpl)
#define DVPL (c->d.dvpl)
#define NL (c->d.n_layers)
#define NS (c->d.s)
#define NV (c->d.v)
#define NORO (c->g.no_of_oro_layers)
#define NCC (c->d.n_cond)
#define NCONST (c->d.n_const)

/* ------------------------------------------------------------------ */
/* gradient_operators.c                                                 */
/* ------------------------------------------------------------------ */
void or_grad_hor_cov(const or_ctx_t *c, const double *in, double *out)
{   /* gradient_operators.c:14-33 */
    #pragma omp parallel for
    for (int h = 0; h < NVH; ++h)
        for (int l = 0; l < NL; ++l) {
            long long vi = NSH + (long long)l*VPL + h;
            out[vi] = (in[c->g.to_index[h] + (long long)l*NSH] - in[c->g.from_index[h] + (long long)l*NSH])
                      /c->g.normal_distance[vi];
        }
}

void or_grad_vert_cov(const or_ctx_t *c, const double *in, double *out)
{   /* gradient_operators.c:35-54 */
    long long n_v_vectors = (long long)(NL + 1)*NSH;
    #pragma omp parallel for
    for (long long i = NSH; i < n_v_vectors - NSH; ++i) {
        int l = (int)(i/NSH), h = (int)(i - (long long)l*NSH);
        long long lower = h + (long long)l*NSH, upper = h + (long long)(l - 1)*NSH, vi = h + (long long)l*VPL;
        out[vi] = (in[upper] - in[lower])/c->g.normal_distance[vi];
    }
}

double or_remap_verpri2horpri_vector(const or_ctx_t *c, const double *f, int l, int h)
{   /* averaging.c:14-39 */
    const double *ipw = c->g.inner_product_weights;
    int fr = c->g.from_index[h], to = c->g.to_index[h];
    double comp = ipw[8*((long long)l*NSH + fr) + 6]*f[(long long)l*VPL + fr];
    comp += ipw[8*((long long)l*NSH + to) + 6]*f[(long long)l*VPL + to];
    if (l < NL - 1) {
        comp += ipw[8*((long long)l*NSH + fr) + 7]*f[(long long)(l + 1)*VPL + fr];
        comp += ipw[8*((long long)l*NSH + to) + 7]*f[(long long)(l + 1)*VPL + to];
    }
    comp = 0.5*comp;
    return comp;
}

void or_vector_field_hor_cov_to_con(const or_ctx_t *c, double *f)
{   /* averaging.c:113-133 */
    #pragma omp parallel for
    for (long long i = 0; i < (long long)NORO*NVH; ++i) {
        int lo = (int)(i/NVH), h = (int)(i - (long long)lo*NVH);
        double vg = or_remap_verpri2horpri_vector(c, f, lo + (NL - NORO), h);
        long long vi = NSH + (long long)(NL - NORO + lo)*VPL + h;
        f[vi] += -c->g.slope[vi]*vg;
    }
}

void or_grad(const or_ctx_t *c, const double *in, double *out)
{   /* gradient_operators.c:56-74 */
    or_grad_hor_cov(c, in, out);
    or_grad_vert_cov(c, in, out);
    or_vector_field_hor_cov_to_con(c, out);
}

double or_vertical_contravariant_corr(const or_ctx_t *c, const double *f, int l, int h)
{   /* averaging.c:41-95 */
    double r = 0;
    int ne = h < N_PENT ? 5 : 6;
    const double *ipw = c->g.inner_product_weights, *slope = c->g.slope;
    const int *adj = c->g.adjacent_vector_indices_h;
    if (l >= NL - NORO) {
        if (l == NL - NORO) {
            for (int i = 0; i < ne; ++i) {
                long long si = (long long)l*NSH + h, vi = NSH + (long long)l*VPL + adj[6*h + i];
                r += -0.5*ipw[8*si + i]*slope[vi]*f[vi];
            }
        } else {
            for (int i = 0; i < ne; ++i) {
                long long si = (long long)(l - 1)*NSH + h, vi = NSH + (long long)(l - 1)*VPL + adj[6*h + i];
                r += -0.5*ipw[8*si + i]*slope[vi]*f[vi];
            }
            for (int i = 0; i < ne; ++i) {
                long long si = (long long)l*NSH + h, vi = NSH + (long long)l*VPL + adj[6*h + i];
                r += -0.5*ipw[8*si + i]*slope[vi]*f[vi];
            }
        }
    }
    return r;
}

double or_horizontal_covariant(const or_ctx_t *c, const double *f, int l, int h)
{   /* averaging.c:97-111 */
    long long vi = NSH + (long long)l*VPL + h;
    double r = f[vi];
    if (l >= NL - NORO) {
        double vc = or_remap_verpri2horpri_vector(c, f, l, h);
        r += c->g.slope[vi]*vc;
    }
    return r;
}

/* ------------------------------------------------------------------ */
/* divergences.c                                                        */
/* ------------------------------------------------------------------ */
void or_divv_h(const or_ctx_t *c, const double *in, double *out)
{   /* divergences.c:14-64 */
    const int *adj = c->g.adjacent_vector_indices_h, *sg = c->g.adjacent_signs_h;
    const double *area = c->g.area;
    #pragma omp parallel for
    for (int h = 0; h < NSH; ++h) {
        int ne = h < N_PENT ? 5 : 6;
        for (int l = 0; l < NL; ++l) {
            long long i = (long long)l*NSH + h;
            double comp_h = 0.0;
            for (int j = 0; j < ne; ++j) {
                long long vi = NSH + (long long)l*VPL + adj[6*h + j];
                comp_h += in[vi]*sg[6*h + j]*area[vi];
            }
            double comp_v = 0.0, cu, cl;
            if (l == NL - NORO - 1) {
                cl = or_vertical_contravariant_corr(c, in, l + 1, h);
                comp_v = -cl*area[h + (long long)(l + 1)*VPL];
            } else if (l == NL - 1) {
                cu = or_vertical_contravariant_corr(c, in, l, h);
                comp_v = cu*area[h + (long long)l*VPL];
            } else if (l > NL - NORO - 1) {
                cu = or_vertical_contravariant_corr(c, in, l, h);
                cl = or_vertical_contravariant_corr(c, in, l + 1, h);
                comp_v = cu*area[h + (long long)l*VPL] - cl*area[h + (long long)(l + 1)*VPL];
            }
            out[i] = 1.0/c->g.volume[i]*(comp_h + comp_v);
        }
    }
}

void or_divv_h_tracer(const or_ctx_t *c, const double *in, const double *dens, const double *wind, double *out)
{   /* divergences.c:66-148 */
    const int *adj = c->g.adjacent_vector_indices_h, *sg = c->g.adjacent_signs_h;
    const double *area = c->g.area;
    #pragma omp parallel for
    for (int h = 0; h < NSH; ++h) {
        int ne = h < N_PENT ? 5 : 6;
        for (int l = 0; l < NL; ++l) {
            long long i = (long long)l*NSH + h;
            double comp_h = 0.0;
            for (int j = 0; j < ne; ++j) {
                long long vi = NSH + (long long)l*VPL + adj[6*h + j];
                comp_h += in[vi]*sg[6*h + j]*area[vi];
            }
            double comp_v = 0.0, cu, cl, du, dl;
            if (l == NL - NORO - 1) {
                cl = or_vertical_contravariant_corr(c, wind, l + 1, h);
                dl = (cl <= 0.0) ? dens[i] : dens[i + NSH];
                comp_v = -dl*cl*area[h + (long long)(l + 1)*VPL];
            } else if (l == NL - 1) {
                cu = or_vertical_contravariant_corr(c, wind, l, h);
                du = (cu <= 0.0) ? dens[i - NSH] : dens[i];
                comp_v = du*cu*area[h + (long long)l*VPL];
            } else if (l > NL - NORO - 1) {
                cu = or_vertical_contravariant_corr(c, wind, l, h);
                du = (cu <= 0.0) ? dens[i - NSH] : dens[i];
                cl = or_vertical_contravariant_corr(c, wind, l + 1, h);
                dl = (cl <= 0.0) ? dens[i] : dens[i + NSH];
                comp_v = du*cu*area[h + (long long)l*VPL] - dl*cl*area[h + (long long)(l + 1)*VPL];
            }
            out[i] = 1.0/c->g.volume[i]*(comp_h + comp_v);
        }
    }
}

/* ------------------------------------------------------------------ */
/* multiplications.c, inner_product.c                                   */
/* ------------------------------------------------------------------ */
void or_scalar_times_vector_h(const or_ctx_t *c, const double *s, const double *v, double *out)
{   /* multiplications.c:45-67 */
    #pragma omp parallel for
    for (int h = 0; h < NVH; ++h)
        for (int l = 0; l < NL; ++l) {
            long long vi = NSH + (long long)l*VPL + h;
            double sv = 0.5*(s[c->g.to_index[h] + (long long)l*NSH] + s[c->g.from_index[h] + (long long)l*NSH]);
            out[vi] = sv*v[vi];
        }
}

void or_scalar_times_vector_h_upstream(const or_ctx_t *c, const double *s, const double *v, double *out)
{   /* multiplications.c:69-96 */
    #pragma omp parallel for
    for (int h = 0; h < NVH; ++h)
        for (int l = 0; l < NL; ++l) {
            long long vi = NSH + (long long)l*VPL + h;
            double sv = (v[vi] >= 0.0) ? s[c->g.from_index[h] + (long long)l*NSH] : s[c->g.to_index[h] + (long long)l*NSH];
            out[vi] = sv*v[vi];
        }
}

void or_scalar_times_vector_v(const or_ctx_t *c, const double *s, const double *v, double *out)
{   /* multiplications.c:98-117 */
    #pragma omp parallel for
    for (int h = 0; h < NSH; ++h)
        for (int l = 1; l < NL; ++l) {
            long long i = (long long)l*VPL + h;
            double sv = 0.5*(s[h + (long long)(l - 1)*NSH] + s[h + (long long)l*NSH]);
            out[i] = sv*v[i];
        }
}

void or_scalar_times_vector(const or_ctx_t *c, const double *s, const double *v, double *out)
{   /* multiplications.c:35-43 */
    or_scalar_times_vector_h(c, s, v, out);
    or_scalar_times_vector_v(c, s, v, out);
}

void or_inner_product(const or_ctx_t *c, const double *a, const double *b, double *out)
{   /* inner_product.c:13-43 */
    const int *adj = c->g.adjacent_vector_indices_h;
    const double *ipw = c->g.inner_product_weights;
    #pragma omp parallel for
    for (int h = 0; h < NSH; ++h) {
        int ne = h < N_PENT ? 5 : 6;
        for (int l = 0; l < NL; ++l) {
            long long i = (long long)l*NSH + h, base = 8*i;
            out[i] = 0;
            for (int j = 0; j < ne; ++j) {
                long long vi = NSH + (long long)l*VPL + adj[6*h + j];
                out[i] += ipw[base + j]*a[vi]*b[vi];
            }
            out[i] += ipw[base + 6]*a[h + (long long)l*VPL]*b[h + (long long)l*VPL];
            out[i] += ipw[base + 7]*a[h + (long long)(l + 1)*VPL]*b[h + (long long)(l + 1)*VPL];
        }
    }
}

/* ------------------------------------------------------------------ */
/* vorticities.c, vorticity_flux.c                                      */
/* ------------------------------------------------------------------ */
void or_calc_rel_vort_on_triangles(const or_ctx_t *c, const double *w, double *res)
{   /* vorticities.c:167-222 */
    const double *zv = c->g.z_vector, *dzv = c->dg.z_vector, *nd = c->g.normal_distance;
    #pragma omp parallel for
    for (long long i = 0; i < (long long)NL*NT; ++i) {
        int l = (int)(i/NT), h = (int)(i - (long long)l*NT);
        res[i] = 0;
        for (int j = 0; j < 3; ++j) {
            long long vi = NSH + (long long)l*VPL + c->dg.vorticity_indices_triangles[3*h + j];
            double vel = w[vi], lrf = 1;
            if (l >= NL - NORO) {
                long long di = NVH + (long long)l*DVPL + h;
                lrf = (c->g.radius + dzv[di])/(c->g.radius + zv[vi]);
                double delta_z = dzv[di] - zv[vi];
                long long ivg;
                if (delta_z > 0) ivg = vi - VPL;
                else if (l == NL - 1) ivg = vi - VPL;
                else ivg = vi + VPL;
                double vg = (w[vi] - w[ivg])/(zv[vi] - zv[ivg]);
                vel += delta_z*vg;
            }
            res[i] += lrf*nd[vi]*c->dg.vorticity_signs_triangles[3*h + j]*vel;
        }
        res[i] = res[i]/c->dg.area[NVH + (long long)l*DVPL + h];
    }
}

void or_calc_rel_vort(const or_ctx_t *c, const double *w, game_diag_view_t *dg_)
{   /* vorticities.c:104-165 */
    or_calc_rel_vort_on_triangles(c, w, dg_->rel_vort_on_triangles);
    const double *rvt = dg_->rel_vort_on_triangles, *da = c->dg.area, *nd = c->g.normal_distance;
    double *rv = dg_->rel_vort;
    long long n = (long long)NL*2*NVH + NVH;
    #pragma omp parallel for
    for (long long i = NVH; i < n; ++i) {
        int l = (int)(i/(2*NVH)), h = (int)(i - (long long)l*2*NVH);
        if (h >= NVH) {
            int e = h - NVH;
            long long base = NVH + (long long)l*DVPL;
            int fd = c->dg.from_index[e], td = c->dg.to_index[e];
            rv[i] = (da[base + fd]*rvt[(long long)l*NT + fd] + da[base + td]*rvt[(long long)l*NT + td])
                    /(da[base + fd] + da[base + td]);
        } else {
            long long base = (long long)l*VPL;
            if (l == NL) {
                long long i2 = base - NVH + h;
                double cv2 = or_horizontal_covariant(c, w, l - 1, h);
                rv[i] = 1/da[h + (long long)l*DVPL]*nd[i2]*cv2;
            } else {
                long long i0 = base + NSH + h, i1 = base + c->g.from_index[h], i2 = base - NVH + h, i3 = base + c->g.to_index[h];
                double cv0 = or_horizontal_covariant(c, w, l, h);
                double cv2 = or_horizontal_covariant(c, w, l - 1, h);
                rv[i] = 1/da[h + (long long)l*DVPL]*(
                    - nd[i0]*cv0
                    + nd[i1]*w[i1]
                    + nd[i2]*cv2
                    - nd[i3]*w[i3]);
            }
        }
    }
    #pragma omp parallel for
    for (int i = 0; i < NVH; ++i) rv[i] = rv[i + 2*NVH];
}

void or_add_f_to_rel_vort(const or_ctx_t *c, const double *rv, double *out)
{   /* vorticities.c:87-102 */
    long long n = (long long)NL*2*NVH + NVH;
    #pragma omp parallel for
    for (long long i = 0; i < n; ++i) {
        int l = (int)(i/(2*NVH)), h = (int)(i - (long long)l*2*NVH);
        out[i] = rv[i] + c->dg.f_vec[h];
    }
}

/* The reference's struct grid places gravity_potential (S doubles) and
   gravity_m (V doubles) right behind z_vector, and calc_pot_vort
   (vorticities.c:55-70) indexes z_vector with an offset of NO_OF_SCALARS, which
   for the lowest level runs past the end of z_vector into those members.
   This helper returns exactly what such an index reads. */
static double ref_z_vector_read(const or_ctx_t *c, long long k)
{
    if (k < NV) return c->g.z_vector[k];
    k -= NV;
    if (k < NS) return c->g.gravity_potential[k];
    k -= NS;
    if (k < NV) return c->g.gravity_m[k];
    fprintf(stderr, "oracle: z_vector index beyond the emulated struct layout\n");
    abort();
}

double or_density_to_rhombus(const or_ctx_t *c, const double *rho, int l, int e)
{   /* vorticities.c:33-43 */
    double dv = 0;
    for (int j = 0; j < 4; ++j)
        dv += c->g.density_to_rhombi_weights[4*e + j]*rho[(long long)l*NSH + c->g.density_to_rhombi_indices[4*e + j]];
    return dv;
}

void or_calc_pot_vort(const or_ctx_t *c, const double *w, const double *rho, game_diag_view_t *dg_)
{   /* vorticities.c:18-85 */
    or_calc_rel_vort(c, w, dg_);
    or_add_f_to_rel_vort(c, dg_->rel_vort, dg_->pot_vort);
    long long n = (long long)NL*2*NVH + NVH;
    const int *fr = c->g.from_index, *to = c->g.to_index;
    #pragma omp parallel for
    for (long long i = 0; i < n; ++i) {
        int l = (int)(i/(2*NVH)), h = (int)(i - (long long)l*2*NVH);
        double dv;
        if (h >= NVH) {
            dv = or_density_to_rhombus(c, rho, l, h - NVH);
        } else {
            if (l == 0) {
                dv = 0.5*(rho[fr[h]] + rho[to[h]])
                     + (0.5*(rho[fr[h]] + rho[to[h]]) - 0.5*(rho[fr[h] + NSH] + rho[to[h] + NSH]))
                     /(ref_z_vector_read(c, NS + h) - ref_z_vector_read(c, NS + VPL + h))
                     *(c->g.z_vector[0] - ref_z_vector_read(c, NS + h));
            } else if (l == NL) {
                long long a = (long long)(l - 1)*NSH, b = (long long)(l - 2)*NSH;
                dv = 0.5*(rho[a + fr[h]] + rho[a + to[h]])
                     + (0.5*(rho[b + fr[h]] + rho[b + to[h]]) - 0.5*(rho[a + fr[h]] + rho[a + to[h]]))
                     /(ref_z_vector_read(c, NS + (long long)(l - 2)*VPL + h) - ref_z_vector_read(c, NS + (long long)(l - 1)*VPL + h))
                     *(0.5*(c->g.z_vector[(long long)l*VPL + fr[h]] + c->g.z_vector[(long long)l*VPL + to[h]])
                       - ref_z_vector_read(c, NS + (long long)(l - 1)*VPL + h));
            } else {
                long long uf = (long long)(l - 1)*NSH + fr[h], ut = (long long)(l - 1)*NSH + to[h];
                dv = 0.25*(rho[uf] + rho[ut] + rho[uf + NSH] + rho[ut + NSH]);
            }
        }
        dg_->pot_vort[i] = dg_->pot_vort[i]/dv;
    }
}

void or_vorticity_flux(const or_ctx_t *c, const double *F, const double *pv, double *out)
{   /* vorticity_flux.c:16-187 */
    const int *fr = c->g.from_index, *to = c->g.to_index, *ti = c->g.trsk_indices, *tm = c->g.trsk_modified_curl_indices;
    const int *adj = c->g.adjacent_vector_indices_h;
    const double *tw = c->g.trsk_weights, *ipw = c->g.inner_product_weights;
    #pragma omp parallel for
    for (int h = 0; h < VPL; ++h) {
        for (int l = 0; l < NL + 1; ++l) {
            long long i = (long long)l*VPL + h;
            if (h >= NSH && l < NL) {
                out[i] = 0;
                int e = h - NSH;
                long long fb = NSH + (long long)l*VPL, pb = NVH + (long long)l*2*NVH;
                if (fr[e] < N_PENT) {
                    for (int j = 0; j < 4; ++j)
                        out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*pv[pb + tm[10*e + j]];
                } else {
                    for (int j = 0; j < 5; ++j) {
                        if (j == 2)
                            out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*0.5*(pv[pb + tm[10*e + j]] + pv[pb + e]);
                        else
                            out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*pv[pb + tm[10*e + j]];
                    }
                }
                if (to[e] < N_PENT) {
                    for (int j = 5; j < 9; ++j)
                        out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*pv[pb + tm[10*e + j]];
                } else {
                    for (int j = 5; j < 10; ++j) {
                        if (j == 7)
                            out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*0.5*(pv[pb + tm[10*e + j]] + pv[pb + e]);
                        else
                            out[i] += tw[10*e + j]*F[fb + ti[10*e + j]]*pv[pb + tm[10*e + j]];
                    }
                }
                long long sf = (long long)l*NSH + fr[e], st = (long long)l*NSH + to[e];
                out[i] -= 0.5*ipw[8*sf + 6]*F[(long long)l*VPL + fr[e]]*pv[e + (long long)l*2*NVH];
                out[i] -= 0.5*ipw[8*st + 6]*F[(long long)l*VPL + to[e]]*pv[e + (long long)l*2*NVH];
                out[i] -= 0.5*ipw[8*sf + 7]*F[(long long)(l + 1)*VPL + fr[e]]*pv[e + (long long)(l + 1)*2*NVH];
                out[i] -= 0.5*ipw[8*st + 7]*F[(long long)(l + 1)*VPL + to[e]]*pv[e + (long long)(l + 1)*2*NVH];
            } else if (h < NSH) {
                out[i] = 0;
                int ne = h < N_PENT ? 5 : 6;
                double vw = 0.5;
                if (l == 0 || l == NL) vw = 1;
                if (l >= 1)
                    for (int j = 0; j < ne; ++j)
                        out[i] += vw*ipw[8*((long long)(l - 1)*NSH + h) + j]
                                  *F[NSH + (long long)(l - 1)*VPL + adj[6*h + j]]
                                  *pv[(long long)l*2*NVH + adj[6*h + j]];
                if (l <= NL - 1)
                    for (int j = 0; j < ne; ++j)
                        out[i] += vw*ipw[8*((long long)l*NSH + h) + j]
                                  *F[NSH + (long long)l*VPL + adj[6*h + j]]
                                  *pv[(long long)l*2*NVH + adj[6*h + j]];
            }
        }
    }
}

void or_linear_combine_two_states(const or_ctx_t *c, const game_state_view_t *s0, const game_state_view_t *s1,
                                  game_state_view_t *o, double c0, double c1)
{   /* linear_combine_two_states.c:14-43 (soil temperature is not carried) */
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i) {
        for (int j = 0; j < NCONST; ++j)
            o->rho[j*NS + i] = c0*s0->rho[j*NS + i] + c1*s1->rho[j*NS + i];
        o->rhotheta_v[i] = c0*s0->rhotheta_v[i] + c1*s1->rhotheta_v[i];
        o->theta_v_pert[i] = o->rhotheta_v[i]/o->rho[NCC*NS + i] - c->g.theta_v_bg[i];
        o->exner_pert[i] = c0*s0->exner_pert[i] + c1*s1->exner_pert[i];
    }
    #pragma omp parallel for
    for (long long i = 0; i < NV; ++i) o->wind[i] = c0*s0->wind[i] + c1*s1->wind[i];
}

/* ------------------------------------------------------------------ */
/* constituents: derived_quantities.c, dictionary.c, phase_trans.c      */
/* ------------------------------------------------------------------ */
void or_temperature_diagnostics(const or_ctx_t *c, const game_state_view_t *s, double *T)
{   /* derived_quantities.c:15-40 */
    if (c->d.moisture_on == 0) {
        #pragma omp parallel for
        for (long long i = 0; i < NS; ++i)
            T[i] = (c->g.theta_v_bg[i] + s->theta_v_pert[i])*(c->g.exner_bg[i] + s->exner_pert[i]);
    } else {
        #pragma omp parallel for
        for (long long i = 0; i < NS; ++i)
            T[i] = (c->g.theta_v_bg[i] + s->theta_v_pert[i])*(c->g.exner_bg[i] + s->exner_pert[i])
                   /(1.0 + s->rho[(NCC + 1)*NS + i]/s->rho[NCC*NS + i]*(M_D/M_V - 1.0));
    }
}

static double density_total(const or_ctx_t *c, const game_state_view_t *s, long long i)
{   /* derived_quantities.c:83-95 */
    double r = 0.0;
    for (int k = 0; k < NCONST; ++k) r += s->rho[k*NS + i];
    return r;
}

static double c_p_ice(double t)
{   /* dictionary.c:111-132 */
    if (t > T_0) t = T_0;
    if (t < 20.0) t = 20.0;
    double r = -2.0572 + 0.14644*t + 0.06163*t*exp(-pow(t/125.1, 2));
    return r/M_V;
}

static double c_p_water(double t)
{   /* dictionary.c:134-171 */
    double tc = t - T_0, r;
    if (tc >= 0.0) {
        if (tc > 35.0) tc = 35.0;
        r = 0.9979 + 3.1e-6*pow(tc - 35.0, 2) + 3.8e-9*pow(tc - 35.0, 4);
    } else {
        if (tc < -37.0) tc = -37.0;
        r = 1.000938 - 2.7052e-3*tc - 2.3235e-5*pow(tc, 2) + 4.3778e-6*pow(tc, 3) + 2.7136e-7*pow(tc, 4);
    }
    return 4186.8*r;
}

double or_c_p_cond(int k, double t)
{   /* dictionary.c:173-190 */
    return (fmod(k, 2) == 0) ? c_p_ice(t) : c_p_water(t);
}

static double enthalpy_evaporation(double t)
{   /* dictionary.c:222-258 */
    double r;
    if (t < T_0) {
        if (t < 30.0) t = 30.0;
        r = 56579.0 - 42.212*t + exp(0.1149*(281.6 - t));
        r = r/M_V;
    } else {
        if (t > T_0 + 100.0) t = T_0 + 100.0;
        r = 3151378.0 - 2386.0*t;
    }
    return r;
}

static double enthalpy_sublimation(double t)
{   /* dictionary.c:260-285 */
    if (t < 30.0) t = 30.0;
    if (t > T_0) t = T_0;
    double r = 46782.5 + 35.8925*t - 0.07414*pow(t, 2) + 541.5*exp(-pow(t/123.75, 2));
    return r/M_V;
}

double or_phase_trans_heat(int dir, double t)
{   /* dictionary.c:192-220 */
    double r = 0.0;
    if (dir == 0) r = enthalpy_evaporation(t);
    if (dir == 1) r = enthalpy_sublimation(t);
    if (dir == 2) r = enthalpy_sublimation(t) - enthalpy_evaporation(t);
    return r;
}

double or_saturation_pressure_over_water(double t)
{   /* dictionary.c:287-333 */
    double tc = t - T_0, r;
    if (tc > 100.0) tc = 100.0;
    if (tc > 0.0) r = exp(34.494 - 4924.99/(tc + 237.1))/pow(tc + 105.0, 1.57);
    else {
        if (tc < -50.0) tc = -50.0;
        r = 6.107799961
            + 4.436518521e-1*tc
            + 1.428945805e-2*pow(tc, 2)
            + 2.650648471e-4*pow(tc, 3)
            + 3.031240396e-6*pow(tc, 4)
            + 2.034080948e-8*pow(tc, 5)
            + 6.136820929e-11*pow(tc, 6);
        r = 100.0*r;
    }
    return r;
}

static double dsat_water_dT(double t)
{   /* dictionary.c:335-359 */
    double tc = t - T_0;
    if (tc > 100.0) tc = 100.0;
    if (tc < 0.0) tc = 0.0;
    return or_saturation_pressure_over_water(t)*(4924.99/pow(tc + 237.1, 2.0) - 1.57/(tc + 105.0));
}

double or_saturation_pressure_over_ice(double t)
{   /* dictionary.c:361-385 */
    double tc = t - T_0;
    if (tc < -80.0) tc = -80.0;
    if (tc > 0.0) tc = 0.0;
    return exp(43.494 - 6545.8/(tc + 278.0))/pow(tc + 868.0, 2);
}

static double dsat_ice_dT(double t)
{   /* dictionary.c:387-412 */
    double tc = t - T_0;
    if (tc < -80.0) tc = -80.0;
    if (tc > 0.0) tc = 0.0;
    return or_saturation_pressure_over_ice(t)*(6545.8/pow(tc + 278.0, 2.0) - 2.0/(tc + 868.0));
}

static double c_v_mass_weighted_air(const or_ctx_t *c, const game_state_view_t *s, const double *T, long long i)
{   /* derived_quantities.c:97-115 */
    double r = 0.0;
    for (int k = 0; k < NCC; ++k) r += s->rho[k*NS + i]*or_c_p_cond(k, T[i]);
    r += s->rho[NCC*NS + i]*C_D_V;
    if (c->d.moisture_on == 1) r += s->rho[(NCC + 1)*NS + i]*C_V_V;
    return r;
}

void or_calc_h2otracers_source_rates(const or_ctx_t *c, const game_state_view_t *s, const game_diag_view_t *dg_,
                                     const game_config_t *cfg, game_irrev_view_t *irrev, double dt)
{   /* phase_trans.c:16-311 (surface phase transitions off) */
    const double *T = dg_->temperature;
    const double *rho = s->rho;
    double *ptr = irrev->phase_trans_rates, *phr = irrev->phase_trans_heating_rate;
    double mcwc = 0.2e-3;
    (void)cfg;
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i) {
        double sp, wvp, dp, ap, ef, a, b, cc, p, q, dd, ptd;
        sp = (T[i] >= T_0) ? or_saturation_pressure_over_water(T[i]) : or_saturation_pressure_over_ice(T[i]);
        wvp = rho[(NCC + 1)*NS + i]*R_V*T[i];
        dp = (rho[NCC*NS + i] - rho[(NCC + 1)*NS + i])*R_D*T[i];
        ap = dp + wvp;
        ef = (T[i] >= T_0) ? 1.00071*exp(0.000000045*ap) : 0.99882*exp(0.00000008*ap);
        sp = ef*sp;
        double cvm = c_v_mass_weighted_air(c, s, T, i);
        if (sp >= wvp) {
            if (T[i] >= T_0) {
                ptr[2*NS + i] = -rho[2*NS + i]/dt;
                a = -R_V*or_phase_trans_heat(0, T[i])/cvm;
                b = R_V*T[i] - R_V*rho[(NCC + 1)*NS + i]*or_phase_trans_heat(0, T[i])/cvm
                    + ef*dsat_water_dT(T[i])*or_phase_trans_heat(0, T[i])/cvm;
                cc = wvp - sp;
                p = b/a;
                q = cc/a;
                dd = -0.5*p - pow(0.25*pow(p, 2) - q, 0.5);
                ptd = fmin(rho[3*NS + i], dd);
                ptr[4*NS + i] = ptd/dt;
                ptr[3*NS + i] = rho[2*NS + i]/dt - ptd/dt;
                phr[i] = ptr[2*NS + i]*or_phase_trans_heat(2, T[i]) - ptd*or_phase_trans_heat(0, T[i])/dt;
            } else {
                ptr[3*NS + i] = -rho[3*NS + i]/dt;
                a = -R_V*or_phase_trans_heat(1, T[i])/cvm;
                b = R_V*T[i] - R_V*rho[(NCC + 1)*NS + i]*or_phase_trans_heat(1, T[i])/cvm
                    + ef*dsat_ice_dT(T[i])*or_phase_trans_heat(1, T[i])/cvm;
                cc = wvp - sp;
                p = b/a;
                q = cc/a;
                dd = -0.5*p - pow(0.25*pow(p, 2) - q, 0.5);
                ptd = fmin(rho[2*NS + i], dd);
                ptr[4*NS + i] = ptd/dt;
                ptr[2*NS + i] = rho[3*NS + i]/dt - ptd/dt;
                phr[i] = -ptr[3*NS + i]*or_phase_trans_heat(2, T[i]) - ptd*or_phase_trans_heat(1, T[i])/dt;
            }
        } else {
            if (T[i] >= T_0) {
                ptr[2*NS + i] = -rho[2*NS + i]/dt;
                a = -R_V*or_phase_trans_heat(0, T[i])/cvm;
                b = R_V*T[i] - R_V*rho[(NCC + 1)*NS + i]*or_phase_trans_heat(0, T[i])/cvm
                    + ef*dsat_water_dT(T[i])*or_phase_trans_heat(0, T[i])/cvm;
                cc = wvp - sp;
                p = b/a;
                q = cc/a;
                dd = -0.5*p - pow(0.25*pow(p, 2) - q, 0.5);
                ptr[4*NS + i] = dd/dt;
                ptr[3*NS + i] = rho[2*NS + i]/dt - dd/dt;
                phr[i] = ptr[2*NS + i]*or_phase_trans_heat(2, T[i]) - dd*or_phase_trans_heat(0, T[i])/dt;
            } else {
                ptr[3*NS + i] = -rho[3*NS + i]/dt;
                a = -R_V*or_phase_trans_heat(1, T[i])/cvm;
                b = R_V*T[i] - R_V*rho[(NCC + 1)*NS + i]*or_phase_trans_heat(1, T[i])/cvm
                    + ef*dsat_ice_dT(T[i])*or_phase_trans_heat(1, T[i])/cvm;
                cc = wvp - sp;
                p = b/a;
                q = cc/a;
                dd = -0.5*p - pow(0.25*pow(p, 2) - q, 0.5);
                ptr[4*NS + i] = dd/dt;
                ptr[2*NS + i] = rho[3*NS + i]/dt - dd/dt;
                phr[i] = -ptr[3*NS + i]*or_phase_trans_heat(2, T[i]) - dd*or_phase_trans_heat(1, T[i])/dt;
            }
        }
        ptr[i] = 0.0;
        ptr[NS + i] = 0.0;
        if (T[i] < T_0) {
            ptr[i] = fmax(rho[2*NS + i] - mcwc*rho[4*NS + i], 0.0)/1000.0;
            ptr[2*NS + i] -= ptr[i];
        } else if (T[i] >= T_0) {
            ptr[NS + i] = fmax(rho[3*NS + i] - mcwc*rho[4*NS + i], 0.0)/1000.0;
            ptr[3*NS + i] -= ptr[NS + i];
        }
        if (T[i] >= T_0 && rho[i] > 0.0) {
            ptr[i] = -rho[i]/dt;
            ptr[NS + i] -= ptr[i];
            phr[i] += ptr[i]*or_phase_trans_heat(2, T[i]);
        }
        if (T[i] < T_0 && rho[NS + i] > 0.0) {
            ptr[NS + i] = -rho[NS + i]/dt;
            ptr[i] -= ptr[NS + i];
            phr[i] += -ptr[NS + i]*or_phase_trans_heat(2, T[i]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* time stepping                                                        */
/* ------------------------------------------------------------------ */
void or_manage_pressure_gradient(const or_ctx_t *c, const game_state_view_t *s, game_diag_view_t *dg_,
                                 game_forcings_view_t *f, game_irrev_view_t *irrev, const game_config_t *cfg)
{   /* manage_pressure_gradient.c:17-72 */
    if (cfg->totally_first_step_bool == 0) {
        #pragma omp parallel for
        for (long long i = 0; i < NV; ++i)
            f->pgrad_acc_old[i] = -f->pressure_gradient_acc_neg_nl[i] - f->pressure_gradient_acc_neg_l[i];
    }
    double *ph = dg_->scalar_field_placeholder;
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i) ph[i] = C_D_P*(c->g.theta_v_bg[i] + s->theta_v_pert[i]);
    or_grad(c, s->exner_pert, f->pressure_gradient_acc_neg_nl);
    or_scalar_times_vector(c, ph, f->pressure_gradient_acc_neg_nl, f->pressure_gradient_acc_neg_nl);
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i) ph[i] = C_D_P*s->theta_v_pert[i];
    or_scalar_times_vector(c, ph, c->g.exner_bg_grad, f->pressure_gradient_acc_neg_l);
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i)
        irrev->pressure_gradient_decel_factor[i] = s->rho[NCC*NS + i]/density_total(c, s, i);
    or_scalar_times_vector(c, irrev->pressure_gradient_decel_factor, f->pressure_gradient_acc_neg_nl, f->pressure_gradient_acc_neg_nl);
    or_scalar_times_vector(c, irrev->pressure_gradient_decel_factor, f->pressure_gradient_acc_neg_l, f->pressure_gradient_acc_neg_l);
    if (cfg->totally_first_step_bool == 1) {
        #pragma omp parallel for
        for (long long i = 0; i < NV; ++i)
            f->pgrad_acc_old[i] = -f->pressure_gradient_acc_neg_nl[i] - f->pressure_gradient_acc_neg_l[i];
    }
}

void or_calc_pressure_grad_condensates_v(const or_ctx_t *c, const game_state_view_t *s, game_forcings_view_t *f,
                                         game_irrev_view_t *irrev)
{   /* manage_pressure_gradient.c:74-87 */
    #pragma omp parallel for
    for (long long i = 0; i < NS; ++i)
        irrev->pressure_gradient_decel_factor[i] = s->rho[NCC*NS + i]/density_total(c, s, i) - 1.0;
    or_scalar_times_vector_v(c, irrev->pressure_gradient_decel_factor, c->g.gravity_m, f->pressure_grad_condensates_v);
}

void or_vector_tendencies_expl(const or_ctx_t *c, const game_state_view_t *s, game_state_view_t *tend,
                               game_diag_view_t *dg_, game_forcings_view_t *f, game_irrev_view_t *irrev,
                               const game_config_t *cfg, int rk_step)
{   /* vector_tendencies_expl.c:18-133 (diffusion and PBL branches off) */
    if (rk_step == 1 || cfg->totally_first_step_bool == 1) {
        const double *rho_gas = &s->rho[NCC*NS];
        or_scalar_times_vector(c, rho_gas, s->wind, dg_->flux_density);
        or_calc_pot_vort(c, s->wind, rho_gas, dg_);
        or_vorticity_flux(c, dg_->flux_density, dg_->pot_vort, f->pot_vort_tend);
        or_inner_product(c, s->wind, s->wind, dg_->v_squared);
        or_grad(c, dg_->v_squared, f->v_squared_grad);
    }
    double new_w = 1.0;
    if (rk_step == 1) new_w = 0.5;
    double old_w = 1.0 - new_w;
    double chw = 0.5 + cfg->impl_thermo_weight;
    double ohw = 1.0 - chw;
    double cvw = 1.0 - cfg->impl_thermo_weight;
    double *tw = tend->wind;
    #pragma omp parallel for
    for (long long i = 0; i < NV; ++i) {
        int l = (int)(i/VPL);
        int h = (int)(i - (long long)l*VPL);
        if (i < NSH || i >= NV - NSH) tw[i] = 0.0;
        else if (h >= NSH)
            tw[i] = old_w*tw[i] + new_w*(
                ohw*f->pgrad_acc_old[i]
                - chw*(f->pressure_gradient_acc_neg_nl[i] + f->pressure_gradient_acc_neg_l[i])
                + f->pot_vort_tend[i]
                - 0.5*f->v_squared_grad[i]
                + irrev->friction_acc[i]);
        else if (h < NSH)
            tw[i] = old_w*tw[i] + new_w*(
                -cvw*(f->pressure_gradient_acc_neg_nl[i] + f->pressure_gradient_acc_neg_l[i])
                + f->pot_vort_tend[i]
                - 0.5*f->v_squared_grad[i]
                + irrev->friction_acc[i]
                + f->pressure_grad_condensates_v[i]);
    }
}

void or_scalar_tendencies_expl(const or_ctx_t *c, const game_state_view_t *s_old, const game_state_view_t *s,
                               game_state_view_t *tend, game_diag_view_t *dg_, game_forcings_view_t *f,
                               game_irrev_view_t *irrev, const game_config_t *cfg, int rk_step)
{   /* scalar_tendencies_expl.c:17-181 (diffusion branches off) */
    (void)s_old; (void)cfg;
    double old_w[6], new_w[6];
    for (int i = 0; i < NCONST; ++i) {
        new_w[i] = 1.0;
        if (rk_step == 1 && i != NCC) new_w[i] = 0.5;
        old_w[i] = 1.0 - new_w[i];
    }
    for (int i = 0; i < NCONST; ++i) {
        long long shift = (long long)i*NS, shift_pt = shift;
        if (c->d.moisture_on == 1 && i == NCC + 1) shift_pt = shift - NS;
        if (i == NCC) {
            or_scalar_times_vector_h(c, &s->rho[shift], s->wind, dg_->flux_density);
            or_divv_h(c, dg_->flux_density, dg_->flux_density_divv);
        } else {
            or_scalar_times_vector_h_upstream(c, &s->rho[shift], s->wind, dg_->flux_density);
            or_divv_h_tracer(c, dg_->flux_density, &s->rho[shift], s->wind, dg_->flux_density_divv);
        }
        #pragma omp parallel for
        for (long long j = 0; j < NS; ++j) {
            long long si = shift + j;
            tend->rho[si] = old_w[i]*tend->rho[si] + new_w[i]*(
                -dg_->flux_density_divv[j]
                + irrev->mass_diff_tendency[shift + j]
                + irrev->phase_trans_rates[shift_pt + j]);
        }
        if (i == NCC) {
            double *ph = dg_->scalar_field_placeholder;
            #pragma omp parallel for
            for (long long j = 0; j < NS; ++j) ph[j] = s->rhotheta_v[j]/s->rho[shift + j];
            or_scalar_times_vector_h(c, ph, dg_->flux_density, dg_->flux_density);
            or_divv_h(c, dg_->flux_density, dg_->flux_density_divv);
            #pragma omp parallel for
            for (long long j = 0; j < NS; ++j) {
                tend->rhotheta_v[j] = old_w[i]*tend->rhotheta_v[j] + new_w[i]*(
                    -dg_->flux_density_divv[j]
                    + C_D_V*s->rho[shift + j]/c_v_mass_weighted_air(c, s, dg_->temperature, j)*(
                        irrev->heating_diss[j]
                        + irrev->temperature_diffusion_heating[j]
                        + f->radiation_tendency[j]
                        + irrev->phase_trans_heating_rate[j]
                        + irrev->condensates_sediment_heat[j]
                    )/(C_D_P*(c->g.exner_bg[j] + s->exner_pert[j]))
                    + (irrev->phase_trans_rates[shift + j] + irrev->mass_diff_tendency[shift + j])*ph[j]);
            }
        }
    }
}

void or_thomas(const double *cv, const double *dv, const double *ev, const double *rv, double *x, int n)
{   /* column_solvers.c:641-669 */
    double ep[64], rp[64];
    ep[0] = ev[0]/dv[0];
    for (int j = 1; j < n - 1; ++j) ep[j] = ev[j]/(dv[j] - ep[j - 1]*cv[j - 1]);
    rp[0] = rv[0]/dv[0];
    for (int j = 1; j < n; ++j) rp[j] = (rv[j] - rp[j - 1]*cv[j - 1])/(dv[j] - ep[j - 1]*cv[j - 1]);
    x[n - 1] = rp[n - 1];
    for (int j = n - 2; j >= 0; --j) x[j] = rp[j] - ep[j]*x[j + 1];
}

void or_three_band_solver_ver_waves(const or_ctx_t *c, const game_state_view_t *so, game_state_view_t *sn,
                                    const game_state_view_t *td, const game_config_t *cfg, double dt, int rk_step)
{   /* column_solvers.c:19-383 (soil and surface sensible heat off) */
    double iw = cfg->impl_thermo_weight;
    double dsh = cfg->damping_start_height_over_toa*c->g.z_vector[0];
    double pdw = 0.5;
    long long gi = (long long)NCC*NS;
    const double *vol = c->g.volume, *exbg = c->g.exner_bg, *zs = c->g.z_scalar, *area = c->g.area, *zv = c->g.z_vector;
    int L = NL;
    #pragma omp parallel for
    for (int i = 0; i < NSH; ++i) {
        double cvv[64], dvv[64], evv[64], rvv[64], re[64], rte[64], tpe[64], epe[64], tin[64], sol[64],
               rio[64], rie[64], al[64], be[64], ga[64];
        for (int j = 0; j < L; ++j) {
            long long b = i + (long long)j*NSH;
            re[j] = so->rho[gi + b] + dt*td->rho[gi + b];
            rte[j] = so->rhotheta_v[b] + dt*td->rhotheta_v[b];
            if (rk_step == 0) {
                al[j] = -so->rhotheta_v[b]/pow(so->rho[gi + b], 2)/vol[b];
                be[j] = 1.0/so->rho[gi + b]/vol[b];
                ga[j] = R_D/(C_D_V*so->rhotheta_v[b])*(exbg[b] + so->exner_pert[b])/vol[b];
            } else {
                double ao = -so->rhotheta_v[b]/pow(so->rho[gi + b], 2);
                double bo = 1.0/so->rho[gi + b];
                double go = R_D/(C_D_V*so->rhotheta_v[b])*(exbg[b] + so->exner_pert[b]);
                double an = -sn->rhotheta_v[b]/pow(sn->rho[gi + b], 2);
                double bn = 1.0/sn->rho[gi + b];
                double gn = R_D/(C_D_V*sn->rhotheta_v[b])*(exbg[b] + sn->exner_pert[b]);
                al[j] = ((1.0 - pdw)*ao + pdw*an)/vol[b];
                be[j] = ((1.0 - pdw)*bo + pdw*bn)/vol[b];
                ga[j] = ((1.0 - pdw)*go + pdw*gn)/vol[b];
            }
            tpe[j] = so->theta_v_pert[b] + dt*vol[b]*(al[j]*td->rho[gi + b] + be[j]*td->rhotheta_v[b]);
            epe[j] = so->exner_pert[b] + dt*vol[b]*ga[j]*td->rhotheta_v[b];
        }
        for (int j = 0; j < L - 1; ++j) {
            long long b = i + (long long)j*NSH, lo = i + (long long)(j + 1)*NSH;
            rio[j] = 0.5*(so->rho[gi + b] + so->rho[gi + lo]);
            rie[j] = 0.5*(re[j] + re[j + 1]);
            tin[j] = 0.5*(sn->rhotheta_v[b]/sn->rho[gi + b] + sn->rhotheta_v[lo]/sn->rho[gi + lo]);
        }
        for (int j = 0; j < L - 1; ++j) {
            long long b = i + (long long)j*NSH, lo = i + (long long)(j + 1)*NSH;
            long long wi = i + (long long)(j + 1)*VPL;
            dvv[j] = -pow(tin[j], 2)*(ga[j] + ga[j + 1])
                     + 0.5*(exbg[b] - exbg[lo])*(al[j + 1] - al[j] + tin[j]*(be[j + 1] - be[j]))
                     - (zs[b] - zs[lo])/(iw*pow(dt, 2)*C_D_P*rio[j])
                       *(2.0/area[wi] + dt*so->wind[wi]*0.5*(-1.0/vol[b] + 1.0/vol[lo]));
            rvv[j] = -(so->wind[wi] + dt*td->wind[wi])*(zs[b] - zs[lo])/(iw*pow(dt, 2)*C_D_P)
                     + tin[j]*(epe[j] - epe[j + 1])/dt
                     + 0.5/dt*(tpe[j] + tpe[j + 1])*(exbg[b] - exbg[lo])
                     - (zs[b] - zs[lo])/(iw*pow(dt, 2)*C_D_P)*so->wind[wi]*rie[j]/rio[j];
        }
        for (int j = 0; j < L - 2; ++j) {
            long long b = i + (long long)j*NSH, lo = i + (long long)(j + 1)*NSH, l2 = (long long)(j + 2)*NSH + i;
            cvv[j] = tin[j + 1]*ga[j + 1]*tin[j]
                     + 0.5*(exbg[lo] - exbg[l2])*(al[j + 1] + be[j + 1]*tin[j])
                     - (zs[lo] - zs[l2])/(iw*dt*C_D_P)*0.5*so->wind[i + (long long)(j + 2)*VPL]/(vol[lo]*rio[j + 1]);
            evv[j] = tin[j]*ga[j + 1]*tin[j + 1]
                     - 0.5*(exbg[b] - exbg[lo])*(al[j + 1] + be[j + 1]*tin[j + 1])
                     + (zs[b] - zs[lo])/(iw*dt*C_D_P)*0.5*so->wind[i + (long long)(j + 1)*VPL]/(vol[lo]*rio[j]);
        }
        or_thomas(cvv, dvv, evv, rvv, sol, L - 1);
        for (int j = 0; j < L - 1; ++j) {
            double zad = zv[i + (long long)(j + 1)*VPL] - dsh, dc;
            if (zad < 0.0) dc = 0.0;
            else dc = cfg->damping_coeff_max*pow(sin(0.5*M_PI*zad/(zv[0] - dsh)), 2);
            sol[j] = sol[j]/(1.0 + dt*dc);
        }
        for (int j = 0; j < L; ++j) {
            long long b = i + (long long)j*NSH;
            if (j == 0) sn->rho[gi + b] = re[j] + dt*(sol[j])/vol[b];
            else if (j == L - 1) sn->rho[gi + b] = re[j] + dt*(-sol[j - 1])/vol[b];
            else sn->rho[gi + b] = re[j] + dt*(-sol[j - 1] + sol[j])/vol[b];
        }
        for (int j = 0; j < L; ++j) {
            long long b = i + (long long)j*NSH;
            if (j == 0) sn->rhotheta_v[b] = rte[j] + dt*(tin[j]*sol[j])/vol[b];
            else if (j == L - 1) sn->rhotheta_v[b] = rte[j] + dt*(-tin[j - 1]*sol[j - 1])/vol[b];
            else sn->rhotheta_v[b] = rte[j] + dt*(-tin[j - 1]*sol[j - 1] + tin[j]*sol[j])/vol[b];
        }
        for (int j = 0; j < L - 1; ++j) {
            long long b = i + (long long)j*NSH, wi = i + (long long)(j + 1)*VPL;
            double din = 0.5*(sn->rho[gi + b] + sn->rho[gi + i + (long long)(j + 1)*NSH]);
            sn->wind[wi] = (2.0*sol[j]/area[wi] - din*so->wind[wi])/rio[j];
        }
        for (int j = 0; j < L; ++j) {
            long long b = i + (long long)j*NSH;
            sn->theta_v_pert[b] = sn->rhotheta_v[b]/sn->rho[gi + b] - c->g.theta_v_bg[b];
        }
        for (int j = 0; j < L; ++j) {
            long long b = i + (long long)j*NSH;
            sn->exner_pert[b] = so->exner_pert[b] + vol[b]*ga[j]*(sn->rhotheta_v[b] - so->rhotheta_v[b]);
        }
    }
}

void or_three_band_solver_gen_densities(const or_ctx_t *c, const game_state_view_t *so, game_state_view_t *sn,
                                        const game_state_view_t *td, const game_diag_view_t *dg_,
                                        game_irrev_view_t *irrev, const game_config_t *cfg, double dt, int rk_step)
{   /* column_solvers.c:385-639 */
    double iw = 0.5, ew = 1.0 - iw;
    int L = NL;
    const double *vol = c->g.volume, *area = c->g.area, *T = dg_->temperature;
    double *sh = irrev->condensates_sediment_heat;
    for (int k = 0; k < NCONST; ++k) {
        if (k == NCC) continue;
        double vsink = 0.0;
        if (k < NCC/4) vsink = cfg->snow_velocity;
        else if (k < NCC/2) vsink = cfg->rain_velocity;
        else if (k < NCC) vsink = cfg->cloud_droplets_velocity;
        #pragma omp parallel for
        for (int i = 0; i < NSH; ++i) {
            double cvv[64], dvv[64], evv[64], rvv[64], fi[64], fr[64], fe[64], sol[64];
            for (int j = 0; j < L - 1; ++j) {
                if (rk_step == 0 && k == 0) sh[(long long)j*NSH + i] = 0.0;
                long long b = i + (long long)j*NSH;
                fi[j] = so->wind[i + (long long)(j + 1)*VPL];
                fr[j] = sn->wind[i + (long long)(j + 1)*VPL];
                long long lo = i + (long long)(j + 1)*NSH, up = b;
                if (k < NCC/4) { fi[j] -= cfg->snow_velocity; fr[j] -= cfg->snow_velocity; }
                else if (k < NCC/2) { fi[j] -= cfg->rain_velocity; fr[j] -= cfg->rain_velocity; }
                else if (k < NCC) { fi[j] -= cfg->cloud_droplets_velocity; fr[j] -= cfg->cloud_droplets_velocity; }
                double a = area[i + (long long)(j + 1)*VPL];
                fi[j] = a*fi[j];
                fr[j] = a*fr[j];
                double doi, toi;
                if (fr[j] >= 0.0) { doi = so->rho[(long long)k*NS + lo]; toi = T[lo]; }
                else { doi = so->rho[(long long)k*NS + up]; toi = T[up]; }
                fr[j] = doi*fr[j];
                fe[j] = or_c_p_cond(k, toi)*toi*fr[j];
            }
            if (rk_step == 0 && k == 0) sh[(long long)(L - 1)*NSH + i] = 0.0;
            for (int j = 0; j < L - 1; ++j) {
                long long b = i + (long long)j*NSH;
                if (fi[j] >= 0.0) { cvv[j] = 0.0; evv[j] = -iw*dt/vol[b]*fi[j]; }
                else { cvv[j] = iw*dt/vol[i + (long long)(j + 1)*NSH]*fi[j]; evv[j] = 0.0; }
            }
            long long sfc_area = i + NV - NSH, low_s = (long long)k*NS + i + NS - NSH, low_t = i + NS - NSH;
            for (int j = 0; j < L; ++j) {
                long long b = i + (long long)j*NSH;
                if (j == 0) {
                    dvv[j] = (fi[0] >= 0.0) ? 1.0 : 1.0 - iw*dt/vol[b]*fi[0];
                } else if (j == L - 1) {
                    dvv[j] = (fi[j - 1] >= 0.0) ? 1.0 + iw*dt/vol[b]*fi[j - 1] : 1.0;
                    if (k < NCC) dvv[j] += iw*vsink*dt*area[sfc_area]/vol[b];
                } else {
                    dvv[j] = 1.0;
                    if (fi[j - 1] >= 0.0) dvv[j] += iw*dt/vol[b]*fi[j - 1];
                    if (fi[j] < 0.0) dvv[j] -= iw*dt/vol[b]*fi[j];
                }
                rvv[j] = so->rho[(long long)k*NS + b] + dt*td->rho[(long long)k*NS + b];
                if (j == 0) {
                    rvv[j] += ew*dt*fr[j]/vol[b];
                    if (rk_step == 0 && k < NCC) sh[b] += fe[j]/vol[b];
                } else if (j == L - 1) {
                    rvv[j] += -ew*dt*fr[j - 1]/vol[b];
                    if (rk_step == 0 && k < NCC) sh[b] += -fe[j - 1]/vol[b];
                    if (k < NCC) {
                        rvv[j] += -ew*vsink*dt*so->rho[low_s]*area[sfc_area]/vol[b];
                        if (rk_step == 0)
                            sh[b] += -vsink*T[low_t]*or_c_p_cond(k, T[low_t])*so->rho[low_s]*area[sfc_area]/vol[b];
                    }
                } else {
                    rvv[j] += ew*dt*(-fr[j - 1] + fr[j])/vol[b];
                    if (rk_step == 0 && k < NCC) sh[b] += (-fe[j - 1] + fe[j])/vol[b];
                }
            }
            or_thomas(cvv, dvv, evv, rvv, sol, L);
            for (int j = 0; j < L; ++j) if (sol[j] < 0.0) sol[j] = 0.0;
            for (int j = 0; j < L; ++j) sn->rho[(long long)k*NS + i + (long long)j*NSH] = sol[j];
        }
    }
}

int or_manage_rkhevi(const or_ctx_t *c, game_state_view_t *so, game_state_view_t *sn, game_state_view_t *td,
                     game_diag_view_t *dg_, game_forcings_view_t *f, game_irrev_view_t *irrev,
                     game_config_t *cfg, double dt)
{   /* manage_rkhevi.c:20-100 */
    if (cfg->momentum_diff_h || cfg->momentum_diff_v || cfg->mass_diff_h || cfg->mass_diff_v
        || cfg->temperature_diff_h || cfg->temperature_diff_v || cfg->rad_on || cfg->pbl_scheme
        || cfg->sfc_sensible_heat_flux || cfg->sfc_phase_trans || cfg->prog_soil_temp)
        return 5;
    if (c->g.no_of_oro_layers < 1 || c->d.n_layers < 3 || c->d.n_layers > 63) return 2;
    or_temperature_diagnostics(c, so, dg_->temperature);
    if (c->d.moisture_on == 1) or_calc_h2otracers_source_rates(c, so, dg_, cfg, irrev, dt);
    for (int rk = 0; rk < 2; ++rk) {
        if (rk == 0) or_manage_pressure_gradient(c, sn, dg_, f, irrev, cfg);
        or_calc_pressure_grad_condensates_v(c, sn, f, irrev);
        or_vector_tendencies_expl(c, sn, td, dg_, f, irrev, cfg, rk);
        #pragma omp parallel for
        for (int h = 0; h < NVH; ++h)
            for (int l = 0; l < NL; ++l) {
                long long vi = NSH + (long long)l*VPL + h;
                sn->wind[vi] = so->wind[vi] + dt*td->wind[vi];
            }
        or_scalar_tendencies_expl(c, so, sn, td, dg_, f, irrev, cfg, rk);
        or_three_band_solver_ver_waves(c, so, sn, td, cfg, dt, rk);
        if (NCONST > 1) or_three_band_solver_gen_densities(c, so, sn, td, dg_, irrev, cfg, dt, rk);
    }
    return 0;
}

int or_run_steps(const or_ctx_t *c, game_state_view_t *so, game_state_view_t *sn, game_state_view_t *td,
                 game_diag_view_t *dg_, game_forcings_view_t *f, game_irrev_view_t *irrev,
                 game_config_t *cfg, double dt, int n_steps)
{   /* coordinator.c:302-324 without output */
    for (int n = 0; n < n_steps; ++n) {
        or_linear_combine_two_states(c, sn, so, so, 1, 0);
        int rc = or_manage_rkhevi(c, so, sn, td, dg_, f, irrev, cfg, dt);
        if (rc) return rc;
        cfg->totally_first_step_bool = 0;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* integrals                                                             */
/* ------------------------------------------------------------------ */
double or_global_scalar_integrator(const or_ctx_t *c, const double *x, int naive)
{   /* write_output.c:1659-1668; compensated variant: Neumaier (1974) */
    const double *vol = c->g.volume;
    if (naive) {
        double r = 0.0;
        for (long long i = 0; i < NS; ++i) r += x[i]*vol[i];
        return r;
    }
    double s = 0.0, comp = 0.0;
    for (long long i = 0; i < NS; ++i) {
        double v = x[i]*vol[i], t = s + v;
        if (fabs(s) >= fabs(v)) comp += (s - t) + v;
        else comp += (v - t) + s;
        s = t;
    }
    return s + comp;
}

void or_integrals(const or_ctx_t *c, const game_state_view_t *s, const double *T, double *out, int naive)
{   /* write_output.c:1515-1596: masses, rhotheta_v, kinetic, potential, internal */
    for (int k = 0; k < NCONST; ++k) out[k] = or_global_scalar_integrator(c, &s->rho[(long long)k*NS], naive);
    out[NCONST] = or_global_scalar_integrator(c, s->rhotheta_v, naive);
    double *tmp = malloc(sizeof(double)*(size_t)NS);
    const double *rg = &s->rho[(long long)NCC*NS];
    or_inner_product(c, s->wind, s->wind, tmp);
    for (long long i = 0; i < NS; ++i) tmp[i] = rg[i]*tmp[i];
    out[NCONST + 1] = 0.5*or_global_scalar_integrator(c, tmp, naive);
    for (long long i = 0; i < NS; ++i) tmp[i] = rg[i]*c->g.gravity_potential[i];
    out[NCONST + 2] = or_global_scalar_integrator(c, tmp, naive);
    for (long long i = 0; i < NS; ++i) tmp[i] = rg[i]*T[i];
    out[NCONST + 3] = C_D_V*or_global_scalar_integrator(c, tmp, naive);
    free(tmp);
}
