/*
 * rkhevi_runtime.hip -- device-resident RKHEVI model behind the C ABI of
 * include/game_rkhevi.h.
 *
 * One process holds one model: the grid tables are uploaded once, the three
 * states (old, new, tendency) and every carried field (forcings, diagnostics,
 * irreversible quantities) stay in HBM, and the step is a fixed sequence of
 * kernel launches on one stream.  The sequence is manage_rkhevi
 * (src/time_stepping/manage_rkhevi.c:20-100).
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <type_traits>
#include <initializer_list>

#include "../../include/game_rkhevi.h"
#include "rkhevi_dev.h"
#include "rkhevi_kernels.h"
#include "rkhevi_adjoint.h"
#include "step_adjoint.h"
#include "adj_tables_host.h"
#include "runtime_internal.h"

KernelTimer *g_kernel_timer = nullptr;

namespace rt {

char g_err[512];

Model *M = nullptr;

int dalloc(void **p, size_t n)
{
    if (n == 0) n = 8;
    HIP_OK(hipMalloc(p, n));
    HIP_OK(hipMemsetAsync(*p, 0, n, M->stream));
    M->allocs.push_back(*p);
    M->bytes += (long long)n;
    return 0;
}

void dfree(void *p)
{
    if (!p || !M) return;
    for (size_t k = 0; k < M->allocs.size(); ++k)
        if (M->allocs[k] == p) {
            (void)hipStreamSynchronize(M->stream);
            (void)hipFree(p);
            M->allocs.erase(M->allocs.begin() + (long)k);
            return;
        }
}

template <class T>
int upload(const T **dst, const T *src, size_t count)
{
    void *p;
    int rc = dalloc(&p, count*sizeof(T));
    if (rc) return rc;
    if (src) HIP_OK(hipMemcpyAsync(p, src, count*sizeof(T), hipMemcpyHostToDevice, M->stream));
    *dst = (const T *)p;
    return 0;
}

int fzeros(float **dst, size_t count)
{
    void *p;
    int rc = dalloc(&p, count*sizeof(float));
    *dst = (float *)p;
    return rc;
}

int dzeros(double **dst, size_t count)
{
    void *p;
    int rc = dalloc(&p, count*sizeof(double));
    *dst = (double *)p;
    return rc;
}

/* what z_vector[k] reads in the reference's struct grid, where gravity_potential
   and gravity_m follow z_vector (src/game_types.h:76-79); see calc_pot_vort,
   src/spatial_operators/vorticities.c:48-71 */
double ref_z_vector_read(const game_dims_t &d, const game_grid_view_t &g, long long k)
{
    if (k < d.v) return g.z_vector[k];
    k -= d.v;
    if (k < d.s) return g.gravity_potential[k];
    k -= d.s;
    return g.gravity_m[k];
}

int check_config(const game_config_t *c)
{
    if (c->momentum_diff_h || c->momentum_diff_v || c->mass_diff_h || c->mass_diff_v || c->temperature_diff_h
        || c->temperature_diff_v || c->rad_on || c->pbl_scheme || c->sfc_sensible_heat_flux || c->sfc_phase_trans
        || c->prog_soil_temp) {
        snprintf(g_err, sizeof g_err, "a physics switch outside the RKHEVI dynamical-core path is on "
                 "(diffusion, radiation, PBL, surface fluxes and soil must be off)");
        return GAME_ERR_UNSUPPORTED;
    }
    return 0;
}

DevState *slot(int k) { return (k >= 0 && k < 3) ? &M->s[k] : nullptr; }

}  // namespace rt

using namespace rt;

extern "C" {

const char *game_gpu_last_error(void) { return g_err; }

long long game_gpu_bytes_allocated(void) { return M ? M->bytes : 0; }

static void drop_graphs(void)
{
    /* a replay still in flight finishes before its executable graph goes */
    if ((M->gexec[0] || M->gexec[1]) && M->stream) (void)hipStreamSynchronize(M->stream);
    for (int p = 0; p < 2; ++p) {
        if (M->gexec[p]) (void)hipGraphExecDestroy(M->gexec[p]);
        M->gexec[p] = nullptr;
        M->graph_key[p] = nullptr;
    }
    M->graph_dt = -1.0;
}

int game_gpu_finalize(void)
{
    if (!M) return 0;
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    drop_graphs();
    for (void *p : M->allocs) (void)hipFree(p);
    if (M->h_partials) (void)hipHostFree(M->h_partials);
    for (auto e : M->timer.pool) (void)hipEventDestroy(e);
    if (M->side) (void)hipStreamSynchronize(M->side);
    if (M->ev_fork) (void)hipEventDestroy(M->ev_fork);
    if (M->ev_join) (void)hipEventDestroy(M->ev_join);
    if (M->side) (void)hipStreamDestroy(M->side);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    g_kernel_timer = nullptr;
    delete M;
    M = nullptr;
    return 0;
}

static int init_model(const game_dims_t *dims, const game_grid_view_t *g, const game_dualgrid_view_t *dg,
                      const game_config_t *config, int precision);

/* M is either fully built or null: a failure part-way through the set-up
   releases what it had allocated, keeping the error message */
int game_gpu_init(const game_dims_t *dims, const game_grid_view_t *g, const game_dualgrid_view_t *dg,
                  const game_config_t *config, int precision)
{
    g_err[0] = 0;
    if (M) game_gpu_finalize();
    int rc = init_model(dims, g, dg, config, precision);
    if (rc && M) game_gpu_finalize();
    return rc;
}

static int init_model(const game_dims_t *dims, const game_grid_view_t *g, const game_dualgrid_view_t *dg,
                      const game_config_t *config, int precision)
{
    if (!dims || !g || !dg || !config) { snprintf(g_err, sizeof g_err, "null argument"); return GAME_ERR_ARG; }
    game_dims_t d = *dims;
    game_dims_derive(&d);
    if (d.n_layers < 3 || d.n_layers > GAME_MAX_LAYERS - 1 || g->no_of_oro_layers < 1
        || g->no_of_oro_layers >= d.n_layers || (d.moisture_on != 0 && d.moisture_on != 1)) {
        snprintf(g_err, sizeof g_err, "unsupported sizes: layers %d (3..63), orography layers %d (1..layers-1)",
                 d.n_layers, g->no_of_oro_layers);
        return GAME_ERR_ARG;
    }
    if (precision != GAME_FP64 && precision != GAME_MIXED) {
        snprintf(g_err, sizeof g_err, "precision mode %d is not available", precision);
        return GAME_ERR_ARG;
    }
    int rc = check_config(config);
    if (rc) return rc;
    int ndev = 0;
    HIP_OK(hipGetDeviceCount(&ndev));
    if (ndev < 1) { snprintf(g_err, sizeof g_err, "no HIP device"); return GAME_ERR_HIP; }
    static long long generations = 0;
    M = new Model();
    M->generation = ++generations;
    M->d = d;
    M->cfg = *config;
    M->precision = precision;
    HIP_OK(hipStreamCreate(&M->stream));
    HIP_OK(hipStreamCreate(&M->side));
    HIP_OK(hipEventCreateWithFlags(&M->ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&M->ev_join, hipEventDisableTiming));
    M->concurrent = getenv("GAME_GPU_ONE_STREAM") ? 0 : 1;
    DevGrid &G = M->G;
    G.nsh = d.nsh; G.nvh = d.nvh; G.nt = d.nt; G.vpl = d.vpl; G.dvpl = d.dvpl; G.L = d.n_layers;
    G.n_oro = g->no_of_oro_layers; G.n_const = d.n_const; G.n_cond = d.n_cond; G.moist = d.moisture_on;
    G.s = d.s; G.v = d.v; G.curl = d.curl; G.radius = g->radius; G.toa = g->toa;
    /* a flat surface gives slope == 0 exactly (grad_hor_cov of identical heights),
       so every terrain-correction term of the reference is an exact zero */
    G.flat = 1;
    for (size_t k = 0; k < (size_t)d.v && G.flat; ++k) if (g->slope[k] != 0.0) G.flat = 0;
    if (getenv("GAME_GPU_NO_FLAT_SHORTCUT")) G.flat = 0;
    /* per-cell kernels: one thread per cell and chunk of about three layers
       -- 163842 columns alone are only 2.5 waves per SIMD at res 7, while a
       thread that walks a few layers reads the cell's tables once and
       carries the values its layers share (measured at res 7 x 26: 9 chunks
       3.80 ms/step, 26 chunks 3.87, 3 chunks 3.89); results do not depend
       on the split */
    G.split_cell = (d.n_layers + 2)/3;
    if (const char *e = getenv("GAME_GPU_SPLIT_CELL")) {
        int v = atoi(e);
        if (v >= 1 && v <= d.n_layers) G.split_cell = v;
    }
    G.gd_streaming = getenv("GAME_GPU_GD_STREAMING") ? 1 : 0;
    G.gd_per_tracer = getenv("GAME_GPU_GD_PER_TRACER") ? 1 : 0;
    G.round_tendencies = 0;
    if (const char *e = getenv("GAME_GPU_ROUND_TENDENCIES")) {
        if (!strcmp(e, "rho")) G.round_tendencies = 1;
        else if (!strcmp(e, "rhotheta_v")) G.round_tendencies = 2;
        else if (!strcmp(e, "both")) G.round_tendencies = 3;
    }
    G.vw_streaming = getenv("GAME_GPU_VW_STREAMING") ? 1 : 0;
    G.vw_rows = getenv("GAME_GPU_VW_ROWS") ? 1 : 0;
    G.sh_per_box = getenv("GAME_GPU_SH_PER_BOX") ? 1 : 0;
    G.unfused_phase = getenv("GAME_GPU_UNFUSED_PHASE") ? 1 : 0;
    M->use_graphs = getenv("GAME_GPU_NO_GRAPH") ? 0 : 1;
    G.split_edge = 2;
    if (const char *e = getenv("GAME_GPU_SPLIT_EDGE")) {
        int v = atoi(e);
        if (v >= 1 && v <= d.n_layers) G.split_edge = v;
    }
    size_t S = (size_t)d.s, V = (size_t)d.v, NVH = (size_t)d.nvh, NSH = (size_t)d.nsh, NT = (size_t)d.nt;
#define UP(dst, src, n) do { if ((rc = upload(&(dst), (src), (n)))) return rc; } while (0)
    UP(G.nd, g->normal_distance, V); UP(G.vol, g->volume, S); UP(G.area, g->area, V); UP(G.zs, g->z_scalar, S);
    UP(G.zv, g->z_vector, V); UP(G.phi, g->gravity_potential, S); UP(G.gm, g->gravity_m, V);
    UP(G.slope, g->slope, V); UP(G.thbg, g->theta_v_bg, S); UP(G.exbg, g->exner_bg, S);
    UP(G.exbg_grad, g->exner_bg_grad, V); UP(G.ipw, g->inner_product_weights, 8*S);
    UP(G.tw, g->trsk_weights, 10*NVH); UP(G.drw, g->density_to_rhombi_weights, 4*NVH);
    UP(G.dnd, dg->normal_distance, (size_t)d.dual_v); UP(G.dzv, dg->z_vector, (size_t)d.dual_v);
    UP(G.darea, dg->area, (size_t)d.curl); UP(G.fvec, dg->f_vec, 2*NVH);
    UP(G.from, g->from_index, NVH); UP(G.to, g->to_index, NVH); UP(G.adj, g->adjacent_vector_indices_h, 6*NSH);
    UP(G.sgn, g->adjacent_signs_h, 6*NSH); UP(G.ti, g->trsk_indices, 10*NVH);
    UP(G.tmc, g->trsk_modified_curl_indices, 10*NVH); UP(G.dri, g->density_to_rhombi_indices, 4*NVH);
    UP(G.dfrom, dg->from_index, NVH); UP(G.dto, dg->to_index, NVH);
    UP(G.vit, dg->vorticity_indices_triangles, 3*NT); UP(G.vst, dg->vorticity_signs_triangles, 3*NT);
    /* per-edge differences of calc_pot_vort's density extrapolation (vorticities.c:48-71) */
    {
        int L = d.n_layers;
        long long Sll = d.s, vpl = d.vpl;
        std::vector<double> t1(NVH), t2(NVH), s1(NVH), s2(NVH);
        for (size_t e = 0; e < NVH; ++e) {
            t1[e] = ref_z_vector_read(d, *g, Sll + (long long)e) - ref_z_vector_read(d, *g, Sll + vpl + (long long)e);
            t2[e] = g->z_vector[0] - ref_z_vector_read(d, *g, Sll + (long long)e);
            s1[e] = ref_z_vector_read(d, *g, Sll + (long long)(L - 2)*vpl + (long long)e)
                    - ref_z_vector_read(d, *g, Sll + (long long)(L - 1)*vpl + (long long)e);
            s2[e] = 0.5*(g->z_vector[(long long)L*vpl + g->from_index[e]] + g->z_vector[(long long)L*vpl + g->to_index[e]])
                    - ref_z_vector_read(d, *g, Sll + (long long)(L - 1)*vpl + (long long)e);
        }
        UP(G.pv_top_dz1, t1.data(), NVH); UP(G.pv_top_dz2, t2.data(), NVH);
        UP(G.pv_sfc_dz1, s1.data(), NVH); UP(G.pv_sfc_dz2, s2.data(), NVH);
        HIP_OK(hipStreamSynchronize(M->stream));
    }
    /* Klemp damping coefficients (column_solvers.c:284-298): a grid constant */
    {
        int L = d.n_layers;
        double dsh = config->damping_start_height_over_toa*g->z_vector[0];
        std::vector<double> dmp((size_t)(L - 1)*NSH);
        for (int j = 0; j < L - 1; ++j)
            for (size_t i = 0; i < NSH; ++i) {
                double zad = g->z_vector[i + (size_t)(j + 1)*d.vpl] - dsh, dc;
                if (zad < 0.0) dc = 0.0;
                else dc = config->damping_coeff_max*pow(sin(0.5*M_PI*zad/(g->z_vector[0] - dsh)), 2);
                dmp[(size_t)j*NSH + i] = dc;
            }
        UP(G.damp, dmp.data(), dmp.size());
        HIP_OK(hipStreamSynchronize(M->stream));
    }
    /* cell-ordered copy of the inner-product weights */
    {
        const int L = d.n_layers;
        std::vector<double> t((size_t)8*S);
        for (int l = 0; l < L; ++l)
            for (int j = 0; j < 8; ++j)
                for (size_t h = 0; h < NSH; ++h)
                    t[((size_t)l*8 + j)*NSH + h] = g->inner_product_weights[8*((size_t)l*NSH + h) + j];
        UP(G.ipwT, t.data(), t.size());
        HIP_OK(hipStreamSynchronize(M->stream));
    }
    if ((rc = dalloc((void **)&G.idv, sizeof(double)*(size_t)d.s))) return rc;
    M->idv_dt = -1.0;
#undef UP
    for (int k = 0; k < 3; ++k) {
        DevState &s = M->s[k];
        if ((rc = dzeros(&s.rho, (size_t)d.n_const*S))) return rc;
        if ((rc = dzeros(&s.rt, S))) return rc;
        if ((rc = dzeros(&s.tp, S))) return rc;
        if ((rc = dzeros(&s.ep, S))) return rc;
        if ((rc = dzeros(&s.wind, V))) return rc;
    }
    if ((rc = dzeros(&M->tmp_s, S))) return rc;
    if (precision == GAME_FP64) {
        DevWork &W = M->W;
        double **vfields[] = {&W.flux, &W.flux2, &W.pgrad_comb, &W.pgrad_old, &W.pgrad_nl, &W.pgrad_l, &W.pgc_v, &W.vsq_grad,
                              &W.pvt};
        for (double **f : vfields) if ((rc = dzeros(f, V))) return rc;
        double **sfields[] = {&W.flux_divv, &W.temp, &W.vsq, &W.phr, &W.sed_heat};
        for (double **f : sfields) if ((rc = dzeros(f, S))) return rc;
        if ((rc = dzeros(&W.rvt, (size_t)d.dual_v_vectors))) return rc;
        if ((rc = dzeros(&W.vcc, 3*(size_t)G.n_oro*NSH))) return rc;
        if ((rc = dzeros(&W.gas_frac, S))) return rc;
        if ((rc = dzeros(&W.pv, (size_t)d.curl))) return rc;
        if ((rc = dzeros(&W.ptr, (size_t)(d.n_cond + 1)*S))) return rc;
        if ((rc = dzeros(&W.cpc, (size_t)(d.n_cond > 0 ? d.n_cond : 1)*S))) return rc;
        W.td_wind = M->s[GAME_STATE_TENDENCY].wind;
    } else {
        /* mixed: every work array FP32; the tendency state keeps FP64 densities
           and rhotheta_v, the wind tendency is FP32 */
        DevWorkF &W = M->WF;
        float **vfields[] = {&W.flux, &W.pgrad_comb, &W.pgrad_old, &W.pgrad_nl, &W.pgrad_l, &W.pgc_v, &W.vsq_grad, &W.pvt,
                             &W.td_wind};
        for (float **f : vfields) if ((rc = fzeros(f, V))) return rc;
        float **sfields[] = {&W.temp, &W.vsq, &W.phr, &W.sed_heat};
        for (float **f : sfields) if ((rc = fzeros(f, S))) return rc;
        if ((rc = fzeros(&W.rvt, (size_t)d.dual_v_vectors))) return rc;
        if ((rc = fzeros(&W.pv, (size_t)d.curl))) return rc;
        if ((rc = fzeros(&W.ptr, (size_t)(d.n_cond + 1)*S))) return rc;
        if ((rc = fzeros(&W.cpc, (size_t)(d.n_cond > 0 ? d.n_cond : 1)*S))) return rc;
        if ((rc = dzeros(&W.vcc, 3*(size_t)G.n_oro*NSH))) return rc;
        if ((rc = dzeros(&W.gas_frac, S))) return rc;
        if ((rc = dzeros(&W.flux2, 8))) return rc;
        if ((rc = dzeros(&W.flux_divv, S))) return rc;
    }
    /* the device address of the allocation-free solver's scratch on this device */
    if (adj::thomas_scratch_resolve(g_err, sizeof g_err)) return GAME_ERR_HIP;
    if ((rc = dzeros(&M->partials, (size_t)2*col::INTEGRAL_BLOCKS*INTEGRALS_MAX))) return rc;
    HIP_OK(hipHostMalloc((void **)&M->h_partials, sizeof(double)*2*col::INTEGRAL_BLOCKS*INTEGRALS_MAX,
                         hipHostMallocDefault));
    StepConsts &K = M->K;
    K.impl_weight = config->impl_thermo_weight;
    K.damping_start = config->damping_start_height_over_toa*g->z_vector[0];
    K.damping_max = config->damping_coeff_max;
    K.snow_v = config->snow_velocity;
    K.rain_v = config->rain_velocity;
    K.cloud_v = config->cloud_droplets_velocity;
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

int game_gpu_set_config(const game_config_t *config)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!config) { snprintf(g_err, sizeof g_err, "null config"); return GAME_ERR_ARG; }
    int rc = check_config(config);
    if (rc) return rc;
    M->cfg = *config;
    drop_graphs();
    M->K.impl_weight = config->impl_thermo_weight;
    M->K.snow_v = config->snow_velocity;
    M->K.rain_v = config->rain_velocity;
    M->K.cloud_v = config->cloud_droplets_velocity;
    return 0;
}

static int copy_state(int k, const game_state_view_t *h, int to_device)
{
    if (!M) return GAME_ERR_NOT_INIT;
    DevState *s = slot(k);
    if (!s || !h) { snprintf(g_err, sizeof g_err, "bad state slot %d", k); return GAME_ERR_ARG; }
    size_t S = (size_t)M->d.s*sizeof(double), V = (size_t)M->d.v*sizeof(double);
    struct { double *dev; double *host; size_t n; } f[] = {
        {s->rho, h->rho, S*(size_t)M->d.n_const}, {s->rt, h->rhotheta_v, S}, {s->tp, h->theta_v_pert, S},
        {s->ep, h->exner_pert, S}, {s->wind, h->wind, V}};
    for (auto &x : f) {
        if (!x.host) continue;
        if (to_device) HIP_OK(hipMemcpyAsync(x.dev, x.host, x.n, hipMemcpyHostToDevice, M->stream));
        else HIP_OK(hipMemcpyAsync(x.host, x.dev, x.n, hipMemcpyDeviceToHost, M->stream));
    }
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

int game_gpu_set_state(int k, const game_state_view_t *h) { return copy_state(k, h, 1); }
int game_gpu_get_state(int k, game_state_view_t *h) { return copy_state(k, h, 0); }

static int d2h(double *host, const double *dev, size_t n)
{
    if (host) HIP_OK(hipMemcpyAsync(host, dev, n*sizeof(double), hipMemcpyDeviceToHost, M->stream));
    return 0;
}

}  /* extern "C" */

/* FP32 work array of the mixed form, widened to the caller's doubles */
static int d2h(double *host, const float *dev, size_t n)
{
    if (!host) return 0;
    std::vector<float> tmp(n);
    HIP_OK(hipMemcpyAsync(tmp.data(), dev, n*sizeof(float), hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    for (size_t i = 0; i < n; ++i) host[i] = (double)tmp[i];
    return 0;
}

template <class WK>
static int get_forcings_t(const WK &W, game_forcings_view_t *h)
{
    size_t V = (size_t)M->d.v;
    int rc;
    if ((rc = d2h(h->pgrad_acc_old, W.pgrad_old, V))) return rc;
    if ((rc = d2h(h->pressure_gradient_acc_neg_nl, W.pgrad_nl, V))) return rc;
    if ((rc = d2h(h->pressure_gradient_acc_neg_l, W.pgrad_l, V))) return rc;
    if ((rc = d2h(h->pressure_grad_condensates_v, W.pgc_v, V))) return rc;
    if ((rc = d2h(h->v_squared_grad, W.vsq_grad, V))) return rc;
    if ((rc = d2h(h->pot_vort_tend, W.pvt, V))) return rc;
    if (h->radiation_tendency) memset(h->radiation_tendency, 0, sizeof(double)*(size_t)M->d.s);
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

template <class WK>
static int get_diagnostics_t(const WK &W, game_diag_view_t *h)
{
    int rc;
    if ((rc = d2h(h->flux_density, W.flux, (size_t)M->d.v))) return rc;
    if ((rc = d2h(h->flux_density_divv, W.flux_divv, (size_t)M->d.s))) return rc;
    if ((rc = d2h(h->rel_vort_on_triangles, W.rvt, (size_t)M->d.dual_v_vectors))) return rc;
    if ((rc = d2h(h->pot_vort, W.pv, (size_t)M->d.curl))) return rc;
    if ((rc = d2h(h->temperature, W.temp, (size_t)M->d.s))) return rc;
    if ((rc = d2h(h->v_squared, W.vsq, (size_t)M->d.s))) return rc;
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

template <class WK>
static int get_irrev_t(const WK &W, game_irrev_view_t *h)
{
    int rc;
    size_t S = (size_t)M->d.s;
    if ((rc = d2h(h->condensates_sediment_heat, W.sed_heat, S))) return rc;
    if ((rc = d2h(h->phase_trans_rates, W.ptr, (size_t)(M->d.n_cond + 1)*S))) return rc;
    if ((rc = d2h(h->phase_trans_heating_rate, W.phr, S))) return rc;
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

/* manage_rkhevi.c:20-100 on the device-resident states; td_wind is the
   tendency state's wind (FP64 form) or the FP32 wind tendency (mixed).
   The step's entry condition is state_new == state_old (src/coordinator.c:305),
   so whatever rk 0 reads of state_new before writing it is read from
   state_old: the step then also runs on a state_new buffer that holds an
   older state, which lets game_gpu_run_steps swap the two buffers instead of
   copying. */
/* the five fields of a state, device to device */
static int copy_state_d2d(const DevState &dst, const DevState &src, hipStream_t st)
{
    const size_t S = (size_t)M->d.s*sizeof(double), V = (size_t)M->d.v*sizeof(double);
    HIP_OK(hipMemcpyAsync(dst.rho, src.rho, S*(size_t)M->d.n_const, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(dst.rt, src.rt, S, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(dst.tp, src.tp, S, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(dst.ep, src.ep, S, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(dst.wind, src.wind, V, hipMemcpyDeviceToDevice, st));
    return 0;
}

static int copy_tendency(double *rho, double *rt_, double *wind, const DevState &td, hipStream_t st)
{
    const size_t S = (size_t)M->d.s*sizeof(double), V = (size_t)M->d.v*sizeof(double);
    HIP_OK(hipMemcpyAsync(rho, td.rho, S, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(rt_, td.rt, S, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemcpyAsync(wind, td.wind, V, hipMemcpyDeviceToDevice, st));
    return 0;
}

template <class WK, typename TW>
static int step_t(const WK &W, TW *td_wind, double dt, int first, const StepTape *tape = nullptr)
{
    hipStream_t st = M->stream;
    const DevGrid &G = M->G;
    DevState &so = M->s[GAME_STATE_OLD], &sn = M->s[GAME_STATE_NEW], &td = M->s[GAME_STATE_TENDENCY];
    M->K.dt = dt;
    int rc;
    if (tape && (rc = copy_state_d2d(tape->s, so, st))) return rc;
    if (G.n_const > 1 && dt != M->idv_dt) {
        col::tracer_quotients(G, dt, st);
        M->idv_dt = dt;
    }
    if (G.moist && G.n_const == 6 && !G.unfused_phase)
        col::temperature_phase_trans(G, so, W.temp, W.cpc, W.gas_frac, W.ptr, W.phr, dt, st);
    else {
        col::temperature(G, so, W.temp, G.n_cond > 0 ? W.cpc : nullptr, W.gas_frac, st);
        if (G.moist) col::phase_trans(G, so, W.temp, W.cpc, W.ptr, W.phr, dt, st);
    }
    /* per-kernel events time each kernel alone: with them on, one stream */
    const bool conc = M->concurrent && !g_kernel_timer;
    for (int rk = 0; rk < 2; ++rk) {
        const DevState &snr = rk == 0 ? so : sn;           /* state_new before this substep writes it */
        DevState sns = snr;                                 /* ... with the horizontal wind just updated */
        sns.wind = sn.wind;
        const double *rho_gas = snr.rho + (i64)G.n_cond*G.s;
        if ((rk == 1 || first) && conc) {
            /* the pressure gradient and the kinetic energy on the side
               stream, the PV and TRSK flux on this one: they read the same
               state and write different arrays */
            HIP_OK(hipEventRecord(M->ev_fork, st));
            HIP_OK(hipStreamWaitEvent(M->side, M->ev_fork, 0));
            if (rk == 0) hor::pgrad(G, snr, W, first, M->cfg.impl_thermo_weight, M->side);
            else hor::pgc_v(G, snr, W, M->side);
            hor::advection_kinetic(G, snr.wind, W, M->side);
            hor::advection_pv(G, snr.wind, rho_gas, W, st);
            HIP_OK(hipEventRecord(M->ev_join, M->side));
            HIP_OK(hipStreamWaitEvent(st, M->ev_join, 0));
        } else {
            if (rk == 0) hor::pgrad(G, snr, W, first, M->cfg.impl_thermo_weight, st);
            else hor::pgc_v(G, snr, W, st);
            if (rk == 1 || first) hor::advection(G, snr.wind, rho_gas, W, st);
        }
        if constexpr (std::is_same<WK, DevWork>::value)
            if (tape && rk == 1) {
                HIP_OK(hipMemcpyAsync(tape->flux, W.flux, (size_t)M->d.v*sizeof(double), hipMemcpyDeviceToDevice, st));
                HIP_OK(hipMemcpyAsync(tape->pv, W.pv, (size_t)M->d.curl*sizeof(double), hipMemcpyDeviceToDevice, st));
            }
        hor::vector_tend(G, W, td_wind, so.wind, sn.wind, rk == 1 ? 0.5 : 1.0, M->cfg.impl_thermo_weight, dt, st);
        hor::scalar_tend(G, sns, td, W, rk, st);
        if (sizeof(TW) == sizeof(double)) col::ver_waves(G, so, snr, sn, td, M->K, rk, st);
        else col::ver_waves(G, so, snr, sn, td, (const float *)td_wind, M->K, rk, st);
        if (G.n_const > 1) {
            if (conc && rk == 0 && !tape) {
                /* the first substep's tracer solve (and sediment heat) on the
                   side stream: it writes only the tracer densities and the
                   sediment heat, which the second substep's PV chain does not
                   read; its pressure-gradient kernel, which reads all the
                   densities, follows on the side stream */
                HIP_OK(hipEventRecord(M->ev_fork, st));
                HIP_OK(hipStreamWaitEvent(M->side, M->ev_fork, 0));
                col::gen_densities(G, so, sn, td, W.temp, W.cpc, W.sed_heat, M->K, rk, M->side);
            } else col::gen_densities(G, so, sn, td, W.temp, W.cpc, W.sed_heat, M->K, rk, st);
        }
        if (tape && rk == 0) {
            if ((rc = copy_state_d2d(tape->s1, sn, st))) return rc;
            if ((rc = copy_tendency(tape->td0_rho, tape->td0_rt, tape->td0_wind, td, st))) return rc;
        }
        if (tape && rk == 1 && (rc = copy_tendency(tape->td1_rho, tape->td1_rt, tape->td1_wind, td, st))) return rc;
    }
    HIP_OK(hipGetLastError());
    return 0;
}

int rt::step_fp64(double dt, int first, const StepTape *tape)
{
    return step_t(M->W, M->s[GAME_STATE_TENDENCY].wind, dt, first, tape);
}

extern "C" {

int game_gpu_get_forcings(game_forcings_view_t *h)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!h) return GAME_ERR_ARG;
    return M->precision == GAME_MIXED ? get_forcings_t(M->WF, h) : get_forcings_t(M->W, h);
}

int game_gpu_get_diagnostics(game_diag_view_t *h)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!h) return GAME_ERR_ARG;
    return M->precision == GAME_MIXED ? get_diagnostics_t(M->WF, h) : get_diagnostics_t(M->W, h);
}

int game_gpu_get_irrev(game_irrev_view_t *h)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!h) return GAME_ERR_ARG;
    return M->precision == GAME_MIXED ? get_irrev_t(M->WF, h) : get_irrev_t(M->W, h);
}

/* manage_rkhevi.c:20-100 on the device-resident states */
int game_gpu_manage_rkhevi(double dt, int first)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (M->precision == GAME_MIXED) return step_t(M->WF, M->WF.td_wind, dt, first);
    return step_t(M->W, M->s[GAME_STATE_TENDENCY].wind, dt, first);
}

int game_gpu_linear_combine_two_states(int k0, int k1, int ko, double c0, double c1)
{
    if (!M) return GAME_ERR_NOT_INIT;
    DevState *a = slot(k0), *b = slot(k1), *o = slot(ko);
    if (!a || !b || !o) { snprintf(g_err, sizeof g_err, "bad state slot"); return GAME_ERR_ARG; }
    col::linear_combine(M->G, *a, *b, *o, c0, c1, M->stream);
    HIP_OK(hipGetLastError());
    return 0;
}

}  /* extern "C" */

/* one steady step (not the first) replayed from a HIP graph: the step's
   launches depend only on the state buffers, dt and the configuration, so
   one captured step per assignment of the two state buffers serves every
   later step with the same dt.  The tracer quotient table of a new dt is
   refreshed before the capture (outside the graph); no memory set or copy is
   recorded.  Returns 1 when the step must run eagerly instead. */
static int step_from_graph(double dt, int *rc)
{
    *rc = 0;
    if (!M->use_graphs || g_kernel_timer) return 1;
    hipStream_t st = M->stream;
    if (M->G.n_const > 1 && dt != M->idv_dt) {
        col::tracer_quotients(M->G, dt, st);
        M->idv_dt = dt;
    }
    if (dt != M->graph_dt) {
        drop_graphs();
        M->graph_dt = dt;
    }
    const void *key = M->s[GAME_STATE_OLD].rho;
    int p = M->graph_key[0] == key ? 0 : (M->graph_key[1] == key ? 1 : -1);
    if (p < 0) {
        p = M->graph_key[0] == nullptr ? 0 : 1;
        if (M->gexec[p]) {
            (void)hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(M->gexec[p]);
            M->gexec[p] = nullptr;
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            snprintf(g_err, sizeof g_err, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
            *rc = GAME_ERR_HIP;
            return 0;
        }
        int r = M->precision == GAME_MIXED ? step_t(M->WF, M->WF.td_wind, dt, 0)
                                           : step_t(M->W, M->s[GAME_STATE_TENDENCY].wind, dt, 0);
        e = hipStreamEndCapture(st, &g);
        if (r) { if (g) (void)hipGraphDestroy(g); *rc = r; return 0; }
        if (e != hipSuccess) {
            snprintf(g_err, sizeof g_err, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
            *rc = GAME_ERR_HIP;
            return 0;
        }
        e = hipGraphInstantiate(&M->gexec[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            M->gexec[p] = nullptr;
            snprintf(g_err, sizeof g_err, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
            *rc = GAME_ERR_HIP;
            return 0;
        }
        M->graph_key[p] = key;
    }
    hipError_t e = hipGraphLaunch(M->gexec[p], st);
    if (e != hipSuccess) {
        snprintf(g_err, sizeof g_err, "hipGraphLaunch failed: %s", hipGetErrorString(e));
        *rc = GAME_ERR_HIP;
    }
    return 0;
}

extern "C" {

int game_gpu_run_steps(int n, double dt, int first)
{
    if (!M) return GAME_ERR_NOT_INIT;
    for (int k = 0; k < n; ++k) {
        /* src/coordinator.c:305 copies state_new into state_old; exchanging the
           two buffers gives the step the same inputs (see step_t): the fields
           the step never writes (top and surface w) are equal in both */
        DevState t = M->s[GAME_STATE_OLD];
        M->s[GAME_STATE_OLD] = M->s[GAME_STATE_NEW];
        M->s[GAME_STATE_NEW] = t;
        int rc = 0;
        /* the first step takes other branches (pgrad_old, rk-0 advection): eager */
        if (first || step_from_graph(dt, &rc)) rc = game_gpu_manage_rkhevi(dt, first);
        if (rc) return rc;
        first = 0;
    }
    return 0;
}

int game_gpu_temperature_diagnostics(int k)
{
    if (!M) return GAME_ERR_NOT_INIT;
    DevState *s = slot(k);
    if (!s) return GAME_ERR_ARG;
    if (M->precision == GAME_MIXED) col::temperature(M->G, *s, M->WF.temp, (float *)nullptr, nullptr, M->stream);
    else col::temperature(M->G, *s, M->W.temp, (double *)nullptr, nullptr, M->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

int game_gpu_inner_product(const double *a, const double *b, double *out)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!a || !b || !out) { snprintf(g_err, sizeof g_err, "null buffer argument"); return GAME_ERR_ARG; }
    size_t V = (size_t)M->d.v, S = (size_t)M->d.s;
    double *da = M->s[GAME_STATE_OLD].wind, *db = M->s[GAME_STATE_NEW].wind;
    HIP_OK(hipMemcpyAsync(da, a, V*sizeof(double), hipMemcpyHostToDevice, M->stream));
    if (b != a) HIP_OK(hipMemcpyAsync(db, b, V*sizeof(double), hipMemcpyHostToDevice, M->stream));
    hor::inner_product(M->G, da, b != a ? db : da, M->tmp_s, M->stream);
    HIP_OK(hipMemcpyAsync(out, M->tmp_s, S*sizeof(double), hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

}  /* extern "C" */

/* the compensated sum of integral q's block partials (host side) */
static double merge_partials(int q)
{
    const double *p = M->h_partials + (size_t)2*col::INTEGRAL_BLOCKS*q;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < col::INTEGRAL_BLOCKS; ++b) {
        double v = p[2*b], t = s + v;
        if (fabs(s) >= fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
        c += p[2*b + 1];
    }
    return s + c;
}

/* write_output.c:1515-1596 with compensated sums: every integral's block
   partials are formed by one launch each, copied back together and merged on
   the host.  The internal energy multiplies by the diagnostics temperature,
   as the reference does (diagnostics->temperature, write_output.c:1584-1588):
   the temperature of the state the last step started from (manage_rkhevi.c:29),
   or of the slot last passed to game_gpu_temperature_diagnostics. */
template <class WK>
static int integrals_t(const WK &W, const DevState &s, double *out)
{
    const DevGrid &G = M->G;
    const int nint = G.n_const + 4;
    double *P = M->partials;
    const size_t stride = (size_t)2*col::INTEGRAL_BLOCKS;
    for (int c = 0; c < G.n_const; ++c)
        col::integral_partials(G.s, s.rho + (i64)c*G.s, (const double *)nullptr, G.vol, P + stride*c, M->stream);
    col::integral_partials(G.s, s.rt, (const double *)nullptr, G.vol, P + stride*G.n_const, M->stream);
    const double *rg = s.rho + (i64)G.n_cond*G.s;
    hor::inner_product(G, s.wind, s.wind, M->tmp_s, M->stream);
    col::integral_partials(G.s, rg, M->tmp_s, G.vol, P + stride*(G.n_const + 1), M->stream);
    col::integral_partials(G.s, rg, G.phi, G.vol, P + stride*(G.n_const + 2), M->stream);
    col::integral_partials(G.s, rg, W.temp, G.vol, P + stride*(G.n_const + 3), M->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(M->h_partials, P, sizeof(double)*stride*(size_t)nint, hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    for (int q = 0; q < nint; ++q) out[q] = merge_partials(q);
    out[G.n_const + 1] *= 0.5;
    out[G.n_const + 3] *= GC_C_D_V;
    return 0;
}

extern "C" {

int game_gpu_integrals(int k, double *out)
{
    if (!M) return GAME_ERR_NOT_INIT;
    DevState *s = slot(k);
    if (!s || !out) return GAME_ERR_ARG;
    return M->precision == GAME_MIXED ? integrals_t(M->WF, *s, out) : integrals_t(M->W, *s, out);
}

/* the records write_out_integral appends (src/io/write_output.c:1515-1596):
   "masses" (time, every constituent's mass), "potential_temperature_density"
   (time, the rhotheta_v integral) and "energy" (time, kinetic, potential,
   internal), tab-separated with the reference's %lf format */
int game_gpu_write_integrals(int k, double time_since_init, const char *directory)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!directory) { snprintf(g_err, sizeof g_err, "null directory"); return GAME_ERR_ARG; }
    double out[INTEGRALS_MAX];
    int rc = game_gpu_integrals(k, out);
    if (rc) return rc;
    const int nc = M->d.n_const;
    const char *names[3] = {"masses", "potential_temperature_density", "energy"};
    for (int f = 0; f < 3; ++f) {
        std::string path = std::string(directory) + "/" + names[f];
        FILE *fp = fopen(path.c_str(), "a");
        if (!fp) { snprintf(g_err, sizeof g_err, "cannot open %s for appending", path.c_str()); return GAME_ERR_ARG; }
        fprintf(fp, "%lf\t", time_since_init);
        if (f == 0)
            for (int c = 0; c < nc; ++c) fprintf(fp, c == nc - 1 ? "%lf\n" : "%lf\t", out[c]);
        else if (f == 1)
            fprintf(fp, "%lf\n", out[nc]);
        else
            fprintf(fp, "%lf\t%lf\t%lf\n", out[nc + 1], out[nc + 2], out[nc + 3]);
        if (fclose(fp)) { snprintf(g_err, sizeof g_err, "cannot write %s", path.c_str()); return GAME_ERR_ARG; }
    }
    return 0;
}

/* the reference's time loop with its integral output (src/coordinator.c:302-336):
   per step the driver's copy and manage_rkhevi (as game_gpu_run_steps), then
   the three records of the new state at time t_start + (k + 1) dt */
int game_gpu_run_steps_write_integrals(int n, double dt, int first, double t_start, const char *directory)
{
    if (!M) return GAME_ERR_NOT_INIT;
    for (int k = 0; k < n; ++k) {
        int rc = game_gpu_run_steps(1, dt, first);
        if (rc) return rc;
        first = 0;
        if ((rc = game_gpu_write_integrals(GAME_STATE_NEW, t_start + (k + 1)*dt, directory))) return rc;
    }
    return 0;
}

}  /* extern "C" */

template <class WK>
static int balance_terms_t(const WK &W, const double *wind, const double *density, double *v_squared, double *pvt)
{
    DevState &t = M->s[GAME_STATE_TENDENCY];   /* used as scratch, cleared after */
    size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    typedef typename std::remove_pointer<decltype(W.vsq)>::type TW;
    HIP_OK(hipMemcpyAsync(t.wind, wind, V*sizeof(double), hipMemcpyHostToDevice, M->stream));
    HIP_OK(hipMemcpyAsync(t.rt, density, S*sizeof(double), hipMemcpyHostToDevice, M->stream));
    hor::advection(M->G, t.wind, t.rt, W, M->stream);
    int rc;
    if ((rc = d2h(v_squared, W.vsq, S))) return rc;
    if ((rc = d2h(pvt, W.pvt, V))) return rc;
    HIP_OK(hipMemsetAsync(t.wind, 0, V*sizeof(double), M->stream));
    HIP_OK(hipMemsetAsync(t.rt, 0, S*sizeof(double), M->stream));
    HIP_OK(hipMemsetAsync(W.pvt, 0, V*sizeof(TW), M->stream));
    HIP_OK(hipMemsetAsync(W.vsq, 0, S*sizeof(TW), M->stream));
    HIP_OK(hipMemsetAsync(W.vsq_grad, 0, V*sizeof(TW), M->stream));
    HIP_OK(hipMemsetAsync(W.flux, 0, V*sizeof(TW), M->stream));
    HIP_OK(hipMemsetAsync(W.pv, 0, (size_t)M->d.curl*sizeof(TW), M->stream));
    HIP_OK(hipMemsetAsync(W.rvt, 0, (size_t)M->d.dual_v_vectors*sizeof(TW), M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

extern "C" {

/* v_squared and pot_vort_tend of a (wind, density) pair, the terms of the
   initial hydrostatic balance (src/io/set_initial_state.c:155-166) */
int game_gpu_init_balance_terms(const double *wind, const double *density, double *v_squared, double *pvt)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!wind || !density) { snprintf(g_err, sizeof g_err, "null wind or density"); return GAME_ERR_ARG; }
    if (M->precision == GAME_MIXED) return balance_terms_t(M->WF, wind, density, v_squared, pvt);
    return balance_terms_t(M->W, wind, density, v_squared, pvt);
}

int game_gpu_enable_timing(int on)
{
    if (!M) return GAME_ERR_NOT_INIT;
    g_kernel_timer = on ? &M->timer : nullptr;
    return 0;
}

int game_gpu_synchronize(void)
{
    if (!M) return GAME_ERR_NOT_INIT;
    HIP_OK(hipStreamSynchronize(M->stream));
    M->timer.collect();
    return 0;
}

int game_gpu_reset_timing(void)
{
    if (!M) return GAME_ERR_NOT_INIT;
    game_gpu_synchronize();
    M->timer.stats.clear();
    return 0;
}

int game_gpu_kernel_count(void)
{
    if (!M) return 0;
    M->timer.collect();
    return (int)M->timer.stats.size();
}

int game_gpu_kernel_stats(int k, const char **name, long long *launches, double *ms)
{
    if (!M) return GAME_ERR_NOT_INIT;
    M->timer.collect();
    if (k < 0 || k >= (int)M->timer.stats.size()) return GAME_ERR_ARG;
    *name = M->timer.stats[k].name.c_str();
    *launches = M->timer.stats[k].launches;
    *ms = M->timer.stats[k].ms;
    return 0;
}


/* ------------------------------------------------------------------ */
/* operator-level interface and adjoints (device pointers, FP64)        */
/* ------------------------------------------------------------------ */
}  /* extern "C" */

namespace rt {
static int upload_ints(const int **dst, const std::vector<int> &v)
{
    return upload<int>(dst, v.data(), v.size());
}
}  // namespace rt

/* reverse index tables of the adjoint gathers, from the forward tables */
int rt::build_adjoint_tables(void)
{
    if (M->adj_ready) return 0;
    const DevGrid &G = M->G;
    const size_t NSH = (size_t)G.nsh, NVH = (size_t)G.nvh, NT = (size_t)G.nt;
    std::vector<int> adj(6*NSH), from(NVH), to(NVH), vit(3*NT), dri(4*NVH), ti(10*NVH), tmc(10*NVH);
    std::vector<double> drw(4*NVH), tw(10*NVH);
    struct { int *h; const int *d; size_t n; } cp[] = {
        {adj.data(), G.adj, 6*NSH}, {from.data(), G.from, NVH}, {to.data(), G.to, NVH},
        {vit.data(), G.vit, 3*NT}, {dri.data(), G.dri, 4*NVH}, {ti.data(), G.ti, 10*NVH}, {tmc.data(), G.tmc, 10*NVH}};
    for (auto &c : cp) HIP_OK(hipMemcpyAsync(c.h, c.d, c.n*sizeof(int), hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipMemcpyAsync(drw.data(), G.drw, 4*NVH*sizeof(double), hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipMemcpyAsync(tw.data(), G.tw, 10*NVH*sizeof(double), hipMemcpyDeviceToHost, M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    AdjTablesHost T;
    if (build_adj_tables_host(G.nsh, G.nvh, G.nt, GAME_N_PENT, adj.data(), from.data(), to.data(), vit.data(),
                              dri.data(), drw.data(), ti.data(), tmc.data(), tw.data(), ADJ_RH_MAX, ADJ_TR_MAX,
                              PVR_MAX, T, g_err, sizeof g_err))
        return GAME_ERR_ARG;
    AdjTables &A = M->A;
    int rc;
    if ((rc = upload_ints(&A.rhp_e, T.rhp_e)) || (rc = upload_ints(&A.trp_e, T.trp_e)) || (rc = upload_ints(&A.trp_m, T.trp_m))
        || (rc = upload(&A.rhp_w, T.rhp_w.data(), T.rhp_w.size())) || (rc = upload(&A.trp_w, T.trp_w.data(), T.trp_w.size())))
        return rc;
    if ((rc = upload_ints(&A.slot_from, T.slot_from)) || (rc = upload_ints(&A.slot_to, T.slot_to))
        || (rc = upload_ints(&A.tri_of_edge, T.tri_of_edge)) || (rc = upload_ints(&A.tri_slot, T.tri_slot))
        || (rc = upload_ints(&A.rh_off, T.rh_off)) || (rc = upload_ints(&A.rh_ent, T.rh_ent))
        || (rc = upload_ints(&A.tr_off, T.tr_off)) || (rc = upload_ints(&A.tr_ent, T.tr_ent)))
        return rc;
    PvReverse &R = M->PR;
    if ((rc = upload_ints(&R.e, T.pvr_e)) || (rc = upload_ints(&R.f, T.pvr_f)) || (rc = upload_ints(&R.n, T.pvr_n))
        || (rc = upload(&R.w, T.pvr_w.data(), T.pvr_w.size())))
        return rc;
    R.n_max = 0;
    for (int c : T.pvr_n) R.n_max = c > R.n_max ? c : R.n_max;
    HIP_OK(hipStreamSynchronize(M->stream));
    M->adj_ready = 1;
    return 0;
}

static int op_done(void)
{
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

/* every buffer argument of an operator entry point must be given: a null one
   is an argument error, not a launch on address zero */
static bool all_given(std::initializer_list<const void *> ps)
{
    for (const void *p : ps)
        if (!p) return false;
    return true;
}
#define NEED(...)                                                                \
    if (!all_given({__VA_ARGS__})) {                                             \
        snprintf(g_err, sizeof g_err, "%s: a buffer argument is null", __func__); \
        return GAME_ERR_ARG;                                                     \
    }

#define OP_PROLOGUE(NEED_TABLES)                                               \
    if (!M) return GAME_ERR_NOT_INIT;                                          \
    if (M->precision != GAME_FP64) {                                           \
        snprintf(g_err, sizeof g_err, "operator interface needs the FP64 form"); \
        return GAME_ERR_ARG;                                                   \
    }                                                                          \
    if (NEED_TABLES) { int rc_ = build_adjoint_tables(); if (rc_) return rc_; }

extern "C" {

int game_gpu_op_grad(const double *f, double *out)
{
    OP_PROLOGUE(0)
    NEED(f, out)
    HIP_OK(hipMemsetAsync(out, 0, (size_t)M->d.v*sizeof(double), M->stream));
    hor::grad(M->G, f, out, M->stream);
    return op_done();
}

int game_gpu_op_grad_adjoint(const double *gbar, double *fbar)
{
    OP_PROLOGUE(1)
    NEED(gbar, fbar)
    adj::grad_adjoint(M->G, M->A, gbar, fbar, M->stream);
    return op_done();
}

int game_gpu_op_divv_h(const double *in, double *out)
{
    OP_PROLOGUE(0)
    NEED(in, out)
    adj::divv_h(M->G, in, out, M->stream);
    return op_done();
}

int game_gpu_op_divv_h_adjoint(const double *dbar, double *inbar)
{
    OP_PROLOGUE(1)
    NEED(dbar, inbar)
    adj::divv_h_adjoint(M->G, M->A, dbar, inbar, M->stream);
    return op_done();
}

int game_gpu_op_curl_triangles(const double *wind, double *out)
{
    OP_PROLOGUE(0)
    NEED(wind, out)
    hor::curl_triangles(M->G, wind, out, M->stream);
    return op_done();
}

int game_gpu_op_curl_triangles_adjoint(const double *zbar, double *windbar)
{
    OP_PROLOGUE(1)
    NEED(zbar, windbar)
    adj::curl_triangles_adjoint(M->G, M->A, zbar, windbar, M->stream);
    return op_done();
}

int game_gpu_op_rhombus_density(const double *rho, double *out)
{
    OP_PROLOGUE(0)
    NEED(rho, out)
    adj::rhombus_density(M->G, rho, out, M->stream);
    return op_done();
}

int game_gpu_op_rhombus_density_adjoint(const double *dbar, double *rhobar)
{
    OP_PROLOGUE(1)
    NEED(dbar, rhobar)
    adj::rhombus_density_adjoint(M->G, M->A, dbar, rhobar, M->stream);
    return op_done();
}

int game_gpu_op_vorticity_flux(const double *flux, const double *pot_vort, double *out)
{
    OP_PROLOGUE(0)
    NEED(flux, pot_vort, out)
    hor::vorticity_flux(M->G, flux, pot_vort, out, M->stream);
    return op_done();
}

int game_gpu_op_vorticity_flux_adjoint(const double *pot_vort, const double *outbar, double *fluxbar)
{
    OP_PROLOGUE(1)
    NEED(pot_vort, outbar, fluxbar)
    adj::vorticity_flux_adjoint(M->G, M->A, pot_vort, outbar, fluxbar, M->stream);
    return op_done();
}

int game_gpu_op_inner_product(const double *a, const double *b, double *out)
{
    OP_PROLOGUE(0)
    NEED(a, b, out)
    hor::inner_product(M->G, a, b, out, M->stream);
    return op_done();
}

int game_gpu_op_inner_product_adjoint(const double *b, const double *outbar, double *abar)
{
    OP_PROLOGUE(1)
    NEED(b, outbar, abar)
    adj::inner_product_adjoint(M->G, M->A, b, outbar, abar, M->stream);
    return op_done();
}

int game_gpu_op_thomas(int n, int ncol, const double *c, const double *d, const double *e, const double *r,
                       double *x, double *work)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (n < 1 || ncol < 1) { snprintf(g_err, sizeof g_err, "bad system size"); return GAME_ERR_ARG; }
    NEED(c, d, e, r, x, work)
    adj::thomas(n, ncol, c, d, e, r, x, work, M->stream);
    return op_done();
}

int game_gpu_op_thomas_noalloc(int n, int ncol, const double *c, const double *d, const double *e, const double *r,
                               double *x)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (n < 1 || ncol < 1) { snprintf(g_err, sizeof g_err, "bad system size"); return GAME_ERR_ARG; }
    NEED(c, d, e, r, x)
    if ((long long)n*(long long)ncol > THOMAS_SCRATCH_DOUBLES) {
        snprintf(g_err, sizeof g_err, "game_gpu_op_thomas_noalloc: %d systems of %d rows need %lld doubles of "
                 "scratch, more than its %lld", ncol, n, (long long)n*(long long)ncol, (long long)THOMAS_SCRATCH_DOUBLES);
        return GAME_ERR_ARG;
    }
    /* resolved at init on the model's device; resolved here for another current device */
    if (adj::thomas_scratch_resolve(g_err, sizeof g_err)) return GAME_ERR_HIP;
    adj::thomas(n, ncol, c, d, e, r, x, adj::thomas_scratch(), M->stream);
    return op_done();
}

int game_gpu_op_thomas_adjoint(int n, int ncol, const double *c, const double *d, const double *e, const double *x,
                               const double *xbar, double *rbar, double *cbar, double *dbar, double *ebar,
                               double *work)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (n < 1 || ncol < 1) { snprintf(g_err, sizeof g_err, "bad system size"); return GAME_ERR_ARG; }
    NEED(c, d, e, x, xbar, rbar, cbar, dbar, ebar, work)
    adj::thomas_adjoint(n, ncol, c, d, e, x, xbar, rbar, cbar, dbar, ebar, work, M->stream);
    return op_done();
}

}  /* extern "C" */
