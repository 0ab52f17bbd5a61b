/*
 * rkhevi_dev.h -- device-side data of the RKHEVI step (shared by the HIP files).
 *
 * Fields keep the reference's element order (see include/game_rkhevi.h), so
 * one thread per (layer, horizontal index) with the horizontal index fastest
 * reads and writes whole cache lines; neighbour gathers go through the
 * reference's own index tables.
 */
#ifndef GAME_RKHEVI_DEV_H
#define GAME_RKHEVI_DEV_H

#include <hip/hip_runtime.h>

#define GAME_N_PENT 12
#define GAME_MAX_LAYERS 64

/* src/game_constants.h */
#define GC_T_0 273.15
#define GC_N_A (6.02214076e23)
#define GC_P_0 100000.0
#define GC_C_D_V 717.942189
#define GC_C_V_V 1396.475121
#define GC_C_D_P 1005.0
#define GC_M_D (GC_N_A*0.004810e-23)
#define GC_M_V (GC_N_A*0.002991e-23)
#define GC_R_D 287.057811
#define GC_R_V 461.524879

typedef long long i64;

/* grid tables and metrics (struct grid / dualgrid members) on the device */
struct DevGrid {
    int nsh, nvh, nt, vpl, dvpl, L, n_oro, n_const, n_cond, moist;
    int flat;   /* slope == 0 everywhere: every terrain-correction term is an exact zero */
    i64 s, v, curl;
    double radius, toa;
    const double *nd, *vol, *area, *zs, *zv, *phi, *gm, *slope, *thbg, *exbg, *exbg_grad, *ipw;
    /* the inner-product weights in layer-major, slot-major order, so that
       lanes of consecutive cells read consecutive addresses:
       ipwT[(l*8 + j)*nsh + h] = ipw[8*(l*nsh + h) + j] */
    const double *ipwT;
    const double *tw, *drw, *dnd, *dzv, *darea, *fvec;
    const int *from, *to, *adj, *sgn, *ti, *tmc, *dri, *dfrom, *dto, *vit, *vst;
    /* per-edge differences that calc_pot_vort forms for its density
       extrapolation to the top and the surface, read exactly as the
       reference's struct layout serves them (vorticities.c:48-71) */
    const double *pv_top_dz1, *pv_top_dz2, *pv_sfc_dz1, *pv_sfc_dz2;
    /* Klemp damping coefficient of every interior interface, (L-1)*NSH */
    const double *damp;
    /* (0.5 dt)/vol of every grid box for the dt of the last step (the tracer
       solver's implicit weight times dt over the volume); the runtime refills
       it when dt changes */
    double *idv;
    /* number of layer chunks the per-cell kernels split a column into
       (gridDim.y); more chunks give more waves to hide gather latency */
    int split_cell;
    /* the same for the per-edge kernels */
    int split_edge;
    /* 1: tracer solver with one (column, tracer) per lane in registers; 0: LDS-staged */
    int gd_streaming;
    /* one tracer per block in the LDS-staged tracer solve instead of all five */
    int gd_per_tracer;
    /* mixed form, a precision probe: bit 0 rounds the density tendencies,
       bit 1 the rhotheta_v tendency, to FP32 after the scalar tendencies
       (the precision of FP32 storage, not its bytes) */
    int round_tendencies;
    /* 1: sound-wave solver with one column per lane; 0: three phases over LDS */
    int vw_streaming;
    /* 1 (26 layers): the row-parallel three-phase sound-wave kernel instead of the layer-run one */
    int vw_rows;
    /* 1: sediment heat with one thread per grid box; 0: one thread per column (26 layers) */
    int sh_per_box;
    /* 1: temperature and phase transitions as two kernels instead of one */
    int unfused_phase;
};

struct DevState {
    double *rho, *rt, *tp, *ep, *wind;   /* rho, rhotheta_v, theta_v_pert, exner_pert, wind */
};

/* work arrays of the step; TW = double in the FP64 form, float in the mixed
   form (register arithmetic stays FP64 in both) */
template <typename TW>
struct DevWorkT {
    /* diagnostics */
    TW *flux, *rvt, *pv, *temp, *vsq;
    double *flux2, *flux_divv;   /* reference diagnostics the step leaves untouched */
    /* forcings */
    TW *pgrad_old, *pgrad_nl, *pgrad_l, *pgc_v, *vsq_grad, *pvt;
    /* the pressure-gradient part of the momentum tendency, formed with the
       gradient: ohw*old - chw*(nl + l) at edges, -cvw*(nl + l) at interior w
       points (vector_tendencies_expl.c's weights for the step's implicit weight) */
    TW *pgrad_comb;
    /* irreversible quantities */
    TW *ptr, *phr, *sed_heat;
    /* vertical contravariant corrections of wind, gas mass flux and rhotheta_v
       flux in the terrain-following levels, 3 x n_oro x NSH (unused when flat) */
    double *vcc;
    /* wind tendency (the FP64 form keeps it in the tendency state) */
    TW *td_wind;
    /* c_p of every condensate at the step's temperature, n_cond x S: the
       temperature is fixed for the whole step, so each value is formed once */
    TW *cpc;
    /* gas density over total density of the step's old state, S (formed with
       the temperature; read by the pressure gradient at the first substep) */
    double *gas_frac;
};
typedef DevWorkT<double> DevWork;
typedef DevWorkT<float> DevWorkF;

struct StepConsts {
    double dt, impl_weight, damping_start, damping_max, snow_v, rain_v, cloud_v;
};

__device__ __forceinline__ double density_total(const DevGrid &G, const double *rho, i64 i)
{   /* derived_quantities.c:83-95: left-to-right sum from 0 */
    double r = 0.0;
    for (int k = 0; k < G.n_const; ++k) r += rho[k*G.s + i];
    return r;
}

/* the same sum with the constituent count known at compile time (NC > 0):
   all loads issue before the first add */
template <int NC>
__device__ __forceinline__ double density_total_n(const DevGrid &G, const double *rho, i64 i)
{
    if (NC <= 0) return density_total(G, rho, i);
    double v[NC > 0 ? NC : 1];
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] = rho[k*G.s + i];
    double r = 0.0;
#pragma unroll
    for (int k = 0; k < NC; ++k) r += v[k];
    return r;
}

/* dictionary.c:111-190 */
__device__ __forceinline__ double c_p_ice(double t)
{
    if (t > GC_T_0) t = GC_T_0;
    if (t < 20.0) t = 20.0;
    double q = t/125.1;
    double r = -2.0572 + 0.14644*t + 0.06163*t*exp(-(q*q));
    return r/GC_M_V;
}

__device__ __forceinline__ double c_p_water(double t)
{
    double tc = t - GC_T_0, r;
    /* a branch, not a select: a wave whose boxes are all above or all below
       freezing evaluates one of the two polynomials */
    /* the reference's integer powers pow(x, 3.0) and pow(x, 4.0) as products
       (within about one ulp of glibc's pow, as the device's pow is) */
    if (tc >= 0.0) {
        if (tc > 35.0) tc = 35.0;
        double a = tc - 35.0;
        const double a2 = a*a;
        r = 0.9979 + 3.1e-6*(a*a) + 3.8e-9*(a2*a2);
    } else {
        if (tc < -37.0) tc = -37.0;
        const double t2 = tc*tc;
        r = 1.000938 - 2.7052e-3*tc - 2.3235e-5*(tc*tc) + 4.3778e-6*(t2*tc) + 2.7136e-7*(t2*t2);
    }
    return 4186.8*r;
}

/* even k: ice phases (snow, cloud ice), odd k: liquid (rain, cloud water) */
__device__ __forceinline__ double c_p_cond(int k, double t)
{
    return (k & 1) == 0 ? c_p_ice(t) : c_p_water(t);
}

/* c_v_mass_weighted_air with the condensates' c_p taken from a table formed
   at the step's temperature (same terms, same order) */
template <typename TW>
__device__ __forceinline__ double c_v_mass_weighted_cp(const DevGrid &G, const double *rho, const TW *cpc, i64 i)
{
    double r = 0.0;
    for (int k = 0; k < G.n_cond; ++k) r += rho[k*G.s + i]*(double)cpc[k*G.s + i];
    r += rho[(i64)G.n_cond*G.s + i]*GC_C_D_V;
    if (G.moist == 1) r += rho[(i64)(G.n_cond + 1)*G.s + i]*GC_C_V_V;
    return r;
}

/* derived_quantities.c:97-115 */
__device__ __forceinline__ double c_v_mass_weighted(const DevGrid &G, const double *rho, double T, i64 i)
{
    double r = 0.0;
    for (int k = 0; k < G.n_cond; ++k) r += rho[k*G.s + i]*c_p_cond(k, T);
    r += rho[(i64)G.n_cond*G.s + i]*GC_C_D_V;
    if (G.moist == 1) r += rho[(i64)(G.n_cond + 1)*G.s + i]*GC_C_V_V;
    return r;
}

/* averaging.c:41-95 */
__device__ __forceinline__ double vertical_contravariant_corr(const DevGrid &G, const double *f, int l, int h)
{
    double r = 0;
    int ne = h < GAME_N_PENT ? 5 : 6;
    if (l >= G.L - G.n_oro) {
        if (l == G.L - G.n_oro) {
            for (int i = 0; i < ne; ++i) {
                i64 si = (i64)l*G.nsh + h, vi = G.nsh + (i64)l*G.vpl + G.adj[6*h + i];
                r += -0.5*G.ipw[8*si + i]*G.slope[vi]*f[vi];
            }
        } else {
            for (int i = 0; i < ne; ++i) {
                i64 si = (i64)(l - 1)*G.nsh + h, vi = G.nsh + (i64)(l - 1)*G.vpl + G.adj[6*h + i];
                r += -0.5*G.ipw[8*si + i]*G.slope[vi]*f[vi];
            }
            for (int i = 0; i < ne; ++i) {
                i64 si = (i64)l*G.nsh + h, vi = G.nsh + (i64)l*G.vpl + G.adj[6*h + i];
                r += -0.5*G.ipw[8*si + i]*G.slope[vi]*f[vi];
            }
        }
    }
    return r;
}

/* divergences.c:14-64 for cell h of layer l */
__device__ __forceinline__ double divv_h_cell(const DevGrid &G, const double *in, int l, int h)
{
    int ne = h < GAME_N_PENT ? 5 : 6;
    i64 i = (i64)l*G.nsh + h;
    double comp_h = 0.0;
    for (int j = 0; j < ne; ++j) {
        i64 vi = G.nsh + (i64)l*G.vpl + G.adj[6*h + j];
        comp_h += in[vi]*G.sgn[6*h + j]*G.area[vi];
    }
    double comp_v = 0.0;
    int L = G.L, no = G.n_oro;
    if (l == L - no - 1) {
        double cl = vertical_contravariant_corr(G, in, l + 1, h);
        comp_v = -cl*G.area[h + (i64)(l + 1)*G.vpl];
    } else if (l == L - 1) {
        double cu = vertical_contravariant_corr(G, in, l, h);
        comp_v = cu*G.area[h + (i64)l*G.vpl];
    } else if (l > L - no - 1) {
        double cu = vertical_contravariant_corr(G, in, l, h);
        double cl = vertical_contravariant_corr(G, in, l + 1, h);
        comp_v = cu*G.area[h + (i64)l*G.vpl] - cl*G.area[h + (i64)(l + 1)*G.vpl];
    }
    return 1.0/G.vol[i]*(comp_h + comp_v);
}

/* per-kernel timing with HIP events on the launch stream (off unless enabled) */
struct KernelTimer {
    virtual void begin(const char *name, hipStream_t st) = 0;
    virtual void end(hipStream_t st) = 0;
};
extern KernelTimer *g_kernel_timer;

#define LAUNCH_SHM(NAME, KERNEL, GRID, BLOCK, SHM, STREAM, ...)                     \
    do {                                                                            \
        if (g_kernel_timer) g_kernel_timer->begin(NAME, STREAM);                    \
        hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, STREAM, __VA_ARGS__);          \
        if (g_kernel_timer) g_kernel_timer->end(STREAM);                            \
    } while (0)
#define LAUNCH(NAME, KERNEL, GRID, BLOCK, STREAM, ...) LAUNCH_SHM(NAME, KERNEL, GRID, BLOCK, 0, STREAM, __VA_ARGS__)

#endif
