/*
 * init_state.c -- analytic balanced initial states for the RKHEVI step.
 *
 * The reference builds idealised states in set_ideal_init
 * (src/io/set_initial_state.c:23-242): analytic T, p and q per cell, the wind
 * projected on the edge normals, then a bottom-up discrete hydrostatic
 * balance that solves a quadratic for the Exner pressure in every layer and
 * includes the kinetic-energy and vorticity-flux terms of the initial wind.
 * This module restates that procedure for the workloads the step is measured
 * on (a lapse-rate atmosphere with a jet and seeded noise).  The balance needs
 * v^2 and the vertical vorticity-flux of the initial wind; they are computed
 * by the caller between game_init_phase1 and game_init_phase2 with the step's
 * own operators (the device kernels in production use).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "init_state.h"

#define T_0_REF 273.15
#define P_0_REF 100000.0
#define C_D_P_REF 1005.0
#define R_D_REF 287.057811
#define R_V_REF 461.524879
#define N_A_REF (6.02214076e23)
#define M_D_REF (N_A_REF*0.004810e-23)
#define M_V_REF (N_A_REF*0.002991e-23)
#define G_MEAN 9.80616

/* counter-based seeded noise keyed by (seed, cell, layer): splitmix64 + Box-Muller */
static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30))*0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27))*0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static uint64_t key3(uint64_t seed, uint64_t cell, uint64_t layer)
{
    return splitmix64(splitmix64(splitmix64(seed) ^ cell) ^ (layer + 0x632BE59BD9B4E019ULL));
}

double game_uniform01(uint64_t seed, long long cell, int layer, int stream)
{
    uint64_t k = splitmix64(key3(seed, (uint64_t)cell, (uint64_t)layer) + (uint64_t)stream);
    return ((k >> 11) + 0.5)*(1.0/9007199254740992.0);   /* (0, 1) */
}

double game_gauss(uint64_t seed, long long cell, int layer)
{
    double u1 = game_uniform01(seed, cell, layer, 1), u2 = game_uniform01(seed, cell, layer, 2);
    return sqrt(-2.0*log(u1))*cos(2.0*M_PI*u2);
}

/* analytic profile: T = T_sfc + gamma*z below z_trop, isothermal above */
static double profile_temp(const game_init_params_t *P, double z)
{
    double zz = z < P->z_tropopause ? z : P->z_tropopause;
    return P->t_sfc + P->lapse_rate*zz;
}

static double profile_pres(const game_init_params_t *P, double z)
{
    double expo = -G_MEAN/(R_D_REF*P->lapse_rate);
    if (z < P->z_tropopause) return P->p_sfc*pow(1 + P->lapse_rate*z/P->t_sfc, expo);
    double p_t = P->p_sfc*pow(1 + P->lapse_rate*P->z_tropopause/P->t_sfc, expo);
    double t_t = P->t_sfc + P->lapse_rate*P->z_tropopause;
    return p_t*exp(-G_MEAN*(z - P->z_tropopause)/(R_D_REF*t_t));
}

/* saturation pressure in Pa (src/constituents/dictionary.c:287-333, 361-385),
   with the enhancement factors of dictionary.c:414-436 */
static double e_sat(double T, double p)
{
    double tc = T - T_0_REF;
    double es;
    if (T >= T_0_REF) {
        if (tc > 100.0) tc = 100.0;
        es = exp(34.494 - 4924.99/(tc + 237.1))/pow(tc + 105.0, 1.57);
        return 1.00071*exp(0.000000045*p)*es;
    }
    if (tc < -80.0) tc = -80.0;
    es = exp(43.494 - 6545.8/(tc + 278.0))/pow(tc + 868.0, 2);
    return 0.99882*exp(0.00000008*p)*es;
}

void game_init_default_params(game_init_params_t *P)
{
    memset(P, 0, sizeof *P);
    P->t_sfc = 288.0;
    P->lapse_rate = -0.0065;
    P->z_tropopause = 11000.0;
    P->p_sfc = 101325.0;
    P->u_max = 35.0;
    P->z_jet_top = 20000.0;
    P->sigma_t = 0.1;
    P->seed = 20261202ULL;
    P->moist = 0;
    P->rh_fraction = 0.8;
    P->q_noise = 0.05;
    P->z_moist_top = 8000.0;
}

int game_init_phase1(const game_init_params_t *P, const game_dims_t *d, const game_grid_view_t *g,
                     const double *latitude_vector, game_state_view_t *state, game_init_work_t *w)
{
    int nsh = d->nsh, vpl = d->vpl, L = d->n_layers;
    if (P->moist && !d->moisture_on) return 1;
    /* 3D scalar fields (set_initial_state.c:48-78) */
    #pragma omp parallel for
    for (long long i = 0; i < d->s; ++i) {
        int l = (int)(i/nsh), h = (int)(i - (long long)l*nsh);
        double z = g->z_scalar[i];
        double T = profile_temp(P, z) + P->sigma_t*game_gauss(P->seed, h, l);
        double p = profile_pres(P, z);
        double q = 0.0;
        if (P->moist && z < P->z_moist_top) {
            double es = e_sat(T, p);
            double eps = R_D_REF/R_V_REF;
            double q_sat = eps*es/(p - (1.0 - eps)*es);
            double u = game_uniform01(P->seed, h, l, 3);
            q = P->rh_fraction*q_sat*(1.0 + P->q_noise*(2.0*u - 1.0));
        }
        w->temperature[i] = T;
        w->pressure[i] = p;
        w->specific_humidity[i] = q;
        w->temperature_v[i] = T*(1.0 + q*(M_D_REF/M_V_REF - 1.0));
        /* moist-air density before the balance (set_initial_state.c:155-158) */
        w->density[i] = p/(R_D_REF*w->temperature_v[i]);
    }
    /* horizontal wind projected on the edge normals (set_initial_state.c:109-141) */
    #pragma omp parallel for
    for (int l = 0; l < L; ++l)
        for (int e = 0; e < d->nvh; ++e) {
            long long vi = nsh + (long long)l*vpl + e;
            double z = g->z_vector[vi];
            double lat = latitude_vector[e];
            double zz = z < P->z_jet_top ? z : P->z_jet_top;
            double u = P->u_max*pow(cos(lat), 2)*sin(M_PI*zz/P->z_jet_top);
            double v = 0.0;
            state->wind[vi] = u*cos(g->direction[e]) + v*sin(g->direction[e]);
        }
    /* vertical wind = 0 (set_initial_state.c:143-151) */
    for (int l = 0; l < d->n_levels; ++l)
        for (int h = 0; h < nsh; ++h) state->wind[(long long)l*vpl + h] = 0.0;
    return 0;
}

int game_init_phase2(const game_init_params_t *P, const game_dims_t *d, const game_grid_view_t *g,
                     const double *v_squared, const double *pot_vort_tend,
                     game_state_view_t *state, game_init_work_t *w)
{
    int nsh = d->nsh, vpl = d->vpl, L = d->n_layers;
    (void)P;
    /* bottom-up discrete hydrostatic balance (set_initial_state.c:167-206) */
    #pragma omp parallel for
    for (int h = 0; h < nsh; ++h)
        for (int l = L - 1; l >= 0; --l) {
            long long si = (long long)l*nsh + h;
            double *ex = state->exner_pert;   /* full Exner pressure until the background is removed */
            if (l == L - 1) {
                ex[si] = pow(w->pressure[si]/P_0_REF, R_D_REF/C_D_P_REF);
            } else {
                double b = -0.5*ex[si + nsh]/w->temperature_v[si + nsh]
                           *(w->temperature_v[si] - w->temperature_v[si + nsh]
                             + 2.0/C_D_P_REF*(g->gravity_potential[si] - g->gravity_potential[si + nsh]
                                              + 0.5*v_squared[si] - 0.5*v_squared[si + nsh]
                                              - (g->z_scalar[si] - g->z_scalar[si + nsh])*pot_vort_tend[h + (long long)(l + 1)*vpl]));
                double c = pow(ex[si + nsh], 2)*w->temperature_v[si]/w->temperature_v[si + nsh];
                ex[si] = b + pow((pow(b, 2) + c), 0.5);
            }
            state->theta_v_pert[si] = w->temperature_v[si]/ex[si];
            w->density[si] = P_0_REF*pow(ex[si], C_D_P_REF/R_D_REF)/(R_D_REF*w->temperature_v[si]);
            state->rhotheta_v[si] = w->density[si]*state->theta_v_pert[si];
        }
    /* remove the background state (set_initial_state.c:211-216) */
    #pragma omp parallel for
    for (long long i = 0; i < d->s; ++i) {
        state->exner_pert[i] = state->exner_pert[i] - g->exner_bg[i];
        state->theta_v_pert[i] = state->theta_v_pert[i] - g->theta_v_bg[i];
    }
    /* densities (set_initial_state.c:218-230): condensates 0, gas, vapour = q*rho_gas */
    long long S = d->s;
    #pragma omp parallel for
    for (long long i = 0; i < S; ++i) {
        for (int k = 0; k < d->n_cond; ++k) state->rho[k*S + i] = 0.0;
        state->rho[(long long)d->n_cond*S + i] = w->density[i];
        if (d->moisture_on) state->rho[(long long)(d->n_cond + 1)*S + i] = w->specific_humidity[i]*w->density[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* restart: the state from densities and temperature                     */
/* ------------------------------------------------------------------ */
/* saturation pressure in Pa over plane water, Huang (2018) Eq. (17) above
   freezing, Pruppacher and Klett (2010) Eq. (A.4-1) for super-cooled water
   (src/constituents/dictionary.c:287-333) */
static double sat_pressure_water(double temperature)
{
    double tc = temperature - T_0_REF;
    if (tc > 100.0) tc = 100.0;
    if (tc > 0.0) return exp(34.494 - 4924.99/(tc + 237.1))/pow(tc + 105.0, 1.57);
    if (tc < -50.0) tc = -50.0;
    double r = 6.107799961 + 4.436518521e-1*tc + 1.428945805e-2*pow(tc, 2) + 2.650648471e-4*pow(tc, 3)
               + 3.031240396e-6*pow(tc, 4) + 2.034080948e-8*pow(tc, 5) + 6.136820929e-11*pow(tc, 6);
    return 100.0*r;
}

/* over plane ice, Huang (2018) Eq. (18) (dictionary.c:361-385) */
static double sat_pressure_ice(double temperature)
{
    double tc = temperature - T_0_REF;
    if (tc < -80.0) tc = -80.0;
    if (tc > 0.0) tc = 0.0;
    return exp(43.494 - 6545.8/(tc + 278.0))/pow(tc + 868.0, 2);
}

double game_rel_humidity(double abs_humidity, double temperature)
{
    double vapour_pressure = abs_humidity*R_V_REF*temperature;
    double saturation = temperature > T_0_REF ? sat_pressure_water(temperature) : sat_pressure_ice(temperature);
    return vapour_pressure/saturation;
}

int game_state_from_temperature(const game_dims_t *d, const game_grid_view_t *g, double *rho,
                                const double *temperature, game_state_view_t *st)
{
    const long long S = (long long)d->nsh*d->n_layers;
    const int n_cond = 4*d->moisture_on, n_const = 1 + d->moisture_on + n_cond;
    double *gas = rho + (long long)n_cond*S, *vap = d->moisture_on ? rho + (long long)(n_cond + 1)*S : NULL;
    if (vap) {
        #pragma omp parallel for
        for (long long i = 0; i < S; ++i) {
            double rh = game_rel_humidity(vap[i], temperature[i]);
            if (rh > 1.0) vap[i] = vap[i]/game_rel_humidity(vap[i], temperature[i]);
        }
    }
    /* dry, the reference reads the member behind rho there (rhotheta_v of the
       state being filled, zero), so the virtual temperature is the temperature */
    #pragma omp parallel for
    for (long long i = 0; i < S; ++i) {
        double tv = vap ? temperature[i]*(1.0 + vap[i]/gas[i]*(M_D_REF/M_V_REF - 1.0)) : temperature[i];
        double pressure = gas[i]*R_D_REF*tv;
        double pot_temp_v = tv*pow(P_0_REF/pressure, R_D_REF/C_D_P_REF);
        st->rhotheta_v[i] = gas[i]*pot_temp_v;
        st->theta_v_pert[i] = pot_temp_v - g->theta_v_bg[i];
        st->exner_pert[i] = tv/(g->theta_v_bg[i] + st->theta_v_pert[i]) - g->exner_bg[i];
    }
    for (long long i = 0; i < (long long)n_const*S; ++i)
        if (rho[i] < 0) return 1;
    return 0;
}
