/*
 * rkhevi_kernels.h -- launchers of the device kernels (one namespace per
 * precision form of the step).
 */
#ifndef GAME_RKHEVI_KERNELS_H
#define GAME_RKHEVI_KERNELS_H

#include "rkhevi_dev.h"

/* horizontal operators and tendencies: work arrays double (FP64 form) or
   float (mixed form), register arithmetic FP64 in both */
namespace hor {
/* the gas fractions come from W.gas_frac: s must be the state the last
   col::temperature(..., W.gas_frac, ...) was formed from */
/* itw: the implicit weight of the momentum tendency (W.pgrad_comb) */
void pgrad(const DevGrid &G, const DevState &s, const DevWork &W, int first, double itw, hipStream_t st);
void pgrad(const DevGrid &G, const DevState &s, const DevWorkF &W, int first, double itw, hipStream_t st);
void pgc_v(const DevGrid &G, const DevState &s, const DevWork &W, hipStream_t st);
void pgc_v(const DevGrid &G, const DevState &s, const DevWorkF &W, hipStream_t st);
void advection(const DevGrid &G, const double *wind, const double *rho_gas, const DevWork &W, hipStream_t st);
void advection(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkF &W, hipStream_t st);
/* the two independent parts of advection: mass flux, PV and the TRSK flux
   (pot_vort_tend); kinetic energy and its gradient (v_squared_grad) */
void advection_pv(const DevGrid &G, const double *wind, const double *rho_gas, const DevWork &W, hipStream_t st);
void advection_pv(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkF &W, hipStream_t st);
void advection_kinetic(const DevGrid &G, const double *wind, const DevWork &W, hipStream_t st);
void advection_kinetic(const DevGrid &G, const double *wind, const DevWorkF &W, hipStream_t st);
void vector_tend(const DevGrid &G, const DevWork &W, double *tend_w, const double *w_old, double *w_new,
                 double new_w, double itw, double dt, hipStream_t st);
void vector_tend(const DevGrid &G, const DevWorkF &W, float *tend_w, const double *w_old, double *w_new,
                 double new_w, double itw, double dt, hipStream_t st);
void scalar_tend(const DevGrid &G, const DevState &s, const DevState &td, const DevWork &W, int rk, hipStream_t st);
void scalar_tend(const DevGrid &G, const DevState &s, const DevState &td, const DevWorkF &W, int rk, hipStream_t st);
void inner_product(const DevGrid &G, const double *a, const double *b, double *out, hipStream_t st);
void grad(const DevGrid &G, const double *f, double *out, hipStream_t st);
void curl_triangles(const DevGrid &G, const double *wind, double *out, hipStream_t st);
void vorticity_flux(const DevGrid &G, const double *F, const double *pv, double *out, hipStream_t st);
}

/* column solvers, point physics and state arithmetic shared by both forms:
   FP64 arithmetic; the work arrays are double (FP64 form) or float (mixed) */
namespace col {
/* cpc (may be null): c_p of every condensate at the temperature, n_cond x S;
   gf (may be null): gas density over total density, S */
void temperature(const DevGrid &G, const DevState &s, double *T, double *cpc, double *gf, hipStream_t st);
void temperature(const DevGrid &G, const DevState &s, float *T, float *cpc, double *gf, hipStream_t st);
void phase_trans(const DevGrid &G, const DevState &s, const double *T, const double *cpc, double *ptr, double *phr,
                 double dt, hipStream_t st);
void phase_trans(const DevGrid &G, const DevState &s, const float *T, const float *cpc, float *ptr, float *phr,
                 double dt, hipStream_t st);
/* moist (6 constituents, 4 condensates): temperature, c_p, gas fraction and
   phase transitions in one pass, the results of temperature + phase_trans */
void temperature_phase_trans(const DevGrid &G, const DevState &s, double *T, double *cpc, double *gf, double *ptr,
                             double *phr, double dt, hipStream_t st);
void temperature_phase_trans(const DevGrid &G, const DevState &s, float *T, float *cpc, double *gf, float *ptr,
                             float *phr, double dt, hipStream_t st);
/* snr: the new state as read before this substep writes it (at rk 0 the old
   state, which equals it by the step's entry condition) */
void ver_waves(const DevGrid &G, const DevState &so, const DevState &snr, const DevState &sn, const DevState &td,
               const StepConsts &K, int rk, hipStream_t st);
void ver_waves(const DevGrid &G, const DevState &so, const DevState &snr, const DevState &sn, const DevState &td,
               const float *td_wind, const StepConsts &K, int rk, hipStream_t st);
void gen_densities(const DevGrid &G, const DevState &so, const DevState &sn, const DevState &td, const double *T,
                   const double *cpc, double *sed_heat, const StepConsts &K, int rk, hipStream_t st);
void gen_densities(const DevGrid &G, const DevState &so, const DevState &sn, const DevState &td, const float *T,
                   const float *cpc, float *sed_heat, const StepConsts &K, int rk, hipStream_t st);
void linear_combine(const DevGrid &G, const DevState &a, const DevState &b, const DevState &o, double c0, double c1,
                    hipStream_t st);
/* G.idv[b] = (0.5 dt)/vol[b] */
void tracer_quotients(const DevGrid &G, double dt, hipStream_t st);
/* per-block compensated (sum, error) pairs of sum_i x[i]*y[i]*vol[i] (y may be
   null: sum_i x[i]*vol[i]) into partials[0 .. 2*INTEGRAL_BLOCKS) */
const int INTEGRAL_BLOCKS = 1024;
int integral_partials(i64 n, const double *x, const double *y, const double *vol, double *partials, hipStream_t st);
int integral_partials(i64 n, const double *x, const float *y, const double *vol, double *partials, hipStream_t st);
}

#endif
