/*
 * rkhevi_window.hip -- forward and reverse sweeps of a window of RKHEVI steps
 * (dry, FP64) behind include/game_rkhevi.h: the gradient of a cost of the
 * final state with respect to the initial control variables (rho,
 * rhotheta_v, wind).  The reference has no adjoint.
 *
 * Checkpointing: the forward sweep stores the state and the carried forcings
 * (the previous step's pressure gradient nl + l and momentum advection) every
 * `checkpoint_every` steps.  The reverse sweep visits the segments last to
 * first; each segment is run forward once more from its checkpoint with every
 * step's inputs and substep results copied to a tape (StepTape), then its steps
 * are reversed last to first.  Memory: one checkpoint per segment plus one
 * segment of tapes.
 *
 * Reverse of one step (manage_rkhevi.c:20-100 backwards; see step_adjoint.h
 * for the functors):
 *   second substep: sound-wave solver, scalar transport, wind update,
 *     momentum tendency, momentum advection (TRSK flux, potential vorticity,
 *     kinetic energy) of the first substep's state;
 *   first substep: sound-wave solver, scalar transport, wind update,
 *     momentum tendency (the previous step's pressure gradient and advection,
 *     or at the first step the current ones);
 *   pressure gradient; theta_v_pert, which the driver's copy re-derives from
 *   rho and rhotheta_v (coordinator.c:305, linear_combine_two_states.c:26).
 */
#include <math.h>
#include <string.h>

#include "runtime_internal.h"

using namespace rt;

#define WTPB 256

template <class F>
__global__ void __launch_bounds__(WTPB) k_apply(F f, i64 n)
{
    const i64 t = (i64)blockIdx.x*WTPB + threadIdx.x;
    if (t < n) f(t);
}

template <class F>
static void launch(const char *name, const F &f, i64 n, hipStream_t st)
{
    if (n <= 0) return;
    LAUNCH(name, k_apply<F>, dim3((unsigned)((n + WTPB - 1)/WTPB)), dim3(WTPB), st, f, n);
}

namespace {

/* the state of the control variables: theta_v_pert = rt/rho - theta_bg and
   exner_pert = (R_d rt/p_0)^(R_d/c_v) - exner_bg (the equation of state of
   set_initial_state.c:302-314, written for rhotheta_v) */
struct EosState {
    DevGrid G;
    const double *rho, *rt;
    double *tp, *ep;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        tp[t] = rt[t]/rho[t] - G.thbg[t];
        ep[t] = pow(GC_R_D*rt[t]/GC_P_0, GC_R_D/GC_C_D_V) - G.exbg[t];
    }
};

/* its reverse: the exner_pert and theta_v_pert cotangents into rt and rho */
struct EosStateAdj {
    DevGrid G;
    const double *rho, *rt;
    const double *b_ep, *b_tp;
    double *b_rho, *b_rt;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        const double r = rho[t], q = rt[t];
        const double ex = pow(GC_R_D*q/GC_P_0, GC_R_D/GC_C_D_V);
        b_rt[t] += b_ep[t]*(GC_R_D/GC_C_D_V)*ex/q + b_tp[t]/r;
        b_rho[t] += -b_tp[t]*q/(r*r);
    }
};

/* edges and interior w of a wind cotangent: edges to ub (written), the top
   and surface w (passed through unchanged by the step) added to b_in */
struct PassThroughAdj {
    DevGrid G;
    const double *b_out;
    double *ub, *b_in;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.v) return;
        const int lev = (int)(t/G.vpl), h = (int)(t - (i64)lev*G.vpl);
        if (h >= G.nsh) ub[t] = b_out[t];
        else {
            ub[t] = 0.0;
            if (lev == 0 || lev == G.L) b_in[t] += b_out[t];
        }
    }
};

struct DiffToCot {
    i64 n;
    const double *x, *y;
    double *b;
    GAME_HD void operator()(i64 t) const { if (t < n) b[t] = x[t] - y[t]; }
};

struct Cot {
    double *rho, *rt, *tp, *ep, *wind;
};

struct Window {
    long long owner = 0;        /* the model generation the buffers belong to */
    std::vector<void *> mem;    /* every buffer of the window, for release */
    int k = 0, n_ckpt = 0;
    std::vector<DevState> ck_state;
    std::vector<double *> ck_nl, ck_l, ck_pvt, ck_vsqg;
    std::vector<StepTape> tape;
    DevState seg_end;
    Cot bS, bSn, bS1;
    double *bt0_rho, *bt0_rt, *bt0_wind, *bt1_rho, *bt1_rt, *bt1_wind;
    double *bP, *bPold, *bA, *bA1, *bAprev, *bU;
    double *Fb, *Fthb, *cvb, *vsqb, *tmpV1, *tmpV2, *fluxb, *pvb, *rvb, *dvb, *zb, *ws, *ones, *vcb;
    DevWork AW;
    /* phase clock of the last gradient: events and the phase each closes */
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_phase;
    int n_ev = 0;
    double phase_ms[4] = {0, 0, 0, 0};
    void mark(int phase, hipStream_t st)
    {
        if (n_ev == (int)ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            ev.push_back(e);
            ev_phase.push_back(0);
        }
        ev_phase[n_ev] = phase;
        (void)hipEventRecord(ev[n_ev++], st);
    }
    /* after the stream is synchronised: the time between consecutive marks,
       summed per phase */
    void read_clock()
    {
        for (double &v : phase_ms) v = 0.0;
        for (int i = 1; i < n_ev; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[i - 1], ev[i]) == hipSuccess) phase_ms[ev_phase[i]] += ms;
        }
        n_ev = 0;
    }
};
enum { PH_FORWARD, PH_RECOMPUTE, PH_REVERSE, PH_FINISH };

Window *Wn = nullptr;
Window *Wbuild = nullptr;   /* the window whose buffers are being allocated */

int zeros(double **p, size_t n)
{
    int rc = dzeros(p, n);
    if (!rc && Wbuild) Wbuild->mem.push_back(*p);
    return rc;
}

/* releases a window's buffers (those of an older model went with it) */
void window_free(Window *w)
{
    if (!w) return;
    if (M && w->owner == M->generation)
        for (void *p : w->mem) dfree(p);
    for (hipEvent_t e : w->ev) (void)hipEventDestroy(e);
    delete w;
}

int alloc_state(DevState &s)
{
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    int rc;
    if ((rc = zeros(&s.rho, S*(size_t)M->d.n_const)) || (rc = zeros(&s.rt, S)) || (rc = zeros(&s.tp, S))
        || (rc = zeros(&s.ep, S)) || (rc = zeros(&s.wind, V)))
        return rc;
    return 0;
}

int alloc_cot(Cot &c)
{
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    int rc;
    if ((rc = zeros(&c.rho, S)) || (rc = zeros(&c.rt, S)) || (rc = zeros(&c.tp, S)) || (rc = zeros(&c.ep, S))
        || (rc = zeros(&c.wind, V)))
        return rc;
    return 0;
}

int window_alloc(Window &w, int k, int n_ckpt);

/* buffers for windows checkpointed every k steps with up to n_ckpt checkpoints;
   they live with the model (freed by game_gpu_finalize) */
int window_buffers(int k, int n_ckpt)
{
    if (Wn && Wn->owner == M->generation && Wn->k >= k && Wn->n_ckpt >= n_ckpt) return 0;
    /* the buffers of a shorter window, or of another model, are replaced */
    const int kk = (Wn && Wn->owner == M->generation && Wn->k > k) ? Wn->k : k;
    const int nn = (Wn && Wn->owner == M->generation && Wn->n_ckpt > n_ckpt) ? Wn->n_ckpt : n_ckpt;
    window_free(Wn);
    Wn = nullptr;
    Window *nw = new Window();
    Wbuild = nw;
    int rc = window_alloc(*nw, kk, nn);
    Wbuild = nullptr;
    if (rc) { window_free(nw); return rc; }     /* a partly built window is never kept */
    Wn = nw;
    return 0;
}

/* allocates every buffer of w (the model keeps them until game_gpu_finalize) */
int window_alloc(Window &w, int k, int n_ckpt)
{
    w.owner = M->generation;
    w.k = k;
    w.n_ckpt = n_ckpt;
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v, L = (size_t)M->d.n_layers, NVH = (size_t)M->d.nvh,
                 NSH = (size_t)M->d.nsh, NT = (size_t)M->d.nt, CURL = (size_t)M->d.curl;
    int rc;
    w.ck_state.resize(n_ckpt);
    w.ck_nl.resize(n_ckpt); w.ck_l.resize(n_ckpt); w.ck_pvt.resize(n_ckpt); w.ck_vsqg.resize(n_ckpt);
    for (int c = 0; c < n_ckpt; ++c)
        if ((rc = alloc_state(w.ck_state[c])) || (rc = zeros(&w.ck_nl[c], V)) || (rc = zeros(&w.ck_l[c], V))
            || (rc = zeros(&w.ck_pvt[c], V)) || (rc = zeros(&w.ck_vsqg[c], V)))
            return rc;
    w.tape.resize(k);
    for (int j = 0; j < k; ++j) {
        StepTape &t = w.tape[j];
        if ((rc = alloc_state(t.s)) || (rc = alloc_state(t.s1)) || (rc = zeros(&t.td0_rho, S))
            || (rc = zeros(&t.td0_rt, S)) || (rc = zeros(&t.td0_wind, V)) || (rc = zeros(&t.td1_rho, S))
            || (rc = zeros(&t.td1_rt, S)) || (rc = zeros(&t.td1_wind, V)) || (rc = zeros(&t.flux, V))
            || (rc = zeros(&t.pv, CURL)))
            return rc;
    }
    if ((rc = alloc_state(w.seg_end)) || (rc = alloc_cot(w.bS)) || (rc = alloc_cot(w.bSn)) || (rc = alloc_cot(w.bS1)))
        return rc;
    double **sv[] = {&w.bt0_rho, &w.bt0_rt, &w.bt1_rho, &w.bt1_rt, &w.vsqb, &w.ones};
    for (double **p : sv) if ((rc = zeros(p, S))) return rc;
    double **vv[] = {&w.bt0_wind, &w.bt1_wind, &w.bP, &w.bPold, &w.bA, &w.bA1, &w.bAprev, &w.bU, &w.tmpV1, &w.tmpV2,
                     &w.fluxb};
    for (double **p : vv) if ((rc = zeros(p, V))) return rc;
    double **ev[] = {&w.Fb, &w.Fthb, &w.cvb};
    for (double **p : ev) if ((rc = zeros(p, L*NVH))) return rc;
    double **cv[] = {&w.pvb, &w.rvb, &w.dvb};
    for (double **p : cv) if ((rc = zeros(p, CURL))) return rc;
    if ((rc = zeros(&w.zb, L*NT)) || (rc = zeros(&w.ws, (size_t)sadj::VW_N*L*NSH))
        || (rc = zeros(&w.vcb, 2*(size_t)M->G.n_oro*NSH + 1)))
        return rc;
    DevWork &A = w.AW;
    if ((rc = zeros(&A.flux, V)) || (rc = zeros(&A.rvt, L*NT)) || (rc = zeros(&A.pv, CURL)) || (rc = zeros(&A.pvt, V))
        || (rc = zeros(&A.vsq, S)) || (rc = zeros(&A.vsq_grad, V)))
        return rc;
    std::vector<double> one(S, 1.0);
    HIP_OK(hipMemcpyAsync(w.ones, one.data(), S*sizeof(double), hipMemcpyHostToDevice, M->stream));
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

int copy(double *dst, const double *src, size_t n)
{
    HIP_OK(hipMemcpyAsync(dst, src, n*sizeof(double), hipMemcpyDeviceToDevice, M->stream));
    return 0;
}

int copy_state(const DevState &dst, const DevState &src)
{
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    int rc;
    if ((rc = copy(dst.rho, src.rho, S)) || (rc = copy(dst.rt, src.rt, S)) || (rc = copy(dst.tp, src.tp, S))
        || (rc = copy(dst.ep, src.ep, S)) || (rc = copy(dst.wind, src.wind, V)))
        return rc;
    return 0;
}

int zero(double *p, size_t n)
{
    HIP_OK(hipMemsetAsync(p, 0, n*sizeof(double), M->stream));
    return 0;
}

/* zeroes up to ZERO_MAX arrays in one launch */
#define ZERO_MAX 20
struct ZeroMany {
    double *p[ZERO_MAX];
    i64 n[ZERO_MAX];
    int cnt;
    GAME_HD void operator()(i64 t) const
    {
        for (int i = 0; i < cnt; ++i)
            if (t < n[i]) p[i][t] = 0.0;
    }
};

struct ZeroList {
    ZeroMany z{};
    i64 longest = 0;
    void add(double *p, size_t n)
    {
        z.p[z.cnt] = p;
        z.n[z.cnt++] = (i64)n;
        if ((i64)n > longest) longest = (i64)n;
    }
    void add_cot(const Cot &c)
    {
        const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
        add(c.rho, S); add(c.rt, S); add(c.tp, S); add(c.ep, S); add(c.wind, V);
    }
    void run(hipStream_t st) const { launch("k_zero_many", z, longest, st); }
};

int zero_cot(const Cot &c)
{
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    int rc;
    if ((rc = zero(c.rho, S)) || (rc = zero(c.rt, S)) || (rc = zero(c.tp, S)) || (rc = zero(c.ep, S))
        || (rc = zero(c.wind, V)))
        return rc;
    return 0;
}

int check_window_model(void)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (M->precision != GAME_FP64 || M->d.n_const != 1) {
        snprintf(g_err, sizeof g_err, "the adjoint window needs the FP64 form of a dry model");
        return GAME_ERR_UNSUPPORTED;
    }
    return 0;
}

/* the state of the control variables into both state slots */
int load_controls(const double *rho0, const double *rt0, const double *wind0)
{
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    int rc;
    DevState &n = M->s[GAME_STATE_NEW];
    if ((rc = copy(n.rho, rho0, S)) || (rc = copy(n.rt, rt0, S)) || (rc = copy(n.wind, wind0, V))) return rc;
    EosState f{M->G, n.rho, n.rt, n.tp, n.ep};
    launch("k_eos_state", f, M->G.s, M->stream);
    return copy_state(M->s[GAME_STATE_OLD], n);
}

/* one step as game_gpu_run_steps takes it: the state buffers exchanged, then
   manage_rkhevi, eagerly when taped */
int window_step(double dt, int first, const StepTape *tape)
{
    DevState t = M->s[GAME_STATE_OLD];
    M->s[GAME_STATE_OLD] = M->s[GAME_STATE_NEW];
    M->s[GAME_STATE_NEW] = t;
    return step_fp64(dt, first, tape);
}

int save_checkpoint(int c)
{
    Window &w = *Wn;
    const size_t V = (size_t)M->d.v;
    int rc;
    if ((rc = copy_state(w.ck_state[c], M->s[GAME_STATE_NEW])) || (rc = copy(w.ck_nl[c], M->W.pgrad_nl, V))
        || (rc = copy(w.ck_l[c], M->W.pgrad_l, V)) || (rc = copy(w.ck_pvt[c], M->W.pvt, V))
        || (rc = copy(w.ck_vsqg[c], M->W.vsq_grad, V)))
        return rc;
    return 0;
}

int restore_checkpoint(int c)
{
    Window &w = *Wn;
    const size_t V = (size_t)M->d.v;
    int rc;
    if ((rc = copy_state(M->s[GAME_STATE_NEW], w.ck_state[c])) || (rc = copy_state(M->s[GAME_STATE_OLD], w.ck_state[c]))
        || (rc = copy(M->W.pgrad_nl, w.ck_nl[c], V)) || (rc = copy(M->W.pgrad_l, w.ck_l[c], V))
        || (rc = copy(M->W.pvt, w.ck_pvt[c], V)) || (rc = copy(M->W.vsq_grad, w.ck_vsqg[c], V)))
        return rc;
    return 0;
}

/* reverse of the momentum advection of (wind, rho) for the cotangent advb of
   pot_vort_tend - 0.5 v_squared_grad (vector_tendencies_expl.c:24-35),
   accumulated into b_wind and b_rho */
/* the reverse of the momentum advection of (wind, rho); flux and pv are the
   forward's mass flux and potential vorticity if taped, else null (they are
   then recomputed) */
int advection_adjoint(const double *wind, const double *rho, const double *flux, const double *pv,
                      const double *advb, double *b_wind, double *b_rho)
{
    Window &w = *Wn;
    const DevGrid &G = M->G;
    const AdjTables &A = M->A;
    hipStream_t st = M->stream;
    if (!flux || !pv) {
        hor::advection(G, wind, rho, w.AW, st);
        flux = w.AW.flux;
        pv = w.AW.pv;
    }
    /* kinetic energy: -0.5 grad(v.v).  With two streams it runs on the side
       stream beside the TRSK and PV reverse below, which does not touch
       b_wind until the curl's adjoint, where the two join */
    const hipStream_t side = (M->concurrent && !g_kernel_timer) ? M->side : nullptr;
    if (side) {
        HIP_OK(hipEventRecord(M->ev_fork, st));
        HIP_OK(hipStreamWaitEvent(side, M->ev_fork, 0));
    }
    adj::grad_adjoint(G, A, advb, w.vsqb, side ? side : st, -0.5);
    adj::inner_product_adjoint(G, A, wind, w.vsqb, b_wind, side ? side : st, 2.0, 1);
    /* TRSK flux: in the mass flux (with two streams on the side stream too: its
       cotangent is read only by the mass flux's reverse after the join) and
       in the PV */
    adj::vorticity_flux_adjoint(G, A, pv, advb, w.fluxb, side ? side : st);
    const int vf_chunks = (G.L + VF_CH - 1)/VF_CH;
    if (M->PR.n_max <= 12)
        launch("k_vflux_pv_adj_rh", sadj::VFluxPvAdjRhN<12>{G, M->PR, flux, advb, w.pvb, vf_chunks},
               (i64)vf_chunks*G.nvh, st);
    else
        launch("k_vflux_pv_adj_rh", sadj::VFluxPvAdjRh{G, M->PR, flux, advb, w.pvb, vf_chunks}, (i64)vf_chunks*G.nvh, st);
    launch("k_vflux_pv_adj_tan", sadj::VFluxPvAdjTan{G, A, flux, advb, w.pvb}, (i64)(G.L + 1)*G.nvh, st);
    /* PV = (rel_vort + f)/density */
    launch("k_pv_div_adj", sadj::PvDivAdj{G, w.pvb, pv, rho, w.rvb, w.dvb}, G.curl, st);
    launch("k_rhombus_dens_adj", sadj::RhombusDensAdj{G, A, w.dvb, b_rho}, G.s, st);
    launch("k_half_dens_adj", sadj::HalfDensAdj{G, w.dvb, b_rho}, G.s, st);
    launch("k_rhombus_vort_adj", sadj::RhombusVortAdj{G, w.rvb, w.zb}, (i64)G.L*G.nt, st);
    if (side) {
        HIP_OK(hipEventRecord(M->ev_join, side));
        HIP_OK(hipStreamWaitEvent(st, M->ev_join, 0));
    }
    adj::curl_triangles_adjoint(G, A, w.zb, b_wind, st, 1);
    launch("k_tan_vort_adj_e", sadj::TanVortAdjE{G, w.rvb, w.cvb, b_wind}, (i64)G.L*G.nvh, st);
    launch("k_tan_vort_adj_c", sadj::TanVortAdjC{G, A, w.rvb, w.cvb, b_wind}, (i64)(G.L + 1)*G.nsh, st);
    /* mass flux = rho * wind */
    launch("k_stv_adj_v", sadj::StvAdjV{G, rho, nullptr, w.fluxb, 1.0, 1, 1, b_wind}, G.v, st);
    launch("k_stv_adj_s", sadj::StvAdjS{G, wind, nullptr, w.fluxb, 1.0, 1, b_rho}, G.s, st);
    return 0;
}

/* reverse of step n from its tape; w.bS holds the cotangent of the state the
   step produced (whose wind edges are next->wind), w.bP the cotangent of the
   step's pressure gradient nl + l, w.bA that of its second-substep advection;
   on return w.bS, w.bP, w.bA hold the same for the step before */
/* the reverse of the gas's scalar tendencies (scalar_tendencies_expl.c:107-176)
   of densities (rho, rt) with the updated wind: the wind's cotangent goes to
   w.bU, the densities' are added to (b_rho, b_rt) */
static void scalar_tend_adjoint(const double *rho, const double *rt, const double *wind, const double *b_trho,
                                const double *b_trt, double *b_rho, double *b_rt)
{
    Window &w = *Wn;
    const DevGrid &G = M->G;
    hipStream_t st = M->stream;
    if (!G.flat) launch("k_vcc_cot", sadj::VccCot{G, b_trho, b_trt, w.vcb}, (i64)G.n_oro*G.nsh, st);
    launch("k_scalar_tend_adj_e", sadj::ScalarTendAdjE{G, M->A, rho, rt, wind, b_trho, b_trt, w.vcb, w.Fb, w.Fthb,
                                                       w.bU}, (i64)G.L*G.nvh, st);
    launch("k_scalar_tend_adj_c", sadj::ScalarTendAdjC{G, rho, rt, w.Fb, w.Fthb, b_rho, b_rt}, G.s, st);
}

/* the reverse of the sound-wave solver of one substep (column_solvers.c:19-383) */
static void ver_waves_adjoint(const sadj::VwArgs &a, hipStream_t st)
{
    const DevGrid &G = a.G;
    const i64 ni = (i64)(G.L - 1)*G.nsh;
    launch("k_vw_layer_fwd", sadj::VwLayerFwd{a}, G.s, st);
    launch("k_vw_iface_fwd", sadj::VwIfaceFwd{a}, ni, st);
    launch("k_vw_solve", sadj::VwSolve{a, 0}, G.nsh, st);
    launch("k_vw_layer_new", sadj::VwLayerNew{a}, G.s, st);
    launch("k_vw_iface_sol_bar", sadj::VwIfaceSolBar{a}, ni, st);
    launch("k_vw_solve_t", sadj::VwSolve{a, 1}, G.nsh, st);
    launch("k_vw_iface_bar", sadj::VwIfaceBar{a}, ni, st);
    launch("k_vw_layer_bar", sadj::VwLayerBar{a}, G.s, st);
}

int step_adjoint(const StepTape &T, const DevState &next, double dt, int first)
{
    Window &w = *Wn;
    const DevGrid &G = M->G;
    const AdjTables &A = M->A;
    hipStream_t st = M->stream;
    const double iw = M->cfg.impl_thermo_weight;
    const double chw = 0.5 + iw, ohw = 1.0 - chw, cvw = 1.0 - iw;
    const size_t S = (size_t)G.s, V = (size_t)G.v;
    int rc;
    {
        ZeroList zl;
        zl.add_cot(w.bSn);
        zl.add_cot(w.bS1);
        zl.add(w.bt0_rho, S); zl.add(w.bt0_rt, S); zl.add(w.bt0_wind, V);
        zl.add(w.bt1_rho, S); zl.add(w.bt1_rt, S); zl.add(w.bt1_wind, V);
        zl.add(w.bPold, V);
        zl.run(st);
    }
    /* the gradient of the step's exner_pert, which the pressure gradient's
       reverse below needs, depends on the tape alone: with two streams it is
       formed on the side stream now (joined by the advection reverse's join
       or at the end).  It writes the edges and interior w points of tmpV1,
       which nothing else writes: its top and surface w points keep the
       allocation's zero */
    const hipStream_t side = (M->concurrent && !g_kernel_timer) ? M->side : nullptr;
    if (side) {
        HIP_OK(hipEventRecord(M->ev_fork, st));
        HIP_OK(hipStreamWaitEvent(side, M->ev_fork, 0));
        hor::grad(G, T.s.ep, w.tmpV1, side);
    }
    /* ---- second substep ---- */
    {
        sadj::VwArgs a{G, 1, dt, iw, T.s.rho, T.s.rt, T.s.tp, T.s.ep, T.s.wind, T.s1.rho, T.s1.rt, T.s1.ep,
                       T.td1_rho, T.td1_rt, T.td1_wind, w.bS.rho, w.bS.rt, w.bS.wind, w.bS.tp, w.bS.ep,
                       w.bSn.rho, w.bSn.rt, w.bSn.tp, w.bSn.ep, w.bSn.wind, w.bS1.rho, w.bS1.rt, w.bS1.ep,
                       w.bt1_rho, w.bt1_rt, w.bt1_wind, w.ws};
        ver_waves_adjoint(a, st);
    }
    launch("k_pass_adj", PassThroughAdj{G, w.bS.wind, w.bU, w.bSn.wind}, G.v, st);
    scalar_tend_adjoint(T.s1.rho, T.s1.rt, next.wind, w.bt1_rho, w.bt1_rt, w.bS1.rho, w.bS1.rt);
    launch("k_wind_update_adj", sadj::WindUpdateAdj{G, dt, w.bU, w.bSn.wind, w.bt1_wind, 0}, G.v, st);
    launch("k_tend_adj_rk1", sadj::TendAdjRk1{G, ohw, chw, cvw, w.bt1_wind, w.bA, w.bt0_wind, w.bPold, w.bP, w.bA1},
           G.v, st);
    if ((rc = advection_adjoint(T.s1.wind, T.s1.rho, T.flux, T.pv, w.bA1, w.bS1.wind, w.bS1.rho))) return rc;
    /* ---- first substep ---- */
    {
        sadj::VwArgs a{G, 0, dt, iw, T.s.rho, T.s.rt, T.s.tp, T.s.ep, T.s.wind, T.s.rho, T.s.rt, T.s.ep,
                       T.td0_rho, T.td0_rt, T.td0_wind, w.bS1.rho, w.bS1.rt, w.bS1.wind, w.bS1.tp, w.bS1.ep,
                       w.bSn.rho, w.bSn.rt, w.bSn.tp, w.bSn.ep, w.bSn.wind, w.bSn.rho, w.bSn.rt, w.bSn.ep,
                       w.bt0_rho, w.bt0_rt, w.bt0_wind, w.ws};
        ver_waves_adjoint(a, st);
    }
    launch("k_pass_adj", PassThroughAdj{G, w.bS1.wind, w.bU, w.bSn.wind}, G.v, st);
    scalar_tend_adjoint(T.s.rho, T.s.rt, T.s1.wind, w.bt0_rho, w.bt0_rt, w.bSn.rho, w.bSn.rt);
    launch("k_wind_update_adj", sadj::WindUpdateAdj{G, dt, w.bU, w.bSn.wind, w.bt0_wind, 1}, G.v, st);
    launch("k_tend_adj_rk0", sadj::TendAdjRk0{G, ohw, chw, cvw, w.bt0_wind, w.bPold, w.bP, w.bAprev}, G.v, st);
    if (first) {
        /* pgrad_old = -(nl + l) of this step; the advection of the first
           substep computed from this step's state */
        launch("k_axpy", sadj::Axpy{G.v, -1.0, w.bPold, w.bP, 1}, G.v, st);
        if ((rc = advection_adjoint(T.s.wind, T.s.rho, nullptr, nullptr, w.bAprev, w.bSn.wind, w.bSn.rho)))
            return rc;
    }
    /* ---- pressure gradient: P = mean(c_p (theta_bg + tp)) grad(ep) + mean(c_p tp) grad(exner_bg) ---- */
    if (side) {
        HIP_OK(hipEventRecord(M->ev_join, side));
        HIP_OK(hipStreamWaitEvent(st, M->ev_join, 0));
    } else hor::grad(G, T.s.ep, w.tmpV1, st);
    launch("k_stv_adj_v", sadj::StvAdjV{G, G.thbg, T.s.tp, w.bP, GC_C_D_P, 1, 0, w.tmpV2}, G.v, st);
    adj::grad_adjoint(G, A, w.tmpV2, w.bSn.ep, st, 1.0, 1);
    launch("k_stv_adj_s", sadj::StvAdjS{G, w.tmpV1, G.exbg_grad, w.bP, GC_C_D_P, 1, w.bSn.tp}, G.s, st);
    /* theta_v_pert is re-derived from rho and rhotheta_v by the driver's copy */
    launch("k_theta_adj", sadj::ThetaBoundaryAdj{G, T.s.rho, T.s.rt, w.bSn.rho, w.bSn.rt, w.bSn.tp}, G.s, st);
    /* hand over to the step before */
    std::swap(w.bS, w.bSn);
    if (!first) launch("k_axpy", sadj::Axpy{G.v, -1.0, w.bPold, w.bP, 0}, G.v, st);
    std::swap(w.bA, w.bAprev);
    HIP_OK(hipGetLastError());
    return 0;
}

/* the forward sweep (checkpoints every k steps when k > 0) and, with a
   terminal cotangent (set by `terminal`), the reverse sweep */
template <class Terminal>
int window_run(int n_steps, double dt, int k, const double *rho0, const double *rt0, const double *wind0,
               const Terminal &terminal, double *rho_bar, double *rt_bar, double *wind_bar)
{
    int rc;
    const int n_ckpt = (n_steps + k - 1)/k;
    if ((rc = window_buffers(k, n_ckpt))) return rc;
    if ((rc = build_adjoint_tables())) return rc;
    Window &w = *Wn;
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    hipStream_t st = M->stream;
    /* forward sweep with checkpoints; the carried forcings of a checkpoint at
       step 0 are not read (the first step forms its own) */
    /* the last segment is taped as the sweep runs it, so the reverse sweep
       starts without recomputing it */
    const int last0 = (n_ckpt - 1)*k;
    w.n_ev = 0;
    w.mark(PH_FORWARD, st);
    if ((rc = load_controls(rho0, rt0, wind0))) return rc;
    for (int n = 0; n < n_steps; ++n) {
        if (n % k == 0 && n < last0 && (rc = save_checkpoint(n/k))) return rc;
        if (n >= last0) rc = window_step(dt, n == 0, &w.tape[n - last0]);
        else rc = game_gpu_run_steps(1, dt, n == 0);               /* graph replays after the first */
        if (rc) return rc;
    }
    /* the terminal cotangent from the final state */
    if ((rc = zero_cot(w.bS)) || (rc = zero(w.bP, V)) || (rc = zero(w.bA, V))) return rc;
    if ((rc = terminal(M->s[GAME_STATE_NEW], w.bS))) return rc;
    /* reverse sweep by segments */
    for (int c = n_ckpt - 1; c >= 0; --c) {
        const int n0 = c*k, n1 = (n0 + k < n_steps) ? n0 + k : n_steps;
        w.mark(c == n_ckpt - 1 ? PH_FORWARD : PH_REVERSE, st);
        if (c < n_ckpt - 1) {
            if ((rc = restore_checkpoint(c))) return rc;
            for (int n = n0; n < n1; ++n)
                if ((rc = window_step(dt, n == 0, &w.tape[n - n0]))) return rc;
        }
        if ((rc = copy_state(w.seg_end, M->s[GAME_STATE_NEW]))) return rc;
        w.mark(PH_RECOMPUTE, st);
        for (int n = n1 - 1; n >= n0; --n) {
            const DevState &next = (n == n1 - 1) ? w.seg_end : w.tape[n - n0 + 1].s;
            if ((rc = step_adjoint(w.tape[n - n0], next, dt, n == 0))) return rc;
        }
    }
    w.mark(PH_REVERSE, st);
    /* the control variables: rho, rhotheta_v, wind */
    {
        const DevState &s0 = w.tape[0].s;
        EosStateAdj f{M->G, s0.rho, s0.rt, w.bS.ep, w.bS.tp, w.bS.rho, w.bS.rt};
        launch("k_eos_adj", f, M->G.s, st);
    }
    if ((rc = copy(rho_bar, w.bS.rho, S)) || (rc = copy(rt_bar, w.bS.rt, S)) || (rc = copy(wind_bar, w.bS.wind, V)))
        return rc;
    w.mark(PH_FINISH, st);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));
    w.read_clock();
    return 0;
}

}  // namespace

extern "C" {

int game_gpu_window_forward(int n_steps, double dt, const double *rho0, const double *rt0, const double *wind0,
                            double *rho_out, double *rt_out, double *exner_out, double *wind_out)
{
    int rc = check_window_model();
    if (rc) return rc;
    if (n_steps < 1) { snprintf(g_err, sizeof g_err, "n_steps must be >= 1"); return GAME_ERR_ARG; }
    if (!rho0 || !rt0 || !wind0) { snprintf(g_err, sizeof g_err, "null initial state"); return GAME_ERR_ARG; }
    if ((rc = load_controls(rho0, rt0, wind0))) return rc;
    if ((rc = game_gpu_run_steps(n_steps, dt, 1))) return rc;
    const DevState &s = M->s[GAME_STATE_NEW];
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    if ((rho_out && (rc = copy(rho_out, s.rho, S))) || (rt_out && (rc = copy(rt_out, s.rt, S)))
        || (exner_out && (rc = copy(exner_out, s.ep, S))) || (wind_out && (rc = copy(wind_out, s.wind, V))))
        return rc;
    HIP_OK(hipStreamSynchronize(M->stream));
    return 0;
}

int game_gpu_window_vjp(int n_steps, double dt, int checkpoint_every, const double *rho0, const double *rt0,
                        const double *wind0, const double *rho_bar_n, const double *rt_bar_n,
                        const double *exner_bar_n, const double *wind_bar_n, double *rho_bar, double *rt_bar,
                        double *wind_bar)
{
    int rc = check_window_model();
    if (rc) return rc;
    if (n_steps < 1 || checkpoint_every < 1) {
        snprintf(g_err, sizeof g_err, "n_steps and checkpoint_every must be >= 1");
        return GAME_ERR_ARG;
    }
    if (!rho0 || !rt0 || !wind0 || !rho_bar || !rt_bar || !wind_bar) {
        snprintf(g_err, sizeof g_err, "null initial state or gradient buffer");
        return GAME_ERR_ARG;
    }
    const size_t S = (size_t)M->d.s, V = (size_t)M->d.v;
    auto terminal = [&](const DevState &, const Cot &b) -> int {
        int r;
        if ((rho_bar_n && (r = copy(b.rho, rho_bar_n, S))) || (rt_bar_n && (r = copy(b.rt, rt_bar_n, S)))
            || (exner_bar_n && (r = copy(b.ep, exner_bar_n, S))) || (wind_bar_n && (r = copy(b.wind, wind_bar_n, V))))
            return r;
        return 0;
    };
    return window_run(n_steps, dt, checkpoint_every, rho0, rt0, wind0, terminal, rho_bar, rt_bar, wind_bar);
}

int game_gpu_window_gradient(int n_steps, double dt, int checkpoint_every, const double *rho0, const double *rt0,
                             const double *wind0, const double *rho_target, double *cost, double *rho_bar,
                             double *rt_bar, double *wind_bar)
{
    int rc = check_window_model();
    if (rc) return rc;
    if (n_steps < 1 || checkpoint_every < 1 || !cost) {
        snprintf(g_err, sizeof g_err, "n_steps and checkpoint_every must be >= 1 and cost non-null");
        return GAME_ERR_ARG;
    }
    if (!rho0 || !rt0 || !wind0 || !rho_target || !rho_bar || !rt_bar || !wind_bar) {
        snprintf(g_err, sizeof g_err, "null initial state, target or gradient buffer");
        return GAME_ERR_ARG;
    }
    const DevGrid &G = M->G;
    auto terminal = [&](const DevState &s, const Cot &b) -> int {
        /* J = 1/2 sum (rho_N - rho_target)^2: its cotangent is rho_N - rho_target */
        launch("k_diff", DiffToCot{G.s, s.rho, rho_target, b.rho}, G.s, M->stream);
        const int blocks = col::integral_partials(G.s, b.rho, b.rho, Wn->ones, M->partials, M->stream);
        HIP_OK(hipMemcpyAsync(M->h_partials, M->partials, sizeof(double)*2*(size_t)blocks, hipMemcpyDeviceToHost,
                              M->stream));
        HIP_OK(hipStreamSynchronize(M->stream));
        double sum = 0.0, comp = 0.0;
        for (int q = 0; q < blocks; ++q) {
            const double v = M->h_partials[2*q], t = sum + v;
            comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
            sum = t;
            comp += M->h_partials[2*q + 1];
        }
        *cost = 0.5*(sum + comp);
        return 0;
    };
    return window_run(n_steps, dt, checkpoint_every, rho0, rt0, wind0, terminal, rho_bar, rt_bar, wind_bar);
}

int game_gpu_window_last_timing(double *ms)
{
    if (!M) return GAME_ERR_NOT_INIT;
    if (!ms) return GAME_ERR_ARG;
    for (int i = 0; i < 4; ++i) ms[i] = (Wn && Wn->owner == M->generation) ? Wn->phase_ms[i] : 0.0;
    return 0;
}

}  /* extern "C" */
