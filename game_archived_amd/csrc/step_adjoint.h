/*
 * step_adjoint.h -- the reverse (adjoint) of the dry RKHEVI step, as
 * element functors: each struct holds its arguments and computes, in
 * operator()(t), every output element owned by index t.  rkhevi_window.hip
 * launches them (one thread per t) in the order of the reverse sweep.  The
 * reference has no adjoint; each functor cites the forward lines it reverses.
 *
 * Conventions: a cotangent array named xb holds d(J)/d(x) for the forward
 * array x (same layout); "+=" functors accumulate, every element owned by one
 * thread, so no atomics are needed.  Everything is FP64.
 */
#ifndef GAME_STEP_ADJOINT_H
#define GAME_STEP_ADJOINT_H

#include "rkhevi_dev.h"
#include "rkhevi_adjoint.h"

#define GAME_HD __host__ __device__

/* reverse table of the rhombus PV read by the TRSK sums (vorticity_flux.c:46-109):
   for rhombus m, the (output edge, weight, flux edge) of every term that reads
   PV m, padded slot-major to PVR_MAX entries with weight 0; the hexagon
   half-averages contribute one entry with half the weight for the modified
   index and one for the edge's own rhombus */
#define PVR_MAX 20
struct PvReverse {
    const int *e;        /* PVR_MAX x NVH: output edge */
    const int *f;        /* PVR_MAX x NVH: flux edge */
    const double *w;     /* PVR_MAX x NVH: weight */
    const int *n;        /* NVH: the number of entries of each rhombus */
    int n_max;           /* the largest of them (12 on the icosahedral grids) */
};

namespace sadj {

GAME_HD inline i64 eidx(const DevGrid &G, int l, int e) { return G.nsh + (i64)l*G.vpl + e; }
GAME_HD inline i64 widx(const DevGrid &G, int lev, int h) { return (i64)lev*G.vpl + h; }
GAME_HD inline i64 sidx(const DevGrid &G, int l, int h) { return (i64)l*G.nsh + h; }
/* (q, r) = (t / n, t % n) for a flat index below 2^32 (every array here) */
GAME_HD inline void split_index(i64 t, int n, int &q, int &r)
{
    const unsigned u = (unsigned)t, m = (unsigned)n;
    q = (int)(u/m);
    r = (int)(u - (unsigned)q*m);
}

/* ------------------------------------------------------------------ */
/* three_band_solver_ver_waves, column_solvers.c:19-383: one column     */
/* (L layers, interfaces j = 0..L-2 between layers j and j+1 at level   */
/* j+1):                                                                */
/*   re = rho_o + dt t_rho, rte = rt_o + dt t_rt,                       */
/*   alpha, beta, gamma (old state at rk 0; the mean of old and new at  */
/*   rk 1), theta_e = tp_o + dt vol (alpha t_rho + beta t_rt),          */
/*   pi_e = ep_o + dt vol gamma t_rt, rho_i = mean of rho_o,            */
/*   rho_ie = mean of re, theta_i = mean of rt_n/rho_n,                 */
/*   the bands c, d, e and right-hand side r, x = A^-1 r, damping,      */
/*   then rho, rt, w, tp, ep of the new state.                          */
/* The reverse runs as seven launches: the per-layer coefficients       */
/* (VwLayerFwd), the per-interface bands (VwIfaceFwd), the solve        */
/* (VwSolve, one thread per column), the cotangent of the solution      */
/* (VwIfaceSolBar), the transposed solve (VwSolveT), the interface      */
/* cotangents (VwIfaceBar) and the per-layer cotangents (VwLayerBar).   */
/* Every launch but the two solves has one thread per (layer, column)   */
/* or (interface, column), and every contribution an interface makes to */
/* its neighbours is gathered by the neighbour, so no element has two  */
/* writers.  The workspace holds VW_N arrays of L doubles per column,   */
/* array k of column c element j at (k*L + j)*ncol + c.                 */
/* ------------------------------------------------------------------ */
enum { VW_AL, VW_BE, VW_GA, VW_TIN, VW_C, VW_D, VW_E, VW_R, VW_X, VW_EP, VW_BX, VW_BR, VW_BTIN, VW_BRIO,
       VW_RN, VW_RTN, VW_BRE, VW_BRTE, VW_N };

struct VwArgs {
    DevGrid G;
    int rk;
    double dt, iw;
    /* forward inputs: old state, new state as read (rk 1; rk 0: the old), tendency */
    const double *rho_o, *rt_o, *tp_o, *ep_o, *wind_o;
    const double *rho_n, *rt_n, *ep_n;
    const double *t_rho, *t_rt, *t_wind;
    /* cotangents of the outputs: rho, rt, w (interior levels), tp, ep of the new state */
    const double *b_rho, *b_rt, *b_wind, *b_tp, *b_ep;
    /* accumulated: cotangents of the old state, of the new state as read (rk 1
       only; at rk 0 they are added to the old state's), of the tendency */
    double *bo_rho, *bo_rt, *bo_tp, *bo_ep, *bo_wind;
    double *bn_rho, *bn_rt, *bn_ep;
    double *bt_rho, *bt_rt, *bt_wind;
    double *ws;

    GAME_HD double &W(int k, int j, int c) const { return ws[((i64)k*G.L + j)*G.nsh + c]; }
    /* (layer or interface, column) of a flat thread index */
    GAME_HD void split(i64 t, int &j, int &c) const
    {
        const unsigned u = (unsigned)t, n = (unsigned)G.nsh;
        j = (int)(u/n);
        c = (int)(u - (unsigned)j*n);
    }
    GAME_HD const double *rn_() const { return rk == 0 ? rho_o : rho_n; }
    GAME_HD const double *rtn_() const { return rk == 0 ? rt_o : rt_n; }
    GAME_HD double rio(int j, int c) const
    {
        return 0.5*(rho_o[sidx(G, j, c)] + rho_o[sidx(G, j + 1, c)]);
    }
    GAME_HD double re(int j, int c) const { const i64 b = sidx(G, j, c); return rho_o[b] + dt*t_rho[b]; }
    GAME_HD double rte(int j, int c) const { const i64 b = sidx(G, j, c); return rt_o[b] + dt*t_rt[b]; }
    GAME_HD double tpe(int j, int c) const
    {
        const i64 b = sidx(G, j, c);
        return tp_o[b] + dt*G.vol[b]*(W(VW_AL, j, c)*t_rho[b] + W(VW_BE, j, c)*t_rt[b]);
    }
    GAME_HD double epe(int j, int c) const
    {
        const i64 b = sidx(G, j, c);
        return ep_o[b] + dt*G.vol[b]*W(VW_GA, j, c)*t_rt[b];
    }
    GAME_HD double damp(int j, int c) const { return 1.0 + dt*G.damp[sidx(G, j, c)]; }
    /* the damped solution at interface j */
    GAME_HD double sol(int j, int c) const { return W(VW_X, j, c)/damp(j, c); }
    /* rho and rt of the new state in layer j */
    GAME_HD void rnew(int j, int c, double &rn, double &rtn) const
    {
        const int n = G.L - 1;
        const double su = j < n ? sol(j, c) : 0.0, sd = j > 0 ? sol(j - 1, c) : 0.0;
        const double tu = j < n ? W(VW_TIN, j, c) : 0.0, td = j > 0 ? W(VW_TIN, j - 1, c) : 0.0;
        const double vol = G.vol[sidx(G, j, c)];
        rn = re(j, c) + dt*(-sd + su)/vol;
        rtn = rte(j, c) + dt*(-td*sd + tu*su)/vol;
    }
    /* w_new[j] = (2 sol/A - din wo)/rho_i, din = mean of rho_new: the
       cotangent it gives din */
    GAME_HD double bdin(int j, int c) const
    {
        const i64 wi = widx(G, j + 1, c);
        return -wind_o[wi]/rio(j, c)*b_wind[wi];
    }
    /* cotangents of rho_new and rt_new in layer j from the new state's rho,
       rt, tp (= rt/rho - thbg), ep (= ep_o + vol ga (rt - rt_o)) and w */
    GAME_HD void bnew(int j, int c, double rn, double rtn, double &bre, double &brte) const
    {
        const int n = G.L - 1;
        const i64 b = sidx(G, j, c);
        const double bep = b_ep[b], btp = b_tp[b];
        brte = b_rt[b] + G.vol[b]*W(VW_GA, j, c)*bep + btp/rn;
        bre = b_rho[b] - btp*rtn/(rn*rn);
        if (j > 0) bre += 0.5*bdin(j - 1, c);
        if (j < n) bre += 0.5*bdin(j, c);
    }
};

/* alpha, beta, gamma of layer j (column_solvers.c:99-121) */
struct VwLayerFwd {
    VwArgs a;
    GAME_HD void operator()(i64 t) const
    {
        const DevGrid &G = a.G;
        if (t >= G.s) return;
        int j, c;
        a.split(t, j, c);
        const i64 b = sidx(G, j, c);
        const double ro = a.rho_o[b], rto = a.rt_o[b], vol = G.vol[b], exbg = G.exbg[b];
        double al, be, ga;
        if (a.rk == 0) {
            al = -rto/(ro*ro)/vol;
            be = 1.0/ro/vol;
            ga = GC_R_D/(GC_C_D_V*rto)*(exbg + a.ep_o[b])/vol;
        } else {
            const double pdw = 0.5;
            const double rn = a.rho_n[b], rtn = a.rt_n[b];
            const double ao = -rto/(ro*ro), bo = 1.0/ro, go = GC_R_D/(GC_C_D_V*rto)*(exbg + a.ep_o[b]);
            const double an = -rtn/(rn*rn), bn = 1.0/rn, gn = GC_R_D/(GC_C_D_V*rtn)*(exbg + a.ep_n[b]);
            al = ((1.0 - pdw)*ao + pdw*an)/vol;
            be = ((1.0 - pdw)*bo + pdw*bn)/vol;
            ga = ((1.0 - pdw)*go + pdw*gn)/vol;
        }
        a.W(VW_AL, j, c) = al;
        a.W(VW_BE, j, c) = be;
        a.W(VW_GA, j, c) = ga;
    }
};

/* theta_i, the bands and the right-hand side of interface j
   (column_solvers.c:123-190) */
struct VwIfaceFwd {
    VwArgs a;
    GAME_HD double tin(int j, int c) const
    {
        const double *rn = a.rn_(), *rtn = a.rtn_();
        const i64 b = sidx(a.G, j, c), lo = sidx(a.G, j + 1, c);
        return 0.5*(rtn[b]/rn[b] + rtn[lo]/rn[lo]);
    }
    GAME_HD void operator()(i64 t) const
    {
        const DevGrid &G = a.G;
        const int n = G.L - 1;
        if (t >= (i64)n*G.nsh) return;
        int j, c;
        a.split(t, j, c);
        const double dt = a.dt, iw = a.iw;
        const i64 b = sidx(G, j, c), lo = sidx(G, j + 1, c), wi = widx(G, j + 1, c);
        const double ti = tin(j, c), rio = a.rio(j, c), wo = a.wind_o[wi];
        const double rie = 0.5*(a.re(j, c) + a.re(j + 1, c));
        const double dz = G.zs[b] - G.zs[lo], dex = G.exbg[b] - G.exbg[lo];
        const double q = dz/(iw*(dt*dt)*GC_C_D_P);
        const double ga0 = a.W(VW_GA, j, c), ga1 = a.W(VW_GA, j + 1, c);
        const double al0 = a.W(VW_AL, j, c), al1 = a.W(VW_AL, j + 1, c);
        const double be0 = a.W(VW_BE, j, c), be1 = a.W(VW_BE, j + 1, c);
        a.W(VW_TIN, j, c) = ti;
        a.W(VW_D, j, c) = -(ti*ti)*(ga0 + ga1) + 0.5*dex*(al1 - al0 + ti*(be1 - be0))
                          - dz/(iw*(dt*dt)*GC_C_D_P*rio)*(2.0/G.area[wi] + dt*wo*0.5*(-1.0/G.vol[b] + 1.0/G.vol[lo]));
        a.W(VW_R, j, c) = -(wo + dt*a.t_wind[wi])*q + ti*(a.epe(j, c) - a.epe(j + 1, c))/dt
                          + 0.5/dt*(a.tpe(j, c) + a.tpe(j + 1, c))*dex - q*wo*rie/rio;
        if (j < n - 1) {
            const i64 l2 = sidx(G, j + 2, c);
            const double dz1 = G.zs[lo] - G.zs[l2], dex1 = G.exbg[lo] - G.exbg[l2];
            const double ti1 = tin(j + 1, c), rio1 = a.rio(j + 1, c);
            a.W(VW_C, j, c) = ti1*ga1*ti + 0.5*dex1*(al1 + be1*ti)
                              - dz1/(iw*dt*GC_C_D_P)*0.5*a.wind_o[widx(G, j + 2, c)]/(G.vol[lo]*rio1);
            a.W(VW_E, j, c) = ti*ga1*ti1 - 0.5*dex*(al1 + be1*ti1) + dz/(iw*dt*GC_C_D_P)*0.5*wo/(G.vol[lo]*rio);
        }
    }
};

/* the new state's rho and rt in layer j and their cotangents from the new
   state's rho, rt, tp, ep and w, formed once for the kernels after the solve */
struct VwLayerNew {
    VwArgs a;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= a.G.s) return;
        int j, c;
        a.split(t, j, c);
        double rn, rtn, bre, brte;
        a.rnew(j, c, rn, rtn);
        a.bnew(j, c, rn, rtn, bre, brte);
        a.W(VW_RN, j, c) = rn;
        a.W(VW_RTN, j, c) = rtn;
        a.W(VW_BRE, j, c) = bre;
        a.W(VW_BRTE, j, c) = brte;
    }
};

/* Thomas algorithm of column_solvers.c:641-667: A x = r (bands c below,
   d, e above) or, with transpose, A^T x = r (e below, c above); the
   quotients go to VW_EP */
struct VwSolve {
    VwArgs a;
    int transpose;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= a.G.nsh) return;
        const int c = (int)t, n = a.G.L - 1;
        const int lo_b = transpose ? VW_E : VW_C, up_b = transpose ? VW_C : VW_E;
        const int rhs = transpose ? VW_BX : VW_R, out = transpose ? VW_BR : VW_X;
        double ep = n > 1 ? a.W(up_b, 0, c)/a.W(VW_D, 0, c) : 0.0, rp = a.W(rhs, 0, c)/a.W(VW_D, 0, c);
        a.W(VW_EP, 0, c) = ep;
        a.W(out, 0, c) = rp;
        for (int j = 1; j < n; ++j) {
            const double lo = a.W(lo_b, j - 1, c);
            const double den = a.W(VW_D, j, c) - ep*lo;
            rp = (a.W(rhs, j, c) - rp*lo)/den;
            if (j < n - 1) ep = a.W(up_b, j, c)/den;
            a.W(VW_EP, j, c) = ep;
            a.W(out, j, c) = rp;
        }
        double x = rp;
        for (int j = n - 2; j >= 0; --j) {
            x = a.W(out, j, c) - a.W(VW_EP, j, c)*x;
            a.W(out, j, c) = x;
        }
    }
};

/* the cotangent of the solution x at interface j from the new state's
   rho, rt (through the fluxes sol and theta_i sol) and w */
struct VwIfaceSolBar {
    VwArgs a;
    GAME_HD void operator()(i64 t) const
    {
        const DevGrid &G = a.G;
        const int n = G.L - 1;
        if (t >= (i64)n*G.nsh) return;
        int j, c;
        a.split(t, j, c);
        const double dt = a.dt;
        const i64 b = sidx(G, j, c), lo = sidx(G, j + 1, c), wi = widx(G, j + 1, c);
        const double bre0 = a.W(VW_BRE, j, c), brte0 = a.W(VW_BRTE, j, c);
        const double bre1 = a.W(VW_BRE, j + 1, c), brte1 = a.W(VW_BRTE, j + 1, c);
        const double grt = brte0/G.vol[b] - brte1/G.vol[lo];
        const double grho = bre0/G.vol[b] - bre1/G.vol[lo];
        double bx = 2.0/(G.area[wi]*a.rio(j, c))*a.b_wind[wi];
        bx += dt*a.W(VW_TIN, j, c)*grt + dt*grho;
        a.W(VW_BX, j, c) = bx/a.damp(j, c);
    }
};

/* the cotangent interface j contributes to theta_i, rho_i and w_o of
   interface k = j + 1 through the bands of row j + 1 (c[j]) and row j
   (e[j]); zero for the last interface */
struct VwNext { double btin, brio, bwo; };

GAME_HD inline VwNext vw_next(const VwArgs &a, int j, int c)
{
    const DevGrid &G = a.G;
    VwNext r{0.0, 0.0, 0.0};
    if (j >= G.L - 2) return r;
    const double dt = a.dt, iw = a.iw;
    const int k = j + 1;
    const i64 b = sidx(G, j, c), lo = sidx(G, k, c), l2 = sidx(G, k + 1, c);
    const double x = a.W(VW_X, j, c), rb = a.W(VW_BR, j, c);
    const double bc = -a.W(VW_BR, k, c)*x, be_ = -rb*a.W(VW_X, k, c);
    const double tin = a.W(VW_TIN, j, c), ga1 = a.W(VW_GA, k, c), be1 = a.W(VW_BE, k, c);
    const double dex = G.exbg[b] - G.exbg[lo];
    const double dz1 = G.zs[lo] - G.zs[l2];
    const double rio1 = a.rio(k, c), wo1 = a.wind_o[widx(G, k + 1, c)];
    const double k1 = dz1/(iw*dt*GC_C_D_P)*0.5;
    r.btin = ga1*tin*bc + (tin*ga1 - 0.5*dex*be1)*be_;
    r.brio = k1*wo1/(G.vol[lo]*rio1*rio1)*bc;
    r.bwo = -k1/(G.vol[lo]*rio1)*bc;
    return r;
}

/* the cotangents of theta_i and rho_i at interface j, and of w_o and the w
   tendency there: the band cotangents -(rb x^T) and the right-hand side rb
   of row j, the parts rows j and j - 1 give interface j, and the new state's
   rho, rt and w */
struct VwIfaceBar {
    VwArgs a;
    GAME_HD void operator()(i64 t) const
    {
        const DevGrid &G = a.G;
        const int n = G.L - 1;
        if (t >= (i64)n*G.nsh) return;
        int j, c;
        a.split(t, j, c);
        const double dt = a.dt, iw = a.iw;
        const i64 b = sidx(G, j, c), lo = sidx(G, j + 1, c), wi = widx(G, j + 1, c);
        /* the new state: theta_i sol in the rt flux, w_new */
        const double rn0 = a.W(VW_RN, j, c), rn1 = a.W(VW_RN, j + 1, c);
        const double brte0 = a.W(VW_BRTE, j, c), brte1 = a.W(VW_BRTE, j + 1, c);
        const double sol = a.sol(j, c), rio = a.rio(j, c), wo = a.wind_o[wi], bw = a.b_wind[wi];
        const double grt = brte0/G.vol[b] - brte1/G.vol[lo];
        const double din = 0.5*(rn0 + rn1);
        const double wnew = (2.0*sol/G.area[wi] - din*wo)/rio;
        double btin = dt*sol*grt;
        double brio = -wnew/rio*bw;
        double bwo = -din/rio*bw;
        /* row j: r and d, then c[j] and e[j] */
        const double x = a.W(VW_X, j, c), rb = a.W(VW_BR, j, c);
        const double bd = -rb*x, br = rb;
        const double tin = a.W(VW_TIN, j, c);
        const double dz = G.zs[b] - G.zs[lo], dex = G.exbg[b] - G.exbg[lo];
        const double q = dz/(iw*(dt*dt)*GC_C_D_P);
        const double rie = 0.5*(a.re(j, c) + a.re(j + 1, c));
        const double ga0 = a.W(VW_GA, j, c), ga1 = a.W(VW_GA, j + 1, c);
        const double be0 = a.W(VW_BE, j, c), be1 = a.W(VW_BE, j + 1, c);
        bwo += -q*br - q*rie/rio*br;
        a.bt_wind[wi] += -dt*q*br;
        btin += (a.epe(j, c) - a.epe(j + 1, c))/dt*br;
        brio += q*wo*rie/(rio*rio)*br;
        const double kk = 2.0/G.area[wi] + dt*wo*0.5*(-1.0/G.vol[b] + 1.0/G.vol[lo]);
        btin += (-2.0*tin*(ga0 + ga1) + 0.5*dex*(be1 - be0))*bd;
        brio += dz/(iw*(dt*dt)*GC_C_D_P*rio*rio)*kk*bd;
        bwo += -dz/(iw*(dt*dt)*GC_C_D_P*rio)*dt*0.5*(-1.0/G.vol[b] + 1.0/G.vol[lo])*bd;
        if (j < n - 1) {
            const i64 l2 = sidx(G, j + 2, c);
            const double bc = -a.W(VW_BR, j + 1, c)*x;
            const double be_ = -rb*a.W(VW_X, j + 1, c);
            const double dex1 = G.exbg[lo] - G.exbg[l2];
            const double tin1 = a.W(VW_TIN, j + 1, c);
            btin += (tin1*ga1 + 0.5*dex1*be1)*bc;
            btin += ga1*tin1*be_;
            const double k0 = dz/(iw*dt*GC_C_D_P)*0.5;
            bwo += k0/(G.vol[lo]*rio)*be_;
            brio -= k0*wo/(G.vol[lo]*rio*rio)*be_;
        }
        if (j > 0) {
            const VwNext p = vw_next(a, j - 1, c);
            btin += p.btin;
            brio += p.brio;
            bwo += p.bwo;
        }
        a.W(VW_BTIN, j, c) = btin;
        a.W(VW_BRIO, j, c) = brio;
        a.bo_wind[wi] += bwo;
    }
};

/* the per-layer cotangents of layer j: alpha, beta, gamma, theta_e, pi_e
   from the rows of interfaces j - 1 and j, re and rte, and through them the
   old state, the new state as read and the tendency */
struct VwLayerBar {
    VwArgs a;
    GAME_HD void operator()(i64 t) const
    {
        const DevGrid &G = a.G;
        if (t >= G.s) return;
        int j, c;
        a.split(t, j, c);
        const int n = G.L - 1, rk = a.rk;
        const double dt = a.dt, iw = a.iw, pdw = 0.5;
        const i64 b = sidx(G, j, c);
        const double ro = a.rho_o[b], rto = a.rt_o[b], vol = G.vol[b], exbg = G.exbg[b];
        const double tr = a.t_rho[b], tt = a.t_rt[b];
        const double al = a.W(VW_AL, j, c), be = a.W(VW_BE, j, c), ga = a.W(VW_GA, j, c);
        const double rtn = a.W(VW_RTN, j, c), brte = a.W(VW_BRTE, j, c);
        double bre = a.W(VW_BRE, j, c);
        double bal = 0.0, bbe = 0.0, bga = vol*(rtn - rto)*a.b_ep[b], btpe = 0.0, bepe = 0.0;
        /* row j (the layer above interface j) */
        if (j < n) {
            const double tin = a.W(VW_TIN, j, c), br = a.W(VW_BR, j, c), bd = -br*a.W(VW_X, j, c);
            const double dex = exbg - G.exbg[sidx(G, j + 1, c)];
            bepe += tin/dt*br;
            btpe += 0.5/dt*dex*br;
            bga += -(tin*tin)*bd;
            bal -= 0.5*dex*bd;
            bbe -= 0.5*dex*tin*bd;
            /* rho_ie = mean of re: its cotangent at interface j */
            const i64 lo = sidx(G, j + 1, c);
            const double q = (G.zs[b] - G.zs[lo])/(iw*(dt*dt)*GC_C_D_P);
            bre += 0.5*(-q*a.wind_o[widx(G, j + 1, c)]/a.rio(j, c)*br);
        }
        /* row j - 1 (the layer below interface j - 1), and its bands c[j-1], e[j-1] */
        if (j > 0) {
            const int p = j - 1;
            const i64 bp = sidx(G, p, c);
            const double tin = a.W(VW_TIN, p, c), br = a.W(VW_BR, p, c), xp = a.W(VW_X, p, c), bd = -br*xp;
            const double dex = G.exbg[bp] - exbg;
            bepe -= tin/dt*br;
            btpe += 0.5/dt*dex*br;
            bga += -(tin*tin)*bd;
            bal += 0.5*dex*bd;
            bbe += 0.5*dex*tin*bd;
            if (j < n) {
                const double bc = -a.W(VW_BR, j, c)*xp, be_ = -br*a.W(VW_X, j, c);
                const double tin1 = a.W(VW_TIN, j, c);
                const double dex1 = exbg - G.exbg[sidx(G, j + 1, c)];
                bga += tin1*tin*bc + tin*tin1*be_;
                bal += 0.5*dex1*bc - 0.5*dex*be_;
                bbe += 0.5*dex1*tin*bc - 0.5*dex*tin1*be_;
            }
            const double q = (G.zs[bp] - G.zs[b])/(iw*(dt*dt)*GC_C_D_P);
            bre += 0.5*(-q*a.wind_o[widx(G, j, c)]/a.rio(p, c)*br);
        }
        double btr = 0.0, btt = 0.0, bro = 0.0, brto = 0.0, btpo = 0.0, bepo = 0.0;
        double brn = 0.0, brtn = 0.0, bepn = 0.0;
        /* ep_new's ep_o and rt_o terms */
        bepo += a.b_ep[b];
        brto += -vol*ga*a.b_ep[b];
        /* theta_e = tp_o + dt vol (al tr + be tt) */
        btpo += btpe;
        bal += dt*vol*tr*btpe;
        bbe += dt*vol*tt*btpe;
        btr += dt*vol*al*btpe;
        btt += dt*vol*be*btpe;
        /* pi_e = ep_o + dt vol ga tt */
        bepo += bepe;
        bga += dt*vol*tt*bepe;
        btt += dt*vol*ga*bepe;
        /* re = rho_o + dt tr, rte = rt_o + dt tt */
        bro += bre;
        btr += dt*bre;
        brto += brte;
        btt += dt*brte;
        /* rho_i = mean of rho_o (interfaces j-1 and j) */
        if (j < n) bro += 0.5*a.W(VW_BRIO, j, c);
        if (j > 0) bro += 0.5*a.W(VW_BRIO, j - 1, c);
        /* theta_i = mean of rt_n/rho_n */
        double btin_sum = 0.0;
        if (j < n) btin_sum += 0.5*a.W(VW_BTIN, j, c);
        if (j > 0) btin_sum += 0.5*a.W(VW_BTIN, j - 1, c);
        const double rnr = rk == 0 ? ro : a.rho_n[b], rtnr = rk == 0 ? rto : a.rt_n[b];
        brtn += btin_sum/rnr;
        brn += -btin_sum*rtnr/(rnr*rnr);
        /* alpha, beta, gamma */
        if (rk == 0) {
            /* al = -rto/ro^2/vol, be = 1/ro/vol, ga = R/(cv rto) (exbg + ep_o)/vol */
            brto += -bal/(ro*ro)/vol;
            bro += 2.0*rto/(ro*ro*ro)/vol*bal;
            bro += -bbe/(ro*ro)/vol;
            brto += -ga/rto*bga;
            bepo += GC_R_D/(GC_C_D_V*rto)/vol*bga;
        } else {
            const double rnn = a.rho_n[b], rtnn = a.rt_n[b];
            const double ho = (1.0 - pdw)/vol, hn = pdw/vol;
            brto += -ho*bal/(ro*ro);
            bro += ho*2.0*rto/(ro*ro*ro)*bal;
            brtn += -hn*bal/(rnn*rnn);
            brn += hn*2.0*rtnn/(rnn*rnn*rnn)*bal;
            bro += -ho*bbe/(ro*ro);
            brn += -hn*bbe/(rnn*rnn);
            const double go = GC_R_D/(GC_C_D_V*rto)*(exbg + a.ep_o[b]);
            const double gn = GC_R_D/(GC_C_D_V*rtnn)*(exbg + a.ep_n[b]);
            brto += -ho*go/rto*bga;
            bepo += ho*GC_R_D/(GC_C_D_V*rto)*bga;
            brtn += -hn*gn/rtnn*bga;
            bepn += hn*GC_R_D/(GC_C_D_V*rtnn)*bga;
        }
        a.bt_rho[b] += btr;
        a.bt_rt[b] += btt;
        a.bo_tp[b] += btpo;
        a.bo_ep[b] += bepo;
        if (rk == 0) {
            a.bo_rho[b] += bro + brn;
            a.bo_rt[b] += brto + brtn;
        } else {
            a.bo_rho[b] += bro;
            a.bo_rt[b] += brto;
            a.bn_rho[b] += brn;
            a.bn_rt[b] += brtn;
            a.bn_ep[b] += bepn;
        }
    }
};

/* ------------------------------------------------------------------ */
/* scalar_tendencies_expl.c:107-176 for the gas (dry): t_rho = -div(F),  */
/* t_rt = -div(theta_v F), F = mean(rho) u, theta_v = rt/rho             */
/* ------------------------------------------------------------------ */
/* cotangent of the vertical contravariant correction at (cell, level):
   it enters the box below the level with +area/vol and the box above with
   -area/vol (divergences.c:41-59) */
GAME_HD inline double vcc_cot_h(const DevGrid &G, const double *db, int c, int lev)
{
    const i64 s0 = (i64)lev*G.nsh + c, s1 = s0 - G.nsh;
    const double a = G.area[c + (i64)lev*G.vpl];
    return db[s0]/G.vol[s0]*a - db[s1]/G.vol[s1]*a;
}

/* the cotangents of the vertical contravariant corrections of both fluxes
   at (cell, terrain level L - n_oro + k): vc[k*NSH + c] for the rhotheta_v
   flux's tendency cotangent, vc[n_oro*NSH + k*NSH + c] for the mass flux's */
struct VccCot {
    DevGrid G;
    const double *b_trho, *b_trt;
    double *vc;
    GAME_HD void operator()(i64 t) const
    {
        const i64 n = (i64)G.n_oro*G.nsh;
        if (t >= n) return;
        const unsigned u = (unsigned)t, ns = (unsigned)G.nsh;
        const int k = (int)(u/ns), c = (int)(u - (unsigned)k*ns);
        const int lev = G.L - G.n_oro + k;
        vc[t] = vcc_cot_h(G, b_trt, c, lev);
        vc[n + t] = vcc_cot_h(G, b_trho, c, lev);
    }
};

/* (divv_h^T db) at edge e of layer l (divergences.c:14-64, averaging.c:41-95),
   with the terrain levels' correction cotangents vc of db (VccCot) */
GAME_HD inline double divv_adj_edge(const DevGrid &G, const AdjTables &A, const double *db, const double *vc, int l,
                                    int e, int fr, int to)
{
    const i64 vi = eidx(G, l, e);
    const i64 sf = sidx(G, l, fr), st = sidx(G, l, to);
    double r = G.area[vi]*(db[sf]/G.vol[sf] - db[st]/G.vol[st]);
    const int top = G.L - G.n_oro;
    if (!G.flat && l >= top) {
        const int cell[2] = {fr, to}, slot[2] = {A.slot_from[e], A.slot_to[e]};
        for (int p = 0; p < 2; ++p) {
            const int c = cell[p];
            const double k = -0.5*G.ipw[8*sidx(G, l, c) + slot[p]]*G.slope[vi];
            double v = vc[(i64)(l - top)*G.nsh + c];
            if (l + 1 <= G.L - 1) v += vc[(i64)(l + 1 - top)*G.nsh + c];
            r += k*v;
        }
    }
    return r;
}

/* per (layer, edge): the cotangents of F and of theta_v F; ub += mean(rho) Fb.
   The cell sums of the densities' cotangents take, per edge,
   P1 = 0.5 F Fthb (theta_v's) and P2 = 0.5 u Fb (rho's) */
struct ScalarTendAdjE {
    DevGrid G;
    AdjTables A;
    const double *rho, *rt, *wind;          /* the substep's densities and updated wind */
    const double *b_trho, *b_trt;           /* cotangents of the two tendencies */
    const double *vc;                       /* VccCot of the two (terrain grids) */
    double *P1, *P2;                        /* L x NVH, written */
    double *b_wind;                         /* accumulated at the edges */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= (i64)G.L*G.nvh) return;
        const unsigned u_ = (unsigned)t, nv = (unsigned)G.nvh;
        const int l = (int)(u_/nv), e = (int)(u_ - (unsigned)l*nv);
        const int fr = G.from[e], to = G.to[e];
        const i64 sf = sidx(G, l, fr), st = sidx(G, l, to), vi = eidx(G, l, e);
        const i64 nvc = (i64)G.n_oro*G.nsh;
        const double fthb = -divv_adj_edge(G, A, b_trt, vc, l, e, fr, to);
        const double rf = rho[sf], rt_ = rho[st];
        const double fb = -divv_adj_edge(G, A, b_trho, vc + nvc, l, e, fr, to)
                          + 0.5*(rt[sf]/rf + rt[st]/rt_)*fthb;
        const double u = wind[vi];
        const double F = 0.5*(rf + rt_)*u;
        P1[t] = 0.5*F*fthb;
        P2[t] = 0.5*u*fb;
        b_wind[vi] += 0.5*(rf + rt_)*fb;
    }
};

/* per (layer, cell): the densities' cotangents from F and theta_v */
struct ScalarTendAdjC {
    DevGrid G;
    const double *rho, *rt;
    const double *P1, *P2;
    double *b_rho, *b_rt;                   /* accumulated */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        const unsigned u_ = (unsigned)t, ns = (unsigned)G.nsh;
        const int l = (int)(u_/ns), c = (int)(u_ - (unsigned)l*ns);
        const int ne = c < GAME_N_PENT ? 5 : 6;
        double thvb = 0.0, rb = 0.0;
        /* a fixed six slots, so every load issues at once; a pentagon's
           sixth reads its first edge and adds nothing (same sums) */
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const i64 k = (i64)l*G.nvh + G.adj[6*c + (j < ne ? j : 0)];
            const double p1 = P1[k], p2 = P2[k];
            if (j < ne) {
                thvb += p1;
                rb += p2;
            }
        }
        const double r = rho[t];
        b_rho[t] += rb - thvb*rt[t]/(r*r);
        b_rt[t] += thvb/r;
    }
};

/* ------------------------------------------------------------------ */
/* scalar_times_vector (multiplications.c:35-117), F = s v: its two     */
/* cotangents                                                            */
/* ------------------------------------------------------------------ */
/* sb (per layer, cell) += factor * sum over the F's that read s of 0.5 v Fb;
   v = v1 (+ v2 when given) */
struct StvAdjS {
    DevGrid G;
    const double *v1, *v2, *Fb;
    double factor;
    int with_w;                             /* 1: scalar_times_vector (edges and interior w); 0: _h */
    double *sb;                             /* accumulated */
    GAME_HD double v(i64 i) const { return v2 ? v1[i] + v2[i] : v1[i]; }
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        int l, c;
        split_index(t, G.nsh, l, c);
        const int ne = c < GAME_N_PENT ? 5 : 6;
        double r = 0.0;
        /* a fixed six slots, so every load issues at once (v2's after v1's,
           outside the sum); a pentagon's sixth reads its first edge and adds
           nothing (same sum) */
        i64 vi[6];
        double a[6], f[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            vi[j] = eidx(G, l, G.adj[6*c + (j < ne ? j : 0)]);
            a[j] = v1[vi[j]];
            f[j] = Fb[vi[j]];
        }
        if (v2) {
#pragma unroll
            for (int j = 0; j < 6; ++j) a[j] = a[j] + v2[vi[j]];
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double term = 0.5*a[j]*f[j];
            if (j < ne) r += term;
        }
        if (with_w) {
            if (l >= 1) r += 0.5*v(widx(G, l, c))*Fb[widx(G, l, c)];
            if (l + 1 <= G.L - 1) r += 0.5*v(widx(G, l + 1, c))*Fb[widx(G, l + 1, c)];
        }
        sb[t] += factor*r;
    }
};

/* vb (V) (+)= factor * mean(s) Fb at the edges and interior w points, s = s1 (+ s2) */
struct StvAdjV {
    DevGrid G;
    const double *s1, *s2, *Fb;
    double factor;
    int with_w, accumulate;
    double *vb;
    GAME_HD double s(i64 i) const { return s2 ? s1[i] + s2[i] : s1[i]; }
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.v) return;
        int lev, h;
        split_index(t, G.vpl, lev, h);
        double r = 0.0;
        if (h >= G.nsh) {
            const int e = h - G.nsh;
            r = factor*0.5*(s(sidx(G, lev, G.from[e])) + s(sidx(G, lev, G.to[e])))*Fb[t];
        } else if (with_w && lev >= 1 && lev <= G.L - 1) {
            r = factor*0.5*(s(sidx(G, lev - 1, h)) + s(sidx(G, lev, h)))*Fb[t];
        }
        if (accumulate) vb[t] += r;
        else vb[t] = r;
    }
};

/* ------------------------------------------------------------------ */
/* vorticity_flux.c:16-187 in the PV, for a fixed mass flux F            */
/* ------------------------------------------------------------------ */
GAME_HD inline double vflux_vert_weight(const DevGrid &G, int lev)
{
    return (lev == 0 || lev == G.L) ? 1.0 : 0.5;
}

/* the rhombus PV rows: thread (chunk, m) walks the layers of its chunk,
   loading each of m's reverse entries once; at most VF_CH layers per chunk */
#define VF_CH 8
struct VFluxPvAdjRh {
    DevGrid G;
    PvReverse R;
    const double *F, *ob;                   /* mass flux (V), cotangent of the output (V) */
    double *pvb;                            /* curl, rhombus rows written */
    int chunks;                             /* ceil(L / VF_CH) or more */
    GAME_HD void operator()(i64 t) const
    {
        const int nvh = G.nvh, L = G.L;
        if (t >= (i64)chunks*nvh) return;
        int ch, m;
        split_index(t, nvh, ch, m);
        const int l0 = (int)(((i64)L*ch)/chunks), l1 = (int)(((i64)L*(ch + 1))/chunks);
        double acc[VF_CH];
#pragma unroll
        for (int q = 0; q < VF_CH; ++q) acc[q] = 0.0;
        const int n = R.n[m];
        /* four entries' gathers in flight at once; the sums keep entry order */
#pragma unroll 4
        for (int k = 0; k < n; ++k) {
            const i64 km = (i64)k*nvh + m;
            const double w = R.w[km];
            const int fe = R.f[km], oe = R.e[km];
#pragma unroll
            for (int q = 0; q < VF_CH; ++q)
                if (l0 + q < l1) acc[q] += w*F[eidx(G, l0 + q, fe)]*ob[eidx(G, l0 + q, oe)];
        }
#pragma unroll
        for (int q = 0; q < VF_CH; ++q)
            if (l0 + q < l1) pvb[(i64)(l0 + q)*2*nvh + nvh + m] = acc[q];
    }
};

/* the same rows with a fixed NE entries per rhombus (NE >= every count;
   the padding entries have weight 0 and read edge 0, valid indices whose
   products add a zero after the real terms): the entries stay in registers
   while the thread walks its chunk, and every gather of a layer is issued
   before the layer's sum, which keeps entry order */
template <int NE>
struct VFluxPvAdjRhN {
    DevGrid G;
    PvReverse R;
    const double *F, *ob;                   /* mass flux (V), cotangent of the output (V) */
    double *pvb;                            /* curl, rhombus rows written */
    int chunks;                             /* layer chunks per rhombus */
    GAME_HD void operator()(i64 t) const
    {
        const int nvh = G.nvh, L = G.L;
        if (t >= (i64)chunks*nvh) return;
        int ch, m;
        split_index(t, nvh, ch, m);
        const int l0 = (int)(((i64)L*ch)/chunks), l1 = (int)(((i64)L*(ch + 1))/chunks);
        double w[NE];
        int fe[NE], oe[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const i64 km = (i64)k*nvh + m;
            w[k] = R.w[km];
            fe[k] = R.f[km];
            oe[k] = R.e[km];
        }
        for (int l = l0; l < l1; ++l) {
            double fv[NE], ov[NE];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                fv[k] = F[eidx(G, l, fe[k])];
                ov[k] = ob[eidx(G, l, oe[k])];
            }
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < NE; ++k) acc += w[k]*fv[k]*ov[k];
            pvb[(i64)l*2*nvh + nvh + m] = acc;
        }
    }
};

/* the tangential (half-level) PV rows, per (level, edge) */
struct VFluxPvAdjTan {
    DevGrid G;
    AdjTables A;
    const double *F, *ob;                   /* mass flux (V), cotangent of the output (V) */
    double *pvb;                            /* curl, tangential rows written */
    GAME_HD void operator()(i64 t) const
    {
        const int nvh = G.nvh, L = G.L;
        if (t >= (i64)(L + 1)*nvh) return;
        int lev, e;
        split_index(t, nvh, lev, e);
        const int fr = G.from[e], to = G.to[e];
        double r = 0.0;
        /* the non-standard term of the edge in layer lev (ipw slot 6, w at
           level lev) and in layer lev - 1 (slot 7, w at level lev) */
        if (lev <= L - 1)
            r += -0.5*ob[eidx(G, lev, e)]*(G.ipw[8*sidx(G, lev, fr) + 6]*F[widx(G, lev, fr)]
                                           + G.ipw[8*sidx(G, lev, to) + 6]*F[widx(G, lev, to)]);
        if (lev >= 1)
            r += -0.5*ob[eidx(G, lev - 1, e)]*(G.ipw[8*sidx(G, lev - 1, fr) + 7]*F[widx(G, lev, fr)]
                                               + G.ipw[8*sidx(G, lev - 1, to) + 7]*F[widx(G, lev, to)]);
        /* the vertical outputs of the edge's two cells at level lev */
        const double vw = vflux_vert_weight(G, lev);
        const int cell[2] = {fr, to}, slot[2] = {A.slot_from[e], A.slot_to[e]};
        for (int p = 0; p < 2; ++p) {
            const int c = cell[p];
            double s = 0.0;
            if (lev >= 1) s += G.ipw[8*sidx(G, lev - 1, c) + slot[p]]*F[eidx(G, lev - 1, e)];
            if (lev <= L - 1) s += G.ipw[8*sidx(G, lev, c) + slot[p]]*F[eidx(G, lev, e)];
            r += vw*s*ob[widx(G, lev, c)];
        }
        pvb[(i64)lev*2*nvh + e] = r;
    }
};

/* ------------------------------------------------------------------ */
/* calc_pot_vort, vorticities.c:18-85: pv = (rel_vort + f)/density        */
/* ------------------------------------------------------------------ */
/* the density of curl element t (rhombus: four-cell weighted mean; half
   level: the edge mean above and below, extrapolated at top and surface) */
GAME_HD inline double pv_density(const DevGrid &G, const double *rho, int lev, int h)
{
    const int nvh = G.nvh, nsh = G.nsh, L = G.L;
    if (h >= nvh) {
        const int e = h - nvh;
        double dv = 0.0;
        for (int j = 0; j < 4; ++j) dv += G.drw[4*e + j]*rho[(i64)lev*nsh + G.dri[4*e + j]];
        return dv;
    }
    const int e = h, fr = G.from[e], to = G.to[e];
    if (lev == 0) {
        const double m0 = 0.5*(rho[fr] + rho[to]), m1 = 0.5*(rho[fr + nsh] + rho[to + nsh]);
        return m0 + (m0 - m1)/G.pv_top_dz1[e]*G.pv_top_dz2[e];
    }
    if (lev == L) {
        const i64 a = (i64)(L - 1)*nsh, b = (i64)(L - 2)*nsh;
        const double ma = 0.5*(rho[a + fr] + rho[a + to]), mb = 0.5*(rho[b + fr] + rho[b + to]);
        return ma + (mb - ma)/G.pv_sfc_dz1[e]*G.pv_sfc_dz2[e];
    }
    const i64 uf = (i64)(lev - 1)*nsh + fr, ut = (i64)(lev - 1)*nsh + to;
    return 0.25*(rho[uf] + rho[ut] + rho[uf + nsh] + rho[ut + nsh]);
}

struct PvDivAdj {
    DevGrid G;
    const double *pvb, *pv, *rho;
    double *rvb, *dvb;                      /* curl, written */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.curl) return;
        int lev, h;
        split_index(t, 2*G.nvh, lev, h);
        const double dv = pv_density(G, rho, lev, h);
        rvb[t] = pvb[t]/dv;
        dvb[t] = -pvb[t]*pv[t]/dv;
    }
};

/* rhombus rows: rel_vort = (A_f zeta_f + A_t zeta_t)/(A_f + A_t)
   (vorticities.c:120-130); per (layer, triangle) over its three edges */
struct RhombusVortAdj {
    DevGrid G;
    const double *rvb;                      /* curl */
    double *zb;                             /* L x NT, written */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= (i64)G.L*G.nt) return;
        int l, tri;
        split_index(t, G.nt, l, tri);
        const i64 base = G.nvh + (i64)l*G.dvpl;
        const double at = G.darea[base + tri];
        double r = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int e = G.vit[3*tri + k];
            r += at/(G.darea[base + G.dfrom[e]] + G.darea[base + G.dto[e]])*rvb[(i64)l*2*G.nvh + G.nvh + e];
        }
        zb[t] = r;
    }
};

/* tangential rows (vorticities.c:132-163): the cotangent of the horizontal
   covariant component of each edge; the top row copies level 1 */
GAME_HD inline double tan_row_cot(const DevGrid &G, const double *rvb, int lev, int e)
{
    double r = rvb[(i64)lev*2*G.nvh + e];
    if (lev == 1) r += rvb[e];
    return r;
}

struct TanVortAdjE {
    DevGrid G;
    const double *rvb;
    double *cvb;                            /* L x NVH, written: cotangent of the covariant component */
    double *b_wind;                         /* accumulated at the edges */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= (i64)G.L*G.nvh) return;
        int l, e;
        split_index(t, G.nvh, l, e);
        const i64 vi = eidx(G, l, e);
        double r = G.nd[vi]/G.darea[(i64)(l + 1)*G.dvpl + e]*tan_row_cot(G, rvb, l + 1, e);
        if (l >= 1) r += -G.nd[vi]/G.darea[(i64)l*G.dvpl + e]*tan_row_cot(G, rvb, l, e);
        cvb[t] = r;
        b_wind[vi] += r;
    }
};

/* per (level, cell): the w cotangents of the tangential rows and of the
   terrain part of the covariant components (averaging.c:14-39, 97-111) */
struct TanVortAdjC {
    DevGrid G;
    AdjTables A;
    const double *rvb, *cvb;
    double *b_wind;                         /* accumulated at the w points */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= (i64)(G.L + 1)*G.nsh) return;
        int lev, c;
        split_index(t, G.nsh, lev, c);
        const int ne = c < GAME_N_PENT ? 5 : 6, L = G.L, o0 = G.L - G.n_oro;
        const i64 wi = widx(G, lev, c);
        const bool rows = lev >= 1 && lev <= L - 1;
        const bool t6 = !G.flat && lev >= o0 && lev <= L - 1, t7 = !G.flat && lev - 1 >= o0 && lev - 1 < L - 1;
        const double ndw = rows ? G.nd[wi] : 0.0;
        const double w6 = t6 ? G.ipw[8*sidx(G, lev, c) + 6] : 0.0, w7 = t7 ? G.ipw[8*sidx(G, lev - 1, c) + 7] : 0.0;
        double r = 0.0;
        /* the cell's edges; the sign is +1 where the cell is the edge's from-cell */
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (j < ne) {
                const int e = G.adj[6*c + j];
                if (rows) {
                    const double s = (double)G.sgn[6*c + j];
                    r += s*ndw/G.darea[(i64)lev*G.dvpl + e]*tan_row_cot(G, rvb, lev, e);
                }
                if (t6) r += 0.5*G.slope[eidx(G, lev, e)]*w6*cvb[(i64)lev*G.nvh + e];
                if (t7) r += 0.5*G.slope[eidx(G, lev - 1, e)]*w7*cvb[(i64)(lev - 1)*G.nvh + e];
            }
        }
        b_wind[wi] += r;
    }
};

/* density cotangents of the half-level (tangential) rows, per (layer, cell) */
struct HalfDensAdj {
    DevGrid G;
    const double *dvb;
    double *b_rho;                          /* accumulated */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        int l, c;
        split_index(t, G.nsh, l, c);
        const int ne = c < GAME_N_PENT ? 5 : 6, L = G.L;
        const i64 nv2 = 2*(i64)G.nvh;
        double r = 0.0;
        /* a fixed six slots with the two interior loads unconditional (the
           rows l and l + 1 <= L exist), so they issue at once; a pentagon's
           sixth slot reads its first edge (same sum) */
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int e = G.adj[6*c + (j < ne ? j : 0)];
            const double da = dvb[(i64)l*nv2 + e], db = dvb[(i64)(l + 1)*nv2 + e];
            /* a pentagon's sixth slot adds -0.0, the exact identity of IEEE
               addition */
            if (l >= 1) r += j < ne ? 0.25*da : -0.0;
            if (l + 1 <= L - 1) r += j < ne ? 0.25*db : -0.0;
            if (j >= ne) continue;
            /* the extrapolations to the top and the surface */
            if (l <= 1) {
                const double kt = G.pv_top_dz2[e]/G.pv_top_dz1[e];
                if (l == 0) r += 0.5*(1.0 + kt)*dvb[e];
                if (l == 1) r += -0.5*kt*dvb[e];
            }
            if (l >= L - 2) {
                const double ks = G.pv_sfc_dz2[e]/G.pv_sfc_dz1[e];
                if (l == L - 1) r += 0.5*(1.0 - ks)*dvb[(i64)L*nv2 + e];
                if (l == L - 2) r += 0.5*ks*dvb[(i64)L*nv2 + e];
            }
        }
        b_rho[t] += r;
    }
};

/* density cotangents of the rhombus rows (vorticities.c:33-43), per (layer, cell) */
struct RhombusDensAdj {
    DevGrid G;
    AdjTables A;
    const double *dvb;
    double *b_rho;                          /* accumulated */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        int l, c;
        split_index(t, G.nsh, l, c);
        double r = 0.0;
        /* every entry's load issues at once: padding entries (weight 0)
           name edge 0, a valid index, and add -0.0, the exact identity of
           IEEE addition (same sum) */
#pragma unroll
        for (int k = 0; k < ADJ_RH_MAX; ++k) {
            const double w = A.rhp_w[(i64)k*G.nsh + c];
            const double d = dvb[(i64)l*2*G.nvh + G.nvh + A.rhp_e[(i64)k*G.nsh + c]];
            r += w != 0.0 ? w*d : -0.0;
        }
        b_rho[t] += r;
    }
};

/* ------------------------------------------------------------------ */
/* vector_tendencies_expl.c:74-130 and the wind update, manage_rkhevi.c:67-75 */
/* ------------------------------------------------------------------ */
/* second substep: tend1 = 0.5 tend0 + 0.5 (ohw pgrad_old - chw P + adv1) at
   edges, 0.5 tend0 + 0.5 (-cvw P + adv1) at interior w points; its cotangent
   tb1 (edges: dt times that of the updated wind) goes to tend0, pgrad_old, P
   and the advection term adv1 (plus the next step's cotangent of adv) */
struct TendAdjRk1 {
    DevGrid G;
    double ohw, chw, cvw;
    const double *tb1;                      /* cotangent of tend1 (V) */
    const double *advb_next;                /* cotangent of this step's adv from the next step */
    double *tb0, *poldb, *Pb;               /* accumulated */
    double *advb;                           /* written: cotangent of adv1 */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.v) return;
        int lev, h;
        split_index(t, G.vpl, lev, h);
        const bool edge = h >= G.nsh, inner_w = !edge && lev >= 1 && lev <= G.L - 1;
        if (!edge && !inner_w) { advb[t] = advb_next[t]; return; }
        const double b = tb1[t];
        tb0[t] += 0.5*b;
        if (edge) {
            poldb[t] += 0.5*ohw*b;
            Pb[t] += -0.5*chw*b;
        } else {
            Pb[t] += -0.5*cvw*b;
        }
        advb[t] = 0.5*b + advb_next[t];
    }
};

/* first substep: tend0 = ohw pgrad_old - chw P + adv_prev (edges),
   -cvw P + adv_prev (interior w) */
struct TendAdjRk0 {
    DevGrid G;
    double ohw, chw, cvw;
    const double *tb0;
    double *poldb, *Pb;                     /* accumulated */
    double *advb_prev;                      /* written: cotangent of the adv the step read */
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.v) return;
        int lev, h;
        split_index(t, G.vpl, lev, h);
        const bool edge = h >= G.nsh, inner_w = !edge && lev >= 1 && lev <= G.L - 1;
        if (!edge && !inner_w) { advb_prev[t] = 0.0; return; }
        const double b = tb0[t];
        if (edge) {
            poldb[t] += ohw*b;
            Pb[t] += -chw*b;
        } else {
            Pb[t] += -cvw*b;
        }
        advb_prev[t] = b;
    }
};

/* ------------------------------------------------------------------ */
/* element-wise pieces                                                   */
/* ------------------------------------------------------------------ */
/* y (+)= a*x over n elements */
struct Axpy {
    i64 n;
    double a;
    const double *x;
    double *y;
    int accumulate;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= n) return;
        y[t] = accumulate ? y[t] + a*x[t] : a*x[t];
    }
};

/* the wind update's edges: b_wind += ub (edges), tb (+)= dt*ub (edges) */
struct WindUpdateAdj {
    DevGrid G;
    double dt;
    const double *ub;
    double *b_wind, *tb;
    int accumulate_tb;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.v) return;
        int lev, h;
        split_index(t, G.vpl, lev, h);
        if (h < G.nsh) return;
        (void)lev;
        b_wind[t] += ub[t];
        tb[t] = accumulate_tb ? tb[t] + dt*ub[t] : dt*ub[t];
    }
};

/* theta_v_pert = rt/rho - theta_bg at a step boundary (the driver's copy
   re-derives it, linear_combine_two_states.c:26): its cotangent moves to
   rho and rt and is cleared */
struct ThetaBoundaryAdj {
    DevGrid G;
    const double *rho, *rt;
    double *b_rho, *b_rt, *b_tp;
    GAME_HD void operator()(i64 t) const
    {
        if (t >= G.s) return;
        const double r = rho[t], bt = b_tp[t];
        b_rho[t] -= bt*rt[t]/(r*r);
        b_rt[t] += bt/r;
        b_tp[t] = 0.0;
    }
};

}  // namespace sadj

#endif
