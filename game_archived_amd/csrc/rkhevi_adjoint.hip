/*
 * rkhevi_adjoint.hip -- hand-written adjoints of the linear operators of the
 * RKHEVI step (FP64), for gfx950.
 *
 * Every adjoint is a gather: each thread owns one output element of the
 * adjoint and sums the contributions of the forward outputs that read it,
 * through reverse index tables built once on the host (no atomics, so the
 * result does not depend on scheduling).  Each coefficient is recomputed
 * exactly as the forward kernel forms it, terrain-following terms included,
 * so <A x, y> = <x, A^T y> holds to rounding.
 *
 *   operator (forward, reference)                    adjoint gather
 *   grad (gradient_operators.c:56-74)                 cell <- adjacent edges, own column
 *   divv_h (divergences.c:14-64)                      edge <- from/to cells
 *   curl on triangles (vorticities.c:167-222)         edge <- its two triangles, layers l-1..l+1
 *   rhombus density (vorticities.c:74-80)             cell <- rhombi that weight it (CSR)
 *   TRSK vorticity flux in F (vorticity_flux.c)       edge <- edges whose TRSK stencil holds it (CSR)
 *                                                     w point <- adjacent edges and own levels
 *   inner product in a (inner_product.c:13-43)        edge <- from/to cells, w <- own column
 *   three-band solve (column_solvers.c:641-667)       transposed three-band solve
 */
#include "rkhevi_dev.h"
#include "rkhevi_adjoint.h"

#define ATPB 256
#define AGRID(n) dim3((unsigned)(((n) + ATPB - 1)/ATPB))
/* kernels that walk layers: one thread per (cell or edge, chunk of layers) */
#define ACHUNKS 4
#define AGRID2(n) dim3((unsigned)(((n) + ATPB - 1)/ATPB), ACHUNKS)

__device__ __forceinline__ void adj_chunk(int n, int &l0, int &l1)
{
    l0 = (int)(((i64)n*blockIdx.y)/gridDim.y);
    l1 = (int)(((i64)n*(blockIdx.y + 1))/gridDim.y);
}

/* ------------------------------------------------------------------ */
/* grad^T: fbar(c, l) from gbar at the cell's edges and w points        */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(ATPB) k_grad_adj(DevGrid G, AdjTables A, const double *gb, double *fb, double gs,
                                                   int acc)
{
    /* the input is gs*gb, formed as it is loaded; acc adds the result to fb */
    /* one thread per (cell, chunk of layers): layer l's edge cotangents feed
       the horizontal term of layer l and the terrain terms of the w points
       l (ipw slot 6) and l + 1 (slot 7); each w point's cotangent is formed
       once and carried to the next layer */
    const int h = blockIdx.x*ATPB + threadIdx.x;
    if (h >= G.nsh) return;
    const int nsh = G.nsh, vpl = G.vpl, L = G.L, top = G.flat ? L : L - G.n_oro;
    const int ne = h < GAME_N_PENT ? 5 : 6;
    int ej[6];
    double sg[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int jj = j < ne ? j : 0;
        ej[j] = G.adj[6*h + jj];
        sg[j] = j < ne ? (double)G.sgn[6*h + jj] : 0.0;
    }
    int l0, l1;
    adj_chunk(L, l0, l1);
    /* layer l: the horizontal sum and the two terrain sums */
    auto layer = [&](int l, double &hs, double &s6, double &s7) {
        double g[6], sl[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const i64 vi = nsh + (i64)l*vpl + ej[j];
            g[j] = gs*gb[vi];
            sl[j] = (l >= top) ? G.slope[vi] : 0.0;
            /* g_e = (f_to - f_from)/d_e: the from-cell (sign +1) gets -1/d_e */
            if (j == 0) hs = 0.0;
            hs += -sg[j]*g[j]/G.nd[vi];
        }
        s6 = 0.0; s7 = 0.0;
        if (l >= top) {
            const double w6 = G.ipw[8*((i64)l*nsh + h) + 6], w7 = G.ipw[8*((i64)l*nsh + h) + 7];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                if (j < ne) {
                    s6 += -sl[j]*0.5*w6*g[j];
                    s7 += -sl[j]*0.5*w7*g[j];
                }
            }
        }
    };
    /* the cotangent of the vertical gradient at w point lev over its normal distance */
    auto wcot = [&](int lev, double s6, double s7_above) {
        double r = gs*gb[h + (i64)lev*vpl] + s6 + s7_above;
        return r/G.nd[h + (i64)lev*vpl];
    };
    double hs, s6, s7, s7p = 0.0, unused_h, unused_s6;
    if (l0 >= 1) layer(l0 - 1, unused_h, unused_s6, s7p);
    layer(l0, hs, s6, s7);
    double wc = l0 >= 1 ? wcot(l0, s6, s7p) : 0.0;
    for (int l = l0; l < l1; ++l) {
        double hn = 0.0, s6n = 0.0, s7n = 0.0, wn = 0.0;
        if (l + 1 <= L - 1) {
            layer(l + 1, hn, s6n, s7n);
            wn = wcot(l + 1, s6n, s7);
        }
        const double v = hs + wn - wc;
        const i64 o = (i64)l*nsh + h;
        fb[o] = acc ? fb[o] + v : v;
        hs = hn; s7 = s7n; wc = wn;
    }
}

/* ------------------------------------------------------------------ */
/* divv_h and divv_h^T                                                   */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(ATPB) k_divv(DevGrid G, const double *in, double *out)
{
    i64 t = (i64)blockIdx.x*ATPB + threadIdx.x;
    if (t >= G.s) return;
    const int l = (int)(t/G.nsh), h = (int)(t - (i64)l*G.nsh);
    if (!G.flat) { out[t] = divv_h_cell(G, in, l, h); return; }
    /* flat surface: the terrain terms are zero; the six edges load in one batch
       (a pentagon's sixth slot reads its first edge and adds -0.0) */
    const int ne = h < GAME_N_PENT ? 5 : 6;
    double v[6], a[6], sg[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int jj = j < ne ? j : 0;
        const i64 vi = G.nsh + (i64)l*G.vpl + G.adj[6*h + jj];
        v[j] = in[vi]; a[j] = G.area[vi]; sg[j] = (double)G.sgn[6*h + jj];
    }
    double comp_h = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) comp_h += j < ne ? v[j]*sg[j]*a[j] : -0.0;
    out[t] = 1.0/G.vol[t]*comp_h;
}

/* cotangent of the vertical contravariant correction vcc(c, lev), lev in
   [L - n_oro, L - 1]: it enters layer lev with +area_w(lev) and layer lev - 1
   with -area_w(lev) (divergences.c:41-59) */
__device__ __forceinline__ double vcc_cot(const DevGrid &G, const double *db, int c, int lev)
{
    i64 s0 = (i64)lev*G.nsh + c, s1 = s0 - G.nsh;
    double a = G.area[c + (i64)lev*G.vpl];
    return db[s0]/G.vol[s0]*a - db[s1]/G.vol[s1]*a;
}

__global__ void __launch_bounds__(ATPB) k_divv_adj(DevGrid G, AdjTables A, const double *db, double *Fb)
{
    i64 t = (i64)blockIdx.x*ATPB + threadIdx.x;
    if (t >= G.v) return;
    const int nsh = G.nsh, vpl = G.vpl, L = G.L, no = G.n_oro;
    const int l = (int)(t/vpl), h = (int)(t - (i64)l*vpl);
    if (h < nsh) { Fb[t] = 0.0; return; }   /* divv_h reads horizontal components only */
    const int e = h - nsh;
    const int fr = G.from[e], to = G.to[e];
    const i64 sf = (i64)l*nsh + fr, st = (i64)l*nsh + to;
    double r = G.area[t]*(db[sf]/G.vol[sf] - db[st]/G.vol[st]);
    if (!G.flat && l >= L - no) {
        const int cell[2] = {fr, to}, slot[2] = {A.slot_from[e], A.slot_to[e]};
        for (int p = 0; p < 2; ++p) {
            const int c = cell[p];
            double k = -0.5*G.ipw[8*((i64)l*nsh + c) + slot[p]]*G.slope[t];
            double v = vcc_cot(G, db, c, l);
            if (l + 1 <= L - 1) v += vcc_cot(G, db, c, l + 1);
            r += k*v;
        }
    }
    Fb[t] = r;
}

/* ------------------------------------------------------------------ */
/* curl on triangles^T                                                  */
/* ------------------------------------------------------------------ */
/* one thread per (edge, chunk of layers) walking down the layers: the two
   triangles' contributions of layer l (to layer l itself and, in the terrain
   layers, to the layer its interpolation reads) are formed once and kept for
   the layers above and below; wind points at w levels get 0 (the forward
   reads none).  The sum for layer lp adds, per triangle, the terms of layers
   lp-1, lp, lp+1 in that order. */
struct CurlAdjTerm {
    double s[2], o[2];   /* base*self (to layer l) and base*other (to layer l -+ 1) */
    int d[2];            /* -1: the other term goes to l-1, +1: to l+1, 0: none */
};

__device__ __forceinline__ CurlAdjTerm curl_adj_term(const DevGrid &G, const double *zb, int e, const int *tri,
                                                     const double *sg, int l)
{
    const int nsh = G.nsh, vpl = G.vpl, L = G.L;
    const i64 vi = nsh + (i64)l*vpl + e;
    const i64 vu = l > 0 ? vi - vpl : vi, vd = l < L - 1 ? vi + vpl : vu;
    const double nd = G.nd[vi], zv = G.zv[vi], zu = G.zv[vu], zd = G.zv[vd];
    double z[2], da[2], dzv[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const i64 di = G.nvh + (i64)l*G.dvpl + tri[p];
        z[p] = zb[(i64)l*G.nt + tri[p]];
        da[p] = G.darea[di];
        dzv[p] = G.dzv[di];
    }
    CurlAdjTerm c;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double lrf = 1.0, self = 1.0, other = 0.0;
        int d = 0;
        if (l >= L - G.n_oro) {
            lrf = (G.radius + dzv[p])/(G.radius + zv);
            const double dz = dzv[p] - zv;
            const bool up = dz > 0 || l == L - 1;
            const double qq = dz/(zv - (up ? zu : zd));
            self = 1.0 + qq;
            other = -qq;
            d = up ? -1 : 1;
        }
        const double base = z[p]*lrf*nd*sg[p]/da[p];
        c.s[p] = base*self;
        c.o[p] = base*other;
        c.d[p] = d;
    }
    return c;
}

__global__ void __launch_bounds__(ATPB) k_curl_adj(DevGrid G, AdjTables A, const double *zb, double *wb, int acc)
{
    /* acc adds the edge cotangents to wb and leaves its w points; otherwise
       the w points are set to zero */
    const int t = blockIdx.x*ATPB + threadIdx.x;
    const int nsh = G.nsh, vpl = G.vpl, L = G.L;
    int l0, l1;
    adj_chunk(L, l0, l1);
    if (!acc && t < nsh) {
        for (int l = l0; l < l1; ++l) wb[(i64)l*vpl + t] = 0.0;
        if (l1 == L) wb[(i64)L*vpl + t] = 0.0;
    }
    if (t >= G.nvh) return;
    const int e = t;
    int tri[2];
    double sg[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        tri[p] = A.tri_of_edge[2*e + p];
        sg[p] = (double)G.vst[3*tri[p] + A.tri_slot[2*e + p]];
    }
    CurlAdjTerm prev = {}, cur = curl_adj_term(G, zb, e, tri, sg, l0), nxt = {};
    if (l0 >= 1) prev = curl_adj_term(G, zb, e, tri, sg, l0 - 1);
    if (l0 + 1 <= L - 1) nxt = curl_adj_term(G, zb, e, tri, sg, l0 + 1);
    for (int lp = l0; lp < l1; ++lp) {
        double r = 0.0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            if (lp >= 1 && prev.d[p] == 1) r += prev.o[p];
            r += cur.s[p];
            if (lp + 1 <= L - 1 && nxt.d[p] == -1) r += nxt.o[p];
        }
        const i64 o = nsh + (i64)lp*vpl + e;
        wb[o] = acc ? wb[o] + r : r;
        prev = cur;
        cur = nxt;
        if (lp + 2 <= L - 1 && lp + 1 < l1) nxt = curl_adj_term(G, zb, e, tri, sg, lp + 2);
    }
}

/* ------------------------------------------------------------------ */
/* rhombus density and its adjoint                                      */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(ATPB) k_rhombus(DevGrid G, const double *rho, double *out)
{
    i64 t = (i64)blockIdx.x*ATPB + threadIdx.x;
    if (t >= (i64)G.L*G.nvh) return;
    const int l = (int)(t/G.nvh), e = (int)(t - (i64)l*G.nvh);
    double dr = 0.0;
    for (int j = 0; j < 4; ++j) dr += G.drw[4*e + j]*rho[(i64)l*G.nsh + G.dri[4*e + j]];
    out[t] = dr;
}

__global__ void __launch_bounds__(ATPB) k_rhombus_adj(DevGrid G, AdjTables A, const double *db, double *rb)
{
    /* one thread per (cell, chunk of layers): the cell's rhombus entries are
       loaded once, in ascending (edge, slot) order, padding weighted 0 */
    const int c = blockIdx.x*ATPB + threadIdx.x;
    if (c >= G.nsh) return;
    const int nsh = G.nsh, nvh = G.nvh;
    int e[ADJ_RH_MAX];
    double w[ADJ_RH_MAX];
#pragma unroll
    for (int k = 0; k < ADJ_RH_MAX; ++k) { e[k] = A.rhp_e[(i64)k*nsh + c]; w[k] = A.rhp_w[(i64)k*nsh + c]; }
    int l0, l1;
    adj_chunk(G.L, l0, l1);
    for (int l = l0; l < l1; ++l) {
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < ADJ_RH_MAX; ++k) r += w[k]*db[(i64)l*nvh + e[k]];
        rb[(i64)l*nsh + c] = r;
    }
}

/* ------------------------------------------------------------------ */
/* TRSK vorticity flux^T in the mass flux F (pv fixed)                   */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ double vflux_w_weight(const DevGrid &G, int lev)
{   /* vorticity_flux.c:373-417: half weights inside, full at top and surface */
    return (lev == 0 || lev == G.L) ? 1.0 : 0.5;
}

__global__ void __launch_bounds__(ATPB) k_vflux_adj_e(DevGrid G, AdjTables A, const double *pv, const double *ob,
                                                      double *Fb)
{
    /* horizontal F of edge e' at layer l, one thread per (edge, chunk of
       layers): its TRSK reverse entries are loaded once */
    const int ep = blockIdx.x*ATPB + threadIdx.x;
    if (ep >= G.nvh) return;
    const int nsh = G.nsh, nvh = G.nvh, vpl = G.vpl;
    int e[ADJ_TR_MAX], m[ADJ_TR_MAX];
    bool hf[ADJ_TR_MAX];
    double w[ADJ_TR_MAX];
#pragma unroll
    for (int k = 0; k < ADJ_TR_MAX; ++k) {
        e[k] = A.trp_e[(i64)k*nvh + ep];
        const int mm = A.trp_m[(i64)k*nvh + ep];
        hf[k] = mm < 0;
        m[k] = mm < 0 ? -mm - 1 : mm;
        w[k] = A.trp_w[(i64)k*nvh + ep];
    }
    const int cell[2] = {G.from[ep], G.to[ep]}, slot[2] = {A.slot_from[ep], A.slot_to[ep]};
    int l0, l1;
    adj_chunk(G.L, l0, l1);
    for (int l = l0; l < l1; ++l) {
        const i64 pb = nvh + (i64)l*2*nvh;
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < ADJ_TR_MAX; ++k) {
            const double pm = pv[pb + m[k]], pe = pv[pb + e[k]];
            const double q = hf[k] ? 0.5*(pm + pe) : pm;
            r += w[k]*q*ob[nsh + (i64)l*vpl + e[k]];
        }
        /* the vertical outputs of the edge's two cells at levels l and l + 1 */
        const double q1 = pv[(i64)(l + 1)*2*nvh + ep], q0 = pv[(i64)l*2*nvh + ep];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int c = cell[p];
            const double wt = G.ipwT[((i64)l*8 + slot[p])*nsh + c];
            r += vflux_w_weight(G, l + 1)*wt*q1*ob[(i64)(l + 1)*vpl + c];
            r += vflux_w_weight(G, l)*wt*q0*ob[(i64)l*vpl + c];
        }
        Fb[nsh + (i64)l*vpl + ep] = r;
    }
}

__global__ void __launch_bounds__(ATPB) k_vflux_adj_w(DevGrid G, const double *pv, const double *ob, double *Fb)
{
    /* vertical F of cell h at level l: the w terms of the cell's edges at
       layers l (ipw slot 6) and l - 1 (slot 7) (vorticity_flux.c:110-187);
       all loads first, layer indices clamped, out-of-range terms masked */
    const i64 t = (i64)blockIdx.x*ATPB + threadIdx.x;
    const int nsh = G.nsh, nvh = G.nvh, vpl = G.vpl, L = G.L;
    if (t >= (i64)(L + 1)*nsh) return;
    const int l = (int)(t/nsh), h = (int)(t - (i64)l*nsh);
    const int ne = h < GAME_N_PENT ? 5 : 6;
    const int la = l <= L - 1 ? l : L - 1, lb = l >= 1 ? l - 1 : 0;
    const double w6 = G.ipwT[((i64)la*8 + 6)*nsh + h], w7 = G.ipwT[((i64)lb*8 + 7)*nsh + h];
    double q[6], oa[6], obb[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int e = G.adj[6*h + (j < ne ? j : 0)];
        q[j] = pv[e + (i64)l*2*nvh];
        oa[j] = ob[nsh + (i64)la*vpl + e];
        obb[j] = ob[nsh + (i64)lb*vpl + e];
    }
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j >= ne) continue;
        if (l <= L - 1) r += -0.5*w6*q[j]*oa[j];
        if (l >= 1) r += -0.5*w7*q[j]*obb[j];
    }
    Fb[(i64)l*vpl + h] = r;
}

/* ------------------------------------------------------------------ */
/* inner product^T in its first argument                                 */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(ATPB) k_ip_adj(DevGrid G, AdjTables A, const double *b, const double *ib,
                                                 double *ab, double scale, int acc)
{
    /* acc: ab += scale*result; otherwise ab = result */
    i64 t = (i64)blockIdx.x*ATPB + threadIdx.x;
    if (t >= G.v) return;
    const int nsh = G.nsh, vpl = G.vpl, L = G.L;
    const int l = (int)(t/vpl), h = (int)(t - (i64)l*vpl);
    double r = 0.0;
    if (h >= nsh) {
        const int e = h - nsh;
        const int fr = G.from[e], to = G.to[e];
        const i64 sf = (i64)l*nsh + fr, st = (i64)l*nsh + to;
        r = b[t]*(G.ipw[8*sf + A.slot_from[e]]*ib[sf] + G.ipw[8*st + A.slot_to[e]]*ib[st]);
    } else {
        double s = 0.0;
        if (l >= 1) s += G.ipw[8*((i64)(l - 1)*nsh + h) + 7]*ib[(i64)(l - 1)*nsh + h];
        if (l <= L - 1) s += G.ipw[8*((i64)l*nsh + h) + 6]*ib[(i64)l*nsh + h];
        r = b[t]*s;
    }
    ab[t] = acc ? ab[t] + scale*r : r;
}

/* ------------------------------------------------------------------ */
/* batched three-band solve and its adjoint (one column per lane;        */
/* element j of column k at j*ncol + k, so lanes read consecutive words) */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(ATPB) k_thomas(int n, int ncol, const double *c, const double *d, const double *e,
                                                 const double *r, double *x, double *ep)
{
    const int k = blockIdx.x*ATPB + threadIdx.x;
    if (k >= ncol) return;
    /* column_solvers.c:641-667 */
    double epj = (n > 1) ? e[k]/d[k] : 0.0, xj = r[k]/d[k];
    ep[k] = epj;
    x[k] = xj;
    for (int j = 1; j < n; ++j) {
        const i64 o = (i64)j*ncol + k, om = o - ncol;
        const double den = d[o] - epj*c[om];
        if (j < n - 1) epj = e[o]/den;
        xj = (r[o] - xj*c[om])/den;
        ep[o] = epj;
        x[o] = xj;
    }
    for (int j = n - 2; j >= 0; --j) {
        const i64 o = (i64)j*ncol + k;
        xj = x[o] - ep[o]*xj;
        x[o] = xj;
    }
}

/* A^T rbar = xbar (A^T has sub-diagonal e and super-diagonal c), then the
   coefficient cotangents -rbar x^T on the three bands */
__global__ void __launch_bounds__(ATPB) k_thomas_adj(int n, int ncol, const double *c, const double *d,
                                                     const double *e, const double *x, const double *xb,
                                                     double *rb, double *cb, double *db, double *eb, double *ep)
{
    const int k = blockIdx.x*ATPB + threadIdx.x;
    if (k >= ncol) return;
    double epj = (n > 1) ? c[k]/d[k] : 0.0, yj = xb[k]/d[k];
    ep[k] = epj;
    rb[k] = yj;
    for (int j = 1; j < n; ++j) {
        const i64 o = (i64)j*ncol + k, om = o - ncol;
        const double den = d[o] - epj*e[om];
        if (j < n - 1) epj = c[o]/den;
        yj = (xb[o] - yj*e[om])/den;
        ep[o] = epj;
        rb[o] = yj;
    }
    for (int j = n - 2; j >= 0; --j) {
        const i64 o = (i64)j*ncol + k;
        yj = rb[o] - ep[o]*yj;
        rb[o] = yj;
    }
    for (int j = 0; j < n; ++j) {
        const i64 o = (i64)j*ncol + k;
        db[o] = -rb[o]*x[o];
        if (j < n - 1) {
            cb[o] = -rb[o + ncol]*x[o];
            eb[o] = -rb[o]*x[o + ncol];
        }
    }
}

/* ------------------------------------------------------------------ */
namespace adj {

void grad_adjoint(const DevGrid &G, const AdjTables &A, const double *gb, double *fb, hipStream_t st, double gs,
                  int acc)
{ LAUNCH("k_grad_adj", k_grad_adj, AGRID2(G.nsh), dim3(ATPB), st, G, A, gb, fb, gs, acc); }

void divv_h(const DevGrid &G, const double *in, double *out, hipStream_t st)
{ LAUNCH("k_divv", k_divv, AGRID(G.s), dim3(ATPB), st, G, in, out); }

void divv_h_adjoint(const DevGrid &G, const AdjTables &A, const double *db, double *Fb, hipStream_t st)
{ LAUNCH("k_divv_adj", k_divv_adj, AGRID(G.v), dim3(ATPB), st, G, A, db, Fb); }

void curl_triangles_adjoint(const DevGrid &G, const AdjTables &A, const double *zb, double *wb, hipStream_t st,
                            int acc)
{ LAUNCH("k_curl_adj", k_curl_adj, AGRID2(G.nvh > G.nsh ? G.nvh : G.nsh), dim3(ATPB), st, G, A, zb, wb, acc); }

void rhombus_density(const DevGrid &G, const double *rho, double *out, hipStream_t st)
{ LAUNCH("k_rhombus", k_rhombus, AGRID((i64)G.L*G.nvh), dim3(ATPB), st, G, rho, out); }

void rhombus_density_adjoint(const DevGrid &G, const AdjTables &A, const double *db, double *rb, hipStream_t st)
{ LAUNCH("k_rhombus_adj", k_rhombus_adj, AGRID2(G.nsh), dim3(ATPB), st, G, A, db, rb); }

void vorticity_flux_adjoint(const DevGrid &G, const AdjTables &A, const double *pv, const double *ob, double *Fb,
                            hipStream_t st)
{
    LAUNCH("k_vflux_adj_e", k_vflux_adj_e, AGRID2(G.nvh), dim3(ATPB), st, G, A, pv, ob, Fb);
    LAUNCH("k_vflux_adj_w", k_vflux_adj_w, AGRID((i64)(G.L + 1)*G.nsh), dim3(ATPB), st, G, pv, ob, Fb);
}

void inner_product_adjoint(const DevGrid &G, const AdjTables &A, const double *b, const double *ib, double *ab,
                           hipStream_t st, double scale, int acc)
{ LAUNCH("k_ip_adj", k_ip_adj, AGRID(G.v), dim3(ATPB), st, G, A, b, ib, ab, scale, acc); }

void thomas(int n, int ncol, const double *c, const double *d, const double *e, const double *r, double *x,
            double *work, hipStream_t st)
{ LAUNCH("k_thomas", k_thomas, AGRID(ncol), dim3(ATPB), st, n, ncol, c, d, e, r, x, work); }

void thomas_adjoint(int n, int ncol, const double *c, const double *d, const double *e, const double *x,
                    const double *xb, double *rb, double *cb, double *db, double *eb, double *work, hipStream_t st)
{ LAUNCH("k_thomas_adj", k_thomas_adj, AGRID(ncol), dim3(ATPB), st, n, ncol, c, d, e, x, xb, rb, cb, db, eb, work); }

}  // namespace adj

/* ------------------------------------------------------------------ */
/* scratch of the allocation-free three-band solve                     */
/* ------------------------------------------------------------------ */
/* The elimination quotients of game_gpu_op_thomas_noalloc live here instead
   of in a caller's work buffer: THOMAS_SCRATCH_DOUBLES doubles (64 MiB) that
   every GPU the library runs on holds from module load to process end.  In
   host code `scratch` names a host-side stand-in, never the device memory:
   only the address hipGetSymbolAddress returns for the current device (kept
   per device below) may be passed to a kernel. */
static __device__ __align__(256) double scratch[THOMAS_SCRATCH_DOUBLES];

namespace {
const int SCRATCH_DEVICES = 64;
double *g_scratch_ptr[SCRATCH_DEVICES] = {nullptr};
}

namespace adj {

int thomas_scratch_resolve(char *err, size_t nerr)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        snprintf(err, nerr, "hipGetDevice failed: %s", hipGetErrorString(e));
        return -1;
    }
    if (dev < 0 || dev >= SCRATCH_DEVICES) {
        snprintf(err, nerr, "device %d is beyond the %d devices the solver scratch supports", dev, SCRATCH_DEVICES);
        return -1;
    }
    if (g_scratch_ptr[dev]) return 0;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        snprintf(err, nerr, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        return -1;
    }
    g_scratch_ptr[dev] = (double *)p;
    return 0;
}

double *thomas_scratch(void)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_DEVICES) return nullptr;
    return g_scratch_ptr[dev];
}

}  // namespace adj
