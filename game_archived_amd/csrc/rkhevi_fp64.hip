/*
 * rkhevi_fp64.hip -- FP64 device kernels of GAME's RKHEVI step for gfx950.
 *
 * Each kernel evaluates one or more loops of the reference with the same
 * floating-point operations in the same order (this file is compiled with
 * -ffp-contract=off), so that the FP64 path reproduces the CPU reference to
 * the last bit wherever no transcendental function is involved.  Fusion only
 * ever recomputes a value with identical arithmetic; no reduction order is
 * changed.  All kernels are HBM-bandwidth bound gathers/stencils; there is no
 * dense contraction, hence no MFMA.
 */
#include "rkhevi_dev.h"
#include "rkhevi_kernels.h"

#define TPB 256
#define GRID1(n) dim3((unsigned)(((n) + TPB - 1)/TPB))
#define GRID2(n, chunks) dim3((unsigned)(((n) + TPB - 1)/TPB), (unsigned)(chunks))

/* Workgroups are dispatched to the eight XCDs round-robin (workgroup i on
   XCD i % 8), each XCD with its own L2.  The gather kernels renumber their
   workgroups so that each XCD takes one contiguous range of the grid: the
   neighbour data a workgroup shares with the next one then sits in the same
   L2 instead of being fetched by several XCDs. */
#define N_XCD 8u
struct Blk { unsigned x, y; };
__device__ __forceinline__ Blk xcd_block()
{
    const unsigned T = gridDim.x*gridDim.y, lin = blockIdx.y*gridDim.x + blockIdx.x;
    const unsigned q = T/N_XCD, r = T%N_XCD, x = lin%N_XCD, k = lin/N_XCD;
    const unsigned m = (x < r) ? x*(q + 1) + k : r*(q + 1) + (x - r)*q + k;
    Blk b;
    b.x = m%gridDim.x;
    b.y = m/gridDim.x;
    return b;
}

/* layers [l0, l1) of chunk `chunk` when n layers are split into gridDim.y chunks */
__device__ __forceinline__ void layer_chunk(int n, unsigned chunk, int &l0, int &l1)
{
    l0 = (int)(((i64)n*chunk)/gridDim.y);
    l1 = (int)(((i64)n*(chunk + 1))/gridDim.y);
}

/* ------------------------------------------------------------------ */
/* small device helpers restating reference functions                   */
/* ------------------------------------------------------------------ */

/* raw vertical covariant gradient at an interior w point (gradient_operators.c:35-54) */
template <typename TF>
__device__ __forceinline__ double grad_w_raw(const DevGrid &G, const TF *f, int lev, int h)
{
    i64 vi = h + (i64)lev*G.vpl;
    return ((double)f[h + (i64)(lev - 1)*G.nsh] - (double)f[h + (i64)lev*G.nsh])/G.nd[vi];
}

/* averaging.c:14-39 on a stored vector field */
__device__ __forceinline__ double remap_verpri2horpri(const DevGrid &G, const double *f, int l, int e)
{
    int fr = G.from[e], to = G.to[e];
    const double *ipw = G.ipw;
    double c = ipw[8*((i64)l*G.nsh + fr) + 6]*f[(i64)l*G.vpl + fr];
    c += ipw[8*((i64)l*G.nsh + to) + 6]*f[(i64)l*G.vpl + to];
    if (l < G.L - 1) {
        c += ipw[8*((i64)l*G.nsh + fr) + 7]*f[(i64)(l + 1)*G.vpl + fr];
        c += ipw[8*((i64)l*G.nsh + to) + 7]*f[(i64)(l + 1)*G.vpl + to];
    }
    return 0.5*c;
}

/* ------------------------------------------------------------------ */
/* kernels                                                              */
/* ------------------------------------------------------------------ */

/* derived_quantities.c:15-40 */
template <typename TO>
__global__ void __launch_bounds__(TPB) k_temperature(DevGrid G, DevState s, TO *T, TO *cpc, double *gf)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= G.s) return;
    /* the gas fraction of the box, as the pressure gradient forms it
       (manage_pressure_gradient.c:17-87) */
    if (gf) gf[i] = s.rho[(i64)G.n_cond*G.s + i]/density_total(G, s.rho, i);
    double t = (G.thbg[i] + s.tp[i])*(G.exbg[i] + s.ep[i]);
    if (G.moist)
        t = t/(1.0 + s.rho[(i64)(G.n_cond + 1)*G.s + i]/s.rho[(i64)G.n_cond*G.s + i]*(GC_M_D/GC_M_V - 1.0));
    T[i] = (TO)t;
    if (cpc) {
        const double tt = (double)(TO)t;
        for (int k = 0; k < G.n_cond; ++k) cpc[(i64)k*G.s + i] = (TO)c_p_cond(k, tt);   /* dictionary.c:111-190 */
    }
}

/* manage_pressure_gradient.c:17-87 at the interior w points (layers 1..L-1),
   fused with calc_pressure_grad_condensates_v */
template <typename TW, int NC>
__global__ void __launch_bounds__(TPB) k_pgrad_w(DevGrid G, DevState s, DevWorkT<TW> W, int first_step, int with_pgrad,
                                                 const double *gf, double cvw)
{
    i64 n = (i64)(G.L - 1)*G.nsh;
    i64 t = (i64)xcd_block().x*TPB + threadIdx.x;
    if (t >= n) return;
    int lev = (int)(t/G.nsh) + 1, h = (int)(t - (i64)(lev - 1)*G.nsh);
    i64 w = h + (i64)lev*G.vpl, up = h + (i64)(lev - 1)*G.nsh, lo = h + (i64)lev*G.nsh;
    i64 gi = (i64)G.n_cond*G.s;
    /* gf: the gas fractions of state s, formed with this arithmetic by k_temperature */
    double rt_up = 0.0, rt_lo = 0.0, d_up, d_lo;
    if (gf) { d_up = gf[up]; d_lo = gf[lo]; }
    else {
        rt_up = density_total_n<NC>(G, s.rho, up);
        rt_lo = density_total_n<NC>(G, s.rho, lo);
        d_up = s.rho[gi + up]/rt_up;
        d_lo = s.rho[gi + lo]/rt_lo;
    }
    if (with_pgrad) {
        if (!first_step) W.pgrad_old[w] = (TW)(-(double)W.pgrad_nl[w] - (double)W.pgrad_l[w]);
        double raw = (s.ep[up] - s.ep[lo])/G.nd[w];
        double p1_up = GC_C_D_P*(G.thbg[up] + s.tp[up]), p1_lo = GC_C_D_P*(G.thbg[lo] + s.tp[lo]);
        double nl = (0.5*(p1_up + p1_lo))*raw;
        double p2_up = GC_C_D_P*s.tp[up], p2_lo = GC_C_D_P*s.tp[lo];
        double ll = (0.5*(p2_up + p2_lo))*G.exbg_grad[w];
        nl = (0.5*(d_up + d_lo))*nl;
        ll = (0.5*(d_up + d_lo))*ll;
        W.pgrad_nl[w] = (TW)nl;
        W.pgrad_l[w] = (TW)ll;
        if (first_step) W.pgrad_old[w] = (TW)(-nl - ll);
        /* the momentum tendency's combination of the stored values
           (vector_tendencies_expl.c) */
        W.pgrad_comb[w] = (TW)(-cvw*((double)(TW)nl + (double)(TW)ll));
    }
    double c_up = d_up - 1.0, c_lo = d_lo - 1.0;   /* s.rho[gi + up]/rt_up - 1.0 */
    W.pgc_v[w] = (TW)((0.5*(c_up + c_lo))*G.gm[w]);
}

/* manage_pressure_gradient.c:17-72 at the edges: one thread per edge walks
   all layers, so the edge's table entries are read once */
template <bool FLAT, typename TF>
__device__ __forceinline__ double grad_edge_t(const DevGrid &G, const TF *f, int l, int e, int fr, int to)
{
    i64 vi = G.nsh + (i64)l*G.vpl + e;
    double g = ((double)f[to + (i64)l*G.nsh] - (double)f[fr + (i64)l*G.nsh])/G.nd[vi];
    if (!FLAT && l >= G.L - G.n_oro) {
        const double *ipw = G.ipw;
        double c = ipw[8*((i64)l*G.nsh + fr) + 6]*grad_w_raw(G, f, l, fr);
        c += ipw[8*((i64)l*G.nsh + to) + 6]*grad_w_raw(G, f, l, to);
        if (l < G.L - 1) {
            c += ipw[8*((i64)l*G.nsh + fr) + 7]*grad_w_raw(G, f, l + 1, fr);
            c += ipw[8*((i64)l*G.nsh + to) + 7]*grad_w_raw(G, f, l + 1, to);
        }
        c = 0.5*c;
        g += -G.slope[vi]*c;
    }
    return g;
}

/* grad at the edges (gradient_operators.c:14-33 with the terrain term of
   averaging.c:113-133) for the successive layers l0, l0 + 1, ... of one
   edge, as manage_pressure_gradient.c:17-72 and the kinetic-energy gradient
   use it.  The two cells' values of layer l + 1 and their vertical
   gradients at the w level between layers l and l + 1, which layer l's
   terrain correction reads, are carried to layer l + 1 instead of loaded
   and formed again; each quotient is the one the layer-wise formula forms,
   so the results are identical.  The next layer's values are loaded before
   this layer's are combined.  Used where terrain terms exist: on a flat
   grid the layer-wise form needs fewer registers (the flat pressure
   gradient keeps five waves per SIMD). */
struct EdgeCarry { double ff, ft, gwf, gwt; };

template <bool FLAT, typename TF>
__device__ __forceinline__ double grad_edge_walk(const DevGrid &G, const TF *f, int l, int l0, int e, int fr, int to,
                                                 EdgeCarry &c)
{
    const int nsh = G.nsh, L = G.L, lt = L - G.n_oro;
    const i64 vi = nsh + (i64)l*G.vpl + e;
    if (l == l0) { c.ff = (double)f[fr + (i64)l*nsh]; c.ft = (double)f[to + (i64)l*nsh]; }
    /* layer l + 1 (the last layer reloads itself, unused) */
    const int ln = l + 1 < L ? l + 1 : l;
    const double nff = (double)f[fr + (i64)ln*nsh], nft = (double)f[to + (i64)ln*nsh];
    double g = (c.ft - c.ff)/G.nd[vi];
    if (!FLAT && l >= lt) {
        /* grad_w_raw at level l of both cells: carried unless layer l - 1
           had no terrain term or lies in another chunk */
        if (l == l0 || l == lt) { c.gwf = grad_w_raw(G, f, l, fr); c.gwt = grad_w_raw(G, f, l, to); }
        const double *ipw = G.ipw;
        double s = ipw[8*((i64)l*nsh + fr) + 6]*c.gwf;
        s += ipw[8*((i64)l*nsh + to) + 6]*c.gwt;
        if (l < L - 1) {
            /* grad_w_raw at level l + 1 */
            const double gf1 = (c.ff - nff)/G.nd[fr + (i64)(l + 1)*G.vpl];
            const double gt1 = (c.ft - nft)/G.nd[to + (i64)(l + 1)*G.vpl];
            s += ipw[8*((i64)l*nsh + fr) + 7]*gf1;
            s += ipw[8*((i64)l*nsh + to) + 7]*gt1;
            c.gwf = gf1;
            c.gwt = gt1;
        }
        s = 0.5*s;
        g += -G.slope[vi]*s;
    }
    c.ff = nff;
    c.ft = nft;
    return g;
}

template <bool FLAT, typename TW, int NC>
__global__ void __launch_bounds__(TPB) k_pgrad_e(DevGrid G, DevState s, DevWorkT<TW> W, int first_step,
                                                 const double *gf, double chw, double ohw)
{
    const Blk bk = xcd_block();
    int e = bk.x*TPB + threadIdx.x;
    if (e >= G.nvh) return;
    const int fr = G.from[e], to = G.to[e];
    const i64 gi = (i64)G.n_cond*G.s;
    int l0, l1;
    layer_chunk(G.L, bk.y, l0, l1);
    EdgeCarry carry;
    for (int l = l0; l < l1; ++l) {
        i64 vi = G.nsh + (i64)l*G.vpl + e;
        TW old_t = 0;
        if (!first_step) {
            old_t = (TW)(-(double)W.pgrad_nl[vi] - (double)W.pgrad_l[vi]);
            W.pgrad_old[vi] = old_t;
        }
        i64 sf = (i64)l*G.nsh + fr, st = (i64)l*G.nsh + to;
        double g;
        if constexpr (FLAT) g = grad_edge_t<FLAT, double>(G, s.ep, l, e, fr, to);
        else g = grad_edge_walk<FLAT, double>(G, s.ep, l, l0, e, fr, to, carry);
        double p1_t = GC_C_D_P*(G.thbg[st] + s.tp[st]), p1_f = GC_C_D_P*(G.thbg[sf] + s.tp[sf]);
        double nl = (0.5*(p1_t + p1_f))*g;
        double p2_t = GC_C_D_P*s.tp[st], p2_f = GC_C_D_P*s.tp[sf];
        double ll = (0.5*(p2_t + p2_f))*G.exbg_grad[vi];
        double d_t, d_f;
        if (gf) { d_t = gf[st]; d_f = gf[sf]; }
        else {
            d_t = s.rho[gi + st]/density_total_n<NC>(G, s.rho, st);
            d_f = s.rho[gi + sf]/density_total_n<NC>(G, s.rho, sf);
        }
        nl = (0.5*(d_t + d_f))*nl;
        ll = (0.5*(d_t + d_f))*ll;
        W.pgrad_nl[vi] = (TW)nl;
        W.pgrad_l[vi] = (TW)ll;
        if (first_step) {
            old_t = (TW)(-nl - ll);
            W.pgrad_old[vi] = old_t;
        }
        /* the momentum tendency's combination of the stored values
           (vector_tendencies_expl.c) */
        W.pgrad_comb[vi] = (TW)(ohw*(double)old_t - chw*((double)(TW)nl + (double)(TW)ll));
    }
}

/* scalar_times_vector(rho_gas, wind) into the mass flux density
   (multiplications.c:35-117).  One thread per edge and chunk of layers (the
   edge's from/to cells are read once), and one per cell and chunk of w
   levels, in the same launch. */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_mass_flux(DevGrid G, const double *rho_gas, const double *wind, TW *flux)
{
    const Blk bk = xcd_block();
    const int t = bk.x*TPB + threadIdx.x;
    const int nsh = G.nsh, vpl = G.vpl, L = G.L;
    int l0, l1;
    layer_chunk(L, bk.y, l0, l1);
    if (t < G.nvh) {
        const int e = t, fr = G.from[e], to = G.to[e];
        for (int l = l0; l < l1; ++l) {
            const i64 vi = nsh + (i64)l*vpl + e;
            const double sv = 0.5*(rho_gas[to + (i64)l*nsh] + rho_gas[fr + (i64)l*nsh]);
            flux[vi] = (TW)(sv*wind[vi]);
        }
    }
    if (t < nsh) {
        /* interior w levels 1..L-1: level lev of this chunk's layers [l0, l1) */
        const int h = t;
        for (int lev = l0 > 1 ? l0 : 1; lev < l1; ++lev) {
            const i64 w = (i64)lev*vpl + h;
            const double sv = 0.5*(rho_gas[h + (i64)(lev - 1)*nsh] + rho_gas[h + (i64)lev*nsh]);
            flux[w] = (TW)(sv*wind[w]);
        }
    }
}

/* vorticities.c:167-222, one thread per triangle walking all layers.
   The loads of the layer below are issued before the current layer is
   formed, and the wind and edge heights of the layers above and below (the
   two candidates of the terrain interpolation) are carried in registers, so
   no load waits on the sign test that selects between them. */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_rel_vort_tri(DevGrid G, const double *w, TW *res)
{
    int h = xcd_block().x*TPB + threadIdx.x;
    if (h >= G.nt) return;
    const int L = G.L, lt = L - G.n_oro, nsh = G.nsh, vpl = G.vpl, nvh = G.nvh;
    const i64 dvpl = G.dvpl;
    int vit[3];
    double sg[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) { vit[j] = G.vit[3*h + j]; sg[j] = (double)G.vst[3*h + j]; }
    struct Lay { double w[3], z[3], nd[3], dz, da; };
    auto load = [&](int l) {
        Lay x;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const i64 vi = nsh + (i64)l*vpl + vit[j];
            x.w[j] = w[vi];
            x.z[j] = G.zv[vi];
            x.nd[j] = G.nd[vi];
        }
        const i64 di = nvh + (i64)l*dvpl + h;
        x.dz = G.dzv[di];
        x.da = G.darea[di];
        return x;
    };
    double wp[3] = {0.0, 0.0, 0.0}, zp[3] = {0.0, 0.0, 0.0};   /* layer l - 1 */
    Lay cur = load(0);
    for (int l = 0; l < L; ++l) {
        const Lay nxt = load(l + 1 < L ? l + 1 : l);
        double r = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double vel = cur.w[j], lrf = 1;
            if (l >= lt) {
                lrf = (G.radius + cur.dz)/(G.radius + cur.z[j]);
                const double delta_z = cur.dz - cur.z[j];
                /* the layer above where delta_z > 0 or at the lowest layer, else the layer below */
                const bool up = delta_z > 0 || l == L - 1;
                const double wg = up ? wp[j] : nxt.w[j], zg = up ? zp[j] : nxt.z[j];
                const double vg = (cur.w[j] - wg)/(cur.z[j] - zg);
                vel += delta_z*vg;
            }
            r += lrf*cur.nd[j]*sg[j]*vel;
        }
        res[(i64)l*G.nt + h] = (TW)(r/cur.da);
#pragma unroll
        for (int j = 0; j < 3; ++j) { wp[j] = cur.w[j]; zp[j] = cur.z[j]; }
        cur = nxt;
    }
}

/* averaging.c:97-111 */
template <bool FLAT>
__device__ __forceinline__ double horizontal_covariant_t(const DevGrid &G, const double *w, int l, int e)
{
    i64 vi = G.nsh + (i64)l*G.vpl + e;
    double r = w[vi];
    if (!FLAT && l >= G.L - G.n_oro) {
        double vc = remap_verpri2horpri(G, w, l, e);
        r += G.slope[vi]*vc;
    }
    return r;
}

/* tangential relative vorticity of edge e at level l >= 1 (vorticities.c:131-156) */
template <bool FLAT>
__device__ __forceinline__ double rel_vort_tangential(const DevGrid &G, const double *w, int l, int e, int fr, int to)
{
    i64 base = (i64)l*G.vpl;
    const double *nd = G.nd;
    if (l == G.L) {
        i64 i2 = base - G.nvh + e;
        double cv2 = horizontal_covariant_t<FLAT>(G, w, l - 1, e);
        return 1/G.darea[e + (i64)l*G.dvpl]*nd[i2]*cv2;
    }
    i64 i0 = base + G.nsh + e, i1 = base + fr, i2 = base - G.nvh + e, i3 = base + to;
    double cv0 = horizontal_covariant_t<FLAT>(G, w, l, e);
    double cv2 = horizontal_covariant_t<FLAT>(G, w, l - 1, e);
    return 1/G.darea[e + (i64)l*G.dvpl]*(
        - nd[i0]*cv0
        + nd[i1]*w[i1]
        + nd[i2]*cv2
        - nd[i3]*w[i3]);
}

/* calc_pot_vort = calc_rel_vort + add_f_to_rel_vort + division by the
   interpolated density (vorticities.c:18-165).  One thread per edge walks all
   levels and writes the tangential element of each level and the rhombus
   element of each layer. */
template <bool FLAT, typename TW>
__global__ void __launch_bounds__(TPB) k_pot_vort(DevGrid G, const double *w, const double *rho, const TW *rvt,
                                                  TW *pv)
{
    const Blk bk = xcd_block();
    int e = bk.x*TPB + threadIdx.x;
    if (e >= G.nvh) return;
    const int nvh = G.nvh, nsh = G.nsh, L = G.L, vpl = G.vpl;
    const int fr = G.from[e], to = G.to[e], fd = G.dfrom[e], td = G.dto[e];
    const double *nd = G.nd;
    int dri[4];
    double drw[4];
    for (int j = 0; j < 4; ++j) { dri[j] = G.dri[4*e + j]; drw[j] = G.drw[4*e + j]; }
    /* the edge's two cells are two of its rhombus cells: their densities come
       with the rhombus loads, and the layer above's are carried from the
       previous level, as are the covariant wind and normal distance of the
       edge in the layer above (each value is the one the reference's
       formula reads, so the arithmetic is unchanged) */
    int pf = 0, pt = 0;
    for (int j = 0; j < 4; ++j) { if (dri[j] == fr) pf = j; if (dri[j] == to) pt = j; }
    const double f_t = G.fvec[e], f_r = G.fvec[nvh + e];
    int l0, l1;
    layer_chunk(L + 1, bk.y, l0, l1);
    double rv1 = (l0 <= 1) ? rel_vort_tangential<FLAT>(G, w, 1, e, fr, to) : 0.0;
    /* layer l-1 at the edge: density of both cells, covariant wind, normal distance */
    double rf_m = 0.0, rt_m = 0.0, cv_m = 0.0, nd_m = 0.0;
    if (l0 >= 1) {
        rf_m = rho[(i64)(l0 - 1)*nsh + fr];
        rt_m = rho[(i64)(l0 - 1)*nsh + to];
        cv_m = horizontal_covariant_t<FLAT>(G, w, l0 - 1, e);
        nd_m = nd[nsh + (i64)(l0 - 1)*vpl + e];
    }
    for (int l = l0; l < l1; ++l) {
        if (l == L) {
            /* the lowest level (vorticities.c:158-163 and the surface density extrapolation) */
            const double rv = 1/G.darea[e + (i64)l*G.dvpl]*nd_m*cv_m;
            const i64 a = (i64)(l - 1)*nsh, b = (i64)(l - 2)*nsh;
            const double dv = 0.5*(rho[a + fr] + rho[a + to])
                              + (0.5*(rho[b + fr] + rho[b + to]) - 0.5*(rho[a + fr] + rho[a + to]))
                              /G.pv_sfc_dz1[e]*G.pv_sfc_dz2[e];
            pv[(i64)l*2*nvh + e] = (TW)((rv + f_t)/dv);
            break;
        }
        /* loads of level l */
        const i64 base = (i64)l*vpl, i0 = base + nsh + e, i1 = base + fr, i3 = base + to;
        const i64 db = nvh + (i64)l*G.dvpl;
        double r4[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) r4[j] = rho[(i64)l*nsh + dri[j]];
        const double nd0 = nd[i0], nd1 = nd[i1], nd3 = nd[i3], w1 = w[i1], w3 = w[i3];
        const double da_e = G.darea[e + (i64)l*G.dvpl], af = G.darea[db + fd], at = G.darea[db + td];
        const double rvf = (double)rvt[(i64)l*G.nt + fd], rvd = (double)rvt[(i64)l*G.nt + td];
        const double cv0 = horizontal_covariant_t<FLAT>(G, w, l, e);
        const double rf = pf == 0 ? r4[0] : (pf == 1 ? r4[1] : (pf == 2 ? r4[2] : r4[3]));
        const double rt = pt == 0 ? r4[0] : (pt == 1 ? r4[1] : (pt == 2 ? r4[2] : r4[3]));
        /* tangential element at level l; the top copies level 1 (vorticities.c:158-163) */
        double rv;
        if (l <= 1) rv = rv1;
        else rv = 1/da_e*(- nd0*cv0 + nd1*w1 + nd_m*cv_m - nd3*w3);
        double dv;
        if (l == 0) {
            const double rf1 = rho[nsh + fr], rt1 = rho[nsh + to];
            dv = 0.5*(rf + rt) + (0.5*(rf + rt) - 0.5*(rf1 + rt1))/G.pv_top_dz1[e]*G.pv_top_dz2[e];
        } else dv = 0.25*(rf_m + rt_m + rf + rt);
        pv[(i64)l*2*nvh + e] = (TW)((rv + f_t)/dv);
        /* rhombus element of layer l */
        const double rr = (af*rvf + at*rvd)/(af + at);
        double dr = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) dr += drw[j]*r4[j];
        pv[(i64)l*2*nvh + nvh + e] = (TW)((rr + f_r)/dr);
        rf_m = rf; rt_m = rt; cv_m = cv0; nd_m = nd0;
    }
}

/* vorticity_flux.c:16-187, horizontal part: one thread per edge, all layers */
template <typename TW>
/* at most two waves per SIMD: the register budget lets every gather of a
   layer issue in one batch (measured faster than four waves with the gathers
   issued in pairs) */
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_vorticity_flux_e(DevGrid G, const TW *F, const TW *pv, TW *out)
{
    const Blk bk = xcd_block();
    int e = bk.x*TPB + threadIdx.x;
    if (e >= G.nvh) return;
    const int nsh = G.nsh, nvh = G.nvh;
    double tw[10];
    int ti[10], tm[10];
    for (int j = 0; j < 10; ++j) { tw[j] = G.tw[10*(i64)e + j]; ti[j] = G.ti[10*(i64)e + j]; tm[j] = G.tmc[10*(i64)e + j]; }
    const int fr = G.from[e], to = G.to[e];
    const bool pent_f = fr < GAME_N_PENT, pent_t = to < GAME_N_PENT;
    /* next to a pentagon a side has four terms: its fifth slot reads the
       first slot's (valid) entries and adds -0.0, the exact identity of IEEE
       addition, so every load of a layer issues without a branch */
    if (pent_f) { ti[4] = ti[0]; tm[4] = tm[0]; }
    if (pent_t) { ti[9] = ti[5]; tm[9] = tm[5]; }
    const double *ipwT = G.ipwT;
    int l0, l1;
    layer_chunk(G.L, bk.y, l0, l1);
    /* the w-point terms of layer l+1 are those of layer l+1's upper level:
       carried to the next layer instead of loaded again */
    double Ff = (double)F[(i64)l0*G.vpl + fr], Ft = (double)F[(i64)l0*G.vpl + to];
    double q0 = (double)pv[e + (i64)l0*2*nvh];
    for (int l = l0; l < l1; ++l) {
        i64 fb = nsh + (i64)l*G.vpl, pb = nvh + (i64)l*2*nvh;
        double Fv[10], pvv[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) { Fv[j] = (double)F[fb + ti[j]]; pvv[j] = (double)pv[pb + tm[j]]; }
        const double pve = (double)pv[pb + e];
        double o = 0;
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const bool pent = j < 5 ? pent_f : pent_t;
            /* slots 2 and 7 average with the edge's own PV, except next to a pentagon */
            double t = ((j == 2 || j == 7) && !pent) ? tw[j]*Fv[j]*0.5*(pvv[j] + pve) : tw[j]*Fv[j]*pvv[j];
            if ((j == 4 || j == 9) && pent) t = -0.0;
            o += t;
        }
        /* inner-product weights 6 and 7 of the two cells in layer l */
        const double w6f = ipwT[((i64)l*8 + 6)*nsh + fr], w6t = ipwT[((i64)l*8 + 6)*nsh + to];
        const double w7f = ipwT[((i64)l*8 + 7)*nsh + fr], w7t = ipwT[((i64)l*8 + 7)*nsh + to];
        const double q1 = (double)pv[e + (i64)(l + 1)*2*nvh];
        const double Ff1 = (double)F[(i64)(l + 1)*G.vpl + fr], Ft1 = (double)F[(i64)(l + 1)*G.vpl + to];
        o -= 0.5*w6f*Ff*q0;
        o -= 0.5*w6t*Ft*q0;
        o -= 0.5*w7f*Ff1*q1;
        o -= 0.5*w7t*Ft1*q1;
        out[nsh + (i64)l*G.vpl + e] = (TW)o;
        Ff = Ff1; Ft = Ft1; q0 = q1;
    }
}

/* vorticity_flux.c:373-417, vertical part: one thread per cell, levels 0..L */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_vorticity_flux_w(DevGrid G, const TW *F, const TW *pv, TW *out)
{
    const Blk bk = xcd_block();
    int h = bk.x*TPB + threadIdx.x;
    if (h >= G.nsh) return;
    const int nsh = G.nsh, nvh = G.nvh;
    const int ne = h < GAME_N_PENT ? 5 : 6;
    /* a fixed six edges: a pentagon's sixth slot reads its first edge and
       adds -0.0, the exact identity of IEEE addition */
    int adj[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) adj[j] = G.adj[6*h + (j < ne ? j : 0)];
    const int L = G.L, vpl = G.vpl;
    const double *ipwT = G.ipwT;
    /* level l sums the terms of layers l - 1 and l; the flux and weights of
       layer l are carried to level l + 1 instead of loaded again */
    auto load_layer = [&](int l, double *Fx, double *Wx) {
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            Fx[j] = (double)F[nsh + (i64)l*vpl + adj[j]];
            Wx[j] = ipwT[((i64)l*8 + j)*nsh + h];
        }
    };
    int l0, l1;
    layer_chunk(L + 1, bk.y, l0, l1);
    double Fm[6], Wm[6];
    load_layer(l0 >= 1 ? l0 - 1 : 0, Fm, Wm);
    auto level = [&](int l) {
        double Fc[6], Wc[6], q[6];
        load_layer(l <= L - 1 ? l : L - 1, Fc, Wc);
#pragma unroll
        for (int j = 0; j < 6; ++j) q[j] = (double)pv[(i64)l*2*nvh + adj[j]];
        double vw = 0.5;
        if (l == 0 || l == L) vw = 1;
        double o = 0;
        if (l >= 1) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                double t = vw*Wm[j]*Fm[j]*q[j];
                if (j >= ne) t = -0.0;
                o += t;
            }
        }
        if (l <= L - 1) {
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                double t = vw*Wc[j]*Fc[j]*q[j];
                if (j >= ne) t = -0.0;
                o += t;
            }
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) { Fm[j] = Fc[j]; Wm[j] = Wc[j]; }
        return o;
    };
    if (l1 - l0 <= 3) {
        /* a short chunk: every level before any store, so no store orders a
           later level's loads (a step past the chunk repeats its last level) */
        double o3[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) o3[t] = level(l0 + t < l1 ? l0 + t : l1 - 1);
#pragma unroll
        for (int t = 0; t < 3; ++t)
            if (l0 + t < l1) out[(i64)(l0 + t)*vpl + h] = (TW)o3[t];
        return;
    }
    for (int l = l0; l < l1; ++l) out[(i64)l*vpl + h] = (TW)level(l);
}

/* inner_product.c:13-43, one thread per cell, all layers */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_inner_product(DevGrid G, const double *a, const double *b, TW *out)
{
    const Blk bk = xcd_block();
    int h = bk.x*TPB + threadIdx.x;
    if (h >= G.nsh) return;
    const int ne = h < GAME_N_PENT ? 5 : 6;
    /* a fixed six edges: a pentagon's sixth slot reads its first edge and
       adds -0.0, the exact identity of IEEE addition */
    int adj[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) adj[j] = G.adj[6*h + (j < ne ? j : 0)];
    int l0, l1;
    layer_chunk(G.L, bk.y, l0, l1);
    /* the w point below layer l is the one above layer l + 1: carried */
    double au = a[h + (i64)l0*G.vpl], bu = b[h + (i64)l0*G.vpl];
    auto layer = [&](int l) {
        const double *w = G.ipwT + (i64)l*8*G.nsh + h;   /* weight j at w[j*nsh] */
        const double al = a[h + (i64)(l + 1)*G.vpl], bl = b[h + (i64)(l + 1)*G.vpl];
        double o = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            i64 vi = G.nsh + (i64)l*G.vpl + adj[j];
            double t = w[j*(i64)G.nsh]*a[vi]*b[vi];
            if (j >= ne) t = -0.0;
            o += t;
        }
        o += w[6*(i64)G.nsh]*au*bu;
        o += w[7*(i64)G.nsh]*al*bl;
        au = al;
        bu = bl;
        return o;
    };
    if (l1 - l0 <= 3) {
        /* a short chunk: every layer before any store, so no store orders a
           later layer's loads (a step past the chunk repeats its last layer) */
        double o3[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) o3[t] = layer(l0 + t < l1 ? l0 + t : l1 - 1);
#pragma unroll
        for (int t = 0; t < 3; ++t)
            if (l0 + t < l1) out[(i64)(l0 + t)*G.nsh + h] = (TW)o3[t];
        return;
    }
    for (int l = l0; l < l1; ++l) out[(i64)l*G.nsh + h] = (TW)layer(l);
}

/* grad (gradient_operators.c:56-74) of a scalar field: edges (one thread per
   edge, all layers) */
template <bool FLAT, typename TW>
__global__ void __launch_bounds__(TPB) k_grad_e(DevGrid G, const TW *f, TW *out)
{
    const Blk bk = xcd_block();
    int e = bk.x*TPB + threadIdx.x;
    if (e >= G.nvh) return;
    const int fr = G.from[e], to = G.to[e];
    int l0, l1;
    layer_chunk(G.L, bk.y, l0, l1);
    EdgeCarry carry;
    for (int l = l0; l < l1; ++l)
        out[G.nsh + (i64)l*G.vpl + e] = FLAT ? (TW)grad_edge_t<FLAT, TW>(G, f, l, e, fr, to)
                                             : (TW)grad_edge_walk<FLAT, TW>(G, f, l, l0, e, fr, to, carry);
}

/* grad, interior w points */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_grad_w(DevGrid G, const TW *f, TW *out)
{
    i64 t = (i64)xcd_block().x*TPB + threadIdx.x;
    if (t >= (i64)(G.L - 1)*G.nsh) return;
    int lev = (int)(t/G.nsh) + 1, h = (int)(t - (i64)(lev - 1)*G.nsh);
    out[h + (i64)lev*G.vpl] = (TW)grad_w_raw(G, f, lev, h);
}

/* vector_tendencies_expl.c:74-131 fused with the horizontal wind update of
   manage_rkhevi.c:67-75 */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_vector_tend(DevGrid G, DevWorkT<TW> W, TW *tend_w, const double *wind_old,
                                                     double *wind_new, double new_w, double chw, double ohw,
                                                     double cvw, double dt)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= G.v) return;
    double old_w = 1.0 - new_w;
    int l = (int)(i/G.vpl);
    int h = (int)(i - (i64)l*G.vpl);
    if (i < G.nsh || i >= G.v - G.nsh) { tend_w[i] = (TW)0.0; return; }
    /* friction_acc is zero with momentum diffusion and PBL off: x + 0.0 == x.
       The old tendency has weight 0 in the first substep, where it is neither
       read nor added (its product with 0 is a zero, the identity of the sum);
       IEEE addition commutes, so adding it last changes no bit */
    if (h >= G.nsh) {
        double tv = new_w*(
            (double)W.pgrad_comb[i]   /* ohw*old - chw*(nl + l) */
            + (double)W.pvt[i]
            - 0.5*(double)W.vsq_grad[i]);
        if (old_w != 0.0) tv = old_w*(double)tend_w[i] + tv;
        tend_w[i] = (TW)tv;
        wind_new[i] = wind_old[i] + dt*tv;
    } else {
        double tv = new_w*(
            (double)W.pgrad_comb[i]   /* -cvw*(nl + l) */
            + (double)W.pvt[i]
            - 0.5*(double)W.vsq_grad[i]
            + (double)W.pgc_v[i]);
        if (old_w != 0.0) tv = old_w*(double)tend_w[i] + tv;
        tend_w[i] = (TW)tv;
    }
}

/* Vertical contravariant corrections (averaging.c:41-95) of the three fields
   divv_h and divv_h_tracer need in the terrain-following levels L-n_oro..L-1:
   the wind (tracers, divergences.c:94-145), the gas mass flux and its
   rhotheta_v flux (divergences.c:41-59); the fluxes are formed on the fly
   exactly as scalar_times_vector_h forms them (multiplications.c:45-67). */
__global__ void __launch_bounds__(TPB) k_vcc(DevGrid G, DevState s, double *vcc)
{
    /* one thread per cell walks a chunk of the terrain levels.  The
       correction at level lev sums the cell's edge terms of layer lev - 1 and
       then of layer lev, starting from zero (from layer lev alone at the top
       terrain level): the running sum of layer lev - 1's terms is carried to
       the next level, so each layer's edge terms are formed once, and the sums
       are the reference's in the reference's order */
    const Blk bk = xcd_block();
    const int h = bk.x*TPB + threadIdx.x;
    if (h >= G.nsh) return;
    const int nsh = G.nsh, L = G.L, no = G.n_oro, top = L - no;
    const i64 n = (i64)no*nsh;
    const int ne = h < GAME_N_PENT ? 5 : 6;
    const double *rg = s.rho + (i64)G.n_cond*G.s;
    /* each edge as (edge, the cell across it): the means of the two cells
       are sums of two terms, which IEEE addition forms alike in either
       order, so the cell's own density and theta_v are loaded and divided
       once per layer */
    int ej[6], nb[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int e = G.adj[6*h + (j < ne ? j : 0)], f = G.from[e];
        ej[j] = e;
        nb[j] = f == h ? G.to[e] : f;
    }
    int c0, c1;
    layer_chunk(no, bk.y, c0, c1);
    /* one layer's inputs: the cell's density and rhotheta_v, and per edge
       the wind, the neighbour's density and rhotheta_v, the weight (from the
       cell-major copy of ipw: coalesced) and the slope */
    struct In { double rh, rt, u[6], rn[6], rtn[6], w[6], sl[6]; };
    auto load = [&](int l) {
        In x;
        const i64 si = (i64)l*nsh + h;
        x.rh = rg[si]; x.rt = s.rt[si];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const i64 vi = nsh + (i64)l*G.vpl + ej[j];
            const i64 sn = (i64)l*nsh + nb[j];
            x.u[j] = s.wind[vi];
            x.rn[j] = rg[sn];
            x.rtn[j] = s.rt[sn];
            x.w[j] = G.ipwT[((i64)l*8 + (j < ne ? j : 0))*nsh + h];
            x.sl[j] = G.slope[vi];
        }
        return x;
    };
    /* the edge terms of a loaded layer added to (aw, a1, a2) and (pw, p1, p2) */
    auto add_layer = [&](const In &x, double &aw, double &a1, double &a2, double &pw, double &p1, double &p2) {
        const double rh = x.rh, th = x.rt/rh;
        double tw[6], t1[6], t2[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double u = x.u[j];
            const double rn = x.rn[j];
            const double f1 = (0.5*(rh + rn))*u;
            const double f2 = (0.5*(th + x.rtn[j]/rn))*f1;
            const double c = -0.5*x.w[j]*x.sl[j];
            tw[j] = c*u;
            t1[j] = c*f1;
            t2[j] = c*f2;
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            if (j < ne) {
                aw += tw[j]; a1 += t1[j]; a2 += t2[j];
                pw += tw[j]; p1 += t1[j]; p2 += t2[j];
            }
        }
    };
    double pw = 0.0, p1 = 0.0, p2 = 0.0;
    /* the layer above the chunk (its terms start the first level's sum),
       then the chunk's layers, each loaded before the one before it is
       summed (a step past the chunk reloads its last layer, unused) */
    const int k_first = c0 > 0 ? c0 - 1 : c0;
    In cur = load(top + k_first);
    for (int k = k_first; k < c1; ++k) {
        const In nxt = load(top + (k + 1 < c1 ? k + 1 : k));
        if (k < c0) {
            double dw = 0.0, d1 = 0.0, d2 = 0.0;
            add_layer(cur, pw, p1, p2, dw, d1, d2);
        } else {
            double rw = pw, r1 = p1, r2 = p2;
            pw = 0.0; p1 = 0.0; p2 = 0.0;
            add_layer(cur, rw, r1, r2, pw, p1, p2);
            const i64 t = (i64)k*nsh + h;
            vcc[t] = rw;
            vcc[n + t] = r1;
            vcc[2*n + t] = r2;
        }
        cur = nxt;
    }
}

/* scalar_tendencies_expl.c:89-177 for every constituent (diffusion off): the
   centred gas mass flux and its rhotheta_v flux, the upstream tracer fluxes,
   their divergences (divergences.c:14-148) and the RK-weighted tendencies.
   One thread per cell walks all layers; the edge fluxes are formed on the fly
   from the reference's own formulas, so no flux array is written. */
template <bool FLAT, typename TW, int NC>
__global__ void __launch_bounds__(TPB) k_scalar_tend(DevGrid G, DevState s, DevWorkT<TW> W, double *td_rho,
                                                     double *td_rt, int rk)
{
    const Blk bk = xcd_block();
    int h = bk.x*TPB + threadIdx.x;
    if (h >= G.nsh) return;
    /* NC: the number of constituents as a compile-time constant (1 dry, 6 moist),
       so the constituent loop unrolls and all neighbour loads can be in flight together */
    const int nsh = G.nsh, vpl = G.vpl, L = G.L, no = G.n_oro, ncc = NC > 0 ? (NC > 1 ? NC - 2 : 0) : G.n_cond;
    const int nc = NC > 0 ? NC : G.n_const;
    const i64 S = G.s, nvcc = (i64)no*nsh;
    const int ne = h < GAME_N_PENT ? 5 : 6;
    /* each edge as (edge, the cell across it, whether h is its from-cell);
       the centred fluxes are symmetric in the two cells and IEEE addition
       commutes, so they come out bit for bit as the reference forms them
       from from/to.  The neighbour loops run over a fixed six: a pentagon's
       sixth slot reads valid entries (its first edge, its own cell) and adds
       -0.0, the exact identity of IEEE addition, so no branch guards a load
       and the per-edge arrays stay in registers. */
    int ej[6], nb[6];
    unsigned from_bits = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int jj = j < ne ? j : 0;
        const int e = G.adj[6*h + jj], f = G.from[e], t = G.to[e];
        ej[j] = e;
        nb[j] = j < ne ? ((f == h) ? t : f) : h;
        if (j < ne && G.sgn[6*h + jj] > 0) from_bits |= 1u << j;
    }
    int l0, l1;
    layer_chunk(L, bk.y, l0, l1);
    /* one layer's tendencies: with NC known they are returned in orho/ort
       (stored by the caller), otherwise stored here */
    auto layer = [&](int l, double *orho, double &ort) {
        const i64 i = (i64)l*nsh + h;
        double a[6], u[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const i64 vi = nsh + (i64)l*vpl + ej[j];
            a[j] = G.area[vi];
            u[j] = s.wind[vi];
        }
        /* which faces carry a terrain correction (divergences.c:41-59) */
        int has_u = 0, has_l = 0;
        if (!FLAT) {
            if (l == L - no - 1) has_l = 1;
            else if (l == L - 1) has_u = 1;
            else if (l > L - no - 1) { has_u = 1; has_l = 1; }
        }
        const double vol = G.vol[i];
        const double area_u = has_u ? G.area[h + (i64)l*vpl] : 0.0;
        const double area_l = has_l ? G.area[h + (i64)(l + 1)*vpl] : 0.0;
        const i64 vu = (i64)(l - (L - no))*nsh + h, vl = (i64)(l + 1 - (L - no))*nsh + h;
        /* with NC known the new tendencies are kept in registers and stored
           after every load of the layer, so no store orders the loads */
        double out_rho[NC > 0 ? NC : 1];
        double out_rt = 0.0;
#pragma unroll
        for (int k = 0; k < nc; ++k) {
            const i64 shift = (i64)k*S;
            const double new_w = (rk == 1 && k != ncc) ? 0.5 : 1.0, old_w = 1.0 - new_w;
            const double *rk_ = s.rho + shift;
            const double rho_h = rk_[i];
            double res;
            if (k == ncc) {
                const double th_h = s.rt[i]/rho_h;
                double c1 = 0.0, c2 = 0.0;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    const i64 sb = (i64)l*nsh + nb[j];
                    const double rho_nb = rk_[sb];
                    /* times the edge sign +-1: exact, so a sign flip */
                    const bool pos = (from_bits >> j) & 1u;
                    double f1 = (0.5*(rho_h + rho_nb))*u[j];
                    double f2 = (0.5*(th_h + s.rt[sb]/rho_nb))*f1;
                    double t1 = (pos ? f1 : -f1)*a[j], t2 = (pos ? f2 : -f2)*a[j];
                    if (j >= ne) { t1 = -0.0; t2 = -0.0; }
                    c1 += t1;
                    c2 += t2;
                }
                double v1 = 0.0, v2 = 0.0;
                if (!FLAT) {
                    if (has_u && has_l) {
                        v1 = W.vcc[nvcc + vu]*area_u - W.vcc[nvcc + vl]*area_l;
                        v2 = W.vcc[2*nvcc + vu]*area_u - W.vcc[2*nvcc + vl]*area_l;
                    } else if (has_l) {
                        v1 = -W.vcc[nvcc + vl]*area_l;
                        v2 = -W.vcc[2*nvcc + vl]*area_l;
                    } else if (has_u) {
                        v1 = W.vcc[nvcc + vu]*area_u;
                        v2 = W.vcc[2*nvcc + vu]*area_u;
                    }
                }
                double div1 = 1.0/vol*(c1 + v1);
                double div2 = 1.0/vol*(c2 + v2);
                const double rg = rho_h;
                /* the gas weights are 1 and 0 in both substeps: the old
                   tendencies are neither read nor added (their product with 0
                   is a zero, the identity of the sum) */
                if (!G.moist) {
                    /* dry: every source term of the reference sum is a zero array */
                    res = new_w*(-div1);
                    out_rt = new_w*(-div2);
                } else {
                    /* the moist-air gas takes the vapour phase-transition rate
                       (scalar_tendencies_expl.c:94-99, 125-131, 174-176) */
                    double ptr_v = (double)W.ptr[shift + i];
                    const double th = th_h;
                    res = new_w*(-div1 + ptr_v);
                    double cv = c_v_mass_weighted_cp(G, s.rho, W.cpc, i);
                    double heat = (double)W.phr[i] + (double)W.sed_heat[i];
                    out_rt = new_w*(
                        -div2
                        + GC_C_D_V*rg/cv*heat/(GC_C_D_P*(G.exbg[i] + s.ep[i]))
                        + ptr_v*th);
                }
            } else {
                double c = 0.0;
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    /* upstream density: the from-cell for u >= 0 */
                    const bool is_from = (from_bits >> j) & 1u;
                    const double rho_nb = rk_[(i64)l*nsh + nb[j]];
                    const double sv = ((u[j] >= 0.0) == is_from) ? rho_h : rho_nb;
                    const double f = sv*u[j];
                    double t = (is_from ? f : -f)*a[j];
                    if (j >= ne) t = -0.0;
                    c += t;
                }
                double v = 0.0;
                if (!FLAT) {
                    if (has_u && has_l) {
                        double cu = W.vcc[vu], cl = W.vcc[vl];
                        double du = (cu <= 0.0) ? rk_[i - nsh] : rk_[i];
                        double dl = (cl <= 0.0) ? rk_[i] : rk_[i + nsh];
                        v = du*cu*area_u - dl*cl*area_l;
                    } else if (has_l) {
                        double cl = W.vcc[vl];
                        double dl = (cl <= 0.0) ? rk_[i] : rk_[i + nsh];
                        v = -dl*cl*area_l;
                    } else if (has_u) {
                        double cu = W.vcc[vu];
                        double du = (cu <= 0.0) ? rk_[i - nsh] : rk_[i];
                        v = du*cu*area_u;
                    }
                }
                double div = 1.0/vol*(c + v);
                i64 shift_pt = (k == ncc + 1) ? shift - S : shift;
                /* the old tendency enters with weight 0.5 in the second
                   substep only (a uniform branch); IEEE addition commutes */
                res = new_w*(-div + (double)W.ptr[shift_pt + i]);
                if (old_w != 0.0) res = old_w*td_rho[shift + i] + res;
            }
            if (NC > 0) out_rho[NC > 0 ? k : 0] = res;
            else td_rho[shift + i] = res;
        }
        if (NC > 0) {
#pragma unroll
            for (int k = 0; k < (NC > 0 ? NC : 1); ++k) orho[k] = out_rho[k];
            ort = out_rt;
        } else td_rt[i] = out_rt;
    };
    auto store = [&](int l, const double *orho, double ort) {
        const i64 i = (i64)l*nsh + h;
#pragma unroll
        for (int k = 0; k < (NC > 0 ? NC : 1); ++k) td_rho[(i64)k*S + i] = orho[k];
        td_rt[i] = ort;
    };
    constexpr int RUN = 3;
    if (NC > 0 && l1 - l0 <= RUN) {
        /* a short chunk: the loads and arithmetic of all its layers come
           before any store, so no store orders a later layer's loads (a step
           past the chunk repeats its last layer and stores nothing) */
        double orho[RUN][NC > 0 ? NC : 1], ort[RUN];
#pragma unroll
        for (int t = 0; t < RUN; ++t) layer(l0 + t < l1 ? l0 + t : l1 - 1, orho[t], ort[t]);
#pragma unroll
        for (int t = 0; t < RUN; ++t)
            if (l0 + t < l1) store(l0 + t, orho[t], ort[t]);
        return;
    }
    for (int l = l0; l < l1; ++l) {
        double orho[NC > 0 ? NC : 1], ort = 0.0;
        layer(l, orho, ort);
        if (NC > 0) store(l, orho, ort);
    }
}

/* three_band_solver_ver_waves (column_solvers.c:19-383), one column per lane.
   The coefficient rows are generated and eliminated in one downward sweep;
   the back substitution walks upwards and writes each layer as soon as both
   of its interface fluxes are known.  Quantities used in both sweeps are
   recomputed with identical arithmetic rather than stored. */
struct ColumnLayer {
    double re, rte, al, be, ga, tpe, epe;
};

__device__ __forceinline__ ColumnLayer column_layer(const DevGrid &G, const DevState &so, const DevState &sn,
                                                    const DevState &td, double dt, int rk, int i, int j)
{
    ColumnLayer c;
    i64 b = i + (i64)j*G.nsh, gi = (i64)G.n_cond*G.s;
    double vol = G.vol[b];
    double rho_o = so.rho[gi + b], rt_o = so.rt[b];
    c.re = rho_o + dt*td.rho[gi + b];
    c.rte = rt_o + dt*td.rt[b];
    if (rk == 0) {
        c.al = -rt_o/(rho_o*rho_o)/vol;
        c.be = 1.0/rho_o/vol;
        c.ga = GC_R_D/(GC_C_D_V*rt_o)*(G.exbg[b] + so.ep[b])/vol;
    } else {
        const double pdw = 0.5;
        double rho_n = sn.rho[gi + b], rt_n = sn.rt[b];
        double ao = -rt_o/(rho_o*rho_o);
        double bo = 1.0/rho_o;
        double go = GC_R_D/(GC_C_D_V*rt_o)*(G.exbg[b] + so.ep[b]);
        double an = -rt_n/(rho_n*rho_n);
        double bn = 1.0/rho_n;
        double gn = GC_R_D/(GC_C_D_V*rt_n)*(G.exbg[b] + sn.ep[b]);
        c.al = ((1.0 - pdw)*ao + pdw*an)/vol;
        c.be = ((1.0 - pdw)*bo + pdw*bn)/vol;
        c.ga = ((1.0 - pdw)*go + pdw*gn)/vol;
    }
    c.tpe = so.tp[b] + dt*vol*(c.al*td.rho[gi + b] + c.be*td.rt[b]);
    c.epe = so.ep[b] + dt*vol*c.ga*td.rt[b];
    return c;
}

/* interface values of the column at interface j (between layers j and j+1) */
__device__ __forceinline__ double theta_int_new(const DevGrid &G, const DevState &sn, int i, int j)
{
    i64 gi = (i64)G.n_cond*G.s, b = i + (i64)j*G.nsh, lo = b + G.nsh;
    return 0.5*(sn.rt[b]/sn.rho[gi + b] + sn.rt[lo]/sn.rho[gi + lo]);
}

__device__ __forceinline__ double rho_int_old(const DevGrid &G, const DevState &so, int i, int j)
{
    i64 gi = (i64)G.n_cond*G.s, b = i + (i64)j*G.nsh;
    return 0.5*(so.rho[gi + b] + so.rho[gi + b + G.nsh]);
}

template <int LT, typename TWIND>
__global__ void __launch_bounds__(128) k_ver_waves(DevGrid G, DevState so, DevState snr, DevState sn, DevState td,
                                                   const TWIND *td_wind, StepConsts K, int rk)
{
    int i = blockIdx.x*128 + threadIdx.x;
    if (i >= G.nsh) return;
    const int L = LT > 0 ? LT : G.L;
    const int nsh = G.nsh, vpl = G.vpl;
    const i64 gi = (i64)G.n_cond*G.s;
    const double dt = K.dt, iw = K.impl_weight;
    const double dt2 = dt*dt;
    const double *zs = G.zs, *exbg = G.exbg, *vol = G.vol, *area = G.area;
    double ep[(LT > 0 ? LT : GAME_MAX_LAYERS) - 1], x[(LT > 0 ? LT : GAME_MAX_LAYERS) - 1];

    /* downward sweep: build row j and eliminate it (column_solvers.c:141-189, 641-660) */
    ColumnLayer ca = column_layer(G, so, snr, td, dt, rk, i, 0);   /* layer j */
    ColumnLayer cb = column_layer(G, so, snr, td, dt, rk, i, 1);   /* layer j+1 */
    double tin = theta_int_new(G, snr, i, 0), rio = rho_int_old(G, so, i, 0);
    double c_prev = 0.0;
#pragma unroll
    for (int j = 0; j < L - 1; ++j) {
        i64 b = i + (i64)j*nsh, lo = b + nsh;
        i64 wi = i + (i64)(j + 1)*vpl;
        double w_old = so.wind[wi];
        double rie = 0.5*(ca.re + cb.re);
        double d = -(tin*tin)*(ca.ga + cb.ga)
                   + 0.5*(exbg[b] - exbg[lo])*(cb.al - ca.al + tin*(cb.be - ca.be))
                   - (zs[b] - zs[lo])/(iw*dt2*GC_C_D_P*rio)
                     *(2.0/area[wi] + dt*w_old*0.5*(-1.0/vol[b] + 1.0/vol[lo]));
        double r = -(w_old + dt*(double)td_wind[wi])*(zs[b] - zs[lo])/(iw*dt2*GC_C_D_P)
                   + tin*(ca.epe - cb.epe)/dt
                   + 0.5/dt*(ca.tpe + cb.tpe)*(exbg[b] - exbg[lo])
                   - (zs[b] - zs[lo])/(iw*dt2*GC_C_D_P)*w_old*rie/rio;
        double e = 0.0, c = 0.0, tin_n = 0.0, rio_n = 0.0;
        ColumnLayer cc = cb;
        if (j < L - 2) {
            i64 l2 = lo + nsh;
            tin_n = theta_int_new(G, snr, i, j + 1);
            rio_n = rho_int_old(G, so, i, j + 1);
            /* upper diagonal e[j] and lower diagonal c[j] (column_solvers.c:173-189) */
            c = tin_n*cb.ga*tin
                + 0.5*(exbg[lo] - exbg[l2])*(cb.al + cb.be*tin)
                - (zs[lo] - zs[l2])/(iw*dt*GC_C_D_P)*0.5*so.wind[i + (i64)(j + 2)*vpl]/(vol[lo]*rio_n);
            e = tin*cb.ga*tin_n
                - 0.5*(exbg[b] - exbg[lo])*(cb.al + cb.be*tin_n)
                + (zs[b] - zs[lo])/(iw*dt*GC_C_D_P)*0.5*w_old/(vol[lo]*rio);
            cc = column_layer(G, so, snr, td, dt, rk, i, j + 2);
        }
        if (j == 0) {
            if (L - 1 > 1) ep[0] = e/d;
            x[0] = r/d;
        } else {
            double den = d - ep[j - 1]*c_prev;
            if (j < L - 2) ep[j] = e/den;
            x[j] = (r - x[j - 1]*c_prev)/den;
        }
        c_prev = c;
        ca = cb; cb = cc;
        tin = tin_n; rio = rio_n;
    }
    /* back substitution (column_solvers.c:662-667) */
#pragma unroll
    for (int j = L - 3; j >= 0; --j) x[j] = x[j] - ep[j]*x[j + 1];
    /* Klemp (2008) damping of the solved fluxes (column_solvers.c:284-298) */
#pragma unroll
    for (int j = 0; j < L - 1; ++j) x[j] = x[j]/(1.0 + dt*G.damp[(i64)j*nsh + i]);
    /* new state, layer by layer, each layer read before it is written (column_solvers.c:300-370) */
    double tin_m = 0.0, rho_new_m = 0.0, rio_m = 0.0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        i64 b = i + (i64)j*nsh;
        ColumnLayer cl = column_layer(G, so, snr, td, dt, rk, i, j);
        double tin_j = (j < L - 1) ? theta_int_new(G, snr, i, j) : 0.0;
        double rio_j = (j < L - 1) ? rho_int_old(G, so, i, j) : 0.0;
        double v = vol[b];
        double rho_n, rt_n;
        if (j == 0) {
            rho_n = cl.re + dt*(x[j])/v;
            rt_n = cl.rte + dt*(tin_j*x[j])/v;
        } else if (j == L - 1) {
            rho_n = cl.re + dt*(-x[j - 1])/v;
            rt_n = cl.rte + dt*(-tin_m*x[j - 1])/v;
        } else {
            rho_n = cl.re + dt*(-x[j - 1] + x[j])/v;
            rt_n = cl.rte + dt*(-tin_m*x[j - 1] + tin_j*x[j])/v;
        }
        double rt_o = so.rt[b];
        sn.rho[gi + b] = rho_n;
        sn.rt[b] = rt_n;
        sn.tp[b] = rt_n/rho_n - G.thbg[b];
        sn.ep[b] = so.ep[b] + v*cl.ga*(rt_n - rt_o);
        if (j >= 1) {
            i64 wi = i + (i64)j*vpl;
            double din = 0.5*(rho_new_m + rho_n);
            sn.wind[wi] = (2.0*x[j - 1]/area[wi] - din*so.wind[wi])/rio_m;
        }
        tin_m = tin_j;
        rho_new_m = rho_n;
        rio_m = rio_j;
    }
}

/* The same solve in three phases over a block of VW_NCOL columns, for
   parallelism: the streaming kernel above gives each column one lane for
   all of its work, 2.5 waves per SIMD at res 7, and its recurrence arrays
   live in scratch.  Here
     1. every thread builds rows (column, interface) of the tridiagonal
        system into LDS, each from the two layers around the interface;
     2. one lane per column runs the Thomas elimination and back
        substitution on LDS, then all threads apply the damping;
     3. every thread forms the new state of (column, layer) boxes into
        registers; after a barrier (the reads of state snr may be of the
        state being written) they are stored.
   Every quantity is formed with the arithmetic of the streaming kernel, so
   the two kernels agree bit for bit. */
/* Thomas algorithm (column_solvers.c:641-667) on rows j*NCOL + col of LDS
   arrays c, d, e, r (n rows): the multipliers overwrite e, the solution r.
   The next row is read before the current one is written, so each step
   waits on its division chain only, not on an LDS round trip. */
template <int NCOL, int NMAX>
__device__ __forceinline__ void thomas_lds(const double *cc, const double *dd, double *ee, double *rr, int n, int col)
{
    double ep = 0.0, x;
    const double d0 = dd[col];
    if (n > 1) ep = ee[col]/d0;
    x = rr[col]/d0;
    ee[col] = ep; rr[col] = x;
    double cp = cc[col], dn = 0.0, en = 0.0, rn = 0.0;
    if (n > 1) { dn = dd[NCOL + col]; en = ee[NCOL + col]; rn = rr[NCOL + col]; }
#pragma unroll
    for (int j = 1; j < NMAX; ++j) {
        if (j >= n) break;
        const int o = j*NCOL + col, on = (j + 1 < n ? j + 1 : j)*NCOL + col;
        const double cpn = cc[o], dnn = dd[on], enn = ee[on], rnn = rr[on];
        const double den = dn - ep*cp;
        if (j < n - 1) ep = en/den;
        x = (rn - x*cp)/den;
        ee[o] = ep; rr[o] = x;
        cp = cpn; dn = dnn; en = enn; rn = rnn;
    }
    /* back substitution, upwards */
    double eb = 0.0, rb = 0.0;
    if (n > 1) { eb = ee[(n - 2)*NCOL + col]; rb = rr[(n - 2)*NCOL + col]; }
#pragma unroll
    for (int t = 0; t < NMAX - 1; ++t) {
        const int j = n - 2 - t;
        if (j < 0) break;
        const int o = j*NCOL + col, om = (j > 0 ? j - 1 : 0)*NCOL + col;
        const double ebn = ee[om], rbn = rr[om];
        x = rb - eb*x;
        rr[o] = x;
        eb = ebn; rb = rbn;
    }
}

/* thomas_lds for systems whose every row has one off-diagonal entry at
   most: interface j (between rows j and j + 1) stores that entry in ce[j]
   and in fl[j] whether it is the upper one (e_j, fl 1) or the lower one
   (c_j, row j + 1's, fl 0); the other is 0.0, as the reference leaves it.
   The arithmetic is thomas_lds's; the multipliers overwrite ce. */
template <int NCOL, int NMAX>
__device__ __forceinline__ void thomas_lds_ce(double *ce, const unsigned char *fl, const double *dd, double *rr,
                                              int n, int col)
{
    const double d0 = dd[col], q0 = ce[col];
    const bool f0 = fl[col];
    double ep = 0.0, x;
    if (n > 1) ep = (f0 ? q0 : 0.0)/d0;
    x = rr[col]/d0;
    ce[col] = ep; rr[col] = x;
    double cp = f0 ? 0.0 : q0, dn = 0.0, rn = 0.0, qn = 0.0;
    bool fn = false;
    if (n > 1) { dn = dd[NCOL + col]; rn = rr[NCOL + col]; qn = ce[NCOL + col]; fn = fl[NCOL + col]; }
#pragma unroll
    for (int j = 1; j < NMAX; ++j) {
        if (j >= n) break;
        const int o = j*NCOL + col, on = (j + 1 < n ? j + 1 : j)*NCOL + col;
        const double dnn = dd[on], rnn = rr[on], qnn = ce[on];
        const bool fnn = fl[on];
        const double en = fn ? qn : 0.0, cpn = fn ? 0.0 : qn;
        const double den = dn - ep*cp;
        if (j < n - 1) ep = en/den;
        x = (rn - x*cp)/den;
        ce[o] = ep; rr[o] = x;
        cp = cpn; dn = dnn; rn = rnn; qn = qnn; fn = fnn;
    }
    /* back substitution, upwards */
    double eb = 0.0, rb = 0.0;
    if (n > 1) { eb = ce[(n - 2)*NCOL + col]; rb = rr[(n - 2)*NCOL + col]; }
#pragma unroll
    for (int t = 0; t < NMAX - 1; ++t) {
        const int j = n - 2 - t;
        if (j < 0) break;
        const int o = j*NCOL + col, om = (j > 0 ? j - 1 : 0)*NCOL + col;
        const double ebn = ce[om], rbn = rr[om];
        x = rb - eb*x;
        rr[o] = x;
        eb = ebn; rb = rbn;
    }
}

#define VW_NCOL 32
#define VW_THREADS 256

/* the state of box (i, j) the solver reads, loaded in one batch: old state,
   tendencies, the state snr (the new state before this substep writes it),
   volume and background exner */
struct VwBox {
    double vol, exbg, rho_o, rt_o, ep_o, tp_o, trho, trt, rho_n, rt_n, ep_n;
};

__device__ __forceinline__ VwBox vw_box(const DevGrid &G, const DevState &so, const DevState &snr,
                                        const DevState &td, int i, int j)
{
    const i64 b = i + (i64)j*G.nsh, gi = (i64)G.n_cond*G.s;
    VwBox x;
    x.vol = G.vol[b]; x.exbg = G.exbg[b];
    x.rho_o = so.rho[gi + b]; x.rt_o = so.rt[b]; x.ep_o = so.ep[b]; x.tp_o = so.tp[b];
    x.trho = td.rho[gi + b]; x.trt = td.rt[b];
    x.rho_n = snr.rho[gi + b]; x.rt_n = snr.rt[b]; x.ep_n = snr.ep[b];
    return x;
}

/* column_layer's arithmetic on loaded values */
__device__ __forceinline__ ColumnLayer vw_layer(const VwBox &x, double dt, int rk)
{
    ColumnLayer c;
    c.re = x.rho_o + dt*x.trho;
    c.rte = x.rt_o + dt*x.trt;
    if (rk == 0) {
        c.al = -x.rt_o/(x.rho_o*x.rho_o)/x.vol;
        c.be = 1.0/x.rho_o/x.vol;
        c.ga = GC_R_D/(GC_C_D_V*x.rt_o)*(x.exbg + x.ep_o)/x.vol;
    } else {
        const double pdw = 0.5;
        double ao = -x.rt_o/(x.rho_o*x.rho_o);
        double bo = 1.0/x.rho_o;
        double go = GC_R_D/(GC_C_D_V*x.rt_o)*(x.exbg + x.ep_o);
        double an = -x.rt_n/(x.rho_n*x.rho_n);
        double bn = 1.0/x.rho_n;
        double gn = GC_R_D/(GC_C_D_V*x.rt_n)*(x.exbg + x.ep_n);
        c.al = ((1.0 - pdw)*ao + pdw*an)/x.vol;
        c.be = ((1.0 - pdw)*bo + pdw*bn)/x.vol;
        c.ga = ((1.0 - pdw)*go + pdw*gn)/x.vol;
    }
    c.tpe = x.tp_o + dt*x.vol*(c.al*x.trho + c.be*x.trt);
    c.epe = x.ep_o + dt*x.vol*c.ga*x.trt;
    return c;
}

/* The same solve in three phases over a block of VW_NCOL columns, for
   parallelism: the streaming kernel above gives each column one lane for
   all of its work, 2.5 waves per SIMD at res 7, and its recurrence arrays
   live in scratch.  Here
     1. every thread builds rows (column, interface) of the tridiagonal
        system into LDS, each from the two layers around the interface;
     2. one lane per column runs the Thomas elimination and back
        substitution on LDS, then all threads apply the damping;
     3. every thread forms the new state of (column, layer) boxes into
        registers; after a barrier (the reads of state snr may be of the
        state being written) they are stored.
   Each row and box loads all of its inputs in one unconditional batch
   (indices clamped in range, unused values discarded), so a thread waits on
   memory once per row, not once per branch.  Every quantity is formed with
   the arithmetic of the streaming kernel, so the two agree bit for bit. */
template <int LT, typename TWIND>
__global__ void __launch_bounds__(VW_THREADS) k_ver_waves_3p(DevGrid G, DevState so, DevState snr, DevState sn,
                                                             DevState td, const TWIND *td_wind, StepConsts K, int rk)
{
    extern __shared__ double lds[];
    const int L = LT > 0 ? LT : G.L, NI = L - 1;
    const int nsh = G.nsh, vpl = G.vpl;
    const i64 gi = (i64)G.n_cond*G.s;
    const double dt = K.dt, iw = K.impl_weight;
    const double dt2 = dt*dt;
    const double *zs = G.zs, *area = G.area;
    const int i0 = xcd_block().x*VW_NCOL;
    double *cc = lds, *dd = lds + NI*VW_NCOL, *ee = lds + 2*NI*VW_NCOL, *rr = lds + 3*NI*VW_NCOL;

    /* phase 1: row j couples interfaces j-1, j, j+1 (column_solvers.c:141-189) */
    constexpr int MAXIT1 = ((LT > 0 ? LT : GAME_MAX_LAYERS) - 1)*VW_NCOL/VW_THREADS + 1;
#pragma unroll
    for (int it = 0; it < MAXIT1; ++it) {
        const int idx = threadIdx.x + it*VW_THREADS;
        if (idx >= NI*VW_NCOL) break;
        const int j = idx/VW_NCOL, ic = i0 + (idx - j*VW_NCOL);
        const int i = ic < nsh ? ic : nsh - 1;
        const int j2 = j + 2 < L ? j + 2 : L - 1;   /* layer j+2, clamped */
        const i64 b = i + (i64)j*nsh, lo = b + nsh, l2 = i + (i64)j2*nsh;
        const i64 wi = i + (i64)(j + 1)*vpl;
        /* one batch of loads */
        const VwBox A = vw_box(G, so, snr, td, i, j), B = vw_box(G, so, snr, td, i, j + 1);
        const double rt_n2 = snr.rt[l2], rho_n2 = snr.rho[gi + l2], rho_o2 = so.rho[gi + l2], exbg2 = G.exbg[l2];
        const double zb = zs[b], zl = zs[lo], z2 = zs[l2];
        const double w_old = so.wind[wi], tdw = (double)td_wind[wi], a_w = area[wi];
        const double w_old2 = so.wind[i + (i64)j2*vpl];   /* w at level j+2 (interface j+1) */
        const ColumnLayer ca = vw_layer(A, dt, rk), cb = vw_layer(B, dt, rk);
        const double tin = 0.5*(A.rt_n/A.rho_n + B.rt_n/B.rho_n), rio = 0.5*(A.rho_o + B.rho_o);
        const double rie = 0.5*(ca.re + cb.re);
        const double d = -(tin*tin)*(ca.ga + cb.ga)
                         + 0.5*(A.exbg - B.exbg)*(cb.al - ca.al + tin*(cb.be - ca.be))
                         - (zb - zl)/(iw*dt2*GC_C_D_P*rio)
                           *(2.0/a_w + dt*w_old*0.5*(-1.0/A.vol + 1.0/B.vol));
        const double r = -(w_old + dt*tdw)*(zb - zl)/(iw*dt2*GC_C_D_P)
                         + tin*(ca.epe - cb.epe)/dt
                         + 0.5/dt*(ca.tpe + cb.tpe)*(A.exbg - B.exbg)
                         - (zb - zl)/(iw*dt2*GC_C_D_P)*w_old*rie/rio;
        double e = 0.0, c = 0.0;
        if (j < L - 2) {
            /* theta_int_new and rho_int_old of interface j+1 */
            const double tin_n = 0.5*(B.rt_n/B.rho_n + rt_n2/rho_n2), rio_n = 0.5*(B.rho_o + rho_o2);
            c = tin_n*cb.ga*tin
                + 0.5*(B.exbg - exbg2)*(cb.al + cb.be*tin)
                - (zl - z2)/(iw*dt*GC_C_D_P)*0.5*w_old2/(B.vol*rio_n);
            e = tin*cb.ga*tin_n
                - 0.5*(A.exbg - B.exbg)*(cb.al + cb.be*tin_n)
                + (zb - zl)/(iw*dt*GC_C_D_P)*0.5*w_old/(B.vol*rio);
        }
        cc[idx] = c; dd[idx] = d; ee[idx] = e; rr[idx] = r;
    }
    __syncthreads();
    /* phase 2: Thomas algorithm (column_solvers.c:641-667), one lane per column;
       the multipliers overwrite e, the solution overwrites r */
    if (threadIdx.x < VW_NCOL && i0 + (int)threadIdx.x < nsh)
        thomas_lds<VW_NCOL, (LT > 0 ? LT : GAME_MAX_LAYERS) - 1>(cc, dd, ee, rr, NI, threadIdx.x);
    __syncthreads();
    /* Klemp (2008) damping of the solved fluxes (column_solvers.c:284-298) */
#pragma unroll
    for (int it = 0; it < MAXIT1; ++it) {
        const int idx = threadIdx.x + it*VW_THREADS;
        if (idx >= NI*VW_NCOL) break;
        const int j = idx/VW_NCOL, ic = i0 + (idx - j*VW_NCOL);
        const int i = ic < nsh ? ic : nsh - 1;
        rr[idx] = rr[idx]/(1.0 + dt*G.damp[(i64)j*nsh + i]);
    }
    __syncthreads();
    /* phase 3: new state of box (column, layer j) (column_solvers.c:300-370) */
    constexpr int MAXIT3 = (LT > 0 ? LT : GAME_MAX_LAYERS)*VW_NCOL/VW_THREADS + 1;
    double o_rho[MAXIT3], o_rt[MAXIT3], o_tp[MAXIT3], o_ep[MAXIT3], o_w[MAXIT3];
#pragma unroll
    for (int it = 0; it < MAXIT3; ++it) {
        o_rho[it] = 0.0; o_rt[it] = 0.0; o_tp[it] = 0.0; o_ep[it] = 0.0; o_w[it] = 0.0;
        const int idx = threadIdx.x + it*VW_THREADS;
        if (idx >= L*VW_NCOL) break;
        const int j = idx/VW_NCOL, col = idx - j*VW_NCOL, ic = i0 + col;
        const int i = ic < nsh ? ic : nsh - 1;
        const int jm = j >= 1 ? j - 1 : 0, jp = j < L - 1 ? j + 1 : L - 1;
        const i64 b = i + (i64)j*nsh, bm = i + (i64)jm*nsh, bp = i + (i64)jp*nsh;
        const i64 wi = i + (i64)(j >= 1 ? j : 1)*vpl;
        /* one batch of loads */
        const VwBox A = vw_box(G, so, snr, td, i, j);
        const double rt_np = snr.rt[bp], rho_np = snr.rho[gi + bp];
        const double rt_nm = snr.rt[bm], rho_nm = snr.rho[gi + bm], rho_om = so.rho[gi + bm];
        const double trho_m = td.rho[gi + bm], vm = G.vol[bm];
        const double thbg = G.thbg[b], a_w = area[wi], w_o = so.wind[wi];
        const double x_j = (j < L - 1) ? rr[j*VW_NCOL + col] : 0.0;
        const double x_m = (j >= 1) ? rr[(j - 1)*VW_NCOL + col] : 0.0;
        const double x_mm = (j >= 2) ? rr[(j - 2)*VW_NCOL + col] : 0.0;
        const ColumnLayer cl = vw_layer(A, dt, rk);
        /* theta_int_new of the interfaces below (j) and above (j-1) the box */
        const double tin_j = (j < L - 1) ? 0.5*(A.rt_n/A.rho_n + rt_np/rho_np) : 0.0;
        const double tin_m = (j >= 1) ? 0.5*(rt_nm/rho_nm + A.rt_n/A.rho_n) : 0.0;
        const double v = A.vol;
        double rho_n, rt_n;
        if (j == 0) {
            rho_n = cl.re + dt*(x_j)/v;
            rt_n = cl.rte + dt*(tin_j*x_j)/v;
        } else if (j == L - 1) {
            rho_n = cl.re + dt*(-x_m)/v;
            rt_n = cl.rte + dt*(-tin_m*x_m)/v;
        } else {
            rho_n = cl.re + dt*(-x_m + x_j)/v;
            rt_n = cl.rte + dt*(-tin_m*x_m + tin_j*x_j)/v;
        }
        o_rho[it] = rho_n;
        o_rt[it] = rt_n;
        o_tp[it] = rt_n/rho_n - thbg;
        o_ep[it] = A.ep_o + v*cl.ga*(rt_n - A.rt_o);
        if (j >= 1) {
            /* the new density of the box above, with the arithmetic of its own thread */
            const double re_m = rho_om + dt*trho_m;
            double rho_m;
            if (j - 1 == 0) rho_m = re_m + dt*(x_m)/vm;
            else rho_m = re_m + dt*(-x_mm + x_m)/vm;
            const double rio_m = 0.5*(rho_om + A.rho_o);   /* rho_int_old(j - 1) */
            const double din = 0.5*(rho_m + rho_n);
            o_w[it] = (2.0*x_m/a_w - din*w_o)/rio_m;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXIT3; ++it) {
        const int idx = threadIdx.x + it*VW_THREADS;
        if (idx >= L*VW_NCOL) break;
        const int j = idx/VW_NCOL, i = i0 + (idx - j*VW_NCOL);
        if (i >= nsh) continue;
        const i64 b = i + (i64)j*nsh;
        sn.rho[gi + b] = o_rho[it];
        sn.rt[b] = o_rt[it];
        sn.tp[b] = o_tp[it];
        sn.ep[b] = o_ep[it];
        if (j >= 1) sn.wind[i + (i64)j*vpl] = o_w[it];
    }
}

/* The 26- and 55-layer sound-wave solve with each wave walking a run of
   layers of a block of vr_ncol(L) columns (lane = column): a row of the three-band system
   needs the layers above and below its interface, so walking down a run
   forms each layer's linearisation (column_layer, nine divisions at rk 1)
   once instead of twice and loads each box once; the inputs of the next
   layer are loaded before the current row is formed.  Phase 2 is the same
   one-lane-per-column elimination (a full wave per block), phase 3 walks
   runs of boxes the same way, carrying the box above's values (its new
   density is the one its own step formed).  Every quantity keeps the
   arithmetic of k_ver_waves. */
#define VR_THREADS 256
/* columns per block: 64 at 26 layers (50 KiB of LDS, three blocks per CU),
   32 at 55 layers (55 KiB, two blocks per CU; the lanes of a column group
   are half a wave) */
__host__ __device__ constexpr int vr_ncol(int lt) { return lt == 26 ? 64 : 32; }

template <int LT, typename TWIND>
__global__ void __launch_bounds__(VR_THREADS) k_ver_waves_run(DevGrid G, DevState so, DevState snr, DevState sn,
                                                              DevState td, const TWIND *td_wind, StepConsts K, int rk)
{
    extern __shared__ double lds[];
    constexpr int VR_NCOL = vr_ncol(LT);
    constexpr int L = LT, NI = LT - 1;
    constexpr int NW = VR_THREADS/VR_NCOL;
    constexpr int PER1 = (NI + NW - 1)/NW, PER3 = (L + NW - 1)/NW;
    /* every wave's run of rows and of boxes starts inside the column, so
       each of its first loads is in range */
    static_assert((NW - 1)*PER1 < NI && (NW - 1)*PER3 < L, "a wave of the block would have an empty run");
    const int nsh = G.nsh, vpl = G.vpl;
    const i64 gi = (i64)G.n_cond*G.s;
    const double dt = K.dt, iw = K.impl_weight;
    const double dt2 = dt*dt;
    const double *zs = G.zs, *area = G.area;
    const int i0 = xcd_block().x*VR_NCOL;
    const int w = threadIdx.x/VR_NCOL, col = threadIdx.x - w*VR_NCOL;
    const bool live = i0 + col < nsh;
    const int i = live ? i0 + col : nsh - 1;
    double *cc = lds, *dd = lds + NI*VR_NCOL, *ee = lds + 2*NI*VR_NCOL, *rr = lds + 3*NI*VR_NCOL;

    /* phase 1: rows j0..j1-1 of this wave (column_solvers.c:141-189) */
    {
        const int j0 = w*PER1, j1 = j0 + PER1 < NI ? j0 + PER1 : NI;
        VwBox A = vw_box(G, so, snr, td, i, j0), B = vw_box(G, so, snr, td, i, j0 + 1);
        double za = zs[i + (i64)j0*nsh], zb = zs[i + (i64)(j0 + 1)*nsh];
        ColumnLayer ca = vw_layer(A, dt, rk), cb = vw_layer(B, dt, rk);
#pragma unroll
        for (int t = 0; t < PER1; ++t) {
            const int j = j0 + t;
            if (j >= j1) break;
            const int j2 = j + 2 < L ? j + 2 : L - 1;   /* layer j+2, clamped */
            const VwBox C = vw_box(G, so, snr, td, i, j2);
            const double zc = zs[i + (i64)j2*nsh];
            const i64 wi = i + (i64)(j + 1)*vpl;
            const double w_old = so.wind[wi], tdw = (double)td_wind[wi], a_w = area[wi];
            const double w_old2 = so.wind[i + (i64)j2*vpl];   /* w at level j+2 (interface j+1) */
            const double tin = 0.5*(A.rt_n/A.rho_n + B.rt_n/B.rho_n), rio = 0.5*(A.rho_o + B.rho_o);
            const double rie = 0.5*(ca.re + cb.re);
            const double d = -(tin*tin)*(ca.ga + cb.ga)
                             + 0.5*(A.exbg - B.exbg)*(cb.al - ca.al + tin*(cb.be - ca.be))
                             - (za - zb)/(iw*dt2*GC_C_D_P*rio)
                               *(2.0/a_w + dt*w_old*0.5*(-1.0/A.vol + 1.0/B.vol));
            const double r = -(w_old + dt*tdw)*(za - zb)/(iw*dt2*GC_C_D_P)
                             + tin*(ca.epe - cb.epe)/dt
                             + 0.5/dt*(ca.tpe + cb.tpe)*(A.exbg - B.exbg)
                             - (za - zb)/(iw*dt2*GC_C_D_P)*w_old*rie/rio;
            double e = 0.0, c = 0.0;
            if (j < L - 2) {
                const double tin_n = 0.5*(B.rt_n/B.rho_n + C.rt_n/C.rho_n), rio_n = 0.5*(B.rho_o + C.rho_o);
                c = tin_n*cb.ga*tin
                    + 0.5*(B.exbg - C.exbg)*(cb.al + cb.be*tin)
                    - (zb - zc)/(iw*dt*GC_C_D_P)*0.5*w_old2/(B.vol*rio_n);
                e = tin*cb.ga*tin_n
                    - 0.5*(A.exbg - B.exbg)*(cb.al + cb.be*tin_n)
                    + (za - zb)/(iw*dt*GC_C_D_P)*0.5*w_old/(B.vol*rio);
            }
            const int idx = j*VR_NCOL + col;
            cc[idx] = c; dd[idx] = d; ee[idx] = e; rr[idx] = r;
            A = B; ca = cb; za = zb;
            B = C; cb = vw_layer(C, dt, rk); zb = zc;
        }
    }
    __syncthreads();
    if (threadIdx.x < VR_NCOL && live) thomas_lds<VR_NCOL, NI>(cc, dd, ee, rr, NI, col);
    __syncthreads();
    /* Klemp (2008) damping of the solved fluxes (column_solvers.c:284-298) */
#pragma unroll
    for (int it = 0; it < (NI*VR_NCOL + VR_THREADS - 1)/VR_THREADS; ++it) {
        const int idx = threadIdx.x + it*VR_THREADS;
        if (idx >= NI*VR_NCOL) break;
        const int j = idx/VR_NCOL, ic = i0 + (idx - j*VR_NCOL);
        const int ii = ic < nsh ? ic : nsh - 1;
        rr[idx] = rr[idx]/(1.0 + dt*G.damp[(i64)j*nsh + ii]);
    }
    __syncthreads();
    /* phase 3: new state of boxes k0..k1-1 of this wave (column_solvers.c:300-370) */
    const int k0 = w*PER3, k1 = k0 + PER3 < L ? k0 + PER3 : L;
    double o_rho[PER3], o_rt[PER3], o_tp[PER3], o_ep[PER3], o_w[PER3];
    {
        VwBox A = vw_box(G, so, snr, td, i, k0);
        /* the box above the run: theta, old density and new density */
        double rt_nm = 0.0, rho_nm = 1.0, rho_om = 0.0, rho_prev = 0.0;
        if (k0 >= 1) {
            const i64 bm = i + (i64)(k0 - 1)*nsh;
            rt_nm = snr.rt[bm]; rho_nm = snr.rho[gi + bm]; rho_om = so.rho[gi + bm];
            const double re_m = rho_om + dt*td.rho[gi + bm], vm = G.vol[bm];
            const double x_m = rr[(k0 - 1)*VR_NCOL + col];
            if (k0 - 1 == 0) rho_prev = re_m + dt*(x_m)/vm;
            else rho_prev = re_m + dt*(-rr[(k0 - 2)*VR_NCOL + col] + x_m)/vm;
        }
#pragma unroll
        for (int t = 0; t < PER3; ++t) {
            o_rho[t] = 0.0; o_rt[t] = 0.0; o_tp[t] = 0.0; o_ep[t] = 0.0; o_w[t] = 0.0;
            const int j = k0 + t;
            if (j >= k1) break;
            const int jp = j + 1 < L ? j + 1 : L - 1;
            const VwBox P = vw_box(G, so, snr, td, i, jp);
            const i64 b = i + (i64)j*nsh, wi = i + (i64)(j >= 1 ? j : 1)*vpl;
            const double thbg = G.thbg[b], a_w = area[wi], w_o = so.wind[wi];
            const double x_j = (j < L - 1) ? rr[j*VR_NCOL + col] : 0.0;
            const double x_m = (j >= 1) ? rr[(j - 1)*VR_NCOL + col] : 0.0;
            const ColumnLayer cl = vw_layer(A, dt, rk);
            const double tin_j = (j < L - 1) ? 0.5*(A.rt_n/A.rho_n + P.rt_n/P.rho_n) : 0.0;
            const double tin_m = (j >= 1) ? 0.5*(rt_nm/rho_nm + A.rt_n/A.rho_n) : 0.0;
            const double v = A.vol;
            double rho_n, rt_n;
            if (j == 0) {
                rho_n = cl.re + dt*(x_j)/v;
                rt_n = cl.rte + dt*(tin_j*x_j)/v;
            } else if (j == L - 1) {
                rho_n = cl.re + dt*(-x_m)/v;
                rt_n = cl.rte + dt*(-tin_m*x_m)/v;
            } else {
                rho_n = cl.re + dt*(-x_m + x_j)/v;
                rt_n = cl.rte + dt*(-tin_m*x_m + tin_j*x_j)/v;
            }
            o_rho[t] = rho_n;
            o_rt[t] = rt_n;
            o_tp[t] = rt_n/rho_n - thbg;
            o_ep[t] = A.ep_o + v*cl.ga*(rt_n - A.rt_o);
            if (j >= 1) {
                const double rio_m = 0.5*(rho_om + A.rho_o);   /* rho_int_old(j - 1) */
                const double din = 0.5*(rho_prev + rho_n);
                o_w[t] = (2.0*x_m/a_w - din*w_o)/rio_m;
            }
            rt_nm = A.rt_n; rho_nm = A.rho_n; rho_om = A.rho_o; rho_prev = rho_n;
            A = P;
        }
    }
    __syncthreads();
    if (!live) return;
#pragma unroll
    for (int t = 0; t < PER3; ++t) {
        const int j = k0 + t;
        if (j >= k1) break;
        const i64 b = i + (i64)j*nsh;
        sn.rho[gi + b] = o_rho[t];
        sn.rt[b] = o_rt[t];
        sn.tp[b] = o_tp[t];
        sn.ep[b] = o_ep[t];
        if (j >= 1) sn.wind[i + (i64)j*vpl] = o_w[t];
    }
}

/* three_band_solver_gen_densities (column_solvers.c:385-639).

   The tracers of a column are independent of each other except through the
   sediment heat, which k_sed_heat forms separately, so one block solves NCOL
   columns of one tracer in three phases:
     1. every thread builds rows (column, layer) of the tridiagonal system
        into LDS -- all loads independent, fully coalesced; an interface has
        one nonzero off-diagonal (the upstream side of its flux), stored
        once with a flag, so a column takes three doubles and a byte per row;
     2. one lane per column runs the Thomas elimination on LDS;
     3. every thread writes the clipped solution back, coalesced.
   Each coefficient is formed with the reference's arithmetic. */
#define GD_THREADS 256
/* columns per block: 64 at 26 layers (41 KiB of LDS, three blocks per CU;
   60 columns and four blocks measured no faster); 32 for deeper columns
   (43 KiB at 55 layers: three blocks per CU where four-array rows allowed
   two, 50 KiB at 64) -- the lanes of a 32-column group are half a wave */
__host__ __device__ constexpr int gd_ncol(int lt) { return lt == 26 ? 64 : 32; }

/* vertical fluxes of tracer k through interface jj of column i
   (column_solvers.c:414-465): fi with the old wind, fr with the new wind
   times the upstream density */
__device__ __forceinline__ void tracer_iface_flux(const DevGrid &G, const DevState &so, const DevState &sn,
                                                  const StepConsts &K, const double *rho_o, int k, int i, int jj,
                                                  double &fi, double &fr)
{
    const int nsh = G.nsh, vpl = G.vpl, ncc = G.n_cond;
    fi = so.wind[i + (i64)(jj + 1)*vpl];
    fr = sn.wind[i + (i64)(jj + 1)*vpl];
    if (k < ncc/4) { fi -= K.snow_v; fr -= K.snow_v; }
    else if (k < ncc/2) { fi -= K.rain_v; fr -= K.rain_v; }
    else if (k < ncc) { fi -= K.cloud_v; fr -= K.cloud_v; }
    double a = G.area[i + (i64)(jj + 1)*vpl];
    fi = a*fi;
    fr = a*fr;
    i64 b = i + (i64)jj*nsh;
    /* both candidates load before the sign is known: no load waits on another */
    const double d_lo = rho_o[b + nsh], d_up = rho_o[b];
    double doi = (fr >= 0.0) ? d_lo : d_up;
    fr = doi*fr;
}

template <int LT>
__global__ void __launch_bounds__(GD_THREADS) k_gen_densities(DevGrid G, DevState so, DevState sn, DevState td,
                                                              StepConsts K)
{
    extern __shared__ double lds[];
    constexpr int GD_NCOL = gd_ncol(LT);
    const int L = LT > 0 ? LT : G.L, nsh = G.nsh, ncc = G.n_cond;
    /* the tracer is the fast block index: the blocks of one group of columns
       run side by side on one XCD and share the wind, area and volume columns in L2 */
    const unsigned ntr = (unsigned)G.n_const - 1, blk = xcd_block().x;
    const int tr = (int)(blk%ntr);
    const int k = tr < ncc ? tr : tr + 1;   /* skip the gas */
    const int i0 = (int)(blk/ntr)*GD_NCOL;
    const double iw = 0.5, ew = 1.0 - iw, dt = K.dt;
    const double *vol = G.vol, *area = G.area;
    const double *rho_o = so.rho + (i64)k*G.s;
    const double *trho = td.rho + (i64)k*G.s;
    double vsink = 0.0;
    if (k < ncc/4) vsink = K.snow_v;
    else if (k < ncc/2) vsink = K.rain_v;
    else if (k < ncc) vsink = K.cloud_v;
    /* one off-diagonal per interface (the flux's sign picks c or e), so three
       arrays of doubles and one of flags */
    double *ce = lds, *dd = lds + L*GD_NCOL, *rr = lds + 2*L*GD_NCOL;
    unsigned char *fl = (unsigned char *)(lds + 3*L*GD_NCOL);

    /* phase 1: rows (column_solvers.c:466-615).  Each wave takes a run of
       layers of the block's columns (lane = column, so every load is
       coalesced) and walks down it: the flux through the interface below
       row j is the flux above row j+1, formed once.  (iw dt)/vol comes from
       the per-box table the runtime keeps for the current dt: the same
       quotient, formed once instead of once per tracer and row.  The inputs
       of every row of the run are loaded before the first row is formed
       (indices clamped in range; a step past the run repeats its last row
       and writes nothing), so the walk waits on memory once, not once per
       row: about fifty loads in flight per wave, within the registers the
       LDS-bound occupancy leaves (206 -> 200 us at 26 layers, res 7, and
       524 -> 516 us at 55, against one row ahead).  Lanes past the last
       column compute the last column's rows; phases 2 and 3 skip them. */
    {
        constexpr int NW = GD_THREADS/GD_NCOL;
        constexpr int PER_MAX = ((LT > 0 ? LT : GAME_MAX_LAYERS) + NW - 1)/NW;
        const int per = (L + NW - 1)/NW;
        const int w = threadIdx.x/GD_NCOL, col = threadIdx.x - w*GD_NCOL;
        const int i = i0 + col < nsh ? i0 + col : nsh - 1;
        const int j0 = w*per, j1 = j0 + per < L ? j0 + per : L;
        const double *idv = G.idv;
        const int vpl = G.vpl;
        const double sv = k < ncc/4 ? K.snow_v : (k < ncc/2 ? K.rain_v : (k < ncc ? K.cloud_v : 0.0));
        /* the bottom row's sedimentation terms: surface area and lowest density */
        const double a_sfc = area[i + G.v - nsh], r_low = rho_o[i + G.s - nsh];
        struct In { double wo, wn, a, dlo, dup, q, qlo, v, rb, tb; };
        /* the loads of row j and of the interface below it */
        auto load = [&](int j) {
            In x;
            const int jc = j < L - 1 ? j : L - 2;
            const i64 wi = i + (i64)(jc + 1)*vpl, bc = i + (i64)jc*nsh, b = i + (i64)j*nsh;
            x.wo = so.wind[wi]; x.wn = sn.wind[wi]; x.a = area[wi];
            x.dlo = rho_o[bc + nsh]; x.dup = rho_o[bc];
            x.q = idv[b]; x.qlo = idv[j < L - 1 ? b + nsh : b];
            x.v = vol[b]; x.rb = rho_o[b]; x.tb = trho[b];
            return x;
        };
        /* the same for row j + 1 of the walk, given row j: its density above
           the interface, its own density and its quotient are row j's
           density below and quotient below (for the last row the first is
           not used), so three of the ten values come from registers */
        auto load_next = [&](int j, const In &prev) {
            In x;
            const int jc = j < L - 1 ? j : L - 2;
            const i64 wi = i + (i64)(jc + 1)*vpl, bc = i + (i64)jc*nsh, b = i + (i64)j*nsh;
            x.wo = so.wind[wi]; x.wn = sn.wind[wi]; x.a = area[wi];
            x.dlo = rho_o[bc + nsh]; x.dup = prev.dlo;
            x.q = prev.qlo; x.qlo = idv[j < L - 1 ? b + nsh : b];
            x.v = vol[b]; x.rb = prev.dlo; x.tb = trho[b];
            return x;
        };
        /* interface fluxes (column_solvers.c:414-465) */
        auto iface = [&](const In &x, int jj, double &fi, double &fr) {
            double wo = x.wo, wn = x.wn;
            if (k < ncc) { wo -= sv; wn -= sv; }
            wo = x.a*wo;
            wn = x.a*wn;
            const double doi = (wn >= 0.0) ? x.dlo : x.dup;
            fi = jj < L - 1 ? wo : 0.0;
            fr = jj < L - 1 ? doi*wn : 0.0;
        };
        /* a wave whose run starts past the last layer (more waves than runs
           when L is small against the waves of the block) has no row: it
           loads nothing */
        if (j0 < j1) {
            double fi_m = 0.0, fr_m = 0.0;
            /* every load of the run is issued before the first row is formed:
               the interface above the run (its flux enters the first row;
               interface 0 stands in, unused, for the top run), then the rows
               (a step past the run's end repeats its last row, unused) */
            In xm;
            {
                const int jm = j0 > 0 ? j0 - 1 : 0;
                const i64 wi = i + (i64)(jm + 1)*vpl, bc = i + (i64)jm*nsh;
                xm.wo = so.wind[wi]; xm.wn = sn.wind[wi]; xm.a = area[wi];
                xm.dlo = rho_o[bc + nsh]; xm.dup = rho_o[bc];
            }
            In xs[PER_MAX];
            xs[0] = load(j0);
#pragma unroll
            for (int t = 1; t < PER_MAX; ++t) xs[t] = load_next(j0 + t < j1 ? j0 + t : j1 - 1, xs[t - 1]);
            if (j0 > 0) iface(xm, j0 - 1, fi_m, fr_m);
#pragma unroll
            for (int t = 0; t < PER_MAX; ++t) {
                const bool valid = j0 + t < j1;
                const int j = valid ? j0 + t : j1 - 1;
                const In &cur = xs[t];
                const int idx = j*GD_NCOL + col;
                double fi, fr;
                iface(cur, j, fi, fr);
                const double q = cur.q, v = cur.v;   /* q = iw*dt/vol[b] */
                double d;
                if (j == 0) d = (fi >= 0.0) ? 1.0 : 1.0 - q*fi;
                else if (j == L - 1) {
                    d = (fi_m >= 0.0) ? 1.0 + q*fi_m : 1.0;
                    if (k < ncc) d += iw*vsink*dt*a_sfc/v;
                } else {
                    d = 1.0;
                    if (fi_m >= 0.0) d += q*fi_m;
                    if (fi < 0.0) d -= q*fi;
                }
                double r = cur.rb + dt*cur.tb;
                if (j == 0) r += ew*dt*fr/v;
                else if (j == L - 1) {
                    r += -ew*dt*fr_m/v;
                    if (k < ncc) r += -ew*vsink*dt*r_low*a_sfc/v;
                } else r += ew*dt*(-fr_m + fr)/v;
                /* off-diagonals of interface j (column_solvers.c:475-488);
                   -iw*dt/vol*fi is (-(iw*dt/vol))*fi exactly; the other one
                   stays 0.0 */
                double o = 0.0;
                const bool up = fi >= 0.0;
                if (j < L - 1) o = up ? (-q)*fi : cur.qlo*fi;
                if (valid) { ce[idx] = o; fl[idx] = up; dd[idx] = d; rr[idx] = r; }
                fi_m = fi;
                fr_m = fr;
            }
        }
    }
    __syncthreads();
    /* phase 2: Thomas algorithm (column_solvers.c:641-667), one lane per column;
       the multipliers overwrite e, the solution overwrites r */
    if (threadIdx.x < GD_NCOL && i0 + (int)threadIdx.x < nsh)
        thomas_lds_ce<GD_NCOL, (LT > 0 ? LT : GAME_MAX_LAYERS)>(ce, fl, dd, rr, L, threadIdx.x);
    __syncthreads();
    /* phase 3: clipped solution (column_solvers.c:616-637) */
    double *rho_n = sn.rho + (i64)k*G.s;
    for (int idx = threadIdx.x; idx < L*GD_NCOL; idx += GD_THREADS) {
        const int j = idx/GD_NCOL, col = idx - j*GD_NCOL, i = i0 + col;
        if (i >= nsh) continue;
        double v = rr[idx];
        if (v < 0.0) v = 0.0;
        rho_n[i + (i64)j*nsh] = v;
    }
}

/* The moist tracer solve with the five tracers of a group of NCOLS columns in
   one block (the default at 26 layers; k_gen_densities above is the
   one-tracer-per-block form, the default otherwise).  A row's inputs that every tracer shares --
   the old and new vertical wind and the area of the interface below it, the
   volume and (iw dt)/volume -- are loaded once for the five tracers, and a
   tracer's density above an interface is the one below the interface
   above, carried down the run, so phase 1 issues 16 loads per row instead
   of 35.  Tracer t is constituent t for the four condensates (snow, rain,
   cloud ice, cloud water: t < 4) and constituent 5 for the vapour; the gas
   (constituent 4) has no tracer solve.  Each coefficient is formed with
   the arithmetic of k_gen_densities, phase 2 runs one lane per (tracer,
   column) system and phase 3 writes the five tracers back, coalesced along
   the columns. */
template <int LT, int NCOLS, int NT>
__global__ void __launch_bounds__(NT) k_gen_densities_all(DevGrid G, DevState so, DevState sn, DevState td,
                                                          StepConsts K)
{
    extern __shared__ double lds[];
    constexpr int NTR = 5, NSYS = NTR*NCOLS, NW = NT/NCOLS, L = LT, PER = (L + NW - 1)/NW;
    static_assert(NSYS <= NT, "one lane per system in phase 2");
    const int nsh = G.nsh, vpl = G.vpl;
    const i64 S = G.s;
    const int i0 = (int)xcd_block().x*NCOLS;
    const double iw = 0.5, ew = 1.0 - iw, dt = K.dt;
    double *ce = lds, *dd = lds + L*NSYS, *rr = lds + 2*L*NSYS;
    unsigned char *fl = (unsigned char *)(lds + 3*L*NSYS);
    /* tracer t's constituent and sinking speed (column_solvers.c:414-465) */
    auto kof = [](int t) { return t < 4 ? t : 5; };
    auto vs = [&](int t) { return t < 1 ? K.snow_v : (t < 2 ? K.rain_v : (t < 4 ? K.cloud_v : 0.0)); };

    /* phase 1: rows j0..j1-1 of the five tracers of column col
       (column_solvers.c:466-615).  Every load of the run is issued before
       the first row is formed (a step past the run's end repeats its last
       row and writes nothing); a run starting past the last layer loads
       nothing.  Lanes past the last column compute the last column's rows;
       phases 2 and 3 skip them. */
    {
        const int w = threadIdx.x/NCOLS, col = threadIdx.x - w*NCOLS;
        const int i = i0 + col < nsh ? i0 + col : nsh - 1;
        const int j0 = w*PER, j1 = j0 + PER < L ? j0 + PER : L;
        if (j0 < j1) {
            const double a_sfc = G.area[i + G.v - nsh];
            /* the interface above the run (its fluxes enter the first row;
               interface 0 stands in, unused, for the top run) */
            const int jm = j0 > 0 ? j0 - 1 : 0;
            const i64 wim = i + (i64)(jm + 1)*vpl;
            const double wo_m = so.wind[wim], wn_m = sn.wind[wim], a_m = G.area[wim];
            /* per tracer: the density of the run's first row, of the row
               above it and of the lowest layer */
            double r0[NTR], dup_m[NTR], r_low[NTR];
#pragma unroll
            for (int t = 0; t < NTR; ++t) {
                const double *ro = so.rho + (i64)kof(t)*S;
                r0[t] = ro[i + (i64)j0*nsh];
                dup_m[t] = ro[i + (i64)jm*nsh];
                r_low[t] = ro[i + S - nsh];
            }
            /* per row j: the interface below it (jc, clamped to the lowest
               interface for the last row, whose fluxes are zero) */
            double wo[PER], wn[PER], ar[PER], q[PER], qlo[PER], v[PER], dlo[PER][NTR], tb[PER][NTR];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int j = j0 + u < j1 ? j0 + u : j1 - 1;
                const int jc = j < L - 1 ? j : L - 2;
                const i64 wi = i + (i64)(jc + 1)*vpl, b = i + (i64)j*nsh;
                wo[u] = so.wind[wi]; wn[u] = sn.wind[wi]; ar[u] = G.area[wi];
                q[u] = G.idv[b]; qlo[u] = G.idv[j < L - 1 ? b + nsh : b]; v[u] = G.vol[b];
#pragma unroll
                for (int t = 0; t < NTR; ++t) {
                    const i64 kS = (i64)kof(t)*S;
                    dlo[u][t] = so.rho[kS + i + (i64)(jc + 1)*nsh];
                    tb[u][t] = td.rho[kS + b];
                }
            }
            double fi_m[NTR], fr_m[NTR];
#pragma unroll
            for (int t = 0; t < NTR; ++t) {
                fi_m[t] = 0.0;
                fr_m[t] = 0.0;
                if (j0 > 0) {
                    double o_ = wo_m, n_ = wn_m;
                    if (t < 4) { o_ -= vs(t); n_ -= vs(t); }
                    o_ = a_m*o_;
                    n_ = a_m*n_;
                    const double doi = (n_ >= 0.0) ? r0[t] : dup_m[t];
                    fi_m[t] = o_;
                    fr_m[t] = doi*n_;
                }
            }
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const bool valid = j0 + u < j1;
                const int j = valid ? j0 + u : j1 - 1;
                const double qq = q[u], vv = v[u];
#pragma unroll
                for (int t = 0; t < NTR; ++t) {
                    const double sv = vs(t);
                    /* the row's own density, which is also the density above
                       its interface below (fluxes of the last row are zero) */
                    const double rb = u == 0 ? r0[t] : dlo[u - 1][t];
                    double wo_ = wo[u], wn_ = wn[u];
                    if (t < 4) { wo_ -= sv; wn_ -= sv; }
                    wo_ = ar[u]*wo_;
                    wn_ = ar[u]*wn_;
                    const double doi = (wn_ >= 0.0) ? dlo[u][t] : rb;
                    const double fi = j < L - 1 ? wo_ : 0.0;
                    const double fr = j < L - 1 ? doi*wn_ : 0.0;
                    double d;
                    if (j == 0) d = (fi >= 0.0) ? 1.0 : 1.0 - qq*fi;
                    else if (j == L - 1) {
                        d = (fi_m[t] >= 0.0) ? 1.0 + qq*fi_m[t] : 1.0;
                        if (t < 4) d += iw*sv*dt*a_sfc/vv;
                    } else {
                        d = 1.0;
                        if (fi_m[t] >= 0.0) d += qq*fi_m[t];
                        if (fi < 0.0) d -= qq*fi;
                    }
                    double r = rb + dt*tb[u][t];
                    if (j == 0) r += ew*dt*fr/vv;
                    else if (j == L - 1) {
                        r += -ew*dt*fr_m[t]/vv;
                        if (t < 4) r += -ew*sv*dt*r_low[t]*a_sfc/vv;
                    } else r += ew*dt*(-fr_m[t] + fr)/vv;
                    /* off-diagonals of interface j (column_solvers.c:475-488) */
                    double o = 0.0;
                    const bool up = fi >= 0.0;
                    if (j < L - 1) o = up ? (-qq)*fi : qlo[u]*fi;
                    const int idx = j*NSYS + t*NCOLS + col;
                    if (valid) { ce[idx] = o; fl[idx] = up; dd[idx] = d; rr[idx] = r; }
                    fi_m[t] = fi;
                    fr_m[t] = fr;
                }
            }
        }
    }
    __syncthreads();
    /* phase 2: Thomas algorithm (column_solvers.c:641-667), lane = system */
    if (threadIdx.x < NSYS && i0 + (int)(threadIdx.x%NCOLS) < nsh)
        thomas_lds_ce<NSYS, L>(ce, fl, dd, rr, L, threadIdx.x);
    __syncthreads();
    /* phase 3: clipped solution (column_solvers.c:616-637) */
    for (int idx = threadIdx.x; idx < L*NSYS; idx += NT) {
        const int j = idx/NSYS, sys = idx - j*NSYS, t = sys/NCOLS, col = sys - t*NCOLS, i = i0 + col;
        if (i >= nsh) continue;
        double x = rr[idx];
        if (x < 0.0) x = 0.0;
        sn.rho[(i64)kof(t)*S + i + (i64)j*nsh] = x;
    }
}

/* The same tracer solve with one (column, tracer) per lane, the rows
   generated and eliminated in one downward sweep through registers (the
   layout k_ver_waves uses); selected with GAME_GPU_GD_STREAMING for
   measurement against the LDS-staged kernel above. */
template <int LT>
__global__ void __launch_bounds__(128) k_gen_densities_stream(DevGrid G, DevState so, DevState sn, DevState td,
                                                              StepConsts K)
{
    const unsigned ntr = (unsigned)G.n_const - 1, blk = xcd_block().x;
    const int tr = (int)(blk%ntr);
    const int ncc = G.n_cond;
    const int k = tr < ncc ? tr : tr + 1;
    const int i = (int)(blk/ntr)*128 + threadIdx.x;
    if (i >= G.nsh) return;
    const int L = LT > 0 ? LT : G.L;
    const int nsh = G.nsh;
    const double iw = 0.5, ew = 1.0 - iw, dt = K.dt;
    const double *vol = G.vol, *area = G.area;
    const double *rho_o = so.rho + (i64)k*G.s, *trho = td.rho + (i64)k*G.s;
    double vsink = 0.0;
    if (k < ncc/4) vsink = K.snow_v;
    else if (k < ncc/2) vsink = K.rain_v;
    else if (k < ncc) vsink = K.cloud_v;
    double ep[LT > 0 ? LT : GAME_MAX_LAYERS], x[LT > 0 ? LT : GAME_MAX_LAYERS];
    double fi_m = 0.0, fr_m = 0.0, c_prev = 0.0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const i64 b = i + (i64)j*nsh;
        double fi = 0.0, fr = 0.0;
        if (j < L - 1) tracer_iface_flux(G, so, sn, K, rho_o, k, i, j, fi, fr);
        double d;
        if (j == 0) d = (fi >= 0.0) ? 1.0 : 1.0 - iw*dt/vol[b]*fi;
        else if (j == L - 1) {
            d = (fi_m >= 0.0) ? 1.0 + iw*dt/vol[b]*fi_m : 1.0;
            if (k < ncc) d += iw*vsink*dt*area[i + G.v - nsh]/vol[b];
        } else {
            d = 1.0;
            if (fi_m >= 0.0) d += iw*dt/vol[b]*fi_m;
            if (fi < 0.0) d -= iw*dt/vol[b]*fi;
        }
        double r = rho_o[b] + dt*trho[b];
        if (j == 0) r += ew*dt*fr/vol[b];
        else if (j == L - 1) {
            r += -ew*dt*fr_m/vol[b];
            if (k < ncc) r += -ew*vsink*dt*rho_o[i + G.s - nsh]*area[i + G.v - nsh]/vol[b];
        } else r += ew*dt*(-fr_m + fr)/vol[b];
        double c = 0.0, e = 0.0;
        if (j < L - 1) {
            if (fi >= 0.0) e = -iw*dt/vol[b]*fi;
            else c = iw*dt/vol[b + nsh]*fi;
        }
        if (j == 0) {
            ep[0] = e/d;
            x[0] = r/d;
        } else {
            const double den = d - ep[j - 1]*c_prev;
            if (j < L - 1) ep[j] = e/den;
            x[j] = (r - x[j - 1]*c_prev)/den;
        }
        c_prev = c;
        fi_m = fi;
        fr_m = fr;
    }
#pragma unroll
    for (int j = L - 2; j >= 0; --j) x[j] = x[j] - ep[j]*x[j + 1];
    double *rho_n = sn.rho + (i64)k*G.s;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        double v = x[j];
        if (v < 0.0) v = 0.0;
        rho_n[i + (i64)j*nsh] = v;
    }
}

/* sediment heat of the condensates at rk 0 (column_solvers.c:414-465 and the
   sh updates of :500-611): one thread per grid box adds the contributions of
   the condensates in the reference's constituent order, starting from 0 */
template <typename TW, int NCC>
__global__ void __launch_bounds__(TPB) k_sed_heat(DevGrid G, DevState so, DevState sn, const TW *T, const TW *cpc,
                                                  TW *sed_heat, StepConsts K)
{
    const i64 b = (i64)xcd_block().x*TPB + threadIdx.x;
    if (b >= G.s) return;
    /* NCC > 0: the condensate count at compile time, so both loops unroll */
    const int nsh = G.nsh, vpl = G.vpl, L = G.L, ncc = NCC > 0 ? NCC : G.n_cond;
    const int j = (int)(b/nsh), i = (int)(b - (i64)j*nsh);
    const double *vol = G.vol, *area = G.area;
    double sh = 0.0;
#pragma unroll
    for (int k = 0; k < ncc; ++k) {
        const double *rho_o = so.rho + (i64)k*G.s;
        double vsink;
        if (k < ncc/4) vsink = K.snow_v;
        else if (k < ncc/2) vsink = K.rain_v;
        else vsink = K.cloud_v;
        double fe = 0.0, fe_m = 0.0;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            const int jj = pass == 0 ? j : j - 1;
            if (jj < 0 || jj > L - 2) continue;
            double fr = sn.wind[i + (i64)(jj + 1)*vpl];
            fr -= vsink;
            fr = area[i + (i64)(jj + 1)*vpl]*fr;
            const i64 bb = i + (i64)jj*nsh;
            /* the upstream box is chosen after both are loaded */
            const bool lo = fr >= 0.0;
            const double d0 = rho_o[bb + nsh], d1 = rho_o[bb];
            const double t0 = (double)T[bb + nsh], t1 = (double)T[bb];
            const double c0 = (double)cpc[(i64)k*G.s + bb + nsh], c1 = (double)cpc[(i64)k*G.s + bb];
            const double doi = lo ? d0 : d1, toi = lo ? t0 : t1, cpi = lo ? c0 : c1;
            fr = doi*fr;
            const double f = cpi*toi*fr;
            if (pass == 0) fe = f; else fe_m = f;
        }
        if (j == 0) sh += fe/vol[b];
        else if (j == L - 1) {
            const i64 low = i + G.s - nsh, sfc = i + G.v - nsh;
            sh += -fe_m/vol[b];
            const double tl = (double)T[low];
            sh += -vsink*tl*(double)cpc[(i64)k*G.s + low]*rho_o[low]*area[sfc]/vol[b];
        } else sh += (-fe_m + fe)/vol[b];
    }
    sed_heat[b] = (TW)sh;
}

/* The same sums with one thread per column walking down its layers: the flux
   through each interface is formed once (by the reference's arithmetic) and
   carried to the box below, and the upstream box's density, temperature and
   c_p are loaded once per box and kept for the next interface.  The helper's
   restrict-qualified pointers let the loads of later layers move ahead of the
   stores of earlier ones. */
template <typename TW, int NCC, int LT>
__device__ __forceinline__ void sed_heat_column(const DevGrid &G, const StepConsts &K, int i,
                                                const double *__restrict__ rho_o, const double *__restrict__ wind_n,
                                                const double *__restrict__ area, const double *__restrict__ vol,
                                                const TW *__restrict__ T, const TW *__restrict__ cpc,
                                                TW *__restrict__ sed_heat)
{
    const int nsh = G.nsh, vpl = G.vpl, L = LT;
    const i64 S = G.s;
    double vsink[NCC];
#pragma unroll
    for (int k = 0; k < NCC; ++k) vsink[k] = k < NCC/4 ? K.snow_v : (k < NCC/2 ? K.rain_v : K.cloud_v);
    /* values of the box above the current interface */
    double d_up[NCC], c_up[NCC], t_up = (double)T[i];
#pragma unroll
    for (int k = 0; k < NCC; ++k) { d_up[k] = rho_o[k*S + i]; c_up[k] = (double)cpc[k*S + i]; }
    double fe_m[NCC];
#pragma unroll
    for (int k = 0; k < NCC; ++k) fe_m[k] = 0.0;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const i64 b = i + (i64)j*nsh;
        const double v = vol[b];
        double fe[NCC], d_lo[NCC], c_lo[NCC], t_lo = 0.0;
#pragma unroll
        for (int k = 0; k < NCC; ++k) { fe[k] = 0.0; d_lo[k] = 0.0; c_lo[k] = 0.0; }
        if (j <= L - 2) {
            /* interface j, below box j (column_solvers.c:414-465) */
            const double wn = wind_n[i + (i64)(j + 1)*vpl], a = area[i + (i64)(j + 1)*vpl];
            t_lo = (double)T[b + nsh];
#pragma unroll
            for (int k = 0; k < NCC; ++k) {
                d_lo[k] = rho_o[k*S + b + nsh];
                c_lo[k] = (double)cpc[k*S + b + nsh];
                double fr = wn;
                fr -= vsink[k];
                fr = a*fr;
                const bool lo = fr >= 0.0;
                const double doi = lo ? d_lo[k] : d_up[k], toi = lo ? t_lo : t_up, cpi = lo ? c_lo[k] : c_up[k];
                fr = doi*fr;
                fe[k] = cpi*toi*fr;
            }
        }
        double sh = 0.0;
#pragma unroll
        for (int k = 0; k < NCC; ++k) {
            if (j == 0) sh += fe[k]/v;
            else if (j == L - 1) {
                const i64 sfc = i + G.v - nsh;
                sh += -fe_m[k]/v;
                /* the lowest box: t_up, d_up, c_up hold its values */
                sh += -vsink[k]*t_up*c_up[k]*d_up[k]*area[sfc]/v;
            } else sh += (-fe_m[k] + fe[k])/v;
            fe_m[k] = fe[k];
            d_up[k] = d_lo[k];
            c_up[k] = c_lo[k];
        }
        t_up = t_lo;
        sed_heat[b] = (TW)sh;
    }
}

template <typename TW, int NCC, int LT>
__global__ void __launch_bounds__(TPB) k_sed_heat_col(DevGrid G, DevState so, DevState sn, const TW *T, const TW *cpc,
                                                      TW *sed_heat, StepConsts K)
{
    const int i = xcd_block().x*TPB + threadIdx.x;
    if (i >= G.nsh) return;
    sed_heat_column<TW, NCC, LT>(G, K, i, so.rho, sn.wind, G.area, G.vol, T, cpc, sed_heat);
}

/* phase_trans.c:16-311 (surface phase transitions off), one thread per grid box */
__device__ __forceinline__ double sat_water(double t)
{   /* dictionary.c:287-333 */
    double tc = t - GC_T_0, r;
    if (tc > 100.0) tc = 100.0;
    if (tc > 0.0) r = exp(34.494 - 4924.99/(tc + 237.1))/pow(tc + 105.0, 1.57);
    else {
        if (tc < -50.0) tc = -50.0;
        r = 6.107799961
            + 4.436518521e-1*tc
            + 1.428945805e-2*(tc*tc)
            + 2.650648471e-4*pow(tc, 3.0)
            + 3.031240396e-6*pow(tc, 4.0)
            + 2.034080948e-8*pow(tc, 5.0)
            + 6.136820929e-11*pow(tc, 6.0);
        r = 100.0*r;
    }
    return r;
}

__device__ __forceinline__ double sat_ice(double t)
{   /* dictionary.c:361-385 */
    double tc = t - GC_T_0;
    if (tc < -80.0) tc = -80.0;
    if (tc > 0.0) tc = 0.0;
    double a = tc + 868.0;
    return exp(43.494 - 6545.8/(tc + 278.0))/(a*a);
}

__device__ __forceinline__ double dsat_water(double t)
{   /* dictionary.c:335-359 */
    double tc = t - GC_T_0;
    if (tc > 100.0) tc = 100.0;
    if (tc < 0.0) tc = 0.0;
    double a = tc + 237.1;
    return sat_water(t)*(4924.99/(a*a) - 1.57/(tc + 105.0));
}

__device__ __forceinline__ double dsat_ice(double t)
{   /* dictionary.c:387-412 */
    double tc = t - GC_T_0;
    if (tc < -80.0) tc = -80.0;
    if (tc > 0.0) tc = 0.0;
    double a = tc + 278.0;
    return sat_ice(t)*(6545.8/(a*a) - 2.0/(tc + 868.0));
}

__device__ __forceinline__ double enthalpy_evaporation(double t)
{   /* dictionary.c:222-258 */
    if (t < GC_T_0) {
        if (t < 30.0) t = 30.0;
        return (56579.0 - 42.212*t + exp(0.1149*(281.6 - t)))/GC_M_V;
    }
    if (t > GC_T_0 + 100.0) t = GC_T_0 + 100.0;
    return 3151378.0 - 2386.0*t;
}

__device__ __forceinline__ double enthalpy_sublimation(double t)
{   /* dictionary.c:260-285 */
    if (t < 30.0) t = 30.0;
    if (t > GC_T_0) t = GC_T_0;
    double q = t/123.75;
    return (46782.5 + 35.8925*t - 0.07414*(t*t) + 541.5*exp(-(q*q)))/GC_M_V;
}

__device__ __forceinline__ double phase_trans_heat(int dir, double t)
{   /* dictionary.c:192-220 */
    if (dir == 0) return enthalpy_evaporation(t);
    if (dir == 1) return enthalpy_sublimation(t);
    return enthalpy_sublimation(t) - enthalpy_evaporation(t);
}

/* the saturation adjustment of one grid box (phase_trans.c:16-311, surface
   phase transitions off) on loaded values: densities r[0..5] (snow, rain,
   cloud ice, cloud water, gas, vapour), temperature T and the condensates'
   c_p cp[0..3] as stored; rates out[0..4] and the heating rate */
__device__ __forceinline__ void phase_trans_point(const double *r, double T, const double *cp, double dt, double *out,
                                                  double &heat_out)
{
    const double rk0 = r[0], rk1 = r[1], rk2 = r[2], rk3 = r[3], rk4 = r[4], rk5 = r[5];
    const double mcwc = 0.2e-3;
    double wvp = rk5*GC_R_V*T;
    double dp = (rk4 - rk5)*GC_R_D*T;
    double ap = dp + wvp;
    const int warm = T >= GC_T_0;
    /* saturation pressure, enhancement factor, the phase's latent heat and the
       saturation pressure's derivative, over water or over ice: a branch, not
       a select, so a wave whose boxes are all above or all below freezing
       evaluates one set of exponentials */
    double sp, ef, pth, ds;
    if (warm) {
        sp = sat_water(T);
        ef = 1.00071*exp(0.000000045*ap);
        pth = phase_trans_heat(0, T);
        ds = dsat_water(T);
    } else {
        sp = sat_ice(T);
        ef = 0.99882*exp(0.00000008*ap);
        pth = phase_trans_heat(1, T);
        ds = dsat_ice(T);
    }
    sp = ef*sp;
    /* c_v_mass_weighted_air with the condensates' c_p (derived_quantities.c:97-115) */
    double cvm = 0.0;
    for (int k = 0; k < 4; ++k) cvm += r[k]*cp[k];
    cvm += rk4*GC_C_D_V;
    cvm += rk5*GC_C_V_V;
    double a = -GC_R_V*pth/cvm;
    double b = GC_R_V*T - GC_R_V*rk5*pth/cvm + ef*ds*pth/cvm;
    double cc = wvp - sp;
    double p = b/a, q = cc/a;
    /* the reference's pow(x, 0.5) (phase_trans.c:92-99): the square root,
       which IEEE rounds correctly (glibc's pow is within 0.52 ulp of it);
       the discriminant is never -0.0 or -inf, where the two differ */
    double dd = -0.5*p - sqrt(0.25*(p*p) - q);
    double r0, r1, r2, r3, r4, heat;
    if (sp >= wvp) {
        if (warm) {
            r2 = -rk2/dt;
            double ptd = fmin(rk3, dd);
            r4 = ptd/dt;
            r3 = rk2/dt - ptd/dt;
            heat = r2*phase_trans_heat(2, T) - ptd*phase_trans_heat(0, T)/dt;
        } else {
            r3 = -rk3/dt;
            double ptd = fmin(rk2, dd);
            r4 = ptd/dt;
            r2 = rk3/dt - ptd/dt;
            heat = -r3*phase_trans_heat(2, T) - ptd*phase_trans_heat(1, T)/dt;
        }
    } else {
        if (warm) {
            r2 = -rk2/dt;
            r4 = dd/dt;
            r3 = rk2/dt - dd/dt;
            heat = r2*phase_trans_heat(2, T) - dd*phase_trans_heat(0, T)/dt;
        } else {
            r3 = -rk3/dt;
            r4 = dd/dt;
            r2 = rk3/dt - dd/dt;
            heat = -r3*phase_trans_heat(2, T) - dd*phase_trans_heat(1, T)/dt;
        }
    }
    r0 = 0.0;
    r1 = 0.0;
    if (T < GC_T_0) {
        r0 = fmax(rk2 - mcwc*rk4, 0.0)/1000.0;
        r2 -= r0;
    } else if (T >= GC_T_0) {
        r1 = fmax(rk3 - mcwc*rk4, 0.0)/1000.0;
        r3 -= r1;
    }
    if (T >= GC_T_0 && rk0 > 0.0) {
        r0 = -rk0/dt;
        r1 -= r0;
        heat += r0*phase_trans_heat(2, T);
    }
    if (T < GC_T_0 && rk1 > 0.0) {
        r1 = -rk1/dt;
        r0 -= r1;
        heat += -r1*phase_trans_heat(2, T);
    }
    out[0] = r0; out[1] = r1; out[2] = r2; out[3] = r3; out[4] = r4;
    heat_out = heat;
}

template <typename TW>
__global__ void __launch_bounds__(TPB) k_phase_trans(DevGrid G, DevState s, const TW *temp, const TW *cpc, TW *ptr,
                                                     TW *phr, double dt)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= G.s) return;
    const i64 S = G.s;
    double r[6], cp[4], out[5], heat;
    for (int k = 0; k < 6; ++k) r[k] = s.rho[k*S + i];
    for (int k = 0; k < 4; ++k) cp[k] = (double)cpc[k*S + i];
    phase_trans_point(r, (double)temp[i], cp, dt, out, heat);
    for (int k = 0; k < 5; ++k) ptr[k*S + i] = (TW)out[k];
    phr[i] = (TW)heat;
}

/* temperature_diagnostics, the condensates' c_p, the gas fraction and the
   saturation adjustment of a moist grid box in one pass: the adjustment
   takes the temperature and c_p as stored (rounded to TW), so its results
   are those of k_temperature followed by k_phase_trans */
template <typename TW>
__global__ void __launch_bounds__(TPB) k_temp_phase(DevGrid G, DevState s, TW *T, TW *cpc, double *gf, TW *ptr,
                                                    TW *phr, double dt)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= G.s) return;
    const i64 S = G.s;
    double r[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = s.rho[k*S + i];
    const double thbg = G.thbg[i], tp = s.tp[i], exbg = G.exbg[i], ep = s.ep[i];
    /* density_total's left-to-right sum from 0 (derived_quantities.c:83-95)
       over the six densities already loaded */
    double rt = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) rt += r[k];
    gf[i] = r[4]/rt;
    double t = (thbg + tp)*(exbg + ep);
    t = t/(1.0 + r[5]/r[4]*(GC_M_D/GC_M_V - 1.0));
    T[i] = (TW)t;
    const double tt = (double)(TW)t;
    /* dictionary.c:111-190: snow and cloud ice share the ice formula, rain and
       cloud water the liquid one; each is evaluated once */
    const TW c_ice = (TW)c_p_ice(tt), c_liq = (TW)c_p_water(tt);
    double cp[4];
    for (int k = 0; k < 4; ++k) {
        const TW c = (k & 1) == 0 ? c_ice : c_liq;
        cpc[k*S + i] = c;
        cp[k] = (double)c;
    }
    double out[5], heat;
    phase_trans_point(r, tt, cp, dt, out, heat);
    for (int k = 0; k < 5; ++k) ptr[k*S + i] = (TW)out[k];
    phr[i] = (TW)heat;
}

/* linear_combine_two_states.c:14-43 */
__global__ void __launch_bounds__(TPB) k_lincomb_scalars(DevGrid G, DevState a, DevState b, DevState o, double c0,
                                                         double c1)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= G.s) return;
    for (int k = 0; k < G.n_const; ++k) o.rho[k*G.s + i] = c0*a.rho[k*G.s + i] + c1*b.rho[k*G.s + i];
    double rt = c0*a.rt[i] + c1*b.rt[i];
    o.rt[i] = rt;
    o.tp[i] = rt/o.rho[(i64)G.n_cond*G.s + i] - G.thbg[i];
    o.ep[i] = c0*a.ep[i] + c1*b.ep[i];
}

__global__ void __launch_bounds__(TPB) k_lincomb_vector(i64 n, const double *a, const double *b, double *o, double c0,
                                                        double c1)
{
    i64 i = (i64)blockIdx.x*TPB + threadIdx.x;
    if (i >= n) return;
    o[i] = c0*a[i] + c1*b[i];
}

/* merge of two compensated partial sums (s, c) += (s2, c2): the sum of the
   leading parts with its exact rounding error (Neumaier's two-sum), the
   compensations added */
__device__ __forceinline__ void neumaier_merge(double &s, double &c, double s2, double c2)
{
    const double t = s + s2;
    double err;
    if (fabs(s) >= fabs(s2)) err = (s - t) + s2;
    else err = (s2 - t) + s;
    s = t;
    c = (c + c2) + err;
}

/* per-block compensated partial sums of x*volume (write_output.c:1659-1668 with
   Neumaier summation).  Each lane sums a grid-strided subset; the 64 lanes of
   a wavefront are merged by butterfly shuffles (six steps of __shfl_xor, no
   LDS), the four wavefronts of the block through LDS; the host adds the
   block partials, also compensated. */
template <typename TY>
__global__ void __launch_bounds__(TPB) k_integral_partials(i64 n, const double *x, const TY *y, const double *vol,
                                                           double *partials)
{
    __shared__ double ss[TPB/64], sc[TPB/64];
    double s = 0.0, c = 0.0;
    for (i64 i = (i64)blockIdx.x*TPB + threadIdx.x; i < n; i += (i64)gridDim.x*TPB) {
        double v = x[i]*vol[i];
        if (y) v = x[i]*(double)y[i]*vol[i];
        neumaier_merge(s, c, v, 0.0);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double s2 = __shfl_xor(s, off, 64), c2 = __shfl_xor(c, off, 64);
        neumaier_merge(s, c, s2, c2);
    }
    const int wv = threadIdx.x/64, lane = threadIdx.x%64;
    if (lane == 0) { ss[wv] = s; sc[wv] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = ss[0], Cc = sc[0];
        for (int k = 1; k < TPB/64; ++k) neumaier_merge(S, Cc, ss[k], sc[k]);
        partials[2*blockIdx.x] = S;
        partials[2*blockIdx.x + 1] = Cc;
    }
}

/* ------------------------------------------------------------------ */
/* host-side launchers                                                  */
/* ------------------------------------------------------------------ */
/* horizontal operators and tendencies; TW is the work-array type */
namespace hor {

#define FLAT_DISPATCH(NAME, KERNEL, GRIDDIM, ...)                                                 \
    do {                                                                                      \
        if (G.flat) LAUNCH(NAME, (KERNEL<true, TW>), GRIDDIM, dim3(TPB), st, __VA_ARGS__);    \
        else LAUNCH(NAME, (KERNEL<false, TW>), GRIDDIM, dim3(TPB), st, __VA_ARGS__);          \
    } while (0)
#define FLAT_DISPATCH_NC(NAME, KERNEL, NC, GRIDDIM, ...)                                          \
    do {                                                                                      \
        if (G.flat) LAUNCH(NAME, (KERNEL<true, TW, NC>), GRIDDIM, dim3(TPB), st, __VA_ARGS__);\
        else LAUNCH(NAME, (KERNEL<false, TW, NC>), GRIDDIM, dim3(TPB), st, __VA_ARGS__);      \
    } while (0)

template <typename TW>
static void pgrad_t(const DevGrid &G, const DevState &s, const DevWorkT<TW> &W, int first, double itw, hipStream_t st)
{
    /* the weights of the momentum tendency, as vector_tend_t forms them */
    const double chw = 0.5 + itw, ohw = 1.0 - chw, cvw = 1.0 - itw;
    if (G.n_const == 6) {
        LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 6>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, first, 1,
               (const double *)W.gas_frac, cvw);
        FLAT_DISPATCH_NC("k_pgrad_e", k_pgrad_e, 6, GRID2(G.nvh, G.split_edge), G, s, W, first,
                         (const double *)W.gas_frac, chw, ohw);
    } else if (G.n_const == 1) {
        LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 1>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, first, 1,
               (const double *)W.gas_frac, cvw);
        FLAT_DISPATCH_NC("k_pgrad_e", k_pgrad_e, 1, GRID2(G.nvh, G.split_edge), G, s, W, first,
                         (const double *)W.gas_frac, chw, ohw);
    } else {
        LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 0>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, first, 1,
               (const double *)W.gas_frac, cvw);
        FLAT_DISPATCH_NC("k_pgrad_e", k_pgrad_e, 0, GRID2(G.nvh, G.split_edge), G, s, W, first,
                         (const double *)W.gas_frac, chw, ohw);
    }
}

template <typename TW>
static void pgc_v_t(const DevGrid &G, const DevState &s, const DevWorkT<TW> &W, hipStream_t st)
{
    /* the constituent count as a compile-time constant where it is 6 or 1, so
       the density sums unroll and their loads issue together */
    if (G.n_const == 6)
        LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 6>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, 0, 0, nullptr, 0.0);
    else if (G.n_const == 1)
        LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 1>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, 0, 0, nullptr, 0.0);
    else LAUNCH("k_pgrad_w", (k_pgrad_w<TW, 0>), GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, s, W, 0, 0, nullptr, 0.0);
}

template <typename TW>
static void advection_pv_t(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkT<TW> &W,
                           hipStream_t st)
{
    LAUNCH("k_mass_flux", k_mass_flux<TW>, GRID2(G.nvh > G.nsh ? G.nvh : G.nsh, G.split_edge), dim3(TPB), st, G,
           rho_gas, wind, W.flux);
    LAUNCH("k_rel_vort_tri", k_rel_vort_tri<TW>, GRID1(G.nt), dim3(TPB), st, G, wind, W.rvt);
    FLAT_DISPATCH("k_pot_vort", k_pot_vort, GRID2(G.nvh, G.split_edge), G, wind, rho_gas, (const TW *)W.rvt, W.pv);
    LAUNCH("k_vorticity_flux_e", k_vorticity_flux_e<TW>, GRID2(G.nvh, G.split_edge), dim3(TPB), st, G, (const TW *)W.flux,
           (const TW *)W.pv, W.pvt);
    LAUNCH("k_vorticity_flux_w", k_vorticity_flux_w<TW>, GRID2(G.nsh, G.split_cell), dim3(TPB), st, G,
           (const TW *)W.flux, (const TW *)W.pv, W.pvt);
}

template <typename TW>
static void advection_kinetic_t(const DevGrid &G, const double *wind, const DevWorkT<TW> &W, hipStream_t st)
{
    LAUNCH("k_inner_product", k_inner_product<TW>, GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, wind, wind, W.vsq);
    FLAT_DISPATCH("k_grad_e", k_grad_e, GRID2(G.nvh, G.split_edge), G, (const TW *)W.vsq, W.vsq_grad);
    LAUNCH("k_grad_w", k_grad_w<TW>, GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, (const TW *)W.vsq, W.vsq_grad);
}

template <typename TW>
static void advection_t(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkT<TW> &W,
                        hipStream_t st)
{
    advection_pv_t<TW>(G, wind, rho_gas, W, st);
    advection_kinetic_t<TW>(G, wind, W, st);
}

template <typename TW>
static void vector_tend_t(const DevGrid &G, const DevWorkT<TW> &W, TW *tend_w, const double *w_old, double *w_new,
                          double new_w, double itw, double dt, hipStream_t st)
{
    double chw = 0.5 + itw, ohw = 1.0 - chw, cvw = 1.0 - itw;
    LAUNCH("k_vector_tend", k_vector_tend<TW>, GRID1(G.v), dim3(TPB), st, G, W, tend_w, w_old, w_new, new_w, chw, ohw,
           cvw, dt);
}

/* the precision of FP32 storage on an FP64 array (a probe of the mixed form) */
__global__ void __launch_bounds__(TPB) k_round_f32(double *x, i64 n)
{
    const i64 t = (i64)blockIdx.x*TPB + threadIdx.x;
    if (t < n) x[t] = (double)(float)x[t];
}

template <typename TW>
static void scalar_tend_t(const DevGrid &G, const DevState &s, const DevState &td, const DevWorkT<TW> &W, int rk,
                          hipStream_t st);

template <typename TW>
static void scalar_tend_probe(const DevGrid &G, const DevState &td, hipStream_t st)
{
    if (sizeof(TW) != sizeof(float) || !G.round_tendencies) return;
    if (G.round_tendencies & 1)
        LAUNCH("k_round_f32", k_round_f32, GRID1((i64)G.n_const*G.s), dim3(TPB), st, td.rho, (i64)G.n_const*G.s);
    if (G.round_tendencies & 2) LAUNCH("k_round_f32", k_round_f32, GRID1(G.s), dim3(TPB), st, td.rt, G.s);
}

template <typename TW>
static void scalar_tend_t(const DevGrid &G, const DevState &s, const DevState &td, const DevWorkT<TW> &W, int rk,
                          hipStream_t st)
{
    if (!G.flat)
        /* chunks of three terrain levels (res 6 x 26, 23 terrain layers: 68 us
           per launch; chunks of 2, 4 and 8 levels 77, 73 and 82 us).  Loading
           each layer before the one above it is summed took the kernel's time
           in the mountain window's gradient from 5.98 to 5.46 ms (176
           registers, two waves per SIMD); chunks of six levels then 6.87 ms,
           and the weights from ipw's interleaved
           layout (eight per box) 5.68 ms */
        LAUNCH("k_vcc", k_vcc, GRID2(G.nsh, (G.n_oro + 2)/3), dim3(TPB), st, G, s, W.vcc);
    if (G.n_const == 6) {
        if (G.flat) LAUNCH("k_scalar_tend", (k_scalar_tend<true, TW, 6>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                           W, td.rho, td.rt, rk);
        else LAUNCH("k_scalar_tend", (k_scalar_tend<false, TW, 6>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                    W, td.rho, td.rt, rk);
    } else if (G.n_const == 1) {
        if (G.flat) LAUNCH("k_scalar_tend", (k_scalar_tend<true, TW, 1>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                           W, td.rho, td.rt, rk);
        else LAUNCH("k_scalar_tend", (k_scalar_tend<false, TW, 1>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                    W, td.rho, td.rt, rk);
    } else {
        if (G.flat) LAUNCH("k_scalar_tend", (k_scalar_tend<true, TW, 0>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                           W, td.rho, td.rt, rk);
        else LAUNCH("k_scalar_tend", (k_scalar_tend<false, TW, 0>), GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, s,
                    W, td.rho, td.rt, rk);
    }
    scalar_tend_probe<TW>(G, td, st);
}

void pgrad(const DevGrid &G, const DevState &s, const DevWork &W, int first, double itw, hipStream_t st)
{ pgrad_t<double>(G, s, W, first, itw, st); }
void pgrad(const DevGrid &G, const DevState &s, const DevWorkF &W, int first, double itw, hipStream_t st)
{ pgrad_t<float>(G, s, W, first, itw, st); }
void pgc_v(const DevGrid &G, const DevState &s, const DevWork &W, hipStream_t st) { pgc_v_t<double>(G, s, W, st); }
void pgc_v(const DevGrid &G, const DevState &s, const DevWorkF &W, hipStream_t st) { pgc_v_t<float>(G, s, W, st); }
void advection(const DevGrid &G, const double *wind, const double *rho_gas, const DevWork &W, hipStream_t st)
{ advection_t<double>(G, wind, rho_gas, W, st); }
void advection(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkF &W, hipStream_t st)
{ advection_t<float>(G, wind, rho_gas, W, st); }
void advection_pv(const DevGrid &G, const double *wind, const double *rho_gas, const DevWork &W, hipStream_t st)
{ advection_pv_t<double>(G, wind, rho_gas, W, st); }
void advection_pv(const DevGrid &G, const double *wind, const double *rho_gas, const DevWorkF &W, hipStream_t st)
{ advection_pv_t<float>(G, wind, rho_gas, W, st); }
void advection_kinetic(const DevGrid &G, const double *wind, const DevWork &W, hipStream_t st)
{ advection_kinetic_t<double>(G, wind, W, st); }
void advection_kinetic(const DevGrid &G, const double *wind, const DevWorkF &W, hipStream_t st)
{ advection_kinetic_t<float>(G, wind, W, st); }
void vector_tend(const DevGrid &G, const DevWork &W, double *tend_w, const double *w_old, double *w_new,
                 double new_w, double itw, double dt, hipStream_t st)
{ vector_tend_t<double>(G, W, tend_w, w_old, w_new, new_w, itw, dt, st); }
void vector_tend(const DevGrid &G, const DevWorkF &W, float *tend_w, const double *w_old, double *w_new,
                 double new_w, double itw, double dt, hipStream_t st)
{ vector_tend_t<float>(G, W, tend_w, w_old, w_new, new_w, itw, dt, st); }
void scalar_tend(const DevGrid &G, const DevState &s, const DevState &td, const DevWork &W, int rk, hipStream_t st)
{ scalar_tend_t<double>(G, s, td, W, rk, st); }
void scalar_tend(const DevGrid &G, const DevState &s, const DevState &td, const DevWorkF &W, int rk, hipStream_t st)
{ scalar_tend_t<float>(G, s, td, W, rk, st); }

void inner_product(const DevGrid &G, const double *a, const double *b, double *out, hipStream_t st)
{ LAUNCH("k_inner_product", k_inner_product<double>, GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, a, b, out); }

/* single operators (FP64), for the operator-level interface and its adjoints */
void grad(const DevGrid &G, const double *f, double *out, hipStream_t st)
{
    typedef double TW;
    FLAT_DISPATCH("k_grad_e", k_grad_e, GRID2(G.nvh, G.split_edge), G, f, out);
    LAUNCH("k_grad_w", k_grad_w<double>, GRID1((i64)(G.L - 1)*G.nsh), dim3(TPB), st, G, f, out);
}

void curl_triangles(const DevGrid &G, const double *wind, double *out, hipStream_t st)
{ LAUNCH("k_rel_vort_tri", k_rel_vort_tri<double>, GRID1(G.nt), dim3(TPB), st, G, wind, out); }

void vorticity_flux(const DevGrid &G, const double *F, const double *pv, double *out, hipStream_t st)
{
    LAUNCH("k_vorticity_flux_e", k_vorticity_flux_e<double>, GRID2(G.nvh, G.split_edge), dim3(TPB), st, G, F, pv, out);
    LAUNCH("k_vorticity_flux_w", k_vorticity_flux_w<double>, GRID2(G.nsh, G.split_cell), dim3(TPB), st, G, F, pv, out);
}

}  // namespace hor

/* the column solvers and point physics, shared by both precision forms (the
   work arrays they read or write are double in the FP64 form, float in the
   mixed form) */
namespace col {

void temperature(const DevGrid &G, const DevState &s, double *T, double *cpc, double *gf, hipStream_t st)
{ LAUNCH("k_temperature", k_temperature<double>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, gf); }
void temperature(const DevGrid &G, const DevState &s, float *T, float *cpc, double *gf, hipStream_t st)
{ LAUNCH("k_temperature", k_temperature<float>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, gf); }

void phase_trans(const DevGrid &G, const DevState &s, const double *T, const double *cpc, double *ptr, double *phr,
                 double dt, hipStream_t st)
{ LAUNCH("k_phase_trans", k_phase_trans<double>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, ptr, phr, dt); }
void phase_trans(const DevGrid &G, const DevState &s, const float *T, const float *cpc, float *ptr, float *phr,
                 double dt, hipStream_t st)
{ LAUNCH("k_phase_trans", k_phase_trans<float>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, ptr, phr, dt); }
void temperature_phase_trans(const DevGrid &G, const DevState &s, double *T, double *cpc, double *gf, double *ptr,
                             double *phr, double dt, hipStream_t st)
{ LAUNCH("k_temp_phase", k_temp_phase<double>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, gf, ptr, phr, dt); }
void temperature_phase_trans(const DevGrid &G, const DevState &s, float *T, float *cpc, double *gf, float *ptr,
                             float *phr, double dt, hipStream_t st)
{ LAUNCH("k_temp_phase", k_temp_phase<float>, GRID1(G.s), dim3(TPB), st, G, s, T, cpc, gf, ptr, phr, dt); }

template <typename TWIND>
static void ver_waves_t(const DevGrid &G, const DevState &so, const DevState &snr, const DevState &sn,
                        const DevState &td, const TWIND *td_wind, const StepConsts &K, int rk, hipStream_t st)
{
    if (G.vw_streaming) {
        dim3 g((G.nsh + 127)/128), b(128);
        if (G.L == 26) LAUNCH("k_ver_waves", (k_ver_waves<26, TWIND>), g, b, st, G, so, snr, sn, td, td_wind, K, rk);
        else if (G.L == 55)
            LAUNCH("k_ver_waves", (k_ver_waves<55, TWIND>), g, b, st, G, so, snr, sn, td, td_wind, K, rk);
        else LAUNCH("k_ver_waves", (k_ver_waves<0, TWIND>), g, b, st, G, so, snr, sn, td, td_wind, K, rk);
        return;
    }
    if ((G.L == 26 || G.L == 55) && !G.vw_rows) {
        const int ncol = vr_ncol(G.L);
        dim3 g((G.nsh + ncol - 1)/ncol), b(VR_THREADS);
        size_t lds = (size_t)4*(G.L - 1)*ncol*sizeof(double);   /* 50 KiB at 26 layers, 55 KiB at 55 */
        if (G.L == 26)
            LAUNCH_SHM("k_ver_waves", (k_ver_waves_run<26, TWIND>), g, b, lds, st, G, so, snr, sn, td, td_wind, K, rk);
        else
            LAUNCH_SHM("k_ver_waves", (k_ver_waves_run<55, TWIND>), g, b, lds, st, G, so, snr, sn, td, td_wind, K, rk);
        return;
    }
    dim3 g((G.nsh + VW_NCOL - 1)/VW_NCOL), b(VW_THREADS);
    size_t lds = (size_t)4*(G.L - 1)*VW_NCOL*sizeof(double);   /* 25 KiB at 26 layers, 63 KiB at 64 */
    if (G.L == 26)
        LAUNCH_SHM("k_ver_waves", (k_ver_waves_3p<26, TWIND>), g, b, lds, st, G, so, snr, sn, td, td_wind, K, rk);
    else if (G.L == 55)
        LAUNCH_SHM("k_ver_waves", (k_ver_waves_3p<55, TWIND>), g, b, lds, st, G, so, snr, sn, td, td_wind, K, rk);
    else LAUNCH_SHM("k_ver_waves", (k_ver_waves_3p<0, TWIND>), g, b, lds, st, G, so, snr, sn, td, td_wind, K, rk);
}

void ver_waves(const DevGrid &G, const DevState &so, const DevState &snr, const DevState &sn, const DevState &td,
               const StepConsts &K, int rk, hipStream_t st)
{ ver_waves_t<double>(G, so, snr, sn, td, td.wind, K, rk, st); }

void ver_waves(const DevGrid &G, const DevState &so, const DevState &snr, const DevState &sn, const DevState &td,
               const float *td_wind, const StepConsts &K, int rk, hipStream_t st)
{ ver_waves_t<float>(G, so, snr, sn, td, td_wind, K, rk, st); }

static void gen_densities_only(const DevGrid &G, const DevState &so, const DevState &sn, const DevState &td,
                               const StepConsts &K, hipStream_t st)
{
    if (G.gd_streaming) {
        dim3 g(((G.nsh + 127)/128)*(G.n_const - 1)), b(128);
        if (G.L == 26) LAUNCH("k_gen_densities", k_gen_densities_stream<26>, g, b, st, G, so, sn, td, K);
        else if (G.L == 55) LAUNCH("k_gen_densities", k_gen_densities_stream<55>, g, b, st, G, so, sn, td, K);
        else LAUNCH("k_gen_densities", k_gen_densities_stream<0>, g, b, st, G, so, sn, td, K);
        return;
    }
    if (G.n_const == 6 && G.n_cond == 4 && !G.gd_per_tracer && G.L == 26) {
        /* five tracers per block at 26 layers: 16 columns and 256 threads
           (51 KiB of LDS, three blocks per CU), res 7: 190 us against 200
           for one tracer per block; 8 columns and 128 threads 224 us, 12
           columns 196 us (256 threads) and 204 us (192).  At 55 layers 8
           columns (54 KiB, two blocks per CU) took 701 us against 516:
           fewer systems in flight for the longer serial elimination */
        constexpr int NCOLS = 16, NT = 256;
        dim3 g((G.nsh + NCOLS - 1)/NCOLS), b(NT);
        size_t lds = (size_t)3*26*5*NCOLS*sizeof(double) + (((size_t)26*5*NCOLS + 7) & ~(size_t)7);
        LAUNCH_SHM("k_gen_densities", (k_gen_densities_all<26, NCOLS, NT>), g, b, lds, st, G, so, sn, td, K);
        return;
    }
    const int ncol = gd_ncol(G.L == 26 ? 26 : (G.L == 55 ? 55 : 0));
    dim3 g(((G.nsh + ncol - 1)/ncol)*(G.n_const - 1)), b(GD_THREADS);
    /* 41 KiB at 26 layers, 43 KiB at 55, 50 KiB at 64 */
    size_t lds = (size_t)3*G.L*ncol*sizeof(double) + (((size_t)G.L*ncol + 7) & ~(size_t)7);
    if (G.L == 26) LAUNCH_SHM("k_gen_densities", k_gen_densities<26>, g, b, lds, st, G, so, sn, td, K);
    else if (G.L == 55) LAUNCH_SHM("k_gen_densities", k_gen_densities<55>, g, b, lds, st, G, so, sn, td, K);
    else LAUNCH_SHM("k_gen_densities", k_gen_densities<0>, g, b, lds, st, G, so, sn, td, K);
}

void gen_densities(const DevGrid &G, const DevState &so, const DevState &sn, const DevState &td, const double *T,
                   const double *cpc, double *sed_heat, const StepConsts &K, int rk, hipStream_t st)
{
    if (rk == 0 && G.n_cond > 0) {
        if (G.n_cond == 4 && G.L == 26 && !G.sh_per_box)
            LAUNCH("k_sed_heat", (k_sed_heat_col<double, 4, 26>), GRID1(G.nsh), dim3(TPB), st, G, so, sn, T, cpc, sed_heat,
                   K);
        else if (G.n_cond == 4 && G.L == 55 && !G.sh_per_box)
            LAUNCH("k_sed_heat", (k_sed_heat_col<double, 4, 55>), GRID1(G.nsh), dim3(TPB), st, G, so, sn, T, cpc, sed_heat,
                   K);
        else if (G.n_cond == 4)
            LAUNCH("k_sed_heat", (k_sed_heat<double, 4>), GRID1(G.s), dim3(TPB), st, G, so, sn, T, cpc, sed_heat, K);
        else LAUNCH("k_sed_heat", (k_sed_heat<double, 0>), GRID1(G.s), dim3(TPB), st, G, so, sn, T, cpc, sed_heat, K);
    }
    gen_densities_only(G, so, sn, td, K, st);
}

void gen_densities(const DevGrid &G, const DevState &so, const DevState &sn, const DevState &td, const float *T,
                   const float *cpc, float *sed_heat, const StepConsts &K, int rk, hipStream_t st)
{
    if (rk == 0 && G.n_cond > 0) {
        if (G.n_cond == 4 && G.L == 26 && !G.sh_per_box)
            LAUNCH("k_sed_heat", (k_sed_heat_col<float, 4, 26>), GRID1(G.nsh), dim3(TPB), st, G, so, sn, T, cpc, sed_heat,
                   K);
        else if (G.n_cond == 4 && G.L == 55 && !G.sh_per_box)
            LAUNCH("k_sed_heat", (k_sed_heat_col<float, 4, 55>), GRID1(G.nsh), dim3(TPB), st, G, so, sn, T, cpc, sed_heat,
                   K);
        else if (G.n_cond == 4)
            LAUNCH("k_sed_heat", (k_sed_heat<float, 4>), GRID1(G.s), dim3(TPB), st, G, so, sn, T, cpc, sed_heat, K);
        else LAUNCH("k_sed_heat", (k_sed_heat<float, 0>), GRID1(G.s), dim3(TPB), st, G, so, sn, T, cpc, sed_heat, K);
    }
    gen_densities_only(G, so, sn, td, K, st);
}

__global__ void __launch_bounds__(TPB) k_tracer_quotients(DevGrid G, double dt)
{
    const i64 b = (i64)blockIdx.x*TPB + threadIdx.x;
    if (b >= G.s) return;
    const double iw = 0.5;   /* the tracer solver's implicit weight */
    G.idv[b] = iw*dt/G.vol[b];
}

void tracer_quotients(const DevGrid &G, double dt, hipStream_t st)
{ LAUNCH("k_tracer_quotients", k_tracer_quotients, GRID1(G.s), dim3(TPB), st, G, dt); }

void linear_combine(const DevGrid &G, const DevState &a, const DevState &b, const DevState &o, double c0, double c1,
                    hipStream_t st)
{
    LAUNCH("k_lincomb_scalars", k_lincomb_scalars, GRID1(G.s), dim3(TPB), st, G, a, b, o, c0, c1);
    LAUNCH("k_lincomb_vector", k_lincomb_vector, GRID1(G.v), dim3(TPB), st, G.v, (const double *)a.wind, (const double *)b.wind,
                       o.wind, c0, c1);
}

int integral_partials(i64 n, const double *x, const double *y, const double *vol, double *partials, hipStream_t st)
{
    LAUNCH("k_integral_partials", k_integral_partials<double>, dim3(INTEGRAL_BLOCKS), dim3(TPB), st, n, x, y, vol,
           partials);
    return INTEGRAL_BLOCKS;
}

int integral_partials(i64 n, const double *x, const float *y, const double *vol, double *partials, hipStream_t st)
{
    LAUNCH("k_integral_partials", k_integral_partials<float>, dim3(INTEGRAL_BLOCKS), dim3(TPB), st, n, x, y, vol,
           partials);
    return INTEGRAL_BLOCKS;
}

}  // namespace col
