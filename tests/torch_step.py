"""The dry RKHEVI step in PyTorch (float64, CPU): a second restatement of the
reference's algorithm, differentiable by autograd.  TEST INFRASTRUCTURE ONLY:
its vector-Jacobian products are the reference against which the hand-written
device adjoints are checked, component by component and for a whole
checkpointed window.  Its forward is checked against the C oracle
(tests/test_torch_step.py).

Fields use split layouts: a scalar field is [L, NSH]; a vector field is the
pair (u [L, NVH] at the edges, w [L+1, NSH] at the levels); a curl field is
(tangential [L+1, NVH], rhombus [L, NVH]).

Covered (dry: one constituent, moisture off, every physics switch off):
  grad, gradient_operators.c:14-74 and averaging.c:14-39, 113-133
  divv_h, divergences.c:14-64 and averaging.c:41-95
  scalar_times_vector(_h/_v), multiplications.c:35-117
  inner_product, inner_product.c:13-43
  calc_pot_vort, vorticities.c:18-222 (with averaging.c:97-111)
  vorticity_flux, vorticity_flux.c:16-187
  manage_pressure_gradient, manage_pressure_gradient.c:17-72 (decel factor 1)
  vector_tendencies_expl, vector_tendencies_expl.c:18-133
  scalar_tendencies_expl, scalar_tendencies_expl.c:17-181 (gas only)
  three_band_solver_ver_waves, column_solvers.c:19-383
  manage_rkhevi, manage_rkhevi.c:20-100; the driver copy, coordinator.c:302-324
"""
import math

import numpy as np
import torch

C_D_P = 1005.0
C_D_V = 717.942189
R_D = 287.057811
P_0 = 100000.0
N_PENT = 12

T = torch.float64


def _t(a, dtype=T):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype)


def _i(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.long)


class TorchGrid:
    """The grid's tables and metrics as tensors in the split layouts."""

    def __init__(self, grid, config=None):
        d = grid.dims
        self.d = d
        L, nsh, nvh, nt, vpl, dvpl = d.n_layers, d.nsh, d.nvh, d.nt, d.vpl, d.dvpl
        self.L, self.nsh, self.nvh, self.nt = L, nsh, nvh, nt
        self.n_oro = d.n_oro_layers
        g, dg = grid.g, grid.dg
        ue = (d.nsh + np.arange(L)[:, None] * vpl + np.arange(nvh)[None, :])
        wl = (np.arange(L + 1)[:, None] * vpl + np.arange(nsh)[None, :])
        self.u_idx, self.w_idx = ue, wl
        split = lambda a: (_t(a[ue]), _t(a[wl]))
        self.nd_u, self.nd_w = split(g["normal_distance"])
        self.area_u, self.area_w = split(g["area"])
        self.zv_u, self.zv_w = split(g["z_vector"])
        self.slope_u = _t(g["slope"][ue])
        self.exbg_grad_u, self.exbg_grad_w = split(g["exner_bg_grad"])
        sc = lambda a: _t(np.asarray(a).reshape(L, nsh))
        self.vol, self.zs, self.thbg, self.exbg = sc(g["volume"]), sc(g["z_scalar"]), sc(g["theta_v_bg"]), sc(g["exner_bg"])
        self.ipw = _t(g["inner_product_weights"].reshape(L, nsh, 8))
        self.fr, self.to = _i(g["from_index"]), _i(g["to_index"])
        adj = g["adjacent_vector_indices_h"].reshape(nsh, 6).copy()
        sgn = g["adjacent_signs_h"].reshape(nsh, 6).astype(np.float64).copy()
        self.ne = np.where(np.arange(nsh) < N_PENT, 5, 6)
        self.adj = _i(adj)
        self.sgn = _t(sgn)                                  # pentagon slot 5: sign 0
        self.mask6 = _t((np.arange(6)[None, :] < self.ne[:, None]).astype(np.float64))
        self.ti = _i(g["trsk_indices"].reshape(nvh, 10))
        self.tm = _i(g["trsk_modified_curl_indices"].reshape(nvh, 10))
        self.tw = _t(g["trsk_weights"].reshape(nvh, 10))
        fr_np, to_np = g["from_index"], g["to_index"]
        # TRSK slots read: 0..3 and 5..8 beside a pentagon, else 0..4 and 5..9;
        # slots 2 and 7 of a hexagon side half-average PV with the edge's own
        used = np.ones((nvh, 10))
        used[fr_np < N_PENT, 4] = 0.0
        used[to_np < N_PENT, 9] = 0.0
        half = np.zeros((nvh, 10))
        half[fr_np >= N_PENT, 2] = 1.0
        half[to_np >= N_PENT, 7] = 1.0
        self.trsk_used, self.trsk_half = _t(used), _t(half)
        self.dri = _i(g["density_to_rhombi_indices"].reshape(nvh, 4))
        self.drw = _t(g["density_to_rhombi_weights"].reshape(nvh, 4))
        self.dfr, self.dto = _i(dg["from_index"]), _i(dg["to_index"])
        self.vit = _i(dg["vorticity_indices_triangles"].reshape(nt, 3))
        self.vst = _t(dg["vorticity_signs_triangles"].reshape(nt, 3).astype(np.float64))
        darea, dzv = dg["area"], dg["z_vector"]
        tri = nvh + np.arange(L)[:, None] * dvpl + np.arange(nt)[None, :]
        tan = np.arange(L + 1)[:, None] * dvpl + np.arange(nvh)[None, :]
        self.darea_tri, self.dzv_tri = _t(darea[tri]), _t(dzv[tri])
        self.darea_tan = _t(darea[tan])
        self.f_tan, self.f_rh = _t(dg["f_vec"][:nvh]), _t(dg["f_vec"][nvh:])
        self.radius = grid.gv.radius
        # the differences calc_pot_vort forms for its top and surface density
        # extrapolation, read as the reference's struct layout serves them
        # (z_vector indexed past its end into gravity_potential, gravity_m)
        zv_all = np.concatenate([g["z_vector"], g["gravity_potential"], g["gravity_m"]])
        S, e = d.s, np.arange(nvh)
        self.pv_t1 = _t(zv_all[S + e] - zv_all[S + vpl + e])
        self.pv_t2 = _t(g["z_vector"][0] - zv_all[S + e])
        self.pv_s1 = _t(zv_all[S + (L - 2) * vpl + e] - zv_all[S + (L - 1) * vpl + e])
        self.pv_s2 = _t(0.5 * (g["z_vector"][L * vpl + fr_np] + g["z_vector"][L * vpl + to_np])
                        - zv_all[S + (L - 1) * vpl + e])
        # Klemp damping of every interior interface (column_solvers.c:284-298)
        c = config
        dsh = (0.53 if c is None else c.damping_start_height_over_toa) * g["z_vector"][0]
        cmax = 0.25 if c is None else c.damping_coeff_max
        zad = g["z_vector"][wl[1:L]] - dsh
        damp = np.where(zad < 0.0, 0.0, cmax * np.sin(0.5 * np.pi * zad / (g["z_vector"][0] - dsh)) ** 2)
        self.damp = _t(damp)
        self.impl_weight = 0.75 if c is None else c.impl_thermo_weight
        self.oro0 = L - self.n_oro                       # first terrain-following layer

    # -- layout conversions --------------------------------------------------
    def split_v(self, x):
        x = _t(x) if not torch.is_tensor(x) else x
        return x[_i(self.u_idx)], x[_i(self.w_idx)]

    def join_v(self, u, w):
        out = torch.zeros(self.d.v, dtype=T)
        out = out.index_put((_i(self.u_idx).reshape(-1),), u.reshape(-1))
        return out.index_put((_i(self.w_idx).reshape(-1),), w.reshape(-1))

    def split_s(self, x):
        x = _t(x) if not torch.is_tensor(x) else x
        return x.reshape(self.L, self.nsh)

    def join_curl(self, tan, rh):
        L, nvh = self.L, self.nvh
        out = torch.zeros(self.d.curl, dtype=T)
        ti = (np.arange(L + 1)[:, None] * 2 * nvh + np.arange(nvh)[None, :]).reshape(-1)
        ri = (np.arange(L)[:, None] * 2 * nvh + nvh + np.arange(nvh)[None, :]).reshape(-1)
        out = out.index_put((_i(ti),), tan.reshape(-1))
        return out.index_put((_i(ri),), rh.reshape(-1))


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------
def remap_w(G, w):
    """averaging.c:14-39 for every edge of every layer: [L, NVH]"""
    ip6, ip7 = G.ipw[:, :, 6], G.ipw[:, :, 7]
    comp = ip6[:, G.fr] * w[:-1][:, G.fr] + ip6[:, G.to] * w[:-1][:, G.to]
    low = ip7[:, G.fr] * w[1:][:, G.fr] + ip7[:, G.to] * w[1:][:, G.to]
    mask = torch.ones(G.L, 1, dtype=T)
    mask[G.L - 1] = 0.0
    return 0.5 * (comp + mask * low)


def oro_mask(G):
    m = torch.zeros(G.L, 1, dtype=T)
    m[G.oro0:] = 1.0
    return m


def grad(G, f):
    """gradient_operators.c:56-74: S -> (u, w); top and surface w are 0"""
    u = (f[:, G.to] - f[:, G.fr]) / G.nd_u
    w = torch.zeros(G.L + 1, G.nsh, dtype=T)
    w = torch.cat([w[:1], (f[:-1] - f[1:]) / G.nd_w[1:G.L], w[:1]], 0)
    u = u - oro_mask(G) * G.slope_u * remap_w(G, w)
    return u, w


def horizontal_covariant(G, u, w):
    """averaging.c:97-111 for every edge of every layer"""
    return u + oro_mask(G) * G.slope_u * remap_w(G, w)


def vcc(G, u):
    """vertical_contravariant_corr (averaging.c:41-95) at every level 0..L, [L+1, NSH]
    (zero above the terrain-following levels and at the surface level)"""
    a = G.ipw[:, :, :6] * (G.slope_u * u)[:, G.adj] * G.mask6     # [L, NSH, 6]
    per_layer = -0.5 * a.sum(2)                                     # [L, NSH]
    out = torch.zeros(G.L + 1, G.nsh, dtype=T)
    rows = []
    for lev in range(G.L + 1):
        if lev < G.oro0 or lev == G.L:
            rows.append(out[lev])
        elif lev == G.oro0:
            rows.append(per_layer[lev])
        else:
            rows.append(per_layer[lev - 1] + per_layer[lev])
    return torch.stack(rows)


def divv_h(G, u):
    """divergences.c:14-64: horizontal field -> S"""
    comp_h = (u[:, G.adj] * G.sgn * G.area_u[:, G.adj]).sum(2)
    c = vcc(G, u)
    comp_v = c[:-1] * G.area_w[:-1] - c[1:] * G.area_w[1:]       # faces of layer l: levels l, l+1
    return (comp_h + comp_v) / G.vol


def stv_h(G, s, u):
    return 0.5 * (s[:, G.fr] + s[:, G.to]) * u


def stv_v(G, s, w):
    """interior levels only; top and surface keep w_out (the reference leaves them)"""
    mid = 0.5 * (s[:-1] + s[1:]) * w[1:G.L]
    return torch.cat([torch.zeros(1, G.nsh, dtype=T), mid, torch.zeros(1, G.nsh, dtype=T)], 0)


def inner_product(G, au, aw, bu, bw):
    """inner_product.c:13-43"""
    h = (G.ipw[:, :, :6] * (au * bu)[:, G.adj] * G.mask6).sum(2)
    return h + G.ipw[:, :, 6] * aw[:-1] * bw[:-1] + G.ipw[:, :, 7] * aw[1:] * bw[1:]


def rel_vort_triangles(G, u):
    """vorticities.c:167-222: [L, NT]"""
    rows = []
    for l in range(G.L):
        e = G.vit                                          # [NT, 3]
        vel = u[l][e]
        lrf = torch.ones_like(vel)
        if l >= G.oro0:
            zt = G.dzv_tri[l][:, None]
            ze = G.zv_u[l][e]
            lrf = (G.radius + zt) / (G.radius + ze)
            dz = zt - ze
            up_l = l - 1
            dn_l = l + 1 if l < G.L - 1 else l - 1
            use_up = (dz > 0) | (l == G.L - 1)
            u_nb = torch.where(use_up, u[up_l][e], u[dn_l][e])
            z_nb = torch.where(use_up, G.zv_u[up_l][e], G.zv_u[dn_l][e])
            vel = vel + dz * (vel - u_nb) / (ze - z_nb)
        rows.append((lrf * G.nd_u[l][e] * G.vst * vel).sum(1) / G.darea_tri[l])
    return torch.stack(rows)


def rel_vort(G, u, w):
    """vorticities.c:104-165: (tangential [L+1, NVH], rhombus [L, NVH])"""
    zt = rel_vort_triangles(G, u)
    af, at = G.darea_tri[:, G.dfr], G.darea_tri[:, G.dto]
    rh = (af * zt[:, G.dfr] + at * zt[:, G.dto]) / (af + at)
    cov = horizontal_covariant(G, u, w)                     # [L, NVH]
    nd_wf, nd_wt = G.nd_w[:, G.fr], G.nd_w[:, G.to]        # [L+1, NVH]
    wf, wt = w[:, G.fr], w[:, G.to]
    rows = [None]
    for lev in range(1, G.L + 1):
        if lev == G.L:
            rows.append(1.0 / G.darea_tan[lev] * G.nd_u[lev - 1] * cov[lev - 1])
        else:
            rows.append(1.0 / G.darea_tan[lev] * (-G.nd_u[lev] * cov[lev] + nd_wf[lev] * wf[lev]
                                                  + G.nd_u[lev - 1] * cov[lev - 1] - nd_wt[lev] * wt[lev]))
    rows[0] = rows[1]
    return torch.stack(rows), rh


def half_level_density(G, rho):
    """the densities calc_pot_vort divides the tangential PV by (vorticities.c:44-80): [L+1, NVH]"""
    em = 0.5 * (rho[:, G.fr] + rho[:, G.to])               # [L, NVH]
    top = em[0] + (em[0] - em[1]) / G.pv_t1 * G.pv_t2
    sfc = em[G.L - 1] + (em[G.L - 2] - em[G.L - 1]) / G.pv_s1 * G.pv_s2
    mid = 0.25 * (rho[:-1][:, G.fr] + rho[:-1][:, G.to] + rho[1:][:, G.fr] + rho[1:][:, G.to])
    return torch.cat([top[None], mid, sfc[None]], 0)


def rhombus_density(G, rho):
    return (G.drw * rho[:, G.dri]).sum(2)


def pot_vort(G, u, w, rho):
    """vorticities.c:18-85"""
    tan, rh = rel_vort(G, u, w)
    return (tan + G.f_tan) / half_level_density(G, rho), (rh + G.f_rh) / rhombus_density(G, rho)


def vorticity_flux(G, Fu, Fw, qt, qr):
    """vorticity_flux.c:16-187: (u, w) output"""
    qm = qr[:, G.tm]                                          # [L, NVH, 10]
    qown = qr[:, :, None]
    q = torch.where(G.trsk_half.bool(), 0.5 * (qm + qown), qm)
    out_u = (G.trsk_used * G.tw * Fu[:, G.ti] * q).sum(2)
    ip6, ip7 = G.ipw[:, :, 6], G.ipw[:, :, 7]
    out_u = out_u - 0.5 * ip6[:, G.fr] * Fw[:-1][:, G.fr] * qt[:-1]
    out_u = out_u - 0.5 * ip6[:, G.to] * Fw[:-1][:, G.to] * qt[:-1]
    out_u = out_u - 0.5 * ip7[:, G.fr] * Fw[1:][:, G.fr] * qt[1:]
    out_u = out_u - 0.5 * ip7[:, G.to] * Fw[1:][:, G.to] * qt[1:]
    # vertical: levels 0..L
    vw = torch.full((G.L + 1, 1), 0.5, dtype=T)
    vw[0] = 1.0
    vw[G.L] = 1.0
    layer_terms = (G.ipw[:, :, :6] * Fu[:, G.adj] * G.mask6)   # [L, NSH, 6]
    above = torch.zeros(G.L + 1, G.nsh, dtype=T)
    above = torch.cat([above[:1], (layer_terms * qt[1:][:, G.adj]).sum(2)], 0)      # level l reads layer l-1
    below = torch.cat([(layer_terms * qt[:-1][:, G.adj]).sum(2), above[:1]], 0)    # level l reads layer l
    out_w = vw * (above + below)
    return out_u, out_w


def advection(G, u, w, rho):
    """the momentum-advection terms of vector_tendencies_expl.c:24-35:
    (pot_vort_tend, v_squared_grad) as (u, w) pairs"""
    Fu, Fw = stv_h(G, rho, u), stv_v(G, rho, w)
    qt, qr = pot_vort(G, u, w, rho)
    pu, pw = vorticity_flux(G, Fu, Fw, qt, qr)
    vsq = inner_product(G, u, w, u, w)
    gu, gw = grad(G, vsq)
    return (pu, pw), (gu, gw)


def pressure_gradient(G, tp, ep):
    """manage_pressure_gradient.c:17-72 (dry: the decel factor is 1): nl + l"""
    a = C_D_P * (G.thbg + tp)
    gu, gw = grad(G, ep)
    b = C_D_P * tp
    Pu = stv_h(G, a, gu) + stv_h(G, b, G.exbg_grad_u)
    Pw = stv_v(G, a, gw) + stv_v(G, b, G.exbg_grad_w)
    return Pu, Pw


def scalar_tendencies(G, rho, rt, u):
    """scalar_tendencies_expl.c (gas only): -div(F), -div(theta_v F)"""
    F = stv_h(G, rho, u)
    trho = -divv_h(G, F)
    thv = rt / rho
    trt = -divv_h(G, stv_h(G, thv, F))
    return trho, trt


def thomas(c, d, e, r):
    """column_solvers.c:641-667 over columns: lists of [NSH] tensors"""
    n = len(d)
    ep, rp = [None] * n, [None] * n
    ep[0] = e[0] / d[0]
    rp[0] = r[0] / d[0]
    for j in range(1, n):
        den = d[j] - ep[j - 1] * c[j - 1]
        if j < n - 1:
            ep[j] = e[j] / den
        rp[j] = (r[j] - rp[j - 1] * c[j - 1]) / den
    x = [None] * n
    x[n - 1] = rp[n - 1]
    for j in range(n - 2, -1, -1):
        x[j] = rp[j] - ep[j] * x[j + 1]
    return x


def ver_waves(G, so, snr, td, dt, rk):
    """three_band_solver_ver_waves (column_solvers.c:19-383) over all columns.
    so, snr: dicts with rho, rt, tp, ep [L, NSH] and w [L+1, NSH] (so);
    td: trho, trt [L, NSH], tw [L+1, NSH].  Returns (rho, rt, w, tp, ep) of the
    new state (w: all levels; top and surface copied from so)."""
    L = G.L
    iw, pdw = G.impl_weight, 0.5
    vol, exbg, zs, area = G.vol, G.exbg, G.zs, G.area_w
    ro, rto, tpo, epo, wo = so["rho"], so["rt"], so["tp"], so["ep"], so["w"]
    rn, rtn, epn = snr["rho"], snr["rt"], snr["ep"]
    re = ro + dt * td["trho"]
    rte = rto + dt * td["trt"]
    if rk == 0:
        al = -rto / (ro * ro) / vol
        be = 1.0 / ro / vol
        ga = R_D / (C_D_V * rto) * (exbg + epo) / vol
    else:
        ao, bo, go = -rto / (ro * ro), 1.0 / ro, R_D / (C_D_V * rto) * (exbg + epo)
        an, bn, gn = -rtn / (rn * rn), 1.0 / rn, R_D / (C_D_V * rtn) * (exbg + epn)
        al = ((1 - pdw) * ao + pdw * an) / vol
        be = ((1 - pdw) * bo + pdw * bn) / vol
        ga = ((1 - pdw) * go + pdw * gn) / vol
    tpe = tpo + dt * vol * (al * td["trho"] + be * td["trt"])
    epe = epo + dt * vol * ga * td["trt"]
    rio = 0.5 * (ro[:-1] + ro[1:])
    rie = 0.5 * (re[:-1] + re[1:])
    tin = 0.5 * (rtn[:-1] / rn[:-1] + rtn[1:] / rn[1:])
    dz = zs[:-1] - zs[1:]
    dex = exbg[:-1] - exbg[1:]
    wi = wo[1:L]
    a_w = area[1:L]
    dvv = (-(tin * tin) * (ga[:-1] + ga[1:]) + 0.5 * dex * (al[1:] - al[:-1] + tin * (be[1:] - be[:-1]))
           - dz / (iw * dt * dt * C_D_P * rio) * (2.0 / a_w + dt * wi * 0.5 * (-1.0 / vol[:-1] + 1.0 / vol[1:])))
    rvv = (-(wi + dt * td["tw"][1:L]) * dz / (iw * dt * dt * C_D_P) + tin * (epe[:-1] - epe[1:]) / dt
           + 0.5 / dt * (tpe[:-1] + tpe[1:]) * dex - dz / (iw * dt * dt * C_D_P) * wi * rie / rio)
    cvv = (tin[1:] * ga[1:-1] * tin[:-1] + 0.5 * dex[1:] * (al[1:-1] + be[1:-1] * tin[:-1])
           - dz[1:] / (iw * dt * C_D_P) * 0.5 * wi[1:] / (vol[1:-1] * rio[1:]))
    evv = (tin[:-1] * ga[1:-1] * tin[1:] - 0.5 * dex[:-1] * (al[1:-1] + be[1:-1] * tin[1:])
           + dz[:-1] / (iw * dt * C_D_P) * 0.5 * wi[:-1] / (vol[1:-1] * rio[:-1]))
    sol = thomas(list(cvv), list(dvv), list(evv), list(rvv))
    sol = torch.stack(sol) / (1.0 + dt * G.damp)
    zero = torch.zeros(1, G.nsh, dtype=T)
    s_up = torch.cat([sol, zero], 0)          # sol[j] at the lower face of layer j
    s_dn = torch.cat([zero, sol], 0)          # sol[j-1] at the upper face
    rho_new = re + dt * (s_up - s_dn) / vol
    tin_up = torch.cat([tin, zero], 0)
    tin_dn = torch.cat([zero, tin], 0)
    rt_new = rte + dt * (tin_up * s_up - tin_dn * s_dn) / vol
    din = 0.5 * (rho_new[:-1] + rho_new[1:])
    w_mid = (2.0 * sol / a_w - din * wi) / rio
    w_new = torch.cat([wo[:1], w_mid, wo[L:]], 0)
    tp_new = rt_new / rho_new - G.thbg
    ep_new = epo + vol * ga * (rt_new - rto)
    return rho_new, rt_new, w_new, tp_new, ep_new


def step(G, s, pgrad_prev, adv_prev, dt, first):
    """One RKHEVI step (manage_rkhevi.c:20-100) after the driver's copy
    new -> old (coordinator.c:305), dry.  s: dict rho, rt, ep, u, w (theta_v_pert
    is re-derived from rho and rt by the driver's copy); pgrad_prev: (u, w) of
    the previous step's nl + l; adv_prev: (u, w) of the previous step's
    pot_vort_tend - 0.5 v_squared_grad.  Returns (state, pgrad, adv)."""
    chw = 0.5 + G.impl_weight
    ohw = 1.0 - chw
    cvw = 1.0 - G.impl_weight
    rho, rt, ep, u, w = s["rho"], s["rt"], s["ep"], s["u"], s["w"]
    tp = rt / rho - G.thbg
    Pu, Pw = pressure_gradient(G, tp, ep)
    if first:
        pold_u = -Pu
        (pu, pw), (gu, gw) = advection(G, u, w, rho)
        adv_u, adv_w = pu - 0.5 * gu, pw - 0.5 * gw
    else:
        pold_u = -pgrad_prev[0]
        adv_u, adv_w = adv_prev
    # rk 0
    tw_u = ohw * pold_u - chw * Pu + adv_u
    tw_w = -cvw * Pw + adv_w
    u0 = u + dt * tw_u
    trho, trt = scalar_tendencies(G, rho, rt, u0)
    so = {"rho": rho, "rt": rt, "tp": tp, "ep": ep, "w": w}
    td = {"trho": trho, "trt": trt, "tw": tw_w}
    r1, rt1, w1, tp1, ep1 = ver_waves(G, so, so, td, dt, 0)
    # rk 1
    (pu, pw), (gu, gw) = advection(G, u0, w1, r1)
    a1u, a1w = pu - 0.5 * gu, pw - 0.5 * gw
    tw1_u = 0.5 * tw_u + 0.5 * (ohw * pold_u - chw * Pu + a1u)
    tw1_w = 0.5 * tw_w + 0.5 * (-cvw * Pw + a1w)
    u2 = u + dt * tw1_u
    trho1, trt1 = scalar_tendencies(G, r1, rt1, u2)
    snr = {"rho": r1, "rt": rt1, "ep": ep1}
    td1 = {"trho": trho1, "trt": trt1, "tw": tw1_w}
    r2, rt2, w2, tp2, ep2 = ver_waves(G, so, snr, td1, dt, 1)
    return {"rho": r2, "rt": rt2, "ep": ep2, "u": u2, "w": w2}, (Pu, Pw), (a1u, a1w)


def eos_state(G, rho, rt, u, w):
    """the state of the control variables (rho, rhotheta_v, w, u): theta_v_pert
    and exner_pert diagnosed from rho and rhotheta_v through the equation of
    state, Exner = (R_d rhotheta_v / p_0)^(R_d / c_v) (set_initial_state.c:302-314
    written for rhotheta_v)"""
    ep = (R_D * rt / P_0) ** (R_D / C_D_V) - G.exbg
    return {"rho": rho, "rt": rt, "ep": ep, "u": u, "w": w}


def run(G, s, dt, n_steps):
    """n_steps steps from a fresh start (the first takes the first-step branches)"""
    pg, adv = None, None
    for n in range(n_steps):
        s, pg, adv = step(G, s, pg, adv, dt, n == 0)
    return s
