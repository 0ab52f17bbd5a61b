/*
 * adj_tables_host.h -- the reverse index tables of the adjoint gathers, built
 * on the host from the forward tables (reference numbering).  Used by the
 * runtime, which uploads them (AdjTables, PvReverse in the device headers).
 */
#ifndef GAME_ADJ_TABLES_HOST_H
#define GAME_ADJ_TABLES_HOST_H

#include <stdio.h>
#include <vector>

struct AdjTablesHost {
    std::vector<int> slot_from, slot_to, tri_of_edge, tri_slot, rh_off, rh_ent, tr_off, tr_ent;
    std::vector<int> rhp_e, trp_e, trp_m;
    std::vector<double> rhp_w, trp_w;
    /* reverse table of the rhombus PV read by the TRSK sums, padded to pvr_max */
    std::vector<int> pvr_e, pvr_f, pvr_n;
    std::vector<double> pvr_w;
};

/* nsh, nvh, nt, n_pent (12); adj 6*NSH, from/to NVH, vit 3*NT, dri/drw 4*NVH,
   ti/tmc/tw 10*NVH; rh_max, tr_max, pvr_max: padded lengths.  0 or an error
   message in err. */
inline int build_adj_tables_host(int nsh, int nvh, int nt, int n_pent, const int *adj, const int *from,
                                 const int *to, const int *vit, const int *dri, const double *drw, const int *ti,
                                 const int *tmc, const double *tw, int rh_max, int tr_max, int pvr_max,
                                 AdjTablesHost &T, char *err, size_t nerr)
{
    const size_t NSH = (size_t)nsh, NVH = (size_t)nvh, NT = (size_t)nt;
    std::vector<int> sf(NVH, -1), stt(NVH, -1);
    for (size_t c = 0; c < NSH; ++c) {
        int ne = (int)c < n_pent ? 5 : 6;
        for (int j = 0; j < ne; ++j) {
            int e = adj[6*c + j];
            if (from[e] == (int)c) sf[e] = j;
            else if (to[e] == (int)c) stt[e] = j;
        }
    }
    std::vector<int> toe(2*NVH, -1), tsl(2*NVH, -1), cnt(NVH, 0);
    for (size_t t = 0; t < NT; ++t)
        for (int k = 0; k < 3; ++k) {
            int e = vit[3*t + k];
            if (e < 0 || (size_t)e >= NVH || cnt[e] >= 2) {
                snprintf(err, nerr, "triangle %zu lists edge %d a third time", t, e);
                return 2;
            }
            toe[2*e + cnt[e]] = (int)t;
            tsl[2*e + cnt[e]] = k;
            ++cnt[e];
        }
    for (size_t e = 0; e < NVH; ++e)
        if (sf[e] < 0 || stt[e] < 0 || cnt[e] != 2) {
            snprintf(err, nerr, "edge %zu: inconsistent adjacency or triangle tables", e);
            return 2;
        }
    /* CSR lists in ascending (edge, slot) order */
    std::vector<int> roff(NSH + 1, 0), rent;
    for (size_t e = 0; e < NVH; ++e) for (int j = 0; j < 4; ++j) ++roff[dri[4*e + j] + 1];
    for (size_t c = 0; c < NSH; ++c) roff[c + 1] += roff[c];
    rent.resize(2*(size_t)roff[NSH]);
    {
        std::vector<int> fill(roff.begin(), roff.end() - 1);
        for (size_t e = 0; e < NVH; ++e)
            for (int j = 0; j < 4; ++j) {
                int c = dri[4*e + j], k = fill[c]++;
                rent[2*(size_t)k] = (int)e;
                rent[2*(size_t)k + 1] = j;
            }
    }
    /* TRSK slots the forward reads: 0..3 / 5..8 beside a pentagon, else 0..4 / 5..9 */
    auto used = [&](size_t e, int j) {
        if (j < 5) return from[e] < n_pent ? j < 4 : true;
        return to[e] < n_pent ? j < 9 : true;
    };
    auto half = [&](size_t e, int j) { return (j == 2 && from[e] >= n_pent) || (j == 7 && to[e] >= n_pent); };
    std::vector<int> toff(NVH + 1, 0), tent;
    for (size_t e = 0; e < NVH; ++e) for (int j = 0; j < 10; ++j) if (used(e, j)) ++toff[ti[10*e + j] + 1];
    for (size_t e = 0; e < NVH; ++e) toff[e + 1] += toff[e];
    tent.resize(2*(size_t)toff[NVH]);
    {
        std::vector<int> fill(toff.begin(), toff.end() - 1);
        for (size_t e = 0; e < NVH; ++e)
            for (int j = 0; j < 10; ++j) {
                if (!used(e, j)) continue;
                int k = fill[ti[10*e + j]]++;
                tent[2*(size_t)k] = (int)e;
                tent[2*(size_t)k + 1] = j;
            }
    }
    /* padded slot-major copies of both lists, with the weights (and PV
       indices) of each entry; a TRSK term that half-averages PV with the
       edge's own (slots 2 and 7 away from pentagons) stores its PV index m
       as -(m + 1) */
    T.rhp_e.assign((size_t)rh_max*NSH, 0);
    T.rhp_w.assign((size_t)rh_max*NSH, 0.0);
    T.trp_e.assign((size_t)tr_max*NVH, 0);
    T.trp_m.assign((size_t)tr_max*NVH, 0);
    T.trp_w.assign((size_t)tr_max*NVH, 0.0);
    for (size_t c = 0; c < NSH; ++c) {
        const int n = roff[c + 1] - roff[c];
        if (n > rh_max) { snprintf(err, nerr, "cell %zu in %d rhombi", c, n); return 2; }
        for (int k = 0; k < n; ++k) {
            const int e = rent[2*(size_t)(roff[c] + k)], j = rent[2*(size_t)(roff[c] + k) + 1];
            T.rhp_e[(size_t)k*NSH + c] = e;
            T.rhp_w[(size_t)k*NSH + c] = drw[4*(size_t)e + j];
        }
    }
    for (size_t ep = 0; ep < NVH; ++ep) {
        const int n = toff[ep + 1] - toff[ep];
        if (n > tr_max) { snprintf(err, nerr, "edge %zu in %d TRSK stencils", ep, n); return 2; }
        for (int k = 0; k < tr_max; ++k) T.trp_e[(size_t)k*NVH + ep] = (int)ep;
        for (int k = 0; k < n; ++k) {
            const int e = tent[2*(size_t)(toff[ep] + k)], j = tent[2*(size_t)(toff[ep] + k) + 1];
            T.trp_e[(size_t)k*NVH + ep] = e;
            T.trp_m[(size_t)k*NVH + ep] = half(e, j) ? -(tmc[10*(size_t)e + j] + 1) : tmc[10*(size_t)e + j];
            T.trp_w[(size_t)k*NVH + ep] = tw[10*(size_t)e + j];
        }
    }
    /* PV reverse: rhombus m <- (output edge e, weight, flux edge) of every
       TRSK term of e reading PV m: the modified index of slot j (half the
       weight at a half-average slot) and, at a half-average slot, the edge's
       own rhombus with the other half */
    std::vector<std::vector<int>> pe(NVH), pf(NVH);
    std::vector<std::vector<double>> pw(NVH);
    for (size_t e = 0; e < NVH; ++e)
        for (int j = 0; j < 10; ++j) {
            if (!used(e, j)) continue;
            const double w = tw[10*e + j];
            const int f = ti[10*e + j];
            const int m = tmc[10*e + j];
            const double share = half(e, j) ? 0.5 : 1.0;
            pe[m].push_back((int)e); pf[m].push_back(f); pw[m].push_back(share*w);
            if (half(e, j)) { pe[e].push_back((int)e); pf[e].push_back(f); pw[e].push_back(0.5*w); }
        }
    T.pvr_e.assign((size_t)pvr_max*NVH, 0);
    T.pvr_f.assign((size_t)pvr_max*NVH, 0);
    T.pvr_w.assign((size_t)pvr_max*NVH, 0.0);
    T.pvr_n.assign(NVH, 0);
    for (size_t m = 0; m < NVH; ++m) {
        const size_t n = pe[m].size();
        if ((int)n > pvr_max) { snprintf(err, nerr, "rhombus %zu read by %zu TRSK terms", m, n); return 2; }
        T.pvr_n[m] = (int)n;
        for (size_t k = 0; k < n; ++k) {
            T.pvr_e[k*NVH + m] = pe[m][k];
            T.pvr_f[k*NVH + m] = pf[m][k];
            T.pvr_w[k*NVH + m] = pw[m][k];
        }
    }
    T.slot_from = sf; T.slot_to = stt; T.tri_of_edge = toe; T.tri_slot = tsl;
    T.rh_off = roff; T.rh_ent = rent; T.tr_off = toff; T.tr_ent = tent;
    return 0;
}

#endif
