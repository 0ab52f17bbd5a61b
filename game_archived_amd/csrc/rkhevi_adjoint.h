/*
 * rkhevi_adjoint.h -- reverse index tables and launchers of the operator
 * adjoints (rkhevi_adjoint.hip).
 */
#ifndef GAME_RKHEVI_ADJOINT_H
#define GAME_RKHEVI_ADJOINT_H

#include "rkhevi_dev.h"

/* reverse tables, built once from the forward tables */
struct AdjTables {
    const int *slot_from, *slot_to;   /* NVH: slot of edge e in the adjacency of its from/to cell */
    const int *tri_of_edge;           /* 2 NVH: the two triangles whose curl reads edge e */
    const int *tri_slot;              /* 2 NVH: the slot of e in those triangles' edge lists */
    const int *rh_off, *rh_ent;       /* CSR over cells: (edge, slot) pairs of the rhombus weights reading the cell */
    const int *tr_off, *tr_ent;       /* CSR over edges: (edge, slot) pairs of the TRSK stencils reading the edge */
    /* the same two lists padded to a fixed length, slot-major so that lanes of
       consecutive cells/edges read consecutive words, with the forward weight
       (and PV index) of each entry stored beside it; padding has weight 0 */
    const int *rhp_e;                 /* ADJ_RH_MAX x NSH: rhombus edge */
    const double *rhp_w;              /* ADJ_RH_MAX x NSH: its weight on the cell */
    const int *trp_e, *trp_m;         /* ADJ_TR_MAX x NVH: stencil edge, its PV index m (-(m + 1) when
                                         the term half-averages PV with the edge's own) */
    const double *trp_w;              /* ADJ_TR_MAX x NVH: TRSK weight */
};
#define ADJ_RH_MAX 12
#define ADJ_TR_MAX 10

namespace adj {
/* grad^T of gs*gb into fb (acc: added to fb) */
void grad_adjoint(const DevGrid &G, const AdjTables &A, const double *gb, double *fb, hipStream_t st,
                  double gs = 1.0, int acc = 0);
void divv_h(const DevGrid &G, const double *in, double *out, hipStream_t st);
void divv_h_adjoint(const DevGrid &G, const AdjTables &A, const double *db, double *Fb, hipStream_t st);
/* acc: the edge cotangents are added to wb, its w points left as they are */
void curl_triangles_adjoint(const DevGrid &G, const AdjTables &A, const double *zb, double *wb, hipStream_t st,
                            int acc = 0);
void rhombus_density(const DevGrid &G, const double *rho, double *out, hipStream_t st);
void rhombus_density_adjoint(const DevGrid &G, const AdjTables &A, const double *db, double *rb, hipStream_t st);
void vorticity_flux_adjoint(const DevGrid &G, const AdjTables &A, const double *pv, const double *ob, double *Fb,
                            hipStream_t st);
/* acc: ab += scale*(the adjoint); otherwise ab = the adjoint */
void inner_product_adjoint(const DevGrid &G, const AdjTables &A, const double *b, const double *ib, double *ab,
                           hipStream_t st, double scale = 1.0, int acc = 0);
void thomas(int n, int ncol, const double *c, const double *d, const double *e, const double *r, double *x,
            double *work, hipStream_t st);
void thomas_adjoint(int n, int ncol, const double *c, const double *d, const double *e, const double *x,
                    const double *xb, double *rb, double *cb, double *db, double *eb, double *work, hipStream_t st);
/* the device address of the solver scratch (THOMAS_SCRATCH_DOUBLES doubles) on
   the current device: resolved once per device by thomas_scratch_resolve
   (0, or -1 with a message), then returned by thomas_scratch (null when not
   resolved on this device) */
int thomas_scratch_resolve(char *err, size_t nerr);
double *thomas_scratch(void);
}
/* capacity of the allocation-free three-band solve, in doubles (64 MiB):
   n*ncol elements, e.g. every column of a res 7 x 26 grid */
#define THOMAS_SCRATCH_DOUBLES (1LL << 23)

#endif
