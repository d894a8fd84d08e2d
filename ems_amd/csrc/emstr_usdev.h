/*
 * Device-side view of one COMPAS window (C POD shared by the host C layer
 * and the HIP kernels).  Columns (wet mesh cells, in surface order) are
 * laid out like a one-row structured grid, so the column kernels of the
 * plain form (set_dz, sub-step elevations, vertical diffusion, mass,
 * scatter/gather) run on them unchanged:
 *   cell arrays   [nz][ncol]     (emstr_dev with nx = ncol, ny = 1)
 *   edge arrays   [nz][ne2]      (ne2 surface-layer edges; the edge of a
 *                                 layer is the surface edge at that depth)
 *   per column    [npem][ncol]   (edges of the cell, in the mesh order)
 * Land and ghost centres are not stored: a ghost is read as the step-start
 * value of its wet centre (set_lateral_BC_tr, tracers.c:7120-7137), the
 * sediment ghost as the bed value (tracers.c:864-867).
 */
#ifndef EMSTR_USDEV_H
#define EMSTR_USDEV_H

#include "emstr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EMSTR_FF_ROWS 16 /* most distinct centres one FFSL reconstruction reads */
#define EMSTR_FF_NPE 8   /* most vertices (edges) of an FFSL centre          */

typedef struct emstr_usdev {
  emstr_dev d;      /* column part: nx = ncol, ny = 1                   */
  int npem, ne2;
  int *npe;         /* [ncol] edges of each cell                         */
  int *ce;          /* [npem][ncol] surface edge j of the cell           */
  int *esc;         /* [npem][ncol] eSc: +1 if edge j points out         */
  /* per surface edge: the columns of e2c[e][0] / e2c[e][1], the
   * columns beyond them through the edge opposite (van Leer's cp / cm,
   * -1 if none), and which cell ULTIMATE reads through ep / em when
   * that far column is wet at the layer (bit 0: cp, else e2c[e][0];
   * bit 1: cm, else e2c[e][1])                                       */
  int *ec1, *ec2, *ecp, *ecm;
  unsigned char *esel;
  double *h1e;      /* [ne2] edge length (h1au1)                        */
  double *h2e;      /* [ne2] centre distance (h2au1)                    */
  /* QUICKEST (tracers.c:592-603, 2080-2123): d3 = 0.5*h1acell/h2au1 and
   * the centre distances of the surface edges em / ep                */
  double *qd3, *qhm, *qhp;
  /* [nz][ne2] edge fields: u1, u1flux3d, dzu1, u2kh                  */
  double *u1, *flux, *dzu, *kh;

  /* FFSL (transport.c:1961-2723): geometry of the surface layer       */
  int ffsl;         /* the window's scheme is FFSL                      */
  int nvl, nvcm;    /* surface vertices; most columns around one (host) */
  double *cx, *cy;  /* [ncol] centre positions                          */
  double *ux, *uy;  /* [ne2] edge midpoints (streamline starts)         */
  double *n1x, *n1y, *t2x, *t2y; /* [ne2] costhu1/sinthu1, costhu2/sinthu2 */
  double *ev;       /* [ne2][4] end vertices x1, y1, x2, y2 (e2v)       */
  double *vxy;      /* [ncol][EMSTR_FF_NPE][2] centre minus vertex j, x
                     * and y, of the column's vertices in order          */
  int *eord;        /* [npem][ncol] edge slots in ascending edge order  */
  /* what the limited reconstruction of wet centre q (wet-list position)
   * reads: the distinct wet centres lr_q[lr_off[q]..lr_off[q+1]-1]
   * (itself first, at most EMSTR_FF_ROWS); its least-squares stencil,
   * entries lsq_off[q]..lsq_off[q+1]-1, as slots lsq_s of that list (a
   * ghost reads its wet centre) with the slope weights lsq_B[2*entry],
   * lsq_B[2*entry + 1] (the constant term is not used); and vertex j's
   * wet centres as the bytes of vsl[q*EMSTR_FF_NPE + j], 0xff ending
   * the list */
  int *lsq_off;
  int *lsq_s;
  int *lsq_gc;      /* the column of each stencil entry's cell at the
                     * centre's layer, ~column for a lateral ghost (read
                     * as the step-start value of that wet centre)      */
  double *lsq_B;
  int *lr_off, *lr_q;
  unsigned *vsl;
  int tqs;          /* row stride of the [nwet][tqs] tr_mod rows        */
  /* fields: [nz][ne2] u2; [ncol] wtop, wbot; [nz][ncol] u, v of the wet
   * centres; [nz][ne2] u, v of the ghost across a land edge (gu0/gv0
   * from the window's arrays, gue/gve rotated from the edge velocity) */
  double *u2, *wtop, *wbot, *uc, *vc, *gu0, *gv0, *gue, *gve;
  /* multi-segment streamlines (hd_trans_interp): the velocity fit of
   * wet centre q reads points vls_off[q]..vls_off[q+1]-1, vls_ent >= 0
   * the dense index of a wet centre, else ~(k * ne2 + e) for the ghost
   * across surface edge e at layer k (the edge velocity rotated), with
   * slope rows vls_bx, vls_by; vfit [6][nz][ncol]: value, x slope and y
   * slope of u, then of v, rebuilt when ffsl_init runs (vfit_ok once it
   * has).  Segments after a streamline's first: lnx[edge] the first
   * record (-1: none), record r in column sg_ci[r] with the offsets and
   * length sg_d[r * 5 + 0..4] (as ld) and the next record sg_nx[r];
   * sg_cnt counts the records of the sub-step, at most sg_cap.        */
  int *vls_off;
  long long *vls_ent;
  double *vls_bx, *vls_by, *vfit;
  int vfit_ok;
  int *lnx, *sg_ci, *sg_nx, *sg_cnt;
  double *sg_d;
  int sg_cap;
} emstr_usdev;

#ifdef __cplusplus
}
#endif
#endif
