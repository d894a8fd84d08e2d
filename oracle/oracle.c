/*
 * ORACLE -- test infrastructure only.  Not part of the product.
 *
 * A CPU restatement, in strict IEEE FP64, of the EMS SHOC tracer step
 * (model/hd/tracers/tracers.c) for one window in z-level coordinates with
 * closed lateral boundaries, no point sources, regions or multi-dt.  Each
 * function follows the reference function cited above it with the same
 * loop order and the same order of floating-point operations, so that a
 * GPU path evaluated the same way agrees bit for bit.  Built with
 * gcc -O3 -fno-fast-math -ffp-contract=off.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load this library, and only as the checker or the timed CPU baseline.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   set_dz             brute-force column walk on small grids; dz sums
 *   van_leer_do        closed forms: linear profile, constants, step (TVD)
 *   quickest(+uniform) closed form on a linear profile, constancy
 *   ulimate_filter     normalised-variable cases by hand
 *   order1/2/4         closed forms on linear profiles
 *   update + surf_conc closed-basin mass conservation, constancy
 *   tr_bounds          clip and copy-up cases by hand
 *   implicit_vdiff_tr  dense solve (numpy) and the discrete decay factor
 *   check_tmass        direct sum
 * The reference's own tests hold no golden vectors for this path, and its
 * executables cannot be built here (no NetCDF), so nothing is compared
 * against reference output directly.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/emstr.h"

#define max(a, b) ((a) > (b) ? (a) : (b))
#define min(a, b) ((a) < (b) ? (a) : (b))

static int oracle_err = 0;
int oracle_last_error(void) { return oracle_err; }

/*-------------------------------------------------------------------*/
/* tracers.c:5574-5593  no-gradient lateral boundary: ghost = bin      */
/*-------------------------------------------------------------------*/
void oracle_set_lateral_BC_tr(geometry_t *g, window_t *wd) {
  int cc, n;
  for (cc = 1; cc <= g->nbpt; cc++) {
    int gb = g->bpt[cc], in = g->bin[cc];
    for (n = 0; n < wd->ntr; n++) wd->tr_wc[n][gb] = wd->tr_wc[n][in];
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:5682-5849  set_dz, z-level branch (thin_merge not taken) */
/*-------------------------------------------------------------------*/
void oracle_set_dz(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int cc, c, cs, zm1, vc, vcs;
  double top, bot;

  /* new surface: walk down while the lower face is at or above eta */
  for (cc = 1; cc <= g->a2_t; cc++) {
    c = cs = g->w2_t[cc];
    g->sur_t[cc] = g->nsur_t[cc];
    zm1 = g->zm1[c];
    while (c != zm1 && g->gridz[c] >= wd->eta[cs]) {
      c = zm1;
      zm1 = g->zm1[c];
    }
    g->nsur_t[cc] = c;
  }

  /* surface cells to process: lower of old and new surface */
  vc = 1;
  for (cc = 1; cc <= g->b2_t; cc++) {
    c = max(g->sur_t[cc], g->nsur_t[cc]);
    cs = g->m2d[c];
    top = min(wd->eta[cs], wc->oldeta[cs]);
    bot = DRY_FRAC * wc->hmin + g->botz[cs];
    if (top > bot && c != g->zm1[c]) {
      wc->s1[vc] = c;
      wc->i1[vc] = g->bot_t[cc];
      wc->i2[vc] = g->nsur_t[cc];
      wc->i3[vc] = g->sur_t[cc];
      if (cc == g->v2_t) wc->vcs1 = vc;
      vc++;
    }
  }
  vcs = vc - 1;
  wc->vcs = vcs;

  /* sub-surface cells, column by column down to the bed */
  for (cc = 1; cc <= vcs; cc++) {
    c = g->zm1[wc->s1[cc]];
    zm1 = g->zm1[c];
    while (c != zm1) {
      wc->s1[vc] = c;
      vc++;
      c = zm1;
      zm1 = g->zm1[c];
    }
    if (cc == wc->vcs1) wc->vc1 = vc - 1;
  }
  vc--;
  wc->vc = vc;

  /* thicknesses below the old surface; bed cell from botz; cells above
   * the old surface take the surface thickness */
  for (cc = 1; cc <= g->b2_t; cc++) {
    int c3, c2, cbot;
    c = c3 = g->sur_t[cc];
    c2 = g->m2d[c3];
    cbot = g->bot_t[cc];
    top = wc->oldeta[c2];
    while (c3 < cbot) {
      bot = g->gridz[c3];
      wc->dz[c3] = top - bot;
      top = bot;
      c3 = g->zm1[c3];
    }
    wc->dz[cbot] = top - g->botz[c2];
    c3 = c;
    while (c > c2) {
      c = g->zp1[c];
      wc->dz[c] = wc->dz[c3];
    }
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:2082-2119  van Leer face value along one direction       */
/*-------------------------------------------------------------------*/
static void vl_do(double *F, const double *tr, const double *vel,
                  const double *cn, int ss, int se, const int *sdo,
                  const int *fmap, const int *bmap) {
  int cc;
  for (cc = ss; cc <= se; cc++) {
    int c = sdo[cc];
    int bm1 = bmap[c];
    int bm2 = bmap[bm1];
    double df;
    if (vel[c] > 0.0) {
      df = (tr[bm1] - tr[bm2]) * (tr[c] - tr[bm1]);
      if (df > 0.0)
        df = 2.0 * df / (tr[c] - tr[bm2]);
      else
        df = 0.0;
      F[c] = tr[bm1] + 0.5 * (1.0 - cn[c]) * df;
    } else if (vel[c] < 0.0) {
      df = (tr[c] - tr[bm1]) * (tr[fmap[c]] - tr[c]);
      if (df > 0.0)
        df = 2.0 * df / (tr[fmap[c]] - tr[bm1]);
      else
        df = 0.0;
      F[c] = tr[c] - 0.5 * (1.0 + cn[c]) * df;
    }
  }
}

/* tracers.c:2037-2072 */
static void vl_faces(geometry_t *g, window_t *wd, win_priv_t *wc,
                     const double *tr, double *Fx, double *Fy, double *Fz) {
  memset(Fx, 0, g->n3_t * sizeof(double));
  vl_do(Fx, tr, wd->u1, wc->w6, 1, g->v3_e1, g->w3_e1, g->xp1, g->xm1);
  vl_do(Fx, tr, wd->u1, wc->w6, g->v3_t + 1, g->a3_t, g->w3_t, g->xp1, g->xm1);
  memset(Fy, 0, g->n3_t * sizeof(double));
  vl_do(Fy, tr, wd->u2, wc->w7, 1, g->v3_e2, g->w3_e2, g->yp1, g->ym1);
  vl_do(Fy, tr, wd->u2, wc->w7, g->v3_t + 1, g->a3_t, g->w3_t, g->yp1, g->ym1);
  memset(Fz, 0, g->n3_t * sizeof(double));
  vl_do(Fz, tr, wd->w, wc->w8, 1, wc->vc, wc->s1, g->zp1, g->zm1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:1819-1931  QUICKEST, non-uniform grid                    */
/*-------------------------------------------------------------------*/
static void qk_faces(geometry_t *g, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fy, double *Fz) {
  int cc;
  memset(Fx, 0, g->n3_t * sizeof(double));
  memset(Fy, 0, g->n3_t * sizeof(double));
  memset(Fz, 0, g->n3_t * sizeof(double));
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    int c2 = g->m2d[c];
    int xp1 = g->xp1[c], yp1 = g->yp1[c];
    int xm1 = g->xm1[c], ym1 = g->ym1[c];
    int xm2 = g->xm1[xm1], ym2 = g->ym1[ym1];
    int xm1s = g->xm1[c2], xp1s = g->xp1[c2];
    int ym1s = g->ym1[c2], yp1s = g->yp1[c2];
    double cx, cy, db, df, face, grad, curv;

    cx = wc->w6[c];
    db = 2.0 / (g->h1au1[xm1s] + g->h1au1[c2]);
    df = 2.0 / (g->h1au1[c2] + g->h1au1[xp1s]);
    face = wc->d3[c2] * (tr[c] - tr[xm1]) + tr[xm1];
    grad = tr[c] - tr[xm1];
    curv = 0.0;
    if (cx > 0.0)
      curv = db * (g->h1au1[xm1s] * tr[c] + g->h1au1[c2] * tr[xm2]) - 2.0 * tr[xm1];
    else if (cx < 0.0)
      curv = df * (g->h1au1[c2] * tr[xp1] + g->h1au1[xp1s] * tr[xm1]) - 2.0 * tr[c];
    Fx[c] = face - 0.5 * cx * grad - (1.0 - cx * cx) * curv / 6.0;

    cy = wc->w7[c];
    db = 2.0 / (g->h2au2[ym1s] + g->h2au2[c2]);
    df = 2.0 / (g->h2au2[c2] + g->h2au2[yp1s]);
    face = wc->d4[c2] * (tr[c] - tr[ym1]) + tr[ym1];
    grad = tr[c] - tr[ym1];
    curv = 0.0;
    if (cy > 0.0)
      curv = db * (g->h2au2[ym1s] * tr[c] + g->h2au2[c2] * tr[ym2]) - 2.0 * tr[ym1];
    else if (cy < 0.0)
      curv = df * (g->h2au2[c2] * tr[yp1] + g->h2au2[yp1s] * tr[ym1]) - 2.0 * tr[c];
    Fy[c] = face - 0.5 * cy * grad - (1.0 - cy * cy) * curv / 6.0;
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zp1 = g->zp1[c], zm1 = g->zm1[c];
    int zm2 = g->zm1[zm1];
    double cz = wc->w8[c];
    double db = 2.0 / (wc->w9[zm1] + wc->w9[c]);
    double df = 2.0 / (wc->w9[c] + wc->w9[zp1]);
    double face = wc->w5[c] * (tr[c] - tr[zm1]) + tr[zm1];
    double grad = tr[c] - tr[zm1];
    double curv = 0.0;
    if (cz > 0.0)
      curv = db * (wc->w9[zm1] * tr[c] + wc->w9[c] * tr[zm2]) - 2.0 * tr[zm1];
    else if (cz < 0.0)
      curv = df * (wc->w9[c] * tr[zp1] + wc->w9[zp1] * tr[zm1]) - 2.0 * tr[c];
    Fz[c] = face - 0.5 * cz * grad - (1.0 - cz * cz) * curv / 6.0;
  }
}

/* tracers.c:1945-2026  QUICKEST, uniform grid */
static void qku_faces(geometry_t *g, win_priv_t *wc, const double *tr,
                      double *Fx, double *Fy, double *Fz) {
  int cc;
  memset(Fx, 0, g->n3_t * sizeof(double));
  memset(Fy, 0, g->n3_t * sizeof(double));
  memset(Fz, 0, g->n3_t * sizeof(double));
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    int xp1 = g->xp1[c], yp1 = g->yp1[c];
    int xm1 = g->xm1[c], ym1 = g->ym1[c];
    int xm2 = g->xm1[xm1], ym2 = g->ym1[ym1];
    double cx = wc->w6[c];
    double face = 0.5 * (tr[c] + tr[xm1]);
    double grad = tr[c] - tr[xm1];
    double curv = 0.0;
    if (cx > 0.0)
      curv = tr[c] - 2.0 * tr[xm1] + tr[xm2];
    else if (cx < 0.0)
      curv = tr[xp1] - 2.0 * tr[c] + tr[xm1];
    Fx[c] = face - 0.5 * cx * grad - (1.0 - cx * cx) * curv / 6.0;
    double cy = wc->w7[c];
    face = 0.5 * (tr[c] + tr[ym1]);
    grad = tr[c] - tr[ym1];
    curv = 0.0;
    if (cy > 0.0)
      curv = tr[c] - 2.0 * tr[ym1] + tr[ym2];
    else if (cy < 0.0)
      curv = tr[yp1] - 2.0 * tr[c] + tr[ym1];
    Fy[c] = face - 0.5 * cy * grad - (1.0 - cy * cy) * curv / 6.0;
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zp1 = g->zp1[c], zm1 = g->zm1[c];
    int zm2 = g->zm1[zm1];
    double cz = wc->w8[c];
    double face = 0.5 * (tr[c] + tr[zm1]);
    double grad = tr[c] - tr[zm1];
    double curv = 0.0;
    if (cz > 0.0)
      curv = tr[c] - 2.0 * tr[zm1] + tr[zm2];
    else if (cz < 0.0)
      curv = tr[zp1] - 2.0 * tr[c] + tr[zm1];
    Fz[c] = face - 0.5 * cz * grad - (1.0 - cz * cz) * curv / 6.0;
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:1586-1620  first-order upwind (fluxes, not face values)  */
/*-------------------------------------------------------------------*/
static void o1_faces(geometry_t *g, window_t *wd, win_priv_t *wc,
                     const double *tr, double *Fx, double *Fy, double *Fz,
                     double dtu) {
  int cc;
  for (cc = 1; cc <= g->n3_t; cc++) {
    int c = g->w3_t[cc];
    Fx[c] = wc->w6[c] * tr[g->xm1[c]] + wc->w7[c] * tr[c];
    Fy[c] = wc->w8[c] * tr[g->ym1[c]] + wc->w9[c] * tr[c];
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zm1 = g->zm1[c];
    double w = wd->w[c];
    Fz[c] = dtu * (0.5 * (w + fabs(w)) * tr[zm1] + 0.5 * (w - fabs(w)) * tr[c]);
  }
}

/* tracers.c:1634-1727  second order, non-uniform */
static void o2_do(double *F, const double *tr, const double *mgs, int ss,
                  int se, const int *sdo, const int *bmap, const int *m2d) {
  int cc;
  for (cc = ss; cc <= se; cc++) {
    int c = sdo[cc];
    int bm1 = bmap[c];
    F[c] = mgs[m2d[c]] * (tr[c] - tr[bm1]) + tr[bm1];
  }
}
static void o2_faces(geometry_t *g, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fy, double *Fz) {
  int cc;
  memset(Fx, 0, g->n3_t * sizeof(double));
  o2_do(Fx, tr, wc->d3, 1, g->v3_e1, g->w3_e1, g->xm1, g->m2d);
  o2_do(Fx, tr, wc->d3, g->v3_t + 1, g->a3_t, g->w3_t, g->xm1, g->m2d);
  memset(Fy, 0, g->n3_t * sizeof(double));
  o2_do(Fy, tr, wc->d4, 1, g->v3_e2, g->w3_e2, g->ym1, g->m2d);
  o2_do(Fy, tr, wc->d4, g->v3_t + 1, g->a3_t, g->w3_t, g->ym1, g->m2d);
  memset(Fz, 0, g->n3_t * sizeof(double));
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zm1 = g->zm1[c];
    Fz[c] = wc->w5[c] * (tr[c] - tr[zm1]) + tr[zm1];
  }
}

/* tracers.c:1741-1810  fourth order, uniform grid */
static void o4_do(double *F, const double *tr, int ss, int se,
                  const int *sdo, const int *bmap, const int *fmap) {
  int cc;
  for (cc = ss; cc <= se; cc++) {
    int c = sdo[cc];
    int bm1 = bmap[c];
    int bm2 = bmap[bm1];
    int fp1 = fmap[c];
    F[c] = (7.0 * (tr[bm1] + tr[c]) - (tr[fp1] + tr[bm2])) / 12.0;
  }
}
static void o4_faces(geometry_t *g, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fy, double *Fz) {
  memset(Fx, 0, g->n3_t * sizeof(double));
  o4_do(Fx, tr, 1, g->v3_e1, g->w3_e1, g->xm1, g->xp1);
  o4_do(Fx, tr, g->v3_t + 1, g->a3_t, g->w3_t, g->xm1, g->xp1);
  memset(Fy, 0, g->n3_t * sizeof(double));
  o4_do(Fy, tr, 1, g->v3_e2, g->w3_e2, g->ym1, g->yp1);
  o4_do(Fy, tr, g->v3_t + 1, g->a3_t, g->w3_t, g->ym1, g->yp1);
  memset(Fz, 0, g->n3_t * sizeof(double));
  o4_do(Fz, tr, 1, wc->vc, wc->s1, g->zm1, g->zp1);
}

/* tracers.c:3010-3036  second-order upwind face value (Leonard 1994) */
static void o2uw_do(double *F, const double *tr, const double *vel, const double *cn,
                    int ss, int se, const int *sdo, const int *fmap, const int *bmap) {
  int cc;
  for (cc = ss; cc <= se; cc++) {
    int c = sdo[cc];
    int bm1 = bmap[c];
    int bm2 = bmap[bm1];
    if (vel[c] > 0.0)
      F[c] = 0.5 * (3.0 - cn[c]) * tr[bm1] - 0.5 * (1.0 - cn[c]) * tr[bm2];
    else if (vel[c] < 0.0)
      F[c] = 0.5 * (3.0 - cn[c]) * tr[c] - 0.5 * (1.0 - cn[c]) * tr[fmap[c]];
    else
      F[c] = 0.0;
  }
}

/* tracers.c:2963-2997  order2_upwind: wet faces, auxiliary faces, s1 */
static void o2uw_faces(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                       double *Fx, double *Fy, double *Fz) {
  memset(Fx, 0, g->n3_t * sizeof(double));
  o2uw_do(Fx, tr, wd->u1, wc->w6, 1, g->v3_e1, g->w3_e1, g->xp1, g->xm1);
  o2uw_do(Fx, tr, wd->u1, wc->w6, g->v3_t + 1, g->a3_t, g->w3_t, g->xp1, g->xm1);
  memset(Fy, 0, g->n3_t * sizeof(double));
  o2uw_do(Fy, tr, wd->u2, wc->w7, 1, g->v3_e2, g->w3_e2, g->yp1, g->ym1);
  o2uw_do(Fy, tr, wd->u2, wc->w7, g->v3_t + 1, g->a3_t, g->w3_t, g->yp1, g->ym1);
  memset(Fz, 0, g->n3_t * sizeof(double));
  o2uw_do(Fz, tr, wd->w, wc->w8, 1, wc->vc, wc->s1, g->zp1, g->zm1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:3501-3557  ULTIMATE limiter on face values               */
/*-------------------------------------------------------------------*/
static void ultimate(const double *cr, double *Fx, const double *F,
                     const int *ctp, int vs, int ve, const int *mapm,
                     const int *mapp) {
  int cc;
  for (cc = vs; cc <= ve; cc++) {
    int c = ctp[cc];
    double tc, tu, td, diff, nc, nu;
    if (cr[c] == 0.0) continue;
    if (cr[c] > 0.0) {
      int m1 = mapm[c];
      tc = F[m1];
      tu = F[mapm[m1]];
      td = F[c];
    } else {
      tc = F[c];
      tu = F[mapp[c]];
      td = F[mapm[c]];
    }
    diff = td - tu;
    nc = nu = 0.0;
    if (diff) {
      nc = (tc - tu) / diff;
      nu = (Fx[c] - tu) / diff;
    }
    if (nc >= 0.0 && nc <= 1.0) {
      if (nu > 1.0)
        Fx[c] = td;
      else if (nu < nc)
        Fx[c] = tu + nc * (td - tu);
      else if (nu > nc / fabs(cr[c]))
        Fx[c] = tu + nc * (td - tu) / fabs(cr[c]);
    } else
      Fx[c] = tu + nc * (td - tu);
  }
}

/*-------------------------------------------------------------------*/
/* Flux-form semi-Lagrangian scheme (FFSL), structured grid.          */
/* Trajectory arrays: per sub-step, tracer independent (the reference */
/* keeps them in win_priv_t; here they are the oracle's own).         */
/*-------------------------------------------------------------------*/
static int ff_n = 0;
static double *ff_nu, *ff_nv, *ff_nw;
static int *ff_clxf, *ff_clyf, *ff_clzf, *ff_clxc, *ff_clyc, *ff_clzc;
static double *ff_crfxf, *ff_crfyf, *ff_crfzf, *ff_crfxc, *ff_crfyc, *ff_crfzc;

static void ff_alloc(int n) {
  if (ff_n == n) return;
  if (ff_n) {
    free(ff_nu); free(ff_nv); free(ff_nw);
    free(ff_clxf); free(ff_clyf); free(ff_clzf);
    free(ff_clxc); free(ff_clyc); free(ff_clzc);
    free(ff_crfxf); free(ff_crfyf); free(ff_crfzf);
    free(ff_crfxc); free(ff_crfyc); free(ff_crfzc);
  }
  ff_nu = calloc(n, sizeof(double)); ff_nv = calloc(n, sizeof(double));
  ff_nw = calloc(n, sizeof(double));
  ff_clxf = calloc(n, sizeof(int)); ff_clyf = calloc(n, sizeof(int));
  ff_clzf = calloc(n, sizeof(int)); ff_clxc = calloc(n, sizeof(int));
  ff_clyc = calloc(n, sizeof(int)); ff_clzc = calloc(n, sizeof(int));
  ff_crfxf = calloc(n, sizeof(double)); ff_crfyf = calloc(n, sizeof(double));
  ff_crfzf = calloc(n, sizeof(double)); ff_crfxc = calloc(n, sizeof(double));
  ff_crfyc = calloc(n, sizeof(double)); ff_crfzc = calloc(n, sizeof(double));
  ff_n = n;
}

/* the trajectory arrays of the last sub-step, for tests */
void oracle_ffsl_maps(int **cl, double **crf) {
  cl[0] = ff_clxf; cl[1] = ff_clyf; cl[2] = ff_clzf;
  cl[3] = ff_clxc; cl[4] = ff_clyc; cl[5] = ff_clzc;
  crf[0] = ff_crfxf; crf[1] = ff_crfyf; crf[2] = ff_crfzf;
  crf[3] = ff_crfxc; crf[4] = ff_crfyc; crf[5] = ff_crfzc;
}

/* tracers.c:2771-2808  horizontal backward trajectory */
static void ff_do_hor(const double *vel, double dt, const double *h, const int *m2d,
                      int ss, int se, const int *sdo, const int *fmap, const int *bmap,
                      int *cl, double *crf) {
  int c, cc, c2, cp;
  double ltraj, lrem, dx;
  for (cc = ss; cc <= se; cc++) {
    c = sdo[cc];
    if (vel[c] == 0.0) continue;
    ltraj = fabs(vel[c]) * dt;
    lrem = ltraj;
    cp = cl[c];
    c2 = m2d[cp];
    dx = h[c2];
    while (dx < lrem) {
      lrem -= dx;
      cp = (vel[c] < 0.0) ? fmap[cp] : bmap[cp];
      c2 = m2d[cp];
      if (h[c2] > 0) dx = h[c2];
    }
    crf[c] = max(min(lrem / dx, 1.0), 0.0);
    cl[c] = cp;
  }
}

/* tracers.c:2820-2859  vertical backward trajectory */
static void ff_do_vert(const double *vel, double dt, const double *h, int ss, int se,
                       const int *sdo, const int *fmap, const int *bmap, int *cl,
                       double *crf) {
  int c, cc, cp, cpb;
  double ltraj, lrem, dz;
  for (cc = ss; cc <= se; cc++) {
    c = sdo[cc];
    if (vel[c] == 0.0) continue;
    ltraj = fabs(vel[c]) * dt;
    lrem = ltraj;
    cp = cpb = cl[c];
    dz = fabs((h[cp] > 0) ? h[cp] : h[c]);
    while (dz < lrem) {
      lrem -= dz;
      cp = (vel[c] < 0.0) ? fmap[cp] : bmap[cp];
      if (cpb == cp) {
        lrem = fmod(lrem, dz);
        break;
      }
      if (h[cp] > 0) dz = fabs(h[cp]);
      cpb = cp;
    }
    crf[c] = max(min(lrem / dz, 1.0), 0.0);
    cl[c] = cp;
  }
}

/* tracers.c:2132-2251  trajectories of one sub-step */
static void oracle_prep_ff_sl(geometry_t *g, window_t *wd, win_priv_t *wc, double dt) {
  int c, cc;
  const size_t nd = (size_t)g->sgsiz * sizeof(double), ni = (size_t)g->sgsiz * sizeof(int);
  memset(ff_crfxf, 0, nd); memset(ff_crfyf, 0, nd); memset(ff_crfzf, 0, nd);
  memset(ff_crfxc, 0, nd); memset(ff_crfyc, 0, nd); memset(ff_crfzc, 0, nd);
  memset(ff_clxf, 0, ni); memset(ff_clyf, 0, ni); memset(ff_clzf, 0, ni);
  memset(ff_clxc, 0, ni); memset(ff_clyc, 0, ni); memset(ff_clzc, 0, ni);
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    ff_clxc[c] = (ff_nu[c] < 0.0) ? g->xp1[c] : g->w3_t[cc];
  }
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    ff_clyc[c] = (ff_nv[c] < 0.0) ? g->yp1[c] : g->w3_t[cc];
  }
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    ff_clzc[c] = (ff_nw[c] < 0.0) ? g->zp1[c] : g->w3_t[cc];
  }
  ff_do_hor(ff_nu, dt, g->h1au1, g->m2d, 1, g->b3_t, g->w3_t, g->xp1, g->xm1, ff_clxc, ff_crfxc);
  ff_do_hor(ff_nv, dt, g->h2au2, g->m2d, 1, g->b3_t, g->w3_t, g->yp1, g->ym1, ff_clyc, ff_crfyc);
  ff_do_vert(ff_nw, dt, wc->w9, 1, g->b3_t, g->w3_t, g->zp1, g->zm1, ff_clzc, ff_crfzc);
  for (cc = 1; cc <= g->b3_e1; cc++) {
    c = g->w3_e1[cc];
    ff_clxf[c] = (wd->u1[c] > 0.0) ? g->xm1[c] : g->w3_e1[cc];
  }
  ff_do_hor(wd->u1, dt, g->h1acell, g->m2d, 1, g->b3_e1, g->w3_e1, g->xp1, g->xm1, ff_clxf,
            ff_crfxf);
  for (cc = 1; cc <= g->b3_e2; cc++) {
    c = g->w3_e2[cc];
    ff_clyf[c] = (wd->u2[c] > 0.0) ? g->ym1[c] : g->w3_e2[cc];
  }
  ff_do_hor(wd->u2, dt, g->h2acell, g->m2d, 1, g->b3_e2, g->w3_e2, g->yp1, g->ym1, ff_clyf,
            ff_crfyf);
  for (cc = 1; cc <= wc->vc; cc++) {
    c = wc->s1[cc];
    ff_clzf[c] = (wd->w[c] > 0.0) ? g->zm1[c] : wc->s1[cc];
  }
  ff_do_vert(wd->w, dt, wc->dz, 1, wc->vc, wc->s1, g->zp1, g->zm1, ff_clzf, ff_crfzf);
}

/* tracers.c:2717-2726  monotone clamp of a transverse interpolation */
static double ff_monoco(int c, int cm, int cp, const double *tr, const double *trmod,
                        double sc) {
  double v, mn, mx;
  mn = min(trmod[cm], trmod[cp]);
  mx = max(trmod[cm], trmod[cp]);
  v = max(mn * sc, tr[c]);
  v = min(mx * sc, v);
  return v;
}

/* tracers.c:2868-2910  van Leer face value on the fractional part */
static void ff_van_leer(double *F, const double *tr, const double *vel, const double *crf,
                        const int *cl, int ss, int se, const int *sdo, const int *fmap,
                        const int *bmap) {
  int c, cc, cp, bm1, bm2;
  double df;
  for (cc = ss; cc <= se; cc++) {
    c = sdo[cc];
    if (crf[c] == 0.0) continue;
    cp = cl[c];
    if (vel[c] > 0.0) cp = fmap[cp];
    bm1 = bmap[cp];
    bm2 = bmap[bm1];
    if (vel[c] > 0.0) {
      df = (tr[bm1] - tr[bm2]) * (tr[cp] - tr[bm1]);
      if (df > 0.0)
        df = 2.0 * df / (tr[cp] - tr[bm2]);
      else
        df = 0.0;
      F[c] = tr[bm1] + 0.5 * (1.0 - crf[c]) * df;
    } else if (vel[c] < 0.0) {
      df = (tr[cp] - tr[bm1]) * (tr[fmap[cp]] - tr[cp]);
      if (df > 0.0)
        df = 2.0 * df / (tr[fmap[cp]] - tr[bm1]);
      else
        df = 0.0;
      F[c] = tr[cp] - 0.5 * (1.0 - crf[c]) * df;
    }
  }
}

/* interpolation of a along the trajectory (cl, crf) of c in the
 * direction of velocity v: the transverse term of tracers.c:2508-2565 */
static double ff_tv(const double *a, double v, int clc, double crfc, const int *bmap,
                    int *upo, int *cpo) {
  int up1 = (v < 0.0) ? clc : bmap[clc];
  int cp = (v < 0.0) ? bmap[up1] : clc;
  *upo = up1;
  *cpo = cp;
  return a[cp] + (crfc * (a[up1] - a[cp]));
}

/* tracers.c:2447-2706  FFSL fluxes of one tracer (final fluxes: the
 * volume flux and the vertical sign are applied here) */
static void oracle_ff_sl3d(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                           double *Fx, double *Fy, double *Fz, double *tr_mod,
                           double *tr_mod_x, double *tr_mod_y, double *tr_mod_z,
                           double dt) {
  const double onesixth = 1.0 / 6.0;
  const double *u1 = wd->u1, *u2 = wd->u2, *w = wd->w;
  const double *u1vm = wd->u1flux3d, *u2vm = wd->u2flux3d;
  const size_t nd = (size_t)g->sgsiz * sizeof(double);
  int c, cc, cp, up1, c2;
  double dx, dy, dz, ltraj, lrem;
  memset(tr_mod, 0, nd);
  memset(Fx, 0, nd);
  memset(Fy, 0, nd);
  memset(Fz, 0, nd);
  /* transverse terms (x, y, z) with the monotone clamp */
  memcpy(tr_mod_x, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    up1 = (ff_nu[c] < 0.0) ? ff_clxc[c] : g->xm1[ff_clxc[c]];
    cp = (ff_nu[c] < 0.0) ? g->xm1[up1] : ff_clxc[c];
    tr_mod_x[c] = tr[cp];
    if (ff_crfxc[c] > 0.) tr_mod_x[c] += (ff_crfxc[c] * (tr[up1] - tr[cp]));
    tr_mod_x[c] = ff_monoco(c, cp, up1, tr_mod_x, tr, 1.0);
  }
  memcpy(tr_mod_y, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    up1 = (ff_nv[c] < 0.0) ? ff_clyc[c] : g->ym1[ff_clyc[c]];
    cp = (ff_nv[c] < 0.0) ? g->ym1[up1] : ff_clyc[c];
    tr_mod_y[c] = tr[cp];
    if (ff_crfyc[c] > 0.) tr_mod_y[c] += (ff_crfyc[c] * (tr[up1] - tr[cp]));
    tr_mod_y[c] = ff_monoco(c, cp, up1, tr_mod_y, tr, 1.0);
  }
  memcpy(tr_mod_z, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    up1 = (ff_nw[c] < 0.0) ? ff_clzc[c] : g->zm1[ff_clzc[c]];
    cp = (ff_nw[c] < 0.0) ? g->zm1[up1] : ff_clzc[c];
    tr_mod_z[c] = tr[cp];
    if (ff_crfzc[c] > 0.) tr_mod_z[c] += (ff_crfzc[c] * (tr[up1] - tr[cp]));
    tr_mod_z[c] = ff_monoco(c, cp, up1, tr_mod_z, tr, 1.0);
  }
  /* x direction */
  memcpy(tr_mod, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    int u, q;
    c = g->w3_t[cc];
    tr_mod[c] += (tr[c] + tr_mod_y[c] + tr_mod_z[c]);
    tr_mod[c] += ff_tv(tr_mod_y, ff_nw[c], ff_clzc[c], ff_crfzc[c], g->zm1, &u, &q);
    tr_mod[c] += ff_tv(tr_mod_z, ff_nv[c], ff_clyc[c], ff_crfyc[c], g->ym1, &u, &q);
    tr_mod[c] *= onesixth;
  }
  ff_van_leer(Fx, tr_mod, u1, ff_crfxf, ff_clxf, 1, g->b3_e1, g->w3_e1, g->xp1, g->xm1);
  for (cc = 1; cc <= g->b3_e1; cc++) {
    c = g->w3_e1[cc];
    if (u1[c] == 0.0) {
      Fx[c] = 0.0;
      continue;
    }
    cp = ff_clxf[c];
    c2 = g->m2d[cp];
    Fx[c] *= (ff_crfxf[c] * g->h1acell[c2]);
    cp = (u1[c] > 0.0) ? g->xm1[c] : g->w3_e1[cc];
    c2 = g->m2d[cp];
    dx = g->h1acell[c2];
    ltraj = fabs(u1[c] * dt);
    lrem = ltraj;
    while (dx < lrem) {
      Fx[c] += (dx * tr_mod[cp]);
      lrem -= dx;
      cp = (u1[c] < 0) ? g->xp1[cp] : g->xm1[cp];
      c2 = g->m2d[cp];
      if (g->h1acell[c2] > 0) dx = g->h1acell[c2];
    }
    Fx[c] *= (u1vm[c] / ltraj);
  }
  /* y direction */
  memcpy(tr_mod, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    int u, q;
    c = g->w3_t[cc];
    tr_mod[c] += (tr[c] + tr_mod_x[c] + tr_mod_z[c]);
    tr_mod[c] += ff_tv(tr_mod_x, ff_nw[c], ff_clzc[c], ff_crfzc[c], g->zm1, &u, &q);
    tr_mod[c] += ff_tv(tr_mod_z, ff_nu[c], ff_clxc[c], ff_crfxc[c], g->xm1, &u, &q);
    tr_mod[c] *= onesixth;
  }
  ff_van_leer(Fy, tr_mod, u2, ff_crfyf, ff_clyf, 1, g->b3_e2, g->w3_e2, g->yp1, g->ym1);
  for (cc = 1; cc <= g->b3_e2; cc++) {
    c = g->w3_e2[cc];
    if (u2[c] == 0.0) {
      Fy[c] = 0.0;
      continue;
    }
    cp = ff_clyf[c];
    c2 = g->m2d[cp];
    Fy[c] *= (ff_crfyf[c] * g->h2acell[c2]);
    cp = (u2[c] > 0.0) ? g->ym1[c] : g->w3_e2[cc];
    c2 = g->m2d[cp];
    dy = g->h2acell[c2];
    ltraj = fabs(u2[c] * dt);
    lrem = ltraj;
    while (dy < lrem) {
      Fy[c] += (dy * tr_mod[cp]);
      lrem -= dy;
      cp = (u2[c] < 0) ? g->yp1[cp] : g->ym1[cp];
      c2 = g->m2d[cp];
      if (g->h2acell[c2] > 0) dy = g->h2acell[c2];
    }
    Fy[c] *= (u2vm[c] / ltraj);
  }
  /* z direction */
  memcpy(tr_mod, tr, nd);
  for (cc = 1; cc <= g->b3_t; cc++) {
    int u, q;
    c = g->w3_t[cc];
    tr_mod[c] += (tr[c] + tr_mod_x[c] + tr_mod_y[c]);
    tr_mod[c] += ff_tv(tr_mod_y, ff_nu[c], ff_clxc[c], ff_crfxc[c], g->xm1, &u, &q);
    tr_mod[c] += ff_tv(tr_mod_x, ff_nv[c], ff_clyc[c], ff_crfyc[c], g->ym1, &u, &q);
    tr_mod[c] *= onesixth;
  }
  ff_van_leer(Fz, tr_mod, w, ff_crfzf, ff_clzf, 1, wc->vc, wc->s1, g->zp1, g->zm1);
  for (cc = 1; cc <= wc->vc; cc++) {
    c = wc->s1[cc];
    if (w[c] == 0.0) {
      Fz[c] = 0.0;
      continue;
    }
    cp = ff_clzf[c];
    Fz[c] *= (ff_crfzf[c] * wc->dz[cp]);
    cp = (w[c] > 0.0) ? g->zm1[c] : wc->s1[cc];
    dz = fabs(wc->dz[cp]);
    lrem = fabs(w[c] * dt);
    while (dz < lrem) {
      Fz[c] += (dz * tr_mod[cp]);
      lrem -= dz;
      cp = (w[c] < 0.0) ? g->zp1[cp] : g->zm1[cp];
      if (wc->dz[cp] > 0) dz = wc->dz[cp];
    }
    if (w[c] < 0) Fz[c] *= -1;
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:1487-1576  face values -> fluxes                         */
/*-------------------------------------------------------------------*/
static void oracle_advect(geometry_t *g, window_t *wd, win_priv_t *wc,
                          const double *tr, double *Fx, double *Fy,
                          double *Fz, double *tm, double *tmx, double *tmy,
                          double *tmz, double dtu) {
  int cc;
  if (wd->dttr == 0.0) return;
  if (wc->trasc == FFSL) {
    oracle_ff_sl3d(g, wd, wc, tr, Fx, Fy, Fz, tm, tmx, tmy, tmz, dtu);
    return;
  }
  if (wc->trasc == ORDER1) {
    o1_faces(g, wd, wc, tr, Fx, Fy, Fz, dtu);
    return;
  }
  if (wc->trasc == ORDER2)
    o2_faces(g, wc, tr, Fx, Fy, Fz);
  else if (wc->trasc == ORDER4)
    o4_faces(g, wc, tr, Fx, Fy, Fz);
  else if (wc->trasc == QUICKEST)
    qk_faces(g, wc, tr, Fx, Fy, Fz);
  else if (wc->trasc == QUICKEST_CO)
    qku_faces(g, wc, tr, Fx, Fy, Fz);
  else if (wc->trasc == VANLEER)
    vl_faces(g, wd, wc, tr, Fx, Fy, Fz);
  else if (wc->trasc == ORDER2_UW)
    o2uw_faces(g, wd, wc, tr, Fx, Fy, Fz);

  if (wc->ultimate) {
    ultimate(wc->w6, Fx, tr, g->w3_e1, 1, g->v3_e1, g->xm1, g->xp1);
    ultimate(wc->w6, Fx, tr, g->w3_t, g->v3_t + 1, g->a3_t, g->xm1, g->xp1);
    ultimate(wc->w7, Fy, tr, g->w3_e2, 1, g->v3_e2, g->ym1, g->yp1);
    ultimate(wc->w7, Fy, tr, g->w3_t, g->v3_t + 1, g->a3_t, g->ym1, g->yp1);
    ultimate(wc->w8, Fz, tr, wc->s1, 1, wc->vc, g->zm1, g->zp1);
  }
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    Fx[c] *= wd->u1flux3d[c];
    Fy[c] *= wd->u2flux3d[c];
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    Fz[c] *= (wd->w[c] * dtu);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:3595-3668  horizontal diffusion (z-levels)               */
/*-------------------------------------------------------------------*/
static void oracle_hor_diffuse(geometry_t *g, window_t *wd, win_priv_t *wc,
                               const double *tr, double *Fx, double *Fy) {
  int cc;
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    int c2 = g->m2d[c];
    int xm1 = g->xm1[c], ym1 = g->ym1[c];
    double csx = wd->dzu1[c] * g->h2au1[c2] * wc->mdx[c2];
    double csy = wd->dzu2[c] * g->h1au2[c2] * wc->mdy[c2];
    Fx[c] -= (csx * wc->u1kh[c] * (tr[c] - tr[xm1]) / g->h1au1[c2]);
    Fy[c] -= (csy * wc->u2kh[c] * (tr[c] - tr[ym1]) / g->h2au2[c2]);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:3047-3207  surface-cell concentration by mass            */
/*-------------------------------------------------------------------*/
static double oracle_surf_conc(geometry_t *g, window_t *wd, win_priv_t *wc,
                               const double *tr, double subeta, int c,
                               int c2, int cc, double fcbot,
                               const double *dtracer) {
  int cb = wc->i1[cc];
  int co = wc->i3[cc];
  double osubeta = wc->d1[c2];
  double waterbot = (c == cb) ? g->botz[c2] : g->gridz[c];
  double dz = subeta - waterbot;
  double svol = dz * g->cellarea[c2];
  double surftr = -fcbot;
  double watertop;
  (void)wd;
  while (c != co) {
    int zp1 = g->zp1[c];
    watertop = g->gridz[zp1];
    surftr += (tr[c] * g->cellarea[c2] * (watertop - waterbot) - dtracer[c]);
    c = zp1;
    waterbot = g->gridz[c];
  }
  watertop = osubeta;
  waterbot = (c == cb) ? g->botz[c2] : g->gridz[c];
  surftr += (tr[c] * g->cellarea[c2] * (watertop - waterbot) - dtracer[c]);
  if (c != c2) {
    do {
      c = g->zp1[c];
      surftr -= (dtracer[c]);
    } while (c != c2);
  }
  if (svol == 0.0 || (surftr < 0.0 && dz <= wc->hmin))
    surftr = 0.0;
  else
    surftr /= svol;
  return surftr;
}

/*-------------------------------------------------------------------*/
/* tracers.c:3217-3324  range repair, clip, copy to layers above      */
/*-------------------------------------------------------------------*/
void oracle_tr_bounds(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int cc, nn;
  double eps = 1e-5;
  double *mn = wc->mintr, *mx = wc->maxtr;
  for (cc = 1; cc <= wc->vcs1; cc++) {
    int c = wc->s1[cc], c3 = c;
    int co = wc->i3[cc];
    int cb = wc->i1[cc];
    if (c < cb && wc->dz[co] < wc->hmin) {
      for (nn = 0; nn < wc->ntbdy; nn++) {
        int n = wc->tbdy[nn];
        double *tr = wd->tr_wc[n];
        if (tr[c] < mn[n] || tr[c] > mx[n]) {
          double surftr = tr[g->zm1[c]];
          c3 = c;
          tr[c3] = surftr;
          while (c3 != g->zp1[c3]) {
            c3 = g->zp1[c3];
            tr[c3] = surftr;
          }
        }
      }
    }
  }
  for (cc = 1; cc <= wc->vc1; cc++) {
    int c = wc->s1[cc];
    for (nn = 0; nn < wc->ntbdy; nn++) {
      int n = wc->tbdy[nn];
      double *tr = wd->tr_wc[n];
      if (mn[n] <= eps && tr[c] < mn[n]) tr[c] = mn[n];
    }
  }
  if (!wc->sigma) {
    for (cc = 1; cc <= wc->vcs; cc++) {
      int c = wc->s1[cc], cs = c;
      while (c != g->zp1[c]) {
        c = g->zp1[c];
        for (nn = 0; nn < wc->ntbdy; nn++) {
          int n = wc->tbdy[nn];
          wd->tr_wc[n][c] = wd->tr_wc[n][cs];
        }
      }
    }
  }
}

/*-------------------------------------------------------------------*/
/* One tracer through one sub-step: the body of the tracer loop of    */
/* advect_diffuse, tracers.c:773-1000 (closed boundaries, z-levels).  */
/*-------------------------------------------------------------------*/
typedef struct {
  double *Fx, *Fy, *Fz, *dtracer;
  double *tm, *tmx, *tmy, *tmz; /* FFSL modified tracers */
} scratch_t;

static void tracer_substep(geometry_t *g, window_t *wd, win_priv_t *wc,
                           int n, double dtu, int slf, int ksf,
                           const double *subeta, scratch_t *s) {
  int cc;
  double *tr = wd->tr_wc[n];
  double *Fx = s->Fx, *Fy = s->Fy, *Fz = s->Fz, *dtracer = s->dtracer;
  const size_t nb = (size_t)g->sgsiz * sizeof(double);
  int kef = wc->vc1, kefs = wc->vcs1;
  tr[0] = (double)n;
  memset(Fx, 0, nb);
  memset(Fy, 0, nb);
  memset(Fz, 0, nb);
  memset(dtracer, 0, nb);
  /* bed: no-gradient into the sediment ghost */
  for (cc = 1; cc <= g->b2_t; cc++) {
    int cb = g->bot_t[cc];
    tr[g->zm1[cb]] = tr[cb];
  }
  /* tracers.c:836-841: FFSL also fills the sediment ghosts from above */
  if (wc->trasc == FFSL) {
    for (cc = 1; cc <= g->ngsed; cc++) {
      int cb = g->gsed_t[cc];
      tr[cb] = tr[g->zp1[cb]];
    }
  }
  if (wc->advect[n])
    oracle_advect(g, wd, wc, tr, Fx, Fy, Fz, s->tm, s->tmx, s->tmy, s->tmz, dtu);
  if (wc->diffuse[n]) oracle_hor_diffuse(g, wd, wc, tr, Fx, Fy);
  for (cc = 1; cc <= g->b3_t; cc++) {
    int c = g->w3_t[cc];
    dtracer[c] = (Fx[g->xp1[c]] - Fx[c] + Fy[g->yp1[c]] - Fy[c]) * dtu;
  }
  /* tracers.c:902-903 */
  if (wc->trasc == FFSL && wd->nstep == 0) memset(dtracer, 0, nb);
  for (cc = ksf; cc <= kef; cc++) {
    int c = wc->s1[cc];
    int c2 = g->m2d[c];
    int zp1 = g->zp1[c];
    if (slf)
      tr[c] *= wc->Ds[c2];
    else
      tr[c] *= wc->Hn1[c2];
    tr[c] -= ((dtracer[c] / g->cellarea[c2] + (Fz[zp1] - Fz[c])) / wc->dz[c]);
    tr[c] /= wc->Hn1[c2];
  }
  if (!wc->sigma && wc->advect[n]) {
    for (cc = 1; cc <= kefs; cc++) {
      int c = wc->s1[cc];
      int c2 = g->m2d[c];
      double top = -Fz[c] * g->cellarea[c2];
      double surftr = oracle_surf_conc(g, wd, wc, tr, subeta[c2], c, c2, cc,
                                       top, dtracer);
      tr[c] = (g->botz[c2] > 0.0 && wc->dz[c] < wc->hmin) ? tr[c] : surftr;
    }
  }
}

/* Per-thread flux work arrays, allocated once per (threads, size) as
 * the reference allocates its w1n..w4n once (tracers.c:788-801). */
static scratch_t *scr = NULL;
static int scr_nt = 0, scr_n = 0;
static scratch_t *scratch_get(int nt, int n) {
  if (scr && scr_nt >= nt && scr_n == n) return scr;
  for (int t = 0; t < scr_nt; t++) {
    free(scr[t].Fx);
    free(scr[t].Fy);
    free(scr[t].Fz);
    free(scr[t].dtracer);
    free(scr[t].tm);
    free(scr[t].tmx);
    free(scr[t].tmy);
    free(scr[t].tmz);
  }
  free(scr);
  scr = (scratch_t *)calloc(nt, sizeof(scratch_t));
  for (int t = 0; t < nt; t++) {
    scr[t].Fx = (double *)calloc(n, sizeof(double));
    scr[t].Fy = (double *)calloc(n, sizeof(double));
    scr[t].Fz = (double *)calloc(n, sizeof(double));
    scr[t].dtracer = (double *)calloc(n, sizeof(double));
    scr[t].tm = (double *)calloc(n, sizeof(double));
    scr[t].tmx = (double *)calloc(n, sizeof(double));
    scr[t].tmy = (double *)calloc(n, sizeof(double));
    scr[t].tmz = (double *)calloc(n, sizeof(double));
  }
  scr_nt = nt;
  scr_n = n;
  return scr;
}

/*-------------------------------------------------------------------*/
/* tracers.c:401-1054  advect_diffuse (z-levels, closed boundaries)   */
/* nthreads > 1 runs the tracer loop with OpenMP and per-thread       */
/* scratch, as tracers.c:770-801 does.                                */
/*-------------------------------------------------------------------*/
int oracle_advect_diffuse_mt(geometry_t *g, window_t *wd, win_priv_t *wc,
                             int nthreads) {
  int cc, nn;
  double *osubeta = wc->d1, *subeta = wc->d2;
  double dtu, dtm, trem;
  double sf = 0.8, minval = 1e-10;
  int itermax = 20, slf = 1, courf, ksf;
  const int vc = wc->vc, vcs = wc->vcs;
  double *u1 = wd->u1, *u2 = wd->u2, *w = wd->w;
  oracle_err = 0;

  /* scheme work arrays, tracers.c:496-578 */
  if (wc->trasc == ORDER1) {
    wc->ultimate = 0;
    for (cc = 1; cc <= g->n3_t; cc++) {
      int c = g->w3_t[cc];
      wc->w6[c] = 0.5 * (wd->u1flux3d[c] + fabs(wd->u1flux3d[c]));
      wc->w7[c] = 0.5 * (wd->u1flux3d[c] - fabs(wd->u1flux3d[c]));
      wc->w8[c] = 0.5 * (wd->u2flux3d[c] + fabs(wd->u2flux3d[c]));
      wc->w9[c] = 0.5 * (wd->u2flux3d[c] - fabs(wd->u2flux3d[c]));
    }
  } else if (wc->trasc == VANLEER) {
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[g->zm1[c]]);
    }
  } else if (wc->trasc == FFSL) {
    /* tracers.c:524-559: cell-centred velocities (Leonard et al. 1996) */
    ff_alloc(g->sgsiz);
    memset(ff_nu, 0, g->sgsiz * sizeof(double));
    memset(ff_nv, 0, g->sgsiz * sizeof(double));
    memset(ff_nw, 0, g->sgsiz * sizeof(double));
    itermax = 200;
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      int xp1, yp1, zp1;
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[g->zm1[c]]);
      xp1 = g->xp1[c];
      if ((u1[c] >= 0.0) && (u1[xp1] > 0.0))
        ff_nu[c] = u1[c];
      else if ((u1[c] < 0.0) && (u1[xp1] <= 0.0))
        ff_nu[c] = u1[xp1];
      yp1 = g->yp1[c];
      if ((u2[c] >= 0.0) && (u2[yp1] > 0.0))
        ff_nv[c] = u2[c];
      else if ((u2[c] < 0.0) && (u2[yp1] <= 0.0))
        ff_nv[c] = u2[yp1];
      zp1 = g->zp1[c];
      if ((w[c] >= 0.0) && (w[zp1] > 0.0))
        ff_nw[c] = w[c];
      else if ((w[c] < 0.0) && (w[zp1] <= 0.0))
        ff_nw[c] = w[zp1];
    }
  } else {
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      int zm1 = g->zm1[c];
      wc->w5[c] = wc->dz[zm1] / (wc->dz[zm1] + wc->dz[c]);
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[zm1]);
    }
    for (cc = 1; cc <= g->n2_t; cc++) {
      int c = g->w2_t[cc];
      int c2 = g->m2d[c];
      int xm1 = g->xm1[c2], ym1 = g->ym1[c2];
      wc->d3[c2] = g->h1acell[xm1] / (g->h1acell[xm1] + g->h1acell[c2]);
      wc->d4[c2] = g->h2acell[ym1] / (g->h2acell[ym1] + g->h2acell[c2]);
    }
  }
  courf = (wc->ultimate || wc->trasc >= QUICKEST) ? 1 : 0;

  /* sub-step, tracers.c:590-668 */
  dtm = dtu = wd->dt;
  if (wc->stab & (SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER)) {
    for (cc = 1; cc <= vc; cc++) {
      int c = wc->s1[cc];
      int c2 = g->m2d[c];
      int xp1 = g->xp1[c], yp1 = g->yp1[c], zp1 = g->zp1[c];
      double d1, d2, d3, d4;
      /* FFSL is limited by the velocity gradients (Lipschitz number) */
      d4 = 0.5 * (u1[c] + u1[xp1]);
      if (wc->trasc == FFSL) d4 = max(fabs(u1[xp1] - u1[c]), wc->u1kh[c] / g->h1acell[c2]);
      if (fabs(d4) < minval) d4 = minval;
      d1 = sf * fabs(g->h1acell[c2] / d4);
      if (d1 < dtm) dtm = d1;
      d4 = 0.5 * (u2[c] + u2[yp1]);
      if (wc->trasc == FFSL) d4 = max(fabs(u2[yp1] - u2[c]), wc->u2kh[c] / g->h2acell[c2]);
      if (fabs(d4) < minval) d4 = minval;
      d2 = sf * fabs(g->h2acell[c2] / d4);
      if (d2 < dtm) dtm = d2;
      d4 = (cc <= vcs) ? 0.0 : 0.5 * (w[c] + w[zp1]);
      if (wc->trasc == FFSL) d4 = (cc <= vcs) ? 0.0 : fabs(w[zp1] - w[c]);
      if (fabs(d4) < minval) d4 = minval;
      d3 = (cc <= vcs) ? dtu : sf * fabs(wc->dz[c] * wc->Ds[c2] / d4);
      if (d3 > 0.0 && d3 < dtm) dtm = d3;
    }
    if (dtm != dtu) dtm = floor(dtm);
  }
  if (dtm == 0.0 || (int)(dtu / dtm) > itermax) {
    oracle_err = 1;
    return 1;
  }
  ksf = vcs + 1;

  for (cc = 1; cc <= vcs; cc++) {
    int c = wc->s1[cc];
    osubeta[g->m2d[c]] = wc->oldeta[g->m2d[c]];
  }

  trem = dtu;
  while (trem > 0) {
    dtu = dtm;
    if (trem < dtu) dtu = trem;
    wc->b1 = dtu;
    if (dtu > 0.0) {
      if (courf) {
        for (cc = 1; cc <= g->n3_t; cc++) {
          int c = g->w3_t[cc];
          int c2 = g->m2d[c];
          wc->w6[c] = u1[c] * dtu / g->h1au1[c2];
          wc->w7[c] = u2[c] * dtu / g->h2au2[c2];
          wc->w8[c] = w[c] * dtu / (wc->w9[c] * wc->Ds[c2]);
        }
      }
      for (cc = 1; cc <= vcs; cc++) {
        int c = wc->s1[cc];
        int c2 = g->m2d[c];
        subeta[c2] = wc->oldeta[c2] +
                     (1.0 + (dtu - trem) / wd->dt) * (wd->eta[c2] - wc->oldeta[c2]);
      }
      /* tracers.c:763-765 */
      if (wc->trasc == FFSL) oracle_prep_ff_sl(g, wd, wc, dtu);
      {
        int nt = nthreads > 0 ? nthreads : 1;
        if (nt > wc->ntbdy) nt = wc->ntbdy > 0 ? wc->ntbdy : 1;
        scratch_t *S = scratch_get(nt, g->sgsiz);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
        for (nn = 0; nn < wc->ntbdy; nn++) {
          int tid = 0;
#ifdef _OPENMP
          tid = omp_get_thread_num();
#endif
          tracer_substep(g, wd, wc, wc->tbdy[nn], dtu, slf, ksf, subeta, &S[tid]);
        }
      }
      oracle_tr_bounds(g, wd, wc);
      for (cc = 1; cc <= vcs; cc++) {
        int c = wc->s1[cc];
        int c2 = g->m2d[c];
        osubeta[c2] = subeta[c2];
      }
    }
    trem -= dtu;
    slf = 0;
  }
  return 0;
}

int oracle_advect_diffuse(geometry_t *g, window_t *wd, win_priv_t *wc) {
  return oracle_advect_diffuse_mt(g, wd, wc, 1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:4067-4214  implicit vertical diffusion, one tracer       */
/*-------------------------------------------------------------------*/
static int oracle_implicit_vdiff_tr(geometry_t *g, window_t *wd,
                                    win_priv_t *wc, double *var,
                                    const double *dzcell, const double *fb,
                                    const double *ft, const int *ctp,
                                    const int *cbt, int vcs,
                                    const double *scale, const double *C,
                                    const double *Cp1, const double *Cm1,
                                    double *work /* 7*(nz+2) */) {
  int cc, k;
  const int nzw = g->nz + 2;
  double *rhs = work, *sol = work + nzw, *ud = work + 2 * nzw;
  double *B = work + 3 * nzw, *Bm1 = work + 4 * nzw, *Bp1 = work + 5 * nzw;
  double *dz = work + 6 * nzw;
  double dt = wd->dt, dzdt, div;
  for (cc = 1; cc <= vcs; cc++) {
    int cs = ctp[cc], c = cs;
    int cb = cbt[cc];
    int c2 = g->m2d[c];
    int zm1 = g->zm1[c];
    int ks = g->s2k[cs], kb = g->s2k[cb];
    if (zm1 == g->zm1[zm1] || !dzcell[zm1]) {
      var[cs] += wd->dt * (fb[c2] - ft[c2]) / max(dzcell[cs], wc->hmin);
      if (c != wc->i4[cc]) var[wc->i4[cc]] = var[cs];
      continue;
    }
    c = cs;
    for (k = ks; k >= kb; k--) {
      B[k] = C[c];
      Bm1[k] = Cm1[c];
      Bp1[k] = Cp1[c];
      dz[k] = dzcell[c];
      c = g->zm1[c];
    }
    c = cb;
    dzdt = dz[kb] / dt;
    rhs[kb] = dzdt * var[cb] + fb[c2];
    c = g->zp1[cb];
    for (k = kb + 1; k < ks; k++) {
      dzdt = dz[k] / dt;
      rhs[k] = dzdt * var[c];
      c = g->zp1[c];
    }
    dzdt = dz[ks] / dt;
    rhs[ks] = dzdt * var[cs] - ft[c2];
    div = B[kb];
    sol[kb] = rhs[kb] / div;
    for (k = kb + 1; k <= ks; k++) {
      ud[k] = Bp1[k - 1] / div;
      div = B[k] - Bm1[k] * ud[k];
      if (div == 0.0) return cs;
      sol[k] = (rhs[k] - Bm1[k] * sol[k - 1]) / div;
    }
    c = cs;
    var[cs] += (sol[ks] - var[cs]) * scale[c2];
    for (k = ks - 1; k >= kb; k--) {
      c = g->zm1[c];
      sol[k] -= ud[k + 1] * sol[k + 1];
      var[c] += (sol[k] - var[c]) * scale[c2];
    }
    c = wc->i4[cc];
    if (c != cs) var[c] += (sol[ks] - var[cs]) * scale[c2];
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* tracers.c:3782-3957  vertical diffusion: coefficients once per     */
/* column, then one solve per tracer (no surface fluxes: topflux = 0) */
/*-------------------------------------------------------------------*/
int oracle_vert_diffuse_3d_mt(geometry_t *g, window_t *wd, win_priv_t *wc,
                              int nthreads) {
  int cc, nn, bad = 0;
  int *ctp = wc->i2, *cbt = wc->i5, *cth = wc->i4;
  double dt = wd->dt;
  double *dzface = wc->w7, *dzcell = wc->w9, *Kz = wc->w10;
  double *Cm1 = wc->w1, *C = wc->w2, *Cp1 = wc->w3;
  memset(wc->d1, 0, g->sgsizS * sizeof(double));
  memset(wc->d2, 0, g->sgsizS * sizeof(double));
  memset(cth, 0, g->sgsizS * sizeof(int));
  memcpy(dzcell, wc->dz, g->sgsiz * sizeof(double));
  memcpy(Kz, wd->Kz, g->sgsiz * sizeof(double));
  oracle_err = 0;

  for (cc = 1; cc <= wc->vcs; cc++) {
    int c, cs, c2, zm1, zp1, cb;
    double dzdt;
    c = cs = ctp[cc];
    c2 = g->m2d[c];
    zm1 = g->zm1[c];
    if (zm1 == g->zm1[zm1] || !dzcell[zm1]) continue;
    while (c != zm1 && dzcell[c]) {
      dzface[c] = 0.5 * (dzcell[zm1] + dzcell[c]);
      dzcell[c] *= wc->Ds[c2];
      Kz[c] /= wc->Ds[c2];
      c = zm1;
      zm1 = g->zm1[c];
    }
    cb = c = g->zp1[c];
    zp1 = g->zp1[c];
    cbt[cc] = cb;
    dzdt = dzcell[c] / dt;
    Cm1[c] = 0.0;
    Cp1[c] = -Kz[zp1] / dzface[zp1];
    C[c] = dzdt - Cp1[c];
    while (zp1 != cs) {
      c = zp1;
      zp1 = g->zp1[c];
      dzdt = dzcell[c] / dt;
      Cm1[c] = -Kz[c] / dzface[c];
      Cp1[c] = -Kz[zp1] / dzface[zp1];
      C[c] = dzdt - Cm1[c] - Cp1[c];
    }
    c = cs;
    dzdt = dzcell[c] / dt;
    Cm1[c] = -Kz[c] / dzface[c];
    Cp1[c] = 0.0;
    C[c] = dzdt - Cm1[c];
  }

  /* topflux = 0 at every column (no heat, salt or generic surface flux);
   * the thin-layer reset of tracers.c:3946-3949 leaves zeros at zero. */
  {
    int nt = nthreads > 0 ? nthreads : 1;
    if (nt > wc->ntdif_v) nt = wc->ntdif_v > 0 ? wc->ntdif_v : 1;
    const int nzw = g->nz + 2;
    double *work = (double *)calloc((size_t)nt * 7 * nzw, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1) reduction(max : bad)
#endif
    for (nn = 0; nn < wc->ntdif_v; nn++) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      int n = wc->tdif_v[nn];
      int r = oracle_implicit_vdiff_tr(g, wd, wc, wd->tr_wc[n], dzcell, wc->d1,
                                       wc->d2, ctp, cbt, wc->vcs, wc->one, C,
                                       Cp1, Cm1, work + (size_t)tid * 7 * nzw);
      if (r > bad) bad = r;
    }
    free(work);
  }
  if (bad) oracle_err = 4;
  return bad;
}

void oracle_vert_diffuse_3d(geometry_t *g, window_t *wd, win_priv_t *wc) {
  oracle_vert_diffuse_3d_mt(g, wd, wc, 1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:5900-5980  reset_dz: surface thicknesses for the updated */
/* elevation, cells to process s2 (z-levels) and the wet-cell mask c1 */
/*-------------------------------------------------------------------*/
void oracle_reset_dz(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int cc, c, cs, cn, zp1, zm1, vc, vcs;
  double top, bot;

  for (cc = 1; cc <= g->b2_t; cc++) {
    c = max(g->sur_t[cc], g->nsur_t[cc]);
    cs = g->m2d[c];
    cn = g->nsur_t[cc];
    zp1 = g->zp1[c];
    bot = max(g->gridz[c], g->botz[cs]);
    while (c > cn) {
      top = g->gridz[zp1];
      wc->dz[c] = top - bot;
      bot = top;
      c = zp1;
      zp1 = g->zp1[c];
    }
    top = wd->eta[cs];
    wc->dz[c] = (top > bot) ? top - bot : wc->dz[c];
  }

  if (!wc->sigma) {
    vc = 1;
    for (cc = 1; cc <= g->b2_t; cc++) {
      c = g->nsur_t[cc];
      cs = g->m2d[c];
      top = wd->eta[cs];
      bot = DRY_FRAC * wc->hmin + g->botz[cs];
      if (top > bot && c != g->zm1[c]) {
        wc->s2[vc] = c;
        wc->i1[vc] = g->bot_t[cc];
        if (cc == g->v2_t) wc->vca2 = vc;
        vc++;
      }
    }
    vcs = vc - 1;
    wc->vcs2 = vcs;
    for (cc = 1; cc <= vcs; cc++) {
      c = g->zm1[wc->s2[cc]];
      zm1 = g->zm1[c];
      while (c != zm1) {
        wc->s2[vc] = c;
        vc++;
        c = zm1;
        zm1 = g->zm1[c];
      }
      if (cc == wc->vca2) wc->vci2 = vc - 1;
    }
    vc--;
    wc->vc2 = vc;
  }

  memset(wc->c1, 0, g->sgsiz * sizeof(char));
  for (cc = 1; cc <= wc->vc2; cc++) wc->c1[wc->s2[cc]] = 1;
}

/* the age tracer's number (agetr points at one of the 3D tracers) */
static int age_tracer(window_t *wd) {
  int n;
  if (!wd->agetr) return -1;
  for (n = 0; n < wd->ntr; n++)
    if (wd->tr_wc[n] == wd->agetr) return n;
  return -1;
}

/*-------------------------------------------------------------------*/
/* tracers.c:6960-6973  calc_age                                      */
/*-------------------------------------------------------------------*/
void oracle_calc_age(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int c, cc;
  double s = 1.0 / 86400.0;
  for (cc = 1; cc <= g->b3_t; cc++) {
    c = g->w3_t[cc];
    if (wc->agemsk[c]) wd->agetr[c] += wd->dttr * s;
  }
}

/* tracer_decay.c:45-52: seconds per unit of a decay timescale */
static double units_scale(const char *u) {
  if (strcmp(u, "days") == 0 || strcmp(u, "day") == 0) return 86400.0;
  if (strcmp(u, "hours") == 0 || strcmp(u, "hour") == 0) return 3600.0;
  if (strcmp(u, "minutes") == 0 || strcmp(u, "minute") == 0) return 60.0;
  return 1.0;
}

/*-------------------------------------------------------------------*/
/* tracer_decay.c:26-63  tracer_decay over the cells to process s1;   */
/* the timescale is the constant, or (DE_TR3) the value of the 3D     */
/* tracer whose number the constant holds, in that tracer's units     */
/* (the reference compares an unset string when neither flag is set:  */
/* taken as seconds)                                                  */
/*-------------------------------------------------------------------*/
void oracle_tracer_decay(geometry_t *g, window_t *wd, win_priv_t *wc, double decay,
                         double *tracer, tracer_info_t *tr) {
  int c, cc;
  double d, sc = 1.0;
  int tn;
  (void)g;
  if (decay == 0.0) return;
  tn = (int)decay;
  if (tr->flag & DE_TR2) {
    oracle_err = 5; /* 2D tracers are not carried */
    return;
  }
  if (tr->flag & DE_TR3) sc = units_scale(wc->trinfo_3d[tn].units);
  for (cc = 1; cc <= wc->vc; cc++) {
    c = wc->s1[cc];
    d = decay;
    if (tr->flag & DE_TR3) d = wd->tr_wc[tn][c];
    tracer[c] -= wd->dt * tracer[c] / (d * sc);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:6122-6207  do_ts_relax (temperature tracer tno, salinity */
/* sno; a negative number: the tracer is absent)                      */
/*-------------------------------------------------------------------*/
static void relax_one(geometry_t *g, window_t *wd, win_priv_t *wc, double *tr, double *rf,
                      double *tc, tracer_info_t *ti) {
  int cc, c, c3;
  double rr = ti->relax_rate;
  double sf = 1.0 / 86400.0;
  int tcf = ti->tctype;
  int ngf = (ti->flag & RLX_GRD) ? 1 : 0;
  (void)wc;
  if (rf && ngf) {
    double depth = ti->relax_dum;
    for (cc = 1; cc <= g->b2_t; cc++) {
      c = g->w2_t[cc];
      while (g->cellz[c] > depth && c < g->bot_t[cc]) c = g->zm1[c];
      c3 = c;
      while (c != g->zp1[c]) {
        c = g->zp1[c];
        rf[c] = rf[c3];
      }
    }
  }
  if (rf && rr) {
    for (cc = 1; cc <= g->b3_t; cc++) {
      c = g->w3_t[cc];
      tr[c] -= wd->dt * rr * (tr[c] - rf[c]);
    }
  }
  if (rf && tc && tcf & (RLX_ADPT | RLX_REG | RLX_OBC)) {
    for (cc = 1; cc <= g->b3_t; cc++) {
      c = g->w3_t[cc];
      rr = (tc[c]) ? sf / tc[c] : 0.0;
      tr[c] -= wd->dt * (tr[c] - rf[c]) * rr;
    }
  }
}

void oracle_do_ts_relax(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int cc, c;
  double *temp = (wd->tno >= 0) ? wd->tr_wc[wd->tno] : NULL;
  double *sal = (wd->sno >= 0) ? wd->tr_wc[wd->sno] : NULL;
  if (temp) relax_one(g, wd, wc, temp, wd->rtemp, wd->temp_tc, &wc->trinfo_3d[wd->tno]);
  if (sal) relax_one(g, wd, wc, sal, wd->rsalt, wd->salt_tc, &wc->trinfo_3d[wd->sno]);
  /* a no-gradient above the surface */
  if (!wc->sigma) {
    for (cc = 1; cc <= wc->vcs2; cc++) {
      int cs;
      c = cs = wc->s2[cc];
      while (c != g->zp1[c]) {
        c = g->zp1[c];
        if (temp) temp[c] = temp[cs];
        if (sal) sal[c] = sal[cs];
      }
    }
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:4296-4490  auxiliary_routines: the parts that change the */
/* window's state (reset_dz, calc_age, tracer_decay, do_ts_relax), in */
/* the reference's order                                              */
/*-------------------------------------------------------------------*/
void oracle_auxiliary_routines(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int nn;
  oracle_reset_dz(g, wd, wc);
  if (wd->agetr && wc->agemsk) oracle_calc_age(g, wd, wc);
  for (nn = 0; nn < wc->ntdec; nn++) {
    int n = wc->tdec[nn];
    oracle_tracer_decay(g, wd, wc, wc->dectr[n], wd->tr_wc[n], &wc->trinfo_3d[n]);
  }
  oracle_do_ts_relax(g, wd, wc);
}

/*-------------------------------------------------------------------*/
/* boundary/bdrycond.c:180-214  bc_leastsq: the least-squares line    */
/* through (1, y1) .. (4, y4), evaluated at 5                         */
/*-------------------------------------------------------------------*/
static double obc_leastsq(double y1, double y2, double y3, double y4) {
  const double x[5] = {0.0, 1.0, 2.0, 3.0, 4.0};
  const double y[5] = {0.0, y1, y2, y3, y4};
  double sxy = 0.0, sx = 0.0, sy = 0.0, sx2 = 0.0, a1, a2;
  const int n = 4;
  int i;
  for (i = 1; i <= n; i++) {
    sx += x[i];
    sy += y[i];
    sxy += (x[i] * y[i]);
    sx2 += (x[i] * x[i]);
  }
  a1 = (sxy - (sx * sy / n)) / (sx2 - (sx * sx) / n);
  a2 = sy / n - a1 * sx / n;
  return a2 + a1 * 5.0;
}

/*-------------------------------------------------------------------*/
/* tracers.c:4569-5165  set_OBC_tr, the conditions built: NOGRAD      */
/* (above trpc when set), NOTHIN, CLAMPD, LINEXT, FILEIN (below trpc  */
/* when set), UPSTRM (Eulerian schemes); newval is wincon->w1 as in   */
/* the reference.  Other settings set error 9.                        */
/*-------------------------------------------------------------------*/
static void oracle_set_obc_tr(geometry_t *g, window_t *wd, win_priv_t *wc, open_bdrys_t *o,
                              int tn) {
  const int bcond = o->bcond_tra[tn];
  const int no3 = o->no3_t;
  double *tr = wd->tr_wc[tn], *newval = wc->w1;
  int cc;
  const int bad = CUSTOM | POLEXT | CYCLIC | STATIS | PROFIL | DEPROF | DESCAL | TRFLUX |
                  TRCONC | TRCONF;
  if ((bcond & bad) || o->bgz > 0 || (o->options & OP_OBCCM) ||
      (o->relax_zone_tra && o->relax_zone_tra[tn] > 0) ||
      !(bcond & (NOGRAD | NOTHIN | CLAMPD | LINEXT | FILEIN)) ||
      (o->trpc[tn] != 0.0 && !((bcond & NOGRAD) && (bcond & FILEIN))) ||
      ((bcond & FILEIN) && (!o->t_transfer || o->trm[tn] < 0)) ||
      ((bcond & UPSTRM) && wc->trasc == FFSL)) {
    oracle_err = 9;
    return;
  }
  if (bcond & NOGRAD) {
    for (cc = 1; cc <= no3; cc++) {
      int c = o->obc_t[cc];
      if (!o->trpc[tn] || g->cellz[c] > o->trpc[tn]) newval[c] = tr[o->oi1_t[cc]];
    }
  } else if (bcond & NOTHIN) {
    for (cc = 1; cc <= no3; cc++) newval[o->obc_t[cc]] = tr[o->obc_t[cc]];
  } else if (bcond & CLAMPD) {
    for (cc = 1; cc <= no3; cc++) newval[o->obc_t[cc]] = o->clampv[tn];
  } else if (bcond & LINEXT) {
    const int *imap = o->nmap;
    for (cc = 1; cc <= no3; cc++) {
      int ei = o->obc_t[cc];
      int ei0 = imap[ei], ei1 = imap[ei0], ei2 = imap[ei1], ei3 = imap[ei2];
      newval[ei] = obc_leastsq(tr[ei3], tr[ei2], tr[ei1], tr[ei0]);
    }
  }
  if (bcond & FILEIN) {
    const double *v = o->t_transfer[o->trm[tn]];
    for (cc = 1; cc <= no3; cc++) {
      int c = o->obc_t[cc];
      if (!o->trpc[tn] || g->cellz[c] <= o->trpc[tn]) newval[c] = v[cc];
    }
  }
  if (bcond & UPSTRM) {
    const int xd = (o->ocodex & (L_EDGE | R_EDGE)) != 0;
    const double *vel = xd ? wd->u1 : wd->u2;
    const double *hat = xd ? g->h1au1 : g->h2au2;
    const int *cv = xd ? o->obc_e1 : o->obc_e2, *civ = xd ? o->oi1_e1 : o->oi1_e2;
    const double sc = ((o->ocodex & L_EDGE) || (o->ocodey & B_EDGE)) ? -1.0 : 1.0;
    for (cc = 1; cc <= no3; cc++) {
      int c = o->obc_t[cc], c1 = o->oi1_t[cc], c2 = g->m2d[c];
      double v;
      if (o->upmeth & CENTER)
        v = 0.5 * (vel[cv[cc]] + vel[civ[cc]]);
      else if (o->upmeth & FACE)
        v = vel[cv[cc]];
      else if (o->upmeth & ADAPTIVE)
        v = -sc * 0.5 * (vel[cv[cc]] + vel[civ[cc]]) > 0.0 ? vel[cv[cc]] : vel[civ[cc]];
      else
        v = vel[c1];
      newval[c] = tr[c] - wd->dt / hat[c2] *
                              (0.5 * (sc * v + fabs(v)) * (tr[c] - tr[c1]) +
                               0.5 * (sc * v - fabs(v)) * (newval[c] - tr[c]));
    }
  }
  if (!(bcond & (NOTHIN | TRCONC | TRFLUX | TRCONF)) || (o->options & OP_OWRITE))
    for (cc = 1; cc <= no3; cc++) tr[o->obc_t[cc]] = newval[o->obc_t[cc]];
}

/* tracers.c:4542-4561  bdry_tracer: every boundary, every tracer of
 * tbdy in order */
void oracle_bdry_tracer(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int n, nn;
  for (n = 0; n < g->nobc; n++)
    for (nn = 0; nn < wc->ntbdy; nn++) oracle_set_obc_tr(g, wd, wc, g->open[n], wc->tbdy[nn]);
}

/*-------------------------------------------------------------------*/
/* tracers.c:239-321  (the parts on the path for a closed window):    */
/* lateral ghosts (tracers.c:153-156), advect_diffuse, vert_diffuse,  */
/* auxiliary_routines                                                 */
/*-------------------------------------------------------------------*/
int oracle_tracer_step_3d_mt(geometry_t *g, window_t *wd, win_priv_t *wc,
                             int nthreads) {
  oracle_set_lateral_BC_tr(g, wd);
  if (oracle_advect_diffuse_mt(g, wd, wc, nthreads)) return 1;
  if (oracle_vert_diffuse_3d_mt(g, wd, wc, nthreads)) return 4;
  oracle_err = 0;
  oracle_auxiliary_routines(g, wd, wc);
  oracle_bdry_tracer(g, wd, wc);
  return oracle_err;
}

void oracle_tracer_step_3d(geometry_t *g, window_t *wd, win_priv_t *wc) {
  oracle_tracer_step_3d_mt(g, wd, wc, 1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:6608-6631  domain tracer mass sum(tr*area*dz) over s1     */
/*-------------------------------------------------------------------*/
double oracle_check_tmass(geometry_t *g, window_t *wd, win_priv_t *wc, int n) {
  int cc;
  double d1 = 0.0;
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int cs = g->m2d[c];
    d1 += wd->tr_wc[n][c] * g->cellarea[cs] * wc->dz[c];
  }
  return d1;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/*-------------------------------------------------------------------*/
/* model/hd/momentum/vel3d.c:4041-4191  set_flux_3d, VEL3D mode (no  */
/* mean fields): face volume fluxes from the face velocities.  Under */
/* FFSL the fluxes are the mean volume fluxes u1vm / u2vm and the     */
/* velocities of the faces past b3 are zeroed.                        */
/*-------------------------------------------------------------------*/
#define FFSL_SCHEME 0x080000
void oracle_set_flux_3d(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int c, cc, cs;
  const double *u = wd->u1, *v = wd->u2;
  memset(wd->u1flux3d, 0, g->sgsiz * sizeof(double));
  memset(wd->u2flux3d, 0, g->sgsiz * sizeof(double));
  if (wc->trasc == FFSL_SCHEME) {
    for (cc = g->b3_e1 + 1; cc <= g->n3_e1; cc++) {
      c = g->w3_e1[cc];
      wd->u1[c] = 0.0;
      wd->u1vm[c] = 0.0;
    }
    for (cc = g->b3_e2 + 1; cc <= g->n3_e2; cc++) {
      c = g->w3_e2[cc];
      wd->u2[c] = 0.0;
      wd->u2vm[c] = 0.0;
    }
    for (cc = 1; cc <= g->b3_e1; cc++) {
      c = g->w3_e1[cc];
      wd->u1flux3d[c] = wd->u1vm[c];
    }
    for (cc = 1; cc <= g->b3_e2; cc++) {
      c = g->w3_e2[cc];
      wd->u2flux3d[c] = wd->u2vm[c];
    }
    return;
  }
  for (cc = 1; cc <= g->b3_e1; cc++) {
    c = g->w3_e1[cc];
    cs = g->m2d[c];
    wd->u1flux3d[c] = u[c] * wd->dzu1[c] * g->h2au1[cs] * wc->mdx[cs];
  }
  for (cc = 1; cc <= wc->ncbot_e1; cc++) {
    c = wc->cbot_e1[cc];
    cs = g->m2d[c];
    wd->u1flux3d[c] = wd->u1flux[cs] / wd->dt;
  }
  for (cc = 1; cc <= g->b2_e1; cc++) {
    c = g->w2_e1[cc];
    cs = g->sur_e1[cc];
    while (c < cs) {
      wd->u1flux3d[c] = 0.0;
      c = g->zm1[c];
    }
  }
  for (cc = 1; cc <= g->b3_e2; cc++) {
    c = g->w3_e2[cc];
    cs = g->m2d[c];
    wd->u2flux3d[c] = v[c] * wd->dzu2[c] * g->h1au2[cs] * wc->mdy[cs];
  }
  for (cc = 1; cc <= wc->ncbot_e2; cc++) {
    c = wc->cbot_e2[cc];
    cs = g->m2d[c];
    wd->u2flux3d[c] = wd->u2flux[cs] / wd->dt;
  }
  for (cc = 1; cc <= g->b2_e2; cc++) {
    c = g->w2_e2[cc];
    cs = g->sur_e2[cc];
    while (c < cs) {
      wd->u2flux3d[c] = 0.0;
      c = g->zm1[c];
    }
  }
}

/*-------------------------------------------------------------------*/
/* model/hd/momentum/vel3d.c:4609-4711  vel_w_update, z-levels: w     */
/* from continuity, bottom up over the processed columns of set_dz,   */
/* from the bed velocity wbot (vel_w_bounds' result, given), without  */
/* sources, relaxation or precipitation (waterss = 0).                */
/*-------------------------------------------------------------------*/
void oracle_vel_w_update(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int c, cc, cs;
  const int *bottom = wc->i1, *sur = wc->i3;
  memset(wd->w, 0, g->sgsiz * sizeof(double));
  for (cc = 1; cc <= wc->vcs; cc++) {
    double fcbot, fctop;
    c = bottom[cc];
    cs = g->m2d[c];
    wd->w[c] = wd->wbot[cs];
    fcbot = wd->w[c] * g->cellarea[cs];
    while (c > sur[cc]) {
      int xp1 = g->xp1[c];
      int yp1 = g->yp1[c];
      int zp1 = g->zp1[c];
      fctop = fcbot + wd->u1flux3d[c] - wd->u1flux3d[xp1] + wd->u2flux3d[c] -
              wd->u2flux3d[yp1] + 0.0;
      wd->w[zp1] = fctop / g->cellarea[cs];
      fcbot = fctop;
      c = zp1;
    }
  }
}

/*-------------------------------------------------------------------*/
/* model/hd/momentum/vel3d.c:760-852  set_dz_at_u1 / set_dz_at_u2,    */
/* z-levels: the surface face of each face column below the higher    */
/* elevation either side (not below the face bed), the face            */
/* thicknesses from it down to the bottom face, the surface face's     */
/* thickness copied to the faces above it.  dir 1: e1, 2: e2.          */
/*-------------------------------------------------------------------*/
static double omax(double a, double b) { return (a > b) ? a : b; }

void oracle_set_dz_at_u(geometry_t *g, window_t *wd, int dir) {
  const int nb2 = (dir == 1) ? g->b2_e1 : g->b2_e2;
  const int *w2 = (dir == 1) ? g->w2_e1 : g->w2_e2;
  int *sur = (dir == 1) ? g->sur_e1 : g->sur_e2;
  const int *bot = (dir == 1) ? g->bot_e1 : g->bot_e2;
  const int *nbm = (dir == 1) ? g->xm1 : g->ym1;
  const double *botzu = (dir == 1) ? g->botzu1 : g->botzu2;
  double *dzu = (dir == 1) ? wd->dzu1 : wd->dzu2;
  double *maxeta = (double *)calloc(g->sgsizS, sizeof(double));
  int cc, c, c3, cs, cb, zm1;
  double top, bt;
  for (cc = 1; cc <= nb2; cc++) {
    cs = c = w2[cc];
    top = maxeta[cs] = omax(wd->eta[cs], wd->eta[nbm[cs]]);
    top = maxeta[cs] = omax(top, botzu[cs]);
    zm1 = g->zm1[c];
    while (c != zm1 && g->gridz[c] > top) {
      c = zm1;
      zm1 = g->zm1[c];
    }
    sur[cc] = c;
  }
  memset(dzu, 0, g->sgsiz * sizeof(double));
  for (cc = 1; cc <= nb2; cc++) {
    c = c3 = sur[cc];
    cs = g->m2d[c];
    cb = bot[cc];
    top = maxeta[cs];
    while (c != cb) {
      bt = g->gridz[c];
      dzu[c] = top - bt;
      top = bt;
      c = g->zm1[c];
    }
    dzu[cb] = top - botzu[cs];
    c = c3;
    top = dzu[c];
    while (c != cs) {
      c = g->zp1[c];
      dzu[c] = top;
    }
  }
  free(maxeta);
}

/*-------------------------------------------------------------------*/
/* One offline transport step of one window: transport_step           */
/* (model/hd/tracers/transport.c:94-530) for an exact grid, closed     */
/* boundaries, no fills, a linear surface and CONS_W, then tracer_step */
/* (tracers.c:142-209) and transport_post (the new oldeta).  The       */
/* stored flow of the step (u1, u2, Kz, eta) is in the window.         */
/* Returns tracer_step_3d's code.                                      */
/*-------------------------------------------------------------------*/
int oracle_transport_step_mt(geometry_t *g, window_t *wd, win_priv_t *wc, int nthreads) {
  int cc, r;
  /* vel_w_bounds, linear (vel3d.c:5144-5216, transport.c:346-354):
   * the elevation's lateral ghosts, wtop = detadt, wbot = 0 */
  for (cc = 1; cc <= g->b2_t; cc++) {
    int c = g->w2_t[cc];
    wd->detadt[c] = (wd->eta[c] - wd->etab[c]) / wd->dttr;
  }
  memset(wd->wbot, 0, g->sgsizS * sizeof(double));
  for (cc = 1; cc <= g->nbpt; cc++)
    if (g->bpt[cc] <= g->sgnumS) wd->eta[g->bpt[cc]] = wd->eta[g->bin[cc]];
  memcpy(wd->wtop, wd->detadt, g->sgsizS * sizeof(double));
  memcpy(wd->etab, wd->eta, g->sgsizS * sizeof(double));
  /* transport.c:438-462 */
  oracle_set_dz_at_u(g, wd, 1);
  oracle_set_dz_at_u(g, wd, 2);
  oracle_set_flux_3d(g, wd, wc);
  oracle_set_dz(g, wd, wc);
  oracle_vel_w_update(g, wd, wc);
  r = oracle_tracer_step_3d_mt(g, wd, wc, nthreads);
  /* transport_post (transport.c:543) */
  memcpy(wc->oldeta, wd->eta, g->sgsizS * sizeof(double));
  return r;
}

