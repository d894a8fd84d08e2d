/*
 * ORACLE -- test infrastructure only.  Not part of the product.
 *
 * A CPU restatement, in strict IEEE FP64, of the EMS COMPAS tracer step
 * (model/hd-us/tracers/tracers.c) for one z-level window with closed
 * lateral boundaries, no point sources, regions or multi-dt, Runge-Kutta
 * with one stage (the default), schemes van Leer, QUICKEST and first
 * order, with or without ULTIMATE, and FFSL (ffsl_don with streamlines
 * that stay in their upstream cell, the linear least-squares
 * interpolant, no ULTIMATE or clipping fill).  Each function follows the
 * reference function cited
 * above it with the same loop order and the same order of floating-point
 * operations.  Built with gcc -O3 -fno-fast-math -ffp-contract=off.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load this library, and only as the checker or the timed CPU baseline.
 *
 * How each function is pinned (tests/test_oracle_us.py):
 *   van_leer (edges)    on a quadrilateral mesh the COMPAS step equals the
 *                       SHOC oracle's step on the same grid (the edge
 *                       stencils cm/cp, signs eSc and the update are all
 *                       exercised); constancy and mass on hexagons
 *   set_dz              same column walk as SHOC; maps checked on meshes
 *   surf_concrk +
 *   runge_kutta         closed-basin mass conservation on hexagons; 2 and
 *                       3 stages equal their compositions of forward-Euler
 *                       steps (tests/test_oracle_us_rk.py), as do 5
 *                       stages
 *   tr_bounds, vert_diffuse_3d, implicit_vdiff_tr
 *                       as SHOC (same maths on the same vertical maps);
 *                       the quadrilateral equivalence covers them
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

#include "../include/emstr_us.h"

#define max(a, b) ((a) > (b) ? (a) : (b))
#define min(a, b) ((a) < (b) ? (a) : (b))

static int ous_err = 0;
int oracle_us_last_error(void) { return ous_err; }

/* Runge-Kutta weights, tracers.c:110-129 (rkweights[3*tc + i][rkstage-1]):
 * for stage tc, the weight of the sub-step's starting value, of the
 * stage's value, of its tendency (also the stage's volume weight); the
 * schemes are forward Euler, the SSP schemes of Shu and Osher with 2 and 3
 * stages, none for 4 stages (the reference's column is zero), and the
 * 5-stage scheme of Spiteri and Ruuth, whose last stage combines stages
 * 2-4 (rkw5_last: weights of the stage-2 and stage-3 values, of both
 * tendencies, of the stage-4 value, and the volume weight; the reference
 * weights the stage-4 tendency like the stage-3 one, tracers.c:1199-1205). */
static const double rkw_1[1][3] = {{1.0, 0.0, 1.0}};
static const double rkw_2[2][3] = {{1.0, 0.0, 1.0}, {0.50, 0.50, 0.50}};
static const double rkw_3[3][3] = {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25},
                                   {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}};
static const double rkw_4[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0},
                                   {0.0, 0.0, 0.0}};
static const double rkw_5[4][3] = {{1.00, 0.00, 0.391752226571890},
                                   {0.444370493651235, 0.555629506348765, 0.368410593050371},
                                   {0.620101851488403, 0.379898148511597, 0.251891774271694},
                                   {0.178079954393132, 0.821920045606868, 0.544974750228521}};
static const double rkw5_last[5] = {0.517231671970585, 0.096059710526147, 0.063692468666290,
                                    0.386708617503269, 0.127384937332581};
static const double *rk_w(int rkstage, int tc) {
  switch (rkstage) {
  case 1: return rkw_1[tc];
  case 2: return rkw_2[tc];
  case 3: return rkw_3[tc];
  case 4: return rkw_4[tc];
  default: return rkw_5[tc];
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:7120-7137  no-gradient lateral boundary: ghost = bin      */
/*-------------------------------------------------------------------*/
void oracle_us_set_lateral_BC_tr(geometry_t *g, window_t *wd) {
  int cc, n;
  for (cc = 1; cc <= g->nbpt; cc++) {
    int gb = g->bpt[cc], in = g->bin[cc];
    for (n = 0; n < wd->ntr; n++) wd->tr_wc[n][gb] = wd->tr_wc[n][in];
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:7231-7380  set_dz, z-level branch (thin_merge not taken) */
/* Differs from SHOC in two places: a column is processed when its   */
/* water top is >= (not >) DRY_FRAC*hmin above the bed, and vcs1 is  */
/* the last processed wet column (cc <= v2_t).                        */
/*-------------------------------------------------------------------*/
void oracle_us_set_dz(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int cc, c, cs, zm1, vc, vcs;
  double top, bot;

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

  vc = 1;
  for (cc = 1; cc <= g->b2_t; cc++) {
    c = max(g->sur_t[cc], g->nsur_t[cc]);
    cs = g->m2d[c];
    top = min(wd->eta[cs], wc->oldeta[cs]);
    bot = DRY_FRAC * wc->hmin + g->botz[cs];
    if (top >= bot && c != g->zm1[c]) {
      wc->s1[vc] = c;
      wc->i1[vc] = g->bot_t[cc];
      wc->i2[vc] = g->nsur_t[cc];
      wc->i3[vc] = g->sur_t[cc];
      if (cc <= g->v2_t) wc->vcs1 = vc;
      vc++;
    }
  }
  vcs = vc - 1;
  wc->vcs = vcs;

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
/* tracers.c:2512-2573  van Leer face value (by cells / by values)    */
/*-------------------------------------------------------------------*/
static void vl_tr(double *F, const double *vel, const double *cn, int e, double tp1,
                  double tp2, double tm1, double tm2) {
  double df;
  if (vel[e] > 0.0) {
    df = (tm1 - tm2) * (tp1 - tm1);
    if (df > 0.0)
      df = 2.0 * df / (tp1 - tm2);
    else
      df = 0.0;
    F[e] = tm1 + 0.5 * (1.0 - cn[e]) * df;
  } else if (vel[e] < 0.0) {
    df = (tp1 - tm1) * (tp2 - tp1);
    if (df > 0.0)
      df = 2.0 * df / (tp2 - tm1);
    else
      df = 0.0;
    F[e] = tp1 - 0.5 * (1.0 + cn[e]) * df;
  }
}

/* tracers.c:2368-2421: edges (second cells beyond each side through
 * the edges opposite), then the vertical faces of s1 */
static void vl_faces(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fz) {
  int ee, cc;
  memset(Fx, 0, g->sze * sizeof(double));
  memset(Fz, 0, g->szc * sizeof(double));
  for (ee = 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    int c1 = g->e2c[e][0];
    int c2 = g->e2c[e][1];
    int cm = g->c2c[g->e2e[e][0]][c2];
    int cp = g->c2c[g->e2e[e][1]][c1];
    vl_tr(Fx, wd->u1, wc->w6, e, tr[c1], tr[cp], tr[c2], tr[cm]);
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zp1 = g->zp1[c], zm1 = g->zm1[c], zm2 = g->zm1[zm1];
    vl_tr(Fz, wd->w, wc->w8, c, tr[c], tr[zp1], tr[zm1], tr[zm2]);
  }
}

/* tracers.c:2055-2151  non-uniform QUICKEST on edges (the spacings of
 * the edges beyond through em/ep of the surface edge) and layers */
static void qk_faces(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fz) {
  int ee, cc;
  (void)wd;
  memset(Fx, 0, g->sze * sizeof(double));
  memset(Fz, 0, g->szc * sizeof(double));
  for (ee = 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    int es = g->m2de[e];
    int em = g->em[es];
    int ep = g->ep[es];
    int c1 = g->e2c[e][0];
    int c2 = g->e2c[e][1];
    int cm = g->c2c[g->e2e[e][0]][c2];
    int cp = g->c2c[g->e2e[e][1]][c1];
    double trp = tr[cp], trm = tr[cm];
    double cx = wc->w6[e];
    double db = 2.0 / (g->h2au1[em] + g->h2au1[es]);
    double df = 2.0 / (g->h2au1[ep] + g->h2au1[es]);
    double face = wc->d3[es] * (tr[c2] - tr[c1]) + tr[c1];
    double grad = tr[c1] - tr[c2];
    double curv = 0.0;
    if (cx > 0.0)
      curv = db * (g->h2au1[em] * tr[c1] + g->h2au1[es] * trm) - 2.0 * tr[c2];
    else if (cx < 0.0)
      curv = df * (g->h2au1[es] * trp + g->h2au1[g->m2de[ep]] * tr[c2]) - 2.0 * tr[c1];
    Fx[e] = face - 0.5 * cx * grad - (1.0 - cx * cx) * curv / 6.0;
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zp1 = g->zp1[c], zm1 = g->zm1[c], zm2 = g->zm1[zm1];
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

/* tracers.c:1753-1788 */
static void o1_faces(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                     double *Fx, double *Fz, double dtu) {
  int ee, cc;
  memset(Fx, 0, g->sze * sizeof(double));
  for (ee = 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    int c1 = g->e2c[e][0];
    int c2 = g->e2c[e][1];
    Fx[e] = wc->w6[e] * tr[c2] + wc->w7[e] * tr[c1];
  }
  memset(Fz, 0, g->szc * sizeof(double));
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zm1 = g->zm1[c];
    double *w = wd->w;
    Fz[c] = dtu * (0.5 * (w[c] + fabs(w[c])) * tr[zm1] + 0.5 * (w[c] - fabs(w[c])) * tr[c]);
  }
}

/* tracers.c:3785-3833 */
static void ult_filter(const double *cr, double *Fx, const double *F, int e, int c,
                       double trp, int cm1, double trm) {
  double nc, nu, tc, tu, td, diff;
  if (cr[e] == 0.0) return;
  if (cr[e] > 0.0) {
    tc = F[cm1];
    tu = trm;
    td = F[c];
  } else {
    tc = F[c];
    tu = trp;
    td = F[cm1];
  }
  diff = td - tu;
  nc = nu = 0.0;
  if (diff) {
    nc = (tc - tu) / diff;
    nu = (Fx[e] - tu) / diff;
  }
  if (nc >= 0.0 && nc <= 1.0) {
    if (nu > 1.0)
      Fx[e] = td;
    else if (nu < nc)
      Fx[e] = tu + nc * (td - tu);
    else if (nu > nc / fabs(cr[e]))
      Fx[e] = tu + nc * (td - tu) / fabs(cr[e]);
  } else
    Fx[e] = tu + nc * (td - tu);
}

/*-------------------------------------------------------------------*/
/* tracers.c:1637-1743  advect: face values, ULTIMATE, times fluxes    */
/*-------------------------------------------------------------------*/
static void oracle_us_advect(geometry_t *g, window_t *wd, win_priv_t *wc, const double *tr,
                             double *Fx, double *Fz, double dtu) {
  int ee, cc;
  if (wd->dttr == 0.0) return;
  if (wc->trasc == ORDER1) {
    o1_faces(g, wd, wc, tr, Fx, Fz, dtu);
    return;
  }
  if (wc->trasc == QUICKEST)
    qk_faces(g, wd, wc, tr, Fx, Fz);
  else
    vl_faces(g, wd, wc, tr, Fx, Fz);
  if (wc->ultimate) {
    /* the outer cells through the edge maps ep/em (tracers.c:1697-1700) */
    for (ee = 1; ee <= g->n3_e1; ee++) {
      int e = g->w3_e1[ee];
      int c = g->e2c[e][0];
      int cm1 = g->e2c[e][1];
      int cp = g->e2c[g->ep[e]][0];
      int cm2 = g->e2c[g->em[e]][1];
      ult_filter(wc->w6, Fx, tr, e, c, tr[cp], cm1, tr[cm2]);
    }
    for (cc = 1; cc <= wc->vc; cc++) {
      int c = wc->s1[cc];
      int cp = g->zp1[c], cm1 = g->zm1[c], cm2 = g->zm1[cm1];
      ult_filter(wc->w8, Fz, tr, c, c, tr[cp], cm1, tr[cm2]);
    }
  }
  for (ee = 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    Fx[e] *= wd->u1flux3d[e];
  }
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    Fz[c] *= (wd->w[c] * dtu);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:5383-5449  horizontal diffusion (z-levels: no limiter)    */
/*-------------------------------------------------------------------*/
static void oracle_us_hor_diffuse(geometry_t *g, window_t *wd, win_priv_t *wc,
                                  const double *tr, double *Fx) {
  int ee;
  for (ee = 1; ee <= g->a3_e1; ee++) {
    int e = g->w3_e1[ee];
    int es = g->m2de[e];
    int c1 = g->e2c[e][0];
    int c2 = g->e2c[e][1];
    double csx = wd->dzu1[e] * g->h1au1[es] * wc->mdx[es];
    Fx[e] -= (csx * wc->u2kh[e] * (tr[c1] - tr[c2]) / g->h2au1[es]);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:1170-1214  runge_kutta: the value of stage tc from the     */
/* stage values ntr[] and tendencies grad[] of cell c                   */
/*-------------------------------------------------------------------*/
static double oracle_us_runge_kutta(win_priv_t *wc, int c2, double ovol, double nvol, int tc,
                                    int c, int slf, double *const *ntr,
                                    double *const *grad) {
  double sf = (slf) ? wc->Ds[c2] : wc->Hn1[c2];
  double tr, vol;
  if (wc->rkstage == 5 && tc == 4) {
    const double *w = rkw5_last;
    vol = ovol + w[4] * (nvol - ovol);
    tr = sf * w[0] * ntr[2][c] + sf * w[1] * ntr[3][c] + w[2] * grad[3][c] +
         sf * w[3] * ntr[4][c] + w[2] * grad[4][c];
  } else {
    const double *w = rk_w(wc->rkstage, tc);
    vol = ovol + w[2] * (nvol - ovol);
    tr = sf * w[0] * ntr[0][c] + sf * w[1] * ntr[tc][c] + w[2] * grad[tc][c];
  }
  tr /= (wc->Hn1[c2] * vol);
  return tr;
}

/*-------------------------------------------------------------------*/
/* tracers.c:3085-3260  surf_concrk: the surface-column mass at t,     */
/* its volumes at t and t+1 and the flux sum grad                      */
/*-------------------------------------------------------------------*/
static double oracle_us_surf_concrk(geometry_t *g, win_priv_t *wc, const double *tr,
                                    double osubeta, double subeta, int c, int c2, int cc,
                                    double *gradc, double *ovol, double *nvol,
                                    double fcbot, const double *dtracer) {
  int cs = c;
  int cb = wc->i1[cc];
  int co = wc->i3[cc];
  double waterbot = (c == cb) ? g->botz[c2] : g->gridz[c];
  double dz = subeta - waterbot;
  double surftr, watertop, vol, grad;
  *nvol = dz * g->cellarea[c2];
  surftr = *ovol = 0.0;
  grad = -fcbot;
  while (c != co) {
    int zp1 = g->zp1[c];
    watertop = g->gridz[zp1];
    vol = (watertop - waterbot) * g->cellarea[c2];
    *ovol += vol;
    surftr += (tr[c] * vol);
    grad -= dtracer[c];
    c = zp1;
    waterbot = g->gridz[c];
  }
  watertop = osubeta;
  waterbot = (c == cb) ? g->botz[c2] : g->gridz[c];
  vol = (watertop - waterbot) * g->cellarea[c2];
  *ovol += vol;
  surftr += tr[c] * vol;
  grad -= dtracer[c];
  if (c != c2) {
    do {
      c = g->zp1[c];
      grad -= dtracer[c];
    } while (c != c2);
  }
  if (*nvol == 0.0 || (surftr < 0.0 && dz <= wc->hmin)) surftr = 0.0;
  (void)cs;
  *gradc = grad;
  return surftr;
}

/*-------------------------------------------------------------------*/
/* tracers.c:3270-3370  tr_bounds (as SHOC)                            */
/*-------------------------------------------------------------------*/
void oracle_us_tr_bounds(geometry_t *g, window_t *wd, win_priv_t *wc) {
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
/* advect_diffuse, tracers.c:803-1098 (one Runge-Kutta stage).        */
/*-------------------------------------------------------------------*/
#define OUS_RKMAX 5 /* most Runge-Kutta stages */
typedef struct {
  double *Fx, *Fz, *dtracer, *rk[OUS_RKMAX], *gr[OUS_RKMAX];
  /* FFSL, per tracer (transport.c:1961-2598) */
  double *trm, *moc, *lww, *lwb, *cx, *cy, *ntr, *dint, *cour;
  int *cl, *mask;
  /* clip_ffsl bounds (CLIP): smin / smax by edge, sminz / smaxz by
   * centre, the cell range mn / mx (transport.c:1988-1991, 4118-4127);
   * zero at allocation, land edges are never written */
  double *smin, *smax, *sminz, *smaxz, *cmn, *cmx;
} uscr_t;

static void us_ffsl_don(geometry_t *g, window_t *wd, win_priv_t *wc, double *tr, double *Fx,
                        double *Fz, double dtu, uscr_t *s);
static void ff_clip(geometry_t *g, win_priv_t *wc, double *tr, uscr_t *s);

static void us_tracer_substep(geometry_t *g, window_t *wd, win_priv_t *wc, int n, double dtu,
                              int slf, int ksf, const double *osubeta,
                              const double *subeta, uscr_t *s) {
  int cc, j, tc;
  double *tr = wd->tr_wc[n];
  double *Fx = s->Fx, *Fz = s->Fz, *dtracer = s->dtracer;
  const int kef = wc->vc1, kefs = wc->vcs1;
  tr[0] = (double)n;
  for (tc = 0; tc < wc->rkstage; tc++) memset(s->gr[tc], 0, g->szc * sizeof(double));
  /* Runge-Kutta stage loop, tracers.c:841-1098 */
  for (tc = 0; tc < wc->rkstage; tc++) {
    memset(Fx, 0, g->sze * sizeof(double));
    memset(Fz, 0, g->szc * sizeof(double));
    memset(dtracer, 0, g->szc * sizeof(double));
    for (cc = 1; cc <= g->b2_t; cc++) {
      int cb = g->bot_t[cc];
      tr[g->zm1[cb]] = tr[cb];
    }
    if (wc->advect[n]) {
      if (wc->trasc & FFSL) {
        if (wd->dttr != 0.0) us_ffsl_don(g, wd, wc, tr, Fx, Fz, dtu, s);
      } else
        oracle_us_advect(g, wd, wc, tr, Fx, Fz, dtu);
    }
    if (wc->diffuse[n]) oracle_us_hor_diffuse(g, wd, wc, tr, Fx);
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      int c2 = g->m2d[c];
      dtracer[c] = 0.0;
      for (j = 1; j <= g->npe[c2]; j++) {
        int e = g->c2e[j][c];
        dtracer[c] += (g->eSc[j][c2] * Fx[e] * dtu);
      }
    }
    for (cc = ksf; cc <= kef; cc++) {
      int c = wc->s1[cc];
      int c2 = g->m2d[c];
      int zp1 = g->zp1[c];
      s->rk[tc][c] = tr[c];
      if (slf)
        tr[c] *= wc->Ds[c2];
      else
        tr[c] *= wc->Hn1[c2];
      s->gr[tc][c] = -((dtracer[c] / g->cellarea[c2] + (Fz[zp1] - Fz[c])) / wc->dz[c]);
      tr[c] = oracle_us_runge_kutta(wc, c2, 1.0, 1.0, tc, c, slf, s->rk, s->gr);
    }
    if (!wc->sigma && wc->advect[n]) {
      for (cc = 1; cc <= kefs; cc++) {
        int c = wc->s1[cc];
        int c2 = g->m2d[c];
        double top = -Fz[c] * g->cellarea[c2];
        double ovol, nvol, grad, surftr;
        s->rk[tc][c] = oracle_us_surf_concrk(g, wc, tr, osubeta[c2], subeta[c2], c, c2, cc,
                                             &grad, &ovol, &nvol, top, dtracer);
        s->gr[tc][c] = grad;
        surftr = oracle_us_runge_kutta(wc, c2, ovol, nvol, tc, c, slf, s->rk, s->gr);
        tr[c] = (g->botz[c2] > 0.0 && wc->dz[c] < wc->hmin) ? tr[c] : surftr;
      }
    }
    /* tracers.c:1092-1093 */
    if ((wc->trasc & FFSL) && (wc->fillf & CLIP)) ff_clip(g, wc, tr, s);
  }
}

static uscr_t *uscr = NULL;
static int uscr_nt = 0, uscr_n = 0;
static uscr_t *uscr_get(int nt, int n) {
  int t;
  if (uscr && uscr_nt >= nt && uscr_n == n) return uscr;
  for (t = 0; t < uscr_nt; t++) {
    for (int q = 0; q < OUS_RKMAX; q++) {
      free(uscr[t].rk[q]);
      free(uscr[t].gr[q]);
    }
    void *p[] = {uscr[t].Fx, uscr[t].Fz, uscr[t].dtracer,
                 uscr[t].trm, uscr[t].moc, uscr[t].lww, uscr[t].lwb, uscr[t].cx,
                 uscr[t].cy, uscr[t].ntr, uscr[t].dint, uscr[t].cour, uscr[t].cl,
                 uscr[t].mask, uscr[t].smin, uscr[t].smax, uscr[t].sminz,
                 uscr[t].smaxz, uscr[t].cmn, uscr[t].cmx};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
  }
  free(uscr);
  uscr = (uscr_t *)calloc(nt, sizeof(uscr_t));
  for (t = 0; t < nt; t++) {
    for (int q = 0; q < OUS_RKMAX; q++) {
      uscr[t].rk[q] = (double *)calloc(n, sizeof(double));
      uscr[t].gr[q] = (double *)calloc(n, sizeof(double));
    }
    double **dp[] = {&uscr[t].Fx, &uscr[t].Fz, &uscr[t].dtracer,
                     &uscr[t].trm, &uscr[t].moc, &uscr[t].lwb, &uscr[t].cx, &uscr[t].cy,
                     &uscr[t].ntr, &uscr[t].dint, &uscr[t].cour, &uscr[t].smin,
                     &uscr[t].smax, &uscr[t].sminz, &uscr[t].smaxz, &uscr[t].cmn,
                     &uscr[t].cmx};
    for (size_t i = 0; i < sizeof(dp) / sizeof(dp[0]); i++)
      *dp[i] = (double *)calloc(n, sizeof(double));
    uscr[t].lww = (double *)calloc((size_t)3 * n, sizeof(double));
    uscr[t].cl = (int *)calloc(n + 1, sizeof(int));
    uscr[t].mask = (int *)calloc(n, sizeof(int));
  }
  uscr_nt = nt;
  uscr_n = n;
  return uscr;
}

/*===================================================================*/
/* FFSL on the unstructured mesh (trasc & FFSL), single window, closed */
/* boundaries, no mean-velocity transport, no ULTIMATE, the CLIP fill */
/* (clip_ffsl): the path taken by ffsl_don with its defaults (tranf = */
/* T_VERT|T_HORZ|T_HV, doint = 1, momo2 = 0, vscheme = V_VC, tf = 0.55) */
/* and the linear least-squares interpolant (L_LSLIN).                */
/*===================================================================*/
#define SMALLV 1e-10  /* transport.c:24 */

/* tracer-independent arrays of one step / sub-step */
static double *ff_nw = NULL, *ff_crfzf = NULL, *ff_crfzc = NULL;
static int *ff_clzf = NULL, *ff_clzc = NULL, *ff_e2ee = NULL;
static int ff_szc = 0, ff_sze = 0;

static void ff_alloc(geometry_t *g) {
  if (ff_szc == g->szc && ff_sze == g->sze) return;
  free(ff_nw);
  free(ff_crfzf);
  free(ff_crfzc);
  free(ff_clzf);
  free(ff_clzc);
  free(ff_e2ee);
  ff_nw = (double *)calloc(g->szc, sizeof(double));
  ff_crfzf = (double *)calloc(g->szc, sizeof(double));
  ff_crfzc = (double *)calloc(g->szc, sizeof(double));
  ff_clzf = (int *)calloc(g->szc, sizeof(int));
  ff_clzc = (int *)calloc(g->szc, sizeof(int));
  /* land edges are never given an index (transport.c:2154-2162), so
   * theirs stays 0 */
  ff_e2ee = (int *)calloc(g->sze, sizeof(int));
  ff_szc = g->szc;
  ff_sze = g->sze;
}

/* The streamline velocity fit (tran_grid_init / delaunay_reinit,
 * transport.c:4317-4658, 4728-4929, with lsql_rebuild, lib/interp/
 * lsq_lin/lsql.c:262-287): the weights are the window's (the table the
 * library builds from the mesh, handed over by oracle_us_set_vlsq); the
 * fit of every wet centre is rebuilt from the centred velocities when
 * ffsl_init runs and kept until it runs again.  ff_fit[6][szc]: value,
 * x slope, y slope of u, then of v.  set_dzz (transport.c:6149-6200)
 * runs there too, into ff_dzz / ff_cellz. */
static const emstr_us_vlsq *ff_vlsq = NULL;
static double *ff_fit = NULL, *ff_dzz = NULL, *ff_cellz = NULL;
static const geometry_t *ff_fit_g = NULL; /* window the fit belongs to */
static int ff_fit_szc = 0;

void oracle_us_set_vlsq(const emstr_us_vlsq *t) { ff_vlsq = t; }
/* a window seen for the first time has no fit (a new context's state) */
void oracle_us_forget_fit(void) { ff_fit_g = NULL; }

/* transport.c:6149-6187  set_dzz (no OBCs).  The reference indexes the
 * 2D oldeta with the 3D surface cell (transport.c:6176), which is its
 * 2D index only in the top layer; the column's 2D index is used here.
 * Only the cellz of a centre minus itself and dzz of wet centres reach
 * the streamlines (hd_trans_interp), so the value does not matter. */
static void ff_set_dzz(geometry_t *g, win_priv_t *wc) {
  int cc;
  memcpy(ff_dzz, wc->dz, g->szc * sizeof(double));
  memcpy(ff_cellz, g->cellz, g->szc * sizeof(double));
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc], c2 = g->m2d[c], zm1 = g->zm1[c];
    ff_dzz[zm1] = 0.5 * (wc->dz[c] + wc->dz[zm1]);
    ff_cellz[zm1] = 0.5 * (g->gridz[c] + max(g->botz[c2], g->gridz[zm1]));
  }
  for (cc = 1; cc <= wc->vcs; cc++) {
    int c = wc->s1[cc], c2 = c, zm1 = g->zm1[c], cb = wc->i1[cc];
    ff_dzz[c] = wc->dz[c] + 0.5 * wc->dz[zm1];
    ff_cellz[c] = 0.5 * (wc->oldeta[g->m2d[c]] + g->gridz[c]);
    while (c2 != g->zp1[c2]) {
      c2 = g->zp1[c2];
      ff_dzz[c2] = ff_dzz[c];
      ff_cellz[c2] = ff_cellz[c];
    }
    ff_dzz[g->zm1[cb]] = wc->dz[cb];
    ff_cellz[g->zm1[cb]] = g->botz[c2];
  }
}

/* delaunay_reinit of u (vid 0) and v (vid 1), lsql_rebuild: the slopes
 * sum v_j B[j][1..2] in the stencil's order from 0, the value the
 * centre's own */
static void ff_fit_rebuild(geometry_t *g, win_priv_t *wc, const double *u, const double *v) {
  int cc, o, q;
  if (ff_fit_szc != g->szc) {
    free(ff_fit);
    free(ff_dzz);
    free(ff_cellz);
    ff_fit = (double *)calloc(6 * (size_t)g->szc, sizeof(double));
    ff_dzz = (double *)calloc(g->szc, sizeof(double));
    ff_cellz = (double *)calloc(g->szc, sizeof(double));
    ff_fit_szc = g->szc;
  }
  ff_set_dzz(g, wc);
  ff_fit_g = NULL;
  if (!ff_vlsq || ff_vlsq->szc != g->szc) return;
  for (q = 0; q < 2; q++) {
    const double *val = q ? v : u;
    double *f = ff_fit + 3 * (size_t)q * g->szc;
    for (cc = 1; cc <= g->a3_t; cc++) {
      int c = g->w3_t[cc];
      double w1 = 0.0, w2 = 0.0;
      for (o = ff_vlsq->vo[c]; o < ff_vlsq->vo[c] + ff_vlsq->vn[c]; o++) {
        double z0 = val[ff_vlsq->cell[o]];
        w1 += z0 * ff_vlsq->bx[o];
        w2 += z0 * ff_vlsq->by[o];
      }
      f[c] = val[c];
      f[g->szc + c] = w1;
      f[2 * (size_t)g->szc + c] = w2;
    }
  }
  ff_fit_g = g;
}

/* lsql_interpolate_point (lsql.c:306-330, beta 1) of centre c's fit */
static double ff_fit_at(geometry_t *g, int q, int c, double x, double y) {
  const double *f = ff_fit + 3 * (size_t)q * g->szc;
  int cs = g->m2d[c];
  double xp = x - g->cellx[cs], yp = y - g->celly[cs];
  return f[c] + 1.0 * (f[g->szc + c] * xp + f[2 * (size_t)g->szc + c] * yp);
}

/* transport.c:5171-5296  hd_trans_interp, L_LSLIN, of velocity q (0: u,
 * 1: v) at (x, y, z) in wet centre c: the fits of the layer above
 * (v2) and of c's layer (v1) at (x, y), then linear in z */
static double ff_trans_interp(geometry_t *g, int q, double x, double y, double z, int c) {
  int zp1 = g->zp1[c];
  double v1, v2, v, d;
  v2 = ff_fit_at(g, q, zp1, x, y);
  /* a NaN v2 takes the point's stored value (transport.c:5237-5239),
   * finite, which the weight d = 0 below then drops */
  v1 = ff_fit_at(g, q, c, x, y);
  if (isnan(v2)) v2 = v1;
  d = z - ff_cellz[c];
  v = d * (v2 - v1) / ff_dzz[c] + v1;
  if (isnan(v)) ous_err = 7;
  return v;
}

/* transport.c:5118-5160  set_tr_nograd (no OBCs) */
static void ff_nograd(geometry_t *g, win_priv_t *wc, double *tr) {
  int cc;
  for (cc = 1; cc <= wc->vcs1; cc++) {
    int c, c2, zp1;
    c = c2 = wc->s1[cc];
    zp1 = g->zp1[c];
    while (c != zp1) {
      tr[zp1] = tr[c2];
      c = zp1;
      zp1 = g->zp1[c];
    }
  }
  for (cc = 1; cc <= g->nbpt; cc++) tr[g->bpt[cc]] = tr[g->bin[cc]];
  for (cc = 1; cc <= g->b2_t; cc++) {
    int c = g->bot_t[cc];
    tr[g->zm1[c]] = tr[c];
  }
}

/* transport.c:6098-6141  vel_center_w */
static void ff_vel_center_w(geometry_t *g, window_t *wd, double *nw) {
  int cc;
  const double *w = wd->w;
  memset(nw, 0, g->szc * sizeof(double));
  for (cc = 1; cc <= g->a2_t; cc++) {
    int c = g->sur_t[cc];
    int cb = g->bot_t[cc];
    int cs = g->m2d[c];
    double wtop = wd->wtop[cs];
    double wbot = w[c];
    nw[c] = 0.5 * (wbot + wtop);
    if (fabs(nw[c]) < SMALLV) nw[c] = SMALLV;
    while (c < cb) {
      c = g->zm1[c];
      wtop = wbot;
      wbot = (c == cb) ? wd->wbot[cs] : w[c];
      nw[c] = 0.5 * (wtop + wbot);
      if (fabs(nw[c]) < SMALLV) nw[c] = SMALLV;
    }
  }
}

/* transport.c:2607-2723  ffsl_init: runs on every other call (the
 * F_FFSL toggle; particles would take the other call).  The centred
 * velocities of the lateral ghosts are the edge velocity rotated to
 * east / north, written into windat->u, v as the reference does; then
 * set_dzz and the velocity fits of the streamlines (delaunay_reinit of
 * u and v).  The sediment values of nu / nv / nw and the fits of w and
 * of the sediment layer are never read on this path.                 */
static void ff_init(geometry_t *g, window_t *wd, win_priv_t *wc, double *nw) {
  int cc, j;
  if (wc->flag & F_FFSL) {
    wc->flag &= ~F_FFSL;
    return;
  }
  ff_vel_center_w(g, wd, nw);
  for (cc = 1; cc <= g->nbpt; cc++) {
    int c = g->bpt[cc], ci = g->bin[cc], e = 0;
    for (j = 1; j <= g->npe[g->m2d[ci]]; j++)
      if (g->c2c[j][ci] == c) e = g->c2e[j][ci];
    int es = g->m2de[e];
    wd->u[c] = wd->u1[e] * g->costhu1[es] + wd->u2[e] * g->costhu2[es];
    wd->v[c] = wd->u1[e] * g->sinthu1[es] + wd->u2[e] * g->sinthu2[es];
    nw[c] = nw[ci];
  }
  ff_fit_rebuild(g, wc, wd->u, wd->v);
  wc->flag |= F_FFSL;
}

/* tracers.c:2732-2770  ff_sl_do_vert */
static void ff_do_vert(const double *vel, double dt, const double *h, int ss, int se,
                       const int *sdo, const int *fmap, const int *bmap, int *cl,
                       double *crf) {
  int cc;
  for (cc = ss; cc <= se; cc++) {
    int c = sdo[cc], cp, cpb;
    double ltraj, lrem, dz;
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

/* tracers.c:2670-2721  prep_ff_sl (vertical maps of a sub-step) */
static void ff_prep(geometry_t *g, window_t *wd, win_priv_t *wc, double dt) {
  int cc;
  memset(ff_crfzf, 0, g->szc * sizeof(double));
  memset(ff_crfzc, 0, g->szc * sizeof(double));
  memset(ff_clzf, 0, g->szc * sizeof(int));
  memset(ff_clzc, 0, g->szc * sizeof(int));
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    ff_clzc[c] = (ff_nw[c] < 0.0) ? g->zp1[c] : g->w3_t[cc];
  }
  ff_do_vert(ff_nw, dt, wc->w9, 1, g->a3_t, g->w3_t, g->zp1, g->zm1, ff_clzc, ff_crfzc);
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    ff_clzf[c] = (wd->w[c] > 0.0) ? g->zm1[c] : wc->s1[cc];
  }
  ff_do_vert(wd->w, dt, wc->dz, 1, wc->vc, wc->s1, g->zp1, g->zm1, ff_clzf, ff_crfzf);
}

/* transport.c:1744-1826  intersect */
static int ff_intersect(geometry_t *g, int c, int j, double xn, double yn, double xs,
                        double ys, double *xi, double *yi) {
  int cs = g->m2d[c];
  int es = g->c2e[j][cs];
  double s1, s2;
  if (es) {
    int v1 = g->e2v[es][0];
    int v2 = g->e2v[es][1];
    if (v1 && v2) {
      double v1x = g->gridx[v1], v1y = g->gridy[v1];
      double v2x = g->gridx[v2], v2y = g->gridy[v2];
      double dx = fabs(v1x - v2x);
      if (dx) {
        s1 = (v2y - v1y) / (v2x - v1x);
        if (xs == xn)
          *xi = xn;
        else {
          s2 = (ys - yn) / (xs - xn);
          *xi = (s2 * xn - s1 * v1x + v1y - yn) / (s2 - s1);
        }
        *yi = s1 * (*xi - v1x) + v1y;
      } else {
        s1 = (v2x - v1x) / (v2y - v1y);
        if (ys == yn)
          *yi = yn;
        else {
          s2 = (xs - xn) / (ys - yn);
          *yi = (s1 * v1y - s2 * yn + xn - v1x) / (s1 - s2);
        }
        *xi = s1 * (*yi - v1y) + v1x;
      }
      if (*xi >= min(v1x, v2x) && *xi <= max(v1x, v2x) && *yi >= min(v1y, v2y) &&
          *yi <= max(v1y, v2y)) {
        if (*xi >= min(xn, xs) && *xi <= max(xn, xs) && *yi >= min(yn, ys) &&
            *yi <= max(yn, ys))
          return 1;
      }
    }
  }
  return 0;
}

/* transport.c:1789-1797  get_dist */
static double ff_dist(double x1, double y1, double x2, double y2) {
  double dx = x2 - x1;
  double dy = y2 - y1;
  return sqrt(dx * dx + dy * dy);
}

/* transport.c:1476-1559  get_posc (metres, not geographic); a crossing
 * into a lateral ghost (wgst != 0, the L_GHOST cells) ends the
 * streamline at the edge */
static int ff_posc(geometry_t *g, int *ei, int ci, double u, double v, double *cx,
                   double *cy, double dt, double *odist) {
  int cn, j, e = 0, found = 0;
  const int cis = g->m2d[ci];
  double xin = *cx, yin = *cy, dist, slon, slat, sinth, costh, xi = 0.0, yi = 0.0;
  /* The reference takes the direction's cosine and sine as
   * cos(atan2(v, u)) and sin(atan2(v, u)) from the C library, whose last
   * bit differs between libraries; their exact values u / s and v / s
   * (s = sqrt(u^2 + v^2), correctly rounded here and on the device) make
   * the end point the same on every platform.  At s = 0, atan2(0, 0) = 0:
   * cosine 1, sine 0. */
  const double s = sqrt(u * u + v * v);
  dist = s * dt * 1.0;
  costh = (s > 0.0) ? u / s : 1.0;
  sinth = (s > 0.0) ? v / s : 0.0;
  slon = xin - dist * costh;
  slat = yin - dist * sinth;
  if (dist == 0.0) {
    *odist = 0.0;
    return -ci;
  }
  cn = ci;
  for (j = 1; j <= g->npe[cis]; j++) {
    e = g->c2e[j][ci];
    if (e != *ei && (found = ff_intersect(g, ci, j, *cx, *cy, slon, slat, &xi, &yi))) {
      cn = g->c2c[j][ci];
      break;
    }
  }
  *ei = e;
  if (g->wgst[cn]) {
    *odist = ff_dist(*cx, *cy, xi, yi) / 1.0;
    *cx = xi;
    *cy = yi;
    return -ci;
  }
  if (!found) {
    *odist = ff_dist(*cx, *cy, slon, slat) / 1.0;
    *cx = slon;
    *cy = slat;
    return -ci;
  }
  *odist = ff_dist(*cx, *cy, xi, yi) / 1.0;
  *cx = xi;
  *cy = yi;
  return cn;
}

/* tracers.c:4682-4708  get_linear_value, df = 2 */
static double ff_lin_value(geometry_t *g, const double *lww, const double *lwb, int co,
                           double xi, double yi) {
  int cs = g->m2d[co];
  double x = g->cellx[cs] - xi;
  double y = g->celly[cs] - yi;
  const double *w = lww + (size_t)3 * co;
  return w[0] + lwb[co] * (w[1] * x + w[2] * y);
}

/* tracers.c:4718-4871  get_linear_limit, df = 2: weights w from the
 * stencil, w[0] reset to the centre value, beta the largest factor
 * that keeps the values at the cell's vertices within the range of the
 * centres around each vertex (lw->beta is 1 while the vertex values are
 * taken); ghost centres take their own value with beta = 0 */
static void ff_linear_limit(geometry_t *g, win_priv_t *wc, const double *tr, double *lww,
                            double *lwb) {
  int cc, j, jj;
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    int cs = g->m2d[c];
    qweights *lw = &wc->lw[c];
    double *w = lww + (size_t)3 * c;
    double beta = 1.0;
    w[0] = w[1] = w[2] = 0.0;
    for (j = 0; j < lw->ncells; j++) {
      int cn = lw->cells[j];
      for (jj = 0; jj < 3; jj++) w[jj] += (tr[cn] * lw->B[j][jj]);
    }
    w[0] = tr[c];
    for (j = 1; j <= g->npe[cs]; j++) {
      int v = g->c2v[j][c];
      double vmin = HUGE_VAL, vmax = -HUGE_VAL;
      if (v) {
        int vs = g->m2dv[v];
        double x = g->cellx[cs] - g->gridx[vs];
        double y = g->celly[cs] - g->gridy[vs];
        double val = w[0] + 1.0 * (w[1] * x + w[2] * y);
        double d = val - w[0];
        for (jj = 1; jj <= g->nvc[vs]; jj++) {
          int c1 = g->v2c[v][jj];
          if (!c1) continue;
          vmin = min(vmin, tr[c1]);
          vmax = max(vmax, tr[c1]);
        }
        if (d != 0.0 && val > vmax) beta = max(min((vmax - w[0]) / d, beta), 0.0);
        if (d != 0.0 && val < vmin) beta = max(min((vmin - w[0]) / d, beta), 0.0);
      }
    }
    lwb[c] = min(max(beta, 0.0), 1.0);
  }
  for (cc = 1; cc <= g->nbpt; cc++) {
    int c = g->bpt[cc];
    double *w = lww + (size_t)3 * c;
    w[0] = tr[c];
    w[1] = w[2] = 0.0;
    lwb[c] = 0.0;
  }
}

/* transport.c:8216-8260  ffsl_van_leer (no monotone diagnostics) */
static void ff_van_leer(double *F, const double *tr, const double *vel, const double *crf,
                        const int *cl, int c, const int *fmap, const int *bmap) {
  int cp, bm1, bm2;
  double df;
  if (crf[c] == 0.0) return;
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

/* test access: the streamline of 3D edge e through its upstream cell
 * for velocity (u, v) over dt; returns the signed cell of get_posc and
 * the end point and length */
int oracle_us_ffsl_posc(geometry_t *g, int e, int ci, double u, double v, double dt,
                        double *out) {
  int ei = e;
  int es = g->m2de[e];
  double cx = g->u1x[es], cy = g->u1y[es], dist = 0.0;
  int r = ff_posc(g, &ei, ci, u, v, &cx, &cy, dt, &dist);
  out[0] = cx;
  out[1] = cy;
  out[2] = dist;
  return r;
}

/* test access: the FFSL maps of the last sub-step (one thread): the
 * upstream centre of every wet edge (cl, by edge), clzf of the cells to
 * process and clzc of the wet centres */
void oracle_us_ffsl_maps(geometry_t *g, win_priv_t *wc, int *cl, int *clzf, int *clzc) {
  int ee, cc;
  for (ee = 1; ee <= g->a3_e1; ee++) {
    int v = uscr ? uscr[0].cl[ee] : 0;
    cl[g->w3_e1[ee]] = v < 0 ? -v : v;
  }
  for (cc = 1; cc <= g->a3_t; cc++) clzc[g->w3_t[cc]] = ff_clzc[g->w3_t[cc]];
  for (cc = 1; cc <= wc->vc; cc++) clzf[wc->s1[cc]] = ff_clzf[wc->s1[cc]];
}

/* test access: get_linear_limit of tr into w [3*szc] and beta [szc] */
void oracle_us_ffsl_limit(geometry_t *g, win_priv_t *wc, const double *tr, double *w,
                          double *beta) {
  ff_linear_limit(g, wc, tr, w, beta);
}

/* transport.c:1961-2598  ffsl_don.  A streamline that leaves its cell
 * into a wet cell goes on from the crossing point with the velocity
 * fitted there (hd_trans_interp); without a fit (no ffsl_init on this
 * window yet, or no velocity weights handed over) that is reported
 * (error 3), past 100 segments error 8, a NaN fitted velocity error 7. */
static void us_ffsl_don(geometry_t *g, window_t *wd, win_priv_t *wc, double *tr, double *Fx,
                        double *Fz, double dtu, uscr_t *s) {
  const double tf = 0.55;
  const double *u1 = wd->u1, *u2 = wd->u2, *w = wd->w, *u1v = wd->u1flux3d;
  double *cour = s->cour, *cx = s->cx, *cy = s->cy, *tr_mod = s->trm, *tr_moc = s->moc;
  double *ntr = s->ntr, *dint = s->dint;
  double *smin = s->smin, *smax = s->smax, *sminz = s->sminz, *smaxz = s->smaxz;
  int *cl = s->cl, *mask = s->mask;
  const int clip = (wc->fillf & CLIP) != 0;
  int cc, ee, n;

  /* the largest Courant number of each cell (transport.c:2058-2068) */
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int c2 = g->m2d[c];
    cour[c] = fabs(wc->w8[c]);
    for (n = 1; n <= g->npe[c2]; n++) {
      int e = g->c2e[n][c];
      cour[c] = max(cour[c], fabs(wc->w6[e]));
    }
  }
  memset(cx, 0, g->sze * sizeof(double));
  memset(cy, 0, g->sze * sizeof(double));
  memset(ntr, 0, g->sze * sizeof(double));
  memset(dint, 0, g->sze * sizeof(double));
  memset(tr_moc, 0, g->szc * sizeof(double));
  memset(mask, 0, g->szc * sizeof(int));

  /* z-direction transverse terms (transport.c:2082-2106) */
  memcpy(tr_mod, tr, g->szc * sizeof(double));
  for (cc = 1; cc <= g->a3_t; cc++) {
    int c = g->w3_t[cc];
    int zp1 = (ff_nw[c] < 0.0) ? ff_clzc[c] : g->zm1[ff_clzc[c]];
    int ci = (ff_nw[c] < 0.0) ? g->zm1[zp1] : ff_clzc[c];
    tr_mod[c] = tf * tr[ci] + (1.0 - tf) * tr[c];
    if (ff_crfzc[c] > 0.) tr_mod[c] += tf * (ff_crfzc[c] * (tr[zp1] - tr[ci]));
  }
  ff_nograd(g, wc, tr_mod);

  /* streamline start points and upstream cells (transport.c:2132-2173) */
  for (ee = 1; ee <= g->a3_e1; ee++) {
    int e = g->w3_e1[ee];
    int e2 = g->m2de[e];
    cx[e] = g->u1x[e2];
    cy[e] = g->u1y[e2];
  }
  for (cc = 1; cc <= g->b3_t; cc++) {
    int c = g->w3_t[cc];
    int c2 = g->m2d[c];
    for (n = 1; n <= g->npe[c2]; n++) {
      int e = g->c2e[n][c];
      ee = ff_e2ee[e];
      if (u1[e] > 0.0)
        cl[ee] = (g->eSc[n][c2] < 0) ? g->c2c[n][c] : c;
      else
        cl[ee] = (g->eSc[n][c2] < 0) ? c : g->c2c[n][c];
    }
  }
  ff_linear_limit(g, wc, tr_mod, s->lww, s->lwb);

  /* trace back and integrate (transport.c:2192-2354) */
  for (ee = 1; ee <= g->a3_e1; ee++) {
    int e = g->w3_e1[ee];
    int es = g->m2de[e];
    int c1 = g->e2c[e][0], c2 = g->e2c[e][1];
    int ei = e, first = 0, nseg = 0;
    double time_left = dtu, dist, trs, tre, px, py;
    double u = u1[e] * g->costhu1[es] + u2[e] * g->costhu2[es];
    double v = u1[e] * g->sinthu1[es] + u2[e] * g->sinthu2[es];
    /* the streamline's bounds start at the upstream value (transport.c:2207) */
    if (clip) smin[e] = smax[e] = tr[cl[ee]];
    while (cl[ee] > 0) {
      int ci = cl[ee];
      double pz = ff_cellz ? ff_cellz[ci] : 0.0;
      /* later segments: the centred velocity fitted at the segment's
       * start (transport.c:2221-2225); none before ffsl_init has run on
       * this window (the reference's would come from tran_grid_init) */
      if (first) {
        if (ff_fit_g != g) {
          ous_err = 3;
          break;
        }
        u = ff_trans_interp(g, 0, cx[e], cy[e], pz, ci);
        v = ff_trans_interp(g, 1, cx[e], cy[e], pz, ci);
      }
      /* transport.c:2330-2333: the reference stops the model */
      if (++nseg > 100) {
        ous_err = 8;
        break;
      }
      first = 1;
      px = cx[e];
      py = cy[e];
      cl[ee] = ff_posc(g, &ei, ci, u, v, &cx[e], &cy[e], time_left, &dist);
      if (cl[ee] > 0) {
        double d1 = sqrt(u * u + v * v);
        time_left -= dist / d1;
      }
      trs = ff_lin_value(g, s->lww, s->lwb, ci, px, py);
      tre = ff_lin_value(g, s->lww, s->lwb, ci, cx[e], cy[e]);
      /* widened by the tracer over the least-squares stencil of the
       * segment's centre (transport.c:2309-2311, get_local_bounds mode 2
       * :3824-3852; called twice there, the second call changes nothing) */
      if (clip) {
        const qweights *lw = &wc->lw[abs(cl[ee])];
        int j;
        for (j = 0; j < lw->ncells; j++) {
          int cn = lw->cells[j];
          smin[e] = min(smin[e], tr[cn]);
          smax[e] = max(smax[e], tr[cn]);
        }
      }
      dint[e] += dist;
      ntr[e] += 0.5 * dist * (trs + tre);
      if (cl[ee] < 0) {
        tre = ff_lin_value(g, s->lww, s->lwb, ci, cx[e], cy[e]);
        tr_moc[c1] += tre;
        tr_moc[c2] += tre;
        mask[c1]++;
        mask[c2]++;
      }
    }
    Fx[e] = (dint[e]) ? ntr[e] / dint[e] : Fx[e];
  }
  for (ee = 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    Fx[e] *= u1v[e];
  }

  /* horizontal transverse terms (transport.c:2412-2420) */
  for (cc = 1; cc <= g->b3_t; cc++) {
    int c = g->w3_t[cc];
    if (mask[c]) tr_moc[c] /= (double)mask[c];
    tr_moc[c] = tf * tr_moc[c] + (1.0 - tf) * tr[c];
  }
  ff_nograd(g, wc, tr_moc);

  /* vertical fluxes, V_VC (transport.c:2441-2490) */
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int zm1 = g->zm1[c];
    int ci, it = 0;
    double dz, dist;
    if (cour[c] < 1.0)
      ff_van_leer(Fz, tr, w, ff_crfzf, ff_clzf, c, g->zp1, g->zm1);
    else
      ff_van_leer(Fz, tr_moc, w, ff_crfzf, ff_clzf, c, g->zp1, g->zm1);
    /* transport.c:2461 */
    if (clip) sminz[c] = smaxz[c] = tr[(w[c] > 0.0) ? zm1 : c];
    if (w[c] == 0.0 || zm1 == g->zm1[zm1]) {
      Fz[c] = 0.0;
      continue;
    }
    ci = ff_clzf[c];
    Fz[c] *= (ff_crfzf[c] * wc->dz[ci]);
    ci = (w[c] > 0.0) ? g->zm1[c] : wc->s1[cc];
    dz = wc->dz[ci];
    dist = fabs(w[c] * dtu);
    while (dz < dist) {
      /* a zero thickness would never end the walk: report it */
      if (++it > 4 * g->nz + 4) {
        ous_err = 4;
        break;
      }
      /* transport.c:2483-2486 */
      if (clip) {
        sminz[c] = min(sminz[c], tr[ci]);
        smaxz[c] = max(smaxz[c], tr[ci]);
      }
      Fz[c] += (dz * tr_moc[ci]);
      dist -= dz;
      ci = (w[c] < 0.0) ? g->zp1[ci] : g->zm1[ci];
      if (wc->dz[ci] > 0) dz = wc->dz[ci];
    }
    if (w[c] < 0) Fz[c] *= -1.0;
  }
}

/* transport.c:4107-4247  clip_ffsl, CLIP branch: each centre's range
 * is its vertical streamline's and those of its edges (a land edge's
 * stays zero, as allocated), and the tracer is clamped to it */
static void ff_clip(geometry_t *g, win_priv_t *wc, double *tr, uscr_t *s) {
  int cc, ee;
  double *mn = s->cmn, *mx = s->cmx;
  for (cc = 1; cc <= g->b3_t; cc++) {
    int c = g->w3_t[cc];
    int c2 = g->m2d[c];
    mn[c] = s->sminz[c];
    mx[c] = s->smaxz[c];
    for (ee = 1; ee <= g->npe[c2]; ee++) {
      int e = g->c2e[ee][c];
      mn[c] = min(mn[c], s->smin[e]);
      mx[c] = max(mx[c], s->smax[e]);
    }
  }
  for (cc = 1; cc <= g->v3_t; cc++) {
    int c = g->w3_t[cc];
    tr[c] = max(min(mx[c], tr[c]), mn[c]);
  }
}

/*-------------------------------------------------------------------*/
/* tracers.c:428-1157  advect_diffuse (z-levels, closed boundaries,   */
/* one Runge-Kutta stage).  Returns 0, or the column (2D centre) at   */
/* which the sub-step limit was exceeded (tracers.c:682-689).         */
/*-------------------------------------------------------------------*/
int oracle_us_advect_diffuse_mt(geometry_t *g, window_t *wd, win_priv_t *wc, int nthreads) {
  int cc, nn, ee;
  double *osubeta = wc->d1, *subeta = wc->d2;
  double dtu, dtm, trem;
  double sf = 0.8, minval = 1e-10;
  int itermax = 20, slf = 1, courf, ksf, ii = 0;
  const int vc = wc->vc, vcs = wc->vcs;
  double *u1 = wd->u1, *w = wd->w;
  ous_err = 0;
  if (wc->trasc == NONE) return 0;
  if (wc->rkstage < 1 || wc->rkstage > OUS_RKMAX) { /* the stage arrays hold 1..5 stages */
    ous_err = 5;
    return -ous_err;
  }

  courf = (wc->ultimate || wc->trasc >= QUICKEST) ? 1 : 0;
  if (wc->trasc & ORDER1) {
    wc->ultimate = 0;
    for (ee = 1; ee <= g->n3_e1; ee++) {
      int e = g->w3_e1[ee];
      wc->w6[e] = 0.5 * (wd->u1flux3d[e] + fabs(wd->u1flux3d[e]));
      wc->w7[e] = 0.5 * (wd->u1flux3d[e] - fabs(wd->u1flux3d[e]));
    }
    courf = 0;
  }
  if (wc->trasc & VANLEER) {
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[g->zm1[c]]);
    }
  } else if (wc->trasc & FFSL) {
    /* tracers.c:567-590: spacings, the centred w of Leonard et al.
     * (1996), then ffsl_init */
    itermax = 200;
    sf = wc->trsf;
    ff_alloc(g);
    memset(ff_nw, 0, g->szc * sizeof(double));
    for (cc = 1; cc <= g->a3_t; cc++) {
      int c = g->w3_t[cc];
      int zp1 = g->zp1[c];
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[g->zm1[c]]);
      ff_nw[c] = 0.0;
      if ((w[c] >= 0.0) && (w[zp1] > 0.0))
        ff_nw[c] = w[c];
      else if ((w[c] < 0.0) && (w[zp1] <= 0.0))
        ff_nw[c] = w[zp1];
    }
    /* FILL_METHOD (tracers.c:1092-1093): CLIP is built where its bounds
     * come from the step's values (one Runge-Kutta stage, every column
     * with water processed: a cell of another column is clamped to the
     * bounds its scratch entry kept from an earlier call); MONOTONIC
     * alone clamps to bounds the tracer step never writes (error 6) */
    if ((wc->fillf & MONOTONIC) && !(wc->fillf & CLIP)) {
      ous_err = 6;
      return -ous_err;
    }
    if (wc->fillf & CLIP) {
      int bad = (wc->rkstage != 1);
      char *pc = (char *)calloc(g->szcS, 1);
      for (cc = 1; cc <= wc->vcs1; cc++) pc[g->m2d[wc->s1[cc]]] = 1;
      for (cc = 1; cc <= g->v3_t && !bad; cc++)
        if (!pc[g->m2d[g->w3_t[cc]]]) bad = 1;
      free(pc);
      if (bad) {
        ous_err = 6;
        return -ous_err;
      }
    }
    ff_init(g, wd, wc, ff_nw);
    /* edge -> position in w3_e1 (transport.c:2160), the same every call */
    for (ee = 1; ee <= g->a3_e1; ee++) ff_e2ee[g->w3_e1[ee]] = ee;
  } else {
    /* tracers.c:592-603 (first order also passes here) */
    for (cc = 1; cc <= g->b3_t; cc++) {
      int c = g->w3_t[cc];
      int zm1 = g->zm1[c];
      wc->w5[c] = wc->dz[zm1] / (wc->dz[zm1] + wc->dz[c]);
      wc->w9[c] = 0.5 * (wc->dz[c] + wc->dz[zm1]);
    }
    for (ee = 1; ee <= g->n2_e1; ee++) {
      int e = g->w2_e1[ee];
      wc->d3[e] = 0.5 * g->h1acell[e] / g->h2au1[e];
    }
  }

  /* sub-step, tracers.c:601-690 */
  dtm = dtu = wd->dt;
  if (wc->stab & (SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER)) {
    for (cc = 1; cc <= vc; cc++) {
      int c = wc->s1[cc];
      int cs = g->m2d[c];
      int zp1 = g->zp1[c];
      int j;
      double d1, d3, d4;
      for (j = 1; j <= g->npe[cs]; j++) {
        int e = g->c2e[j][c];
        int e2 = g->m2de[e];
        if (wc->trasc & FFSL) {
          /* the one-sided Lipschitz number (tracers.c:624-631) */
          int c1 = g->e2c[e][0], c2 = g->e2c[e][1];
          d4 = (wd->u[c1] * g->costhu1[e2] + wd->v[c1] * g->sinthu1[e2]) -
               (wd->u[c2] * g->costhu1[e2] + wd->v[c2] * g->sinthu1[e2]);
          d4 = (d4 < 0.0) ? -d4 : 0.0;
        } else
          d4 = u1[e];
        if (fabs(d4) < minval) d4 = minval;
        d1 = sf * fabs(g->h2au1[e2] / d4);
        if (d1 < dtm) {
          dtm = d1;
          ii = cs;
        }
      }
      if (wc->trasc & FFSL) {
        /* tracers.c:652-656; the bed cell (W_BOT) has none */
        d4 = (cc <= vcs) ? 0.0 : w[zp1] - w[c];
        if (g->zm1[g->zm1[c]] == g->zm1[c]) d4 = 0.0;
        d4 = (d4 < 0.0) ? -d4 : 0.0;
      } else
        d4 = (cc <= vcs) ? 0.0 : 0.5 * (w[c] + w[zp1]);
      if (fabs(d4) < minval) d4 = minval;
      /* the reference scales by Ds of the last edge's centre; Ds = 1
       * at z-levels, so the column's own Ds is used */
      d3 = (cc <= vcs) ? dtu : sf * fabs(wc->dz[c] * wc->Ds[cs] / d4);
      if (d3 > 0.0 && d3 < dtm) {
        dtm = d3;
        ii = c;
      }
    }
    if (dtm != dtu) dtm = floor(dtm);
  }
  if (dtm == 0.0 || (int)(dtu / dtm) > itermax) {
    ous_err = 1;
    return g->m2d[ii] ? g->m2d[ii] : 1;
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
        for (ee = 1; ee <= g->n3_e1; ee++) {
          int e = g->w3_e1[ee];
          int e2 = g->m2de[e];
          wc->w6[e] = fmod(u1[e] * dtu / g->h2au1[e2], 1.0);
        }
        for (cc = 1; cc <= g->n3_t; cc++) {
          int c = g->w3_t[cc];
          int c2 = g->m2d[c];
          wc->w8[c] = fmod(w[c] * dtu / (wc->w9[c] * wc->Ds[c2]), 1.0);
        }
      }
      for (cc = 1; cc <= vcs; cc++) {
        int c = wc->s1[cc];
        int c2 = g->m2d[c];
        subeta[c2] = wc->oldeta[c2] + (1.0 + (dtu - trem) / wd->dt) *
                                          (wd->eta[c2] - wc->oldeta[c2]);
      }
      if (wc->trasc & FFSL) ff_prep(g, wd, wc, dtu);
      {
        int nt = nthreads > 0 ? nthreads : 1;
        int nsz = g->szc > g->sze ? g->szc : g->sze;
        if (nt > wc->ntbdy) nt = wc->ntbdy > 0 ? wc->ntbdy : 1;
        uscr_t *S = uscr_get(nt, nsz);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
#endif
        for (nn = 0; nn < wc->ntbdy; nn++) {
          int tid = 0;
#ifdef _OPENMP
          tid = omp_get_thread_num();
#endif
          us_tracer_substep(g, wd, wc, wc->tbdy[nn], dtu, slf, ksf, osubeta, subeta,
                            &S[tid]);
        }
      }
      oracle_us_tr_bounds(g, wd, wc);
      for (cc = 1; cc <= vcs; cc++) {
        int c = wc->s1[cc];
        int c2 = g->m2d[c];
        osubeta[c2] = subeta[c2];
      }
    }
    trem -= dtu;
    slf = 0;
  }
  /* FFSL: a multi-segment streamline (3) or an endless vertical walk (4) */
  return ous_err ? -ous_err : 0;
}

int oracle_us_advect_diffuse(geometry_t *g, window_t *wd, win_priv_t *wc) {
  return oracle_us_advect_diffuse_mt(g, wd, wc, 1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:5753-5896  implicit vertical diffusion, one tracer       */
/*-------------------------------------------------------------------*/
static int oracle_us_implicit_vdiff_tr(geometry_t *g, window_t *wd, win_priv_t *wc,
                                       double *var, const double *dzcell, const double *fb,
                                       const double *ft, const int *ctp, const int *cbt,
                                       int vcs, const double *scale, const double *C,
                                       const double *Cp1, const double *Cm1, double *work) {
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
/* tracers.c:5562-5744  vertical diffusion (no surface fluxes)         */
/*-------------------------------------------------------------------*/
int oracle_us_vert_diffuse_3d_mt(geometry_t *g, window_t *wd, win_priv_t *wc, int nthreads) {
  int cc, nn, bad = 0;
  int *ctp = wc->i2, *cbt = wc->i5, *cth = wc->i4;
  double dt = wd->dt;
  double *dzface = wc->w7, *dzcell = wc->w9, *Kz = wc->w10;
  double *Cm1 = wc->w1, *C = wc->w2, *Cp1 = wc->w3;
  if (wc->trasc == NONE) return 0;
  memset(wc->d1, 0, g->szcS * sizeof(double));
  memset(wc->d2, 0, g->szcS * sizeof(double));
  memset(cth, 0, g->szcS * sizeof(int));
  memcpy(dzcell, wc->dz, g->szc * sizeof(double));
  memcpy(Kz, wd->Kz, g->szc * sizeof(double));
  ous_err = 0;

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
      int r = oracle_us_implicit_vdiff_tr(g, wd, wc, wd->tr_wc[n], dzcell, wc->d1, wc->d2,
                                          ctp, cbt, wc->vcs, wc->one, C, Cp1, Cm1,
                                          work + (size_t)tid * 7 * nzw);
      if (r > bad) bad = r;
    }
    free(work);
  }
  if (bad) ous_err = 4;
  return bad;
}

void oracle_us_vert_diffuse_3d(geometry_t *g, window_t *wd, win_priv_t *wc) {
  oracle_us_vert_diffuse_3d_mt(g, wd, wc, 1);
}

/*-------------------------------------------------------------------*/
/* tracers.c:185-200 (lateral ghosts) and tracers.c:261-330            */
/*-------------------------------------------------------------------*/
int oracle_us_tracer_step_3d_mt(geometry_t *g, window_t *wd, win_priv_t *wc, int nthreads) {
  oracle_us_set_lateral_BC_tr(g, wd);
  if (oracle_us_advect_diffuse_mt(g, wd, wc, nthreads)) return 1;
  if (oracle_us_vert_diffuse_3d_mt(g, wd, wc, nthreads)) return 4;
  return 0;
}

void oracle_us_tracer_step_3d(geometry_t *g, window_t *wd, win_priv_t *wc) {
  oracle_us_tracer_step_3d_mt(g, wd, wc, 1);
}

/* domain tracer mass sum(tr*area*dz) over s1 (tracers.c:6608-6631 of
 * SHOC, the same sum in COMPAS) */
double oracle_us_tmass(geometry_t *g, window_t *wd, win_priv_t *wc, int n) {
  int cc;
  double d1 = 0.0;
  for (cc = 1; cc <= wc->vc; cc++) {
    int c = wc->s1[cc];
    int cs = g->m2d[c];
    d1 += wd->tr_wc[n][c] * g->cellarea[cs] * wc->dz[c];
  }
  return d1;
}
