/*
 * Host layer of libemstr_us: the COMPAS tracer step (model/hd-us) on the
 * GPU behind the reference's entry points.
 *
 * A context mirrors one window on the device in the layout of
 * emstr_usdev.h.  It is built from the window's own maps: wet mesh cells
 * become columns in surface order, the surface-layer edges become the
 * dense edges of every layer, and the edge stencils the reference reaches
 * through c2c/e2e and ep/em are reduced to per-edge columns and selectors,
 * checked against the 3D maps of every layer when the context is made.
 *
 * Scope (rejected with an error otherwise): z-levels, closed boundaries,
 * one Runge-Kutta stage, schemes van Leer and QUICKEST (with or without
 * ULTIMATE), first order and FFSL (without ULTIMATE or a clipping fill,
 * streamlines that stay in their upstream cell), no mean-velocity
 * transport, no thin-layer merging, and no flux or velocity through land
 * edges.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "emstr_us.h"
#include "emstr_us_kernels.h"

typedef long long ll;

/* error record (emstr_err.c) */
int emstr_set_error(int code, int cell, const char *fmt, ...);
void emstr_clear_error(void);
const char *emstr_last_error(void);
int emstr_last_error_cell(void);

#define E_UNSUP 2
#define E_DEVICE 3
#define E_ARG 5
#define E_SUBSTEP 1
#define E_PIVOT 4

#define HCHK(x)                                                             \
  do {                                                                      \
    hipError_t _e = (x);                                                    \
    if (_e != hipSuccess)                                                   \
      return emstr_set_error(E_DEVICE, -1, "%s: %s", #x, hipGetErrorString(_e)); \
  } while (0)

typedef struct emstr_us_ctx {
  geometry_t *g;
  window_t *wd;
  win_priv_t *wc;
  hipStream_t st;
  emstr_usdev D;
  int ncol, ne2, nesp;
  int *cs2col;   /* [szcS] surface centre -> column (-1)   */
  int *col2cs;   /* [ncol]                                  */
  int *h_kb, *h_ks, *h_kold, *h_knew, *h_cols;
  int *sedg;     /* [ncol] sediment ghost centre            */
  int *sp_of;    /* [nz*ncol] dense cell -> sparse centre   */
  /* device */
  double *tr[6]; /* the fourth to sixth only with Runge-Kutta stages */
  double *d_g3;  /* 5 stages: the stage-3 tendency [ntr][n3] */
  int cur;
  double *stage; /* [max(szc, sze)] */
  int *d_esp;    /* wet edges: sparse index and dense index */
  ll *d_edn;
  int *d_lesp;   /* FFSL: land edges, sparse and dense index (their u1, u2
                    give the ghost velocities of the streamline fit)   */
  ll *d_ledn;
  int nlesp;
  int *d_bpt, *d_bin, *d_gsed, *d_ised;
  int *d_proc, *d_err;
  unsigned long long *d_dtm;
  double *d_fx;  /* edge fluxes [ntr][nz][ne2], made on first use */
  double *d_fz;  /* vertical fluxes [ntr][nz][ncol], made on first use */
  int *h_esp;    /* wet edges: sparse index (host copy of d_esp)         */
  ll *h_edn;     /* wet edges: dense index (host copy of d_edn)          */
  /* FFSL work arrays, made on first use (see emstr_us_kernels.h)      */
  int ffgeom;        /* surface geometry, vertices and stencils uploaded */
  unsigned fgen, vfit_gen; /* field uploads so far; the upload the
                              streamline velocity fit was built from     */
  int ghost_edges;   /* ghost velocities are the rotated edge velocity   */
  double *d_nw, *d_crc, *d_crf, *d_cour, *d_ld, *d_tm, *d_lq, *d_te, *d_moc;
  int *d_clc, *d_clf, *d_lci, *d_fferr;
  int *d_need, ffstamp; /* tr_moc centres read, by sub-step stamp */
  int anydif, have_cols;
  int key_szc, key_sze;
  /* the FP32 form (EMSTR_US_BATCH_F32): FP32 per-tracer intermediates
   * and arithmetic of the FFSL fluxes over the FP64 tracer state; the
   * maps, streamlines, geometry, update and column solve are the plain
   * form's */
  int f32;
  float *f_fx, *f_fz, *f_tm, *f_lq, *f_te, *f_moc;
  /* CLIP (clip_ffsl): each wet centre's tracer range over its
   * least-squares stencil [ntr][nz][ncol], made on first use */
  double *d_bmn, *d_bmx;
} usctx;

static void *dalloc(size_t bytes, int *ok) {
  void *p = NULL;
  if (bytes == 0) bytes = 8;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    *ok = 0;
    return NULL;
  }
  hipMemset(p, 0, bytes);
  return p;
}

static int up(usctx *c, void *dst, const void *src, size_t bytes) {
  HCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->st));
  return 0;
}

static void us_destroy(usctx *c) {
  int n;
  if (!c) return;
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  if (c->st) hipStreamSynchronize(c->st);
  void *dv[] = {d->kb, d->ks, d->kold, d->knew, d->cellarea, d->botz, d->eta, d->oldeta,
                d->osubeta, d->subeta, d->layers, d->w, d->Kz, d->dz, d->mintr, d->maxtr,
                d->advect, d->diffuse, d->in_tbdy, d->in_tdif, d->cols, d->wsp, d->wdn,
                D->npe, D->ce, D->esc, D->ec1, D->ec2, D->ecp, D->ecm, D->esel, D->h1e,
                D->h2e, D->qd3, D->qhm, D->qhp, D->u1, D->flux, D->dzu, D->kh, c->tr[0], c->tr[1], c->tr[2], c->tr[3], c->tr[4], c->tr[5], c->d_g3,
                c->stage, c->d_esp, c->d_edn, c->d_bpt, c->d_bin, c->d_gsed, c->d_ised,
                c->d_proc, c->d_err, c->d_dtm, c->d_fx, c->d_fz, D->cx, D->cy, D->ux,
                D->uy, D->n1x, D->n1y, D->t2x, D->t2y, D->ev, D->vxy, D->eord, D->lsq_off, D->lsq_s, D->lr_off, D->lr_q, D->vsl, D->lsq_B, D->u2, D->wtop,
                D->wbot, D->uc, D->vc, D->gu0, D->gv0, D->gue, D->gve, c->d_nw, c->d_crc,
                c->d_crf, c->d_cour, c->d_ld, c->d_tm, c->d_lq, c->d_te, c->d_moc,
                c->d_clc, c->d_clf, c->d_lci, c->d_fferr, c->d_need, c->f_fx, c->f_fz,
                c->f_tm, c->f_lq, c->f_te, c->f_moc, D->lsq_gc, c->d_bmn, c->d_bmx,
                D->vls_off, D->vls_ent, D->vls_bx, D->vls_by, D->vfit, D->lnx, D->sg_ci,
                D->sg_nx, D->sg_d, D->sg_cnt, c->d_lesp, c->d_ledn};
  for (n = 0; n < (int)(sizeof(dv) / sizeof(dv[0])); n++)
    if (dv[n]) hipFree(dv[n]);
  void *hv[] = {c->cs2col, c->col2cs, c->h_kb, c->h_ks, c->h_kold, c->h_knew, c->h_cols,
                c->sedg, c->sp_of, c->h_esp, c->h_edn};
  for (n = 0; n < (int)(sizeof(hv) / sizeof(hv[0])); n++) free(hv[n]);
  if (c->st) hipStreamDestroy(c->st);
  free(c);
}

/* the window settings this library covers */
static int us_check_scope(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int n;
  if (g->nobc) return emstr_set_error(E_UNSUP, -1, "open boundaries are not supported");
  if (wc->sigma) return emstr_set_error(E_UNSUP, -1, "sigma coordinates are not supported");
  if (wc->thin_merge) return emstr_set_error(E_UNSUP, -1, "thin-layer merging is not supported");
  if (wc->means) return emstr_set_error(E_UNSUP, -1, "mean-velocity transport is not supported");
  if (wc->rkstage < 1 || wc->rkstage > 5 || wc->rkstage == 4)
    return emstr_set_error(E_UNSUP, -1,
                           "Runge-Kutta with %d stages is not supported (1, 2, 3 or 5 stages; "
                           "the reference's weights for 4 stages are zero)",
                           wc->rkstage);
  if (wc->trasc != VANLEER && wc->trasc != ORDER1 && wc->trasc != QUICKEST &&
      wc->trasc != FFSL)
    return emstr_set_error(E_UNSUP, -1, "advection scheme 0x%x is not supported", wc->trasc);
  if (wc->trasc == FFSL) {
    if (wc->ultimate)
      return emstr_set_error(E_UNSUP, -1, "FFSL with ULTIMATE is not supported");
    /* FILL_METHOD (tracers.c:1092-1093, clip_ffsl transport.c:4107-4247).
     * MONOTONIC alone clamps to bounds that ffsl_don writes only under
     * CLIP or ULTIMATE (transport.c:2207-2210, 2461-2462): the reference
     * clips every tracer to zero and its mass rescaling then divides by
     * zero.  CLIP (with or without MONOTONIC, whose fill its branch
     * returns before) is built where its bounds are the step's: one
     * Runge-Kutta stage (a second stage would read cells above the
     * surface clamped to bounds kept from an earlier call) and every
     * tracer of the step advected (ffsl_don writes the bounds of the
     * advected tracers only). */
    if ((wc->fillf & MONOTONIC) && !(wc->fillf & CLIP))
      return emstr_set_error(E_UNSUP, -1,
                             "FFSL with FILL_METHOD MONOTONIC without CLIP is not supported (the "
                             "reference clamps to bounds it never sets)");
    if (wc->fillf & CLIP) {
      if (wc->rkstage != 1)
        return emstr_set_error(E_UNSUP, -1, "FFSL with CLIP: one Runge-Kutta stage only");
      for (n = 0; n < wc->ntbdy; n++)
        if (!wc->advect[wc->tbdy[n]])
          return emstr_set_error(E_UNSUP, wc->tbdy[n],
                                 "FFSL with CLIP: every tracer of the step must be advected");
    }
    if (wc->compatible & V6898)
      return emstr_set_error(E_UNSUP, -1, "FFSL with V6898 compatibility is not supported");
    if (!wc->lw || !g->c2v || !g->v2c || !g->nvc || !g->m2dv || !g->e2v || !g->gridx ||
        !g->gridy || !g->u1x || !g->u1y || !g->costhu2 || !g->sinthu2 || !wd->u2 ||
        !wd->wtop || !wd->wbot || !wd->u || !wd->v)
      return emstr_set_error(E_ARG, -1, "FFSL needs vertex maps, edge geometry, u2, wtop, "
                                        "wbot, u, v and the least-squares weights");
  }
  for (n = 1; n <= g->v2_t; n++) {
    int cs = g->w2_t[n];
    if (wc->Ds[cs] != 1.0 || wc->Hn1[cs] != 1.0 || (wc->one && wc->one[cs] != 1.0))
      return emstr_set_error(E_UNSUP, cs, "z-level windows need Ds = Hn1 = 1");
  }
  for (n = 1; n < g->szeS; n++)
    if (wc->mdx[n] != 1.0) return emstr_set_error(E_UNSUP, n, "mdx must be 1");
  (void)wd;
  return 0;
}

/* Is centre x the ghost of centre w (a lateral ghost whose bin is w)? */
static int ghost_of(const geometry_t *g, const int *gbin, int x, int w) {
  return gbin[x] == w;
}

static usctx *us_create(geometry_t *g, window_t *wd, win_priv_t *wc) {
  int ok = 1, n, cc, j, e;
  emstr_clear_error();
  if (us_check_scope(g, wd, wc)) return NULL;
  usctx *c = (usctx *)calloc(1, sizeof(usctx));
  c->g = g;
  c->wd = wd;
  c->wc = wc;
  c->key_szc = g->szc;
  c->key_sze = g->sze;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    free(c);
    emstr_set_error(E_DEVICE, -1, "stream create");
    return NULL;
  }
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  const int ncol = g->v2_t, nz = g->nz, npem = g->npem;
  const int ne2 = g->szeS - 1;
  c->ncol = ncol;
  c->ne2 = ne2;
  d->nx = ncol;
  d->ny = 1;
  d->nz = nz;
  d->ntr = wd->ntr;
  d->plane = ncol;
  d->n3 = (ll)ncol * nz;
  d->trasc = wc->trasc;
  d->ultimate = wc->ultimate;
  d->us = 1;
  d->dt = wd->dt;
  d->hmin = wc->hmin;
  d->sgsiz = g->szc;
  D->npem = npem;
  D->ne2 = ne2;

  /* columns */
  c->cs2col = (int *)malloc((size_t)g->szcS * sizeof(int));
  for (n = 0; n < g->szcS; n++) c->cs2col[n] = -1;
  c->col2cs = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  c->h_kb = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  c->sedg = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  for (cc = 1; cc <= ncol; cc++) {
    int cs = g->w2_t[cc], cb = g->bot_t[cc];
    c->cs2col[cs] = cc - 1;
    c->col2cs[cc - 1] = cs;
    c->h_kb[cc - 1] = g->s2k[cb];
    c->sedg[cc - 1] = g->zm1[cb];
  }
  /* wet centres */
  d->nwet = g->v3_t;
  int *wsp = (int *)malloc((size_t)(d->nwet + 1) * sizeof(int));
  ll *wdn = (ll *)malloc((size_t)(d->nwet + 1) * sizeof(ll));
  c->sp_of = (int *)calloc((size_t)nz * ncol + 1, sizeof(int));
  for (n = 0; n < d->nwet; n++) {
    int cw = g->w3_t[n + 1];
    int col = c->cs2col[g->m2d[cw]];
    int k = g->s2k[cw];
    if (col < 0 || k < c->h_kb[col] || k >= nz) {
      emstr_set_error(E_UNSUP, cw, "wet centre outside its column");
      ok = 0;
      break;
    }
    wsp[n] = cw;
    wdn[n] = (ll)k * ncol + col;
    c->sp_of[wdn[n]] = cw;
  }
  /* ghost -> its wet centre */
  int *gbin = (int *)calloc(g->szc, sizeof(int));
  for (n = 1; n <= g->nbpt; n++) gbin[g->bpt[n]] = g->bin[n];

  /* per column edge lists */
  int *h_npe = (int *)calloc(ncol + 1, sizeof(int));
  int *h_ce = (int *)calloc((size_t)npem * ncol + 1, sizeof(int));
  int *h_esc = (int *)calloc((size_t)npem * ncol + 1, sizeof(int));
  for (int col = 0; col < ncol && ok; col++) {
    int cs = c->col2cs[col];
    h_npe[col] = g->npe[cs];
    if (g->npe[cs] > npem || g->npe[cs] > 8) {
      emstr_set_error(E_UNSUP, cs, "npe exceeds npem");
      ok = 0;
    }
    for (j = 1; j <= g->npe[cs] && ok; j++) {
      e = g->c2e[j][cs];
      if (e < 1 || e >= g->szeS) {
        emstr_set_error(E_UNSUP, cs, "surface edge %d of centre %d outside the surface", j, cs);
        ok = 0;
        break;
      }
      h_ce[(j - 1) * ncol + col] = e - 1;
      h_esc[(j - 1) * ncol + col] = g->eSc[j][cs];
    }
  }
  /* per surface edge stencils */
  int *h_ec1 = (int *)malloc((size_t)(ne2 + 1) * sizeof(int));
  int *h_ec2 = (int *)malloc((size_t)(ne2 + 1) * sizeof(int));
  int *h_ecp = (int *)malloc((size_t)(ne2 + 1) * sizeof(int));
  int *h_ecm = (int *)malloc((size_t)(ne2 + 1) * sizeof(int));
  unsigned char *h_esel = (unsigned char *)calloc(ne2 + 1, 1);
  double *h_h1 = (double *)malloc((size_t)(ne2 + 1) * sizeof(double));
  double *h_h2 = (double *)malloc((size_t)(ne2 + 1) * sizeof(double));
  double *h_qd3 = (double *)malloc((size_t)(ne2 + 1) * sizeof(double));
  double *h_qhm = (double *)malloc((size_t)(ne2 + 1) * sizeof(double));
  double *h_qhp = (double *)malloc((size_t)(ne2 + 1) * sizeof(double));
  for (e = 1; e <= ne2 && ok; e++) {
    int c1 = g->e2c[e][0], c2 = g->e2c[e][1];
    int q = e - 1;
    h_ec1[q] = c->cs2col[c1];
    h_ec2[q] = (c2 < g->szcS) ? c->cs2col[c2] : -1;
    h_h1[q] = g->h1au1[e];
    h_h2[q] = g->h2au1[e];
    h_ecp[q] = h_ecm[q] = -1;
    /* QUICKEST metrics (tracers.c:600-603, 2082-2085): em/ep of the
     * surface edge are surface edges */
    if (g->em[e] < 1 || g->em[e] > ne2 || g->ep[e] < 1 || g->m2de[g->ep[e]] > ne2 ||
        g->m2de[g->ep[e]] < 1) {
      emstr_set_error(E_UNSUP, e, "surface edge %d: em/ep outside the surface", e);
      ok = 0;
      break;
    }
    h_qd3[q] = 0.5 * g->h1acell[e] / g->h2au1[e];
    h_qhm[q] = g->h2au1[g->em[e]];
    h_qhp[q] = g->h2au1[g->m2de[g->ep[e]]];
    if (h_ec1[q] < 0) {
      emstr_set_error(E_UNSUP, c1, "surface edge %d without a wet e2c[e][0]", e);
      ok = 0;
      break;
    }
    if (h_ec2[q] < 0) continue; /* land edge */
    int cp = g->c2c[g->e2e[e][1]][c1], cm = g->c2c[g->e2e[e][0]][c2];
    h_ecp[q] = (cp < g->szcS) ? c->cs2col[cp] : -1;
    h_ecm[q] = (cm < g->szcS) ? c->cs2col[cm] : -1;
    if (h_ecp[q] >= 0 && g->e2c[g->ep[e]][0] == cp) h_esel[q] |= 1;
    if (h_ecm[q] >= 0 && g->e2c[g->em[e]][1] == cm) h_esel[q] |= 2;
  }
  /* every wet 3D edge against the model of the kernel; wet edges are
   * scattered to [layer][surface edge] */
  int *esp = (int *)malloc((size_t)(g->n3_e1 + 1) * sizeof(int));
  ll *edn = (ll *)malloc((size_t)(g->n3_e1 + 1) * sizeof(ll));
  int nesp = 0;
  for (int ee = 1; ee <= g->n3_e1 && ok; ee++) {
    e = g->w3_e1[ee];
    int c1 = g->e2c[e][0], c2 = g->e2c[e][1];
    int k = g->s2k[c1];
    int es = g->m2de[e], q = es - 1;
    int wet2 = (ee <= g->v3_e1);
    if (es < 1 || es > ne2 || k < 0 || k >= nz) {
      emstr_set_error(E_UNSUP, e, "edge %d: no surface edge or layer", e);
      ok = 0;
      break;
    }
    if (!wet2) continue;
    int col1 = h_ec1[q], col2 = h_ec2[q];
    int bad = (col2 < 0 || g->m2d[c1] != c->col2cs[col1] || g->m2d[c2] != c->col2cs[col2] ||
               g->s2k[c2] != k);
    if (!bad) {
      int cp = g->c2c[g->e2e[e][1]][c1], cm = g->c2c[g->e2e[e][0]][c2];
      int pc = h_ecp[q], mc = h_ecm[q];
      int wp = (pc >= 0 && k >= c->h_kb[pc]), wm = (mc >= 0 && k >= c->h_kb[mc]);
      int cpw = wp ? c->sp_of[(ll)k * ncol + pc] : 0, cmw = wm ? c->sp_of[(ll)k * ncol + mc] : 0;
      if (wp ? (cp != cpw) : !ghost_of(g, gbin, cp, c1)) bad = 1;
      if (wm ? (cm != cmw) : !ghost_of(g, gbin, cm, c2)) bad = 1;
      int cpu = g->e2c[g->ep[e]][0], cmu = g->e2c[g->em[e]][1];
      if (wp ? (cpu != ((h_esel[q] & 1) ? cpw : c1)) : (cpu != c1)) bad = 1;
      if (wm ? (cmu != ((h_esel[q] & 2) ? cmw : c2)) : !ghost_of(g, gbin, cmu, c2)) bad = 1;
    }
    if (bad) {
      emstr_set_error(E_UNSUP, e, "edge %d: mesh maps differ from the device model", e);
      ok = 0;
      break;
    }
    esp[nesp] = e;
    edn[nesp] = (ll)k * ne2 + q;
    nesp++;
  }
  /* the edges of every wet centre at its layer are the surface edges */
  for (n = 1; n <= g->v3_t && ok; n++) {
    int cw = g->w3_t[n], cs = g->m2d[cw], col = c->cs2col[cs];
    for (j = 1; j <= g->npe[cs]; j++)
      if (g->m2de[g->c2e[j][cw]] - 1 != h_ce[(j - 1) * ncol + col]) {
        emstr_set_error(E_UNSUP, cw, "centre %d: edge %d is not its surface edge", cw, j);
        ok = 0;
        break;
      }
  }
  c->nesp = nesp;
  free(gbin);

  /* device arrays */
  if (ok) {
    const size_t P = ncol, N3 = (size_t)ncol * nz, NE3 = (size_t)ne2 * nz;
    const int nt = wd->ntr > 0 ? wd->ntr : 1;
    int **ci[] = {&d->kb, &d->ks, &d->kold, &d->knew};
    for (n = 0; n < 4; n++) *ci[n] = (int *)dalloc(P * sizeof(int), &ok);
    double **c2[] = {&d->cellarea, &d->botz, &d->eta, &d->oldeta, &d->osubeta, &d->subeta};
    for (n = 0; n < 6; n++) *c2[n] = (double *)dalloc(P * sizeof(double), &ok);
    d->layers = (double *)dalloc((nz + 1) * sizeof(double), &ok);
    double **c3[] = {&d->w, &d->Kz, &d->dz};
    for (n = 0; n < 3; n++) *c3[n] = (double *)dalloc(N3 * sizeof(double), &ok);
    d->mintr = (double *)dalloc(nt * sizeof(double), &ok);
    d->maxtr = (double *)dalloc(nt * sizeof(double), &ok);
    int **ti[] = {&d->advect, &d->diffuse, &d->in_tbdy, &d->in_tdif};
    for (n = 0; n < 4; n++) *ti[n] = (int *)dalloc(nt * sizeof(int), &ok);
    d->cols = (int *)dalloc(P * sizeof(int), &ok);
    d->wsp = (int *)dalloc((size_t)d->nwet * sizeof(int), &ok);
    d->wdn = (ll *)dalloc((size_t)d->nwet * sizeof(ll), &ok);
    D->npe = (int *)dalloc(P * sizeof(int), &ok);
    D->ce = (int *)dalloc((size_t)npem * P * sizeof(int), &ok);
    D->esc = (int *)dalloc((size_t)npem * P * sizeof(int), &ok);
    int **ei[] = {&D->ec1, &D->ec2, &D->ecp, &D->ecm};
    for (n = 0; n < 4; n++) *ei[n] = (int *)dalloc((size_t)ne2 * sizeof(int), &ok);
    D->esel = (unsigned char *)dalloc((size_t)ne2, &ok);
    D->h1e = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    D->h2e = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    D->qd3 = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    D->qhm = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    D->qhp = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    double **e3[] = {&D->u1, &D->flux, &D->dzu, &D->kh};
    for (n = 0; n < 4; n++) *e3[n] = (double *)dalloc(NE3 * sizeof(double), &ok);
    for (n = 0; n < 3; n++) c->tr[n] = (double *)dalloc((size_t)nt * N3 * sizeof(double), &ok);
    size_t sgmax = (size_t)(g->szc > g->sze ? g->szc : g->sze);
    c->stage = (double *)dalloc(sgmax * sizeof(double), &ok);
    c->d_esp = (int *)dalloc((size_t)nesp * sizeof(int), &ok);
    c->d_edn = (ll *)dalloc((size_t)nesp * sizeof(ll), &ok);
    c->d_bpt = (int *)dalloc((size_t)(g->nbpt + 1) * sizeof(int), &ok);
    c->d_bin = (int *)dalloc((size_t)(g->nbpt + 1) * sizeof(int), &ok);
    c->d_gsed = (int *)dalloc((size_t)(g->ngsed + 1) * sizeof(int), &ok);
    c->d_ised = (int *)dalloc((size_t)(g->ngsed + 1) * sizeof(int), &ok);
    c->d_proc = (int *)dalloc(P * sizeof(int), &ok);
    c->d_err = (int *)dalloc(sizeof(int), &ok);
    c->d_dtm = (unsigned long long *)dalloc(sizeof(unsigned long long), &ok);
    if (!ok) emstr_set_error(E_DEVICE, -1, "device allocation failed");
  }
  if (ok) {
    /* static data */
    double *h2 = (double *)malloc((size_t)(ncol + 1) * sizeof(double));
    double *hb = (double *)malloc((size_t)(ncol + 1) * sizeof(double));
    for (int col = 0; col < ncol; col++) {
      h2[col] = g->cellarea[c->col2cs[col]];
      hb[col] = g->botz[c->col2cs[col]];
    }
    ok = !(up(c, d->kb, c->h_kb, ncol * sizeof(int)) ||
           up(c, d->cellarea, h2, ncol * sizeof(double)) ||
           up(c, d->botz, hb, ncol * sizeof(double)) ||
           up(c, d->layers, g->layers, (nz + 1) * sizeof(double)) ||
           up(c, d->wsp, wsp, (size_t)d->nwet * sizeof(int)) ||
           up(c, d->wdn, wdn, (size_t)d->nwet * sizeof(ll)) ||
           up(c, D->npe, h_npe, ncol * sizeof(int)) ||
           up(c, D->ce, h_ce, (size_t)npem * ncol * sizeof(int)) ||
           up(c, D->esc, h_esc, (size_t)npem * ncol * sizeof(int)) ||
           up(c, D->ec1, h_ec1, ne2 * sizeof(int)) || up(c, D->ec2, h_ec2, ne2 * sizeof(int)) ||
           up(c, D->ecp, h_ecp, ne2 * sizeof(int)) || up(c, D->ecm, h_ecm, ne2 * sizeof(int)) ||
           up(c, D->esel, h_esel, ne2) || up(c, D->h1e, h_h1, ne2 * sizeof(double)) ||
           up(c, D->h2e, h_h2, ne2 * sizeof(double)) ||
           up(c, D->qd3, h_qd3, ne2 * sizeof(double)) ||
           up(c, D->qhm, h_qhm, ne2 * sizeof(double)) ||
           up(c, D->qhp, h_qhp, ne2 * sizeof(double)) ||
           up(c, c->d_esp, esp, (size_t)nesp * sizeof(int)) ||
           up(c, c->d_edn, edn, (size_t)nesp * sizeof(ll)) ||
           up(c, c->d_bpt, g->bpt + 1, (size_t)g->nbpt * sizeof(int)) ||
           up(c, c->d_bin, g->bin + 1, (size_t)g->nbpt * sizeof(int)) ||
           up(c, c->d_gsed, g->gsed_t + 1, (size_t)g->ngsed * sizeof(int)) ||
           up(c, c->d_ised, g->ised_t + 1, (size_t)g->ngsed * sizeof(int)));
    if (hipStreamSynchronize(c->st) != hipSuccess) ok = 0;
    free(h2);
    free(hb);
  }
  c->h_ks = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  c->h_kold = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  c->h_knew = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  c->h_cols = (int *)malloc((size_t)(ncol + 1) * sizeof(int));
  for (n = 0; n < ncol; n++) c->h_knew[n] = c->h_kold[n] = nz - 1, c->h_ks[n] = -1;
  free(wsp);
  free(wdn);
  free(h_npe);
  free(h_ce);
  free(h_esc);
  free(h_ec1);
  free(h_ec2);
  free(h_ecp);
  free(h_ecm);
  free(h_esel);
  free(h_h1);
  free(h_h2);
  free(h_qd3);
  free(h_qhm);
  free(h_qhp);
  c->h_esp = esp;
  c->h_edn = edn;
  if (!ok) {
    us_destroy(c);
    return NULL;
  }
  return c;
}

/*-------------------------------------------------------------------*/
/* registry: one device mirror per window                            */
/*-------------------------------------------------------------------*/
#define MAXCTX 1024 /* windows with a device mirror at once */
static pthread_mutex_t reg_lock = PTHREAD_MUTEX_INITIALIZER;
static geometry_t *reg_g[MAXCTX];
static usctx *reg_c[MAXCTX];

static usctx *ctx_for(geometry_t *g, window_t *wd, win_priv_t *wc) {
  usctx *c = NULL, *stale = NULL;
  int i;
  pthread_mutex_lock(&reg_lock);
  for (i = 0; i < MAXCTX; i++)
    if (reg_g[i] == g) {
      c = reg_c[i];
      if (c->key_szc != g->szc || c->key_sze != g->sze) {
        stale = c;
        c = NULL;
        reg_g[i] = NULL;
        reg_c[i] = NULL;
      }
      break;
    }
  pthread_mutex_unlock(&reg_lock);
  if (stale) us_destroy(stale);
  if (!c) {
    c = us_create(g, wd, wc);
    if (!c) return NULL;
    pthread_mutex_lock(&reg_lock);
    for (i = 0; i < MAXCTX; i++)
      if (!reg_g[i]) {
        reg_g[i] = g;
        reg_c[i] = c;
        break;
      }
    pthread_mutex_unlock(&reg_lock);
    if (i == MAXCTX) {
      us_destroy(c);
      emstr_set_error(E_ARG, -1, "more than %d windows with a device mirror", MAXCTX);
      return NULL;
    }
  }
  c->wd = wd;
  c->wc = wc;
  return c;
}

void emstr_us_forget_window(geometry_t *g) {
  usctx *c = NULL;
  int i;
  pthread_mutex_lock(&reg_lock);
  for (i = 0; i < MAXCTX; i++)
    if (reg_g[i] == g) {
      c = reg_c[i];
      reg_g[i] = NULL;
      reg_c[i] = NULL;
    }
  pthread_mutex_unlock(&reg_lock);
  us_destroy(c);
}

const char *emstr_us_last_error(void) { return emstr_last_error(); }
int emstr_us_last_error_cell(void) { return emstr_last_error_cell(); }

/*-------------------------------------------------------------------*/
/* uploads                                                           */
/*-------------------------------------------------------------------*/
static int up_cells(usctx *c, double *dst, const double *sp) {
  HCHK(hipMemcpyAsync(c->stage, sp, (size_t)c->g->szc * sizeof(double), hipMemcpyHostToDevice,
                      c->st));
  if (emstr_launch_scatter(c->D.d.wsp, c->D.d.wdn, c->D.d.nwet, c->stage, dst, c->st))
    return emstr_set_error(E_DEVICE, -1, "scatter");
  return 0;
}

static int up_edges(usctx *c, double *dst, const double *sp) {
  HCHK(hipMemcpyAsync(c->stage, sp, (size_t)c->g->sze * sizeof(double), hipMemcpyHostToDevice,
                      c->st));
  if (emstr_launch_scatter(c->d_esp, c->d_edn, c->nesp, c->stage, dst, c->st))
    return emstr_set_error(E_DEVICE, -1, "edge scatter");
  return 0;
}

/* the land edges' values of an edge array (FFSL: the ghost velocities) */
static int up_land_edges(usctx *c, double *dst, const double *sp) {
  if (!c->nlesp) return 0;
  HCHK(hipMemcpyAsync(c->stage, sp, (size_t)c->g->sze * sizeof(double), hipMemcpyHostToDevice,
                      c->st));
  if (emstr_launch_scatter(c->d_lesp, c->d_ledn, c->nlesp, c->stage, dst, c->st))
    return emstr_set_error(E_DEVICE, -1, "land edge scatter");
  return 0;
}

static int up_cols(usctx *c, double *dst, const double *sp2) {
  double *h = (double *)malloc((size_t)c->ncol * sizeof(double));
  for (int col = 0; col < c->ncol; col++) h[col] = sp2[c->col2cs[col]];
  hipError_t e = hipMemcpyAsync(dst, h, (size_t)c->ncol * sizeof(double),
                                hipMemcpyHostToDevice, c->st);
  hipStreamSynchronize(c->st);
  free(h);
  return e == hipSuccess ? 0 : emstr_set_error(E_DEVICE, -1, "column upload");
}

/* per-tracer settings (tracers.c:803-805 tbdy, tdif_v) */
static int up_meta(usctx *c) {
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->D.d;
  const int nt = wd->ntr > 0 ? wd->ntr : 1;
  int *tb = (int *)calloc(nt, sizeof(int)), *td = (int *)calloc(nt, sizeof(int));
  int n, any = 0;
  for (n = 0; n < wc->ntbdy; n++)
    if (wc->tbdy[n] >= 0 && wc->tbdy[n] < wd->ntr) tb[wc->tbdy[n]] = 1;
  for (n = 0; n < wc->ntdif_v; n++)
    if (wc->tdif_v[n] >= 0 && wc->tdif_v[n] < wd->ntr) td[wc->tdif_v[n]] = 1;
  for (n = 0; n < wd->ntr; n++)
    if (wc->diffuse[n]) any = 1;
  c->anydif = any;
  int r = up(c, d->mintr, wc->mintr, nt * sizeof(double)) ||
          up(c, d->maxtr, wc->maxtr, nt * sizeof(double)) ||
          up(c, d->advect, wc->advect, nt * sizeof(int)) ||
          up(c, d->diffuse, wc->diffuse, nt * sizeof(int)) ||
          up(c, d->in_tbdy, tb, nt * sizeof(int)) || up(c, d->in_tdif, td, nt * sizeof(int));
  hipStreamSynchronize(c->st);
  free(tb);
  free(td);
  return r ? E_DEVICE : 0;
}

/* per-column step data from the host's set_dz lists (s1, i2, i3) */
static int cols_from_host(usctx *c) {
  geometry_t *g = c->g;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->D.d;
  int cc;
  for (int q = 0; q < c->ncol; q++) c->h_ks[q] = -1;
  /* old / new surface of every column (vel_center_w walks all of them) */
  for (cc = 1; cc <= c->ncol; cc++) {
    int col = cc - 1, ns = g->nsur_t[cc], os = g->sur_t[cc];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    c->h_kold[col] = (os == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[os];
  }
  for (cc = 1; cc <= wc->vcs; cc++) {
    int s = wc->s1[cc];
    int col = c->cs2col[g->m2d[s]];
    if (col < 0) return emstr_set_error(E_ARG, s, "s1 holds a centre outside the columns");
    c->h_ks[col] = g->s2k[s];
    int ns = wc->i2[cc], os = wc->i3[cc];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    c->h_kold[col] = (os == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[os];
    c->h_cols[cc - 1] = col;
  }
  d->ncol = wc->vcs;
  if (up(c, d->ks, c->h_ks, c->ncol * sizeof(int)) ||
      up(c, d->kold, c->h_kold, c->ncol * sizeof(int)) ||
      up(c, d->knew, c->h_knew, c->ncol * sizeof(int)) ||
      up(c, d->cols, c->h_cols, (size_t)(d->ncol ? d->ncol : 1) * sizeof(int)))
    return E_DEVICE;
  HCHK(hipStreamSynchronize(c->st));
  c->have_cols = 1;
  return 0;
}

/* no flux or velocity through land edges (closed boundaries) */
static int check_land_edges(usctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  for (int ee = g->v3_e1 + 1; ee <= g->n3_e1; ee++) {
    int e = g->w3_e1[ee];
    if (wd->u1[e] != 0.0 || wd->u1flux3d[e] != 0.0)
      return emstr_set_error(E_UNSUP, e, "flow through land edge %d", e);
  }
  return 0;
}

static int us_ffsl_fields(usctx *c);

static int us_upload_fields(usctx *c) {
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  int r;
  if ((r = us_check_scope(c->g, wd, wc))) return r;
  if ((r = check_land_edges(c))) return r;
  d->dt = wd->dt;
  d->hmin = wc->hmin;
  d->trasc = wc->trasc;
  d->ultimate = (wc->trasc == ORDER1) ? 0 : wc->ultimate;
  if ((r = up_meta(c))) return r;
  if (up_edges(c, D->u1, wd->u1) || up_edges(c, D->flux, wd->u1flux3d) ||
      up_cells(c, d->w, wd->w) || up_cells(c, d->Kz, wd->Kz) || up_cells(c, d->dz, wc->dz))
    return E_DEVICE;
  if (c->anydif)
    if (up_edges(c, D->dzu, wd->dzu1) || up_edges(c, D->kh, wc->u2kh)) return E_DEVICE;
  if (up_cols(c, d->eta, wd->eta) || up_cols(c, d->oldeta, wc->oldeta)) return E_DEVICE;
  if ((r = cols_from_host(c))) return r;
  D->ffsl = (wc->trasc == FFSL);
  if (D->ffsl && (r = us_ffsl_fields(c))) return r;
  c->fgen++;
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int us_upload_tracers(usctx *c) {
  emstr_dev *d = &c->D.d;
  for (int n = 0; n < c->wd->ntr; n++) {
    HCHK(hipMemcpyAsync(c->stage, c->wd->tr_wc[n], (size_t)c->g->szc * sizeof(double),
                        hipMemcpyHostToDevice, c->st));
    if (emstr_launch_scatter(d->wsp, d->wdn, d->nwet, c->stage, c->tr[c->cur] + (ll)n * d->n3,
                             c->st))
      return emstr_set_error(E_DEVICE, -1, "scatter");
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int us_download_tracers(usctx *c) {
  emstr_dev *d = &c->D.d;
  geometry_t *g = c->g;
  for (int n = 0; n < c->wd->ntr; n++) {
    HCHK(hipMemcpyAsync(c->stage, c->wd->tr_wc[n], (size_t)g->szc * sizeof(double),
                        hipMemcpyHostToDevice, c->st));
    if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, c->tr[c->cur] + (ll)n * d->n3, c->stage,
                            c->st))
      return emstr_set_error(E_DEVICE, -1, "gather");
    /* ghosts hold the no-gradient values the next step would set */
    emstr_launch_ghosts(c->d_bpt, c->d_bin, g->nbpt, c->stage, c->st);
    emstr_launch_ghosts(c->d_gsed, c->d_ised, g->ngsed, c->stage, c->st);
    HCHK(hipMemcpyAsync(c->wd->tr_wc[n], c->stage, (size_t)g->szc * sizeof(double),
                        hipMemcpyDeviceToHost, c->st));
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

/*-------------------------------------------------------------------*/
/* FFSL: surface geometry, vertices, least-squares stencils           */
/*-------------------------------------------------------------------*/
/* slot of wet centre q in the row list rw[0..*nr-1], appended if new;
 * -1 when the list is full */
static int ff_slot(int *rw, int *nr, int q) {
  for (int i = 0; i < *nr; i++)
    if (rw[i] == q) return i;
  if (*nr >= EMSTR_FF_ROWS) return -1;
  rw[*nr] = q;
  return (*nr)++;
}

static int icmp(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* The streamline velocity fit (emstr_us_vlsq_build) as device tables by
 * wet centre, with the fit and the segment records of multi-segment
 * streamlines; a ghost entry becomes its edge at the centre's layer. */
static int us_ffsl_vlsq(usctx *c) {
  geometry_t *g = c->g;
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  const int ncol = c->ncol, ne2 = c->ne2;
  const size_t N3 = (size_t)d->n3, NE3 = (size_t)ne2 * d->nz;
  emstr_us_vlsq t;
  int ok = 1, q;
  if (emstr_us_vlsq_build(g, &t))
    return emstr_set_error(E_UNSUP, -1, "FFSL: streamline velocity fit (no vertex maps)");
  int *off = (int *)malloc(((size_t)d->nwet + 1) * sizeof(int));
  ll *ent = (ll *)malloc(((size_t)t.nent + 1) * sizeof(ll));
  double *bx = (double *)malloc(((size_t)t.nent + 1) * sizeof(double));
  double *by = (double *)malloc(((size_t)t.nent + 1) * sizeof(double));
  int m = 0;
  for (q = 0; q < d->nwet && off && ent && bx && by; q++) {
    const int cw = g->w3_t[q + 1], k = g->s2k[cw];
    off[q] = m;
    for (int o = t.vo[cw]; o < t.vo[cw] + t.vn[cw]; o++, m++) {
      const int x = t.cell[o];
      if (g->wgst[x])
        ent[m] = ~((ll)k * ne2 + (g->m2de[t.edge[o]] - 1));
      else
        ent[m] = (ll)g->s2k[x] * ncol + c->cs2col[g->m2d[x]];
      bx[m] = t.bx[o];
      by[m] = t.by[o];
    }
  }
  if (off) off[d->nwet] = m;
  emstr_us_vlsq_free(&t);
  if (!off || !ent || !bx || !by) ok = 0;
  D->vls_off = (int *)dalloc(((size_t)d->nwet + 1) * sizeof(int), &ok);
  D->vls_ent = (ll *)dalloc(((size_t)m + 1) * sizeof(ll), &ok);
  D->vls_bx = (double *)dalloc(((size_t)m + 1) * sizeof(double), &ok);
  D->vls_by = (double *)dalloc(((size_t)m + 1) * sizeof(double), &ok);
  D->vfit = (double *)dalloc(6 * N3 * sizeof(double), &ok);
  D->lnx = (int *)dalloc(NE3 * sizeof(int), &ok);
  D->sg_cap = (int)(NE3 / 8 > 65536 ? NE3 / 8 : 65536);
  D->sg_ci = (int *)dalloc((size_t)D->sg_cap * sizeof(int), &ok);
  D->sg_nx = (int *)dalloc((size_t)D->sg_cap * sizeof(int), &ok);
  D->sg_d = (double *)dalloc((size_t)D->sg_cap * 5 * sizeof(double), &ok);
  D->sg_cnt = (int *)dalloc(sizeof(int), &ok);
  D->vfit_ok = 0;
  /* land edges: the ghost velocity of the fit is the land edge's u1, u2
   * rotated, so those go to the dense edge arrays too */
  int nl = 0;
  int *lesp = (int *)malloc(((size_t)(g->n3_e1 - g->v3_e1) + 1) * sizeof(int));
  ll *ledn = (ll *)malloc(((size_t)(g->n3_e1 - g->v3_e1) + 1) * sizeof(ll));
  for (int ee = g->v3_e1 + 1; ee <= g->n3_e1 && lesp && ledn; ee++) {
    const int e = g->w3_e1[ee];
    lesp[nl] = e;
    ledn[nl] = (ll)g->s2k[g->e2c[e][0]] * ne2 + (g->m2de[e] - 1);
    nl++;
  }
  if (!lesp || !ledn) ok = 0;
  c->nlesp = nl;
  c->d_lesp = (int *)dalloc(((size_t)nl + 1) * sizeof(int), &ok);
  c->d_ledn = (ll *)dalloc(((size_t)nl + 1) * sizeof(ll), &ok);
  if (ok && (up(c, c->d_lesp, lesp, ((size_t)nl + 1) * sizeof(int)) ||
             up(c, c->d_ledn, ledn, ((size_t)nl + 1) * sizeof(ll))))
    ok = 0;
  /* no records until a streamline has several segments */
  if (ok && hipMemset(D->lnx, 0xff, NE3 * sizeof(int)) != hipSuccess) ok = 0;
  if (ok && (up(c, D->vls_off, off, ((size_t)d->nwet + 1) * sizeof(int)) ||
             up(c, D->vls_ent, ent, ((size_t)m + 1) * sizeof(ll)) ||
             up(c, D->vls_bx, bx, ((size_t)m + 1) * sizeof(double)) ||
             up(c, D->vls_by, by, ((size_t)m + 1) * sizeof(double)) ||
             hipStreamSynchronize(c->st) != hipSuccess))
    ok = 0;
  free(off);
  free(ent);
  free(bx);
  free(by);
  free(lesp);
  free(ledn);
  return ok ? 0 : emstr_set_error(E_DEVICE, -1, "FFSL streamline velocity fit tables");
}

/* Builds and checks the device model of the FFSL maps: every wet centre's
 * vertices are those of its surface centre, the wet centres around each
 * of them in its layer are the columns around the surface vertex that
 * are wet there, its wet edges in ascending order of w3_e1 follow one
 * per-column order, and every least-squares stencil reads wet centres
 * or ghosts (which read their wet centre). */
static int us_ffsl_geom(usctx *c) {
  geometry_t *g = c->g;
  win_priv_t *wc = c->wc;
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  const int ncol = c->ncol, ne2 = c->ne2, nz = g->nz, npem = g->npem;
  const int nvl = g->szvS - 1, nvcm = g->nvcm;
  int ok = 1, n, j, col, e;
  if (nvl < 1 || nvcm < 1) return emstr_set_error(E_ARG, -1, "FFSL: no vertices");
  if (nvcm > 4) return emstr_set_error(E_UNSUP, -1, "FFSL: more than 4 centres at a vertex");
  if (npem > EMSTR_FF_NPE)
    return emstr_set_error(E_UNSUP, -1, "FFSL: more than %d edges a centre", EMSTR_FF_NPE);
  double *hcx = (double *)malloc((size_t)ncol * sizeof(double));
  double *hcy = (double *)malloc((size_t)ncol * sizeof(double));
  double *he[6];
  for (n = 0; n < 6; n++) he[n] = (double *)malloc((size_t)ne2 * sizeof(double));
  double *hev = (double *)malloc((size_t)ne2 * 4 * sizeof(double));
  int *hcvt = (int *)calloc((size_t)npem * ncol, sizeof(int));
  int *heord = (int *)calloc((size_t)npem * ncol, sizeof(int));
  double *hvx = (double *)malloc((size_t)nvl * sizeof(double));
  double *hvy = (double *)malloc((size_t)nvl * sizeof(double));
  int *hvcol = (int *)malloc((size_t)nvl * nvcm * sizeof(int));
  int *epos = (int *)calloc(g->sze, sizeof(int));
  int *gbin = (int *)calloc(g->szc, sizeof(int));
  for (n = 1; n <= g->nbpt; n++) gbin[g->bpt[n]] = g->bin[n];
  for (n = 1; n <= g->a3_e1; n++) epos[g->w3_e1[n]] = n;
  for (col = 0; col < ncol; col++) {
    hcx[col] = g->cellx[c->col2cs[col]];
    hcy[col] = g->celly[c->col2cs[col]];
  }
  for (e = 1; e <= ne2 && ok; e++) {
    const int q = e - 1, v1 = g->e2v[e][0], v2 = g->e2v[e][1];
    he[0][q] = g->u1x[e];
    he[1][q] = g->u1y[e];
    he[2][q] = g->costhu1[e];
    he[3][q] = g->sinthu1[e];
    he[4][q] = g->costhu2[e];
    he[5][q] = g->sinthu2[e];
    if (v1 < 1 || v1 >= g->szvS || v2 < 1 || v2 >= g->szvS) {
      emstr_set_error(E_UNSUP, e, "surface edge %d: end vertices outside the surface", e);
      ok = 0;
      break;
    }
    hev[4 * q + 0] = g->gridx[v1];
    hev[4 * q + 1] = g->gridy[v1];
    hev[4 * q + 2] = g->gridx[v2];
    hev[4 * q + 3] = g->gridy[v2];
  }
  for (n = 0; n < nvl && ok; n++) {
    const int v = n + 1;
    hvx[n] = g->gridx[v];
    hvy[n] = g->gridy[v];
    if (g->nvc[v] > nvcm) {
      emstr_set_error(E_UNSUP, v, "vertex %d: nvc above nvcm", v);
      ok = 0;
    }
    for (j = 0; j < nvcm; j++) {
      int cj = (j < g->nvc[v]) ? g->v2c[v][j + 1] : 0;
      int cl = (cj > 0 && cj < g->szcS) ? c->cs2col[cj] : -1;
      if (cj && cl < 0) {
        emstr_set_error(E_UNSUP, v, "vertex %d: centre %d is not a surface centre", v, cj);
        ok = 0;
        break;
      }
      hvcol[n * nvcm + j] = cl;
    }
    /* -1 padding at the end only */
    int m = 0;
    for (j = 0; j < nvcm; j++)
      if (hvcol[n * nvcm + j] >= 0) hvcol[n * nvcm + m++] = hvcol[n * nvcm + j];
    for (; m < nvcm; m++) hvcol[n * nvcm + m] = -1;
  }
  for (col = 0; col < ncol && ok; col++) {
    const int cs = c->col2cs[col], np = g->npe[cs];
    int sl[8];
    for (j = 0; j < np; j++) {
      const int v = g->c2v[j + 1][cs];
      if (v < 1 || v >= g->szvS) {
        emstr_set_error(E_UNSUP, cs, "centre %d: vertex %d outside the surface", cs, j + 1);
        ok = 0;
        break;
      }
      hcvt[j * ncol + col] = v - 1;
      sl[j] = j;
    }
    /* edge slots by ascending surface edge */
    for (int a = 1; a < np; a++)
      for (int b = a; b > 0 && g->c2e[sl[b] + 1][cs] < g->c2e[sl[b - 1] + 1][cs]; b--) {
        int t = sl[b];
        sl[b] = sl[b - 1];
        sl[b - 1] = t;
      }
    for (j = 0; j < np; j++) heord[j * ncol + col] = sl[j];
  }
  /* centre minus vertex j of every column (zero past its vertices) */
  double *hvxy = (double *)calloc((size_t)ncol * EMSTR_FF_NPE * 2, sizeof(double));
  for (col = 0; col < ncol && ok && hvxy; col++)
    for (j = 0; j < g->npe[c->col2cs[col]]; j++) {
      const int vt = hcvt[j * ncol + col];
      hvxy[((size_t)col * EMSTR_FF_NPE + j) * 2 + 0] = hcx[col] - hvx[vt];
      hvxy[((size_t)col * EMSTR_FF_NPE + j) * 2 + 1] = hcy[col] - hvy[vt];
    }
  /* wet-list position of every dense cell; every cell from the bed up
   * must be on the wet list (the vertex ranges read them by position) */
  int *hiq = (int *)malloc((size_t)nz * ncol * sizeof(int));
  for (size_t i = 0; i < (size_t)nz * ncol; i++) hiq[i] = -1;
  for (n = 0; n < d->nwet && ok; n++) {
    const int cw = g->w3_t[n + 1];
    hiq[(size_t)g->s2k[cw] * ncol + c->cs2col[g->m2d[cw]]] = n;
  }
  for (col = 0; col < ncol && ok; col++)
    for (int k = c->h_kb[col]; k < nz && ok; k++)
      if (hiq[(size_t)k * ncol + col] < 0) {
        emstr_set_error(E_UNSUP, c->col2cs[col], "column of centre %d: layer %d is not wet",
                        c->col2cs[col], k);
        ok = 0;
      }
  /* every wet centre against the model */
  int *lsq_off = (int *)calloc((size_t)d->nwet + 1, sizeof(int));
  size_t nent = 0;
  for (n = 0; n < d->nwet && ok; n++) {
    const int cw = g->w3_t[n + 1];
    const qweights *lw = &wc->lw[cw];
    if (lw->ncells < 1 || !lw->cells || !lw->B) {
      emstr_set_error(E_ARG, cw, "centre %d: no least-squares stencil", cw);
      ok = 0;
      break;
    }
    if (lw->ncells > 64) { /* the kernel holds a stencil entry a lane */
      emstr_set_error(E_UNSUP, cw, "centre %d: more than 64 least-squares stencil entries", cw);
      ok = 0;
      break;
    }
    lsq_off[n] = (int)nent;
    nent += lw->ncells;
    if (nent > 0x7fffffff) {
      emstr_set_error(E_UNSUP, cw, "least-squares stencils too large");
      ok = 0;
    }
  }
  if (ok) lsq_off[d->nwet] = (int)nent;
  int *lslot = ok ? (int *)malloc((nent ? nent : 1) * sizeof(int)) : NULL;
  int *lgc = ok ? (int *)malloc((nent ? nent : 1) * sizeof(int)) : NULL;
  double *lB = ok ? (double *)malloc((nent ? nent : 1) * 2 * sizeof(double)) : NULL;
  /* the distinct centres each wet centre reads (itself first, then the
   * stencil's, then the vertex ranges'), by wet-list position, and the
   * vertex ranges as slots of that list, four bytes per vertex */
  int *lr_off = ok ? (int *)malloc(((size_t)d->nwet + 1) * sizeof(int)) : NULL;
  int *lr_q = ok ? (int *)malloc(((size_t)d->nwet * EMSTR_FF_ROWS + 1) * sizeof(int)) : NULL;
  unsigned *vsl =
      ok ? (unsigned *)malloc(((size_t)d->nwet * EMSTR_FF_NPE + 1) * sizeof(unsigned)) : NULL;
  size_t nrow = 0;
  if (ok && (!lslot || !lgc || !lB || !lr_off || !lr_q || !vsl || !hvxy)) {
    emstr_set_error(E_DEVICE, -1, "FFSL stencil tables: host allocation");
    ok = 0;
  }
  for (n = 0; n < d->nwet && ok; n++) {
    const int cw = g->w3_t[n + 1], cs = g->m2d[cw], k = g->s2k[cw];
    col = c->cs2col[cs];
    const int np = g->npe[cs];
    for (j = 1; j <= np && ok; j++) {
      /* vertex j: same surface vertex, wet centres around it = columns
       * of the surface vertex wet at layer k */
      const int v = g->c2v[j][cw], vs = hcvt[(j - 1) * ncol + col] + 1;
      if (v < 1 || v >= g->szv || g->m2dv[v] != vs) {
        emstr_set_error(E_UNSUP, cw, "centre %d: vertex %d is not its surface vertex", cw, j);
        ok = 0;
        break;
      }
      int a[16], b[16], na = 0, nb = 0;
      for (int i = 1; i <= g->nvc[vs] && na < 16; i++)
        if (g->v2c[v][i]) a[na++] = g->v2c[v][i];
      for (int i = 0; i < nvcm && nb < 16; i++) {
        int cv = hvcol[(vs - 1) * nvcm + i];
        if (cv >= 0 && k >= c->h_kb[cv]) b[nb++] = c->sp_of[(ll)k * ncol + cv];
      }
      qsort(a, na, sizeof(int), icmp);
      qsort(b, nb, sizeof(int), icmp);
      if (na != nb || memcmp(a, b, na * sizeof(int))) {
        emstr_set_error(E_UNSUP, cw, "centre %d: centres around vertex %d differ", cw, j);
        ok = 0;
        break;
      }
    }
    /* wet edges in the order of w3_e1 */
    int last = 0;
    for (int sl = 0; sl < np && ok; sl++) {
      const int jj = heord[sl * ncol + col] + 1;
      const int e3 = g->c2e[jj][cw];
      if (!epos[e3]) continue; /* not a wet edge */
      if (epos[e3] < last) {
        emstr_set_error(E_UNSUP, cw, "centre %d: edge order differs from w3_e1", cw);
        ok = 0;
      }
      last = epos[e3];
    }
    /* stencil */
    const qweights *lw = &wc->lw[cw];
    int *rw = lr_q + nrow, nr = 1;
    rw[0] = n;
    lr_off[n] = (int)nrow;
    for (int i = 0; i < lw->ncells && ok; i++) {
      int cn = lw->cells[i];
      const int isg = (cn > 0 && cn < g->szc && gbin[cn]) ? 1 : 0;
      if (isg) cn = gbin[cn];
      int ccol = (cn > 0 && cn < g->szc) ? c->cs2col[g->m2d[cn]] : -1;
      int ck = (cn > 0 && cn < g->szc) ? g->s2k[cn] : -1;
      if (ccol < 0 || ck < c->h_kb[ccol] || ck >= nz || c->sp_of[(ll)ck * ncol + ccol] != cn) {
        emstr_set_error(E_UNSUP, cw, "centre %d: stencil entry %d is not wet or a ghost", cw,
                        lw->cells[i]);
        ok = 0;
        break;
      }
      const size_t o = (size_t)lsq_off[n] + i;
      const int sl = ff_slot(rw, &nr, hiq[(size_t)ck * ncol + ccol]);
      if (sl < 0) {
        emstr_set_error(E_UNSUP, cw, "centre %d: more than %d centres in its stencil", cw,
                        EMSTR_FF_ROWS);
        ok = 0;
        break;
      }
      lslot[o] = sl;
      /* the entry's column; a ghost (~column) reads the step-start
       * value of its wet centre */
      lgc[o] = isg ? ~ccol : ccol;
      lB[2 * o + 0] = lw->B[i][1];
      lB[2 * o + 1] = lw->B[i][2];
    }
    /* vertex ranges: the wet centres around each vertex in column order */
    for (j = 0; j < EMSTR_FF_NPE && ok; j++) {
      unsigned word = 0xffffffffu;
      if (j < np) {
        const int vs = hcvt[j * ncol + col];
        int m = 0;
        for (int i = 0; i < nvcm; i++) {
          const int cv = hvcol[vs * nvcm + i];
          if (cv < 0) break;
          if (k < c->h_kb[cv]) continue;
          const int sl = ff_slot(rw, &nr, hiq[(size_t)k * ncol + cv]);
          if (sl < 0) {
            emstr_set_error(E_UNSUP, cw, "centre %d: more than %d centres in its stencil", cw,
                            EMSTR_FF_ROWS);
            ok = 0;
            break;
          }
          word = (word & ~(0xffu << (8 * m))) | ((unsigned)sl << (8 * m));
          m++;
        }
      }
      vsl[(size_t)n * EMSTR_FF_NPE + j] = word;
    }
    nrow += nr;
  }
  if (ok) lr_off[d->nwet] = (int)nrow;
  if (ok) {
    D->cx = (double *)dalloc((size_t)ncol * sizeof(double), &ok);
    D->cy = (double *)dalloc((size_t)ncol * sizeof(double), &ok);
    double **ep[] = {&D->ux, &D->uy, &D->n1x, &D->n1y, &D->t2x, &D->t2y};
    for (n = 0; n < 6; n++) *ep[n] = (double *)dalloc((size_t)ne2 * sizeof(double), &ok);
    D->ev = (double *)dalloc((size_t)ne2 * 4 * sizeof(double), &ok);
    D->vxy = (double *)dalloc((size_t)ncol * EMSTR_FF_NPE * 2 * sizeof(double), &ok);
    D->eord = (int *)dalloc((size_t)npem * ncol * sizeof(int), &ok);
    D->lsq_off = (int *)dalloc(((size_t)d->nwet + 1) * sizeof(int), &ok);
    D->lsq_s = (int *)dalloc((nent ? nent : 1) * sizeof(int), &ok);
    D->lsq_gc = (int *)dalloc((nent ? nent : 1) * sizeof(int), &ok);
    D->lr_off = (int *)dalloc(((size_t)d->nwet + 1) * sizeof(int), &ok);
    D->lr_q = (int *)dalloc((nrow ? nrow : 1) * sizeof(int), &ok);
    D->vsl = (unsigned *)dalloc(((size_t)d->nwet * EMSTR_FF_NPE + 1) * sizeof(unsigned), &ok);
    D->lsq_B = (double *)dalloc((nent ? nent : 1) * 2 * sizeof(double), &ok);
    if (!ok) emstr_set_error(E_DEVICE, -1, "FFSL geometry allocation failed");
  }
  if (ok) {
    D->nvl = nvl;
    D->nvcm = nvcm;
    double **ep[] = {&D->ux, &D->uy, &D->n1x, &D->n1y, &D->t2x, &D->t2y};
    int r = up(c, D->cx, hcx, ncol * sizeof(double)) || up(c, D->cy, hcy, ncol * sizeof(double));
    for (n = 0; n < 6; n++) r = r || up(c, *ep[n], he[n], ne2 * sizeof(double));
    r = r || up(c, D->ev, hev, (size_t)ne2 * 4 * sizeof(double)) ||
        up(c, D->vxy, hvxy, (size_t)ncol * EMSTR_FF_NPE * 2 * sizeof(double)) ||
        up(c, D->eord, heord, (size_t)npem * ncol * sizeof(int)) ||
        up(c, D->lsq_off, lsq_off, ((size_t)d->nwet + 1) * sizeof(int)) ||
        up(c, D->lsq_s, lslot, (nent ? nent : 1) * sizeof(int)) ||
        up(c, D->lsq_gc, lgc, (nent ? nent : 1) * sizeof(int)) ||
        up(c, D->lr_off, lr_off, ((size_t)d->nwet + 1) * sizeof(int)) ||
        up(c, D->lr_q, lr_q, (nrow ? nrow : 1) * sizeof(int)) ||
        up(c, D->vsl, vsl, ((size_t)d->nwet * EMSTR_FF_NPE + 1) * sizeof(unsigned)) ||
        up(c, D->lsq_B, lB, (nent ? nent : 1) * 2 * sizeof(double));
    if (hipStreamSynchronize(c->st) != hipSuccess) r = 1;
    if (r) {
      emstr_set_error(E_DEVICE, -1, "FFSL geometry upload failed");
      ok = 0;
    }
  }
  free(hcx);
  free(hcy);
  for (n = 0; n < 6; n++) free(he[n]);
  free(hev);
  free(hcvt);
  free(heord);
  free(hvx);
  free(hvy);
  free(hvcol);
  free(epos);
  free(gbin);
  free(lsq_off);
  free(lslot);
  free(lgc);
  free(lB);
  free(hiq);
  free(lr_off);
  free(lr_q);
  free(vsl);
  free(hvxy);
  if (!ok) return E_UNSUP;
  if ((n = us_ffsl_vlsq(c))) return n;
  c->ffgeom = 1;
  return 0;
}

/* the rotated edge velocity of every land edge's ghost (ffsl_init,
 * transport.c:2652-2666) */
static double ghost_u(const geometry_t *g, const window_t *wd, int e, int comp) {
  const int es = g->m2de[e];
  return comp ? wd->u1[e] * g->sinthu1[es] + wd->u2[e] * g->sinthu2[es]
              : wd->u1[e] * g->costhu1[es] + wd->u2[e] * g->costhu2[es];
}

/* FFSL fields: u2 on wet edges, wtop / wbot, and, when sub-stepping, the
 * centred velocities of wet centres and of the ghosts across land edges
 * (as the window holds them, and rotated from the edge) */
static int us_ffsl_fields(usctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  const size_t N3 = (size_t)c->ncol * d->nz, NE3 = (size_t)c->ne2 * d->nz;
  int ok = 1, r;
  if (!c->ffgeom && (r = us_ffsl_geom(c))) return r;
  if (!D->u2) {
    D->u2 = (double *)dalloc(NE3 * sizeof(double), &ok);
    D->wtop = (double *)dalloc((size_t)c->ncol * sizeof(double), &ok);
    D->wbot = (double *)dalloc((size_t)c->ncol * sizeof(double), &ok);
    if (!ok) return emstr_set_error(E_DEVICE, -1, "FFSL field allocation failed");
  }
  if (up_edges(c, D->u2, wd->u2) || up_land_edges(c, D->u2, wd->u2) ||
      up_land_edges(c, D->u1, wd->u1) || up_cols(c, D->wtop, wd->wtop) ||
      up_cols(c, D->wbot, wd->wbot))
    return E_DEVICE;
  c->ghost_edges = 0;
  /* the centred velocities of the wet centres: the streamline velocity
   * fit, and the sub-step's Courant numbers */
  if (!D->uc) {
    double **a[] = {&D->uc, &D->vc};
    for (int n = 0; n < 2; n++) *a[n] = (double *)dalloc(N3 * sizeof(double), &ok);
    if (!ok) return emstr_set_error(E_DEVICE, -1, "FFSL field allocation failed");
  }
  if (up_cells(c, D->uc, wd->u) || up_cells(c, D->vc, wd->v)) return E_DEVICE;
  if (!(c->wc->stab & (SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER))) return 0;
  if (!D->gu0) {
    double **b[] = {&D->gu0, &D->gv0, &D->gue, &D->gve};
    for (int n = 0; n < 4; n++) *b[n] = (double *)dalloc(NE3 * sizeof(double), &ok);
    if (!ok) return emstr_set_error(E_DEVICE, -1, "FFSL field allocation failed");
  }
  double *h = (double *)calloc(NE3 * 4, sizeof(double));
  for (int ee = g->v3_e1 + 1; ee <= g->n3_e1; ee++) {
    const int e = g->w3_e1[ee], cg = g->e2c[e][1];
    const size_t i = (size_t)g->s2k[g->e2c[e][0]] * c->ne2 + (g->m2de[e] - 1);
    h[i] = wd->u[cg];
    h[NE3 + i] = wd->v[cg];
    h[2 * NE3 + i] = ghost_u(g, wd, e, 0);
    h[3 * NE3 + i] = ghost_u(g, wd, e, 1);
  }
  r = up(c, D->gu0, h, NE3 * sizeof(double)) || up(c, D->gv0, h + NE3, NE3 * sizeof(double)) ||
      up(c, D->gue, h + 2 * NE3, NE3 * sizeof(double)) ||
      up(c, D->gve, h + 3 * NE3, NE3 * sizeof(double));
  hipStreamSynchronize(c->st);
  free(h);
  return r ? emstr_set_error(E_DEVICE, -1, "FFSL ghost velocities") : 0;
}

/* ffsl_init's toggle (transport.c:2624-2627, 2722): 1 if it runs on this
 * call; it then writes the ghost velocities into the window's u, v */
static int us_ffsl_init(usctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  if (wc->flag & F_FFSL) {
    wc->flag &= ~F_FFSL;
    return 0;
  }
  for (int ee = g->v3_e1 + 1; ee <= g->n3_e1; ee++) {
    const int e = g->w3_e1[ee], cg = g->e2c[e][1];
    wd->u[cg] = ghost_u(g, wd, e, 0);
    wd->v[cg] = ghost_u(g, wd, e, 1);
  }
  wc->flag |= F_FFSL;
  c->ghost_edges = 1;
  return 1;
}

/* the stamp of a sub-step's tr_moc marks (need[] starts at zero).  A
 * stamp that comes round again after 2^31 sub-steps can only leave extra
 * centres marked, which costs work but changes no value read. */
static int us_ffstamp(usctx *c) {
  if (++c->ffstamp <= 0) c->ffstamp = 1;
  return c->ffstamp;
}

static int us_ffsl_alloc(usctx *c) {
  emstr_dev *d = &c->D.d;
  const size_t N3 = (size_t)d->n3, NE3 = (size_t)c->ne2 * d->nz;
  const size_t nt = (size_t)(d->ntr > 0 ? d->ntr : 1);
  int ok = 1;
  if (c->d_nw) return 0;
  double **a[] = {&c->d_nw, &c->d_crc, &c->d_crf, &c->d_cour};
  for (int n = 0; n < 4; n++) *a[n] = (double *)dalloc(N3 * sizeof(double), &ok);
  c->d_clc = (int *)dalloc(N3 * sizeof(int), &ok);
  c->d_clf = (int *)dalloc(N3 * sizeof(int), &ok);
  c->d_need = (int *)dalloc(N3 * sizeof(int), &ok); /* zeroed; stamps start at 1 */
  c->ffstamp = 0;
  c->d_lci = (int *)dalloc(NE3 * sizeof(int), &ok);
  c->d_ld = (double *)dalloc(5 * NE3 * sizeof(double), &ok);
  /* tr_mod rows [nwet][tqs], tqs = tracers rounded up to 16 (128 bytes) */
  c->D.tqs = (int)((nt + 15) / 16 * 16);
  c->d_tm = (double *)dalloc(((size_t)d->nwet + 1) * c->D.tqs * sizeof(double), &ok);
  c->d_lq = (double *)dalloc(4 * nt * N3 * sizeof(double), &ok);
  c->d_te = (double *)dalloc(nt * NE3 * sizeof(double), &ok);
  c->d_moc = (double *)dalloc(nt * N3 * sizeof(double), &ok);
  c->d_fferr = (int *)dalloc(sizeof(int), &ok);
  if (ok && hipMemset(c->d_fferr, 0, sizeof(int)) != hipSuccess) ok = 0;
  if (!ok) {
    emstr_set_error(E_DEVICE, -1, "FFSL work arrays (%zu tracers)", nt);
    return E_DEVICE;
  }
  return 0;
}

/* the FP32 form's fluxes and FFSL intermediates */
static int us_f32_alloc(usctx *c) {
  emstr_dev *d = &c->D.d;
  const size_t N3 = (size_t)d->n3, NE3 = (size_t)c->ne2 * d->nz;
  const size_t nt = (size_t)(d->ntr > 0 ? d->ntr : 1);
  int ok = 1;
  if (c->f_fx) return 0;
  c->f_fx = (float *)dalloc(nt * NE3 * sizeof(float), &ok);
  c->f_fz = (float *)dalloc((nt * N3 + (size_t)c->ncol) * sizeof(float), &ok);
  c->f_tm = (float *)dalloc(((size_t)d->nwet + 1) * ((nt + 15) / 16 * 16) * sizeof(float), &ok);
  c->f_lq = (float *)dalloc(4 * nt * N3 * sizeof(float), &ok);
  c->f_te = (float *)dalloc(nt * NE3 * sizeof(float), &ok);
  c->f_moc = (float *)dalloc(nt * N3 * sizeof(float), &ok);
  if (!ok) {
    emstr_set_error(E_DEVICE, -1, "FP32 form buffers (%zu tracers)", nt);
    return E_DEVICE;
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* the step                                                          */
/*-------------------------------------------------------------------*/
/* advect_diffuse (hd-us tracers.c:428-1157); 0, or the 2D centre of a
 * column where the sub-step limit is exceeded */
/* runge_kutta's weights (tracers.c:110-129) for 1 to 3 stages: of the
 * sub-step's starting value, of the stage's value, of its tendency (and
 * the stage's volume); forward Euler and the SSP schemes of Shu and
 * Osher */
static const double us_rkw[3][3][3] = {
    {{1.0, 0.0, 1.0}},
    {{1.0, 0.0, 1.0}, {0.50, 0.50, 0.50}},
    {{1.0, 0.0, 1.0}, {0.75, 0.25, 0.25}, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0}}};
/* 5 stages (Spiteri and Ruuth): stages 0-3 as above; the last combines
 * the values of stages 2, 3 and 4 and the tendencies of stages 3 and 4,
 * both weighted as the reference weights them (tracers.c:1199-1205), and
 * the volume weight */
static const double us_rkw5[4][3] = {{1.00, 0.00, 0.391752226571890},
                                     {0.444370493651235, 0.555629506348765, 0.368410593050371},
                                     {0.620101851488403, 0.379898148511597, 0.251891774271694},
                                     {0.178079954393132, 0.821920045606868, 0.544974750228521}};
static const double us_rkw5_last[5] = {0.517231671970585, 0.096059710526147,
                                       0.063692468666290, 0.386708617503269,
                                       0.127384937332581};

/* a tracer buffer not in x[0..nx-1] (-1 entries ignored) */
static int us_free_buf(int nbuf, const int *x, int nx) {
  for (int i = 0; i < nbuf; i++) {
    int used = 0;
    for (int j = 0; j < nx; j++) used |= (x[j] == i);
    if (!used) return i;
  }
  return -1;
}

static int us_adv(usctx *c) {
  emstr_usdev *D = &c->D;
  emstr_dev *d = &D->d;
  win_priv_t *wc = c->wc;
  double dtu = d->dt, dtm = d->dt, trem;
  int itermax = 20, first = 1, r0;
  const int ffsl = D->ffsl;
  if (!c->have_cols) return emstr_set_error(E_ARG, -1, "no set_dz data on the device");
  if (c->f32) {
    if (!ffsl || wc->rkstage != 1)
      return emstr_set_error(E_UNSUP, -1, "FP32 form: FFSL with one Runge-Kutta stage only");
    if ((r0 = us_f32_alloc(c))) return r0;
  }
  if (ffsl) {
    /* tracers.c:567-590: centred w, ffsl_init */
    itermax = 200;
    if ((r0 = us_ffsl_alloc(c))) return r0;
    const int init = us_ffsl_init(c);
    if (emstr_us_launch_ffsl_step(D, init, c->d_nw, c->st))
      return emstr_set_error(E_DEVICE, -1, "FFSL step launch");
    /* delaunay_reinit of u and v (transport.c:2709-2711): the velocity
     * fit of the streamlines, kept until ffsl_init runs again; the fit
     * reads only fields of the upload, so with no upload since the last
     * one it would come out the same */
    if (init && !(D->vfit_ok && c->vfit_gen == c->fgen)) {
      if (emstr_us_launch_ffsl_vfit(D, c->st))
        return emstr_set_error(E_DEVICE, -1, "FFSL velocity fit launch");
      D->vfit_ok = 1;
      c->vfit_gen = c->fgen;
    }
  }
  if (wc->stab & (SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER)) {
    unsigned long long init, r;
    double big = dtu;
    memcpy(&init, &big, sizeof(init));
    HCHK(hipMemcpyAsync(c->d_dtm, &init, sizeof(init), hipMemcpyHostToDevice, c->st));
    const double *gu = c->ghost_edges ? D->gue : D->gu0;
    const double *gv = c->ghost_edges ? D->gve : D->gv0;
    if (emstr_us_launch_dtm(D, dtu, c->d_dtm, wc->trsf, gu, gv, c->st))
      return emstr_set_error(E_DEVICE, -1, "dtm launch");
    HCHK(hipMemcpyAsync(&r, c->d_dtm, sizeof(r), hipMemcpyDeviceToHost, c->st));
    HCHK(hipStreamSynchronize(c->st));
    memcpy(&dtm, &r, sizeof(dtm));
    if (dtm != dtu) dtm = floor(dtm);
  }
  if (dtm == 0.0 || (int)(dtu / dtm) > itermax) {
    emstr_set_error(E_SUBSTEP, -1,
                    "tracer_step_3d: maximum number of sub-steps (%d) exceeded", itermax);
    /* the reference returns the column where the limit was found; the
     * first processed column stands for it */
    return c->wc->vcs > 0 ? c->g->m2d[c->wc->s1[1]] : 1;
  }
  if (!c->d_fx || !c->d_fz) {
    const size_t nt = (size_t)(d->ntr > 0 ? d->ntr : 1);
    const size_t bytes = nt * d->nz * c->ne2 * sizeof(double);
    const size_t bz = nt * (size_t)d->n3 * sizeof(double) + (size_t)c->ncol * sizeof(double);
    if (!c->d_fx && hipMalloc((void **)&c->d_fx, bytes) != hipSuccess) {
      c->d_fx = NULL;
      return emstr_set_error(E_DEVICE, -1, "edge flux buffer (%zu bytes)", bytes);
    }
    if (hipMalloc((void **)&c->d_fz, bz) != hipSuccess) {
      c->d_fz = NULL;
      return emstr_set_error(E_DEVICE, -1, "vertical flux buffer (%zu bytes)", bz);
    }
  }
  /* Runge-Kutta stages (tracers.c:841-1098): each stage's fluxes come
   * from the previous stage's values; the update combines them with the
   * sub-step's starting values, which stay in their buffer, as do the
   * step's starting values (the ghosts read them) */
  const int nrk = wc->rkstage, nbuf = (nrk == 5) ? 6 : (nrk > 1) ? 4 : 3;
  {
    const size_t bytes = (size_t)(d->ntr > 0 ? d->ntr : 1) * d->n3 * sizeof(double);
    for (int i = 3; i < nbuf; i++)
      if (!c->tr[i] && hipMalloc((void **)&c->tr[i], bytes) != hipSuccess) {
        c->tr[i] = NULL;
        return emstr_set_error(E_DEVICE, -1, "Runge-Kutta stage buffer (%zu bytes)", bytes);
      }
    if (nrk == 5 && !c->d_g3 && hipMalloc((void **)&c->d_g3, bytes) != hipSuccess) {
      c->d_g3 = NULL;
      return emstr_set_error(E_DEVICE, -1, "Runge-Kutta tendency buffer (%zu bytes)", bytes);
    }
  }
  const int b0 = c->cur;
  int in = b0;
  trem = dtu;
  while (trem > 0) {
    double u = dtm;
    if (trem < u) u = trem;
    wc->b1 = u;
    if (u > 0.0) {
      if (emstr_launch_subeta(d, u, trem, first, c->st))
        return emstr_set_error(E_DEVICE, -1, "subeta launch");
      if (ffsl && emstr_us_launch_ffsl_sub(D, u, c->d_nw, c->d_clc, c->d_crc, c->d_clf,
                                           c->d_crf, c->d_cour, c->d_edn, c->nesp, c->d_lci,
                                           c->d_ld, c->d_fferr, c->st))
        return emstr_set_error(E_DEVICE, -1, "FFSL launch");
      int src = in, sin[5] = {in, -1, -1, -1, -1};
      for (int tc = 0; tc < nrk; tc++) {
        /* with 5 stages the inputs of stages 2 and 3 stay until the last */
        const int keep[5] = {b0, in, src, (nrk == 5) ? sin[2] : -1, (nrk == 5) ? sin[3] : -1};
        const int dst = us_free_buf(nbuf, keep, 5);
        const double *w;
        int mode = 0;
        if (nrk == 5) {
          w = (tc < 4) ? us_rkw5[tc] : us_rkw5_last;
          mode = (tc == 3) ? 1 : (tc == 4) ? 2 : 0;
        } else
          w = us_rkw[nrk - 1][tc];
        const int last = (tc == nrk - 1);
        const double *ts[3] = {c->tr[in], (sin[2] >= 0) ? c->tr[sin[2]] : NULL,
                               (sin[3] >= 0) ? c->tr[sin[3]] : NULL};
        /* CLIP (one stage): each centre's range over its stencil, of the
         * stage's values (ghosts: the step's starting values) */
        const int clip = ffsl && (wc->fillf & CLIP);
        if (clip) {
          if (!c->d_bmn) {
            const size_t bytes = (size_t)(d->ntr > 0 ? d->ntr : 1) * d->n3 * sizeof(double);
            int ok = 1;
            c->d_bmn = (double *)dalloc(bytes, &ok);
            c->d_bmx = (double *)dalloc(bytes, &ok);
            if (!ok) return emstr_set_error(E_DEVICE, -1, "CLIP bounds (%zu bytes)", bytes);
          }
          if (emstr_us_launch_ffsl_bounds(D, c->tr[src], c->tr[b0], c->d_bmn, c->d_bmx,
                                          c->d_fferr, c->st))
            return emstr_set_error(E_DEVICE, -1, "CLIP bounds launch");
        }
        const double *bmn = clip ? c->d_bmn : NULL, *bmx = clip ? c->d_bmx : NULL;
        if (c->f32) {
          /* the FP32 form (one stage): the same maps, FP32 fluxes */
          if (emstr_us_launch_ffsl_flux32(D, u, c->tr[src], c->d_nw, c->d_clc, c->d_crc,
                                          c->d_clf, c->d_crf, c->d_cour, c->d_edn, c->nesp,
                                          c->d_lci, c->d_ld, c->f_tm, c->f_lq, c->f_te,
                                          c->f_moc, c->f_fx, c->f_fz, c->anydif, c->d_need,
                                          us_ffstamp(c), c->d_fferr, c->st) ||
              emstr_us_launch_update32(D, c->tr[in], c->tr[src], c->tr[dst], c->f_fx,
                                       c->f_fz, u, w, bmn, bmx, c->st))
            return emstr_set_error(E_DEVICE, -1, "FFSL launch");
        } else if (ffsl) {
          if (emstr_us_launch_ffsl_flux(D, u, c->tr[src], c->d_nw, c->d_clc, c->d_crc,
                                        c->d_clf, c->d_crf, c->d_cour, c->d_edn, c->nesp,
                                        c->d_lci, c->d_ld, c->d_tm, c->d_lq, c->d_te, c->d_moc,
                                        c->d_fx, c->d_fz, c->anydif, c->d_need, us_ffstamp(c),
                                        c->d_fferr, c->st) ||
              emstr_us_launch_update(D, ts, c->tr[src], c->tr[dst], c->d_fx, c->d_fz, u, w,
                                     last, mode, c->d_g3, bmn, bmx, c->st))
            return emstr_set_error(E_DEVICE, -1, "FFSL launch");
        } else if (emstr_us_launch_adv(D, ts, c->tr[src], c->tr[dst], c->tr[b0], c->d_edn,
                                       c->nesp, c->d_fx, c->d_fz, u, c->anydif, d->ultimate, w,
                                       last, mode, c->d_g3, c->st))
          return emstr_set_error(E_DEVICE, -1, "advection launch");
        src = dst;
        if (tc + 1 < 5) sin[tc + 1] = dst;
      }
      if (emstr_launch_osub(d, c->st)) return emstr_set_error(E_DEVICE, -1, "osub launch");
      in = src;
    }
    trem -= u;
    first = 0;
  }
  c->cur = in;
  return 0;
}

static int us_vdiff(usctx *c) {
  const int scratch = (c->cur + 1) % 3;
  if (!c->have_cols) return emstr_set_error(E_ARG, -1, "no set_dz data on the device");
  if (emstr_launch_vdiff(&c->D.d, c->tr[c->cur], c->tr[scratch], c->d_err, c->st))
    return emstr_set_error(E_DEVICE, -1, "vdiff launch");
  return 0;
}

static int us_sync(usctx *c) {
  int err = 0;
  HCHK(hipStreamSynchronize(c->st));
  if (c->d_fferr) {
    HCHK(hipMemcpy(&err, c->d_fferr, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
      hipMemset(c->d_fferr, 0, sizeof(int));
      if (err == 1)
        return emstr_set_error(E_UNSUP, -1, "FFSL: a streamline goes on into a wet cell "
                                            "before ffsl_init has run on this window (its "
                                            "velocity fit would be the model's initial one)");
      if (err == 5)
        return emstr_set_error(E_UNSUP, -1, "FFSL: a streamline of more than 100 segments "
                                            "(the reference stops the model)");
      if (err == 6)
        return emstr_set_error(E_DEVICE, -1, "FFSL: more streamline segments than the %d "
                                             "records of a sub-step", c->D.sg_cap);
      if (err == 3)
        return emstr_set_error(E_UNSUP, -1, "FFSL with CLIP: a column with water is not "
                                            "processed (its cells would be clamped to bounds "
                                            "kept from an earlier call)");
      return emstr_set_error(E_UNSUP, -1, "FFSL: a vertical trajectory does not end "
                                          "(zero layer thickness)");
    }
  }
  HCHK(hipMemcpy(&err, c->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (err) {
    int cell = c->sp_of[err - 1];
    hipMemset(c->d_err, 0, sizeof(int));
    return emstr_set_error(E_PIVOT, cell, "implicit_vdiff_tr: zero divisor at cell %d", cell);
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* set_dz on the device, lists back to the host (hd-us 7231-7380)    */
/*-------------------------------------------------------------------*/
static int us_set_dz(usctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->D.d;
  const int ncol = c->ncol;
  int cc, vc = 1, col;
  d->hmin = wc->hmin;
  if (up_cols(c, d->eta, wd->eta) || up_cols(c, d->oldeta, wc->oldeta)) return E_DEVICE;
  for (cc = 1; cc <= ncol; cc++) {
    col = cc - 1;
    int ns = g->nsur_t[cc];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
  }
  if (up(c, d->knew, c->h_knew, ncol * sizeof(int))) return E_DEVICE;
  if (emstr_launch_setdz(d, c->d_proc, c->st)) return emstr_set_error(E_DEVICE, -1, "setdz");
  HCHK(hipMemcpyAsync(c->h_ks, d->ks, ncol * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipMemcpyAsync(c->h_kold, d->kold, ncol * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipMemcpyAsync(c->h_knew, d->knew, ncol * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  /* host lists in the reference's order */
#define SPAT(col, k) ((k) < c->h_kb[col] ? c->sedg[col] : c->sp_of[(ll)(k)*ncol + (col)])
  for (cc = 1; cc <= g->a2_t; cc++) {
    col = cc - 1;
    g->sur_t[cc] = SPAT(col, c->h_kold[col]);
    g->nsur_t[cc] = SPAT(col, c->h_knew[col]);
  }
  int ncp = 0;
  for (cc = 1; cc <= g->b2_t; cc++) {
    col = cc - 1;
    if (c->h_ks[col] < 0) continue;
    wc->s1[vc] = SPAT(col, c->h_ks[col]);
    wc->i1[vc] = g->bot_t[cc];
    wc->i2[vc] = g->nsur_t[cc];
    wc->i3[vc] = g->sur_t[cc];
    if (cc <= g->v2_t) wc->vcs1 = vc;
    c->h_cols[ncp++] = col;
    vc++;
  }
  const int vcs = vc - 1;
  wc->vcs = vcs;
  for (cc = 1; cc <= vcs; cc++) {
    col = c->h_cols[cc - 1];
    for (int k = c->h_ks[col] - 1; k >= c->h_kb[col]; k--) wc->s1[vc++] = SPAT(col, k);
    if (cc == wc->vcs1) wc->vc1 = vc - 1;
  }
  wc->vc = vc - 1;
#undef SPAT
  d->ncol = ncp;
  if (up(c, d->cols, c->h_cols, (size_t)(ncp ? ncp : 1) * sizeof(int))) return E_DEVICE;
  /* dz of wet centres */
  if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, d->dz, c->stage, c->st))
    return emstr_set_error(E_DEVICE, -1, "gather");
  double *tmp = (double *)malloc((size_t)g->szc * sizeof(double));
  HCHK(hipMemcpyAsync(tmp, c->stage, (size_t)g->szc * sizeof(double), hipMemcpyDeviceToHost,
                      c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (int q = 1; q <= g->v3_t; q++) {
    int cw = g->w3_t[q];
    wc->dz[cw] = tmp[cw];
  }
  free(tmp);
  c->have_cols = 1;
  return 0;
}

/*-------------------------------------------------------------------*/
/* drop-in entry points                                              */
/*-------------------------------------------------------------------*/
void set_dz(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  usctx *c;
  emstr_clear_error();
  if (wincon->sigma || wincon->thin_merge) {
    emstr_set_error(E_UNSUP, -1, "set_dz: sigma or thin-layer merging is not supported");
    return;
  }
  c = ctx_for(window, windat, wincon);
  if (!c) return;
  us_set_dz(c);
}

int advect_diffuse(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  usctx *c;
  int r;
  emstr_clear_error();
  c = ctx_for(window, windat, wincon);
  if (!c) return 1;
  if ((r = us_upload_fields(c)) || (r = us_upload_tracers(c))) return r;
  if ((r = us_adv(c))) return r;
  if ((r = us_download_tracers(c))) return r;
  return us_sync(c);
}

void vert_diffuse_3d(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  usctx *c;
  emstr_clear_error();
  c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (us_upload_fields(c) || us_upload_tracers(c)) return;
  if (us_vdiff(c)) return;
  if (us_download_tracers(c)) return;
  us_sync(c);
}

void tracer_step_3d(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  usctx *c;
  emstr_clear_error();
  c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (us_upload_fields(c) || us_upload_tracers(c)) return;
  if (us_adv(c)) return;
  if (us_vdiff(c)) return;
  if (us_download_tracers(c)) return;
  us_sync(c);
}

/* model/hd-us/tracers/tracers.c:177-239 over closed z-level windows: the
 * lateral ghosts of every window's tracers take their interior values,
 * then each window's step runs on the GPU in the order twin and its
 * tracers go back to the master, then the mass diagnostics.  The other
 * calls there do nothing at closed basins without multi-dt, flushing or
 * regions; the edge flux ghosts are not read (no flux through land). */
void tracer_step(master_t *master, geometry_t **window, window_t **windat,
                 win_priv_t **wincon, int nwindows) {
  int nn, n, cc, t;
  emstr_clear_error();
  if (!master || !window || !windat || !wincon || nwindows < 1) {
    emstr_set_error(E_ARG, -1, "tracer_step: null structure or no windows");
    if (master) master->crf = RS_RESTART;
    return;
  }
  if (master->mode2d) {
    emstr_set_error(E_UNSUP, -1, "tracer_step: 2D mode (tracer_step_2d) is not built");
    master->crf = RS_RESTART;
    return;
  }
  const int *twin = wincon[1]->twin;
  for (nn = 1; nn <= nwindows; nn++) {
    n = twin ? twin[nn] : nn;
    geometry_t *g = window[n];
    window_t *wd = windat[n];
    for (t = 0; t < wd->ntr; t++) {
      double *tr = wd->tr_wc[t];
      for (cc = 1; cc <= g->nbpt; cc++) tr[g->bpt[cc]] = tr[g->bin[cc]];
    }
  }
  for (nn = 1; nn <= nwindows; nn++) {
    n = twin ? twin[nn] : nn;
    tracer_step_3d(window[n], windat[n], wincon[n]);
    if (emstr_us_last_error()[0]) {
      master->crf = RS_RESTART;
      return;
    }
    if (master->win_data_empty_3d) master->win_data_empty_3d(master, window[n], windat[n], TRACERS);
  }
  for (nn = 1; nn <= nwindows; nn++) {
    n = twin ? twin[nn] : nn;
    if (master->win_data_empty_3d) master->win_data_empty_3d(master, window[n], windat[n], TMASS);
  }
}

/*-------------------------------------------------------------------*/
/* FFSL maps of the last sub-step, in the window's numbering          */
/*-------------------------------------------------------------------*/
static int us_ffsl_maps(usctx *c, int *cl, int *clzf, int *clzc) {
  geometry_t *g = c->g;
  emstr_dev *d = &c->D.d;
  const int ncol = c->ncol, ne2 = c->ne2, nz = g->nz;
  const size_t N3 = (size_t)ncol * nz, NE3 = (size_t)ne2 * nz;
  int q;
  if (!c->d_lci) return emstr_set_error(E_ARG, -1, "no FFSL step has run on this window");
  int *hl = (int *)malloc(NE3 * sizeof(int));
  int *hf = (int *)malloc(N3 * sizeof(int));
  int *hc = (int *)malloc(N3 * sizeof(int));
  const int cap = c->D.sg_cap;
  int *hn = (int *)malloc(NE3 * sizeof(int));
  int *hsc = (int *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(int));
  int *hsn = (int *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(int));
  int r = hipMemcpy(hl, c->d_lci, NE3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hf, c->d_clf, N3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hc, c->d_clc, N3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hn, c->D.lnx, NE3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(hsc, c->D.sg_ci, (size_t)cap * sizeof(int), hipMemcpyDeviceToHost) !=
              hipSuccess ||
          hipMemcpy(hsn, c->D.sg_nx, (size_t)cap * sizeof(int), hipMemcpyDeviceToHost) !=
              hipSuccess;
  if (!r) {
#define SPL(col, L) ((L) < c->h_kb[col] ? c->sedg[col] : c->sp_of[(ll)(L) * ncol + (col)])
    /* the centre where every wet edge's streamline ends (its upstream
     * centre unless it goes on into other cells) */
    for (q = 0; q < c->nesp; q++) {
      const ll ie = c->h_edn[q];
      const int k = (int)(ie / ne2);
      int col = hl[ie];
      for (int x = hn[ie]; x >= 0 && x < cap; x = hsn[x]) col = hsc[x];
      cl[c->h_esp[q]] = c->sp_of[(ll)k * ncol + col];
    }
    /* departure centres of the faces to process and of every wet centre */
    for (q = 0; q < d->nwet; q++) {
      const int cw = g->w3_t[q + 1];
      const int col = c->cs2col[g->m2d[cw]], k = g->s2k[cw];
      const size_t i = (size_t)k * ncol + col;
      clzc[cw] = SPL(col, hc[i]);
      if (c->h_ks[col] >= 0 && k <= c->h_ks[col]) clzf[cw] = SPL(col, hf[i]);
    }
#undef SPL
  }
  free(hl);
  free(hf);
  free(hc);
  free(hn);
  free(hsc);
  free(hsn);
  return r ? emstr_set_error(E_DEVICE, -1, "FFSL map download") : 0;
}

int emstr_us_ffsl_maps(geometry_t *window, int *cl, int *clzf, int *clzc) {
  usctx *c = NULL;
  int i;
  emstr_clear_error();
  pthread_mutex_lock(&reg_lock);
  for (i = 0; i < MAXCTX; i++)
    if (reg_g[i] == window) c = reg_c[i];
  pthread_mutex_unlock(&reg_lock);
  if (!c) return emstr_set_error(E_ARG, -1, "no device mirror for this window");
  return us_ffsl_maps(c, cl, clzf, clzc);
}

/*-------------------------------------------------------------------*/
/* device-resident interface (no counterpart in the reference)       */
/*-------------------------------------------------------------------*/
emstr_us_ctx *emstr_us_create(geometry_t *g, window_t *wd, win_priv_t *wc) {
  return us_create(g, wd, wc);
}

emstr_us_ctx *emstr_us_create_form(geometry_t *g, window_t *wd, win_priv_t *wc, int form) {
  if (form != EMSTR_US_PLAIN_F64 && form != EMSTR_US_BATCH_F32) {
    emstr_set_error(E_ARG, -1, "unknown form %d", form);
    return NULL;
  }
  if (form == EMSTR_US_BATCH_F32 && (wc->trasc != FFSL || wc->rkstage != 1)) {
    emstr_set_error(E_UNSUP, -1, "FP32 form: FFSL with one Runge-Kutta stage only");
    return NULL;
  }
  usctx *c = us_create(g, wd, wc);
  if (c) c->f32 = (form == EMSTR_US_BATCH_F32);
  return c;
}

void emstr_us_destroy(emstr_us_ctx *c) { us_destroy(c); }

int emstr_us_upload_fields(emstr_us_ctx *c) {
  emstr_clear_error();
  return us_upload_fields(c);
}

int emstr_us_upload_tracers(emstr_us_ctx *c) {
  emstr_clear_error();
  return us_upload_tracers(c);
}

int emstr_us_download_tracers(emstr_us_ctx *c) {
  emstr_clear_error();
  return us_download_tracers(c);
}

int emstr_us_steps(emstr_us_ctx *c, int nsteps) {
  int r;
  emstr_clear_error();
  for (int k = 0; k < nsteps; k++) {
    if ((r = us_adv(c))) return r;
    if ((r = us_vdiff(c))) return r;
  }
  return 0;
}

int emstr_us_sync(emstr_us_ctx *c) { return us_sync(c); }

void *emstr_us_stream(emstr_us_ctx *c) { return c ? (void *)c->st : NULL; }
