/*
 * Host side of libemstr: device mirrors of a window (contexts), transfers
 * between the reference's sparse arrays and the dense device layout, the
 * step driver and the drop-in entry points with the reference's
 * signatures:
 *   tracer_step_3d   model/hd/tracers/tracers.c:239-321
 *   advect_diffuse   model/hd/tracers/tracers.c:401-1054
 *   vert_diffuse_3d  model/hd/tracers/tracers.c:3782-3957
 *   set_dz           model/hd/tracers/tracers.c:5682-5849
 * All arithmetic of the step runs in the HIP kernels; this file only
 * moves data, sequences launches and rebuilds the reference's index
 * lists from the per-column results of the device set_dz.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "emstr.h"
#include "emstr_dev.h"
#include "emstr_internal.h"
#include "emstr_kernels.h"
#include "emstr_batch.h"
#include "emstr_prec.h"
#include "emstr_prof.h"
#include "emstr_scratch.h"

#define HCHK(x)                                                              \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      return emstr_set_error(EMSTR_EDEVICE, -1, "%s: %s", #x,                \
                             hipGetErrorString(e_));                         \
  } while (0)

typedef long long ll;

struct emstr_ctx {
  geometry_t *g;
  window_t *wd;
  win_priv_t *wc;
  int form;
  hipStream_t st;
  emstr_dev d;
  /* host-side index helpers */
  int *h_kb;      /* [plane]   */
  int *cc2col;    /* [a2_t+1]  */
  int *col2cc;    /* [plane]   */
  int *sedg;      /* [plane]   sediment ghost sparse index */
  int *topc;      /* [plane]   surface-layer sparse index */
  int *h_ks, *h_kold, *h_knew, *h_proc;
  int *h_cols;
  int *sp_of;     /* [n3] sparse index of a dense wet cell (0 else) */
  /* device buffers */
  double *tr[3];  /* tracer buffers (FP64 plain form) */
  int cur;        /* buffer holding the current state */
  double *stage;  /* device sparse staging [sgsiz] */
  int *d_bpt, *d_bin, *d_gsed, *d_ised;
  int *d_proc;
  int *d_err;
  int *d_fferr;   /* FFSL trajectory walk bound exceeded */
  double *d_face; /* face fluxes [3][ntr][n3] of the face-once update */
  unsigned long long *d_dtm;
  double *d_part, *h_part;
  double *d_wbot;
  int anydif;
  int have_kh;   /* dzu1/dzu2/u1kh/u2kh on the device */
  int have_cols; /* per-column step data present on the device */
  /* identity of the grid the context was built for */
  int key_sgsiz, key_nx, key_ny, key_nz, key_v3;
  const int *key_w3, *key_xp1;
  struct emstr_prec *prec; /* precision form state (EMSTR_PREC_F32) */
  struct emstr_batch *batch; /* batched form state (EMSTR_BATCH_F32) */
  int use_graph;
  emstr_prof *prof;        /* per-kernel event timing                 */
  /* set_flux_3d on the device, made on first use: the dense positions
   * of the window's wet e1 / e2 faces, the surface face layer of each
   * column, the faces one cell deep, staging for dzu and the 2D terms */
  unsigned char *d_fw1, *d_fw2;
  int *h_dn;               /* [sgsiz] dense index of a wet cell (-1)  */
  int *h_ksur, *d_ksur;    /* [plane]                                 */
  ll *h_cb, *d_cb;         /* [sgsizS] dense faces one cell deep      */
  double *d_fa, *d_h, *d_m, *d_uf;
  int *d_ksur2;            /* the e2 faces' surface layers [plane]    */
  /* the offline transport driver: bottom face layer of each face
   * column, face beds, face thicknesses; knew_dev: the new surface of
   * the last device set_dz is the device's (not the window's nsur_t) */
  int *d_kbf1, *d_kbf2;
  double *d_bzu1, *d_bzu2, *d_dzu1, *d_dzu2;
  int knew_dev;
  /* auxiliary_routines, made on first use: relaxation fields and time
   * constants, the age mask (0/1), cell centres, the decay tables, the
   * per-column flag of reset_dz's cells to process */
  double *d_rtemp, *d_rsalt, *d_ttc, *d_stc, *d_agem, *d_cellz;
  int *d_tdec, *d_dtr3;     /* d_dtr3, d_dsc: second halves of d_tdec, d_dec */
  double *d_dec, *d_dsc;
  int ndec_cap;
  int *d_p2, *h_p2;
  /* open boundaries: per boundary its cells (dense), one cell in, the
   * LINEXT chain, cellz and width (device, made once), and the per-call
   * velocities, tracer settings and FILEIN values */
  int nob, b0; /* b0: the buffer of the step's starting values */
  ObcArgs *ob;
  void **obm; /* device allocations of the boundaries */
  int nobm;
  /* host staging of the sediment ghosts' dz, [plane] */
  double *h_dzg;
};

/*-------------------------------------------------------------------*/
/* Context registry: one device mirror per window (drop-in calls).   */
/*-------------------------------------------------------------------*/
#define MAXCTX 1024 /* windows with a device mirror at once */
static pthread_mutex_t reg_lock = PTHREAD_MUTEX_INITIALIZER;
static geometry_t *reg_g[MAXCTX];
static emstr_ctx *reg_c[MAXCTX];

/* A context matches a window only if the structure still describes the
 * same grid (a freed window's address can be reused by a new one). */
static int ctx_matches(const emstr_ctx *c, const geometry_t *g) {
  return c->g == g && c->key_sgsiz == g->sgsiz && c->key_nx == g->nce1 &&
         c->key_ny == g->nce2 && c->key_nz == g->nz && c->key_v3 == g->b3_t &&
         c->key_w3 == g->w3_t && c->key_xp1 == g->xp1;
}

/* Drop the drop-in context of a window (called when its arrays are freed). */
void emstr_forget_window(geometry_t *g) {
  int i;
  emstr_ctx *c = NULL;
  pthread_mutex_lock(&reg_lock);
  for (i = 0; i < MAXCTX; i++)
    if (reg_g[i] == g) {
      c = reg_c[i];
      reg_g[i] = NULL;
      reg_c[i] = NULL;
    }
  pthread_mutex_unlock(&reg_lock);
  if (c) emstr_destroy(c);
}

static emstr_ctx *ctx_for(geometry_t *g, window_t *wd, win_priv_t *wc) {
  emstr_ctx *c = NULL, *stale = NULL;
  int i;
  pthread_mutex_lock(&reg_lock);
  for (i = 0; i < MAXCTX; i++)
    if (reg_g[i] == g) {
      c = reg_c[i];
      if (!ctx_matches(c, g)) {
        stale = c;
        c = NULL;
        reg_g[i] = NULL;
        reg_c[i] = NULL;
      }
      break;
    }
  pthread_mutex_unlock(&reg_lock);
  if (stale) emstr_destroy(stale);
  pthread_mutex_lock(&reg_lock);
  if (!c) {
    c = emstr_create(g, wd, wc, EMSTR_PLAIN_F64);
    if (c) {
      for (i = 0; i < MAXCTX; i++)
        if (!reg_g[i]) {
          reg_g[i] = g;
          reg_c[i] = c;
          break;
        }
      if (i == MAXCTX) {
        pthread_mutex_unlock(&reg_lock);
        emstr_destroy(c);
        emstr_set_error(EMSTR_EARG, -1, "more than %d windows with a device mirror", MAXCTX);
        return NULL;
      }
    }
  }
  pthread_mutex_unlock(&reg_lock);
  if (c) {
    c->wd = wd;
    c->wc = wc;
  }
  return c;
}

/*-------------------------------------------------------------------*/
/* helpers                                                           */
/*-------------------------------------------------------------------*/
static void *dalloc(size_t bytes, int *ok) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
    *ok = 0;
    return NULL;
  }
  hipMemset(p, 0, bytes ? bytes : 8);
  return p;
}

static int upload2d(emstr_ctx *c, double *dst, const double *sp2) {
  /* 2D sparse array -> dense [plane]: every 2D sparse cell (wet or ghost)
   * carries the value of its (i,j) location */
  const geometry_t *g = c->g;
  const ll plane = c->d.plane;
  double *h = (double *)calloc(plane, sizeof(double));
  int n;
  for (n = 1; n <= g->sgnumS; n++) {
    ll col = (ll)g->s2j[n] * c->d.nx + g->s2i[n];
    if (col >= 0 && col < plane) h[col] = sp2[n];
  }
  /* wet columns take their own surface cell's value */
  for (n = 1; n <= g->a2_t; n++) h[c->cc2col[n]] = sp2[g->w2_t[n]];
  hipError_t e = hipMemcpyAsync(dst, h, plane * sizeof(double), hipMemcpyHostToDevice, c->st);
  hipStreamSynchronize(c->st);
  free(h);
  return e == hipSuccess ? 0 : emstr_set_error(EMSTR_EDEVICE, -1, "upload2d");
}

static int upload3d(emstr_ctx *c, double *dst, const double *sp) {
  HCHK(hipMemcpyAsync(c->stage, sp, (size_t)c->g->sgsiz * sizeof(double),
                      hipMemcpyHostToDevice, c->st));
  if (emstr_launch_scatter(c->d.wsp, c->d.wdn, c->d.nwet, c->stage, dst, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "scatter launch");
  return 0;
}

/*-------------------------------------------------------------------*/
/* context creation                                                  */
/*-------------------------------------------------------------------*/
emstr_ctx *emstr_create(geometry_t *g, window_t *wd, win_priv_t *wc, int form) {
  int n, ok = 1;
  emstr_clear_error();
  if (!g || !wd || !wc) {
    emstr_set_error(EMSTR_EARG, -1, "emstr_create: null structure");
    return NULL;
  }
  if (wc->sigma) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1, "sigma coordinates are not supported");
    return NULL;
  }
  if (g->nobc != 0 &&
      (!g->open || form != EMSTR_PLAIN_F64 || wc->trasc == FFSL)) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                    "open boundaries: the plain form with an Eulerian scheme and no "
                    "sub-stepping only");
    return NULL;
  }
  if (form != EMSTR_PLAIN_F64 && form != EMSTR_PREC_F32 && form != EMSTR_BATCH_F32) {
    emstr_set_error(EMSTR_EARG, -1, "unknown form %d", form);
    return NULL;
  }
  for (n = 1; n < g->sgsizS; n++)
    if (wc->Ds[n] != 1.0 || wc->Hn1[n] != 1.0 || wc->mdx[n] != 1.0 || wc->mdy[n] != 1.0) {
      emstr_set_error(EMSTR_EUNSUPPORTED, n, "z-level windows need Ds = Hn1 = mdx = mdy = 1");
      return NULL;
    }
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
      emstr_set_error(EMSTR_EDEVICE, -1, "no HIP device");
      return NULL;
    }
  }
  emstr_ctx *c = (emstr_ctx *)calloc(1, sizeof(emstr_ctx));
  c->g = g;
  c->wd = wd;
  c->wc = wc;
  c->form = form;
  c->key_sgsiz = g->sgsiz;
  c->key_nx = g->nce1;
  c->key_ny = g->nce2;
  c->key_nz = g->nz;
  c->key_v3 = g->b3_t;
  c->key_w3 = g->w3_t;
  c->key_xp1 = g->xp1;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    free(c);
    emstr_set_error(EMSTR_EDEVICE, -1, "stream create");
    return NULL;
  }
  emstr_dev *d = &c->d;
  d->nx = g->nce1;
  d->ny = g->nce2;
  d->nz = g->nz;
  d->ntr = wd->ntr;
  d->plane = (ll)d->nx * d->ny;
  d->n3 = d->plane * d->nz;
  d->trasc = wc->trasc;
  d->ultimate = (wc->trasc == ORDER1) ? 0 : wc->ultimate;
  d->sigma = 0;
  d->dt = wd->dt;
  d->hmin = wc->hmin;
  d->sgsiz = g->sgsiz;
  const ll plane = d->plane, n3 = d->n3;

  /* columns: bed layer, surface cell, sediment ghost */
  c->h_kb = (int *)malloc(plane * sizeof(int));
  for (ll q = 0; q < plane; q++) c->h_kb[q] = d->nz;
  c->cc2col = (int *)calloc(g->a2_t + 2, sizeof(int));
  c->col2cc = (int *)calloc(plane, sizeof(int));
  c->sedg = (int *)calloc(plane, sizeof(int));
  c->topc = (int *)calloc(plane, sizeof(int));
  for (n = 1; n <= g->a2_t; n++) {
    int cs = g->w2_t[n];
    int cb = g->bot_t[n];
    ll col = (ll)g->s2j[cs] * d->nx + g->s2i[cs];
    c->cc2col[n] = (int)col;
    c->col2cc[col] = n;
    c->h_kb[col] = g->s2k[cb];
    c->sedg[col] = g->zm1[cb];
    c->topc[col] = cs;
    int ii = g->s2i[cs], jj = g->s2j[cs];
    if (ii <= 0 || jj <= 0 || ii >= d->nx - 1 || jj >= d->ny - 1) {
      emstr_set_error(EMSTR_EUNSUPPORTED, cs, "wet column on the grid edge (%d,%d)", ii, jj);
      free(c);
      return NULL;
    }
  }
  /* wet cells: sparse <-> dense */
  d->nwet = g->b3_t; /* wet cells, open-boundary cells included */
  int *wsp = (int *)malloc((size_t)d->nwet * sizeof(int));
  ll *wdn = (ll *)malloc((size_t)d->nwet * sizeof(ll));
  c->sp_of = (int *)calloc(n3, sizeof(int));
  for (n = 0; n < d->nwet; n++) {
    int cw = g->w3_t[n + 1];
    ll idx = ((ll)g->s2k[cw] * d->ny + g->s2j[cw]) * d->nx + g->s2i[cw];
    wsp[n] = cw;
    wdn[n] = idx;
    c->sp_of[idx] = cw;
  }
  c->h_ks = (int *)malloc(plane * sizeof(int));
  c->h_kold = (int *)malloc(plane * sizeof(int));
  c->h_knew = (int *)malloc(plane * sizeof(int));
  c->h_proc = (int *)calloc(plane, sizeof(int));
  c->h_cols = (int *)malloc((g->a2_t + 1) * sizeof(int));
  for (ll q = 0; q < plane; q++) c->h_ks[q] = -1, c->h_kold[q] = c->h_knew[q] = d->nz - 1;
  for (n = 1; n <= g->a2_t; n++) {
    ll col = c->cc2col[n];
    int ns = g->nsur_t[n], os = g->sur_t[n];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    c->h_kold[col] = (os == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[os];
  }

  /* device arrays */
  d->kb = (int *)dalloc(plane * sizeof(int), &ok);
  d->ks = (int *)dalloc(plane * sizeof(int), &ok);
  d->kold = (int *)dalloc(plane * sizeof(int), &ok);
  d->knew = (int *)dalloc(plane * sizeof(int), &ok);
  double **d2[] = {&d->h1au1, &d->h2au2, &d->h1acell, &d->h2acell, &d->h2au1,
                   &d->h1au2, &d->cellarea, &d->botz, &d->eta, &d->oldeta,
                   &d->osubeta, &d->subeta};
  for (n = 0; n < (int)(sizeof(d2) / sizeof(d2[0])); n++)
    *d2[n] = (double *)dalloc(plane * sizeof(double), &ok);
  c->d_wbot = (double *)dalloc(plane * sizeof(double), &ok);
  d->dzg = (double *)dalloc(plane * sizeof(double), &ok);
  c->h_dzg = (double *)calloc(plane, sizeof(double));
  d->layers = (double *)dalloc((d->nz + 1) * sizeof(double), &ok);
  double **d3[] = {&d->u1, &d->u2, &d->w, &d->u1flux3d, &d->u2flux3d, &d->Kz, &d->dz};
  for (n = 0; n < (int)(sizeof(d3) / sizeof(d3[0])); n++)
    *d3[n] = (double *)dalloc(n3 * sizeof(double), &ok);
  c->anydif = 0;
  for (n = 0; n < wd->ntr; n++)
    if (wc->diffuse[n]) c->anydif = 1;
  if (c->anydif || wc->trasc == FFSL) {
    double **dd[] = {&d->dzu1, &d->dzu2, &d->u1kh, &d->u2kh};
    for (n = 0; n < 4; n++) *dd[n] = (double *)dalloc(n3 * sizeof(double), &ok);
    c->have_kh = 1;
  }
  const int nt = wd->ntr > 0 ? wd->ntr : 1;
  d->mintr = (double *)dalloc(nt * sizeof(double), &ok);
  d->maxtr = (double *)dalloc(nt * sizeof(double), &ok);
  d->advect = (int *)dalloc(nt * sizeof(int), &ok);
  d->diffuse = (int *)dalloc(nt * sizeof(int), &ok);
  d->in_tbdy = (int *)dalloc(nt * sizeof(int), &ok);
  d->in_tdif = (int *)dalloc(nt * sizeof(int), &ok);
  d->cols = (int *)dalloc((g->a2_t + 1) * sizeof(int), &ok);
  d->wsp = (int *)dalloc((size_t)d->nwet * sizeof(int), &ok);
  d->wdn = (ll *)dalloc((size_t)d->nwet * sizeof(ll), &ok);
  c->stage = (double *)dalloc((size_t)g->sgsiz * sizeof(double), &ok);
  c->d_bpt = (int *)dalloc((g->nbpt + 1) * sizeof(int), &ok);
  c->d_bin = (int *)dalloc((g->nbpt + 1) * sizeof(int), &ok);
  c->d_gsed = (int *)dalloc((g->ngsed + 1) * sizeof(int), &ok);
  c->d_ised = (int *)dalloc((g->ngsed + 1) * sizeof(int), &ok);
  c->d_proc = (int *)dalloc(plane * sizeof(int), &ok);
  c->d_err = (int *)dalloc(sizeof(int), &ok);
  c->d_fferr = (int *)dalloc(sizeof(int), &ok);
  c->d_dtm = (unsigned long long *)dalloc(sizeof(unsigned long long), &ok);
  {
    int nb = (g->a2_t + 255) / 256 + 1;
    c->d_part = (double *)dalloc((size_t)2 * nb * nt * sizeof(double), &ok);
    c->h_part = (double *)malloc((size_t)2 * nb * nt * sizeof(double));
  }
  if (form == EMSTR_PLAIN_F64)
    for (n = 0; n < 3; n++) c->tr[n] = (double *)dalloc((size_t)nt * n3 * sizeof(double), &ok);
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "device allocation failed");
    emstr_destroy(c);
    return NULL;
  }
  hipMemcpy(d->kb, c->h_kb, plane * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->kold, c->h_kold, plane * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->knew, c->h_knew, plane * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->wsp, wsp, (size_t)d->nwet * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->wdn, wdn, (size_t)d->nwet * sizeof(ll), hipMemcpyHostToDevice);
  hipMemcpy(c->d_bpt, g->bpt + 1, g->nbpt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(c->d_bin, g->bin + 1, g->nbpt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(c->d_gsed, g->gsed_t + 1, g->ngsed * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(c->d_ised, g->ised_t + 1, g->ngsed * sizeof(int), hipMemcpyHostToDevice);
  free(wsp);
  free(wdn);
  {
    double *lay = (double *)malloc((d->nz + 1) * sizeof(double));
    if (g->layers) {
      memcpy(lay, g->layers, (d->nz + 1) * sizeof(double));
    } else {
      emstr_set_error(EMSTR_EARG, -1, "geometry->layers is required");
      free(lay);
      emstr_destroy(c);
      return NULL;
    }
    hipMemcpy(d->layers, lay, (d->nz + 1) * sizeof(double), hipMemcpyHostToDevice);
    free(lay);
  }
  /* static 2D metrics */
  if (upload2d(c, d->h1au1, g->h1au1) || upload2d(c, d->h2au2, g->h2au2) ||
      upload2d(c, d->h1acell, g->h1acell) || upload2d(c, d->h2acell, g->h2acell) ||
      upload2d(c, d->h2au1, g->h2au1) || upload2d(c, d->h1au2, g->h1au2) ||
      upload2d(c, d->cellarea, g->cellarea) || upload2d(c, d->botz, g->botz)) {
    emstr_destroy(c);
    return NULL;
  }
  if (form == EMSTR_PREC_F32) {
    c->prec = emstr_prec_create(d, c->st);
    if (!c->prec) {
      emstr_destroy(c);
      return NULL;
    }
  }
  if (form == EMSTR_BATCH_F32) {
    c->batch = emstr_batch_create(d, c->h_kb, c->st);
    if (!c->batch) {
      emstr_destroy(c);
      return NULL;
    }
    c->use_graph = 1;
  }
  c->prof = (emstr_prof *)calloc(1, sizeof(emstr_prof));
  if (c->prec) emstr_prec_set_prof(c->prec, c->prof);
  if (c->batch) emstr_batch_set_prof(c->batch, c->prof);
  c->cur = 0;
  c->have_cols = 0;
  return c;
}

int emstr_profile(emstr_ctx *c, int on) {
  if (!c || !c->prof) return EMSTR_EARG;
  prof_collect(c->prof);
  c->prof->on = on;
  for (int k = 0; k < KC_N; k++) c->prof->ms[k] = 0.0, c->prof->cnt[k] = 0;
  return 0;
}

int emstr_kernel_times(emstr_ctx *c, double *ms, long long *count) {
  if (!c || !c->prof) return EMSTR_EARG;
  prof_collect(c->prof);
  for (int k = 0; k < KC_N; k++) {
    if (ms) ms[k] = c->prof->ms[k];
    if (count) count[k] = c->prof->cnt[k];
  }
  return 0;
}

void emstr_destroy(emstr_ctx *c) {
  int n;
  if (!c) return;
  emstr_dev *d = &c->d;
  if (c->st) hipStreamSynchronize(c->st);
  if (c->prec) emstr_prec_destroy(c->prec);
  if (c->batch) emstr_batch_destroy(c->batch);
  if (c->prof) {
    prof_collect(c->prof);
    prof_free(c->prof);
    free(c->prof);
  }
  void *dv[] = {d->kb, d->ks, d->kold, d->knew, d->h1au1, d->h2au2, d->h1acell,
                d->h2acell, d->h2au1, d->h1au2, d->cellarea, d->botz, d->eta,
                d->oldeta, d->osubeta, d->subeta, d->layers, d->u1, d->u2, d->w,
                d->u1flux3d, d->u2flux3d, d->Kz, d->dz, d->dzu1, d->dzu2, d->u1kh,
                d->u2kh, d->mintr, d->maxtr, d->advect, d->diffuse, d->in_tbdy,
                d->in_tdif, d->cols, d->wsp, d->wdn, c->stage, c->d_bpt, c->d_bin,
                c->d_gsed, c->d_ised, c->d_proc, c->d_err, c->d_dtm, c->d_part,
                c->d_wbot, c->tr[0], c->tr[1], c->tr[2], c->d_fferr, d->fnu, d->fnv,
                d->fnw, d->fw9, d->fcl, d->fcrf, d->ftmx, d->ftmy, d->ftmz, d->ftm,
                d->ffx, d->ffy, d->ffz, c->d_face, c->d_fw1, c->d_fw2, c->d_ksur,
                c->d_cb, c->d_fa, c->d_h, c->d_m, c->d_uf, c->d_ksur2, c->d_kbf1,
                c->d_kbf2, c->d_bzu1, c->d_bzu2, c->d_dzu1, c->d_dzu2, c->d_rtemp,
                c->d_rsalt, c->d_ttc, c->d_stc, c->d_agem, c->d_cellz, c->d_tdec,
                c->d_dec, c->d_p2, d->dzg};
  for (n = 0; n < (int)(sizeof(dv) / sizeof(dv[0])); n++)
    if (dv[n]) hipFree(dv[n]);
  for (n = 0; n < c->nobm; n++)
    if (c->obm[n]) hipFree(c->obm[n]);
  free(c->obm);
  free(c->ob);
  void *hv[] = {c->h_kb, c->cc2col, c->col2cc, c->sedg, c->topc, c->h_ks,
                c->h_kold, c->h_knew, c->h_proc, c->h_cols, c->sp_of, c->h_part,
                c->h_dn, c->h_ksur, c->h_cb, c->h_p2, c->h_dzg};
  for (n = 0; n < (int)(sizeof(hv) / sizeof(hv[0])); n++) free(hv[n]);
  if (c->st) hipStreamDestroy(c->st);
  pthread_mutex_lock(&reg_lock);
  for (n = 0; n < MAXCTX; n++)
    if (reg_c[n] == c) reg_c[n] = NULL, reg_g[n] = NULL;
  pthread_mutex_unlock(&reg_lock);
  free(c);
}

void *emstr_stream(emstr_ctx *c) { return c ? (void *)c->st : NULL; }

int emstr_sync(emstr_ctx *c) {
  int err = 0;
  HCHK(hipStreamSynchronize(c->st));
  HCHK(hipMemcpy(&err, c->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (err) {
    int cell = c->sp_of[err - 1];
    hipMemset(c->d_err, 0, sizeof(int));
    return emstr_set_error(EMSTR_EZEROPIVOT, cell,
                           "implicit_vdiff_tr: zero divisor at cell %d", cell);
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* per-column step data from the host's set_dz results (drop-in)     */
/*-------------------------------------------------------------------*/
/* the FP32 forms rebuild their coefficients when fields change */
static int form_fields(emstr_ctx *c) {
  if (c->prec) return emstr_prec_fields(c->prec, &c->d, c->st);
  if (c->batch) {
    /* the fused kernel advects and diffuses every tracer */
    win_priv_t *wc = c->wc;
    int n;
    if (wc->ntbdy != c->wd->ntr || wc->ntdif_v != c->wd->ntr)
      return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                             "batched form: every tracer must be advected and diffused");
    for (n = 0; n < c->wd->ntr; n++)
      if (!wc->advect[n] || wc->diffuse[n])
        return emstr_set_error(EMSTR_EUNSUPPORTED, n,
                               "batched form: tracer %d: advect on, horizontal diffusion off", n);
    return emstr_batch_fields(c->batch, &c->d, c->st);
  }
  return 0;
}

static int columns_from_host(emstr_ctx *c) {
  geometry_t *g = c->g;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  const ll plane = d->plane;
  int cc;
  for (ll q = 0; q < plane; q++) c->h_ks[q] = -1;
  /* the old and new surfaces of every wet column (reset_dz reads the
   * columns that are not processed too) */
  for (cc = 1; cc <= g->b2_t; cc++) {
    const ll col = c->cc2col[cc];
    const int ns = g->nsur_t[cc], os = g->sur_t[cc];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    c->h_kold[col] = (os == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[os];
  }
  for (cc = 1; cc <= wc->vcs; cc++) {
    int s = wc->s1[cc];
    ll col = (ll)g->s2j[s] * d->nx + g->s2i[s];
    c->h_ks[col] = g->s2k[s];
    int ns = wc->i2[cc], os = wc->i3[cc];
    c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    c->h_kold[col] = (os == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[os];
    c->h_cols[cc - 1] = (int)col;
  }
  d->ncol = wc->vcs;
  HCHK(hipMemcpyAsync(d->ks, c->h_ks, plane * sizeof(int), hipMemcpyHostToDevice, c->st));
  HCHK(hipMemcpyAsync(d->kold, c->h_kold, plane * sizeof(int), hipMemcpyHostToDevice, c->st));
  HCHK(hipMemcpyAsync(d->knew, c->h_knew, plane * sizeof(int), hipMemcpyHostToDevice, c->st));
  HCHK(hipMemcpyAsync(d->cols, c->h_cols, (size_t)d->ncol * sizeof(int), hipMemcpyHostToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  c->have_cols = 1;
  return 0;
}

static int upload_tracer_meta(emstr_ctx *c) {
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  int n, nt = wd->ntr;
  int *tb = (int *)calloc(nt + 1, sizeof(int));
  int *td = (int *)calloc(nt + 1, sizeof(int));
  for (n = 0; n < wc->ntbdy; n++) tb[wc->tbdy[n]] = 1;
  for (n = 0; n < wc->ntdif_v; n++) td[wc->tdif_v[n]] = 1;
  hipMemcpy(d->in_tbdy, tb, nt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->in_tdif, td, nt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->advect, wc->advect, nt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->diffuse, wc->diffuse, nt * sizeof(int), hipMemcpyHostToDevice);
  hipMemcpy(d->mintr, wc->mintr, nt * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(d->maxtr, wc->maxtr, nt * sizeof(double), hipMemcpyHostToDevice);
  free(tb);
  free(td);
  for (n = 0; n < nt; n++)
    if (wc->diffuse[n] && !c->anydif)
      return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                             "horizontal diffusion enabled after the context was created");
  return 0;
}

static int obc_setup(emstr_ctx *c);

int emstr_upload_fields(emstr_ctx *c) {
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  emstr_clear_error();
  d->dt = wd->dt;
  d->hmin = wc->hmin;
  d->trasc = wc->trasc;
  d->ultimate = (wc->trasc == ORDER1) ? 0 : wc->ultimate;
  if (upload3d(c, d->u1, wd->u1) || upload3d(c, d->u2, wd->u2) ||
      upload3d(c, d->w, wd->w) || upload3d(c, d->u1flux3d, wd->u1flux3d) ||
      upload3d(c, d->u2flux3d, wd->u2flux3d) || upload3d(c, d->Kz, wd->Kz) ||
      upload3d(c, d->dz, wc->dz))
    return EMSTR_EDEVICE;
  /* open-boundary columns are left out of the advective update */
  {
    const int r0 = obc_setup(c);
    if (r0) return r0;
  }
  /* the sediment ghosts' dz (reset_dz writes them, set_dz copies them up
   * a column whose old surface is the ghost) */
  for (int cc = 1; cc <= c->g->a2_t; cc++) {
    const ll col = c->cc2col[cc];
    c->h_dzg[col] = wc->dz[c->sedg[col]];
  }
  HCHK(hipMemcpyAsync(d->dzg, c->h_dzg, (size_t)d->plane * sizeof(double),
                      hipMemcpyHostToDevice, c->st));
  d->nstep0 = (wd->nstep == 0);
  if (d->trasc == FFSL && !c->have_kh)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "FFSL selected after the context was created");
  if (c->have_kh)
    if (upload3d(c, d->dzu1, wd->dzu1) || upload3d(c, d->dzu2, wd->dzu2) ||
        upload3d(c, d->u1kh, wc->u1kh) || upload3d(c, d->u2kh, wc->u2kh))
      return EMSTR_EDEVICE;
  if (upload2d(c, d->eta, wd->eta) || upload2d(c, d->oldeta, wc->oldeta) ||
      upload2d(c, c->d_wbot, wd->wbot))
    return EMSTR_EDEVICE;
  if (upload_tracer_meta(c)) return EMSTR_EUNSUPPORTED;
  if (columns_from_host(c)) return EMSTR_EDEVICE;
  HCHK(hipStreamSynchronize(c->st));
  return form_fields(c);
}

int emstr_upload_tracers(emstr_ctx *c) {
  int n;
  emstr_dev *d = &c->d;
  emstr_clear_error();
  for (n = 0; n < c->wd->ntr; n++) {
    HCHK(hipMemcpyAsync(c->stage, c->wd->tr_wc[n], (size_t)c->g->sgsiz * sizeof(double),
                        hipMemcpyHostToDevice, c->st));
    if (c->batch) {
      if (emstr_batch_put(c->batch, d, n, c->stage, c->st)) return EMSTR_EDEVICE;
    } else if (c->prec) {
      if (emstr_prec_put_tracer(c->prec, d, n, c->stage, c->st)) return EMSTR_EDEVICE;
    } else if (emstr_launch_scatter(d->wsp, d->wdn, d->nwet, c->stage,
                                    c->tr[c->cur] + (ll)n * d->n3, c->st)) {
      return emstr_set_error(EMSTR_EDEVICE, -1, "scatter");
    }
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

int emstr_download_tracers(emstr_ctx *c) {
  int n;
  emstr_dev *d = &c->d;
  geometry_t *g = c->g;
  emstr_clear_error();
  for (n = 0; n < c->wd->ntr; n++) {
    HCHK(hipMemcpyAsync(c->stage, c->wd->tr_wc[n], (size_t)g->sgsiz * sizeof(double),
                        hipMemcpyHostToDevice, c->st));
    if (c->batch) {
      if (emstr_batch_get(c->batch, d, n, c->stage, c->st)) return EMSTR_EDEVICE;
    } else if (c->prec) {
      if (emstr_prec_get_tracer(c->prec, d, n, c->stage, c->st)) return EMSTR_EDEVICE;
    } else if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, c->tr[c->cur] + (ll)n * d->n3,
                                   c->stage, c->st)) {
      return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
    }
    /* ghosts hold the no-gradient values the next step would set */
    emstr_launch_ghosts(c->d_bpt, c->d_bin, g->nbpt, c->stage, c->st);
    emstr_launch_ghosts(c->d_gsed, c->d_ised, g->ngsed, c->stage, c->st);
    HCHK(hipMemcpyAsync(c->wd->tr_wc[n], c->stage, (size_t)g->sgsiz * sizeof(double),
                        hipMemcpyDeviceToHost, c->st));
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

/*-------------------------------------------------------------------*/
/* the step                                                          */
/*-------------------------------------------------------------------*/
/* FFSL work arrays (trajectories, transverse tracers, fluxes), made on
 * first use: 184 bytes per dense cell */
static int ffsl_alloc(emstr_ctx *c) {
  emstr_dev *d = &c->d;
  const size_t n3 = (size_t)d->n3;
  int ok = 1, n;
  if (d->fnu) return 0;
  double **a[] = {&d->fnu, &d->fnv, &d->fnw, &d->fw9, &d->ftmx, &d->ftmy,
                  &d->ftmz, &d->ftm, &d->ffx, &d->ffy, &d->ffz};
  for (n = 0; n < (int)(sizeof(a) / sizeof(a[0])); n++)
    *a[n] = (double *)dalloc(n3 * sizeof(double), &ok);
  d->fcl = (ll *)dalloc(6 * n3 * sizeof(ll), &ok);
  d->fcrf = (double *)dalloc(6 * n3 * sizeof(double), &ok);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "FFSL work arrays: allocation failed");
  return 0;
}

static int adv_plain(emstr_ctx *c) {
  emstr_dev *d = &c->d;
  win_priv_t *wc = c->wc;
  double dtu = d->dt, dtm = d->dt, trem;
  int itermax = 20, first = 1, n;
  const int ffsl = (d->trasc == FFSL);
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (ffsl) {
    /* tracers.c:524-559 */
    itermax = 200;
    if (ffsl_alloc(c)) return EMSTR_EDEVICE;
  }
  /* face fluxes of the face-once update, made on first use */
  const size_t fn3 = (size_t)(d->ntr > 0 ? d->ntr : 1) * (size_t)d->n3;
  if (!ffsl && !c->d_face) {
    if (hipMalloc((void **)&c->d_face, 3 * fn3 * sizeof(double)) != hipSuccess) {
      c->d_face = NULL;
      return emstr_set_error(EMSTR_EDEVICE, -1, "face flux buffers (%zu bytes)",
                             3 * fn3 * sizeof(double));
    }
  }
  if (wc->stab & (SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER)) {
    unsigned long long init;
    double big = dtu;
    memcpy(&init, &big, sizeof(init));
    HCHK(hipMemcpyAsync(c->d_dtm, &init, sizeof(init), hipMemcpyHostToDevice, c->st));
    if (emstr_launch_dtm(d, dtu, c->d_dtm, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "dtm launch");
    unsigned long long r;
    HCHK(hipMemcpyAsync(&r, c->d_dtm, sizeof(r), hipMemcpyDeviceToHost, c->st));
    HCHK(hipStreamSynchronize(c->st));
    memcpy(&dtm, &r, sizeof(dtm));
    if (dtm != dtu) dtm = floor(dtm);
  }
  if (dtm == 0.0 || (int)(dtu / dtm) > itermax)
    return emstr_set_error(EMSTR_ESUBSTEP, -1,
                           "tracer_step_3d: maximum number of sub-steps (%d) exceeded", itermax);
  /* buffers: start = cur (kept for lateral ghosts), then rotate */
  const int b0 = c->cur;
  c->b0 = b0;
  int in = b0, out = (b0 + 1) % 3;
  trem = dtu;
  while (trem > 0) {
    double u = dtm;
    if (trem < u) u = trem;
    wc->b1 = u;
    if (u > 0.0) {
      int pi, bad;
      if (emstr_launch_subeta(d, u, trem, first, c->st))
        return emstr_set_error(EMSTR_EDEVICE, -1, "subeta launch");
      pi = prof_begin(c->prof, KC_ADV, c->st);
      if (ffsl) {
        /* tracers.c:763-765, then the fluxes and update of each tracer */
        bad = emstr_launch_ffsl_traj(d, u, c->d_fferr, c->st);
        for (n = 0; n < d->ntr && !bad; n++) {
          bad = emstr_launch_ffsl_fluxes(d, c->tr[in] + (ll)n * d->n3,
                                         c->tr[b0] + (ll)n * d->n3, u, c->d_fferr, c->st);
          if (!bad)
            bad = emstr_launch_adv(d, c->tr[in], c->tr[out], c->tr[b0], u, trem, c->anydif,
                                   n, n + 1, c->st);
        }
      } else {
        bad = emstr_launch_adv_faces(d, c->tr[in], c->tr[out], c->tr[b0], u, c->anydif,
                                     c->d_face, c->d_face + fn3, c->d_face + 2 * fn3, c->st);
      }
      prof_end(c->prof, pi, c->st);
      if (bad || emstr_launch_osub(d, c->st))
        return emstr_set_error(EMSTR_EDEVICE, -1, "advection launch");
      in = out;
      out = (out + 1) % 3;
      if (out == b0) out = (out + 1) % 3;
    }
    trem -= u;
    first = 0;
  }
  c->cur = in;
  if (ffsl) {
    int e = 0;
    HCHK(hipMemcpyAsync(&e, c->d_fferr, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HCHK(hipStreamSynchronize(c->st));
    if (e) {
      HCHK(hipMemsetAsync(c->d_fferr, 0, sizeof(int), c->st));
      return emstr_set_error(EMSTR_ESUBSTEP, -1,
                             "ff_sl: a trajectory crosses more than the walk bound of cells");
    }
  }
  return 0;
}

static int vdiff_plain(emstr_ctx *c) {
  const int scratch = (c->cur + 1) % 3;
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  int pi = prof_begin(c->prof, KC_VDIFF, c->st);
  int bad = emstr_launch_vdiff(&c->d, c->tr[c->cur], c->tr[scratch], c->d_err, c->st);
  prof_end(c->prof, pi, c->st);
  if (bad) return emstr_set_error(EMSTR_EDEVICE, -1, "vdiff launch");
  return 0;
}

int emstr_advect_diffuse(emstr_ctx *c) {
  emstr_clear_error();
  if (c->batch)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "batched form: advection and diffusion are fused");
  if (c->prec) return emstr_prec_advect(c->prec, &c->d, c->wc->stab, c->st);
  return adv_plain(c);
}

int emstr_vert_diffuse(emstr_ctx *c) {
  emstr_clear_error();
  if (c->batch)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "batched form: advection and diffusion are fused");
  if (c->prec) return emstr_prec_vdiff(c->prec, &c->d, c->d_err, c->st);
  return vdiff_plain(c);
}

int emstr_steps(emstr_ctx *c, int nsteps) {
  int r, k;
  emstr_clear_error();
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (c->batch)
    return emstr_batch_steps(c->batch, &c->d, c->wc->stab, nsteps, c->st, c->use_graph);
  for (k = 0; k < nsteps; k++)
    if ((r = emstr_step(c))) return r;
  return 0;
}

/*-------------------------------------------------------------------*/
/* Ensembles: one hipGraph with a branch per batched context          */
/*-------------------------------------------------------------------*/
struct emstr_ensemble {
  int n;
  emstr_ctx **c;
  hipStream_t s0;
  hipStream_t *br;
  hipEvent_t fork, *join;
  hipGraph_t graph;
  hipGraphExec_t gexec;
};

void emstr_ensemble_destroy(emstr_ensemble *e) {
  if (!e) return;
  if (e->s0) hipStreamSynchronize(e->s0);
  if (e->gexec) hipGraphExecDestroy(e->gexec);
  if (e->graph) hipGraphDestroy(e->graph);
  for (int i = 0; i < e->n; i++) {
    if (e->br && e->br[i]) hipStreamDestroy(e->br[i]);
    if (e->join && e->join[i]) hipEventDestroy(e->join[i]);
  }
  if (e->fork) hipEventDestroy(e->fork);
  if (e->s0) hipStreamDestroy(e->s0);
  free(e->br);
  free(e->join);
  free(e->c);
  free(e);
}

emstr_ensemble *emstr_ensemble_create(emstr_ctx **ctxs, int n) {
  int i;
  emstr_clear_error();
  if (!ctxs || n < 1) {
    emstr_set_error(EMSTR_EARG, -1, "ensemble: no contexts");
    return NULL;
  }
  for (i = 0; i < n; i++) {
    emstr_ctx *c = ctxs[i];
    if (!c || !c->batch || !c->have_cols) {
      emstr_set_error(EMSTR_EARG, i, "ensemble: context %d is not a batched context with "
                                     "uploaded fields", i);
      return NULL;
    }
    if (hipStreamSynchronize(c->st) != hipSuccess) {
      emstr_set_error(EMSTR_EDEVICE, i, "ensemble: context %d stream", i);
      return NULL;
    }
  }
  emstr_ensemble *e = (emstr_ensemble *)calloc(1, sizeof(emstr_ensemble));
  e->n = n;
  e->c = (emstr_ctx **)malloc((size_t)n * sizeof(emstr_ctx *));
  memcpy(e->c, ctxs, (size_t)n * sizeof(emstr_ctx *));
  e->br = (hipStream_t *)calloc(n, sizeof(hipStream_t));
  e->join = (hipEvent_t *)calloc(n, sizeof(hipEvent_t));
  int ok = hipStreamCreateWithFlags(&e->s0, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
  for (i = 0; i < n && ok; i++)
    ok = hipStreamCreateWithFlags(&e->br[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "ensemble: streams");
    emstr_ensemble_destroy(e);
    return NULL;
  }
  /* fork: every branch waits on the capture stream, records its
   * window's two steps, and the capture stream joins them all */
  if (hipStreamBeginCapture(e->s0, hipStreamCaptureModeThreadLocal) != hipSuccess ||
      hipEventRecord(e->fork, e->s0) != hipSuccess) {
    emstr_set_error(EMSTR_EDEVICE, -1, "ensemble: capture");
    emstr_ensemble_destroy(e);
    return NULL;
  }
  int rec = 0;
  for (i = 0; i < n && !rec; i++) {
    emstr_ctx *c = e->c[i];
    if (hipStreamWaitEvent(e->br[i], e->fork, 0) != hipSuccess) rec = EMSTR_EDEVICE;
    if (!rec) rec = emstr_batch_record2(c->batch, &c->d, c->wc->stab, e->br[i]);
    if (!rec && (hipEventRecord(e->join[i], e->br[i]) != hipSuccess ||
                 hipStreamWaitEvent(e->s0, e->join[i], 0) != hipSuccess))
      rec = EMSTR_EDEVICE;
  }
  hipGraph_t g = NULL;
  hipError_t ec = hipStreamEndCapture(e->s0, &g);
  e->graph = g;
  if (rec || ec != hipSuccess || !g ||
      hipGraphInstantiate(&e->gexec, e->graph, NULL, NULL, 0) != hipSuccess) {
    if (!rec) emstr_set_error(EMSTR_EDEVICE, -1, "ensemble: graph instantiate");
    emstr_ensemble_destroy(e);
    return NULL;
  }
  return e;
}

int emstr_ensemble_steps(emstr_ensemble *e, int nsteps) {
  emstr_clear_error();
  if (!e) return emstr_set_error(EMSTR_EARG, -1, "ensemble: null");
  if (nsteps < 0 || (nsteps & 1))
    return emstr_set_error(EMSTR_EARG, -1, "ensemble: nsteps must be even");
  /* work the contexts queued on their own streams comes first */
  for (int i = 0; i < e->n; i++) HCHK(hipStreamSynchronize(e->c[i]->st));
  for (int k = 0; k < nsteps; k += 2)
    if (hipGraphLaunch(e->gexec, e->s0) != hipSuccess)
      return emstr_set_error(EMSTR_EDEVICE, -1, "ensemble: graph launch");
  return 0;
}

int emstr_ensemble_sync(emstr_ensemble *e) {
  if (!e) return EMSTR_EARG;
  HCHK(hipStreamSynchronize(e->s0));
  return 0;
}

void *emstr_ensemble_stream(emstr_ensemble *e) { return e ? (void *)e->s0 : NULL; }

int emstr_launch_shape(emstr_ctx *c, int *tile_x, int *tile_y, int *grid, int *grid_full) {
  if (!c || !c->batch || !tile_x || !tile_y || !grid || !grid_full) return EMSTR_EARG;
  emstr_batch_shape(c->batch, tile_x, tile_y, grid, grid_full);
  return 0;
}

int emstr_tune(emstr_ctx *c, int tile_x, int tile_y, int grid, int use_graph) {
  if (!c) return EMSTR_EARG;
  if (use_graph >= 0) c->use_graph = use_graph;
  if (c->batch) return emstr_batch_tune(c->batch, c->h_kb, tile_x, tile_y, grid);
  return 0;
}

int emstr_step(emstr_ctx *c) {
  int r;
  emstr_clear_error();
  if (c->batch) return emstr_steps(c, 1);
  if (c->prec) return emstr_prec_step(c->prec, &c->d, c->wc->stab, c->d_err, c->st);
  if ((r = adv_plain(c))) return r;
  return vdiff_plain(c);
}

int emstr_tracer_mass(emstr_ctx *c, double *mass) {
  emstr_dev *d = &c->d;
  int n, b;
  const int nb = emstr_mass_blocks(d);
  emstr_clear_error();
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (c->batch) {
    if (emstr_batch_mass(c->batch, d, c->d_part, c->st)) return EMSTR_EDEVICE;
  } else if (c->prec) {
    if (emstr_prec_mass(c->prec, d, c->d_part, c->st)) return EMSTR_EDEVICE;
  } else if (emstr_launch_mass(d, c->tr[c->cur], NULL, c->d_part, c->st)) {
    return emstr_set_error(EMSTR_EDEVICE, -1, "mass launch");
  }
  HCHK(hipMemcpyAsync(c->h_part, c->d_part, (size_t)2 * nb * d->ntr * sizeof(double),
                      hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (n = 0; n < d->ntr; n++) {
    /* compensated sum of the block (sum, err) pairs in block order */
    double s = 0.0, e = 0.0;
    for (b = 0; b < nb; b++) {
      double a = c->h_part[2 * ((ll)n * nb + b)];
      double t = s + a;
      double bb = t - s;
      e += (s - (t - bb)) + (a - bb);
      s = t;
      e += c->h_part[2 * ((ll)n * nb + b) + 1];
    }
    mass[n] = s + e;
  }
  return 0;
}

/* emstr_tracer_mass without any allocation: the partials and the totals
 * in the library's fixed device scratch (emstr_scratch.hip), the final
 * reduction on the device in the host reduction's order. */
int emstr_tracer_mass_noalloc(emstr_ctx *c, double *mass) {
  emstr_dev *d = &c->d;
  const long long nb = emstr_mass_blocks(d);
  const long long need = 2 * nb * d->ntr + d->ntr;
  emstr_clear_error();
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (need > EMSTR_SCRATCH_DOUBLES)
    return emstr_set_error(EMSTR_EARG, -1,
                           "mass without allocation: %lld partials for %d tracers exceed the "
                           "%d doubles of scratch (2 * ceil(columns / 256) * ntr + ntr)",
                           need, d->ntr, EMSTR_SCRATCH_DOUBLES);
  double *s = emstr_scratch_ptr();
  if (!s) return EMSTR_EDEVICE;
  double *tot = s + 2 * nb * d->ntr;
  if (c->batch) {
    if (emstr_batch_mass(c->batch, d, s, c->st)) return EMSTR_EDEVICE;
  } else if (c->prec) {
    if (emstr_prec_mass(c->prec, d, s, c->st)) return EMSTR_EDEVICE;
  } else if (emstr_launch_mass(d, c->tr[c->cur], NULL, s, c->st)) {
    return emstr_set_error(EMSTR_EDEVICE, -1, "mass launch");
  }
  if (emstr_launch_mass_reduce(s, (int)nb, d->ntr, tot, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "mass reduction launch");
  HCHK(hipMemcpyAsync(mass, tot, (size_t)d->ntr * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

/*-------------------------------------------------------------------*/
/* device set_dz and w, and the reference's index lists              */
/*-------------------------------------------------------------------*/
/* set_dz on the device; knew_dev: the previous new surface is the one
 * the device's last set_dz left, else the window's nsur_t */
static int setdz_run(emstr_ctx *c, int knew_dev) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  const ll plane = d->plane;
  int cc;
  d->hmin = wc->hmin;
  if (upload2d(c, d->eta, wd->eta) || upload2d(c, d->oldeta, wc->oldeta))
    return EMSTR_EDEVICE;
  if (!knew_dev) {
    for (cc = 1; cc <= g->a2_t; cc++) {
      ll col = c->cc2col[cc];
      int ns = g->nsur_t[cc];
      c->h_knew[col] = (ns == c->sedg[col]) ? c->h_kb[col] - 1 : g->s2k[ns];
    }
    HCHK(hipMemcpyAsync(d->knew, c->h_knew, plane * sizeof(int), hipMemcpyHostToDevice, c->st));
  }
  if (emstr_launch_setdz(d, c->d_proc, c->st)) return emstr_set_error(EMSTR_EDEVICE, -1, "setdz");
  HCHK(hipMemcpyAsync(c->h_ks, d->ks, plane * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipMemcpyAsync(c->h_kold, d->kold, plane * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipMemcpyAsync(c->h_knew, d->knew, plane * sizeof(int), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  /* processed columns in w2_t order */
  int ncol = 0;
  for (cc = 1; cc <= g->b2_t; cc++) {
    ll col = c->cc2col[cc];
    if (c->h_ks[col] >= 0) c->h_cols[ncol++] = (int)col;
  }
  d->ncol = ncol;
  HCHK(hipMemcpyAsync(d->cols, c->h_cols, (size_t)ncol * sizeof(int), hipMemcpyHostToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  c->have_cols = 1;
  return 0;
}

int emstr_device_set_dz(emstr_ctx *c) {
  int r;
  emstr_clear_error();
  if ((r = setdz_run(c, 0))) return r;
  c->knew_dev = 0;
  return form_fields(c);
}

/* Sparse index of layer k of column col; kb-1 is the sediment ghost. */
static int sp_at(emstr_ctx *c, ll col, int k) {
  if (k < c->h_kb[col]) return c->sedg[col];
  return c->sp_of[(ll)k * c->d.plane + col];
}

/* Rebuild sur_t/nsur_t, s1, i1..i3, vc.. and dz on the host structures
 * from the device set_dz (same order as tracers.c:5762-5816). */
static int lists_to_host(emstr_ctx *c) {
  geometry_t *g = c->g;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  int cc, vc = 1;
  for (cc = 1; cc <= g->a2_t; cc++) {
    ll col = c->cc2col[cc];
    g->sur_t[cc] = sp_at(c, col, c->h_kold[col]);
    g->nsur_t[cc] = sp_at(c, col, c->h_knew[col]);
  }
  for (cc = 1; cc <= g->b2_t; cc++) {
    ll col = c->cc2col[cc];
    if (c->h_ks[col] < 0) continue;
    wc->s1[vc] = sp_at(c, col, c->h_ks[col]);
    wc->i1[vc] = g->bot_t[cc];
    wc->i2[vc] = g->nsur_t[cc];
    wc->i3[vc] = g->sur_t[cc];
    if (cc == g->v2_t) wc->vcs1 = vc;
    vc++;
  }
  const int vcs = vc - 1;
  wc->vcs = vcs;
  for (cc = 1; cc <= vcs; cc++) {
    ll col = c->h_cols[cc - 1];
    for (int k = c->h_ks[col] - 1; k >= c->h_kb[col]; k--) wc->s1[vc++] = sp_at(c, col, k);
    if (cc == wc->vcs1) wc->vc1 = vc - 1;
  }
  wc->vc = vc - 1;
  /* dz of wet cells */
  if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, d->dz, c->stage, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
  double *tmp = (double *)malloc((size_t)g->sgsiz * sizeof(double));
  HCHK(hipMemcpyAsync(tmp, c->stage, (size_t)g->sgsiz * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (int q = 1; q <= g->b3_t; q++) {
    int cw = g->w3_t[q];
    wc->dz[cw] = tmp[cw];
  }
  free(tmp);
  return 0;
}

/* w from continuity on the device for the processed columns; the
 * host's windat->w receives it. */

/* set_flux_3d (model/hd/momentum/vel3d.c:4041-4191, VEL3D mode) on
 * the device: the face volume fluxes from the window's face velocities
 * and thicknesses (or, under FFSL, its mean volume fluxes), into the
 * context's fields and windat->u1flux3d / u2flux3d. */
static int flux3d_setup(emstr_ctx *c) {
  geometry_t *g = c->g;
  emstr_dev *d = &c->d;
  const ll n3 = d->n3, plane = d->plane;
  int ok = 1, cc, dir;
  c->h_dn = (int *)malloc((size_t)g->sgsiz * sizeof(int));
  c->h_ksur = (int *)malloc((size_t)plane * sizeof(int));
  c->h_cb = (ll *)malloc((size_t)(g->sgsiz + 1) * sizeof(ll));
  unsigned char *fw = (unsigned char *)malloc((size_t)n3);
  if (!c->h_dn || !c->h_ksur || !c->h_cb || !fw) {
    free(fw);
    return emstr_set_error(EMSTR_EDEVICE, -1, "set_flux_3d: host allocation");
  }
  for (ll i = 0; i < g->sgsiz; i++) c->h_dn[i] = -1;
  for (ll i = 0; i < n3; i++)
    if (c->sp_of[i] > 0) c->h_dn[c->sp_of[i]] = (int)i;
  c->d_fw1 = (unsigned char *)dalloc((size_t)n3, &ok);
  c->d_fw2 = (unsigned char *)dalloc((size_t)n3, &ok);
  c->d_ksur = (int *)dalloc((size_t)plane * sizeof(int), &ok);
  c->d_ksur2 = (int *)dalloc((size_t)plane * sizeof(int), &ok);
  c->d_cb = (ll *)dalloc((size_t)(g->sgsiz + 1) * sizeof(ll), &ok);
  c->d_fa = (double *)dalloc((size_t)n3 * sizeof(double), &ok);
  c->d_h = (double *)dalloc((size_t)plane * sizeof(double), &ok);
  c->d_m = (double *)dalloc((size_t)plane * sizeof(double), &ok);
  c->d_uf = (double *)dalloc((size_t)plane * sizeof(double), &ok);
  for (dir = 1; dir <= 2 && ok; dir++) {
    const int *w3 = (dir == 1) ? g->w3_e1 : g->w3_e2;
    const int nb = (dir == 1) ? g->b3_e1 : g->b3_e2;
    memset(fw, 0, (size_t)n3);
    for (cc = 1; cc <= nb; cc++) {
      const int f = w3[cc];
      if (f < 1 || f >= g->sgsiz || c->h_dn[f] < 0) {
        free(fw);
        return emstr_set_error(EMSTR_EUNSUPPORTED, f,
                               "set_flux_3d: e%d face %d is not the face of a wet cell", dir, f);
      }
      fw[c->h_dn[f]] = 1;
    }
    if (hipMemcpy(dir == 1 ? c->d_fw1 : c->d_fw2, fw, (size_t)n3, hipMemcpyHostToDevice) !=
        hipSuccess)
      ok = 0;
  }
  free(fw);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "set_flux_3d: device allocation");
  return 0;
}

/* One face direction of set_flux_3d.  ksur_dev: the surface face layer
 * of each column is already on the device (the transport driver's
 * set_dz_at_u), else it comes from the window's sur_e1 / sur_e2; dzu:
 * the device face thicknesses, else the window's are uploaded; to_host:
 * copy the fluxes into windat->u1flux3d / u2flux3d. */
static int flux3d_dir(emstr_ctx *c, int dir, int ksur_dev, const double *dzu, int to_host) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  const int ffsl = (wc->trasc == FFSL);
  const int *w3 = (dir == 1) ? g->w3_e1 : g->w3_e2;
  const int nb = (dir == 1) ? g->b3_e1 : g->b3_e2, nn = (dir == 1) ? g->n3_e1 : g->n3_e2;
  double *u = (dir == 1) ? wd->u1 : wd->u2;
  double *uvm = (dir == 1) ? wd->u1vm : wd->u2vm;
  double *out = (dir == 1) ? d->u1flux3d : d->u2flux3d;
  double *hout = (dir == 1) ? wd->u1flux3d : wd->u2flux3d;
  int *ks = (dir == 1) ? c->d_ksur : c->d_ksur2;
  int cc, ncb = 0;
  if (ffsl) {
    /* the velocities past b3 are zeroed on the window (vel3d.c:4069-4078) */
    for (cc = nb + 1; cc <= nn; cc++) {
      u[w3[cc]] = 0.0;
      uvm[w3[cc]] = 0.0;
    }
    if (upload3d(c, c->d_fa, uvm)) return EMSTR_EDEVICE;
  } else {
    const int *cbl = (dir == 1) ? wc->cbot_e1 : wc->cbot_e2;
    const int ncbl = (dir == 1) ? wc->ncbot_e1 : wc->ncbot_e2;
    if (!ksur_dev) {
      const int *w2 = (dir == 1) ? g->w2_e1 : g->w2_e2;
      const int *sur = (dir == 1) ? g->sur_e1 : g->sur_e2;
      const int nb2 = (dir == 1) ? g->b2_e1 : g->b2_e2;
      for (ll q = 0; q < d->plane; q++) c->h_ksur[q] = d->nz;
      for (cc = 1; cc <= nb2; cc++) {
        const int cs = w2[cc], f = sur[cc];
        if (f < 1 || f >= g->sgsiz)
          return emstr_set_error(EMSTR_EARG, cc, "set_flux_3d: surface face of column %d", cc);
        c->h_ksur[(ll)g->s2j[cs] * d->nx + g->s2i[cs]] = g->s2k[f];
      }
      HCHK(hipMemcpyAsync(ks, c->h_ksur, (size_t)d->plane * sizeof(int), hipMemcpyHostToDevice,
                          c->st));
    }
    for (cc = 1; cc <= ncbl; cc++) {
      const int f = cbl[cc];
      if (f < 1 || f >= g->sgsiz || c->h_dn[f] < 0)
        return emstr_set_error(EMSTR_EUNSUPPORTED, f,
                               "set_flux_3d: face %d one cell deep is not a wet cell's", f);
      c->h_cb[ncb++] = c->h_dn[f];
    }
    if (ncb)
      HCHK(hipMemcpyAsync(c->d_cb, c->h_cb, (size_t)ncb * sizeof(ll), hipMemcpyHostToDevice,
                          c->st));
    if (upload3d(c, (dir == 1) ? d->u1 : d->u2, u) ||
        (!dzu && upload3d(c, c->d_fa, (dir == 1) ? wd->dzu1 : wd->dzu2)) ||
        upload2d(c, c->d_h, (dir == 1) ? g->h2au1 : g->h1au2) ||
        upload2d(c, c->d_m, (dir == 1) ? wc->mdx : wc->mdy))
      return EMSTR_EDEVICE;
    if (ncb && upload2d(c, c->d_uf, (dir == 1) ? wd->u1flux : wd->u2flux))
      return EMSTR_EDEVICE;
  }
  if (emstr_launch_flux3d(d, (dir == 1) ? c->d_fw1 : c->d_fw2, ks, (dir == 1) ? d->u1 : d->u2,
                          dzu ? dzu : c->d_fa, c->d_h, c->d_m, ffsl ? c->d_fa : NULL, ncb,
                          c->d_cb, c->d_uf, wd->dt, out, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "set_flux_3d launch");
  if (to_host) {
    /* the window's array: zero, then the wet cells' faces */
    HCHK(hipMemsetAsync(c->stage, 0, (size_t)g->sgsiz * sizeof(double), c->st));
    if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, out, c->stage, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
    HCHK(hipMemcpyAsync(hout, c->stage, (size_t)g->sgsiz * sizeof(double),
                        hipMemcpyDeviceToHost, c->st));
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int flux3d_check(emstr_ctx *c, int need_sur) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  const int ffsl = (wc->trasc == FFSL);
  if (ffsl ? (!wd->u1vm || !wd->u2vm)
           : ((need_sur && (!g->sur_e1 || !g->sur_e2 || !wd->dzu1 || !wd->dzu2)) ||
              ((wc->ncbot_e1 > 0 || wc->ncbot_e2 > 0) &&
               (!wd->u1flux || !wd->u2flux || !wc->cbot_e1 || !wc->cbot_e2))))
    return emstr_set_error(EMSTR_EARG, -1,
                           "set_flux_3d needs sur_e1/e2, dzu1/2 and, for faces one cell deep, "
                           "u1flux/u2flux and cbot_e1/e2 (FFSL: u1vm/u2vm)");
  if (!c->d_fw1) return flux3d_setup(c);
  return 0;
}

int emstr_device_set_flux_3d(emstr_ctx *c) {
  int r;
  emstr_clear_error();
  if ((r = flux3d_check(c, 1))) return r;
  if ((r = flux3d_dir(c, 1, 0, NULL, 1)) || (r = flux3d_dir(c, 2, 0, NULL, 1))) return r;
  return form_fields(c);
}

/* w from continuity of the device's fluxes (no upload of the fluxes). */
static int velw_run(emstr_ctx *c, int to_host) {
  emstr_dev *d = &c->d;
  geometry_t *g = c->g;
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (upload2d(c, c->d_wbot, c->wd->wbot)) return EMSTR_EDEVICE;
  HCHK(hipMemsetAsync(d->w, 0, (size_t)d->n3 * sizeof(double), c->st));
  if (emstr_launch_velw(d, c->d_wbot, c->st)) return emstr_set_error(EMSTR_EDEVICE, -1, "velw");
  if (to_host) {
    HCHK(hipMemsetAsync(c->stage, 0, (size_t)g->sgsiz * sizeof(double), c->st));
    if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, d->w, c->stage, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
    HCHK(hipMemcpyAsync(c->wd->w, c->stage, (size_t)g->sgsiz * sizeof(double),
                        hipMemcpyDeviceToHost, c->st));
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

int emstr_device_vel_w(emstr_ctx *c) {
  int r;
  emstr_clear_error();
  if (!c->have_cols) return emstr_set_error(EMSTR_EARG, -1, "no set_dz data on the device");
  if (upload3d(c, c->d.u1flux3d, c->wd->u1flux3d) ||
      upload3d(c, c->d.u2flux3d, c->wd->u2flux3d))
    return EMSTR_EDEVICE;
  if ((r = velw_run(c, 1))) return r;
  return form_fields(c);
}

/* The offline transport driver's per-window tables, made on first use:
 * the bottom face layer of each e1 / e2 face column, the face beds. */
static int trans_setup(emstr_ctx *c) {
  geometry_t *g = c->g;
  emstr_dev *d = &c->d;
  const ll plane = d->plane;
  int ok = 1, cc, dir;
  int *kbf = (int *)malloc((size_t)plane * sizeof(int));
  if (!kbf) return emstr_set_error(EMSTR_EDEVICE, -1, "transport: host allocation");
  for (dir = 1; dir <= 2; dir++) {
    const int *w2 = (dir == 1) ? g->w2_e1 : g->w2_e2;
    const int *bot = (dir == 1) ? g->bot_e1 : g->bot_e2;
    const double *bz = (dir == 1) ? g->botzu1 : g->botzu2;
    const int nb2 = (dir == 1) ? g->b2_e1 : g->b2_e2;
    for (ll q = 0; q < plane; q++) kbf[q] = -1;
    for (cc = 1; cc <= nb2; cc++) {
      const int cs = w2[cc], cb = bot[cc];
      const ll col = (ll)g->s2j[cs] * d->nx + g->s2i[cs];
      const int k = (cb >= 1 && cb < g->sgsiz) ? g->s2k[cb] : -1;
      if (k < c->h_kb[col] || k >= d->nz || (ll)g->s2j[cb] * d->nx + g->s2i[cb] != col ||
          bz[cs] < g->layers[k] || bz[cs] > g->layers[k + 1]) {
        free(kbf);
        return emstr_set_error(EMSTR_EUNSUPPORTED, cs,
                               "transport: e%d face column %d: bottom face %d is not a wet cell "
                               "of the column holding the face bed",
                               dir, cs, cb);
      }
      kbf[col] = k;
    }
    int **dk = (dir == 1) ? &c->d_kbf1 : &c->d_kbf2;
    double **db = (dir == 1) ? &c->d_bzu1 : &c->d_bzu2;
    double **dz = (dir == 1) ? &c->d_dzu1 : &c->d_dzu2;
    *dk = (int *)dalloc((size_t)plane * sizeof(int), &ok);
    *db = (double *)dalloc((size_t)plane * sizeof(double), &ok);
    *dz = (double *)dalloc((size_t)d->n3 * sizeof(double), &ok);
    if (!ok) break;
    if (hipMemcpy(*dk, kbf, (size_t)plane * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        upload2d(c, *db, bz))
      ok = 0;
  }
  free(kbf);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "transport: device tables");
  return 0;
}

/* One offline transport step of the window (model/hd/tracers/transport.c
 * :94-530 for an exact grid, closed boundaries, no fills, a linear
 * surface and CONS_W; then tracer_step and transport_post): the stored
 * flow of the step is in windat (u1, u2, Kz, eta). */
static int emstr_bdry(emstr_ctx *c);

int emstr_transport_step(emstr_ctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  int cc, r, dir;
  emstr_clear_error();
  if (wc->nonlinear)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "transport: vel_w_bounds with surface slope terms (nonlinear)");
  if (!wd->etab || !wd->detadt || !wd->wtop || !wd->wbot || !g->botzu1 || !g->botzu2 ||
      !g->bot_e1 || !g->bot_e2 || wd->dttr == 0.0)
    return emstr_set_error(EMSTR_EARG, -1,
                           "transport: needs etab, detadt, wtop, wbot, botzu1/2, bot_e1/2 and "
                           "a tracer step dttr");
  if ((r = flux3d_check(c, 0))) return r;
  if (!c->d_kbf1 && (r = trans_setup(c))) return r;
  /* vel_w_bounds, linear (vel3d.c:5144-5216; transport.c:346-354) on
   * the window's 2D arrays: detadt, the elevation's lateral ghosts,
   * wtop = detadt, wbot = 0, etab = eta */
  for (cc = 1; cc <= g->b2_t; cc++) {
    const int cw = g->w2_t[cc];
    wd->detadt[cw] = (wd->eta[cw] - wd->etab[cw]) / wd->dttr;
  }
  memset(wd->wbot, 0, (size_t)g->sgsizS * sizeof(double));
  for (cc = 1; cc <= g->nbpt; cc++)
    if (g->bpt[cc] <= g->sgnumS) wd->eta[g->bpt[cc]] = wd->eta[g->bin[cc]];
  memcpy(wd->wtop, wd->detadt, (size_t)g->sgsizS * sizeof(double));
  memcpy(wd->etab, wd->eta, (size_t)g->sgsizS * sizeof(double));
  /* the stored fields */
  d->dt = wd->dt;
  if (upload3d(c, d->Kz, wd->Kz) || upload2d(c, d->eta, wd->eta)) return EMSTR_EDEVICE;
  /* set_dz_at_u1 / u2, set_flux_3d (transport.c:438-440) */
  for (dir = 1; dir <= 2; dir++) {
    double *dzu = (dir == 1) ? c->d_dzu1 : c->d_dzu2;
    if (emstr_launch_dzu(d, dir, (dir == 1) ? c->d_kbf1 : c->d_kbf2,
                         (dir == 1) ? c->d_bzu1 : c->d_bzu2, dzu,
                         (dir == 1) ? c->d_ksur : c->d_ksur2, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "set_dz_at_u launch");
    if ((r = flux3d_dir(c, dir, 1, dzu, 0))) return r;
    /* horizontal diffusion and FFSL read the face thicknesses */
    if (c->have_kh)
      HCHK(hipMemcpyAsync((dir == 1) ? d->dzu1 : d->dzu2, dzu, (size_t)d->n3 * sizeof(double),
                          hipMemcpyDeviceToDevice, c->st));
  }
  /* set_dz, vel_w_update (transport.c:457-462) */
  if ((r = setdz_run(c, c->knew_dev))) return r;
  c->knew_dev = 1;
  if ((r = velw_run(c, 0))) return r;
  if ((r = form_fields(c))) return r;
  /* tracer_step (tracer_step_3d: the transport, then auxiliary_routines) */
  if ((r = emstr_step(c))) return r;
  if ((r = emstr_auxiliary(c))) return r;
  if ((r = emstr_bdry(c))) return r;
  /* transport_post: the new oldeta (transport.c:543) */
  memcpy(wc->oldeta, wd->eta, (size_t)g->sgsizS * sizeof(double));
  HCHK(hipMemcpyAsync(d->oldeta, d->eta, (size_t)d->plane * sizeof(double),
                      hipMemcpyDeviceToDevice, c->st));
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

/* The flow of the last transport step into the window's host arrays
 * (dzu1, dzu2, u1flux3d, u2flux3d, w and dz at the wet cells). */
int emstr_transport_download(emstr_ctx *c) {
  geometry_t *g = c->g;
  emstr_dev *d = &c->d;
  emstr_clear_error();
  if (!c->d_kbf1) return emstr_set_error(EMSTR_EARG, -1, "no transport step on the device");
  const double *src[6] = {c->d_dzu1, c->d_dzu2, d->u1flux3d, d->u2flux3d, d->w, d->dz};
  double *dst[6] = {c->wd->dzu1, c->wd->dzu2, c->wd->u1flux3d, c->wd->u2flux3d, c->wd->w,
                    c->wc->dz};
  for (int i = 0; i < 6; i++) {
    HCHK(hipMemsetAsync(c->stage, 0, (size_t)g->sgsiz * sizeof(double), c->st));
    if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, src[i], c->stage, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
    HCHK(hipMemcpyAsync(dst[i], c->stage, (size_t)g->sgsiz * sizeof(double),
                        hipMemcpyDeviceToHost, c->st));
  }
  HCHK(hipStreamSynchronize(c->st));
  return 0;
}

/*-------------------------------------------------------------------*/
/* auxiliary_routines (tracers.c:4296-4490): reset_dz, calc_age,      */
/* tracer_decay, do_ts_relax                                          */
/*-------------------------------------------------------------------*/
/* tracer_decay.c:45-52: seconds per unit of a decay timescale */
static double units_scale(const char *u) {
  if (strcmp(u, "days") == 0 || strcmp(u, "day") == 0) return 86400.0;
  if (strcmp(u, "hours") == 0 || strcmp(u, "hour") == 0) return 3600.0;
  if (strcmp(u, "minutes") == 0 || strcmp(u, "minute") == 0) return 60.0;
  return 1.0;
}

/* a 3D field into a device buffer made on first use */
static int aux_field(emstr_ctx *c, double **dst, const double *sp) {
  int ok = 1;
  if (!*dst) *dst = (double *)dalloc((size_t)c->d.n3 * sizeof(double), &ok);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "auxiliary_routines: device allocation");
  return upload3d(c, *dst, sp);
}

int emstr_auxiliary(emstr_ctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  int n, nn, ok = 1, r;
  AuxArgs a;
  memset(&a, 0, sizeof(a));
  a.age = a.tno = a.sno = -1;
  a.dttr = wd->dttr;
  /* the age tracer: agetr points at one of the 3D tracers */
  if (wd->agetr && wc->agemsk) {
    for (n = 0; n < wd->ntr; n++)
      if (wd->tr_wc[n] == wd->agetr) a.age = n;
    if (a.age < 0)
      return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                             "auxiliary_routines: the age tracer is not a 3D tracer of the window");
    double *m = (double *)malloc((size_t)g->sgsiz * sizeof(double));
    for (n = 0; n < g->sgsiz; n++) m[n] = wc->agemsk[n] ? 1.0 : 0.0;
    r = aux_field(c, &c->d_agem, m);
    HCHK(hipStreamSynchronize(c->st));
    free(m);
    if (r) return r;
    a.agemsk = c->d_agem;
  }
  /* the decay tables (tracer_decay.c:40-52) */
  if (wc->ntdec > 0) {
    if (!wc->tdec || !wc->dectr)
      return emstr_set_error(EMSTR_EARG, -1, "auxiliary_routines: ntdec without tdec/dectr");
    int *td = (int *)malloc(sizeof(int) * 2 * (size_t)wc->ntdec);
    double *dv = (double *)malloc(sizeof(double) * 2 * (size_t)wc->ntdec);
    int *t3 = td + wc->ntdec;
    double *sc = dv + wc->ntdec;
    for (nn = 0; nn < wc->ntdec; nn++) {
      n = wc->tdec[nn];
      const int flag = wc->trinfo_3d ? wc->trinfo_3d[n].flag : 0;
      if (n < 0 || n >= wd->ntr || ((flag & DE_TR3) && ((int)wc->dectr[n] < 0 ||
                                                        (int)wc->dectr[n] >= wd->ntr))) {
        free(td), free(dv);
        return emstr_set_error(EMSTR_EARG, -1, "auxiliary_routines: decay tracer out of range");
      }
      if (flag & DE_TR2) {
        free(td), free(dv);
        return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                               "auxiliary_routines: decay timescale from a 2D tracer (DE_TR2)");
      }
      td[nn] = n;
      dv[nn] = wc->dectr[n];
      t3[nn] = (flag & DE_TR3) ? (int)wc->dectr[n] : -1;
      sc[nn] = (flag & DE_TR3) ? units_scale(wc->trinfo_3d[t3[nn]].units) : 1.0;
    }
    if (wc->ntdec > c->ndec_cap) {
      if (c->d_tdec) hipFree(c->d_tdec), hipFree(c->d_dec), c->d_tdec = NULL;
      c->d_tdec = (int *)dalloc(sizeof(int) * 2 * (size_t)wc->ntdec, &ok);
      c->d_dec = (double *)dalloc(sizeof(double) * 2 * (size_t)wc->ntdec, &ok);
      if (!ok) {
        free(td), free(dv);
        return emstr_set_error(EMSTR_EDEVICE, -1, "auxiliary_routines: device allocation");
      }
      c->ndec_cap = wc->ntdec;
    }
    c->d_dtr3 = c->d_tdec + wc->ntdec;
    c->d_dsc = c->d_dec + wc->ntdec;
    HCHK(hipMemcpyAsync(c->d_tdec, td, sizeof(int) * 2 * (size_t)wc->ntdec,
                        hipMemcpyHostToDevice, c->st));
    HCHK(hipMemcpyAsync(c->d_dec, dv, sizeof(double) * 2 * (size_t)wc->ntdec,
                        hipMemcpyHostToDevice, c->st));
    HCHK(hipStreamSynchronize(c->st));
    free(td);
    free(dv);
    a.ntdec = wc->ntdec;
    a.tdec = c->d_tdec;
    a.dtr3 = c->d_tdec + wc->ntdec;
    a.dec = c->d_dec;
    a.dsc = c->d_dec + wc->ntdec;
  }
  /* temperature and salinity relaxation (tracers.c:6129-6190) */
  for (int q = 0; q < 2; q++) {
    const int tn = q ? wd->sno : wd->tno;
    if (tn < 0 || tn >= wd->ntr) continue;
    if (q) a.sno = tn; else a.tno = tn;
    double *rf = q ? wd->rsalt : wd->rtemp, *tc = q ? wd->salt_tc : wd->temp_tc;
    if (!rf || !wc->trinfo_3d) continue;
    const tracer_info_t *ti = &wc->trinfo_3d[tn];
    if ((r = aux_field(c, q ? &c->d_rsalt : &c->d_rtemp, rf))) return r;
    const int tcf = (tc && (ti->tctype & (RLX_ADPT | RLX_REG | RLX_OBC))) ? 1 : 0;
    if (tcf && (r = aux_field(c, q ? &c->d_stc : &c->d_ttc, tc))) return r;
    if (q) {
      a.rsalt = c->d_rsalt, a.salt_tc = tcf ? c->d_stc : NULL, a.rrS = ti->relax_rate;
      a.tcS = tcf, a.grdS = (ti->flag & RLX_GRD) ? 1 : 0, a.depthS = ti->relax_dum;
    } else {
      a.rtemp = c->d_rtemp, a.temp_tc = tcf ? c->d_ttc : NULL, a.rrT = ti->relax_rate;
      a.tcT = tcf, a.grdT = (ti->flag & RLX_GRD) ? 1 : 0, a.depthT = ti->relax_dum;
    }
  }
  if ((a.grdT || a.grdS) && (r = aux_field(c, &c->d_cellz, g->cellz))) return r;
  a.cellz = c->d_cellz;
  a.tracers = a.age >= 0 || a.ntdec > 0 ||
              ((a.tno >= 0 || a.sno >= 0) && (!wc->sigma || a.rtemp || a.rsalt));
  if (a.tracers && c->form != EMSTR_PLAIN_F64)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "auxiliary_routines: age, decay and T/S relaxation in the plain form only");
  if (!c->d_p2) {
    c->d_p2 = (int *)dalloc((size_t)d->plane * sizeof(int), &ok);
    c->h_p2 = (int *)calloc((size_t)d->plane, sizeof(int));
    if (!ok || !c->h_p2) return emstr_set_error(EMSTR_EDEVICE, -1, "auxiliary_routines: allocation");
  }
  if (emstr_launch_aux(d, &a, a.tracers ? c->tr[c->cur] : NULL, c->d_p2, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "auxiliary_routines launch");
  /* the reference fills the relaxation fields in place */
  const double *fsrc[2] = {a.grdT ? c->d_rtemp : NULL, a.grdS ? c->d_rsalt : NULL};
  double *fdst[2] = {wd->rtemp, wd->rsalt};
  for (int q = 0; q < 2; q++) {
    if (!fsrc[q]) continue;
    if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, fsrc[q], c->stage, c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
    double *tmp = (double *)malloc((size_t)g->sgsiz * sizeof(double));
    HCHK(hipMemcpyAsync(tmp, c->stage, (size_t)g->sgsiz * sizeof(double), hipMemcpyDeviceToHost,
                        c->st));
    HCHK(hipStreamSynchronize(c->st));
    for (int q3 = 1; q3 <= g->b3_t; q3++) fdst[q][g->w3_t[q3]] = tmp[g->w3_t[q3]];
    free(tmp);
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* open boundaries (tracers.c:4542-5165)                              */
/*-------------------------------------------------------------------*/
static ll dense_of(const emstr_ctx *c, int sp) {
  const geometry_t *g = c->g;
  return (ll)g->s2k[sp] * c->d.plane + (ll)g->s2j[sp] * c->d.nx + g->s2i[sp];
}

static void *obc_up(emstr_ctx *c, const void *h, size_t bytes, int *ok) {
  void *p = dalloc(bytes ? bytes : 8, ok);
  if (*ok && h && bytes && hipMemcpy(p, h, bytes, hipMemcpyHostToDevice) != hipSuccess) *ok = 0;
  c->obm = (void **)realloc(c->obm, (size_t)(c->nobm + 1) * sizeof(void *));
  c->obm[c->nobm++] = p;
  return p;
}

/* the geometry of every boundary on the device, and the column flags */
static int obc_setup(emstr_ctx *c) {
  geometry_t *g = c->g;
  emstr_dev *d = &c->d;
  int ok = 1, n, cc;
  if (c->ob || !g->nobc) return 0;
  c->nob = g->nobc;
  c->ob = (ObcArgs *)calloc((size_t)g->nobc, sizeof(ObcArgs));
  int *hm = (int *)calloc((size_t)d->plane, sizeof(int));
  for (cc = g->v2_t + 1; cc <= g->b2_t; cc++) hm[c->cc2col[cc]] = 1;
  /* a lateral ghost among the cells inwards holds the step's starting
   * value of its wet cell (set_lateral_BC_tr, tracers.c:5574-5593):
   * such a reference is ~(that cell's dense index), read from the
   * starting buffer */
  int *gbin = (int *)calloc((size_t)g->sgsiz, sizeof(int));
  for (cc = 1; cc <= g->nbpt; cc++) gbin[g->bpt[cc]] = g->bin[cc];
#define OBC_REF(sp) (gbin[sp] ? ~dense_of(c, gbin[sp]) : dense_of(c, sp))
  d->obcm = (int *)obc_up(c, hm, (size_t)d->plane * sizeof(int), &ok);
  free(hm);
  for (n = 0; n < g->nobc && ok; n++) {
    const open_bdrys_t *o = g->open[n];
    const int m = o->no3_t, xd = (o->ocodex & (L_EDGE | R_EDGE)) != 0;
    ll *cell = (ll *)malloc(((size_t)m + 1) * sizeof(ll));
    ll *in1 = (ll *)malloc(((size_t)m + 1) * sizeof(ll));
    ll *ch = (ll *)malloc(((size_t)m + 1) * 4 * sizeof(ll));
    double *cz = (double *)malloc(((size_t)m + 1) * sizeof(double));
    double *hat = (double *)malloc(((size_t)m + 1) * sizeof(double));
    for (cc = 1; cc <= m; cc++) {
      const int sp = o->obc_t[cc], q = cc - 1;
      int e = sp;
      cell[q] = dense_of(c, sp);
      in1[q] = OBC_REF(o->oi1_t[cc]);
      for (int j = 0; j < 4; j++) {
        e = o->nmap[e];
        ch[4 * q + j] = OBC_REF(e);
      }
      cz[q] = g->cellz[sp];
      hat[q] = xd ? g->h1au1[g->m2d[sp]] : g->h2au2[g->m2d[sp]];
    }
    ObcArgs *a = &c->ob[n];
    a->n = m;
    a->cell = (const ll *)obc_up(c, cell, (size_t)m * sizeof(ll), &ok);
    a->in1 = (const ll *)obc_up(c, in1, (size_t)m * sizeof(ll), &ok);
    a->chain = (const ll *)obc_up(c, ch, (size_t)m * 4 * sizeof(ll), &ok);
    a->cz = (const double *)obc_up(c, cz, (size_t)m * sizeof(double), &ok);
    a->hat = (const double *)obc_up(c, hat, (size_t)m * sizeof(double), &ok);
    a->vel = (const double *)obc_up(c, NULL, (size_t)m * sizeof(double), &ok);
    const size_t nt = (size_t)(c->wd->ntr > 0 ? c->wd->ntr : 1);
    a->bcond = (const int *)obc_up(c, NULL, nt * sizeof(int), &ok);
    a->frow = (const int *)obc_up(c, NULL, nt * sizeof(int), &ok);
    a->clampv = (const double *)obc_up(c, NULL, nt * sizeof(double), &ok);
    a->trpc = (const double *)obc_up(c, NULL, nt * sizeof(double), &ok);
    a->fv = (const double *)obc_up(c, NULL, nt * (size_t)m * sizeof(double), &ok);
    free(cell);
    free(in1);
    free(ch);
    free(cz);
    free(hat);
  }
#undef OBC_REF
  free(gbin);
  return ok ? 0 : emstr_set_error(EMSTR_EDEVICE, -1, "open boundary tables");
}

/* bdry_tracer (tracers.c:4542-4561) after auxiliary_routines: each
 * boundary in order, its settings checked and uploaded, then k_obc */
static int emstr_bdry(emstr_ctx *c) {
  geometry_t *g = c->g;
  window_t *wd = c->wd;
  win_priv_t *wc = c->wc;
  int n, cc, r;
  if (!g->nobc) return 0;
  if ((r = obc_setup(c))) return r;
  const int ntr = wd->ntr, nt = ntr > 0 ? ntr : 1;
  const int bad = CUSTOM | POLEXT | CYCLIC | STATIS | PROFIL | DEPROF | DESCAL | TRFLUX |
                  TRCONC | TRCONF;
  for (n = 0; n < g->nobc; n++) {
    const open_bdrys_t *o = g->open[n];
    ObcArgs *a = &c->ob[n];
    const int m = o->no3_t, xd = (o->ocodex & (L_EDGE | R_EDGE)) != 0;
    int *bc = (int *)calloc((size_t)nt, sizeof(int)), *fr = (int *)calloc((size_t)nt, sizeof(int));
    double *cv = (double *)calloc((size_t)nt, sizeof(double));
    double *pc = (double *)calloc((size_t)nt, sizeof(double));
    double *fv = (double *)calloc((size_t)nt * (m ? m : 1), sizeof(double));
    double *vel = (double *)calloc((size_t)(m ? m : 1), sizeof(double));
    int nf = 0, err = 0;
    if (o->bgz > 0 || (o->options & OP_OBCCM)) err = 1;
    for (int k = 0; k < wc->ntbdy && !err; k++) {
      const int tn = wc->tbdy[k], b = o->bcond_tra[tn];
      if ((b & bad) || (o->relax_zone_tra && o->relax_zone_tra[tn] > 0) ||
          !(b & (NOGRAD | NOTHIN | CLAMPD | LINEXT | FILEIN)) ||
          (o->trpc[tn] != 0.0 && !((b & NOGRAD) && (b & FILEIN))) ||
          ((b & FILEIN) && (!o->t_transfer || !o->trm || o->trm[tn] < 0)))
        err = 1;
      bc[tn] = b;
      cv[tn] = o->clampv[tn];
      pc[tn] = o->trpc[tn];
      if (!err && (b & FILEIN)) {
        fr[tn] = nf;
        for (cc = 1; cc <= m; cc++) fv[(size_t)nf * m + cc - 1] = o->t_transfer[o->trm[tn]][cc];
        nf++;
      }
    }
    /* the UPSTRM velocity of every cell, as upmeth picks it */
    const double *u = xd ? wd->u1 : wd->u2;
    const int *fb = xd ? o->obc_e1 : o->obc_e2, *fi = xd ? o->oi1_e1 : o->oi1_e2;
    const double sc = ((o->ocodex & L_EDGE) || (o->ocodey & B_EDGE)) ? -1.0 : 1.0;
    for (cc = 1; cc <= m && !err; cc++) {
      if (o->upmeth & CENTER)
        vel[cc - 1] = 0.5 * (u[fb[cc]] + u[fi[cc]]);
      else if (o->upmeth & FACE)
        vel[cc - 1] = u[fb[cc]];
      else if (o->upmeth & ADAPTIVE)
        vel[cc - 1] = -sc * 0.5 * (u[fb[cc]] + u[fi[cc]]) > 0.0 ? u[fb[cc]] : u[fi[cc]];
      else
        vel[cc - 1] = u[o->oi1_t[cc]];
    }
    if (!err) {
      r = hipMemcpyAsync((void *)a->bcond, bc, nt * sizeof(int), hipMemcpyHostToDevice, c->st) ||
          hipMemcpyAsync((void *)a->frow, fr, nt * sizeof(int), hipMemcpyHostToDevice, c->st) ||
          hipMemcpyAsync((void *)a->clampv, cv, nt * sizeof(double), hipMemcpyHostToDevice,
                         c->st) ||
          hipMemcpyAsync((void *)a->trpc, pc, nt * sizeof(double), hipMemcpyHostToDevice,
                         c->st) ||
          (m && hipMemcpyAsync((void *)a->vel, vel, (size_t)m * sizeof(double),
                               hipMemcpyHostToDevice, c->st)) ||
          (nf && m && hipMemcpyAsync((void *)a->fv, fv, (size_t)nf * m * sizeof(double),
                                     hipMemcpyHostToDevice, c->st));
      if (!r) r = hipStreamSynchronize(c->st) != hipSuccess;
    }
    free(bc);
    free(fr);
    free(cv);
    free(pc);
    free(fv);
    free(vel);
    if (err)
      return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                             "open boundary %d: conditions other than NOGRAD, NOTHIN, CLAMPD, "
                             "LINEXT, FILEIN, UPSTRM (or trpc without NOGRAD and FILEIN, flow "
                             "relaxation, ghost zones, corner averaging)",
                             n);
    if (r) return emstr_set_error(EMSTR_EDEVICE, -1, "open boundary upload");
    a->dt = wd->dt;
    a->sc = sc;
    a->owrite = (o->options & OP_OWRITE) ? 1 : 0;
    if (emstr_launch_obc(&c->d, a, c->tr[c->cur], c->tr[c->b0], c->st))
      return emstr_set_error(EMSTR_EDEVICE, -1, "open boundary launch");
  }
  return 0;
}

/* reset_dz's outputs on the host structures: the surface thicknesses
 * (dz of every wet cell from the device), the cells to process s2 with
 * vc2, vcs2, vca2, vci2 and i1 (same order as tracers.c:5944-5972) and
 * the mask c1 */
static int aux_to_host(emstr_ctx *c) {
  geometry_t *g = c->g;
  win_priv_t *wc = c->wc;
  emstr_dev *d = &c->d;
  int cc, vc = 1;
  HCHK(hipMemcpyAsync(c->h_p2, c->d_p2, (size_t)d->plane * sizeof(int), hipMemcpyDeviceToHost,
                      c->st));
  if (emstr_launch_gather(d->wsp, d->wdn, d->nwet, d->dz, c->stage, c->st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "gather");
  double *tmp = (double *)malloc((size_t)g->sgsiz * sizeof(double));
  HCHK(hipMemcpyAsync(tmp, c->stage, (size_t)g->sgsiz * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (int q = 1; q <= g->b3_t; q++) wc->dz[g->w3_t[q]] = tmp[g->w3_t[q]];
  free(tmp);
  HCHK(hipMemcpyAsync(c->h_dzg, d->dzg, (size_t)d->plane * sizeof(double),
                      hipMemcpyDeviceToHost, c->st));
  HCHK(hipStreamSynchronize(c->st));
  for (cc = 1; cc <= g->a2_t; cc++) {
    const ll col = c->cc2col[cc];
    wc->dz[c->sedg[col]] = c->h_dzg[col];
  }
  if (wc->sigma || !wc->s2) return 0;
  ll *cols = (ll *)malloc(sizeof(ll) * ((size_t)g->b2_t + 1));
  for (cc = 1; cc <= g->b2_t; cc++) {
    const ll col = c->cc2col[cc];
    if (!c->h_p2[col]) continue;
    wc->s2[vc] = sp_at(c, col, c->h_knew[col]);
    wc->i1[vc] = g->bot_t[cc];
    if (cc == g->v2_t) wc->vca2 = vc;
    cols[vc] = col;
    vc++;
  }
  const int vcs = vc - 1;
  wc->vcs2 = vcs;
  for (cc = 1; cc <= vcs; cc++) {
    const ll col = cols[cc];
    for (int k = c->h_knew[col] - 1; k >= c->h_kb[col]; k--) wc->s2[vc++] = sp_at(c, col, k);
    if (cc == wc->vca2) wc->vci2 = vc - 1;
  }
  wc->vc2 = vc - 1;
  free(cols);
  if (wc->c1) {
    memset(wc->c1, 0, (size_t)g->sgsiz);
    for (cc = 1; cc <= wc->vc2; cc++) wc->c1[wc->s2[cc]] = 1;
  }
  return 0;
}

/*-------------------------------------------------------------------*/
/* drop-in entry points                                              */
/*-------------------------------------------------------------------*/
void set_dz(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  emstr_ctx *c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (emstr_device_set_dz(c)) return;
  lists_to_host(c);
}

int advect_diffuse(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  emstr_ctx *c = ctx_for(window, windat, wincon);
  int r;
  if (!c) return 1;
  if ((r = emstr_upload_fields(c)) || (r = emstr_upload_tracers(c))) return r;
  if ((r = emstr_advect_diffuse(c))) return r;
  if ((r = emstr_download_tracers(c))) return r;
  return emstr_sync(c);
}

void vert_diffuse_3d(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  emstr_ctx *c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (emstr_upload_fields(c) || emstr_upload_tracers(c)) return;
  if (emstr_vert_diffuse(c)) return;
  if (emstr_download_tracers(c)) return;
  emstr_sync(c);
}

void tracer_step_3d(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  emstr_ctx *c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (emstr_upload_fields(c) || emstr_upload_tracers(c)) return;
  if (emstr_step(c)) return;
  /* tracers.c:289, 294 */
  if (emstr_auxiliary(c)) return;
  if (emstr_bdry(c)) return;
  if (emstr_download_tracers(c)) return;
  if (aux_to_host(c)) return;
  emstr_sync(c);
}

void auxiliary_routines(geometry_t *window, window_t *windat, win_priv_t *wincon) {
  emstr_ctx *c = ctx_for(window, windat, wincon);
  if (!c) return;
  if (emstr_upload_fields(c) || emstr_upload_tracers(c)) return;
  if (emstr_auxiliary(c)) return;
  if (emstr_download_tracers(c)) return;
  if (aux_to_host(c)) return;
  emstr_sync(c);
}

/* model/hd/tracers/tracers.c:142-209 over closed z-level windows: the
 * lateral ghosts of every window's tracers take their interior values
 * (set_lateral_BC_tr, tracers.c:5574-5593), then each window's step runs
 * on the GPU in the order twin and its tracers go back to the master, and
 * the mass diagnostics after all windows (tracers.c:186-189).  The
 * reference's other calls there do nothing at closed basins without
 * multi-dt, zooming, flushing or regions.  The lateral flux ghosts
 * (set_lateral_BC_vel) are not read: a face on land carries no flux. */
void tracer_step(master_t *master, geometry_t **window, window_t **windat,
                 win_priv_t **wincon, int nwindows) {
  int nn, n, cc, t;
  emstr_clear_error();
  if (!master || !window || !windat || !wincon || nwindows < 1) {
    emstr_set_error(EMSTR_EARG, -1, "tracer_step: null structure or no windows");
    if (master) master->crf = RS_RESTART;
    return;
  }
  if (master->mode2d) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1, "tracer_step: 2D mode (tracer_step_2d) is not built");
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
    if (emstr_last_error()[0]) {
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
