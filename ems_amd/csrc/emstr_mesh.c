/*
 * Builder of z-level COMPAS windows from a 2-D mesh description, with
 * the numbering and maps of the COMPAS preprocessor
 * (model/hd-us/master/pp_us.c:700-830 centres, :960-1200 edges,
 * :1374-1386 edge signs): centres layer by layer from the surface, wet
 * centres before the ghosts of their layer; c2c of a ghost points back
 * to its wet centre in the direction jo(j) and to itself otherwise; an
 * edge is created from the first-visited wet centre, which becomes
 * e2c[e][0]; land edges join a wet centre (e2c[e][0]) to its ghost.
 * Single window, closed boundaries: the cells and edges to process are
 * the wet ones ([1..v3] = [1..a3]), the ghosts follow up to n3.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/emstr_mesh.h"
#include "emstr_lsq.h"

static void *zalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

static int **alloc2(int rows, int cols) {
  int **a = (int **)zalloc(rows, sizeof(int *));
  int *d = (int *)zalloc((size_t)rows * cols, sizeof(int));
  for (int r = 0; r < rows; r++) a[r] = d + (size_t)r * cols;
  return a;
}
static void free2(int **a) {
  if (a) {
    free(a[0]);
    free(a);
  }
}

/* pp_us.c:5168 */
static int jo(int j, int npe) {
  int n = j + npe / 2;
  return (n > npe) ? n - npe : n;
}

emstr_mesh *emstr_mesh_build(int nmc, int npem, const int *npe, const int *nbr,
                             const int *nbrj, const int *kbot, int nz,
                             const double *layers, int ntr) {
  int m, j, k, n;
  if (nmc <= 0 || npem <= 0 || nz <= 0) return NULL;
#define NB(m, j) nbr[(size_t)(m) * npem + (j) - 1]
#define NBJ(m, j) nbrj[(size_t)(m) * npem + (j) - 1]
#define WETK(m, k) ((m) >= 0 && (k) >= kbot[m])
  emstr_mesh *M = (emstr_mesh *)zalloc(1, sizeof(emstr_mesh));
  M->nmc = nmc;
  M->npem = npem;
  M->nz = nz;
  M->ntr = ntr;
  M->kbot = (int *)zalloc(nmc, sizeof(int));
  memcpy(M->kbot, kbot, nmc * sizeof(int));
  M->idx3 = (int *)zalloc((size_t)nz * nmc, sizeof(int));
  M->gidx = (int *)zalloc((size_t)nz * nmc * npem, sizeof(int));
  M->eidx = (int *)zalloc((size_t)nz * nmc * npem, sizeof(int));
#define IDX(k, m) M->idx3[(size_t)(k) * nmc + (m)]
#define GIDX(k, m, j) M->gidx[((size_t)(k) * nmc + (m)) * npem + (j) - 1]
#define EIDX(k, m, j) M->eidx[((size_t)(k) * nmc + (m)) * npem + (j) - 1]

  /* Count and number the centres: per layer from the surface, wet
   * centres then their lateral ghosts; sediment ghosts of the columns
   * whose bed is in layer k+1 after layer k, those with bed in layer 0
   * at the end (cells deeper in the column have larger indices). */
  int c = 1, nwet = 0, ngh = 0, nsed = 0, szcS = 0;
  int *sedg = (int *)zalloc(nmc, sizeof(int));
  for (k = nz - 1; k >= 0; k--) {
    for (m = 0; m < nmc; m++)
      if (WETK(m, k)) {
        IDX(k, m) = c++;
        nwet++;
      }
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      for (j = 1; j <= npe[m]; j++)
        if (!WETK(NB(m, j), k)) {
          GIDX(k, m, j) = c++;
          ngh++;
        }
    }
    if (k == nz - 1) szcS = c;
    if (k >= 1)
      for (m = 0; m < nmc; m++)
        if (kbot[m] == k) {
          sedg[m] = c++;
          nsed++;
        }
  }
  for (m = 0; m < nmc; m++)
    if (kbot[m] == 0) {
      sedg[m] = c++;
      nsed++;
    }
  const int szc = c;
  M->sedg = sedg;

  geometry_t *g = (geometry_t *)zalloc(1, sizeof(geometry_t));
  window_t *wd = (window_t *)zalloc(1, sizeof(window_t));
  win_priv_t *wc = (win_priv_t *)zalloc(1, sizeof(win_priv_t));
  M->window = g;
  M->windat = wd;
  M->wincon = wc;
  g->szc = szc;
  g->szcS = szcS;
  g->nz = nz;
  g->npem = npem;
  g->wn = 1;
  g->layers = (double *)zalloc(nz + 1, sizeof(double));
  memcpy(g->layers, layers, (nz + 1) * sizeof(double));
  g->s2i = (int *)zalloc(szc, sizeof(int));
  g->s2k = (int *)zalloc(szc, sizeof(int));
  g->npe = (int *)zalloc(szcS, sizeof(int));
  g->c2c = alloc2(npem + 1, szc);
  g->c2e = alloc2(npem + 1, szc);
  g->m2d = (int *)zalloc(szc, sizeof(int));
  g->zp1 = (int *)zalloc(szc, sizeof(int));
  g->zm1 = (int *)zalloc(szc, sizeof(int));
  for (n = 0; n < szc; n++) {
    g->zp1[n] = g->zm1[n] = n;
    g->s2i[n] = -1;
    for (j = 1; j <= npem; j++) g->c2c[j][n] = n;
  }

  /* centre maps (pp_us.c:730-812) */
  g->nbpt = ngh;
  g->bpt = (int *)zalloc(ngh + 1, sizeof(int));
  g->bin = (int *)zalloc(ngh + 1, sizeof(int));
  int nb = 0;
  for (k = nz - 1; k >= 0; k--)
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      int cw = IDX(k, m), cs = IDX(nz - 1, m);
      g->s2i[cw] = m;
      g->s2k[cw] = k;
      g->m2d[cw] = cs;
      if (k == nz - 1) g->npe[cw] = npe[m];
      g->zp1[cw] = (k + 1 < nz) ? IDX(k + 1, m) : cw;
      g->zm1[cw] = (k > kbot[m]) ? IDX(k - 1, m) : sedg[m];
      for (j = 1; j <= npe[m]; j++) {
        int q = NB(m, j);
        if (WETK(q, k)) {
          g->c2c[j][cw] = IDX(k, q);
        } else {
          int gc = GIDX(k, m, j);
          g->c2c[j][cw] = gc;
          g->c2c[jo(j, npe[m])][gc] = cw;
          /* the ghost's column: the neighbour's surface centre when
           * the neighbour column is wet, else the surface-layer ghost
           * of the same edge */
          g->m2d[gc] = WETK(q, nz - 1) ? IDX(nz - 1, q) : GIDX(nz - 1, m, j);
          g->s2k[gc] = k;
          g->s2i[gc] = (q >= 0) ? q : -1;
          if (k == nz - 1) g->npe[gc] = npe[m];
          nb++;
          g->bpt[nb] = gc;
          g->bin[nb] = cw;
        }
      }
    }
  g->ngsed = nsed;
  g->gsed_t = (int *)zalloc(nsed + 1, sizeof(int));
  g->ised_t = (int *)zalloc(nsed + 1, sizeof(int));
  n = 0;
  for (m = 0; m < nmc; m++)
    if (kbot[m] < nz) {
      int cb = IDX(kbot[m], m);
      n++;
      g->gsed_t[n] = sedg[m];
      g->ised_t[n] = cb;
      g->zp1[sedg[m]] = cb;
      g->m2d[sedg[m]] = IDX(nz - 1, m);
      g->s2i[sedg[m]] = m;
      g->s2k[sedg[m]] = kbot[m] - 1;
    }

  /* Edges (pp_us.c:1084-1200): per layer from the surface, wet edges
   * created from the first-visited centre, then land edges. */
  int ne = 0, nwe = 0, nle = 0;
  for (k = nz - 1; k >= 0; k--)
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      for (j = 1; j <= npe[m]; j++) {
        int q = NB(m, j);
        if (WETK(q, k)) {
          if (q > m) ne++, nwe++;
        } else {
          ne++, nle++;
        }
      }
    }
  const int sze = ne + 1;
  g->sze = sze;
  g->e2c = alloc2(sze, 2);
  g->e2e = alloc2(sze, 2);
  g->m2de = (int *)zalloc(sze, sizeof(int));
  g->zp1e = (int *)zalloc(sze, sizeof(int));
  g->zm1e = (int *)zalloc(sze, sizeof(int));
  M->e2m = (int *)zalloc(sze, sizeof(int));
  int e = 1, szeS = 0;
  for (k = nz - 1; k >= 0; k--) {
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      int cw = IDX(k, m);
      for (j = 1; j <= npe[m]; j++) {
        int q = NB(m, j);
        if (!WETK(q, k) || q < m) continue;
        int cn = IDX(k, q), jj = NBJ(m, j);
        g->e2c[e][0] = cw;
        g->e2c[e][1] = cn;
        g->c2e[j][cw] = e;
        g->c2e[jj][cn] = e;
        g->e2e[e][0] = j;
        g->e2e[e][1] = jj;
        EIDX(k, m, j) = e;
        EIDX(k, q, jj) = e;
        M->e2m[e] = m * npem + j - 1;
        e++;
      }
    }
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      int cw = IDX(k, m);
      for (j = 1; j <= npe[m]; j++) {
        if (WETK(NB(m, j), k)) continue;
        int gc = GIDX(k, m, j);
        g->e2c[e][0] = cw;
        g->e2c[e][1] = gc;
        g->c2e[j][cw] = e;
        g->c2e[jo(j, npe[m])][gc] = e;
        g->e2e[e][0] = j;
        g->e2e[e][1] = jo(j, npe[m]);
        EIDX(k, m, j) = e;
        M->e2m[e] = m * npem + j - 1;
        e++;
      }
    }
    if (k == nz - 1) szeS = e;
  }
  g->szeS = szeS;
  /* surface and vertical edge maps */
  for (k = nz - 1; k >= 0; k--)
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      for (j = 1; j <= npe[m]; j++) {
        int ee = EIDX(k, m, j);
        if (!ee) continue;
        g->m2de[ee] = EIDX(nz - 1, m, j);
        int up = (k + 1 < nz) ? EIDX(k + 1, m, j) : 0;
        int dn = (k > 0 && WETK(m, k - 1)) ? EIDX(k - 1, m, j) : 0;
        g->zp1e[ee] = up ? up : ee;
        g->zm1e[ee] = dn ? dn : ee;
      }
    }

  /* ep/em (pp_us.c:1272-1282 with e2e() of pp_us.c:4974-5005): the
   * edge with e's index in the centre across edge j of e2c[e][0] */
  g->ep = (int *)zalloc(sze, sizeof(int));
  g->em = (int *)zalloc(sze, sizeof(int));
  for (e = 1; e < sze; e++) {
    for (int side = 0; side < 2; side++) {
      int jdir = g->e2e[e][1 - side];
      int c0 = g->e2c[e][0], nn;
      int np0 = g->npe[g->m2d[c0]];
      for (nn = 1; nn <= np0; nn++)
        if (g->c2e[nn][c0] == e) break;
      int cn = g->c2c[jdir][c0];
      int npn = g->npe[g->m2d[cn]];
      if (nn > npn) nn = npn;
      int en = (nn >= 1) ? g->c2e[nn][cn] : 0;
      if (side == 0)
        g->ep[e] = en ? en : e;
      else
        g->em[e] = en ? en : e;
    }
  }

  /* cells and edges to process */
#define ISWET(cx)                                                          \
  (g->s2i[cx] >= 0 && g->s2k[cx] >= kbot[g->s2i[cx]] && g->s2k[cx] < nz && \
   IDX(g->s2k[cx], g->s2i[cx]) == (cx))
  g->w3_t = (int *)zalloc(szc, sizeof(int));
  n = 0;
  for (c = 1; c < szc; c++)
    if (ISWET(c)) g->w3_t[++n] = c;
  g->v3_t = g->b3_t = g->a3_t = n;
  for (c = 1; c < szc; c++)
    if (!ISWET(c)) g->w3_t[++n] = c;
  g->n3_t = n;
  g->w3_e1 = (int *)zalloc(sze, sizeof(int));
  n = 0;
  for (e = 1; e < sze; e++)
    if (ISWET(g->e2c[e][1])) g->w3_e1[++n] = e;
  g->v3_e1 = g->b3_e1 = g->a3_e1 = n;
  for (e = 1; e < sze; e++)
    if (!ISWET(g->e2c[e][1])) g->w3_e1[++n] = e;
  g->n3_e1 = n;
  /* 2D vectors: surface wet centres / edges, then surface ghosts */
  g->w2_t = (int *)zalloc(szcS, sizeof(int));
  g->w2_e1 = (int *)zalloc(szeS, sizeof(int));
  n = 0;
  for (c = 1; c < szcS; c++)
    if (g->s2i[c] >= 0 && IDX(nz - 1, g->s2i[c]) == c) g->w2_t[++n] = c;
  g->v2_t = g->b2_t = g->a2_t = n;
  for (c = 1; c < szcS; c++)
    if (!(g->s2i[c] >= 0 && IDX(nz - 1, g->s2i[c]) == c)) g->w2_t[++n] = c;
  g->n2_t = n;
  n = 0;
  for (int ee = 1; ee <= g->v3_e1; ee++)
    if (g->w3_e1[ee] < szeS) g->w2_e1[++n] = g->w3_e1[ee];
  g->v2_e1 = g->b2_e1 = g->a2_e1 = n;
  for (int ee = g->v3_e1 + 1; ee <= g->n3_e1; ee++)
    if (g->w3_e1[ee] < szeS) g->w2_e1[++n] = g->w3_e1[ee];
  g->n2_e1 = n;

  /* surface / bed vectors */
  g->sur_t = (int *)zalloc(szcS, sizeof(int));
  g->nsur_t = (int *)zalloc(szcS, sizeof(int));
  g->bot_t = (int *)zalloc(szcS, sizeof(int));
  for (int cc = 1; cc <= g->v2_t; cc++) {
    int cs = g->w2_t[cc];
    m = g->s2i[cs];
    g->sur_t[cc] = g->nsur_t[cc] = cs;
    g->bot_t[cc] = IDX(kbot[m], m);
  }

  /* edge signs (pp_us.c:1374-1386) */
  g->eSc = alloc2(npem + 1, szcS);
  for (int cc = 1; cc <= g->b2_t; cc++) {
    int cs = g->w2_t[cc];
    for (j = 1; j <= g->npe[cs]; j++) {
      e = g->c2e[j][cs];
      g->eSc[j][cs] = (cs == g->e2c[e][1]) ? 1 : -1;
    }
  }

  /* metrics and fields (filled by the caller) */
  g->h1au1 = (double *)zalloc(szeS, sizeof(double));
  g->h2au1 = (double *)zalloc(szeS, sizeof(double));
  g->h1acell = (double *)zalloc(szeS, sizeof(double));
  g->costhu1 = (double *)zalloc(szeS, sizeof(double));
  g->sinthu1 = (double *)zalloc(szeS, sizeof(double));
  g->cellarea = (double *)zalloc(szcS, sizeof(double));
  g->cellx = (double *)zalloc(szcS, sizeof(double));
  g->celly = (double *)zalloc(szcS, sizeof(double));
  g->botz = (double *)zalloc(szcS, sizeof(double));
  g->gridz = (double *)zalloc(szc, sizeof(double));
  g->cellz = (double *)zalloc(szc, sizeof(double));
  for (c = 1; c < szc; c++) {
    int kk = g->s2k[c];
    if (kk >= 0 && kk < nz) {
      g->gridz[c] = layers[kk];
      g->cellz[c] = 0.5 * (layers[kk] + layers[kk + 1]);
    } else if (kk < 0) {
      g->gridz[c] = layers[0];
      g->cellz[c] = layers[0];
    }
  }

  const int mx = szc > sze ? szc : sze;
  const int mxS = szcS > szeS ? szcS : szeS;
  wd->ntr = ntr;
  wd->tr_wc = (double **)zalloc(ntr > 0 ? ntr : 1, sizeof(double *));
  for (n = 0; n < ntr; n++) wd->tr_wc[n] = (double *)zalloc(szc, sizeof(double));
  wd->u1 = (double *)zalloc(sze, sizeof(double));
  wd->u1flux3d = (double *)zalloc(sze, sizeof(double));
  wd->dzu1 = (double *)zalloc(sze, sizeof(double));
  wd->u = (double *)zalloc(szc, sizeof(double));
  wd->v = (double *)zalloc(szc, sizeof(double));
  wd->w = (double *)zalloc(szc, sizeof(double));
  wd->Kz = (double *)zalloc(szc, sizeof(double));
  wd->eta = (double *)zalloc(szcS, sizeof(double));
  wd->topz = (double *)zalloc(szcS, sizeof(double));
  wd->u2 = (double *)zalloc(sze, sizeof(double));
  wd->wtop = (double *)zalloc(szcS, sizeof(double));
  wd->wbot = (double *)zalloc(szcS, sizeof(double));
  /* lateral ghosts -> their wet centre */
  g->wgst = (int *)zalloc(szc, sizeof(int));
  for (n = 1; n <= g->nbpt; n++) g->wgst[g->bpt[n]] = g->bin[n];
  wd->dt = wd->dttr = 1.0;

  wc->trasc = VANLEER;
  wc->rkstage = 1;
  wc->stab = NONE;
  wc->hmin = 0.07;
  wc->trsf = 0.8;
  const int nt = ntr > 0 ? ntr : 1;
  wc->advect = (int *)zalloc(nt, sizeof(int));
  wc->diffuse = (int *)zalloc(nt, sizeof(int));
  wc->tbdy = (int *)zalloc(nt, sizeof(int));
  wc->tdif_v = (int *)zalloc(nt, sizeof(int));
  wc->mintr = (double *)zalloc(nt, sizeof(double));
  wc->maxtr = (double *)zalloc(nt, sizeof(double));
  for (n = 0; n < ntr; n++) {
    wc->advect[n] = 1;
    wc->tbdy[n] = n;
    wc->tdif_v[n] = n;
    wc->mintr[n] = 0.0;
    wc->maxtr[n] = 1e35;
  }
  wc->ntbdy = wc->ntdif_v = ntr;
  wc->oldeta = (double *)zalloc(szcS, sizeof(double));
  wc->dz = (double *)zalloc(szc, sizeof(double));
  wc->Ds = (double *)zalloc(szcS, sizeof(double));
  wc->Hn1 = (double *)zalloc(szcS, sizeof(double));
  wc->Hs = (double *)zalloc(szcS, sizeof(double));
  wc->one = (double *)zalloc(szcS, sizeof(double));
  for (n = 0; n < szcS; n++) wc->Ds[n] = wc->Hn1[n] = wc->one[n] = 1.0;
  wc->u1kh = (double *)zalloc(szc, sizeof(double));
  wc->u2kh = (double *)zalloc(sze, sizeof(double));
  wc->mdx = (double *)zalloc(szeS, sizeof(double));
  for (n = 0; n < szeS; n++) wc->mdx[n] = 1.0;
  wc->s1 = (int *)zalloc(szc, sizeof(int));
  int **iv[] = {&wc->i1, &wc->i2, &wc->i3, &wc->i4, &wc->i5};
  for (n = 0; n < 5; n++) *iv[n] = (int *)zalloc(szcS, sizeof(int));
  double **wv[] = {&wc->w1, &wc->w2, &wc->w3, &wc->w4, &wc->w5,
                   &wc->w6, &wc->w7, &wc->w8, &wc->w9, &wc->w10};
  for (n = 0; n < 10; n++) *wv[n] = (double *)zalloc(mx, sizeof(double));
  double **dv[] = {&wc->d1, &wc->d2, &wc->d3, &wc->d4};
  for (n = 0; n < 4; n++) *dv[n] = (double *)zalloc(mxS, sizeof(double));
  double **vv[] = {&wc->v1, &wc->v2, &wc->v3, &wc->v4, &wc->v5, &wc->v6, &wc->v7};
  for (n = 0; n < 7; n++) *vv[n] = (double *)zalloc(nz + 2, sizeof(double));
  (void)nwe;
  (void)nle;
  (void)nwet;
  return M;
#undef NB
#undef NBJ
#undef WETK
#undef IDX
#undef GIDX
#undef EIDX
#undef ISWET
}

void emstr_mesh_free(emstr_mesh *M) {
  int n;
  if (!M) return;
  geometry_t *g = M->window;
  window_t *wd = M->windat;
  win_priv_t *wc = M->wincon;
  const int szc = g ? g->szc : 0;
  if (g) {
    void *p[] = {g->layers, g->s2i, g->s2k, g->npe, g->m2d, g->m2de, g->zp1, g->zm1,
                 g->zp1e, g->zm1e, g->ep, g->em, g->w3_t, g->w3_e1, g->w2_t, g->w2_e1, g->bpt,
                 g->bin, g->gsed_t, g->ised_t, g->sur_t, g->nsur_t, g->bot_t,
                 g->h1au1, g->h2au1, g->h1acell, g->costhu1, g->sinthu1, g->cellarea, g->cellx,
                 g->celly, g->botz, g->gridz, g->cellz, g->wgst, g->nvc, g->m2dv,
                 g->gridx, g->gridy, g->u1x, g->u1y, g->costhu2, g->sinthu2};
    for (n = 0; n < (int)(sizeof(p) / sizeof(p[0])); n++) free(p[n]);
    free2(g->c2v);
    free2(g->v2c);
    free2(g->e2v);
    free2(g->c2c);
    free2(g->c2e);
    free2(g->e2c);
    free2(g->e2e);
    free2(g->eSc);
    free(g);
  }
  if (wd) {
    for (n = 0; n < wd->ntr; n++) free(wd->tr_wc[n]);
    void *p[] = {wd->tr_wc, wd->u1, wd->u1flux3d, wd->dzu1, wd->u, wd->v, wd->w,
                 wd->Kz, wd->eta, wd->topz, wd->u2, wd->wtop, wd->wbot};
    for (n = 0; n < (int)(sizeof(p) / sizeof(p[0])); n++) free(p[n]);
    free(wd);
  }
  if (wc) {
    if (wc->lw) {
      for (n = 0; n < szc; n++) {
        free(wc->lw[n].cells);
        if (wc->lw[n].B) {
          free(wc->lw[n].B[0]);
          free(wc->lw[n].B);
        }
      }
      free(wc->lw);
    }
    void *p[] = {wc->advect, wc->diffuse, wc->tbdy, wc->tdif_v, wc->mintr, wc->maxtr,
                 wc->oldeta, wc->dz, wc->Ds, wc->Hn1, wc->Hs, wc->one, wc->u1kh,
                 wc->u2kh, wc->mdx, wc->s1, wc->i1, wc->i2, wc->i3, wc->i4, wc->i5,
                 wc->w1, wc->w2, wc->w3, wc->w4, wc->w5, wc->w6, wc->w7, wc->w8,
                 wc->w9, wc->w10, wc->d1, wc->d2, wc->d3, wc->d4, wc->v1, wc->v2,
                 wc->v3, wc->v4, wc->v5, wc->v6, wc->v7};
    for (n = 0; n < (int)(sizeof(p) / sizeof(p[0])); n++) free(p[n]);
    free(wc);
  }
  free(M->kbot);
  free(M->idx3);
  free(M->gidx);
  free(M->eidx);
  free(M->e2m);
  free(M->sedg);
  free(M->v2m);
  free(M);
}

/*-------------------------------------------------------------------*/
/* Vertices (pp_us.c:1900-2100): vertex j of a centre is the corner  */
/* where its edge j starts (counter-clockwise), e2v[e] runs along edge */
/* j of e2c[e][0]; a vertex exists in layer k when a wet centre of    */
/* that layer has it as a corner; v2c lists the wet centres of the    */
/* vertex's layer around it (ghosts are not mapped, pp_us.c:2027).    */
/*-------------------------------------------------------------------*/
int emstr_mesh_vertices(emstr_mesh *M, int nv, const int *vid) {
  geometry_t *g = M->window;
  const int nmc = M->nmc, npem = M->npem, nz = M->nz;
  const int *kbot = M->kbot;
  int k, m, j, n;
  if (!g || nv <= 0 || g->c2v) return -1;
#define VID(m, j) vid[(size_t)(m) * npem + (j) - 1]
#define WETK(m, k) ((m) >= 0 && (k) >= kbot[m])
#define IDX(k, m) M->idx3[(size_t)(k) * nmc + (m)]
  for (m = 0; m < nmc; m++) {
    if (kbot[m] >= nz) continue;
    for (j = 1; j <= g->npe[IDX(nz - 1, m)]; j++)
      if (VID(m, j) < 0 || VID(m, j) >= nv) return -2;
  }
  int *num = (int *)zalloc((size_t)nz * nv, sizeof(int)); /* [k][vertex] */
  int *cnt = (int *)zalloc((size_t)nz * nv, sizeof(int));
  int v = 1, szvS = 0;
  for (k = nz - 1; k >= 0; k--) {
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      const int np = g->npe[IDX(nz - 1, m)];
      for (j = 1; j <= np; j++) {
        int *t = &num[(size_t)k * nv + VID(m, j)];
        if (!*t) *t = v++;
        cnt[(size_t)k * nv + VID(m, j)]++;
      }
    }
    if (k == nz - 1) szvS = v;
  }
  const int szv = v;
  int nvcm = 0;
  for (n = 0; n < nv; n++)
    if (cnt[(size_t)(nz - 1) * nv + n] > nvcm) nvcm = cnt[(size_t)(nz - 1) * nv + n];
  g->szv = szv;
  g->szvS = szvS;
  g->nvcm = nvcm;
  g->c2v = alloc2(npem + 1, g->szc);
  g->v2c = alloc2(szv, nvcm + 1);
  g->nvc = (int *)zalloc(szvS, sizeof(int));
  g->m2dv = (int *)zalloc(szv, sizeof(int));
  g->e2v = alloc2(g->sze, 2);
  M->v2m = (int *)zalloc(szvS, sizeof(int));
  for (k = nz - 1; k >= 0; k--)
    for (m = 0; m < nmc; m++) {
      if (!WETK(m, k)) continue;
      const int c = IDX(k, m), np = g->npe[IDX(nz - 1, m)];
      for (j = 1; j <= np; j++) {
        const int mv = VID(m, j);
        const int v3 = num[(size_t)k * nv + mv], vs = num[(size_t)(nz - 1) * nv + mv];
        g->c2v[j][c] = v3;
        g->m2dv[v3] = vs;
        if (k == nz - 1) {
          g->nvc[vs]++;
          M->v2m[vs] = mv;
        }
        for (n = 1; n <= nvcm; n++)
          if (!g->v2c[v3][n]) {
            g->v2c[v3][n] = c;
            break;
          }
      }
    }
  for (int e = 1; e < g->sze; e++) {
    const int c0 = g->e2c[e][0], jj = g->e2e[e][0];
    const int m0 = g->s2i[c0], np = g->npe[g->m2d[c0]];
    if (m0 < 0 || jj < 1) continue;
    g->e2v[e][0] = g->c2v[jj][c0];
    g->e2v[e][1] = g->c2v[jj % np + 1][c0];
  }
  g->gridx = (double *)zalloc(szvS, sizeof(double));
  g->gridy = (double *)zalloc(szvS, sizeof(double));
  g->u1x = (double *)zalloc(g->szeS, sizeof(double));
  g->u1y = (double *)zalloc(g->szeS, sizeof(double));
  g->costhu2 = (double *)zalloc(g->szeS, sizeof(double));
  g->sinthu2 = (double *)zalloc(g->szeS, sizeof(double));
  M->nv = nv;
  free(num);
  free(cnt);
  return 0;
#undef VID
#undef WETK
#undef IDX
}

/*-------------------------------------------------------------------*/
/* Least-squares weights of every wet centre (build_linear_weights,   */
/* tracers.c:4278-4440, with df = 2, and get_linear_metrics,          */
/* tracers.c:4602-4673): the stencil is the centre and the wet        */
/* centres around its vertices in its layer (stencil() ST_SIZED of    */
/* size 3, diagnostics/monitor.c:2015-2050), each followed by the     */
/* ghosts across its edges; rows (1, cellx[c]-x, celly[c]-y) at the   */
/* surface positions of the stencil centres.  B is the pseudo-inverse */
/* of those rows; the reference takes it from an SVD (svd_lsq_B), here */
/* from the eigen-decomposition of the 3x3 normal matrix with the same */
/* cut of near-zero singular values.                                   */
/*-------------------------------------------------------------------*/
int emstr_mesh_lsq(emstr_mesh *M) {
  geometry_t *g = M->window;
  win_priv_t *wc = M->wincon;
  int cc, i, n;
  if (!g || !g->c2v || wc->lw) return -1;
  wc->lw = (qweights *)zalloc(g->szc, sizeof(qweights));
  int *mask = (int *)zalloc(g->szcS, sizeof(int));
  int *st = (int *)zalloc(64, sizeof(int));
  int *cells = (int *)zalloc(64 * (g->npem + 1), sizeof(int));
  double *px = (double *)zalloc(64 * (g->npem + 1), sizeof(double));
  double *py = (double *)zalloc(64 * (g->npem + 1), sizeof(double));
  double *Bw = (double *)zalloc(3 * 64 * (g->npem + 1), sizeof(double));
  for (cc = 1; cc <= g->a3_t; cc++) {
    const int c = g->w3_t[cc], cs = g->m2d[c];
    int nc = 0;
    /* stencil: the centre, then the wet centres around its vertices */
    const int ns = emstr_lsq_stencil(g, c, st, 64, mask);
    /* each stencil centre, then the ghosts across its edges */
    for (i = 0; i < ns; i++) {
      const int cn = st[i], cns = g->m2d[cn];
      cells[nc] = cn;
      px[nc] = g->cellx[cns];
      py[nc] = g->celly[cns];
      nc++;
      for (n = 1; n <= g->npe[cns]; n++) {
        const int cg = g->c2c[n][cn];
        if (g->wgst[cg]) {
          cells[nc] = cg;
          px[nc] = g->cellx[g->m2d[cg]];
          py[nc] = g->celly[g->m2d[cg]];
          nc++;
        }
      }
    }
    /* pseudo-inverse of the rows (1, x, y) */
    for (i = 0; i < nc; i++) {
      px[i] = g->cellx[cs] - px[i];
      py[i] = g->celly[cs] - py[i];
    }
    emstr_lsq_pinv3(nc, px, py, Bw);
    qweights *lw = &wc->lw[c];
    lw->ncells = nc;
    lw->cells = (int *)zalloc(nc, sizeof(int));
    lw->B = (double **)zalloc(nc, sizeof(double *));
    lw->B[0] = (double *)zalloc((size_t)nc * 3, sizeof(double));
    for (i = 0; i < nc; i++) {
      lw->cells[i] = cells[i];
      lw->B[i] = lw->B[0] + (size_t)i * 3;
      for (int a = 0; a < 3; a++) lw->B[i][a] = Bw[3 * i + a];
    }
    lw->beta = 1.0;
  }
  free(mask);
  free(st);
  free(cells);
  free(px);
  free(py);
  free(Bw);
  return 0;
}
