/*
 * Structured z-level grid builder: fills the reference's sparse
 * structures (geometry_t / window_t / win_priv_t) for one window.
 * Numbering and ghost rules: model/hd/master/preprocess.c:108-720 (see
 * emstr_grid.h).  Also w from continuity (momentum/vel3d.c:4609-4720).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "emstr_grid.h"
#include "emstr_internal.h"

static void *zalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  return p;
}

#define DIRS 4 /* W, E, S, N */

emstr_grid *emstr_grid_build(int nx, int ny, int nz, const double *layers,
                             const double *botz, const double *h1acell,
                             const double *h2acell, const double *h1au1,
                             const double *h2au2, const double *h2au1,
                             const double *h1au2, const double *cellarea,
                             int ntr) {
  int i, j, k, n, d;
  const int di[DIRS] = {-1, 1, 0, 0};
  const int dj[DIRS] = {0, 0, -1, 1};
  emstr_clear_error();
  if (nx < 3 || ny < 3 || nz < 1 || ntr < 0 || !layers || !botz) {
    emstr_set_error(EMSTR_EARG, -1, "emstr_grid_build: bad arguments");
    return NULL;
  }
  for (k = 0; k < nz; k++)
    if (!(layers[k + 1] > layers[k])) {
      emstr_set_error(EMSTR_EARG, -1, "layers must increase");
      return NULL;
    }
  const double top = layers[nz];
  const size_t nxy = (size_t)nx * ny;
  int *kbot = (int *)zalloc(nxy, sizeof(int));
  for (j = 0; j < ny; j++)
    for (i = 0; i < nx; i++) {
      size_t q = (size_t)j * nx + i;
      double b = botz[q];
      if (!(b < top) || isnan(b)) {
        kbot[q] = nz;
      } else {
        int kb = 0;
        for (k = 0; k < nz; k++)
          if (layers[k] <= b) kb = k;
        kbot[q] = kb;
      }
      if ((i == 0 || j == 0 || i == nx - 1 || j == ny - 1) && kbot[q] < nz) {
        free(kbot);
        emstr_set_error(EMSTR_EARG, -1,
                        "column (%d,%d) on the grid edge must be land", i, j);
        return NULL;
      }
    }
#define WET(ii, jj, kk) ((kk) >= kbot[(size_t)(jj)*nx + (ii)])

  /* Count wet cells, lateral ghosts and sediment ghosts per layer. */
  size_t nwet = 0, nlat = 0, nsed = 0;
  for (k = nz - 1; k >= 0; k--)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        if (!WET(i, j, k)) continue;
        nwet++;
        for (d = 0; d < DIRS; d++)
          if (!WET(i + di[d], j + dj[d], k)) nlat++;
        if (k == kbot[(size_t)j * nx + i]) nsed++;
      }
  const int sgnum = (int)(nwet + nlat + nsed);
  const int sgsiz = sgnum + 1;

  /* Sparse index of each wet (i,j,k); ghosts are numbered as created. */
  int *idx3 = (int *)zalloc(nxy * nz, sizeof(int));
  int *xp1 = (int *)zalloc(sgsiz, sizeof(int));
  int *xm1 = (int *)zalloc(sgsiz, sizeof(int));
  int *yp1 = (int *)zalloc(sgsiz, sizeof(int));
  int *ym1 = (int *)zalloc(sgsiz, sizeof(int));
  int *zp1 = (int *)zalloc(sgsiz, sizeof(int));
  int *zm1 = (int *)zalloc(sgsiz, sizeof(int));
  int *s2i = (int *)zalloc(sgsiz, sizeof(int));
  int *s2j = (int *)zalloc(sgsiz, sizeof(int));
  int *s2k = (int *)zalloc(sgsiz, sizeof(int));
  int *m2d = (int *)zalloc(sgsiz, sizeof(int));
  int *bpt = (int *)zalloc(nlat + 1, sizeof(int));
  int *bin = (int *)zalloc(nlat + 1, sizeof(int));
  int *gsed = (int *)zalloc(nsed + 1, sizeof(int));
  int *ised = (int *)zalloc(nsed + 1, sizeof(int));
  int *gdir = (int *)zalloc(sgsiz, sizeof(int)); /* ghost direction+1 */
  int *w3_t = (int *)zalloc(sgsiz, sizeof(int));
  for (n = 0; n <= sgnum; n++) {
    xp1[n] = xm1[n] = yp1[n] = ym1[n] = zp1[n] = zm1[n] = n;
  }

  int c = 1, nb = 0, ns = 0, sgnumS = 0;
  int *sedpending = (int *)zalloc(nxy, sizeof(int));
  for (k = nz - 1; k >= 0; k--) {
    /* wet cells of layer k */
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++)
        if (WET(i, j, k)) {
          idx3[((size_t)k * ny + j) * nx + i] = c;
          s2i[c] = i;
          s2j[c] = j;
          s2k[c] = k;
          c++;
        }
    /* lateral ghosts of layer k, in wet-cell order, W E S N */
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        if (!WET(i, j, k)) continue;
        int cw = idx3[((size_t)k * ny + j) * nx + i];
        for (d = 0; d < DIRS; d++) {
          int ii = i + di[d], jj = j + dj[d];
          if (WET(ii, jj, k)) continue;
          int g = c++;
          s2i[g] = ii;
          s2j[g] = jj;
          s2k[g] = k;
          gdir[g] = d + 1;
          nb++;
          bpt[nb] = g;
          bin[nb] = cw;
          if (d == 0) { xm1[cw] = g; xp1[g] = cw; }
          if (d == 1) { xp1[cw] = g; xm1[g] = cw; }
          if (d == 2) { ym1[cw] = g; yp1[g] = cw; }
          if (d == 3) { yp1[cw] = g; ym1[g] = cw; }
        }
      }
    if (k == nz - 1) sgnumS = c - 1;
    /* sediment ghosts of columns whose bed cell is at layer k+1 */
    if (k + 1 < nz)
      for (j = 0; j < ny; j++)
        for (i = 0; i < nx; i++)
          if (kbot[(size_t)j * nx + i] == k + 1) {
            int cb = idx3[((size_t)(k + 1) * ny + j) * nx + i];
            int g = c++;
            s2i[g] = i;
            s2j[g] = j;
            s2k[g] = k;
            ns++;
            gsed[ns] = g;
            ised[ns] = cb;
            zm1[cb] = g;
            zp1[g] = cb;
          }
  }
  /* sediment ghosts of columns whose bed cell is in layer 0 */
  for (j = 0; j < ny; j++)
    for (i = 0; i < nx; i++)
      if (kbot[(size_t)j * nx + i] == 0) {
        int cb = idx3[(size_t)j * nx + i];
        int g = c++;
        s2i[g] = i;
        s2j[g] = j;
        s2k[g] = -1;
        ns++;
        gsed[ns] = g;
        ised[ns] = cb;
        zm1[cb] = g;
        zp1[g] = cb;
      }
  free(sedpending);
  (void)nb;

  /* Wet-wet links */
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        if (!WET(i, j, k)) continue;
        int cw = idx3[((size_t)k * ny + j) * nx + i];
        if (WET(i + 1, j, k)) xp1[cw] = idx3[((size_t)k * ny + j) * nx + i + 1];
        if (WET(i - 1, j, k)) xm1[cw] = idx3[((size_t)k * ny + j) * nx + i - 1];
        if (WET(i, j + 1, k)) yp1[cw] = idx3[((size_t)k * ny + j + 1) * nx + i];
        if (WET(i, j - 1, k)) ym1[cw] = idx3[((size_t)k * ny + j - 1) * nx + i];
        if (k + 1 < nz) zp1[cw] = idx3[((size_t)(k + 1) * ny + j) * nx + i];
        if (k > kbot[(size_t)j * nx + i])
          zm1[cw] = idx3[((size_t)(k - 1) * ny + j) * nx + i];
      }

  /* 2D map: wet cells and sediment ghosts map to their column's surface
   * cell; a lateral ghost maps to the surface cell at its (i,j): the wet
   * surface cell if (i,j) is a wet column, else the surface-layer ghost
   * created in the same direction from the same column. */
  for (n = 1; n <= sgnum; n++) {
    i = s2i[n];
    j = s2j[n];
    if (n <= sgnumS) {
      m2d[n] = n;
      continue;
    }
    if (!gdir[n]) {
      m2d[n] = idx3[((size_t)(nz - 1) * ny + j) * nx + i];
      continue;
    }
    if (kbot[(size_t)j * nx + i] < nz) {
      m2d[n] = idx3[((size_t)(nz - 1) * ny + j) * nx + i];
    } else {
      d = gdir[n] - 1;
      int ci = i - di[d], cj = j - dj[d]; /* the wet column it came from */
      int cs = idx3[((size_t)(nz - 1) * ny + cj) * nx + ci];
      int g = (d == 0) ? xm1[cs] : (d == 1) ? xp1[cs] : (d == 2) ? ym1[cs] : yp1[cs];
      m2d[n] = g;
    }
  }

  /* Allocate the structures */
  emstr_grid *G = (emstr_grid *)zalloc(1, sizeof(emstr_grid));
  geometry_t *gw = (geometry_t *)zalloc(1, sizeof(geometry_t));
  window_t *wd = (window_t *)zalloc(1, sizeof(window_t));
  win_priv_t *wc = (win_priv_t *)zalloc(1, sizeof(win_priv_t));
  G->window = gw;
  G->windat = wd;
  G->wincon = wc;
  G->nx = nx;
  G->ny = ny;
  G->nz = nz;
  G->ntr = ntr;
  G->kbot = kbot;

  gw->sgnum = sgnum;
  gw->sgsiz = sgsiz;
  gw->sgnumS = sgnumS;
  gw->sgsizS = sgnumS + 1;
  gw->nobc = 0;
  gw->nce1 = nx;
  gw->nce2 = ny;
  gw->nz = nz;
  gw->wn = 1;
  gw->layers = (double *)zalloc(nz + 1, sizeof(double));
  memcpy(gw->layers, layers, (nz + 1) * sizeof(double));
  gw->s2i = s2i;
  gw->s2j = s2j;
  gw->s2k = s2k;
  gw->xp1 = xp1;
  gw->xm1 = xm1;
  gw->yp1 = yp1;
  gw->ym1 = ym1;
  gw->zp1 = zp1;
  gw->zm1 = zm1;
  gw->m2d = m2d;
  gw->nbpt = (int)nlat;
  gw->bpt = bpt;
  gw->bin = bin;
  gw->ngsed = (int)nsed;
  gw->gsed_t = gsed;
  gw->ised_t = ised;

  /* w3_t: wet cells in sparse order, then ghosts */
  int nv = 0;
  for (n = 1; n <= sgnum; n++)
    if (!gdir[n] && s2k[n] >= 0 && idx3[((size_t)s2k[n] * ny + s2j[n]) * nx + s2i[n]] == n)
      w3_t[++nv] = n;
  int nt = nv;
  for (n = 1; n <= sgnum; n++)
    if (!(!gdir[n] && s2k[n] >= 0 && idx3[((size_t)s2k[n] * ny + s2j[n]) * nx + s2i[n]] == n))
      w3_t[++nt] = n;
  gw->w3_t = w3_t;
  gw->v3_t = gw->b3_t = gw->a3_t = nv;
  gw->n3_t = nt;
  /* e1 faces: west face of every wet cell, then east-wall ghosts;
   * e2 faces: south face of every wet cell, then north-wall ghosts. */
  int *w3_e1 = (int *)zalloc(sgsiz, sizeof(int));
  int *w3_e2 = (int *)zalloc(sgsiz, sizeof(int));
  int ne1 = nv, ne2 = nv;
  for (n = 1; n <= nv; n++) w3_e1[n] = w3_e2[n] = w3_t[n];
  for (n = 1; n <= sgnum; n++) {
    if (gdir[n] == 2) w3_e1[++ne1] = n;
    if (gdir[n] == 4) w3_e2[++ne2] = n;
  }
  gw->w3_e1 = w3_e1;
  gw->w3_e2 = w3_e2;
  gw->v3_e1 = gw->b3_e1 = gw->a3_e1 = nv;
  gw->v3_e2 = gw->b3_e2 = gw->a3_e2 = nv;
  gw->n3_e1 = ne1;
  gw->n3_e2 = ne2;

  /* 2D vectors: surface wet cells, then surface ghosts */
  int nv2 = 0;
  for (n = 1; n <= sgnumS; n++)
    if (!gdir[n]) nv2++;
  int *w2_t = (int *)zalloc(gw->sgsizS, sizeof(int));
  int *w2_e1 = (int *)zalloc(gw->sgsizS, sizeof(int));
  int *w2_e2 = (int *)zalloc(gw->sgsizS, sizeof(int));
  {
    int a = 0, b = nv2;
    for (n = 1; n <= sgnumS; n++) {
      if (!gdir[n]) w2_t[++a] = n;
      else w2_t[++b] = n;
    }
    for (n = 1; n <= nv2; n++) w2_e1[n] = w2_e2[n] = w2_t[n];
    int e1 = nv2, e2 = nv2;
    for (n = 1; n <= sgnumS; n++) {
      if (gdir[n] == 2) w2_e1[++e1] = n;
      if (gdir[n] == 4) w2_e2[++e2] = n;
    }
    gw->n2_t = sgnumS;
    gw->n2_e1 = e1;
    gw->n2_e2 = e2;
  }
  gw->w2_t = w2_t;
  gw->w2_e1 = w2_e1;
  gw->w2_e2 = w2_e2;
  gw->v2_t = gw->b2_t = gw->a2_t = nv2;
  gw->v2_e1 = gw->b2_e1 = gw->a2_e1 = nv2;
  gw->v2_e2 = gw->b2_e2 = gw->a2_e2 = nv2;

  gw->sur_t = (int *)zalloc(gw->sgsizS, sizeof(int));
  gw->nsur_t = (int *)zalloc(gw->sgsizS, sizeof(int));
  gw->bot_t = (int *)zalloc(gw->sgsizS, sizeof(int));
  for (n = 1; n <= nv2; n++) {
    int cs = w2_t[n];
    int cb = idx3[((size_t)kbot[(size_t)s2j[cs] * nx + s2i[cs]] * ny + s2j[cs]) * nx + s2i[cs]];
    gw->bot_t[n] = cb;
    gw->sur_t[n] = gw->nsur_t[n] = cs;
  }

  /* 2D metrics at every 2D sparse cell, from its (i,j) */
  const int S = gw->sgsizS;
  gw->h1acell = (double *)zalloc(S, sizeof(double));
  gw->h2acell = (double *)zalloc(S, sizeof(double));
  gw->h1au1 = (double *)zalloc(S, sizeof(double));
  gw->h2au2 = (double *)zalloc(S, sizeof(double));
  gw->h2au1 = (double *)zalloc(S, sizeof(double));
  gw->h1au2 = (double *)zalloc(S, sizeof(double));
  gw->cellarea = (double *)zalloc(S, sizeof(double));
  gw->botz = (double *)zalloc(S, sizeof(double));
  for (n = 1; n <= sgnumS; n++) {
    size_t q = (size_t)s2j[n] * nx + s2i[n];
    if (h1acell) gw->h1acell[n] = h1acell[q];
    if (h2acell) gw->h2acell[n] = h2acell[q];
    if (h1au1) gw->h1au1[n] = h1au1[q];
    if (h2au2) gw->h2au2[n] = h2au2[q];
    if (h2au1) gw->h2au1[n] = h2au1[q];
    if (h1au2) gw->h1au2[n] = h1au2[q];
    if (cellarea) gw->cellarea[n] = cellarea[q];
    gw->botz[n] = (kbot[q] < nz) ? botz[q] : top;
  }
  gw->gridz = (double *)zalloc(sgsiz, sizeof(double));
  gw->cellz = (double *)zalloc(sgsiz, sizeof(double));
  for (n = 1; n <= sgnum; n++) {
    k = s2k[n];
    if (k >= 0) {
      gw->gridz[n] = layers[k];
      gw->cellz[n] = 0.5 * (layers[k] + layers[k + 1]);
    } else {
      gw->gridz[n] = layers[0];
      gw->cellz[n] = layers[0];
    }
  }

  /* window_t */
  wd->ntr = ntr;
  wd->tr_wc = (double **)zalloc(ntr > 0 ? ntr : 1, sizeof(double *));
  for (n = 0; n < ntr; n++) wd->tr_wc[n] = (double *)zalloc(sgsiz, sizeof(double));
  wd->dt = wd->dttr = 1.0;
  wd->u1 = (double *)zalloc(sgsiz, sizeof(double));
  wd->u2 = (double *)zalloc(sgsiz, sizeof(double));
  wd->w = (double *)zalloc(sgsiz, sizeof(double));
  wd->u1flux3d = (double *)zalloc(sgsiz, sizeof(double));
  wd->u2flux3d = (double *)zalloc(sgsiz, sizeof(double));
  wd->dzu1 = (double *)zalloc(sgsiz, sizeof(double));
  wd->dzu2 = (double *)zalloc(sgsiz, sizeof(double));
  wd->Kz = (double *)zalloc(sgsiz, sizeof(double));
  wd->eta = (double *)zalloc(S, sizeof(double));
  wd->topz = (double *)zalloc(S, sizeof(double));
  wd->wbot = (double *)zalloc(S, sizeof(double));
  wd->u1flux = (double *)zalloc(S, sizeof(double));
  wd->u2flux = (double *)zalloc(S, sizeof(double));
  wd->u1vm = (double *)zalloc(sgsiz, sizeof(double));
  wd->u2vm = (double *)zalloc(sgsiz, sizeof(double));

  /* win_priv_t: defaults of model/hd/inputs/readparam.c:133-145 */
  wc->trasc = VANLEER;
  wc->ultimate = 0;
  wc->stab = NONE;
  wc->sigma = 0;
  wc->thin_merge = 0;
  {
    double dtop = layers[nz] - layers[nz - 1];
    wc->hmin = (0.1 < 0.07 * dtop) ? 0.1 : 0.07 * dtop;
  }
  const int nt1 = ntr > 0 ? ntr : 1;
  wc->advect = (int *)zalloc(nt1, sizeof(int));
  wc->diffuse = (int *)zalloc(nt1, sizeof(int));
  wc->tbdy = (int *)zalloc(nt1, sizeof(int));
  wc->tdif_v = (int *)zalloc(nt1, sizeof(int));
  wc->mintr = (double *)zalloc(nt1, sizeof(double));
  wc->maxtr = (double *)zalloc(nt1, sizeof(double));
  for (n = 0; n < ntr; n++) {
    wc->advect[n] = 1;
    wc->diffuse[n] = 0;
    wc->tbdy[n] = n;
    wc->tdif_v[n] = n;
    wc->mintr[n] = 0.0;
    wc->maxtr[n] = 1e35;
  }
  wc->ntbdy = ntr;
  wc->ntdif_v = ntr;
  wc->oldeta = (double *)zalloc(S, sizeof(double));
  wc->dz = (double *)zalloc(sgsiz, sizeof(double));
  wc->Ds = (double *)zalloc(S, sizeof(double));
  wc->Hn1 = (double *)zalloc(S, sizeof(double));
  wc->Hs = (double *)zalloc(S, sizeof(double));
  wc->mdx = (double *)zalloc(S, sizeof(double));
  wc->mdy = (double *)zalloc(S, sizeof(double));
  wc->one = (double *)zalloc(S, sizeof(double));
  for (n = 0; n < S; n++) wc->Ds[n] = wc->Hn1[n] = wc->mdx[n] = wc->mdy[n] = wc->one[n] = 1.0;
  wc->u1kh = (double *)zalloc(sgsiz, sizeof(double));
  wc->u2kh = (double *)zalloc(sgsiz, sizeof(double));
  wc->s1 = (int *)zalloc(sgsiz, sizeof(int));
  wc->i1 = (int *)zalloc(S, sizeof(int));
  wc->i2 = (int *)zalloc(S, sizeof(int));
  wc->i3 = (int *)zalloc(S, sizeof(int));
  wc->i4 = (int *)zalloc(S, sizeof(int));
  wc->i5 = (int *)zalloc(S, sizeof(int));
  double **wv[] = {&wc->w1, &wc->w2, &wc->w3, &wc->w4, &wc->w5,
                   &wc->w6, &wc->w7, &wc->w8, &wc->w9, &wc->w10};
  for (n = 0; n < 10; n++) *wv[n] = (double *)zalloc(sgsiz, sizeof(double));
  double **dv[] = {&wc->d1, &wc->d2, &wc->d3, &wc->d4};
  for (n = 0; n < 4; n++) *dv[n] = (double *)zalloc(S, sizeof(double));
  double **vv[] = {&wc->v1, &wc->v2, &wc->v3, &wc->v4, &wc->v5, &wc->v6, &wc->v7};
  for (n = 0; n < 7; n++) *vv[n] = (double *)zalloc(nz + 2, sizeof(double));
  /* the momentum step's face lists: surface faces at the top layer, no
   * column one cell deep (callers set them from the elevation) */
  gw->sur_e1 = (int *)zalloc(S, sizeof(int));
  gw->sur_e2 = (int *)zalloc(S, sizeof(int));
  for (n = 1; n <= gw->a2_e1; n++) gw->sur_e1[n] = gw->w2_e1[n];
  for (n = 1; n <= gw->a2_e2; n++) gw->sur_e2[n] = gw->w2_e2[n];
  wc->cbot_e1 = (int *)zalloc(S, sizeof(int));
  wc->cbot_e2 = (int *)zalloc(S, sizeof(int));
  wc->ncbot_e1 = wc->ncbot_e2 = 0;
  /* face beds: the shallower bed either side of the face (the cell's own
   * where the neighbour is land), and the bottom face of each face
   * column in the layer of that bed */
  gw->botzu1 = (double *)zalloc(S, sizeof(double));
  gw->botzu2 = (double *)zalloc(S, sizeof(double));
  gw->bot_e1 = (int *)zalloc(S, sizeof(int));
  gw->bot_e2 = (int *)zalloc(S, sizeof(int));
  for (n = 1; n <= nv2; n++) {
    const int cs = w2_t[n];
    const int ii = s2i[cs], jj = s2j[cs];
    for (int dir = 1; dir <= 2; dir++) {
      const int in = (dir == 1) ? ii - 1 : ii, jn = (dir == 1) ? jj : jj - 1;
      int kb = kbot[(size_t)jj * nx + ii];
      double bz = gw->botz[cs];
      if (in >= 0 && jn >= 0) {
        const size_t qn = (size_t)jn * nx + in;
        if (kbot[qn] < nz) {
          const int cn = idx3[((size_t)(nz - 1) * ny + jn) * nx + in];
          if (gw->botz[cn] > bz) bz = gw->botz[cn];
          if (kbot[qn] > kb) kb = kbot[qn];
        }
      }
      const int cb = idx3[((size_t)kb * ny + jj) * nx + ii];
      if (dir == 1) {
        gw->botzu1[cs] = bz;
        gw->bot_e1[n] = cb;
      } else {
        gw->botzu2[cs] = bz;
        gw->bot_e2[n] = cb;
      }
    }
  }
  wd->etab = (double *)zalloc(S, sizeof(double));
  wd->wtop = (double *)zalloc(S, sizeof(double));
  wd->detadt = (double *)zalloc(S, sizeof(double));
  wc->nonlinear = 0;
  /* auxiliary_routines: no temperature / salinity / age tracer, nothing
   * to decay, no relaxation (the caller sets them); reset_dz's lists */
  wd->tno = wd->sno = -1;
  wc->ntdec = 0;
  wc->tdec = (int *)zalloc(nt1, sizeof(int));
  wc->dectr = (double *)zalloc(nt1, sizeof(double));
  wc->trinfo_3d = (tracer_info_t *)zalloc(nt1, sizeof(tracer_info_t));
  wc->s2 = (int *)zalloc(sgsiz, sizeof(int));
  wc->c1 = (char *)zalloc(sgsiz, sizeof(char));
  free(gdir);
  free(idx3);
  return G;
#undef WET
}

void emstr_forget_window(geometry_t *g);

static void free_open(geometry_t *g) {
  for (int n = 0; g->open && n < g->nobc; n++) {
    open_bdrys_t *o = g->open[n];
    if (!o) continue;
    void *a[] = {o->nmap, o->obc_t, o->oi1_t, o->oi2_t, o->bot_t, o->obc_e1, o->oi1_e1,
                 o->obc_e2, o->oi1_e2, o->trm, o->bcond_tra, o->clampv, o->trpc,
                 o->relax_zone_tra};
    for (size_t q = 0; q < sizeof(a) / sizeof(a[0]); q++) free(a[q]);
    free(o);
  }
  free(g->open);
  g->open = NULL;
  g->nobc = 0;
}

int emstr_grid_open(emstr_grid *G, int side) {
  geometry_t *g = G->window;
  const int nx = G->nx, ny = G->ny, ntr = G->ntr;
  int cc, n, c;
  const int xs = (side == L_EDGE || side == R_EDGE);
  const int line = (side == L_EDGE || side == B_EDGE) ? 1 : (xs ? nx - 2 : ny - 2);
  if (side != L_EDGE && side != R_EDGE && side != B_EDGE && side != F_EDGE) return -1;
  /* the wet columns on that side; none may be open already */
  char *on = (char *)zalloc(g->sgsizS, 1);
  int nb = 0;
  for (cc = 1; cc <= g->b2_t; cc++) {
    const int cs = g->w2_t[cc];
    if ((xs ? g->s2i[cs] : g->s2j[cs]) != line) continue;
    if (cc > g->v2_t) continue; /* a corner column of an earlier boundary */
    on[cs] = 1;
    nb++;
  }
  if (!nb) {
    free(on);
    return -1;
  }
  /* w2_t: interior columns, the earlier boundaries' columns, this one's;
   * the per-column vectors follow */
  int *w2 = (int *)zalloc(g->sgsizS, sizeof(int));
  int *bt = (int *)zalloc(g->sgsizS, sizeof(int));
  int *st = (int *)zalloc(g->sgsizS, sizeof(int));
  int *ns = (int *)zalloc(g->sgsizS, sizeof(int));
  int m = 0;
  for (int pass = 0; pass < 3; pass++)
    for (cc = 1; cc <= g->b2_t; cc++) {
      const int cs = g->w2_t[cc];
      const int grp = on[cs] ? 2 : (cc > g->v2_t ? 1 : 0);
      if (grp != pass) continue;
      m++;
      w2[m] = cs;
      bt[m] = g->bot_t[cc];
      st[m] = g->sur_t[cc];
      ns[m] = g->nsur_t[cc];
    }
  for (cc = 1; cc <= g->b2_t; cc++) {
    g->w2_t[cc] = w2[cc];
    g->bot_t[cc] = bt[cc];
    g->sur_t[cc] = st[cc];
    g->nsur_t[cc] = ns[cc];
  }
  g->v2_t -= nb;
  /* w3_t in the same grouping */
  int *w3 = (int *)zalloc(g->sgsiz, sizeof(int));
  int nb3 = 0;
  m = 0;
  for (int pass = 0; pass < 3; pass++)
    for (cc = 1; cc <= g->b3_t; cc++) {
      c = g->w3_t[cc];
      const int grp = on[g->m2d[c]] ? 2 : (cc > g->v3_t ? 1 : 0);
      if (grp != pass) continue;
      w3[++m] = c;
      nb3 += (grp == 2);
    }
  for (cc = 1; cc <= g->b3_t; cc++) g->w3_t[cc] = w3[cc];
  g->v3_t -= nb3;
  /* the boundary: surface cells, then each column's cells below */
  open_bdrys_t *o = (open_bdrys_t *)zalloc(1, sizeof(open_bdrys_t));
  o->ocodex = xs ? side : 0;
  o->ocodey = xs ? 0 : side;
  int *in = (side == L_EDGE) ? g->xp1 : (side == R_EDGE) ? g->xm1 : (side == B_EDGE) ? g->yp1 : g->ym1;
  o->nmap = (int *)zalloc(g->sgsiz, sizeof(int));
  memcpy(o->nmap, in, (size_t)g->sgsiz * sizeof(int));
  o->no2_t = nb;
  o->no3_t = nb3;
  const size_t n3 = (size_t)nb3 + 1;
  o->obc_t = (int *)zalloc(n3, sizeof(int));
  o->oi1_t = (int *)zalloc(n3, sizeof(int));
  o->oi2_t = (int *)zalloc(n3, sizeof(int));
  o->bot_t = (int *)zalloc((size_t)nb + 1, sizeof(int));
  int *fb = (int *)zalloc(n3, sizeof(int)), *fi = (int *)zalloc(n3, sizeof(int));
  m = 0;
  for (cc = g->b2_t - nb + 1; cc <= g->b2_t; cc++) {
    o->obc_t[++m] = g->w2_t[cc];
    o->bot_t[m] = g->bot_t[cc];
  }
  for (n = 1; n <= nb; n++) {
    c = o->obc_t[n];
    while (c != o->bot_t[n]) {
      c = g->zm1[c];
      o->obc_t[++m] = c;
    }
  }
  for (cc = 1; cc <= nb3; cc++) {
    c = o->obc_t[cc];
    o->oi1_t[cc] = o->nmap[c];
    o->oi2_t[cc] = o->nmap[o->oi1_t[cc]];
    /* the face on the boundary and the one a cell in (a face is its
     * cell's west / south face) */
    if (side == L_EDGE || side == B_EDGE) {
      fb[cc] = c;
      fi[cc] = o->nmap[c];
    } else {
      fb[cc] = (side == R_EDGE) ? g->xp1[c] : g->yp1[c];
      fi[cc] = c;
    }
  }
  if (xs) {
    o->obc_e1 = fb;
    o->oi1_e1 = fi;
  } else {
    o->obc_e2 = fb;
    o->oi1_e2 = fi;
  }
  const int nt = ntr > 0 ? ntr : 1;
  o->trm = (int *)zalloc(nt, sizeof(int));
  o->bcond_tra = (int *)zalloc(nt, sizeof(int));
  o->clampv = (double *)zalloc(nt, sizeof(double));
  o->trpc = (double *)zalloc(nt, sizeof(double));
  o->relax_zone_tra = (int *)zalloc(nt, sizeof(int));
  for (n = 0; n < ntr; n++) {
    o->bcond_tra[n] = NOGRAD;
    o->trm[n] = -1;
  }
  g->open = (open_bdrys_t **)realloc(g->open, (size_t)(g->nobc + 1) * sizeof(open_bdrys_t *));
  g->open[g->nobc++] = o;
  free(on);
  free(w2);
  free(bt);
  free(st);
  free(ns);
  free(w3);
  return 0;
}

void emstr_grid_free(emstr_grid *G) {
  int n;
  if (!G) return;
  geometry_t *g = G->window;
  if (g) emstr_forget_window(g);
  if (g) free_open(g);
  window_t *w = G->windat;
  win_priv_t *p = G->wincon;
  if (g) {
    void *a[] = {g->layers, g->s2i, g->s2j, g->s2k, g->w3_t, g->w3_e1, g->w3_e2,
                 g->w2_t, g->w2_e1, g->w2_e2, g->bpt, g->bin, g->gsed_t, g->ised_t,
                 g->sur_t, g->nsur_t, g->bot_t, g->m2d, g->xp1, g->xm1, g->yp1,
                 g->ym1, g->zp1, g->zm1, g->h1acell, g->h2acell, g->h1au1, g->h2au2,
                 g->h2au1, g->h1au2, g->cellarea, g->botz, g->gridz, g->cellz,
                 g->sur_e1, g->sur_e2, g->botzu1, g->botzu2, g->bot_e1, g->bot_e2};
    for (n = 0; n < (int)(sizeof(a) / sizeof(a[0])); n++) free(a[n]);
    free(g);
  }
  if (w) {
    for (n = 0; n < w->ntr; n++) free(w->tr_wc[n]);
    void *a[] = {w->tr_wc, w->u1, w->u2, w->w, w->u1flux3d, w->u2flux3d, w->dzu1,
                 w->dzu2, w->Kz, w->eta, w->topz, w->wbot, w->u1flux, w->u2flux,
                 w->u1vm, w->u2vm, w->etab, w->wtop, w->detadt};
    for (n = 0; n < (int)(sizeof(a) / sizeof(a[0])); n++) free(a[n]);
    free(w);
  }
  if (p) {
    void *a[] = {p->advect, p->diffuse, p->tbdy, p->tdif_v, p->mintr, p->maxtr,
                 p->oldeta, p->dz, p->Ds, p->Hn1, p->Hs, p->mdx, p->mdy, p->one,
                 p->u1kh, p->u2kh, p->s1, p->i1, p->i2, p->i3, p->i4, p->i5,
                 p->w1, p->w2, p->w3, p->w4, p->w5, p->w6, p->w7, p->w8, p->w9,
                 p->w10, p->d1, p->d2, p->d3, p->d4, p->v1, p->v2, p->v3, p->v4,
                 p->v5, p->v6, p->v7, p->cbot_e1, p->cbot_e2, p->tdec, p->dectr,
                 p->trinfo_3d, p->s2, p->c1};
    for (n = 0; n < (int)(sizeof(a) / sizeof(a[0])); n++) free(a[n]);
    free(p);
  }
  free(G->kbot);
  free(G);
}

/*
 * Vertical velocity from continuity, bottom up, for the processed
 * columns of set_dz (vel3d.c:4680-4711): w at the bed is wbot, and the
 * flux through the top of each cell below the (lower) surface is the
 * flux through its bottom plus the net lateral inflow.
 */
void emstr_vel_w_update(geometry_t *window, window_t *windat,
                        win_priv_t *wincon) {
  int cc, c, cs;
  memset(windat->w, 0, (size_t)window->sgsiz * sizeof(double));
  for (cc = 1; cc <= wincon->vcs; cc++) {
    c = wincon->i1[cc];
    cs = window->m2d[c];
    windat->w[c] = windat->wbot[cs];
    double fcbot = windat->w[c] * window->cellarea[cs];
    while (c > wincon->i3[cc]) {
      int xp1 = window->xp1[c];
      int yp1 = window->yp1[c];
      int zp1 = window->zp1[c];
      double fctop = fcbot + windat->u1flux3d[c] - windat->u1flux3d[xp1] +
                     windat->u2flux3d[c] - windat->u2flux3d[yp1];
      windat->w[zp1] = fctop / window->cellarea[cs];
      fcbot = fctop;
      c = zp1;
    }
  }
}
