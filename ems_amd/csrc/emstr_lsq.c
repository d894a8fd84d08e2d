/*
 * Least-squares helpers of the COMPAS path (libemstr_us):
 *
 *   emstr_lsq_pinv3    rows of the pseudo-inverse of an n x 3 matrix with
 *                      rows (1, dx, dy): the reference computes these from
 *                      a singular value decomposition (svd_lsq_B,
 *                      lib/interp/lsq/svd.c:593-660); here from the
 *                      eigen-decomposition of the 3x3 normal matrix
 *                      (cyclic Jacobi), with singular values below 1e-12
 *                      of the largest cut as the SVD cuts zero ones.
 *   emstr_us_vlsq_build  the streamline velocity fit of an FFSL window.
 *
 * The velocity fit (tran_grid_init, model/hd-us/tracers/transport.c:
 * 4317-4658, with the default interpolation rule L_LSLIN, "linearlsq"):
 * every wet centre of layer k is a point of that layer's interpolation
 * structure at its centre position, every lateral ghost a point at the
 * midpoint of the edge joining it to its wet centre.  The points a
 * centre's fit uses (set_interp_points_n, transport.c:5537-5640, size 3)
 * are the centres of its ST_SIZED stencil (the centre, then the wet
 * centres around its vertices in its layer) followed by, for each of
 * those in order, the ghosts across its edges.  lsql_build
 * (lib/interp/lsq_lin/lsql.c:47-210) fits t = c0 + c1 (x - xr) + c2
 * (y - yr) at those points, xr, yr the centre's own position, through
 * the pseudo-inverse B of the rows (1, x - xr, y - yr); lsql_rebuild
 * (:262-287) sets w[1], w[2] to sum_j v_j B[j][1..2] and w[0] to the
 * centre's own value, and lsql_interpolate_point (:306-330) returns
 * w[0] + beta (w[1] (x - xr) + w[2] (y - yr)) with beta 1.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/emstr_us.h"
#include "emstr_lsq.h"

static void jacobi3(double A[3][3], double V[3][3], double ev[3]) {
  int i, j, p, q, sweep;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (sweep = 0; sweep < 50; sweep++) {
    double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    if (off == 0.0) break;
    for (p = 0; p < 2; p++)
      for (q = p + 1; q < 3; q++) {
        if (A[p][q] == 0.0) continue;
        double th = 0.5 * (A[q][q] - A[p][p]) / A[p][q];
        double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (i = 0; i < 3; i++) {
          double aip = A[i][p], aiq = A[i][q];
          A[i][p] = c * aip - sn * aiq;
          A[i][q] = sn * aip + c * aiq;
        }
        for (i = 0; i < 3; i++) {
          double api = A[p][i], aqi = A[q][i];
          A[p][i] = c * api - sn * aqi;
          A[q][i] = sn * api + c * aqi;
        }
        for (i = 0; i < 3; i++) {
          double vip = V[i][p], viq = V[i][q];
          V[i][p] = c * vip - sn * viq;
          V[i][q] = sn * vip + c * viq;
        }
      }
  }
  for (i = 0; i < 3; i++) ev[i] = A[i][i];
}

void emstr_lsq_pinv3(int n, const double *dx, const double *dy, double *B) {
  double A[3][3] = {{0}}, V[3][3], ev[3], Ainv[3][3] = {{0}};
  int i, a, b, l;
  for (i = 0; i < n; i++) {
    const double r[3] = {1.0, dx[i], dy[i]};
    for (a = 0; a < 3; a++)
      for (b = 0; b < 3; b++) A[a][b] += r[a] * r[b];
  }
  jacobi3(A, V, ev);
  const double emax = fmax(fabs(ev[0]), fmax(fabs(ev[1]), fabs(ev[2])));
  for (a = 0; a < 3; a++)
    for (b = 0; b < 3; b++)
      for (l = 0; l < 3; l++)
        if (fabs(ev[l]) > 1e-12 * emax) Ainv[a][b] += V[a][l] * V[b][l] / ev[l];
  for (i = 0; i < n; i++) {
    const double r[3] = {1.0, dx[i], dy[i]};
    for (a = 0; a < 3; a++) B[3 * i + a] = Ainv[a][0] * r[0] + Ainv[a][1] * r[1] + Ainv[a][2] * r[2];
  }
}

/* ST_SIZED stencil of size 3 of centre c: the centre, then the wet
 * centres around its vertices in its layer, each once; returns the
 * count (at most max) */
int emstr_lsq_stencil(const geometry_t *g, int c, int *st, int max, int *mask) {
  const int cs = g->m2d[c];
  int ns = 0, j, i;
  st[ns++] = c;
  mask[cs] = 1;
  for (j = 1; j <= g->npe[cs]; j++) {
    const int v = g->c2v[j][c];
    if (!v) continue;
    const int vs = g->m2dv[v];
    for (i = 1; i <= g->nvc[vs]; i++) {
      const int c1 = g->v2c[v][i];
      if (c1 && !mask[g->m2d[c1]] && ns < max) {
        st[ns++] = c1;
        mask[g->m2d[c1]] = 1;
      }
    }
  }
  for (i = 0; i < ns; i++) mask[g->m2d[st[i]]] = 0;
  return ns;
}

void emstr_us_vlsq_free(emstr_us_vlsq *t) {
  if (!t) return;
  free(t->vo);
  free(t->vn);
  free(t->cell);
  free(t->edge);
  free(t->bx);
  free(t->by);
  memset(t, 0, sizeof(*t));
}

int emstr_us_vlsq_build(const geometry_t *g, emstr_us_vlsq *t) {
  enum { MAXS = 64 };
  int cc, i, n;
  if (!g || !t || !g->c2v || !g->v2c || !g->nvc || !g->m2dv || !g->u1x || !g->u1y) return -1;
  memset(t, 0, sizeof(*t));
  t->szc = g->szc;
  t->vo = (int *)calloc((size_t)g->szc, sizeof(int));
  t->vn = (int *)calloc((size_t)g->szc, sizeof(int));
  int *mask = (int *)calloc((size_t)g->szcS, sizeof(int));
  int *st = (int *)malloc(MAXS * sizeof(int));
  const int emax = MAXS * (g->npem + 1);
  int *pc = (int *)malloc((size_t)emax * sizeof(int));
  int *pe = (int *)malloc((size_t)emax * sizeof(int));
  double *px = (double *)malloc((size_t)emax * sizeof(double));
  double *py = (double *)malloc((size_t)emax * sizeof(double));
  double *B = (double *)malloc((size_t)emax * 3 * sizeof(double));
  size_t cap = (size_t)g->a3_t * 8 + 16, used = 0;
  t->cell = (int *)malloc(cap * sizeof(int));
  t->edge = (int *)malloc(cap * sizeof(int));
  t->bx = (double *)malloc(cap * sizeof(double));
  t->by = (double *)malloc(cap * sizeof(double));
  int ok = t->vo && t->vn && mask && st && pc && pe && px && py && B && t->cell && t->edge &&
           t->bx && t->by;
  for (cc = 1; ok && cc <= g->a3_t; cc++) {
    const int c = g->w3_t[cc], cs = g->m2d[c];
    const double xr = g->cellx[cs], yr = g->celly[cs];
    const int ns = emstr_lsq_stencil(g, c, st, MAXS, mask);
    int np = 0;
    /* the stencil's centres, then the ghosts across their edges */
    for (i = 0; i < ns; i++) {
      pc[np] = st[i];
      pe[np] = 0;
      px[np] = g->cellx[g->m2d[st[i]]] - xr;
      py[np] = g->celly[g->m2d[st[i]]] - yr;
      np++;
    }
    for (i = 0; i < ns; i++) {
      const int cn = st[i], cns = g->m2d[cn];
      for (n = 1; n <= g->npe[cns]; n++) {
        const int cg = g->c2c[n][cn];
        if (!g->wgst[cg]) continue;
        const int e = g->c2e[n][cn], es = g->m2de[e];
        pc[np] = cg;
        pe[np] = e;
        px[np] = g->u1x[es] - xr;
        py[np] = g->u1y[es] - yr;
        np++;
      }
    }
    emstr_lsq_pinv3(np, px, py, B);
    if (used + (size_t)np > cap) {
      cap = 2 * cap + (size_t)np;
      t->cell = (int *)realloc(t->cell, cap * sizeof(int));
      t->edge = (int *)realloc(t->edge, cap * sizeof(int));
      t->bx = (double *)realloc(t->bx, cap * sizeof(double));
      t->by = (double *)realloc(t->by, cap * sizeof(double));
      if (!t->cell || !t->edge || !t->bx || !t->by) {
        ok = 0;
        break;
      }
    }
    t->vo[c] = (int)used;
    t->vn[c] = np;
    for (i = 0; i < np; i++) {
      t->cell[used + i] = pc[i];
      t->edge[used + i] = pe[i];
      t->bx[used + i] = B[3 * i + 1];
      t->by[used + i] = B[3 * i + 2];
    }
    used += (size_t)np;
  }
  t->nent = (int)used;
  free(mask);
  free(st);
  free(pc);
  free(pe);
  free(px);
  free(py);
  free(B);
  if (!ok) {
    emstr_us_vlsq_free(t);
    return -1;
  }
  return 0;
}
