/*
 * Least-squares helpers shared by the COMPAS window builder and the FFSL
 * host code (emstr_lsq.c).
 */
#ifndef EMSTR_LSQ_H
#define EMSTR_LSQ_H

#include "../../include/emstr_us.h"

#ifdef __cplusplus
extern "C" {
#endif

/* B[3 i + a], a = 0..2: the pseudo-inverse of the n rows (1, dx[i], dy[i]) */
void emstr_lsq_pinv3(int n, const double *dx, const double *dy, double *B);
/* ST_SIZED stencil of size 3 of centre c into st (at most max entries);
 * mask: [szcS] zeros, left zero */
int emstr_lsq_stencil(const geometry_t *g, int c, int *st, int max, int *mask);

#ifdef __cplusplus
}
#endif
#endif
