/*
 * Device-side view of one window (C POD shared by the host C layer and the
 * HIP kernels).  The GPU keeps the window in a dense layout:
 *   3D arrays are [nz][ny][nx] with x fastest (one coalesced plane per
 *   layer); tracers are [ntr][nz][ny][nx] (structure of arrays);
 *   2D arrays are [ny][nx].
 * Cells below the bed (k < kb[col]) and land columns (kb = nz) are never
 * read or written.  The reference's ghost cells are not stored: a
 * neighbour that is not wet is read as the value of the cell the stencil
 * walked from, which is what a ghost holds after set_lateral_BC_tr()
 * (model/hd/tracers/tracers.c:5574-5593) and the bed copy (tracers.c:827).
 */
#ifndef EMSTR_DEV_H
#define EMSTR_DEV_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emstr_dev {
  int nx, ny, nz, ntr;
  long long plane; /* nx*ny    */
  long long n3;    /* nz*plane */
  int trasc, ultimate, sigma;
  int us;       /* COMPAS window: set_dz processes a column whose water top
                   is >= (not >) DRY_FRAC*hmin above the bed            */
  double dt, hmin;

  /* per column [ny][nx] */
  int *kb;      /* bed layer, nz for land                          */
  int *ks;      /* layer of the processed surface cell s1, -1 if the
                   column is not processed                          */
  int *kold;    /* layer of the old surface (i3)                    */
  int *knew;    /* layer of the new surface (i2)                    */
  double *h1au1, *h2au2, *h1acell, *h2acell, *h2au1, *h1au2;
  double *cellarea, *botz, *eta, *oldeta;
  double *osubeta, *subeta; /* per sub-step surface elevations     */
  int *obcm;    /* [ny][nx] 1 for an open-boundary column (not advected,
                   tracers.c:811-822); NULL: none                     */
  double *layers;           /* [nz+1] layer faces                  */

  /* 3D fields [nz][ny][nx] */
  double *u1, *u2, *w, *u1flux3d, *u2flux3d, *Kz, *dz;
  double *dzg; /* [ny][nx] dz of the column's sediment ghost (layer kb-1,
                  written by reset_dz, copied up by set_dz); NULL: 0      */
  double *dzu1, *dzu2, *u1kh, *u2kh; /* horizontal diffusion        */

  /* per tracer [ntr] */
  double *mintr, *maxtr;
  int *advect, *diffuse, *in_tbdy, *in_tdif;

  /* processed columns: dense 2D index in s1 order, ncol of them      */
  int ncol;
  int *cols;

  /* sparse <-> dense for wet cells: wsp[i] (1-based sparse index) and
   * wdn[i] (dense 3D index), i < nwet                               */
  int nwet, sgsiz;
  int *wsp;
  long long *wdn;

  /* FFSL (trasc == FFSL only): cell-centred velocities, w9, trajectory
   * heads (positions, see emstr_plain.hip) and fractions of the x, y, z
   * faces and x, y, z centres, and the one-tracer work arrays          */
  int nstep0; /* windat->nstep == 0: no horizontal divergence (FFSL) */
  double *fnu, *fnv, *fnw, *fw9;
  long long *fcl; /* [6][n3] */
  double *fcrf;   /* [6][n3] */
  double *ftmx, *ftmy, *ftmz, *ftm, *ffx, *ffy, *ffz;
} emstr_dev;

#ifdef __cplusplus
}
#endif
#endif
