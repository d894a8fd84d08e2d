/*
 * emstr_grid.h -- builds the reference's sparse structures for a
 * structured (i,j,k) z-level grid with closed lateral boundaries.
 *
 * The reference builds these structures from a parameter file in
 * build_sparse_map() (model/hd/master/preprocess.c:108-720) and
 * window_build() (model/hd/slaves/windows.c:152, one window).  This
 * builder follows the same rules:
 *   - cells are numbered layer by layer from the surface layer down
 *     (preprocess.c:343-363); within a layer wet cells come in (j,i) order
 *     and are followed by that layer's ghost cells, so the 2D sparse index
 *     of a column equals the 3D index of its surface-layer cell;
 *   - all maps start self-mapping (preprocess.c:608-617); wet neighbours
 *     are linked (preprocess.c:640-675);
 *   - every (wet cell, direction) whose neighbour is not wet gets its own
 *     ghost cell that maps back to the wet cell and is listed in bpt/bin
 *     (preprocess.c:677-720), so a ghost holds the value of the cell it
 *     was reached from after set_lateral_BC_tr (tracers.c:5574-5593);
 *   - each column's bed cell maps down to a sediment ghost that maps to
 *     itself.
 * The caller must give a land ring (every column on the grid edge is land).
 */
#ifndef EMSTR_GRID_H
#define EMSTR_GRID_H

#include "emstr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A built window: the three reference structures and their arrays.   */
typedef struct emstr_grid {
  geometry_t *window;
  window_t *windat;
  win_priv_t *wincon;
  int nx, ny, nz, ntr;
  int *kbot; /* [ny*nx] first wet layer, nz for land                  */
} emstr_grid;

/*
 * nx, ny, nz : grid size (including the land ring)
 * layers     : layer faces layers[0..nz], increasing (layers[nz] = top)
 * botz       : [ny*nx] bed height; a column is land if botz >= layers[nz]
 * h1acell ... cellarea : [ny*nx] metrics at every (i,j) (land included)
 * ntr        : number of tracers
 * Returns NULL on bad arguments (message in emstr_last_error()).
 */
emstr_grid *emstr_grid_build(int nx, int ny, int nz, const double *layers,
                             const double *botz, const double *h1acell,
                             const double *h2acell, const double *h1au1,
                             const double *h2au2, const double *h2au1,
                             const double *h1au2, const double *cellarea,
                             int ntr);
void emstr_grid_free(emstr_grid *g);

/* Makes the wet columns on one side of the grid an open boundary:
 * side L_EDGE (columns i = 1), R_EDGE (i = nx - 2), B_EDGE (j = 1) or
 * F_EDGE (j = ny - 2).  The boundary columns move behind the interior
 * ones in w2_t (v2_t < b2_t) and their cells in w3_t (v3_t < b3_t); the
 * boundary's cells are the surface cells then each column's cells
 * below, nmap the inward map; every tracer starts as NOGRAD.  A column
 * of an earlier boundary (a corner) stays with it.  Returns 0, or -1
 * when no wet column of that side is left. */
int emstr_grid_open(emstr_grid *g, int side);

/* w from continuity, bottom up: model/hd/momentum/vel3d.c:4609-4720,
 * without sources, relaxation or precipitation.  Needs set_dz() first.  */
void emstr_vel_w_update(geometry_t *window, window_t *windat,
                        win_priv_t *wincon);

#ifdef __cplusplus
}
#endif
#endif
