/*
 * Builder of z-level COMPAS windows (libemstr_us): the structures of
 * emstr_us.h with the preprocessor's numbering and maps, from a 2-D mesh
 * given as cells, their edge counts and their neighbours.  Used by the
 * tests, the benchmark and the smoke run to make seeded meshes.
 */
#ifndef EMSTR_MESH_H
#define EMSTR_MESH_H

#include "emstr_us.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emstr_mesh {
  geometry_t *window;
  window_t *windat;
  win_priv_t *wincon;
  int nmc, npem, nz, ntr;
  int *kbot;  /* [nmc] bed layer of each mesh cell (nz: land)          */
  int *idx3;  /* [nz][nmc] centre of (layer, mesh cell), 0 if dry      */
  int *gidx;  /* [nz][nmc][npem] ghost across edge j of a wet centre   */
  int *eidx;  /* [nz][nmc][npem] edge j of a wet centre                */
  int *e2m;   /* [sze] mesh cell * npem + j - 1 of e2c[e][0]'s edge    */
  int *sedg;  /* [nmc] sediment ghost below the bed                    */
  int nv;     /* mesh vertices (emstr_mesh_vertices)                   */
  int *v2m;   /* [szvS] surface vertex -> mesh vertex                  */
} emstr_mesh;

/* nbr[m*npem + j-1]: mesh cell across edge j of cell m (-1: none);
 * nbrj: the index of that edge among the neighbour's edges (1-based).
 * kbot[m]: bed layer (nz for land).  layers[0..nz]: layer faces.     */
emstr_mesh *emstr_mesh_build(int nmc, int npem, const int *npe, const int *nbr,
                             const int *nbrj, const int *kbot, int nz,
                             const double *layers, int ntr);
void emstr_mesh_free(emstr_mesh *m);
/* vertex maps for FFSL from vid[m*npem + j-1], the mesh vertex where
 * edge j of cell m starts (counter-clockwise), 0 <= vid < nv; the
 * caller then fills gridx/gridy (through v2m), u1x/u1y, costhu2/sinthu2
 * and the ghost centre positions                                     */
int emstr_mesh_vertices(emstr_mesh *m, int nv, const int *vid);
/* least-squares weights wincon->lw from the centre positions         */
int emstr_mesh_lsq(emstr_mesh *m);

#ifdef __cplusplus
}
#endif
#endif
