/* Batched form (all tracers and columns in one fused kernel): internal
 * interface used by emstr_host.c.  Implemented in emstr_batch.hip. */
#ifndef EMSTR_BATCH_H
#define EMSTR_BATCH_H

#include "emstr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

struct emstr_batch;
struct emstr_prof;

struct emstr_batch *emstr_batch_create(const emstr_dev *d, const int *h_kb, void *stream);
void emstr_batch_destroy(struct emstr_batch *b);
void emstr_batch_set_prof(struct emstr_batch *b, struct emstr_prof *p);
/* column tile shape and persistent grid (0 keeps the current value) */
int emstr_batch_tune(struct emstr_batch *b, const int *h_kb, int tile_x, int tile_y, int grid);
int emstr_batch_fields(struct emstr_batch *b, const emstr_dev *d, void *stream);
/* nsteps tracer steps; with use_graph pairs of steps replay a hipGraph */
int emstr_batch_steps(struct emstr_batch *b, const emstr_dev *d, int stab, int nsteps,
                      void *stream, int use_graph);
void emstr_batch_shape(const struct emstr_batch *b, int *tile_x, int *tile_y, int *grid,
                       int *grid_full);
int emstr_batch_record2(struct emstr_batch *b, const emstr_dev *d, int stab, void *stream);
int emstr_batch_put(struct emstr_batch *b, const emstr_dev *d, int n, const double *sparse,
                    void *stream);
int emstr_batch_get(struct emstr_batch *b, const emstr_dev *d, int n, double *sparse,
                    void *stream);
int emstr_batch_mass(struct emstr_batch *b, const emstr_dev *d, double *part, void *stream);

#ifdef __cplusplus
}
#endif
#endif
