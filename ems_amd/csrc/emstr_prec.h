/* Precision (FP32) and batched forms: internal interface used by
 * emstr_host.c.  Implemented in emstr_prec.hip. */
#ifndef EMSTR_PREC_H
#define EMSTR_PREC_H

#include "emstr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

struct emstr_prec;
struct emstr_prof;

struct emstr_prec *emstr_prec_create(const emstr_dev *d, void *stream);
void emstr_prec_destroy(struct emstr_prec *p);
void emstr_prec_set_prof(struct emstr_prec *p, struct emstr_prof *prof);
/* Fields or set_dz results changed: rebuild the FP32 coefficients. */
int emstr_prec_fields(struct emstr_prec *p, const emstr_dev *d, void *stream);
/* Tracer n: device sparse FP64 array <-> FP32 dense state. */
int emstr_prec_put_tracer(struct emstr_prec *p, const emstr_dev *d, int n,
                          const double *sparse, void *stream);
int emstr_prec_get_tracer(struct emstr_prec *p, const emstr_dev *d, int n,
                          double *sparse, void *stream);
int emstr_prec_advect(struct emstr_prec *p, const emstr_dev *d, int stab,
                      void *stream);
int emstr_prec_vdiff(struct emstr_prec *p, const emstr_dev *d, int *err,
                     void *stream);
int emstr_prec_step(struct emstr_prec *p, const emstr_dev *d, int stab,
                    int *err, void *stream);
/* Per-block (sum, err) FP64 mass partials, as emstr_launch_mass. */
int emstr_prec_mass(struct emstr_prec *p, const emstr_dev *d, double *part,
                    void *stream);

#ifdef __cplusplus
}
#endif
#endif
