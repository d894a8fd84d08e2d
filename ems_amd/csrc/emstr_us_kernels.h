/* C prototypes of the COMPAS kernel launchers (emstr_us.hip); the column
 * launchers of the plain form (emstr_kernels.h) are shared.  Launches are
 * asynchronous on the given stream; they return 0 or a negative error. */
#ifndef EMSTR_US_KERNELS_H
#define EMSTR_US_KERNELS_H

#include "emstr_kernels.h"
#include "emstr_usdev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* advection, horizontal diffusion, update, surf_concrk and tr_bounds of
 * every tracer over one sub-step (van Leer or QUICKEST with optional
 * ULTIMATE, or first order): the fluxes of the nesp wet edges (dense
 * indices edn) into fx [ntr][nz][ne2], the vertical fluxes into fz
 * [ntr][nz][ncol], then the cell updates */
int emstr_us_launch_adv(const emstr_usdev *D, const double *tin, double *tout,
                        const double *tstart, const long long *edn, int nesp, double *fx,
                        double *fz, double dtu, int anydif, int ult, void *stream);
/* minimum sub-step; *out must hold the bits of dtu0 before the launch */
int emstr_us_launch_dtm(const emstr_usdev *D, double dtu0, unsigned long long *out,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif
