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

/* one Runge-Kutta stage of a sub-step: fluxes of tin into fx/fz
 * (scratch, [ntr][nz][ne2] and [ntr][n3]), then the cell update
 * tin -> tout with runge_kutta's weights rkw (of the sub-step's starting
 * values ts[0], of the stage's values tin, of the tendency); tr_bounds
 * follows the last stage (last != 0), the layers above the surface keep
 * their values before it (tstart: step-start values for ghosts).  With 5
 * stages, mode 1 (stage 3) stores the tendency in g3 [ntr][n3] and mode 2
 * (stage 4) combines the values of stages 2 and 3 (ts[1], ts[2]), g3, the
 * stage's values and tendency with rkw[0..4] */
int emstr_us_launch_adv(const emstr_usdev *D, const double *const *ts, const double *tin,
                        double *tout, const double *tstart, const long long *edn, int nesp,
                        double *fx, double *fz, double dtu, int anydif, int ult,
                        const double *rkw, int last, int mode, double *g3, void *stream);
/* minimum sub-step; *out must hold the bits of dtu0 before the launch;
 * sf, gu, gv (ghost velocities [nz][ne2]) are read under FFSL only     */
int emstr_us_launch_dtm(const emstr_usdev *D, double dtu0, unsigned long long *out, double sf,
                        const double *gu, const double *gv, void *stream);
/* the cell update, surf_concrk and tr_bounds from fluxes fx, fz; with
 * bmn, bmx (not NULL) the FFSL CLIP of clip_ffsl before tr_bounds, the
 * centre ranges from emstr_us_launch_ffsl_bounds (one stage only) */
int emstr_us_launch_update(const emstr_usdev *D, const double *const *ts, const double *tin,
                           double *tout, const double *fx, const double *fz, double dtu,
                           const double *rkw, int last, int mode, double *g3,
                           const double *bmn, const double *bmx, void *stream);
/* The FP32 form: one forward-Euler update of the FP64 tracers from FP32
 * edge and vertical fluxes (bmn, bmx as above). */
int emstr_us_launch_update32(const emstr_usdev *D, const double *t0, const double *tin,
                             double *tout, const float *fx, const float *fz, double dtu,
                             const double *rkw, const double *bmn, const double *bmx,
                             void *stream);
/* FFSL CLIP: the range of every wet centre's tracer values over its
 * least-squares stencil (itself first; ghosts from tstart) into bmn, bmx
 * [ntr][n3]; err gets 3 where a column with water is not processed */
int emstr_us_launch_ffsl_bounds(const emstr_usdev *D, const double *tin, const double *tstart,
                                double *bmn, double *bmx, int *err, void *stream);

/* FFSL (emstr_us_ffsl.hip).  Per step: the centred w into nw [nz][ncol]
 * (init: vel_center_w, else the Leonard et al. choice).  Per sub-step:
 * the vertical departure layers / fractions of centres (clc, crc) and of
 * the faces of the cells to process (clf, crf), their Courant numbers,
 * and the streamline of every wet edge (lci [nz][ne2], ld [5][nz][ne2]).
 * Per sub-step and tracer: tm (rows [nwet][tqs]), lq [4][ntr][nz*ncol]
 * (w1, w2, beta, tr_mod), te, moc, then fx and fz.  need [nz][ncol]
 * gets stamp (distinct per call) at the centres whose tr_moc the vertical
 * fluxes read; te and moc are computed for those only.  err
 * gets 1 for a streamline that leaves its cell into a wet cell, 2 for a
 * vertical walk that does not end.                                     */
int emstr_us_launch_ffsl_step(const emstr_usdev *D, int init, double *nw, void *stream);
int emstr_us_launch_ffsl_vfit(const emstr_usdev *D, void *stream);
int emstr_us_launch_ffsl_sub(const emstr_usdev *D, double dtu, const double *nw, int *clc,
                             double *crc, int *clf, double *crf, double *cour,
                             const long long *edn, int nesp, int *lci, double *ld, int *err,
                             void *stream);
int emstr_us_launch_ffsl_flux(const emstr_usdev *D, double dtu, const double *tin,
                              const double *nw, const int *clc, const double *crc,
                              const int *clf, const double *crf, const double *cour,
                              const long long *edn, int nesp, const int *lci, const double *ld,
                              double *tm, double *lq, double *te, double *moc, double *fx,
                              double *fz, int anydif, int *need, int stamp, int *err,
                              void *stream);
int emstr_us_launch_ffsl_flux32(const emstr_usdev *D, double dtu, const double *tin,
                                const double *nw, const int *clc, const double *crc,
                                const int *clf, const double *crf, const double *cour,
                                const long long *edn, int nesp, const int *lci, const double *ld,
                                float *tm, float *lq, float *te, float *moc, float *fx,
                                float *fz, int anydif, int *need, int stamp, int *err,
                                void *stream);

#ifdef __cplusplus
}
#endif
#endif
