/* C prototypes of the kernel launchers (defined in the .hip sources).
 * All launches are asynchronous on the given stream (hipStream_t as
 * void*); they return 0 or a negative launch error. */
#ifndef EMSTR_KERNELS_H
#define EMSTR_KERNELS_H

#include "emstr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* plain FP64 form (emstr_plain.hip): the FFSL update of tracers n0..n1-1
 * from the trajectory fluxes */
int emstr_launch_adv(const emstr_dev *d, const double *tin, double *tout,
                     const double *tstart, double dtu, double trem, int anydif,
                     int n0, int n1, void *stream);
/* the same update face-once (schemes other than FFSL): the west, south
 * and bottom face fluxes of every cell and tracer into fw, fs, fb
 * ([ntr][nz][ny][nx] each), then the cell updates from them */
int emstr_launch_adv_faces(const emstr_dev *d, const double *tin, double *tout,
                           const double *tstart, double dtu, int anydif, double *fw,
                           double *fs, double *fb, void *stream);
/* FFSL: cell-centred velocities and trajectories of one sub-step, then
 * the fluxes of one tracer (t: its buffer, ts: its step-start buffer);
 * err is set when a trajectory crosses more cells than the walk bound */
int emstr_launch_ffsl_traj(const emstr_dev *d, double dtu, int *err, void *stream);
int emstr_launch_ffsl_fluxes(const emstr_dev *d, const double *t, const double *ts,
                             double dtu, int *err, void *stream);
int emstr_launch_vdiff(const emstr_dev *d, double *tr, double *scratch,
                       int *err, void *stream);
int emstr_launch_setdz(const emstr_dev *d, int *proc, void *stream);
/* auxiliary_routines' state changes (emstr_plain.hip): reset_dz of dz
 * (per-column flag p2: a cell to process for the updated elevation),
 * then, if tracers: the no-gradient relaxation fields, per cell the age,
 * decay and T/S relaxation, the T/S copy above the new surface.  The
 * decay tables, relaxation fields and age mask are device arrays. */
typedef struct AuxArgs {
  int tracers;             /* any tracer-changing part                */
  int age;                 /* age tracer number, -1: none            */
  const double *agemsk;    /* [n3] 0/1                                */
  double dttr;
  int ntdec;
  const int *tdec;         /* [ntdec] tracers to decay, in order      */
  const double *dec;       /* [ntdec] decay constants                 */
  const int *dtr3;         /* [ntdec] 3D tracer of the timescale, -1  */
  const double *dsc;       /* [ntdec] seconds per unit                */
  int tno, sno;            /* -1: absent                              */
  double *rtemp, *rsalt;   /* [n3] relaxation fields (NULL: none)     */
  const double *temp_tc, *salt_tc; /* [n3] time constants (NULL)     */
  double rrT, rrS;         /* relaxation rates                        */
  int tcT, tcS;            /* time-constant relaxation on             */
  int grdT, grdS;          /* no-gradient relaxation field            */
  double depthT, depthS;
  const double *cellz;     /* [n3] cell centres (no-gradient fields)  */
} AuxArgs;
int emstr_launch_aux(const emstr_dev *d, const AuxArgs *a, double *tr, int *p2, void *stream);
int emstr_launch_subeta(const emstr_dev *d, double dtu, double trem, int init,
                        void *stream);
int emstr_launch_osub(const emstr_dev *d, void *stream);
int emstr_launch_dtm(const emstr_dev *d, double dtu0, unsigned long long *out,
                     void *stream);
int emstr_launch_scatter(const int *wsp, const long long *wdn, int nwet,
                         const double *sp, double *dn, void *stream);
int emstr_launch_gather(const int *wsp, const long long *wdn, int nwet,
                        const double *dn, double *sp, void *stream);
int emstr_launch_ghosts(const int *g, const int *src, int n, double *a,
                        void *stream);
int emstr_mass_blocks(const emstr_dev *d);
int emstr_launch_mass(const emstr_dev *d, const double *tr, double *colmass,
                      double *part, void *stream);
/* w from continuity on the device (per processed column) */
int emstr_launch_velw(const emstr_dev *d, const double *wbot, void *stream);
/* set_dz_at_u1 (dir 1) / set_dz_at_u2 (dir 2) into dzu [n3] and the
 * surface face layer of each column into ksur [plane]; kbf [plane] the
 * bottom face layer of each face column (-1: none), botzu [plane] the
 * face beds                                                            */
int emstr_launch_dzu(const emstr_dev *d, int dir, const int *kbf, const double *botzu,
                     double *dzu, int *ksur, void *stream);
/* set_flux_3d of one face direction into out [n3]: wet faces fw, the
 * surface face layer of each column ksur [plane], face velocity u and
 * thickness dzu [n3], face length hf and metric md [plane]; the ncb
 * faces one cell deep (dense indices cb) take uf [plane] / dt; uvm not
 * NULL: the FFSL copy of the mean volume fluxes                        */
int emstr_launch_flux3d(const emstr_dev *d, const unsigned char *fw, const int *ksur,
                        const double *u, const double *dzu, const double *hf, const double *md,
                        const double *uvm, int ncb, const long long *cb, const double *uf,
                        double dt, double *out, void *stream);

/* one open boundary for k_obc: its n cells (dense), one cell inwards,
 * the four inward cells of LINEXT, cellz, the UPSTRM velocity and cell
 * width; per tracer the condition bits, CLAMPD value, trpc and FILEIN
 * row of fv ([row][n]) */
#define OBC_FILEIN 0x000001
#define OBC_NOGRAD 0x000004
#define OBC_LINEXT 0x000008
#define OBC_NOTHIN 0x000800
#define OBC_CLAMPD 0x004000
#define OBC_UPSTRM 0x010000
typedef struct ObcArgs {
  int n, owrite;
  const long long *cell, *in1, *chain;
  const double *cz, *vel, *hat;
  double dt, sc;
  const int *bcond, *frow;
  const double *clampv, *trpc, *fv;
} ObcArgs;
/* tr: the tracers, ts: the step's starting values (read for a lateral
 * ghost among the cells inwards, encoded ~index) */
int emstr_launch_obc(const emstr_dev *d, const ObcArgs *a, double *tr, const double *ts,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif
