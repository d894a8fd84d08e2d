/*
 * emstr.h -- C-ABI of the MI355X tracer-transport step for the EMS
 * structured (SHOC) model.
 *
 * The three structures below carry the fields of the reference's
 * geometry_t (model/hd/include/sparse.h:660-1004), win_priv_t
 * (sparse.h:246-654) and window_t (sparse.h:2393-2794) that the tracer
 * step reads or writes, under the same names and with the same meaning:
 * 1-based sparse arrays (index 0 unused), layer-major cell numbering with
 * the surface layer first, self-mapping neighbour maps at boundaries and a
 * separate ghost cell per (wet cell, direction).  Host code that already
 * fills the reference structures can be compiled against this header (or
 * compile emstr_host.c against the reference's sparse.h) because every
 * access is by field name.
 *
 * Entry points keep the reference signatures:
 *   tracer_step_3d   model/hd/tracers/tracers.c:239
 *   advect_diffuse   model/hd/tracers/tracers.c:401
 *   vert_diffuse_3d  model/hd/tracers/tracers.c:3782
 *   set_dz           model/hd/tracers/tracers.c:5682
 * They run on the GPU.  The reference aborts through hd_quit(); these
 * entry points never exit: they record an error that emstr_last_error()
 * returns, and advect_diffuse() returns non-zero as the reference does.
 */
#ifndef EMSTR_H
#define EMSTR_H

#ifdef __cplusplus
extern "C" {
#endif

/* Advection scheme flags: model/hd/include/hd_params.h:331-350 */
#define ORDER1 0x000002
#define ORDER2 0x000004
#define QUICKEST_AD 0x000008
#define ORDER4 0x000010
#define QUICKEST 0x000020
#define QUICKEST_CO 0x000040
#define VANLEER 0x000080
#define LAGRANGE 0x000100
#define ORDER2_UW 0x002000
#define FFSL 0x080000

/* Stability flags: hd_params.h:236-242; NONE: hd_params.h:104 */
#define ORIGINAL 1
#define SUB_STEP 2
#define SUB_STEP_NOSURF 4
#define SUB_STEP_TRACER 16
#define MULTI_DT 32
#ifndef NONE
#define NONE 0x80000000
#endif

/* Fraction of hmin above the bed for a column to be processed:
 * hd_params.h:42 */
#define DRY_FRAC 0.05

/* Tracer attribute flags read by auxiliary_routines: decay timescale
 * from a 2D / 3D tracer (hd_params.h:674-675), relaxation time-constant
 * types and the no-gradient relaxation field (hd_params.h:705-714) */
#define DE_TR2 0x0400
#define DE_TR3 0x0800
#define RLX_ADPT 0x0008
#define RLX_GRD 0x0400
#define RLX_REG 0x0800
#define RLX_OBC 0x1000

/* Tracer attributes: the fields of the reference tracer_info_t
 * (lib/include/tracer.h:50-125) that auxiliary_routines reads.  The
 * units string is shorter than the reference's MAXSTRLEN; only "day",
 * "days", "hour", "hours", "minute", "minutes" are told apart. */
#define EMSTR_UNITS_LEN 64
typedef struct tracer_info {
  char units[EMSTR_UNITS_LEN];
  int flag;          /* DE_TR2 / DE_TR3, RLX_GRD                         */
  double relax_rate; /* relaxation rate (1/s)                            */
  double relax_dum;  /* depth below which the relaxation field is kept  */
  int tctype;        /* relaxation time-constant type (RLX_*)            */
} tracer_info_t;

/*-------------------------------------------------------------------*/
/* Open boundary (the fields of the reference open_bdrys_t,          */
/* model/hd/include/boundary.h:203-318, that the tracer step reads). */
/* Built: tracer conditions NOGRAD (with the depth trpc), NOTHIN,    */
/* CLAMPD, LINEXT, FILEIN and UPSTRM, OP_OWRITE; other conditions,   */
/* ghost zones (bgz), flow relaxation zones and scaling are refused. */
/*-------------------------------------------------------------------*/
#define L_EDGE 0x00000080 /* model/hd/include/hd_params.h:53-60     */
#define R_EDGE 0x00000100
#define B_EDGE 0x00000400
#define F_EDGE 0x00000800
#define FILEIN 0x000001 /* hd_params.h:136-168 */
#define CUSTOM 0x000002
#define NOGRAD 0x000004
#define LINEXT 0x000008
#define POLEXT 0x000010
#define CYCLIC 0x000400
#define NOTHIN 0x000800
#define CLAMPD 0x004000
#define STATIS 0x008000
#define UPSTRM 0x010000
#define PROFIL 0x100000
#define DEPROF 0x200000
#define DESCAL 0x400000
#define TRFLUX 0x1000000
#define TRCONC 0x2000000
#define TRCONF 0x4000000
#define OP_OWRITE 0x20000 /* hd_params.h:212 */
#define OP_OBCCM 0x40000
#define CENTER 1 /* hd_params.h:411-414 (upmeth) */
#define FACE 2
#define ADAPTIVE 8
typedef struct open_bdrys {
  int ocodex, ocodey; /* L_EDGE / R_EDGE, B_EDGE / F_EDGE            */
  int *nmap;          /* [sgsiz] map one cell inwards, normal to it  */
  int no3_t, no2_t;   /* boundary cells (3D), of them in the surface */
  int *obc_t;         /* [1..no3_t] boundary cells, surface first    */
  int *oi1_t, *oi2_t; /* [1..no3_t] one and two cells inwards        */
  int *cyc_t;         /* [1..no3_t] cyclic partners (CYCLIC: refused) */
  int *bot_t;         /* [1..no2_t] bed cell of each boundary column */
  int *obc_e1, *oi1_e1; /* [1..no3_t] u1 face on the boundary, one in */
  int *obc_e2, *oi1_e2; /* [1..no3_t] u2 face on the boundary, one in */
  double **t_transfer; /* FILEIN values: t_transfer[trm[n]][1..no3_t] */
  int bgz;            /* ghost zone (refused above 0)                */
  int *trm;           /* [ntr] row of t_transfer of a FILEIN tracer  */
  int upmeth;         /* UPSTRM velocities: CENTER, FACE, ADAPTIVE or
                         0 (one cell inwards)                        */
  int *bcond_tra;     /* [ntr] tracer condition bits                 */
  double *clampv;     /* [ntr] CLAMPD value                          */
  double *trpc;       /* [ntr] depth splitting NOGRAD (above) from
                         FILEIN (below); 0: none                     */
  int *relax_zone_tra; /* [ntr] flow relaxation zone (refused above 0) */
  int options;        /* OP_OWRITE; OP_OBCCM refused                 */
} open_bdrys_t;

/*-------------------------------------------------------------------*/
/* Grid structure (fields of the reference geometry_t)              */
/*-------------------------------------------------------------------*/
typedef struct geometry {
  int sgnum;      /* number of 3D sparse cells                      */
  int sgnumS;     /* number of 2D sparse cells                      */
  int sgsiz;      /* sgnum + 1                                      */
  int sgsizS;     /* sgnumS + 1                                     */
  int nobc;       /* open boundaries (0: closed basin)              */
  open_bdrys_t **open; /* [nobc]                                    */
  int nce1, nce2; /* grid dimensions                                */
  int nz;         /* number of layers                               */
  int wn;         /* window number                                  */
  double *layers; /* layer faces, layers[0..nz] (bottom to top)     */

  int *s2i, *s2j, *s2k; /* sparse -> (i,j,k)                          */

  /* 3D cells-to-process vectors: [1..v3] wet, [..b3] +OBC,
   * [..a3] +auxiliary, [..n3] +ghosts                                */
  int *w3_t, *w3_e1, *w3_e2;
  int n3_t, n3_e1, n3_e2;
  int v3_t, v3_e1, v3_e2;
  int b3_t, b3_e1, b3_e2;
  int a3_t, a3_e1, a3_e2;
  /* 2D cells-to-process vectors                                     */
  int *w2_t, *w2_e1, *w2_e2;
  int n2_t, n2_e1, n2_e2;
  int v2_t, v2_e1, v2_e2;
  int b2_t, b2_e1, b2_e2;
  int a2_t, a2_e1, a2_e2;

  /* Lateral ghost cells: tr[bpt[cc]] = tr[bin[cc]], cc = 1..nbpt    */
  int nbpt;
  int *bpt, *bin;
  /* Sediment ghosts below the bed and their interior cells          */
  int ngsed;
  int *gsed_t, *ised_t;

  /* Surface / bottom vectors (size a2_t + 1)                        */
  int *sur_t, *nsur_t, *bot_t;
  int *m2d; /* 3D -> 2D (surface layer) sparse map                     */

  /* Neighbour maps (self-mapping where no neighbour exists)         */
  int *xp1, *xm1, *yp1, *ym1, *zp1, *zm1;

  /* Horizontal metrics (2D sparse arrays)                           */
  double *h1acell, *h2acell, *h1au1, *h2au2, *h2au1, *h1au2;
  double *cellarea;
  /* Vertical geometry                                               */
  double *botz;  /* 2D: bed height (negative down)                    */
  double *gridz; /* 3D: height of the lower face of each cell         */
  double *cellz; /* 3D: height of the cell centre                     */
  /* Surface e1 / e2 face of each column of w2_e1 / w2_e2 (size
   * a2_e1 + 1 / a2_e2 + 1; set by the momentum step from the higher
   * elevation either side of the face), read by set_flux_3d          */
  int *sur_e1, *sur_e2;
  /* Bed height of each e1 / e2 face (2D) and the bottom face of each
   * column of w2_e1 / w2_e2 (size a2_e1 + 1 / a2_e2 + 1), read by
   * set_dz_at_u1 / u2                                                 */
  double *botzu1, *botzu2;
  int *bot_e1, *bot_e2;
} geometry_t;

/*-------------------------------------------------------------------*/
/* Window data (fields of the reference window_t)                   */
/*-------------------------------------------------------------------*/
typedef struct window {
  int ntr;        /* number of 3D tracers                             */
  double **tr_wc; /* tr_wc[n][1..sgnum]                               */
  double dt;      /* 3D time step                                     */
  double dttr;    /* tracer time step (0: no advection)               */
  double t;       /* model time (s)                                   */
  int nstep;
  double *u1, *u2, *w;           /* 3D velocities (w on the lower face) */
  double *u1flux3d, *u2flux3d;   /* 3D volume fluxes (m3/s)           */
  double *dzu1, *dzu2;           /* face thicknesses                 */
  double *Kz;                    /* vertical diffusivity, lower face */
  double *eta;                   /* 2D elevation at t+1              */
  double *topz;                  /* 2D surface height                */
  double *wbot;                  /* 2D w at the bed                  */
  double *u1flux, *u2flux;       /* 2D e1 / e2 volume fluxes summed
                                    over the 2D steps of the 3D step
                                    (m3; set_flux_3d's faces one cell
                                    deep)                             */
  double *u1vm, *u2vm;           /* 3D mean volume fluxes (m3/s): the
                                    fluxes set_flux_3d copies under
                                    FFSL                              */
  double *etab;                  /* 2D elevation at the last transport
                                    step (vel_w_bounds' detadt)       */
  double *wtop, *detadt;         /* 2D surface w and d(eta)/dt        */
  /* auxiliary_routines (tracers.c:4296-4490): temperature and salinity
   * (tracer numbers; -1: absent), their relaxation fields and time
   * constants (3D, NULL: none), the age tracer (NULL: none)          */
  int tno, sno;
  double *rtemp, *rsalt;
  double *temp_tc, *salt_tc;
  double *agetr;
} window_t;

/*-------------------------------------------------------------------*/
/* Private window data (fields of the reference win_priv_t)         */
/*-------------------------------------------------------------------*/
typedef struct win_priv {
  int trasc;      /* advection scheme flag                            */
  int ultimate;   /* ULTIMATE limiter on/off                          */
  int stab;       /* stability method flags                           */
  int sigma;      /* sigma coordinates (0: z-levels)                  */
  int thin_merge; /* merge thin surface layers                        */
  double hmin;    /* minimum layer thickness                          */
  int *advect;    /* per tracer: advect                               */
  int *diffuse;   /* per tracer: horizontal diffusion                 */
  int *tbdy;      /* tracers to advect                                */
  int ntbdy;
  int *tdif_v;    /* tracers to diffuse vertically                    */
  int ntdif_v;
  double *mintr, *maxtr; /* per tracer valid range                   */

  double *oldeta; /* 2D elevation at t                                */
  double *dz;     /* 3D cell thickness (set_dz)                       */
  double *Ds, *Hn1, *Hs; /* sigma scaling (1 for z-levels)             */
  double *u1kh, *u2kh;   /* 3D horizontal diffusivities               */
  double *mdx, *mdy;     /* 2D metric factors                         */
  double *one;           /* 2D array of ones (diffusion scale)        */

  /* Cells to process (set_dz)                                       */
  int *s1;
  int vc, vcs, vc1, vcs1;
  int *i1, *i2, *i3, *i4, *i5;

  /* Work arrays                                                     */
  double *w1, *w2, *w3, *w4, *w5, *w6, *w7, *w8, *w9, *w10;
  double *d1, *d2, *d3, *d4;
  double *v1, *v2, *v3, *v4, *v5, *v6, *v7;
  double b1; /* current sub-step (set in advect_diffuse)             */
  int *twin; /* order to process the windows, twin[1..nwindows]
                (tracer_step reads wincon[1]->twin; NULL: 1..nwindows) */
  /* e1 / e2 faces whose column is one cell deep, cbot_e1[1..ncbot_e1]
   * (set by the momentum step; set_flux_3d gives them the 2D flux)  */
  int *cbot_e1, *cbot_e2;
  int ncbot_e1, ncbot_e2;
  int nonlinear;  /* surface slope terms in vel_w_bounds (rejected)    */
  /* auxiliary_routines: tracers to decay tdec[0..ntdec-1] with their
   * decay constants dectr[n] (s; or the 3D tracer number holding the
   * timescale, DE_TR3), tracer attributes trinfo_3d[ntr], the age mask
   * (3D, NULL: none); reset_dz's cells to process for the updated
   * elevation s2[1..vc2] (surface cells first: 1..vca2 wet, ..vcs2 OBC;
   * then sub-surface: ..vci2 wet, ..vc2 OBC) and the wet-cell mask c1 */
  int ntdec;
  int *tdec;
  double *dectr;
  tracer_info_t *trinfo_3d;
  short *agemsk;
  int *s2;
  int vc2, vcs2, vca2, vci2;
  char *c1;
} win_priv_t;

/*-------------------------------------------------------------------*/
/* Master data (the fields of the reference master_t,               */
/* model/hd/include/sparse.h:1050-2390, that tracer_step reads)     */
/*-------------------------------------------------------------------*/
/* master_t copy modes: hd_params.h:295-299 ; crash flag: :735      */
#define TMASS 16
#define TRACERS 256
#define RS_RESTART 0x0004
typedef struct master master_t;
struct master {
  int mode2d; /* 2D mode: only the 3D step is built (rejected)        */
  int crf;    /* crash recovery flag: set to RS_RESTART on a failure  */
  /* master <- window copies after each window's step (TRACERS) and
   * of the mass diagnostics (TMASS); NULL when the caller keeps no
   * master (one window)                                              */
  void (*win_data_empty_3d)(master_t *master, geometry_t *window, window_t *windat,
                            int mode);
};

/*-------------------------------------------------------------------*/
/* Drop-in entry points (reference signatures).  They run the step  */
/* on the GPU: fields are uploaded, the tracers are updated on the  */
/* device and copied back into windat->tr_wc in place.              */
/*-------------------------------------------------------------------*/
/* model/hd/tracers/tracers.c:142-209: every window of the step, 1-based
 * arrays window[1..nwindows] in the order wincon[1]->twin; per window the
 * lateral ghost fill of the tracers (set_lateral_BC_tr, tracers.c:5574),
 * tracer_step_3d on the GPU and master->win_data_empty_3d(TRACERS), then
 * the TMASS copies.  A failure sets master->crf = RS_RESTART and returns,
 * as the reference's crash path does (tracers.c:119, 175).            */
void tracer_step(master_t *master, geometry_t **window, window_t **windat,
                 win_priv_t **wincon, int nwindows);
/* model/hd/tracers/tracers.c:239 */
void tracer_step_3d(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd/tracers/tracers.c:401 ; returns 0 on success */
int advect_diffuse(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd/tracers/tracers.c:3782 */
void vert_diffuse_3d(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd/tracers/tracers.c:5682 (z-level branch) */
void set_dz(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd/tracers/tracers.c:4296-4490, the parts that change the
 * window's state after the transport: reset_dz (tracers.c:5900-5980:
 * surface thicknesses for the updated elevation, the cells-to-process
 * vector s2 and the mask c1), calc_age (:6960-6973), tracer_decay
 * (tracer_decay.c:26-63) and do_ts_relax (:6122-6207), in the
 * reference's order, on the GPU.  The diagnostics and process
 * libraries it also calls (mixed layer, flushing, percentiles, steric
 * height, vorticity, numbers, profiles, DHW, alerts, region and total
 * mass, sediments, ecology, waves, T/S increments) are not built.  A
 * decay timescale from a 2D tracer (DE_TR2) is refused.  tracer_step_3d
 * calls it after vert_diffuse_3d, as tracers.c:289 does.             */
void auxiliary_routines(geometry_t *window, window_t *windat, win_priv_t *wincon);

/*-------------------------------------------------------------------*/
/* Device-resident interface (no counterpart in the reference: the  */
/* reference keeps its state on the host).  A context mirrors one   */
/* window on the GPU; fields and tracers stay in HBM between steps. */
/*-------------------------------------------------------------------*/
typedef struct emstr_ctx emstr_ctx;

/* Arithmetic forms */
#define EMSTR_PLAIN_F64 0   /* reference arithmetic, FP64 throughout   */
#define EMSTR_PREC_F32 1    /* FP32 stencils/sweeps, FP64 budgets/IR   */
#define EMSTR_BATCH_F32 2   /* precision form, all tracers per cell in
                               one lane-per-tracer pass: cell-parallel
                               advection + per-column solve, hipGraph   */

/* Error codes */
#define EMSTR_OK 0
#define EMSTR_ESUBSTEP 1    /* sub-step limit exceeded (tracers.c:664) */
#define EMSTR_EUNSUPPORTED 2
#define EMSTR_EDEVICE 3
#define EMSTR_EZEROPIVOT 4  /* zero divisor in the column solve        */
#define EMSTR_EARG 5

/* Create a device mirror of a window.  form selects the arithmetic.   */
emstr_ctx *emstr_create(geometry_t *window, window_t *windat,
                        win_priv_t *wincon, int form);
void emstr_destroy(emstr_ctx *ctx);
/* Host -> device: geometry-dependent data, fields, tracers.           */
int emstr_upload_fields(emstr_ctx *ctx);
int emstr_upload_tracers(emstr_ctx *ctx);
/* Device -> host: tracers into windat->tr_wc (in place).              */
int emstr_download_tracers(emstr_ctx *ctx);
/* One full tracer step (advect_diffuse + vert_diffuse_3d) on device.  */
int emstr_step(emstr_ctx *ctx);
/* nsteps tracer steps (the batched form replays a captured hipGraph).  */
int emstr_steps(emstr_ctx *ctx, int nsteps);
/* Batched form launch shape: tile_x the tile width (16 columns, one
 * wave's row), tile_y the strip width (columns) of the tile order dealt
 * to the 8 XCDs, grid the number of depth bins the tiles of each XCD's
 * chunk are ordered by (deepest first; 1: strip order), graph replay
 * on/off; 0 (or -1 for use_graph) keeps a value.  Results do not depend
 * on the shape.  ems_amd/autotune.py picks it by measurement. */
int emstr_tune(emstr_ctx *ctx, int tile_x, int tile_y, int grid, int use_graph);
/* The batched form's current shape (tile width, strip width, depth
 * bins) and its workgroups per tracer group (EMSTR_EARG for the other
 * forms). */
int emstr_launch_shape(emstr_ctx *ctx, int *tile_x, int *tile_y, int *grid, int *grid_full);
/* Only the advection / only the vertical diffusion part.              */
int emstr_advect_diffuse(emstr_ctx *ctx);
int emstr_vert_diffuse(emstr_ctx *ctx);
/* auxiliary_routines (tracers.c:4296-4490, the parts listed at the
 * drop-in) on the device state of a plain-form context: dz, age, decay
 * and T/S relaxation; the relaxation fields, time constants and age mask
 * are uploaded from the window each call.  EMSTR_EUNSUPPORTED for the
 * FP32 forms when a decay, relaxation or age tracer is set (their
 * reset_dz runs).                                                      */
int emstr_auxiliary(emstr_ctx *ctx);
/* set_dz on the device (z-level branch of tracers.c:5682-5849) from
 * windat->eta, wincon->oldeta and window->nsur_t; keeps the per-column
 * results on the device (the drop-in set_dz() also writes the host lists). */
int emstr_device_set_dz(emstr_ctx *ctx);
/* w from continuity on the device (model/hd/momentum/vel3d.c:4680-4711)
 * from windat->u1flux3d/u2flux3d/wbot; copies w into windat->w.         */
int emstr_device_vel_w(emstr_ctx *ctx);
/* set_flux_3d on the device (model/hd/momentum/vel3d.c:4041-4191, VEL3D
 * mode, no mean fields): u1flux3d = u1 dzu1 h2au1 mdx over the wet e1
 * faces (likewise e2), the 2D flux / dt at the faces one cell deep
 * (cbot_e1/e2), zero above each column's surface face (sur_e1/e2) and
 * at every other face; under FFSL the mean volume fluxes u1vm / u2vm,
 * with the velocities of the faces past b3 zeroed on the window.  The
 * fluxes stay on the device for the step and are copied into
 * windat->u1flux3d / u2flux3d.                                        */
int emstr_device_set_flux_3d(emstr_ctx *ctx);
/* One offline transport step of the window on the device: what
 * hd_step_trans (model/hd/master/hd_step.c:319-507) runs per window for
 * an exact transport grid, closed boundaries, no fills and CONS_W:
 * transport_step (model/hd/tracers/transport.c:94-530: the linear
 * vel_w_bounds, set_dz_at_u1/u2, set_flux_3d, set_dz, vel_w_update),
 * tracer_step and transport_post (oldeta = eta).  The caller puts the
 * step's stored flow in the window (u1, u2, Kz, eta: what TRANS_DATA
 * holds for the step); windat->etab/detadt/wtop/wbot and wincon->oldeta
 * are updated on the host as the reference updates them; the face
 * thicknesses, fluxes, w, dz and tracers stay on the device
 * (emstr_transport_download / emstr_download_tracers copy them out).
 * Surface slope terms (wincon->nonlinear) are refused.                */
int emstr_transport_step(emstr_ctx *ctx);
/* The last transport step's dzu1, dzu2, u1flux3d, u2flux3d, w and dz at
 * the wet cells into the window's host arrays.                        */
int emstr_transport_download(emstr_ctx *ctx);
/* Domain tracer mass sum(tr*cellarea*dz) over s1, FP64 compensated.   */
int emstr_tracer_mass(emstr_ctx *ctx, double *mass /* [ntr] */);
/* emstr_tracer_mass that allocates nothing (no hipMalloc, hipFree or
 * allocating library call): no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  The partial sums and totals
 * live in a fixed 4 MiB device array of the library (one per GPU, kept
 * until the process ends), which limits 2 * ceil(columns / 256) * ntr +
 * ntr to 524288 doubles (EMSTR_EARG beyond).  Same values, bit for bit,
 * as emstr_tracer_mass.  Calls that overlap on one GPU share that array
 * and race.                                                            */
int emstr_tracer_mass_noalloc(emstr_ctx *ctx, double *mass /* [ntr] */);
/* Per-kernel timing with HIP events on the context's stream: enable
 * (and reset), then read accumulated milliseconds and launch counts per
 * kernel class [advection, vertical diffusion, prep, other, fused].   */
int emstr_profile(emstr_ctx *ctx, int on);
int emstr_kernel_times(emstr_ctx *ctx, double *ms /* [5] */, long long *count /* [5] */);
/* Synchronise the context's stream.                                   */
int emstr_sync(emstr_ctx *ctx);
/* Stream the context runs on (hipStream_t as void*).                  */
void *emstr_stream(emstr_ctx *ctx);

/* Ensemble of batched-form contexts (EMSTR_BATCH_F32 windows of any
 * sizes, e.g. independent domains): every context's step is captured as
 * one branch of a single hipGraph, so one launch advances all of them
 * and small windows run side by side on the GPU.  No counterpart in the
 * reference, which steps windows one after another (tracers.c:142-209).
 * The contexts must have their fields and tracers uploaded; while the
 * ensemble exists they are stepped only through it.  nsteps must be
 * even (the graph holds two steps).                                    */
typedef struct emstr_ensemble emstr_ensemble;
emstr_ensemble *emstr_ensemble_create(emstr_ctx **ctxs, int n);
int emstr_ensemble_steps(emstr_ensemble *ens, int nsteps);
int emstr_ensemble_sync(emstr_ensemble *ens);
void *emstr_ensemble_stream(emstr_ensemble *ens); /* hipStream_t */
void emstr_ensemble_destroy(emstr_ensemble *ens);

/* Release the device mirror the drop-in entry points keep for a window
 * (call before freeing or reusing the window's arrays).               */
void emstr_forget_window(geometry_t *window);

/* Last error text (thread-local) and the failing cell, if any.        */
const char *emstr_last_error(void);
int emstr_last_error_cell(void);

#ifdef __cplusplus
}
#endif
#endif
