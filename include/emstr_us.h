/*
 * libemstr_us: the tracer step of EMS COMPAS (unstructured meshes,
 * model/hd-us) on one MI355X, as a C-ABI shared library.
 *
 * The structures carry the fields of the COMPAS geometry_t, window_t
 * and win_priv_t (model/hd-us/include/sparse.h:880-1213, 2393-2567,
 * 246-807) that the tracer path reads, with the same meaning: 1-based
 * arrays (index 0 unused), cell centres numbered surface layer first,
 * edges of the surface layer first, ghosts after the wet cells.  COMPAS
 * is a separate model from SHOC (its own executable), so this library
 * is separate from libemstr and exports the same entry-point names;
 * include one of emstr.h / emstr_us.h, not both.
 */
#ifndef EMSTR_US_H
#define EMSTR_US_H

#ifdef __cplusplus
extern "C" {
#endif

/* Advection scheme flags: model/hd-us/include/hd_params.h */
#define ORDER1 0x000002
#define ORDER2 0x000004
#define QUICKEST 0x000020
#define VANLEER 0x000080
#define LAGRANGE 0x000100
#define FFSL 0x080000

/* FFSL settings: model/hd-us/include/hd_params.h */
#define TRANSPORT 0x0040  /* means: mean-velocity transport             */
#define MONOTONIC 0x0010  /* fillf                                      */
#define CLIP 0x8000000    /* fillf                                      */
#define V6898 0x000800    /* compatible                                 */
#define F_FFSL 0x100      /* flag: ffsl_init ran on the previous call    */

/* Stability flags */
#define SUB_STEP 2
#define SUB_STEP_NOSURF 4
#define SUB_STEP_TRACER 16
#define MULTI_DT 32
#ifndef NONE
#define NONE 0x80000000
#endif

/* Fraction of hmin above the bed for a column to be processed */
#define DRY_FRAC 0.05

/*-------------------------------------------------------------------*/
/* Mesh structure (fields of the COMPAS geometry_t)                  */
/*-------------------------------------------------------------------*/
typedef struct geometry {
  int szc, sze;     /* 3D cell-centre / edge array sizes              */
  int szcS, szeS;   /* 2D sizes (surface layer)                       */
  int nobc;         /* open boundaries (0: closed basin)              */
  int nz;           /* number of layers                               */
  int npem;         /* maximum edges per cell                         */
  int wn;           /* window number                                  */
  double *layers;   /* layer faces, layers[0..nz] (bottom to top)     */
  int *s2i, *s2k;   /* centre -> mesh cell, layer                     */

  /* Connectivity.  An edge e separates e2c[e][0] and e2c[e][1]; a
   * positive u1[e] flows from e2c[e][1] into e2c[e][0]; e2e[e][0] is
   * the index of e among the edges of e2c[e][0], e2e[e][1] among those
   * of e2c[e][1].  c2c[j][c] is the centre across edge j of c (a ghost
   * at land), c2e[j][c] that edge; eSc[j][c2] = 1 if edge j points out
   * of the cell (c2 == e2c[e][1]), else -1.                          */
  int *npe;         /* [szcS] edges of each cell                      */
  int **c2c;        /* [npem+1][szc]                                  */
  int **c2e;        /* [npem+1][szc]                                  */
  int **e2c;        /* [sze][2]                                       */
  int **e2e;        /* [sze][2]                                       */
  int **eSc;        /* [npem+1][szcS]                                 */
  int *ep, *em;     /* [sze] edge beyond e2c[e][0] / e2c[e][1] with
                     * the index of e (pp_us.c:1272-1282, e2e())       */
  int *m2d;         /* [szc] centre -> surface centre                 */
  int *m2de;        /* [sze] edge -> surface edge                     */
  int *zp1, *zm1;   /* [szc] vertical maps (self-mapping at the ends) */
  int *zp1e, *zm1e; /* [sze]                                          */

  /* Cells and edges to process: [1..v3] wet, [..b3] +OBC, [..a3]
   * +auxiliary, [..n3] +ghosts                                       */
  int *w3_t, *w3_e1;
  int n3_t, v3_t, b3_t, a3_t;
  int n3_e1, v3_e1, b3_e1, a3_e1;
  int *w2_t, *w2_e1;
  int n2_t, v2_t, b2_t, a2_t;
  int n2_e1, v2_e1, b2_e1, a2_e1;

  /* Ghosts: lateral tr[bpt[cc]] = tr[bin[cc]]; sediment below beds  */
  int nbpt;
  int *bpt, *bin;
  int ngsed;
  int *gsed_t, *ised_t;

  int *sur_t, *nsur_t, *bot_t; /* [szcS] surface / bed centres        */

  /* Metrics                                                          */
  double *h1au1;    /* [szeS] edge length                             */
  double *h2au1;    /* [szeS] distance between the centres of an edge */
  double *h1acell;  /* [szeS] distance between edge e and the opposite
                     * edge of e2c[e][0] (pp_us.c:1495-1508)            */
  double *cellarea; /* [szcS]                                         */
  double *costhu1, *sinthu1; /* [szeS] edge normal direction          */
  double *cellx, *celly;     /* [szcS] centre coordinates             */
  double *botz;     /* [szcS] bed height (negative down)              */
  double *gridz;    /* [szc] height of the lower face of each cell    */
  double *cellz;    /* [szc] height of the cell centre                */

  /* Vertices and edge geometry for FFSL (transport.c:1476-1826,
   * tracers.c:4682-4871; pp_us.c:1900-2100).  A vertex of layer k is a
   * corner of the centres of that layer; vertices of the surface layer
   * are numbered first (1..szvS-1).                                    */
  int szv, szvS;    /* 3D / 2D vertex array sizes                     */
  int nvcm;         /* most centres around a vertex                   */
  int **c2v;        /* [npem+1][szc] vertex j of a centre (0: none)   */
  int **v2c;        /* [szv][nvcm+1] wet centres of the vertex's layer
                     * around it, 0-padded; ghosts are not mapped     */
  int *nvc;         /* [szvS] centres around a surface vertex         */
  int *m2dv;        /* [szv] vertex -> surface vertex                 */
  int **e2v;        /* [sze][2] end vertices of an edge               */
  double *gridx, *gridy;     /* [szvS] vertex coordinates             */
  double *u1x, *u1y;         /* [szeS] edge midpoints (u1 points)      */
  double *costhu2, *sinthu2; /* [szeS] edge tangential direction      */
  int *wgst;        /* [szc] wet centre of a lateral ghost (0: none)  */
} geometry_t;

/* Least-squares linear reconstruction of one centre (lib/include/
 * lsqq.h:33-45): w = sum_j tr[cells[j]] * B[j][.], value at (x, y) =
 * w[0] + beta * (w[1] * (cellx - x) + w[2] * (celly - y))             */
typedef struct {
  double **B;       /* [ncells][3]                                    */
  double w[6];
  int ncells;
  int *cells;       /* [ncells] stencil centres (ghosts included)     */
  double tmx, tmn;
  int orf;
  double xr, yr;
  double beta, alpha;
} qweights;

/*-------------------------------------------------------------------*/
/* Window data (fields of the COMPAS window_t)                       */
/*-------------------------------------------------------------------*/
typedef struct window {
  int ntr;
  double **tr_wc;   /* tr_wc[n][1..szc-1]                             */
  double dt, dttr, t;
  int nstep;
  double *u1;       /* [sze] edge-normal velocity                     */
  double *u1flux3d; /* [sze] edge volume flux (m3/s)                  */
  double *dzu1;     /* [sze] edge face thickness                      */
  double *u, *v;    /* [szc] cell-centred east / north velocity       */
  double *w;        /* [szc] vertical velocity on the lower face      */
  double *Kz;       /* [szc] vertical diffusivity, lower face         */
  double *eta;      /* [szcS] elevation at t+1                        */
  double *topz;     /* [szcS]                                         */
  double *u2;       /* [sze] edge-tangential velocity (FFSL)          */
  double *wtop, *wbot; /* [szcS] vertical velocity at the surface / bed */
} window_t;

/*-------------------------------------------------------------------*/
/* Private window data (fields of the COMPAS win_priv_t)             */
/*-------------------------------------------------------------------*/
typedef struct win_priv {
  int trasc, ultimate, stab, sigma, thin_merge;
  int rkstage;      /* Runge-Kutta stages (1: forward Euler)          */
  int means;        /* mean-velocity transport flags (0 here)         */
  double hmin;
  double trsf;      /* FFSL sub-step safety factor                    */
  int *advect, *diffuse;
  int *tbdy, ntbdy;
  int *tdif_v, ntdif_v;
  double *mintr, *maxtr;

  double *oldeta;   /* [szcS]                                         */
  double *dz;       /* [szc] cell thickness (set_dz)                  */
  double *Ds, *Hn1, *Hs; /* [szcS] sigma scaling (1 for z-levels)     */
  double *u1kh;     /* [szc] cell-centred horizontal diffusivity      */
  double *u2kh;     /* [sze] edge diffusivity used by hor_diffuse     */
  double *mdx;      /* [szeS] metric factor                           */
  double *one;      /* [szcS] ones (diffusion scale)                  */

  int *s1;
  int vc, vcs, vc1, vcs1;
  int *i1, *i2, *i3, *i4, *i5;

  double *w1, *w2, *w3, *w4, *w5, *w6, *w7, *w8, *w9, *w10; /* [max(szc,sze)] */
  double *d1, *d2, *d3, *d4;              /* [max(szcS,szeS)]         */
  double *v1, *v2, *v3, *v4, *v5, *v6, *v7; /* [nz+1]                 */
  double b1;        /* current sub-step                               */
  /* FFSL (transport.c:1961-2723)                                     */
  int flag;         /* F_FFSL toggled by ffsl_init                    */
  int fillf;        /* FILL_METHOD (MONOTONIC / CLIP: rejected)       */
  int compatible;   /* V6898: rejected with FFSL                      */
  qweights *lw;     /* [szc] weights of build_linear_weights (2D fit) */
  int *twin;        /* order to process the windows, twin[1..nwindows]
                       (tracer_step reads wincon[1]->twin; NULL: 1..n) */
} win_priv_t;

/*-------------------------------------------------------------------*/
/* Master data (the fields of the COMPAS master_t,                    */
/* model/hd-us/include/sparse.h:1300-2390, that tracer_step reads)    */
/*-------------------------------------------------------------------*/
/* copy modes: model/hd-us/include/hd_params.h ; crash flag RS_RESTART */
#define TMASS 16
#define TRACERS 256
#define RS_RESTART 0x0004
typedef struct master master_t;
struct master {
  int mode2d; /* 2D mode: only the 3D step is built (rejected)          */
  int crf;    /* crash recovery flag: set to RS_RESTART on a failure    */
  /* master <- window copies (TRACERS after each window, then TMASS);
   * NULL when the caller keeps no master                              */
  void (*win_data_empty_3d)(master_t *master, geometry_t *window, window_t *windat,
                            int mode);
};

/*-------------------------------------------------------------------*/
/* Streamline velocity fit of an FFSL window (no counterpart struct in */
/* the reference, which keeps these weights in its interpolation       */
/* library's structures: tran_grid_init, model/hd-us/tracers/          */
/* transport.c:4317-4658, set_interp_points_n :5537-5640, lsql_build   */
/* lib/interp/lsq_lin/lsql.c:47-210).  Wet centre c's fit reads the    */
/* points cell[vo[c] .. vo[c] + vn[c] - 1]: wet centres, then lateral  */
/* ghosts (edge[]: the ghost's edge, whose midpoint is the point; 0    */
/* for a wet centre), with the slope rows bx, by of the pseudo-inverse */
/* of the rows (1, x - xr, y - yr).  A fitted value at (x, y) is        */
/* v[c] + (sum v_j bx_j (x - xr) + sum v_j by_j (y - yr)), (xr, yr) the */
/* centre's position.  Built by the library for FFSL windows; exported */
/* so that a checker can use the same weights.                         */
/*-------------------------------------------------------------------*/
typedef struct emstr_us_vlsq {
  int szc, nent;
  int *vo, *vn;     /* [szc] first entry and count of a wet centre      */
  int *cell, *edge; /* [nent]                                           */
  double *bx, *by;  /* [nent]                                           */
} emstr_us_vlsq;
/* 0 on success, -1 without vertex maps or memory */
int emstr_us_vlsq_build(const geometry_t *window, emstr_us_vlsq *t);
void emstr_us_vlsq_free(emstr_us_vlsq *t);

/*-------------------------------------------------------------------*/
/* Drop-in entry points (COMPAS signatures)                          */
/*-------------------------------------------------------------------*/
/* model/hd-us/tracers/tracers.c:177-239: every window, 1-based arrays in
 * the order wincon[1]->twin: lateral ghost fill of the tracers
 * (set_lateral_BC_tr), tracer_step_3d on the GPU, win_data_empty_3d
 * (TRACERS), then the TMASS copies; a failure sets master->crf =
 * RS_RESTART and returns.                                             */
void tracer_step(master_t *master, geometry_t **window, window_t **windat,
                 win_priv_t **wincon, int nwindows);
/* model/hd-us/tracers/tracers.c:261 */
void tracer_step_3d(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd-us/tracers/tracers.c:428; 0 on success, else the column
 * (2D centre) where the sub-step limit was exceeded (tracers.c:688) */
int advect_diffuse(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd-us/tracers/tracers.c:5562 */
void vert_diffuse_3d(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* model/hd-us/tracers/tracers.c:7231 (z-level branch) */
void set_dz(geometry_t *window, window_t *windat, win_priv_t *wincon);

/*-------------------------------------------------------------------*/
/* Device-resident interface (no counterpart in the reference, which  */
/* keeps its state on the host): a context mirrors one window on the  */
/* GPU; fields and tracers stay in HBM between steps.                 */
/*-------------------------------------------------------------------*/
typedef struct emstr_us_ctx emstr_us_ctx;
emstr_us_ctx *emstr_us_create(geometry_t *window, window_t *windat, win_priv_t *wincon);
/* Forms: EMSTR_US_PLAIN_F64 (emstr_us_create's: the reference's FP64
 * operations, bit-exact), EMSTR_US_BATCH_F32 (FFSL with one Runge-Kutta
 * stage: the per-tracer FFSL work, tr_mod, least-squares slopes and
 * limiter, streamline means, tr_moc, edge and vertical fluxes, in FP32
 * for every tracer of the step per launch, over the plain form's FP64
 * maps, streamlines and geometry; the tracer state, the update and the
 * column solve stay FP64; other settings are refused: NULL and
 * emstr_us_last_error). */
#define EMSTR_US_PLAIN_F64 0
#define EMSTR_US_BATCH_F32 2
emstr_us_ctx *emstr_us_create_form(geometry_t *window, window_t *windat, win_priv_t *wincon,
                                   int form);
void emstr_us_destroy(emstr_us_ctx *ctx);
/* fields, per-column lists and dz from the host structures          */
int emstr_us_upload_fields(emstr_us_ctx *ctx);
int emstr_us_upload_tracers(emstr_us_ctx *ctx);
int emstr_us_download_tracers(emstr_us_ctx *ctx);
/* nsteps of advect_diffuse + vert_diffuse_3d on the device           */
int emstr_us_steps(emstr_us_ctx *ctx, int nsteps);
int emstr_us_sync(emstr_us_ctx *ctx);
void *emstr_us_stream(emstr_us_ctx *ctx); /* hipStream_t of the context */

/* The FFSL maps of the window's last sub-step on the device, in the
 * window's numbering (prep_ff_sl and ffsl_don's upstream cells): cl[e]
 * the centre a wet edge's streamline starts in (sze entries), clzf[c]
 * the departure centre of the lower face of every cell to process and
 * clzc[c] of every wet centre (szc entries; the sediment ghost may
 * appear).  No counterpart in the reference, which keeps them in
 * scratch; for checking the maps. */
int emstr_us_ffsl_maps(geometry_t *window, int *cl, int *clzf, int *clzc);

/* Errors of the last call on this thread ("" when none) */
const char *emstr_us_last_error(void);
int emstr_us_last_error_cell(void);
/* Drop the device mirror of a window (after its arrays are freed) */
void emstr_us_forget_window(geometry_t *window);

#ifdef __cplusplus
}
#endif
#endif
