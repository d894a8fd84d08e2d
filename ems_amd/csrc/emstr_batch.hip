/*
 * Batched form of the EMS tracer step on MI355X (gfx950): all tracers of
 * the step over all water columns, FP32 state with FP64 budgets.
 *
 * Layout: tracers are stored tracer-innermost, [column][layer][tracer] in
 * FP32, so the tracers of one cell are one contiguous 4*ntr-byte row and
 * a wave with one lane per tracer reads a cell with one coalesced load.
 * Everything that does not depend on the tracer (face transports, van
 * Leer factors, the normalised column matrix and its LU factors) is
 * prepared per cell once per step (k_prepB, k_colB) and read by all lanes.
 *
 * Per step:
 *   k_prepB / k_colB  per-cell coefficients, per-column information (the
 *                     FP64 surface volume balance, diffusion extent,
 *                     stiffness, LU factors of the column matrix);
 *   k_advB            cell-parallel advection: face values (van Leer,
 *                     tracers.c:2097-2118) on the 13-point stencil and the
 *                     flux divergence and sub-surface update
 *                     (tracers.c:888-925) in anomaly form
 *                       tr - (sum_f s_f (F_f - tr) q_f + tr*D) / V,
 *                     D the FP64 net volume divergence of the cell, so that
 *                     rounding the transports q to FP32 only perturbs the
 *                     small (F - tr) terms; the processed surface cell by
 *                     surf_conc (tracers.c:3047-3207) as an FP64 volume
 *                     balance;
 *   k_vertB           per column: tr_bounds (tracers.c:3217-3324), then the
 *                     implicit vertical diffusion (tracers.c:3782-3957,
 *                     4067-4214) in increment form, A (x - b) = dzdt*b - A b,
 *                     with one FP64 residual pass of iterative refinement
 *                     in columns whose diffusion number exceeds IR_DMAX.
 * Every FP32 state write is stochastically rounded from an FP64 value with
 * a deterministic hash of (cell, tracer, step), which keeps sub-ulp
 * increments from biasing the state and the tracer mass.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "emstr_batch.h"
#include "emstr_dev.h"
#include "emstr_internal.h"
#include "emstr_kernels.h"
#include "emstr_prof.h"

#define VANLEER 0x000080
#define IR_DMAX 1.0
#define NZMAX 64

typedef long long ll;

/* coefficient slots per cell */
enum { C_QW, C_CW, C_QE, C_CE, C_QS, C_CS, C_QN, C_CN, C_QB, C_CB, C_RV, C_D,
       C_GM, C_GP, C_UD, C_RD, C_N };

/* per-column information (two 64-byte scalar loads) */
struct ColInfo {
  int kb, ks, kold, cs, cb, flags, kw1, kw2, ke1, ke2, ks1, ks2, kn1, kn2, pad0, pad1;
  double vc;   /* consistent old volume of the surface cells + inflow (FP64) */
  double svol; /* new surface volume                                        */
  double dzs;  /* subeta - waterbot of the processed surface cell           */
  double area, osub, botz, pad2, pad3;
};
enum { F_PROC = 1, F_SINGLE = 2, F_STIFF = 4, F_THIN = 8, F_KEEPOLD = 16, F_MOVE = 32 };

struct emstr_batch {
  int ntr, nt64, nz, nx, ny, ncol;
  ll plane;
  float *tr[2];      /* [col][k][ntr] FP32                                    */
  int cur;
  float *co;         /* [col][k][C_N] coefficients                            */
  ColInfo *ci;       /* [plane]                                               */
  int *order;        /* [nord] column processing order (-1: none)             */
  int nord;
  float *scr;        /* refinement scratch [col][k][ntr] (stiff columns)      */
  float *mintr, *maxtr; /* [ntr]                                               */
  int grid;          /* persistent grid size                                  */
  int tile_x, tile_y;
  emstr_prof *prof;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int graph_ok;
  int graph_cur; /* buffer parity the graph was captured for */
  int *stepctr;  /* device step counter (stochastic rounding keys) */
  /* kernel A: per wave slot lists of cell descriptors and of their
   * columns' (processed surface + 1) | flags << 8 */
  unsigned long long *cells;
  int *kfl;
  int niter;  /* list length per slot (a multiple of 64)               */
  int gridA;  /* kernel A workgroups (4 wave slots each)               */
  ll nlist;   /* gridA * 4 * niter                                     */
  int grid_full, gridA_full; /* the grids that fill the GPU          */
};

__device__ __forceinline__ ll cidx(ll col, int nz, int k) { return col * nz + k; }

/*-------------------------------------------------------------------*/
/* prep: per-cell FP32 coefficients from the FP64 fields              */
/*-------------------------------------------------------------------*/
/* the factor of the limited term of the van Leer face value for
 * transport q with Courant number cn: 1 - cn if q > 0, 1 + cn if q < 0 */
__device__ __forceinline__ float vlfac(double q, double cn) {
  return (q > 0.0) ? (float)(1.0 - cn) : (q < 0.0) ? (float)(1.0 + cn) : 1.0f;
}

__global__ void k_prepB(emstr_dev d, emstr_batch b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const int ksc = d.ks[col];
  const ll P = d.plane, nx = d.nx;
  const ll idx = (ll)k * P + col;
  const double dt = d.dt, area = d.cellarea[col];
  const double dz = d.dz[idx];
  const double dzb = (k > kbc) ? d.dz[idx - P] : 0.0;
  const bool we = k >= d.kb[col + 1], wn = k >= d.kb[col + nx];
  const bool top = (ksc >= 0 && k < ksc);
  double qW = d.u1flux3d[idx] * dt, qS = d.u2flux3d[idx] * dt;
  double qE = we ? d.u1flux3d[idx + 1] * dt : 0.0;
  double qN = wn ? d.u2flux3d[idx + nx] * dt : 0.0;
  double qB = d.w[idx] * dt * area;
  /* the top face's transport and Courant number are the bottom face's
   * of the cell above (identical expressions) */
  const double qT = top ? d.w[idx + P] * dt * area : 0.0;
  /* the van Leer face value needs 0.5 (1 - cn) for q > 0 and 0.5 (1 + cn)
   * for q < 0 (tracers.c:2097-2118); the factor 2 of the limiter is
   * folded in, so the slot holds 1 - cn or 1 + cn */
  const double cW = d.u1[idx] * dt / d.h1au1[col];
  const double cE = we ? d.u1[idx + 1] * dt / d.h1au1[col + 1] : 0.0;
  const double cS = d.u2[idx] * dt / d.h2au2[col];
  const double cN = wn ? d.u2[idx + nx] * dt / d.h2au2[col + nx] : 0.0;
  const double cB = d.w[idx] * dt / (0.5 * (dz + dzb));
  float c[C_N];
  c[C_QW] = (float)qW;
  c[C_CW] = vlfac(qW, cW);
  c[C_QE] = (float)qE;
  c[C_CE] = vlfac(qE, cE);
  c[C_QS] = (float)qS;
  c[C_CS] = vlfac(qS, cS);
  c[C_QN] = (float)qN;
  c[C_CN] = vlfac(qN, cN);
  c[C_QB] = (float)qB;
  c[C_CB] = vlfac(qB, cB);
  c[C_RV] = (dz != 0.0) ? (float)(1.0 / (area * dz)) : 0.0f;
  c[C_D] = (float)(qE - qW + qN - qS + qT - qB);
  /* vertical diffusion rows of the new surface column, normalised by
   * dzdt: gm = Cm1/dzdt, gp = Cp1/dzdt (tracers.c:3841-3889) */
  float gm = 0.0f, gp = 0.0f;
  const int cs = d.knew[col];
  if (ksc >= 0 && k <= cs && cs - 1 >= kbc) {
    const double dzdt = dz / dt;
    if (k > kbc) gm = (float)(-d.Kz[idx] / (0.5 * (dzb + dz)) / dzdt);
    if (k < cs) gp = (float)(-d.Kz[idx + P] / (0.5 * (dz + d.dz[idx + P])) / dzdt);
  }
  c[C_GM] = gm;
  c[C_GP] = gp;
  c[C_UD] = 0.0f;
  c[C_RD] = 1.0f;
  float4 *c4 = (float4 *)(b.co + cidx(col, d.nz, k) * C_N);
  for (int w = 0; w < C_N / 4; w++)
    c4[w] = make_float4(c[4 * w], c[4 * w + 1], c[4 * w + 2], c[4 * w + 3]);
}

/* per-column information, one thread per processed-or-wet column */
__global__ void k_colB(emstr_dev d, emstr_batch b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  ColInfo c;
  memset(&c, 0, sizeof(c));
  c.kb = kbc;
  c.ks = -1;
  if (kbc >= d.nz) {
    b.ci[col] = c;
    return;
  }
  const ll P = d.plane, nx = d.nx;
  const int ksc = d.ks[col];
  c.ks = ksc;
  c.kold = d.kold[col];
  c.cs = d.knew[col];
  c.kw1 = d.kb[col - 1];
  c.ke1 = d.kb[col + 1];
  c.ks1 = d.kb[col - nx];
  c.kn1 = d.kb[col + nx];
  c.kw2 = (c.kw1 < d.nz) ? d.kb[col - 2] : d.nz;
  c.ke2 = (c.ke1 < d.nz) ? d.kb[col + 2] : d.nz;
  c.ks2 = (c.ks1 < d.nz) ? d.kb[col - 2 * nx] : d.nz;
  c.kn2 = (c.kn1 < d.nz) ? d.kb[col + 2 * nx] : d.nz;
  c.area = d.cellarea[col];
  c.botz = d.botz[col];
  c.osub = d.osubeta[col];
  int flags = 0;
  if (ksc >= 0) {
    flags |= F_PROC;
    const double dt = d.dt, area = c.area;
    /* surface volume balance in FP64 (surf_conc): old volumes of the cells
     * from the processed surface to the old surface, minus the net
     * horizontal outflow of every cell from it to the top, plus the inflow
     * through its bottom face */
    const int kold = c.kold, top = d.nz - 1;
    double waterbot = (ksc == kbc) ? c.botz : d.layers[ksc];
    const double subeta = d.subeta[col];
    c.dzs = subeta - waterbot;
    c.svol = c.dzs * area;
    double vol = 0.0;
    int cc = ksc;
    double wb = waterbot;
    while (cc != kold && cc < top) {
      vol += area * (d.layers[cc + 1] - wb);
      cc++;
      wb = d.layers[cc];
    }
    vol += area * (c.osub - ((cc == kbc) ? c.botz : d.layers[cc]));
    double out = 0.0;
    bool above = false; /* transport through a face above the processed cell */
#pragma unroll 4
    for (int k = ksc; k <= top; k++) {
      const ll idx = (ll)k * P + col;
      const bool we = k >= d.kb[col + 1], wn = k >= d.kb[col + nx];
      double qE = we ? d.u1flux3d[idx + 1] * dt : 0.0;
      double qN = wn ? d.u2flux3d[idx + nx] * dt : 0.0;
      out += qE - d.u1flux3d[idx] * dt + qN - d.u2flux3d[idx] * dt;
      if (k > ksc && (qE != 0.0 || qN != 0.0 || d.u1flux3d[idx] != 0.0 || d.u2flux3d[idx] != 0.0))
        above = true;
    }
    const double qB = d.w[(ll)ksc * P + col] * dt * area;
    c.vc = vol - out + qB;
    /* cells other than the processed one enter the balance when the old
     * surface is above it or any layer lies above it */
    if (kold != ksc || above) flags |= F_MOVE;
    /* thin old surface layer (tracers.c:3250) */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * P + col] : 0.0;
    if (ksc > kbc && dzco < d.hmin) flags |= F_THIN;
    if (c.botz > 0.0 && d.dz[(ll)ksc * P + col] < d.hmin) flags |= F_KEEPOLD;
    /* vertical diffusion extent (tracers.c:3836-3866) */
    const int cs = c.cs;
    if (cs < kbc || cs - 1 < kbc || d.dz[(ll)(cs - 1) * P + col] == 0.0) {
      flags |= F_SINGLE;
      c.cb = cs;
    } else {
      /* one above the highest zero-thickness layer at or below cs (the
       * run of open layers ending at cs); loads independent of each other */
      int zq = kbc - 1;
#pragma unroll 8
      for (int k = kbc; k <= cs; k++)
        if (d.dz[(ll)k * P + col] == 0.0) zq = k;
      c.cb = zq + 1;
      if (c.cb >= cs) flags |= F_SINGLE;
#pragma unroll 4
      for (int k = c.cb + 1; k <= cs; k++) {
        const double dzc = d.dz[(ll)k * P + col];
        const double dzm = d.dz[(ll)(k - 1) * P + col];
        if (d.Kz[(ll)k * P + col] * dt / (dzc * 0.5 * (dzm + dzc)) > IR_DMAX) flags |= F_STIFF;
      }
      /* LU factorisation of the normalised column matrix (rows gm, 1 - gm -
       * gp, gp): ud_k = gp_{k-1} / div_{k-1}, div_k = 1 - gm - gp - gm ud_k;
       * stored as ud and 1/div */
      if (!(flags & F_SINGLE)) {
        float divp = 1.0f, gpp = 0.0f;
#pragma unroll 4
        for (int k = c.cb; k <= cs; k++) {
          float *cf = b.co + cidx(col, d.nz, k) * C_N;
          const float gm = (k > c.cb) ? cf[C_GM] : 0.0f, gp = cf[C_GP];
          float ud = 0.0f, dv = 1.0f - gm - gp;
          if (k > c.cb) {
            ud = gpp / divp;
            dv = dv - gm * ud;
          }
          cf[C_UD] = ud;
          cf[C_RD] = 1.0f / dv;
          divp = dv;
          gpp = gp;
        }
      }
    }
  }
  c.flags = flags;
  b.ci[col] = c;
}

/*-------------------------------------------------------------------*/
/* the fused step                                                    */
/*-------------------------------------------------------------------*/
/* van Leer face value (tracers.c:2097-2118) for transport q through the
 * face between bm1 and c, returned as its anomaly F - ref against a
 * reference value close to it; f is 1 - cn for q > 0 and 1 + cn for
 * q < 0 (the prep folds the Courant number and the limiter's factor 2
 * into it).  The upwind value's anomaly is an exact FP32 difference for
 * neighbouring values within a factor of two, so the only rounding left
 * is relative to the small limited term and to the anomaly itself, not to
 * the tracer value.  The limiter's division uses the hardware reciprocal
 * (about 1 ulp). */
__device__ __forceinline__ float vla(float bm2, float bm1, float c, float fp1,
                                     float q, float f, float ref) {
  if (q > 0.0f) {
    const float df = (bm1 - bm2) * (c - bm1);
    const float t = df * __builtin_amdgcn_rcpf(c - bm2);
    return (df > 0.0f) ? (bm1 - ref) + f * t : bm1 - ref;
  } else if (q < 0.0f) {
    const float df = (c - bm1) * (fp1 - c);
    const float t = df * __builtin_amdgcn_rcpf(fp1 - bm1);
    return (df > 0.0f) ? (c - ref) - f * t : c - ref;
  }
  return c - ref; /* no transport: its flux term vanishes */
}

/* vla of a cell's own faces without branches: for q < 0 the stencil is
 * mirrored, (bm2, bm1, c) -> (fp1, c, bm1), which leaves df unchanged
 * and negates the reciprocal, so both signs share one expression with the
 * same values bit for bit; q == 0 keeps the upwind anomaly (c - ref).
 * ref is the cell's own value: c on its west, south and bottom faces, bm1
 * on its east, north and top faces. */
__device__ __forceinline__ float vla_m(float bm2, float bm1, float c, float fp1, float q,
                                       float f, float ref) {
  const bool pos = q > 0.0f;
  const float p2 = pos ? bm2 : fp1, p1 = pos ? bm1 : c, p0 = pos ? c : bm1;
  const float df = (p1 - p2) * (p0 - p1);
  const float t = df * __builtin_amdgcn_rcpf(p0 - p2);
  const float r0 = p1 - ref;
  const float r = __builtin_fmaf(f, (df > 0.0f) ? t : 0.0f, r0);
  return (q != 0.0f) ? r : r0;
}
__device__ __forceinline__ float vla_c(float bm2, float bm1, float c, float fp1, float q,
                                       float f) {
  return vla_m(bm2, bm1, c, fp1, q, f, c);
}
__device__ __forceinline__ float vla_b(float bm2, float bm1, float c, float fp1, float q,
                                       float f) {
  return vla_m(bm2, bm1, c, fp1, q, f, bm1);
}

/*-------------------------------------------------------------------*/
/* deterministic stochastic rounding                                 */
/*-------------------------------------------------------------------*/
/* 32-bit integer hash (xorshift-multiply finaliser) */
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

/* per-cell key: one multiply round over the cell index and the lane's
 * (tracer, step) hash */
__device__ __forceinline__ uint32_t srkey(uint32_t cellkey, uint32_t hl) {
  uint32_t x = cellkey ^ hl;
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  return x;
}

/* Round the FP64 value x to one of its two neighbouring FP32 values, the
 * upper one with probability equal to the remainder's fraction of an FP32
 * ulp: a uniform 29-bit number is added to the 29 mantissa bits that FP32
 * drops, which are then cleared.  Unbiased, so a state that changes by less
 * than an ulp per step does not drift systematically.  r is a uniform
 * 32-bit hash of (cell, tracer, step, stage). */
__device__ __forceinline__ float sround(double x, uint32_t r) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  u += (unsigned long long)(r >> 3);
  u &= ~0x1FFFFFFFULL;
  return (float)__longlong_as_double((long long)u);
}

/*-------------------------------------------------------------------*/
/* the fused step: one wave walks one water column                    */
/*-------------------------------------------------------------------*/
#define WPB 4 /* waves per workgroup, each on its own columns */

struct KArgs {
  int nz, ntr, nord, g; /* g: which 64 tracers this launch takes */
  int nact;             /* active lanes: min(64, ntr - 64 g)       */
  int pad;
  ll nx;
  double hmin, dt;
  ll plane;
};

/* lane access: p is a wave-uniform byte address, o the lane's byte offset */
__device__ __forceinline__ float ldb(const char *p, uint32_t o) {
  return *(const float *)(p + o);
}
__device__ __forceinline__ void stb(char *p, uint32_t o, float v) { *(float *)(p + o) = v; }
__device__ __forceinline__ float rdl(int v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(v, l));
}

/* Base addresses of the columns around one column; a neighbour that is
 * land or below the bed at layer k is replaced by the cell the map walk
 * came from, whose ghost holds that value after the lateral boundary fill
 * (tracers.c:5574-5593).  All wave-uniform. */
struct Nbr {
  const char *c, *w1, *w2, *e1, *e2, *s1, *s2, *n1, *n2;
  int kw1, kw2, ke1, ke2, ks1, ks2, kn1, kn2;
};

struct HVals {
  float w1, w2, e1, e2, s1, s2, n1, n2;
};

/* the eight horizontal neighbours at layer k; o = lane offset + k * row */
__device__ __forceinline__ void load_h(HVals &h, const Nbr &nb, int k, uint32_t o) {
  const bool w1 = k >= nb.kw1, e1 = k >= nb.ke1, s1 = k >= nb.ks1, n1 = k >= nb.kn1;
  const char *pw1 = w1 ? nb.w1 : nb.c, *pe1 = e1 ? nb.e1 : nb.c;
  const char *ps1 = s1 ? nb.s1 : nb.c, *pn1 = n1 ? nb.n1 : nb.c;
  const char *pw2 = (w1 && k >= nb.kw2) ? nb.w2 : pw1;
  const char *pe2 = (e1 && k >= nb.ke2) ? nb.e2 : pe1;
  const char *ps2 = (s1 && k >= nb.ks2) ? nb.s2 : ps1;
  const char *pn2 = (n1 && k >= nb.kn2) ? nb.n2 : pn1;
  h.w1 = ldb(pw1, o);
  h.w2 = ldb(pw2, o);
  h.e1 = ldb(pe1, o);
  h.e2 = ldb(pe2, o);
  h.s1 = ldb(ps1, o);
  h.s2 = ldb(ps2, o);
  h.n1 = ldb(pn1, o);
  h.n2 = ldb(pn2, o);
}

/* horizontal anomaly outflow sum_f s_f (F_f - ref) q_f of a cell whose
 * value is v0, from the face transports q and van Leer factors f */
__device__ __forceinline__ float hflux(const HVals &h, bool ew, bool nw, float v0, float ref,
                                       float qW, float fW, float qE, float fE, float qS,
                                       float fS, float qN, float fN) {
  const float AW = vla(h.w2, h.w1, v0, h.e1, qW, fW, ref);
  const float AE = ew ? vla(h.w1, v0, h.e1, h.e2, qE, fE, ref) : 0.0f;
  const float AS = vla(h.s2, h.s1, v0, h.n1, qS, fS, ref);
  const float AN = nw ? vla(h.s1, v0, h.n1, h.n2, qN, fN, ref) : 0.0f;
  return AE * qE - AW * qW + AN * qN - AS * qS;
}

/* hflux of a cell against its own value (ref == v0) */
__device__ __forceinline__ float hflux_own(const HVals &h, bool ew, bool nw, float v0, float qW,
                                           float fW, float qE, float fE, float qS, float fS,
                                           float qN, float fN) {
  const float AW = vla_c(h.w2, h.w1, v0, h.e1, qW, fW);
  const float AE = ew ? vla_b(h.w1, v0, h.e1, h.e2, qE, fE) : 0.0f;
  const float AS = vla_c(h.s2, h.s1, v0, h.n1, qS, fS);
  const float AN = nw ? vla_b(h.s1, v0, h.n1, h.n2, qN, fN) : 0.0f;
  return AE * qE - AW * qW + AN * qN - AS * qS;
}

/*-------------------------------------------------------------------*/
/* kernel A: cell-parallel advection                                  */
/*-------------------------------------------------------------------*/
/* Cell descriptor: column (bits 0-31), layer k (32-39), open neighbours
 * at layer k (40-47, M_* below; a second neighbour is open only if the
 * first one is), bed layer kb (48-55).  ~0 marks padding. */
enum { M_W1 = 1, M_W2 = 2, M_E1 = 4, M_E2 = 8, M_S1 = 16, M_S2 = 32, M_N1 = 64, M_N2 = 128 };
#define NO_CELL (~0ULL)

struct AArgs {
  int nz, ntr, g, nact; /* g: which 64 tracers this launch takes */
  int niter;            /* iterations per wave slot (a multiple of 64) */
  int pad;
  ll nx;
  double hmin;
};

/* the 13-point stencil of one cell for one tracer lane */
struct St13 {
  float v0, zm1, zm2, zp1, zp2, w1, w2, e1, e2, s1, s2, n1, n2;
};

__device__ __forceinline__ unsigned long long rdl64(unsigned long long v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

/* Issue the loads of the stencil of cell dsc (wave-uniform) and of its
 * coefficient words (lanes 0..15: cell k, 16..31: cell k + 1). */
__device__ __forceinline__ void st13_load(St13 &v, int &cw, const char *tin,
                                          const float *co, unsigned long long dsc, ll colb,
                                          ll jb, uint32_t lb, uint32_t rowb, int nz,
                                          int lane) {
  const ll col = (ll)(uint32_t)dsc;
  const int k = (int)((dsc >> 32) & 255), m = (int)((dsc >> 40) & 255);
  const int kb = (int)((dsc >> 48) & 255), top = nz - 1;
  const char *c = tin + col * colb;
  const uint32_t o = lb + k * rowb;
  const char *pw1 = (m & M_W1) ? c - colb : c, *pe1 = (m & M_E1) ? c + colb : c;
  const char *ps1 = (m & M_S1) ? c - jb : c, *pn1 = (m & M_N1) ? c + jb : c;
  const char *pw2 = (m & M_W2) ? c - 2 * colb : pw1, *pe2 = (m & M_E2) ? c + 2 * colb : pe1;
  const char *ps2 = (m & M_S2) ? c - 2 * jb : ps1, *pn2 = (m & M_N2) ? c + 2 * jb : pn1;
  v.v0 = ldb(c, o);
  v.zm1 = ldb(c, lb + max(k - 1, kb) * rowb);
  v.zm2 = ldb(c, lb + max(k - 2, kb) * rowb);
  v.zp1 = ldb(c, lb + min(k + 1, top) * rowb);
  v.zp2 = ldb(c, lb + min(k + 2, top) * rowb);
  v.w1 = ldb(pw1, o);
  v.w2 = ldb(pw2, o);
  v.e1 = ldb(pe1, o);
  v.e2 = ldb(pe2, o);
  v.s1 = ldb(ps1, o);
  v.s2 = ldb(ps2, o);
  v.n1 = ldb(pn1, o);
  v.n2 = ldb(pn2, o);
  cw = __float_as_int(co[(col * nz + k) * C_N + (lane & 31)]);
}

/* The update of one wet cell k <= ks of column col for one tracer lane,
 * from its 13-point stencil v and coefficient words cw (lanes 0..15: cell
 * k, 16..31: cell k + 1); m: open neighbours at layer k (M_*).  Writes the
 * advected value b of a sub-surface cell, or the raw new surface value
 * (surf_conc) of the processed surface cell, to tout. */
__device__ __forceinline__ void adv_cell(const St13 &v, int cw, ll col, int k, int m, int ks,
                                         int flags, const char *tinb, float *tout,
                                         const ColInfo *__restrict__ cinfo, const float *co,
                                         const double *__restrict__ layers, float *scr,
                                         uint32_t lb, uint32_t rowb, ll colb, ll jb, int nz,
                                         uint32_t hl, float mn, bool clip, bool act,
                                         double hmin) {
  const int top = nz - 1;
  const uint32_t o = lb + k * rowb;
  char *oc = (char *)tout + col * colb;
  const float v0 = v.v0;
  const float qB = rdl(cw, C_QB);
  const float fba = vla_c(v.zm2, v.zm1, v0, v.zp1, qB, rdl(cw, C_CB)) * qB;
  HVals h;
  h.w1 = v.w1;
  h.w2 = v.w2;
  h.e1 = v.e1;
  h.e2 = v.e2;
  h.s1 = v.s1;
  h.s2 = v.s2;
  h.n1 = v.n1;
  h.n2 = v.n2;
  const float hh = hflux_own(h, (m & M_E1) != 0, (m & M_N1) != 0, v0, rdl(cw, C_QW),
                             rdl(cw, C_CW), rdl(cw, C_QE), rdl(cw, C_CE), rdl(cw, C_QS),
                             rdl(cw, C_CS), rdl(cw, C_QN), rdl(cw, C_CN));
  const uint32_t key = srkey((uint32_t)(col * nz) + (uint32_t)k, hl);
  if (k < ks) {
    /* sub-surface update (tracers.c:912-925) in anomaly form; the top
     * face is the bottom face of the cell above */
    const float qT = rdl(cw, C_N + C_QB);
    const float ht = hh + vla_b(v.zm1, v0, v.zp1, v.zp2, qT, rdl(cw, C_N + C_CB)) * qT - fba;
    const double an = (double)v0 - (double)((ht + v0 * rdl(cw, C_D)) * rdl(cw, C_RV));
    float bk = sround(an, key);
    /* the thin-layer repair of tr_bounds needs the unclipped value */
    if (k == ks - 1 && (flags & F_THIN) && act) stb((char *)scr + col * colb, o, bk);
    if (clip && bk < mn) bk = mn;
    if (act) stb(oc, o, bk);
  } else {
    /* processed surface cell: surf_conc (tracers.c:3047-3207) as an
     * FP64 volume balance, v0 * Vc + anomaly terms, Vc in FP64 */
    const ColInfo ci = cinfo[col];
    const char *tc = tinb + col * colb;
    double m2 = (double)fba - (double)hh;
    if (flags & F_MOVE) {
      const int kb = ci.kb;
      int cc = ks;
      double wb = (ks == kb) ? ci.botz : layers[ks];
      while (cc != ci.kold && cc < top) {
        const double vol = ci.area * (layers[cc + 1] - wb);
        if (cc > ks) m2 += ((double)ldb(tc, lb + cc * rowb) - (double)v0) * vol;
        cc++;
        wb = layers[cc];
      }
      if (cc > ks)
        m2 += ((double)ldb(tc, lb + cc * rowb) - (double)v0) *
              (ci.area * (ci.osub - ((cc == kb) ? ci.botz : layers[cc])));
      Nbr nb;
      nb.c = tc;
      nb.w1 = tc - colb;
      nb.w2 = tc - 2 * colb;
      nb.e1 = tc + colb;
      nb.e2 = tc + 2 * colb;
      nb.s1 = tc - jb;
      nb.s2 = tc - 2 * jb;
      nb.n1 = tc + jb;
      nb.n2 = tc + 2 * jb;
      nb.kw1 = ci.kw1;
      nb.kw2 = ci.kw2;
      nb.ke1 = ci.ke1;
      nb.ke2 = ci.ke2;
      nb.ks1 = ci.ks1;
      nb.ks2 = ci.ks2;
      nb.kn1 = ci.kn1;
      nb.kn2 = ci.kn2;
      const float *cob = co + col * nz * C_N;
      for (int ka = ks + 1; ka <= top; ka++) {
        HVals ha;
        load_h(ha, nb, ka, lb + ka * rowb);
        const float *ca = cob + ka * C_N;
        m2 -= (double)hflux(ha, ka >= ci.ke1, ka >= ci.kn1, ldb(tc, lb + ka * rowb), v0,
                            ca[C_QW], ca[C_CW], ca[C_QE], ca[C_CE], ca[C_QS], ca[C_CS],
                            ca[C_QN], ca[C_CN]);
      }
    }
    double surftr = (double)v0 * ci.vc + m2;
    if (ci.svol == 0.0 || (surftr < 0.0 && ci.dzs <= hmin))
      surftr = 0.0;
    else
      surftr /= ci.svol;
    /* tracers.c:953-954; tr_bounds is applied in kernel B */
    const float trs = (flags & F_KEEPOLD) ? v0 : sround(surftr, key);
    if (act) stb(oc, o, trs);
  }
}

/*
 * Kernel A: every wet cell of every column (all tracers of one cell per
 * wave, lanes = tracers), cells taken in layer-synchronous tile order so
 * that the neighbours a cell reads were read by other waves moments
 * before.  Each wave slot has its own list of cells (descriptors and the
 * column's processed surface and flags, 64 at a time, lane-distributed),
 * and the stencil and coefficients of its next cell are loaded while the
 * current one is computed.  Writes the advected value b of sub-surface
 * cells and the raw new surface value (surf_conc) of the processed surface
 * cell to the output; tr_bounds and the vertical solve follow in kernel B.
 */
template <int SCH>
__global__ void __launch_bounds__(256)
k_advB(const AArgs a, const float *__restrict__ tin, float *__restrict__ tout,
       const float *__restrict__ co, const unsigned long long *__restrict__ cells,
       const int *__restrict__ kfl, const ColInfo *__restrict__ cinfo,
       const float *__restrict__ mintr, const double *__restrict__ layers,
       float *__restrict__ scr, const int *__restrict__ stepctr) {
  const int nz = a.nz, NT = a.ntr, nact = a.nact;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = min(lane, nact - 1); /* inactive lanes shadow the last tracer */
  const int ln = a.g * 64 + li;
  const bool act = lane < nact;
  const uint32_t lb = (uint32_t)ln * 4u;
  const uint32_t rowb = (uint32_t)NT * 4u;
  const ll colb = (ll)nz * NT * 4;
  const ll jb = a.nx * colb;
  const float mn = mintr[ln];
  const bool clip = mn <= 1e-5f;
  const uint32_t hl = mix32((uint32_t)ln * 0x85EBCA77U ^ (uint32_t)stepctr[0] * 0xC2B2AE3DU);
  const char *tinb = (const char *)tin;
  const int slot = blockIdx.x * 4 + wv;
  const unsigned long long *myc = cells + (ll)slot * a.niter;
  const int *myk = kfl + (ll)slot * a.niter;
  const int niter = a.niter;
  if (niter == 0) return;

  unsigned long long dcur = myc[lane], dnxt = NO_CELL;
  int kcur = myk[lane], knxt = 0;
  if (niter > 64) {
    dnxt = myc[64 + lane];
    knxt = myk[64 + lane];
  }
  /* Two cells per iteration with fixed stencil registers: the loads of
   * one cell are in flight while the other is computed.  Padding entries
   * load the rows of a stand-in cell (column 0, layer 0, no open
   * neighbours: always in bounds) so that every load is issued
   * unconditionally and is never computed. */
  auto live = [](unsigned long long d) -> unsigned long long { return d == NO_CELL ? 0ULL : d; };
  auto one = [&](const St13 &v, int cw, unsigned long long d, int kf) {
    if (d == NO_CELL) return;
    const int k = (int)((d >> 32) & 255);
    const int ks = (kf & 255) - 1, flags = kf >> 8;
    if (k > ks) return; /* above the processed surface, or not processed */
    adv_cell(v, cw, (ll)(uint32_t)d, k, (int)((d >> 40) & 255), ks, flags, tinb, tout, cinfo,
             co, layers, scr, lb, rowb, colb, jb, nz, hl, mn, clip, act, a.hmin);
  };
  St13 va, vb;
  int cwa = 0, cwb = 0;
  st13_load(va, cwa, tinb, co, live(rdl64(dcur, 0)), colb, jb, lb, rowb, nz, lane);
  for (int j = 0; j < niter; j += 2) { /* niter is a multiple of 64 */
    const int jl = j & 63;
    const unsigned long long dA = rdl64(dcur, jl), dB = rdl64(dcur, jl + 1);
    const int kfA = __builtin_amdgcn_readlane(kcur, jl);
    const int kfB = __builtin_amdgcn_readlane(kcur, jl + 1);
    st13_load(vb, cwb, tinb, co, live(dB), colb, jb, lb, rowb, nz, lane);
    one(va, cwa, dA, kfA);
    if (jl == 62) {
      dcur = dnxt;
      kcur = knxt;
      if (j + 66 < niter) {
        dnxt = myc[j + 66 + lane];
        knxt = myk[j + 66 + lane];
      }
    }
    const unsigned long long dN = (j + 2 < niter) ? rdl64(dcur, (jl + 2) & 63) : NO_CELL;
    st13_load(va, cwa, tinb, co, live(dN), colb, jb, lb, rowb, nz, lane);
    one(vb, cwb, dB, kfB);
  }
}

/* per list entry: the processed surface (+1) and flags of its column */
__global__ void k_kflB(const unsigned long long *__restrict__ cells, ll n,
                       const ColInfo *__restrict__ cinfo, int *__restrict__ kfl) {
  const ll i = (ll)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long d = cells[i];
  int v = 0;
  if (d != NO_CELL) {
    const ColInfo *c = cinfo + (uint32_t)d;
    v = (c->ks + 1) | (c->flags << 8);
  }
  kfl[i] = v;
}

/*-------------------------------------------------------------------*/
/* kernel B: tr_bounds and the vertical solve, one wave per column     */
/*-------------------------------------------------------------------*/
/*
 * Lanes = tracers.  Applies tr_bounds at the processed surface (thin old
 * surface layer, clip; tracers.c:3217-3324), fills the layers above with
 * the surface value, then solves (A/dzdt) d = b - (A/dzdt) b over the
 * diffusion extent (tracers.c:3782-3957, 4067-4214) with the LU factors
 * from the prep and writes x = b + d.  The advected values b are read
 * eight layers at a time; s_k and ud_k live in the wave's LDS slice.
 */

template <int SCH>
__global__ void __launch_bounds__(64 * WPB)
k_vertB(const KArgs a, const float *__restrict__ tin, float *__restrict__ tout,
        const float *__restrict__ co, const ColInfo *__restrict__ cinfo,
        const int *__restrict__ order, const float *__restrict__ mintr,
        const float *__restrict__ maxtr, const double *__restrict__ Kz,
        const double *__restrict__ dzf, float *__restrict__ scr,
        const int *__restrict__ stepctr) {
  extern __shared__ float lds[];
  const int nz = a.nz, NT = a.ntr, top = nz - 1, nact = a.nact;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = min(lane, nact - 1);
  float *Ls = lds + wv * nz * (nact + 1);
  float *Lud = Ls + nz * nact;
  const int ln = a.g * 64 + li;
  const bool act = lane < nact;
  const uint32_t lb = (uint32_t)ln * 4u;
  const uint32_t rowb = (uint32_t)NT * 4u;
  const ll colb = (ll)nz * NT * 4;
  const float mn = mintr[ln];
  const float mx = maxtr[ln];
  const bool clip = mn <= 1e-5f;
  const uint32_t hl = mix32((uint32_t)ln * 0x85EBCA77U ^ (uint32_t)stepctr[0] * 0xC2B2AE3DU) ^
                      0x5bd1e995U;
  const int nslot = gridDim.x * WPB;

  for (int q = blockIdx.x * WPB + wv; q < a.nord; q += nslot) {
    const int colq = order[q];
    if (colq < 0) continue;
    const ll col = colq;
    const ColInfo ci = cinfo[col];
    const int kb = ci.kb, ks = ci.ks;
    char *oc = (char *)tout + col * colb;
    const char *tc = (const char *)tin + col * colb;
    if (!(ci.flags & F_PROC)) {
      for (int k = kb; k <= top; k++)
        if (act) stb(oc, lb + k * rowb, ldb(tc, lb + k * rowb));
      continue;
    }
    /* tr_bounds at the surface */
    float trs = ldb(oc, lb + ks * rowb);
    if ((ci.flags & F_THIN) && (trs < mn || trs > mx))
      trs = ldb((const char *)scr + col * colb, lb + (ks - 1) * rowb);
    if (clip && trs < mn) trs = mn;
    for (int k = ks; k <= top; k++)
      if (act) stb(oc, lb + k * rowb, trs);
    if (ci.flags & F_SINGLE) continue;
    const int cb = ci.cb, cs = ci.cs;
    const float *cob = co + col * nz * C_N;
    const uint32_t hcol = (uint32_t)(col * nz);
    /* forward sweep: s_k = (r_k - gm_k s_{k-1}) / div_k,
     * r_k = -(gm (b_{k-1} - b_k) + gp (b_{k+1} - b_k)) */
    float sp = 0.0f, bkm = 0.0f, bk = 0.0f;
    for (int k0 = cb; k0 <= cs; k0 += 8) {
      /* coefficient words of layers k0..k0+3 and k0+4..k0+7, lane l holding
       * word l, issued with the b loads of the chunk */
      const int cA = __float_as_int(cob[k0 * C_N + lane]);
      const int cB = __float_as_int(cob[(k0 + 4) * C_N + lane]);
      float bb[9];
#pragma unroll
      for (int j = 0; j < 9; j++) {
        const int k = min(k0 + j, cs);
        bb[j] = (k < ks) ? ldb(oc, lb + k * rowb) : trs;
      }
      if (k0 == cb) {
        bk = bb[0];
        bkm = bk;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = k0 + j;
        if (k <= cs) {
          const int cv = (j < 4) ? cA : cB;
          const int w = 16 * (j & 3);
          const float gm = (k > cb) ? rdl(cv, w + C_GM) : 0.0f, gp = rdl(cv, w + C_GP);
          const float bkp = (k < cs) ? bb[j + 1] : bk;
          const float r = -(gm * (bkm - bk) + gp * (bkp - bk));
          sp = (r - gm * sp) * rdl(cv, w + C_RD);
          Ls[k * nact + li] = sp;
          Lud[k] = rdl(cv, w + C_UD);
          bkm = bk;
          bk = bkp;
        }
      }
    }
    /* back substitution: d_k = s_k - ud_{k+1} d_{k+1} */
    float dn = 0.0f;
    const bool stiff = (ci.flags & F_STIFF) != 0;
    for (int k0 = cs; k0 >= cb; k0 -= 8) {
      float bb[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = max(k0 - j, cb);
        bb[j] = (k < ks) ? ldb(oc, lb + k * rowb) : trs;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k = k0 - j;
        if (k >= cb) {
          const float sk = Ls[k * nact + li];
          dn = (k == cs) ? sk : sk - Lud[k + 1] * dn;
          if (stiff)
            Ls[k * nact + li] = dn;
          else if (act)
            stb(oc, lb + k * rowb,
                sround((double)bb[j] + (double)dn, srkey(hcol + (uint32_t)k, hl)));
        }
      }
    }
    if (stiff) {
      /* one FP64 residual pass of iterative refinement, the matrix and
       * right-hand side in FP64 from Kz and dz (tracers.c:3841-3889) */
      const ll P = a.plane;
      const double dt = a.dt;
      char *sc = (char *)scr + col * colb;
      float e = 0.0f;
      for (int k = cb; k <= cs; k++) {
        const ll ix = (ll)k * P + col;
        const double dzc = dzf[ix], dzdt = dzc / dt;
        double cm1 = 0.0, cp1 = 0.0;
        if (k > cb) cm1 = -Kz[ix] / (0.5 * (dzf[ix - P] + dzc));
        if (k < cs) cp1 = -Kz[ix + P] / (0.5 * (dzc + dzf[ix + P]));
        const double bv = (double)((k < ks) ? ldb(oc, lb + k * rowb) : trs);
        const double dk = (double)Ls[k * nact + li];
        const double bvm = (k > cb) ? (double)((k - 1 < ks) ? ldb(oc, lb + (k - 1) * rowb) : trs) : bv;
        const double bvp = (k < cs) ? (double)((k + 1 < ks) ? ldb(oc, lb + (k + 1) * rowb) : trs) : bv;
        const double dkm = (k > cb) ? (double)Ls[(k - 1) * nact + li] : 0.0;
        const double dkp = (k < cs) ? (double)Ls[(k + 1) * nact + li] : 0.0;
        const double rhs = -(cm1 * (bvm - bv) + cp1 * (bvp - bv));
        const double Ad = cm1 * dkm + (dzdt - cm1 - cp1) * dk + cp1 * dkp;
        const float rho = (float)((rhs - Ad) / dzdt);
        const float *cm = cob + k * C_N;
        const float gm = (k > cb) ? cm[C_GM] : 0.0f;
        e = (rho - gm * e) * cm[C_RD];
        stb(sc, lb + k * rowb, e);
      }
      float ec = 0.0f;
      for (int k = cs; k >= cb; k--) {
        const float ek = ldb(sc, lb + k * rowb);
        ec = (k == cs) ? ek : ek - Lud[k + 1] * ec;
        stb(sc, lb + k * rowb, ec);
      }
      /* every residual is formed before any b_k is replaced */
      for (int k = cb; k <= cs; k++) {
        const float bv = (k < ks) ? ldb(oc, lb + k * rowb) : trs;
        const float ek = ldb(sc, lb + k * rowb);
        if (act)
          stb(oc, lb + k * rowb,
              sround((double)bv + (double)Ls[k * nact + li] + (double)ek,
                     srkey(hcol + (uint32_t)k, hl)));
      }
    }
  }
}

/* advance the stochastic-rounding step counter (runs after the step) */
__global__ void k_tick(int *ctr) { *ctr += 1; }

/*-------------------------------------------------------------------*/
/* layout conversion and mass                                        */
/*-------------------------------------------------------------------*/
__global__ void k_putB(const int *__restrict__ wsp, const ll *__restrict__ wdn, int nwet,
                       ll plane, int nz, int NT, int n, const double *__restrict__ sp,
                       float *__restrict__ a, float *__restrict__ a2) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nwet) return;
  const ll idx = wdn[q];
  const ll col = idx % plane;
  const int k = (int)(idx / plane);
  const ll e = cidx(col, nz, k) * NT + n;
  const float v = (float)sp[wsp[q]];
  a[e] = v;
  a2[e] = v;
}

__global__ void k_getB(const int *__restrict__ wsp, const ll *__restrict__ wdn, int nwet,
                       ll plane, int nz, int NT, int n, const float *__restrict__ a,
                       double *__restrict__ sp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nwet) return;
  const ll idx = wdn[q];
  const ll col = idx % plane;
  const int k = (int)(idx / plane);
  sp[wsp[q]] = (double)a[cidx(col, nz, k) * NT + n];
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

/* per-block (sum, err) of sum(tr*area*dz) over the processed cells */
__global__ void k_massB(emstr_dev d, emstr_batch b, const float *__restrict__ tr,
                        double *__restrict__ part) {
  const int n = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double ss[256], se[256];
  double s = 0.0, e = 0.0;
  if (q < d.ncol) {
    const ll col = d.cols[q];
    const int kbc = d.kb[col], ksc = d.ks[col];
    const double a = d.cellarea[col];
    for (int k = ksc; k >= kbc; k--) {
      double m = (double)tr[cidx(col, b.nz, k) * b.ntr + n] * a * d.dz[(ll)k * d.plane + col];
      double s2, e2;
      two_sum(s, m, s2, e2);
      s = s2;
      e += e2;
    }
  }
  ss[threadIdx.x] = s;
  se[threadIdx.x] = e;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double s2, e2;
      two_sum(ss[threadIdx.x], ss[threadIdx.x + w], s2, e2);
      ss[threadIdx.x] = s2;
      se[threadIdx.x] += se[threadIdx.x + w] + e2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ((ll)n * gridDim.x + blockIdx.x)] = ss[0];
    part[2 * ((ll)n * gridDim.x + blockIdx.x) + 1] = se[0];
  }
}

/*-------------------------------------------------------------------*/
/* host interface                                                    */
/*-------------------------------------------------------------------*/
/* dynamic LDS of kernel B: one [nz][nact + 1] slice per wave */
static size_t vert_lds(int nz, int nact) {
  return (size_t)WPB * nz * (nact + 1) * sizeof(float);
}

/* Cell lists of kernel A.  Cells are ordered tile by tile (tile_x x
 * tile_y columns, row-major tiles), inside a tile layer by layer from the
 * deepest, inside a layer row-major; the sequence is split into 8
 * contiguous bands, one per XCD (workgroups are dealt to the XCDs round
 * robin).  The S wave slots of XCD x take band x in turns: cell c of the
 * band is iteration c / S of the slot of rank c % S, so at any time the
 * slots of one XCD work on one stretch of one tile layer. */
static int build_cells(emstr_batch *b, const int *kb, int nz) {
  const int nx = b->nx, ny = b->ny, tx = b->tile_x, ty = b->tile_y;
  ll n = 0;
  for (ll c = 0; c < b->plane; c++)
    if (kb[c] < nz) n += nz - kb[c];
  unsigned long long *seq = (unsigned long long *)malloc(sizeof(unsigned long long) * (size_t)(n + 1));
  ll q = 0;
  for (int J = 0; J < ny; J += ty)
    for (int I = 0; I < nx; I += tx)
      for (int k = 0; k < nz; k++)
        for (int j = J; j < J + ty && j < ny; j++)
          for (int i = I; i < I + tx && i < nx; i++) {
            const ll col = (ll)j * nx + i;
            if (kb[col] > k) continue;
            const bool w1 = kb[col - 1] <= k, e1 = kb[col + 1] <= k;
            const bool s1 = kb[col - nx] <= k, n1 = kb[col + nx] <= k;
            const bool w2 = w1 && kb[col - 2] <= k, e2 = e1 && kb[col + 2] <= k;
            const bool s2 = s1 && kb[col - 2 * nx] <= k, n2 = n1 && kb[col + 2 * nx] <= k;
            const unsigned m = (w1 ? M_W1 : 0) | (w2 ? M_W2 : 0) | (e1 ? M_E1 : 0) |
                               (e2 ? M_E2 : 0) | (s1 ? M_S1 : 0) | (s2 ? M_S2 : 0) |
                               (n1 ? M_N1 : 0) | (n2 ? M_N2 : 0);
            seq[q++] = (unsigned long long)(uint32_t)col | ((unsigned long long)k << 32) |
                       ((unsigned long long)m << 40) | ((unsigned long long)kb[col] << 48);
          }
  const ll nslot = (ll)b->gridA * 4, S = nslot / 8;
  const ll P = (n + 7) / 8;
  ll niter = (P + S - 1) / S;
  niter = (niter + 63) / 64 * 64;
  if (niter == 0) niter = 64;
  const ll total = nslot * niter;
  unsigned long long *lst = (unsigned long long *)malloc(sizeof(unsigned long long) * (size_t)total);
  for (ll t = 0; t < total; t++) lst[t] = NO_CELL;
  for (int x = 0; x < 8; x++)
    for (ll c = 0; c < P; c++) {
      const ll src = x * P + c;
      if (src >= n) break;
      const ll r = c % S, it = c / S;
      const ll slot = ((r / 4) * 8 + x) * 4 + (r % 4);
      lst[slot * niter + it] = seq[src];
    }
  if (b->cells) (void)hipFree(b->cells);
  if (b->kfl) (void)hipFree(b->kfl);
  b->cells = NULL;
  b->kfl = NULL;
  bool ok = hipMalloc(&b->cells, sizeof(unsigned long long) * (size_t)total) == hipSuccess;
  ok &= hipMalloc(&b->kfl, sizeof(int) * (size_t)total) == hipSuccess;
  if (ok)
    ok = hipMemcpy(b->cells, lst, sizeof(unsigned long long) * (size_t)total,
                   hipMemcpyHostToDevice) == hipSuccess;
  free(lst);
  free(seq);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "batched form: cell lists");
  b->niter = (int)niter;
  b->nlist = total;
  return 0;
}

/* Column order: tiles of tile_x x tile_y columns in row-major tile
 * order, columns row-major inside a tile; the wet columns of that order
 * are split into 8 contiguous bands.  Position q of the persistent loop is
 * taken by wave q % WPB of workgroup q / WPB, and workgroups are dealt to
 * the 8 XCDs round robin, so position q maps to band (q / WPB) % 8 and the
 * waves resident on one XCD walk one band together. */
static void build_order(emstr_batch *b, const int *kb, int nz) {
  const int nx = b->nx, ny = b->ny, tx = b->tile_x, ty = b->tile_y;
  int *lin = (int *)malloc(sizeof(int) * (size_t)b->plane);
  int n = 0;
  for (int J = 0; J < ny; J += ty)
    for (int I = 0; I < nx; I += tx)
      for (int j = J; j < J + ty && j < ny; j++)
        for (int i = I; i < I + tx && i < nx; i++) {
          const ll col = (ll)j * nx + i;
          if (kb[col] < nz) lin[n++] = (int)col;
        }
  b->ncol = n;
  /* positions per band, a multiple of WPB */
  const int per = ((n + 7) / 8 + WPB - 1) / WPB * WPB;
  b->nord = 8 * per;
  int *ord = (int *)malloc(sizeof(int) * (size_t)(b->nord > 0 ? b->nord : 1));
  for (int q = 0; q < b->nord; q++) {
    const int m = q / WPB;
    const int band = m % 8, idx = (m / 8) * WPB + q % WPB;
    const int src = band * ((n + 7) / 8) + idx;
    ord[q] = (idx < (n + 7) / 8 && src < n) ? lin[src] : -1;
  }
  hipMemcpy(b->order, ord, sizeof(int) * (size_t)b->nord, hipMemcpyHostToDevice);
  free(ord);
  free(lin);
}

extern "C" struct emstr_batch *emstr_batch_create(const emstr_dev *d, const int *h_kb,
                                                  void *stream) {
  (void)stream;
  if (d->nz > NZMAX) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: at most %d layers", NZMAX);
    return NULL;
  }
  emstr_batch *b = (emstr_batch *)calloc(1, sizeof(emstr_batch));
  b->ntr = d->ntr;
  b->nt64 = (d->ntr + 63) / 64;
  b->nz = d->nz;
  b->nx = d->nx;
  b->ny = d->ny;
  b->plane = d->plane;
  b->tile_x = 32;
  b->tile_y = 32;
  const size_t ncell = (size_t)d->plane * d->nz;
  bool ok = true;
  for (int q = 0; q < 2; q++)
    ok &= hipMalloc(&b->tr[q], ncell * d->ntr * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->scr, ncell * d->ntr * sizeof(float)) == hipSuccess;
  /* 8 layers of padding: the walk reads coefficients in 4-layer chunks */
  ok &= hipMalloc(&b->co, (ncell + 8) * C_N * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->ci, (size_t)d->plane * sizeof(ColInfo)) == hipSuccess;
  ok &= hipMalloc(&b->order, ((size_t)d->plane + 8 * WPB + 8) * sizeof(int)) == hipSuccess;
  ok &= hipMalloc(&b->mintr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->maxtr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->stepctr, sizeof(int)) == hipSuccess;
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "batched form: device allocation failed");
    emstr_batch_destroy(b);
    return NULL;
  }
  hipMemset(b->co, 0, (ncell + 8) * C_N * sizeof(float));
  hipMemset(b->stepctr, 0, sizeof(int));
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  /* persistent grids: as many workgroups as can be resident at once */
  int per_cu = 0;
  const size_t lds = vert_lds(b->nz, b->ntr < 64 ? b->ntr : 64);
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void *)k_vertB<VANLEER>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_vertB<VANLEER>,
                                                   64 * WPB, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  b->grid = ncu * per_cu;
  per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_advB<VANLEER>, 256,
                                                   0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  b->gridA = (ncu * per_cu + 7) / 8 * 8;
  b->grid_full = b->grid;
  b->gridA_full = b->gridA;
  build_order(b, h_kb, d->nz);
  if (build_cells(b, h_kb, d->nz)) {
    emstr_batch_destroy(b);
    return NULL;
  }
  return b;
}

extern "C" void emstr_batch_destroy(struct emstr_batch *b) {
  if (!b) return;
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  void *v[] = {b->tr[0], b->tr[1], b->scr,     b->co,    b->ci,  b->order,
               b->mintr, b->maxtr, b->stepctr, b->cells, b->kfl};
  for (int q = 0; q < 11; q++)
    if (v[q]) hipFree(v[q]);
  free(b);
}

extern "C" void emstr_batch_set_prof(struct emstr_batch *b, struct emstr_prof *p) { b->prof = p; }

extern "C" int emstr_batch_tune(struct emstr_batch *b, const int *h_kb, int tile_x, int tile_y,
                                int grid) {
  if (tile_x > 0) b->tile_x = tile_x;
  if (tile_y > 0) b->tile_y = tile_y;
  if (grid > 0) {
    /* the advection grid as asked; the column grid in proportion */
    b->gridA = (grid + 7) / 8 * 8;
    long long g2 = (long long)b->grid_full * b->gridA / b->gridA_full;
    b->grid = (int)(g2 < 1 ? 1 : (g2 > b->grid_full ? b->grid_full : g2));
  }
  build_order(b, h_kb, b->nz);
  if (build_cells(b, h_kb, b->nz)) return -1;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec), b->gexec = NULL;
  if (b->graph) (void)hipGraphDestroy(b->graph), b->graph = NULL;
  b->graph_ok = 0;
  return 0;
}

extern "C" void emstr_batch_shape(const struct emstr_batch *b, int *tile_x, int *tile_y,
                                  int *grid, int *grid_full) {
  *tile_x = b->tile_x;
  *tile_y = b->tile_y;
  *grid = b->gridA;
  *grid_full = b->gridA_full;
}

extern "C" int emstr_batch_fields(struct emstr_batch *b, const emstr_dev *d, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (d->trasc != VANLEER || d->ultimate)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: van Leer without ULTIMATE");
  float *mn = (float *)malloc(sizeof(float) * (d->ntr + 1));
  float *mx = (float *)malloc(sizeof(float) * (d->ntr + 1));
  double *h = (double *)malloc(sizeof(double) * (d->ntr + 1));
  hipMemcpy(h, d->mintr, sizeof(double) * d->ntr, hipMemcpyDeviceToHost);
  for (int n = 0; n < d->ntr; n++) mn[n] = (float)h[n];
  hipMemcpy(h, d->maxtr, sizeof(double) * d->ntr, hipMemcpyDeviceToHost);
  for (int n = 0; n < d->ntr; n++) mx[n] = (float)h[n];
  hipMemcpy(b->mintr, mn, sizeof(float) * d->ntr, hipMemcpyHostToDevice);
  hipMemcpy(b->maxtr, mx, sizeof(float) * d->ntr, hipMemcpyHostToDevice);
  free(mn);
  free(mx);
  free(h);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec), b->gexec = NULL;
  if (b->graph) (void)hipGraphDestroy(b->graph), b->graph = NULL;
  b->graph_ok = 0;
  if (emstr_launch_subeta(d, d->dt, d->dt, 1, st)) return -1;
  return 0;
}

/* the per-step work: prep + fused kernel (captured in a graph) */
static int batch_launch(emstr_batch *b, const emstr_dev *d, hipStream_t st) {
  dim3 bp(64, 4, 1);
  dim3 gp((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
  int pi = prof_begin(b->prof, KC_PREP, st);
  k_prepB<<<gp, bp, 0, st>>>(*d, *b);
  k_colB<<<dim3((d->nx + 63) / 64, (d->ny + 3) / 4, 1), bp, 0, st>>>(*d, *b);
  const float *in = b->tr[b->cur];
  float *out = b->tr[1 - b->cur];
  if (b->nlist > 0)
    k_kflB<<<(unsigned)((b->nlist + 255) / 256), 256, 0, st>>>(b->cells, b->nlist, b->ci,
                                                                 b->kfl);
  prof_end(b->prof, pi, st);
  KArgs ka;
  ka.nz = b->nz;
  ka.ntr = b->ntr;
  ka.nord = b->nord;
  ka.nx = b->nx;
  ka.hmin = d->hmin;
  ka.dt = d->dt;
  ka.plane = d->plane;
  ka.pad = 0;
  AArgs aa;
  aa.nz = b->nz;
  aa.ntr = b->ntr;
  aa.niter = b->niter;
  aa.pad = 0;
  aa.nx = b->nx;
  aa.hmin = d->hmin;
  for (int g = 0; g < b->nt64; g++) {
    ka.g = aa.g = g;
    ka.nact = aa.nact = b->ntr - 64 * g < 64 ? b->ntr - 64 * g : 64;
    pi = prof_begin(b->prof, KC_ADV, st);
    k_advB<VANLEER><<<b->gridA, 256, 0, st>>>(aa, in, out, b->co, b->cells, b->kfl, b->ci,
                                              b->mintr, d->layers, b->scr, b->stepctr);
    prof_end(b->prof, pi, st);
    pi = prof_begin(b->prof, KC_VDIFF, st);
    k_vertB<VANLEER><<<b->grid, 64 * WPB, vert_lds(b->nz, ka.nact), st>>>(
        ka, in, out, b->co, b->ci, b->order, b->mintr, b->maxtr, d->Kz, d->dz, b->scr,
        b->stepctr);
    prof_end(b->prof, pi, st);
  }
  k_tick<<<1, 1, 0, st>>>(b->stepctr);
  b->cur = 1 - b->cur;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_batch_steps(struct emstr_batch *b, const emstr_dev *d, int stab,
                                 int nsteps, void *stream, int use_graph) {
  hipStream_t st = (hipStream_t)stream;
  if (stab & 22) return emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: no sub-stepping");
  const bool graph = use_graph && !(b->prof && b->prof->on);
  while (nsteps > 0) {
    if (graph && nsteps >= 2 && (!b->graph_ok || b->graph_cur == b->cur)) {
      /* the graph holds two steps so that the buffers end where they
       * started; kernel arguments (buffer pointers) are fixed at capture */
      if (!b->graph_ok) {
        const int cur0 = b->cur;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return emstr_set_error(EMSTR_EDEVICE, -1, "graph capture");
        batch_launch(b, d, st);
        batch_launch(b, d, st);
        if (hipStreamEndCapture(st, &b->graph) != hipSuccess ||
            hipGraphInstantiate(&b->gexec, b->graph, NULL, NULL, 0) != hipSuccess)
          return emstr_set_error(EMSTR_EDEVICE, -1, "graph instantiate");
        b->cur = cur0;
        b->graph_cur = cur0;
        b->graph_ok = 1;
      }
      if (hipGraphLaunch(b->gexec, st) != hipSuccess)
        return emstr_set_error(EMSTR_EDEVICE, -1, "graph launch");
      nsteps -= 2;
    } else {
      if (batch_launch(b, d, st)) return emstr_set_error(EMSTR_EDEVICE, -1, "batch launch");
      nsteps -= 1;
    }
  }
  return 0;
}

/* two steps recorded onto a stream that is being captured (an ensemble
 * graph branch); the buffers end where they started */
extern "C" int emstr_batch_record2(struct emstr_batch *b, const emstr_dev *d, int stab,
                                   void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (stab & 22) return emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: no sub-stepping");
  if (b->prof && b->prof->on)
    return emstr_set_error(EMSTR_EARG, -1, "ensemble capture with kernel timing on");
  const int cur0 = b->cur;
  if (batch_launch(b, d, st) || batch_launch(b, d, st))
    return emstr_set_error(EMSTR_EDEVICE, -1, "ensemble capture");
  b->cur = cur0;
  return 0;
}

extern "C" int emstr_batch_put(struct emstr_batch *b, const emstr_dev *d, int n,
                               const double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_putB<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      d->wsp, d->wdn, d->nwet, d->plane, d->nz, b->ntr, n, sparse, b->tr[b->cur],
      b->tr[1 - b->cur]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_batch_get(struct emstr_batch *b, const emstr_dev *d, int n,
                               double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_getB<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      d->wsp, d->wdn, d->nwet, d->plane, d->nz, b->ntr, n, b->tr[b->cur], sparse);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_batch_mass(struct emstr_batch *b, const emstr_dev *d, double *part,
                                void *stream) {
  int nb = (d->ncol + 255) / 256;
  if (nb == 0) return 0;
  k_massB<<<dim3(nb, d->ntr, 1), 256, 0, (hipStream_t)stream>>>(*d, *b, b->tr[b->cur], part);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
