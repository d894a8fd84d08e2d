/*
 * Batched form of the EMS tracer step on MI355X (gfx950): every tracer of
 * the step and every water column in one fused persistent kernel.
 *
 * Mapping: one wavefront per water column, one lane per tracer.  Tracers
 * are stored tracer-innermost, [column][layer][tracer] in FP32, so the
 * tracers of one cell are one contiguous 4*ntr-byte load and a column
 * walk reads each layer once.  Everything that does not depend on the
 * tracer (face transports, Courant numbers, the column matrix, wetness of
 * neighbours, the surface volume balance) is per cell or per column and
 * is read with scalar loads, once, for all tracers.
 *
 * Per column the wave walks up from the bed:
 *   - face values (van Leer, tracers.c:2097-2118) on the 13-point
 *     stencil; the vertical part of the stencil is a register window;
 *   - the flux divergence and sub-surface update (tracers.c:888-925) in
 *     anomaly form:  tr - (sum_f s_f (F_f - tr) q_f + tr*D) / V
 *     with D the FP64 net volume divergence of the cell; this is the
 *     reference's tr - (sum_f s_f F_f q_f)/V, but rounding the transports
 *     q to FP32 only perturbs the small (F - tr) terms;
 *   - surf_conc (tracers.c:3047-3207) as an FP64 volume balance with the
 *     column's consistent volume precomputed in FP64;
 *   - tr_bounds (tracers.c:3217-3324): thin-layer repair, clip, copy up;
 *   - implicit vertical diffusion (tracers.c:3782-3957, 4067-4214) in
 *     increment form: A (x - b) = dzdt*b - A b, so constants are kept
 *     exactly; the forward sweep runs inside the walk, the increments are
 *     held in LDS and the back substitution runs on the way down;
 *   - one FP64 residual pass of iterative refinement in columns whose
 *     diffusion number exceeds IR_DMAX.
 * FP64 stays in the coefficient prep (layer thicknesses from set_dz, the
 * divergence D, the surface volume balance, the refinement residual) and
 * in the mass budgets.
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
};

__device__ __forceinline__ ll cidx(ll col, int nz, int k) { return col * nz + k; }

/*-------------------------------------------------------------------*/
/* prep: per-cell FP32 coefficients from the FP64 fields              */
/*-------------------------------------------------------------------*/
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
  float *c = b.co + cidx(col, d.nz, k) * C_N;
  c[C_QW] = (float)qW;
  c[C_CW] = (float)(d.u1[idx] * dt / d.h1au1[col]);
  c[C_QE] = (float)qE;
  c[C_CE] = we ? (float)(d.u1[idx + 1] * dt / d.h1au1[col + 1]) : 0.0f;
  c[C_QS] = (float)qS;
  c[C_CS] = (float)(d.u2[idx] * dt / d.h2au2[col]);
  c[C_QN] = (float)qN;
  c[C_CN] = wn ? (float)(d.u2[idx + nx] * dt / d.h2au2[col + nx]) : 0.0f;
  c[C_QB] = (float)qB;
  c[C_CB] = (float)(d.w[idx] * dt / (0.5 * (dz + dzb)));
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
      int q = cs;
      while (q >= kbc && d.dz[(ll)q * P + col] != 0.0) q--;
      c.cb = q + 1;
      if (c.cb >= cs) flags |= F_SINGLE;
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
 * face between bm1 and c, with Courant number cn.  The limiter's division
 * uses the hardware reciprocal (about 1 ulp). */
__device__ __forceinline__ float vlf(float bm2, float bm1, float c, float fp1,
                                     float q, float cn) {
  float df;
  if (q > 0.0f) {
    df = (bm1 - bm2) * (c - bm1);
    df = (df > 0.0f) ? __fdividef(2.0f * df, c - bm2) : 0.0f;
    return bm1 + 0.5f * (1.0f - cn) * df;
  } else if (q < 0.0f) {
    df = (c - bm1) * (fp1 - c);
    df = (df > 0.0f) ? __fdividef(2.0f * df, fp1 - bm1) : 0.0f;
    return c - 0.5f * (1.0f + cn) * df;
  }
  return c; /* no transport: F - c = 0 */
}

/* Offsets (in floats) from a cell to its twelve stencil neighbours.  A
 * neighbour that is land or below the bed is replaced by the cell the map
 * walk came from, whose ghost holds that value after the lateral boundary
 * fill (tracers.c:5574-5593); the bed maps down to its sediment ghost
 * (the bed value, tracers.c:827) and the top layer maps up to itself.
 * Everything here is uniform across the wave. */
struct Sten {
  ll w1, w2, e1, e2, s1, s2, n1, n2, zm1, zm2, zp1, zp2;
  bool ew, nw; /* east / north face open */
};

__device__ __forceinline__ void sten(Sten &s, int k, int kb, int top, const ColInfo &ci,
                                     ll zstep, ll cstep, ll rstep) {
  const bool w1 = k >= ci.kw1, e1 = k >= ci.ke1, s1 = k >= ci.ks1, n1 = k >= ci.kn1;
  s.w1 = w1 ? -cstep : 0;
  s.e1 = e1 ? cstep : 0;
  s.w2 = w1 ? ((k >= ci.kw2) ? -2 * cstep : s.w1) : 0;
  s.e2 = e1 ? ((k >= ci.ke2) ? 2 * cstep : s.e1) : 0;
  s.s1 = s1 ? -rstep : 0;
  s.n1 = n1 ? rstep : 0;
  s.s2 = s1 ? ((k >= ci.ks2) ? -2 * rstep : s.s1) : 0;
  s.n2 = n1 ? ((k >= ci.kn2) ? 2 * rstep : s.n1) : 0;
  s.zm1 = (ll)(max(k - 1, kb) - k) * zstep;
  s.zm2 = (ll)(max(k - 2, kb) - k) * zstep;
  s.zp1 = (ll)(min(k + 1, top) - k) * zstep;
  s.zp2 = (ll)(min(k + 2, top) - k) * zstep;
  s.ew = e1;
  s.nw = n1;
}

/* The stencil values of one cell for one tracer lane; pc is the cell's
 * address for lane 0, the lane reads pc[off + ln].  All loads are issued
 * together. */
struct Vals {
  float v0, xw1, xw2, xe1, xe2, ys1, ys2, yn1, yn2, zm1, zm2, zp1, zp2;
};

__device__ __forceinline__ void load_vals(Vals &v, const float *__restrict__ pc, unsigned ln,
                                          const Sten &s, bool vert) {
  v.v0 = pc[ln];
  v.xw1 = pc[s.w1 + ln];
  v.xw2 = pc[s.w2 + ln];
  v.xe1 = pc[s.e1 + ln];
  v.xe2 = pc[s.e2 + ln];
  v.ys1 = pc[s.s1 + ln];
  v.ys2 = pc[s.s2 + ln];
  v.yn1 = pc[s.n1 + ln];
  v.yn2 = pc[s.n2 + ln];
  if (vert) {
    v.zm1 = pc[s.zm1 + ln];
    v.zm2 = pc[s.zm2 + ln];
    v.zp1 = pc[s.zp1 + ln];
    v.zp2 = pc[s.zp2 + ln];
  }
}

/* Horizontal anomaly outflow sum_f s_f (F_f - ref) q_f of a cell */
__device__ __forceinline__ float hsum(const Vals &v, const Sten &s,
                                      const float *__restrict__ c, float ref) {
  const float FW = vlf(v.xw2, v.xw1, v.v0, v.xe1, c[C_QW], c[C_CW]);
  const float FE = s.ew ? vlf(v.xw1, v.v0, v.xe1, v.xe2, c[C_QE], c[C_CE]) : ref;
  const float FS = vlf(v.ys2, v.ys1, v.v0, v.yn1, c[C_QS], c[C_CS]);
  const float FN = s.nw ? vlf(v.ys1, v.v0, v.yn1, v.yn2, c[C_QN], c[C_CN]) : ref;
  return (FE - ref) * c[C_QE] - (FW - ref) * c[C_QW] + (FN - ref) * c[C_QN] -
         (FS - ref) * c[C_QS];
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
/* the fused step: one workgroup (CW waves) per water column          */
/*-------------------------------------------------------------------*/
#define CW 4 /* waves per column */

struct KArgs {
  int nz, ntr, nord, g; /* g: which 64 tracers this launch takes */
  ll nx;
  double hmin, dt;
  ll plane;
};

template <int SCH>
__global__ void __launch_bounds__(64 * CW)
k_col(const KArgs a, const float *__restrict__ tin, float *__restrict__ tout,
      const float *__restrict__ co, const ColInfo *__restrict__ cinfo,
      const int *__restrict__ order, const float *__restrict__ mintr,
      const float *__restrict__ maxtr, const double *__restrict__ layers,
      const double *__restrict__ Kz, const double *__restrict__ dzf,
      float *__restrict__ scr, const int *__restrict__ stepctr) {
  /* LDS per column: Lb[nz][64] advected values, Ls[nz][64] forward sweep
   * then increments, Lpre[64] unclipped update of the layer below the
   * surface, Ltrs[64] new surface value */
  extern __shared__ float lds[];
  const int nz = a.nz, NT = a.ntr, top = nz - 1;
  float *Lb = lds;
  float *Ls = lds + nz * 64;
  float *Lpre = lds + 2 * nz * 64;
  float *Ltrs = Lpre + 64;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = a.g * 64 + lane;
  const bool act = n < NT;
  const unsigned ln = act ? n : NT - 1; /* inactive lanes shadow the last tracer */
  const ll zstep = NT;
  const ll cstep = (ll)nz * NT;
  const ll rstep = a.nx * cstep;
  const float mn = mintr[ln];
  const float mx = maxtr[ln];
  const bool clip = mn <= 1e-5f;
  const uint32_t hl = mix32(ln * 0x85EBCA77U ^ (uint32_t)stepctr[0] * 0xC2B2AE3DU);

  for (int q = blockIdx.x; q < a.nord; q += gridDim.x) {
    const int colq = order[q];
    if (colq < 0) continue;
    const ll col = colq;
    const ColInfo ci = cinfo[col];
    const int kb = ci.kb, ks = ci.ks;
    const float *tc = tin + col * cstep; /* the column, lane 0 */
    float *oc = tout + col * cstep;
    const float *cob = co + col * nz * C_N;
    const uint32_t hcol = (uint32_t)(col * nz) * 0x9E3779B1U;
    if (!(ci.flags & F_PROC)) {
      for (int k = kb + wv; k <= top; k += CW)
        if (act) oc[k * zstep + ln] = tc[k * zstep + ln];
      continue;
    }
    /* ---- phase 1: the layers up to the processed surface, spread over
     * the waves ---- */
    for (int k = kb + wv; k <= ks; k += CW) {
      const float *c = cob + k * C_N;
      const float *pc = tc + k * zstep;
      Sten s;
      sten(s, k, kb, top, ci, zstep, cstep, rstep);
      Vals v;
      load_vals(v, pc, ln, s, true);
      const float v0 = v.v0;
      const float FB = vlf(v.zm2, v.zm1, v0, v.zp1, c[C_QB], c[C_CB]);
      const float fba = (FB - v0) * c[C_QB];
      const float h = hsum(v, s, c, v0);
      const uint32_t key = mix32((hcol + (uint32_t)k * 0x9E3779B1U) ^ hl);
      if (k < ks) {
        /* sub-surface update (tracers.c:912-925) in anomaly form; the
         * top face is the bottom face of the cell above */
        const float qT = c[C_N + C_QB];
        const float FT = vlf(v.zm1, v0, v.zp1, v.zp2, qT, c[C_N + C_CB]);
        const float ht = h + (FT - v0) * qT - fba;
        const double an = (double)v0 - (double)((ht + v0 * c[C_D]) * c[C_RV]);
        float af = sround(an, key);
        if (k == ks - 1) Lpre[lane] = af;
        if (clip && af < mn) af = mn;
        Lb[k * 64 + lane] = af;
      } else {
        /* processed surface cell: surf_conc (tracers.c:3047-3207) as an
         * FP64 volume balance, v0 * Vc + anomaly terms, Vc in FP64 */
        double m = (double)fba - (double)h;
        if (ci.flags & F_MOVE) {
          int cc = ks;
          double wb = (ks == kb) ? ci.botz : layers[ks];
          while (cc != ci.kold && cc < top) {
            const double vol = ci.area * (layers[cc + 1] - wb);
            if (cc > ks) m += ((double)tc[cc * zstep + ln] - (double)v0) * vol;
            cc++;
            wb = layers[cc];
          }
          if (cc > ks)
            m += ((double)tc[cc * zstep + ln] - (double)v0) *
                 (ci.area * (ci.osub - ((cc == kb) ? ci.botz : layers[cc])));
          for (int ka = ks + 1; ka <= top; ka++) {
            Sten sa;
            sten(sa, ka, kb, top, ci, zstep, cstep, rstep);
            Vals va;
            load_vals(va, tc + ka * zstep, ln, sa, false);
            m -= (double)hsum(va, sa, cob + ka * C_N, v0);
          }
        }
        double surftr = (double)v0 * ci.vc + m;
        if (ci.svol == 0.0 || (surftr < 0.0 && ci.dzs <= a.hmin))
          surftr = 0.0;
        else
          surftr /= ci.svol;
        /* tracers.c:953-954 */
        Ltrs[lane] = (ci.flags & F_KEEPOLD) ? v0 : sround(surftr, key);
      }
    }
    __syncthreads();
    /* ---- phase 2 (one wave): tr_bounds at the surface
     * (tracers.c:3217-3324), then implicit vertical diffusion in increment
     * form: (A/dzdt) d = b - (A/dzdt) b, x = b + d ---- */
    if (wv == 0) {
      float trs = Ltrs[lane];
      if ((ci.flags & F_THIN) && (trs < mn || trs > mx)) trs = Lpre[lane];
      if (clip && trs < mn) trs = mn;
      for (int k = ks; k <= top; k++) Lb[k * 64 + lane] = trs;
      if (ci.flags & F_SINGLE) {
        for (int k = kb; k <= top; k++)
          if (act) oc[k * zstep + ln] = Lb[k * 64 + lane];
      } else {
        const int cb = ci.cb, cs = ci.cs;
        for (int k = kb; k < cb; k++)
          if (act) oc[k * zstep + ln] = Lb[k * 64 + lane];
        for (int k = cs + 1; k <= top; k++)
          if (act) oc[k * zstep + ln] = Lb[k * 64 + lane];
        /* forward sweep with the factorisation from the prep:
         * s_k = (r_k - gm_k s_{k-1}) / div_k */
        float sp = 0.0f;
        float bkm = Lb[cb * 64 + lane], bk = bkm;
        for (int k = cb; k <= cs; k++) {
          const float *cm = cob + k * C_N;
          const float bkp = (k < cs) ? Lb[(k + 1) * 64 + lane] : bk;
          const float gm = (k > cb) ? cm[C_GM] : 0.0f, gp = cm[C_GP];
          const float r = -(gm * (bkm - bk) + gp * (bkp - bk));
          sp = (r - gm * sp) * cm[C_RD];
          Ls[k * 64 + lane] = sp;
          bkm = bk;
          bk = bkp;
        }
        const bool stiff = (ci.flags & F_STIFF) != 0;
        float dn = sp;
        for (int k = cs; k >= cb; k--) {
          if (k < cs) {
            dn = Ls[k * 64 + lane] - cob[(k + 1) * C_N + C_UD] * dn;
            Ls[k * 64 + lane] = dn;
          }
          if (!stiff && act)
            oc[k * zstep + ln] =
                sround((double)Lb[k * 64 + lane] + (double)dn,
                       mix32((hcol + (uint32_t)k * 0x9E3779B1U) ^ hl ^ 0x5bd1e995U));
        }
        if (stiff) {
          /* one FP64 residual pass of iterative refinement, the matrix
           * and right-hand side in FP64 from Kz and dz
           * (tracers.c:3841-3889) */
          const ll P = a.plane;
          const double dt = a.dt;
          float *sc = scr + col * cstep;
          float e = 0.0f;
          for (int k = cb; k <= cs; k++) {
            const ll ix = (ll)k * P + col;
            const double dzc = dzf[ix], dzdt = dzc / dt;
            double cm1 = 0.0, cp1 = 0.0;
            if (k > cb) cm1 = -Kz[ix] / (0.5 * (dzf[ix - P] + dzc));
            if (k < cs) cp1 = -Kz[ix + P] / (0.5 * (dzc + dzf[ix + P]));
            const double bv = (double)Lb[k * 64 + lane];
            const double dk = (double)Ls[k * 64 + lane];
            const double bvm = (k > cb) ? (double)Lb[(k - 1) * 64 + lane] : bv;
            const double bvp = (k < cs) ? (double)Lb[(k + 1) * 64 + lane] : bv;
            const double dkm = (k > cb) ? (double)Ls[(k - 1) * 64 + lane] : 0.0;
            const double dkp = (k < cs) ? (double)Ls[(k + 1) * 64 + lane] : 0.0;
            const double rhs = -(cm1 * (bvm - bv) + cp1 * (bvp - bv));
            const double Ad = cm1 * dkm + (dzdt - cm1 - cp1) * dk + cp1 * dkp;
            const float rho = (float)((rhs - Ad) / dzdt);
            const float *cm = cob + k * C_N;
            const float gm = (k > cb) ? cm[C_GM] : 0.0f;
            e = (rho - gm * e) * cm[C_RD];
            sc[k * zstep + ln] = e;
          }
          float ec = 0.0f;
          for (int k = cs; k >= cb; k--) {
            const float ek = sc[k * zstep + ln];
            ec = (k == cs) ? ek : ek - cob[(k + 1) * C_N + C_UD] * ec;
            if (act)
              oc[k * zstep + ln] =
                  sround((double)Lb[k * 64 + lane] + (double)Ls[k * 64 + lane] + (double)ec,
                         mix32((hcol + (uint32_t)k * 0x9E3779B1U) ^ hl ^ 0x5bd1e995U));
          }
        }
      }
    }
    __syncthreads();
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
/* dynamic LDS of the fused kernel: Lb, Ls [nz][64], Lpre, Ltrs [64] */
static size_t col_lds(int nz) { return ((size_t)2 * nz * 64 + 128) * sizeof(float); }

/* Column order: tiles of tile_x x tile_y columns in row-major tile
 * order, columns row-major inside a tile; the wet columns of that order
 * are split into 8 contiguous bands, and position q of the persistent
 * loop maps to band q % 8, so the waves resident on one XCD walk one band
 * (workgroups are dealt to the 8 XCDs round robin). */
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
  const int per = (n + 7) / 8;
  b->nord = 8 * per;
  int *ord = (int *)malloc(sizeof(int) * (size_t)(b->nord > 0 ? b->nord : 1));
  for (int q = 0; q < b->nord; q++) {
    const int src = (q % 8) * per + q / 8;
    ord[q] = (src < n) ? lin[src] : -1;
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
  b->tile_x = 64;
  b->tile_y = 8;
  const size_t ncell = (size_t)d->plane * d->nz;
  bool ok = true;
  for (int q = 0; q < 2; q++)
    ok &= hipMalloc(&b->tr[q], ncell * d->ntr * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->scr, ncell * d->ntr * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->co, ncell * C_N * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->ci, (size_t)d->plane * sizeof(ColInfo)) == hipSuccess;
  ok &= hipMalloc(&b->order, ((size_t)d->plane + 8) * sizeof(int)) == hipSuccess;
  ok &= hipMalloc(&b->mintr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->maxtr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->stepctr, sizeof(int)) == hipSuccess;
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "batched form: device allocation failed");
    emstr_batch_destroy(b);
    return NULL;
  }
  hipMemset(b->co, 0, ncell * C_N * sizeof(float));
  hipMemset(b->stepctr, 0, sizeof(int));
  build_order(b, h_kb, d->nz);
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  /* persistent grid: as many workgroups as can be resident at once */
  int per_cu = 0;
  const size_t lds = col_lds(b->nz);
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void *)k_col<VANLEER>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_col<VANLEER>,
                                                   64 * CW, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  b->grid = ncu * per_cu;
  return b;
}

extern "C" void emstr_batch_destroy(struct emstr_batch *b) {
  if (!b) return;
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->graph) hipGraphDestroy(b->graph);
  void *v[] = {b->tr[0], b->tr[1], b->scr, b->co, b->ci, b->order, b->mintr, b->maxtr,
               b->stepctr};
  for (int q = 0; q < 9; q++)
    if (v[q]) hipFree(v[q]);
  free(b);
}

extern "C" void emstr_batch_set_prof(struct emstr_batch *b, struct emstr_prof *p) { b->prof = p; }

extern "C" int emstr_batch_tune(struct emstr_batch *b, const int *h_kb, int tile_x, int tile_y,
                                int grid) {
  if (tile_x > 0) b->tile_x = tile_x;
  if (tile_y > 0) b->tile_y = tile_y;
  if (grid > 0) b->grid = grid;
  build_order(b, h_kb, b->nz);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec), b->gexec = NULL;
  if (b->graph) (void)hipGraphDestroy(b->graph), b->graph = NULL;
  b->graph_ok = 0;
  return 0;
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
  prof_end(b->prof, pi, st);
  const float *in = b->tr[b->cur];
  float *out = b->tr[1 - b->cur];
  pi = prof_begin(b->prof, KC_FUSED, st);
  KArgs ka;
  ka.nz = b->nz;
  ka.ntr = b->ntr;
  ka.nord = b->nord;
  ka.nx = b->nx;
  ka.hmin = d->hmin;
  ka.dt = d->dt;
  ka.plane = d->plane;
  for (int g = 0; g < b->nt64; g++) {
    ka.g = g;
    k_col<VANLEER><<<b->grid, 64 * CW, col_lds(b->nz), st>>>(
        ka, in, out, b->co, b->ci, b->order, b->mintr, b->maxtr, d->layers, d->Kz, d->dz,
        b->scr, b->stepctr);
  }
  prof_end(b->prof, pi, st);
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
