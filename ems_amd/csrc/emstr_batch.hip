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
 * face between bm1 and c; f is 1 - cn for q > 0 and 1 + cn for q < 0 (the
 * prep folds the Courant number and the limiter's factor 2 into it).  The
 * limiter's division uses the hardware reciprocal (about 1 ulp). */
__device__ __forceinline__ float vlf(float bm2, float bm1, float c, float fp1,
                                     float q, float f) {
  if (q > 0.0f) {
    const float df = (bm1 - bm2) * (c - bm1);
    const float t = df * __builtin_amdgcn_rcpf(c - bm2);
    return (df > 0.0f) ? bm1 + f * t : bm1;
  } else if (q < 0.0f) {
    const float df = (c - bm1) * (fp1 - c);
    const float t = df * __builtin_amdgcn_rcpf(fp1 - bm1);
    return (df > 0.0f) ? c - f * t : c;
  }
  return c; /* no transport: F - c = 0 */
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
  const float FW = vlf(h.w2, h.w1, v0, h.e1, qW, fW);
  const float FE = ew ? vlf(h.w1, v0, h.e1, h.e2, qE, fE) : ref;
  const float FS = vlf(h.s2, h.s1, v0, h.n1, qS, fS);
  const float FN = nw ? vlf(h.s1, v0, h.n1, h.n2, qN, fN) : ref;
  return (FE - ref) * qE - (FW - ref) * qW + (FN - ref) * qN - (FS - ref) * qS;
}

/*
 * One wave per water column, lanes = tracers.  The wave walks up from the
 * bed: the vertical part of the stencil is a sliding register window, the
 * horizontal neighbours of the next layer are loaded while the current
 * layer is computed, the cell coefficients arrive four layers per vector
 * register (lane l holds word l) a chunk ahead and are read with readlane,
 * and the flux through a cell's top face is the next cell's bottom-face
 * flux.  The advected value b_k of each layer is written to the output (it
 * stays in L2 until the final value replaces it) and the forward sweep of
 * the vertical solve runs one layer behind the walk, keeping s_k in the
 * wave's LDS slice.  The back substitution runs on the way down and writes
 * the final values.
 */
template <int SCH>
__global__ void __launch_bounds__(64 * WPB)
k_walk(const KArgs a, const float *__restrict__ tin, float *__restrict__ tout,
       const float *__restrict__ co, const ColInfo *__restrict__ cinfo,
       const int *__restrict__ order, const float *__restrict__ mintr,
       const float *__restrict__ maxtr, const double *__restrict__ layers,
       const double *__restrict__ Kz, const double *__restrict__ dzf,
       float *__restrict__ scr, const int *__restrict__ stepctr) {
  extern __shared__ float lds[];
  const int nz = a.nz, NT = a.ntr, top = nz - 1, nact = a.nact;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = min(lane, nact - 1); /* inactive lanes shadow the last tracer */
  /* LDS slice of this wave: s_k (then d_k) [nz][nact], ud_k [nz] */
  float *Ls = lds + wv * nz * (nact + 1);
  float *Lud = Ls + nz * nact;
  const int ln = a.g * 64 + li;
  const bool act = lane < nact;
  const uint32_t lb = (uint32_t)ln * 4u;
  const uint32_t rowb = (uint32_t)NT * 4u; /* bytes per layer of a column */
  const ll colb = (ll)nz * NT * 4;         /* bytes per column */
  const ll jb = a.nx * colb;               /* bytes per row of columns */
  const float mn = mintr[ln];
  const float mx = maxtr[ln];
  const bool clip = mn <= 1e-5f;
  const uint32_t hl = mix32((uint32_t)ln * 0x85EBCA77U ^ (uint32_t)stepctr[0] * 0xC2B2AE3DU);
  const int nslot = gridDim.x * WPB;

  for (int q = blockIdx.x * WPB + wv; q < a.nord; q += nslot) {
    const int colq = order[q];
    if (colq < 0) continue;
    const ll col = colq;
    const ColInfo ci = cinfo[col];
    const int kb = ci.kb, ks = ci.ks;
    Nbr nb;
    nb.c = (const char *)tin + col * colb;
    nb.w1 = nb.c - colb;
    nb.w2 = nb.c - 2 * colb;
    nb.e1 = nb.c + colb;
    nb.e2 = nb.c + 2 * colb;
    nb.s1 = nb.c - jb;
    nb.s2 = nb.c - 2 * jb;
    nb.n1 = nb.c + jb;
    nb.n2 = nb.c + 2 * jb;
    nb.kw1 = ci.kw1;
    nb.kw2 = ci.kw2;
    nb.ke1 = ci.ke1;
    nb.ke2 = ci.ke2;
    nb.ks1 = ci.ks1;
    nb.ks2 = ci.ks2;
    nb.kn1 = ci.kn1;
    nb.kn2 = ci.kn2;
    char *oc = (char *)tout + col * colb;
    const char *tc = nb.c;
    const float *cob = co + col * nz * C_N;
    const uint32_t hcol = (uint32_t)(col * nz);
    if (!(ci.flags & F_PROC)) {
      for (int k = kb; k <= top; k++)
        if (act) stb(oc, lb + k * rowb, ldb(tc, lb + k * rowb));
      continue;
    }
    const bool solve = !(ci.flags & F_SINGLE);
    const int cb = solve ? ci.cb : nz, cs = solve ? ci.cs : -1;

    /* coefficient chunks: layers kc..kc+3 and kc+4..kc+7 */
    int kc = kb;
    int ch0 = __float_as_int(cob[kc * C_N + lane]);
    int ch1 = __float_as_int(cob[(kc + 4) * C_N + lane]);
    /* vertical window t[k-2 .. k+2], clamped to [kb, top] */
    float zm2, zm1, v0, zp1, zp2;
    v0 = ldb(tc, lb + kb * rowb);
    zm1 = zm2 = v0;
    zp1 = ldb(tc, lb + min(kb + 1, top) * rowb);
    zp2 = ldb(tc, lb + min(kb + 2, top) * rowb);
    HVals h;
    load_h(h, nb, kb, lb + kb * rowb);
    /* bottom-face flux of the bed cell */
    float FB = vlf(zm2, zm1, v0, zp1, rdl(ch0, C_QB), rdl(ch0, C_CB));
    /* forward-sweep state: b_{k-2}, b_{k-1}, s_{k-2}, coefficients of k-1 */
    float bm2 = 0.0f, bm1 = 0.0f, sprev = 0.0f;
    float gm1 = 0.0f, gp1 = 0.0f, rd1 = 1.0f;
    float pre = 0.0f; /* unclipped update of the layer below the surface */
    float trs = 0.0f;

    for (int k = kb; k <= ks; k++) {
      const int r = 16 * (k - kc); /* word of layer k in ch0 */
      /* prefetch the next layer's horizontal neighbours and t[k+3] */
      HVals hn;
      float zp3 = zp2;
      if (k < ks) {
        load_h(hn, nb, k + 1, lb + (k + 1) * rowb);
        zp3 = ldb(tc, lb + min(k + 3, top) * rowb);
      }
      const float qW = rdl(ch0, r + C_QW), fW = rdl(ch0, r + C_CW);
      const float qE = rdl(ch0, r + C_QE), fE = rdl(ch0, r + C_CE);
      const float qS = rdl(ch0, r + C_QS), fS = rdl(ch0, r + C_CS);
      const float qN = rdl(ch0, r + C_QN), fN = rdl(ch0, r + C_CN);
      const float qB = rdl(ch0, r + C_QB);
      const float gmk = rdl(ch0, r + C_GM), gpk = rdl(ch0, r + C_GP);
      const float rdk = rdl(ch0, r + C_RD), udk = rdl(ch0, r + C_UD);
      const float fba = (FB - v0) * qB;
      const float hh = hflux(h, k >= ci.ke1, k >= ci.kn1, v0, v0, qW, fW, qE, fE, qS, fS,
                             qN, fN);
      const uint32_t key = srkey(hcol + (uint32_t)k, hl);
      float bk;
      if (k < ks) {
        /* sub-surface update (tracers.c:912-925) in anomaly form; the top
         * face is the bottom face of the cell above */
        const int nsrc = (r < 48) ? ch0 : ch1;
        const int rn = (r + 16) & 63;
        const float qT = rdl(nsrc, rn + C_QB);
        const float FT = vlf(zm1, v0, zp1, zp2, qT, rdl(nsrc, rn + C_CB));
        const float ht = hh + (FT - v0) * qT - fba;
        const double an =
            (double)v0 - (double)((ht + v0 * rdl(ch0, r + C_D)) * rdl(ch0, r + C_RV));
        bk = sround(an, key);
        pre = bk;
        if (clip && bk < mn) bk = mn;
        FB = FT;
      } else {
        /* processed surface cell: surf_conc (tracers.c:3047-3207) as an
         * FP64 volume balance, v0 * Vc + anomaly terms, Vc in FP64 */
        double m = (double)fba - (double)hh;
        if (ci.flags & F_MOVE) {
          int cc = ks;
          double wb = (ks == kb) ? ci.botz : layers[ks];
          while (cc != ci.kold && cc < top) {
            const double vol = ci.area * (layers[cc + 1] - wb);
            if (cc > ks) m += ((double)ldb(tc, lb + cc * rowb) - (double)v0) * vol;
            cc++;
            wb = layers[cc];
          }
          if (cc > ks)
            m += ((double)ldb(tc, lb + cc * rowb) - (double)v0) *
                 (ci.area * (ci.osub - ((cc == kb) ? ci.botz : layers[cc])));
          for (int ka = ks + 1; ka <= top; ka++) {
            HVals ha;
            load_h(ha, nb, ka, lb + ka * rowb);
            const float *ca = cob + ka * C_N;
            m -= (double)hflux(ha, ka >= ci.ke1, ka >= ci.kn1, ldb(tc, lb + ka * rowb), v0,
                               ca[C_QW], ca[C_CW], ca[C_QE], ca[C_CE], ca[C_QS], ca[C_CS],
                               ca[C_QN], ca[C_CN]);
          }
        }
        double surftr = (double)v0 * ci.vc + m;
        if (ci.svol == 0.0 || (surftr < 0.0 && ci.dzs <= a.hmin))
          surftr = 0.0;
        else
          surftr /= ci.svol;
        /* tracers.c:953-954 */
        trs = (ci.flags & F_KEEPOLD) ? v0 : sround(surftr, key);
        /* tr_bounds (tracers.c:3217-3324): thin old surface layer, clip */
        if ((ci.flags & F_THIN) && (trs < mn || trs > mx)) trs = pre;
        if (clip && trs < mn) trs = mn;
        bk = trs;
      }
      if (act) stb(oc, lb + k * rowb, bk);
      Lud[k] = udk;
      /* forward sweep for layer k-1, now that b_k is known:
       * s = (r - gm s_prev) / div, r = -(gm (b_{k-2} - b_{k-1}) +
       * gp (b_k - b_{k-1})) */
      {
        const int kf = k - 1;
        if (kf >= cb && kf <= cs) {
          const float gm = (kf > cb) ? gm1 : 0.0f;
          const float rr = -(gm * (bm2 - bm1) + gp1 * (bk - bm1));
          sprev = (rr - gm * sprev) * rd1;
          Ls[kf * nact + li] = sprev;
        }
      }
      gm1 = gmk;
      gp1 = gpk;
      rd1 = rdk;
      bm2 = bm1;
      bm1 = bk;
      /* slide */
      zm2 = zm1;
      zm1 = v0;
      v0 = zp1;
      zp1 = zp2;
      zp2 = zp3;
      if (k < ks) h = hn;
      if (r == 48) {
        /* next chunk */
        kc += 4;
        ch0 = ch1;
        ch1 = __float_as_int(cob[(kc + 4) * C_N + lane]);
      }
    }
    /* the surface value fills every layer above (tracers.c:3306-3322) and
     * the forward sweep runs on to the new surface */
    for (int k = ks + 1; k <= top; k++) {
      if (act) stb(oc, lb + k * rowb, trs);
      const float *cf = cob + k * C_N;
      const float gmk = cf[C_GM], gpk = cf[C_GP], rdk = cf[C_RD];
      Lud[k] = cf[C_UD];
      const int kf = k - 1;
      if (kf >= cb && kf <= cs) {
        const float gm = (kf > cb) ? gm1 : 0.0f;
        const float rr = -(gm * (bm2 - bm1) + gp1 * (trs - bm1));
        sprev = (rr - gm * sprev) * rd1;
        Ls[kf * nact + li] = sprev;
      }
      gm1 = gmk;
      gp1 = gpk;
      rd1 = rdk;
      bm2 = bm1;
      bm1 = trs;
    }
    if (cs == top) {
      /* last row: b_{k+1} does not enter (gp = 0) */
      const float gm = (cs > cb) ? gm1 : 0.0f;
      const float rr = -(gm * (bm2 - bm1));
      sprev = (rr - gm * sprev) * rd1;
      Ls[cs * nact + li] = sprev;
    }
    if (!solve) continue;
    /* back substitution: d_k = s_k - ud_{k+1} d_{k+1}, x_k = b_k + d_k */
    const bool stiff = (ci.flags & F_STIFF) != 0;
    float dn = 0.0f;
    float bnext = ldb(oc, lb + cs * rowb);
    for (int k = cs; k >= cb; k--) {
      const float bv = bnext;
      if (k > cb) bnext = ldb(oc, lb + (k - 1) * rowb);
      const float sk = Ls[k * nact + li];
      dn = (k == cs) ? sk : sk - Lud[k + 1] * dn;
      if (stiff) {
        Ls[k * nact + li] = dn;
      } else if (act) {
        stb(oc, lb + k * rowb,
            sround((double)bv + (double)dn, srkey(hcol + (uint32_t)k, hl ^ 0x5bd1e995U)));
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
        const double bv = (double)ldb(oc, lb + k * rowb);
        const double dk = (double)Ls[k * nact + li];
        const double bvm = (k > cb) ? (double)ldb(oc, lb + (k - 1) * rowb) : bv;
        const double bvp = (k < cs) ? (double)ldb(oc, lb + (k + 1) * rowb) : bv;
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
      /* b_k is still needed below k while the residual is formed: write
       * the results after the pass */
      for (int k = cb; k <= cs; k++) {
        const float bv = ldb(oc, lb + k * rowb);
        const float ek = ldb(sc, lb + k * rowb);
        if (act)
          stb(oc, lb + k * rowb,
              sround((double)bv + (double)Ls[k * nact + li] + (double)ek,
                     srkey(hcol + (uint32_t)k, hl ^ 0x5bd1e995U)));
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
/* dynamic LDS of the fused kernel: one [nz][nact + 1] slice per wave */
static size_t walk_lds(int nz, int nact) {
  return (size_t)WPB * nz * (nact + 1) * sizeof(float);
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
  b->tile_x = 64;
  b->tile_y = 8;
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
  build_order(b, h_kb, d->nz);
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  /* persistent grid: as many workgroups as can be resident at once */
  int per_cu = 0;
  const size_t lds = walk_lds(b->nz, b->ntr < 64 ? b->ntr : 64);
  if (lds > 48 * 1024)
    (void)hipFuncSetAttribute((const void *)k_walk<VANLEER>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_walk<VANLEER>,
                                                   64 * WPB, lds) != hipSuccess ||
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
  ka.pad = 0;
  for (int g = 0; g < b->nt64; g++) {
    ka.g = g;
    ka.nact = b->ntr - 64 * g < 64 ? b->ntr - 64 * g : 64;
    k_walk<VANLEER><<<b->grid, 64 * WPB, walk_lds(b->nz, ka.nact), st>>>(
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
