/*
 * Precision form of the EMS tracer step on MI355X (gfx950).
 *
 * Tracers, fluxes, limiters and the tridiagonal sweeps run in FP32.
 * Three steps stay in FP64:
 *   (1) the per-column and domain tracer-mass budgets (TwoSum-compensated
 *       FP64 accumulation of tr*area*dz, as check_tmass,
 *       model/hd/tracers/tracers.c:6608-6631);
 *   (2) one FP64 residual pass of iterative refinement of the column solve
 *       (tracers.c:4067-4214) in every column whose largest diffusion
 *       number D = Kz*dt/(dz*dzface) exceeds IR_DMAX;
 *   (3) the layer thicknesses (dz from set_dz, tracers.c:5682-5849) and the
 *       surface-cell volume balance of surf_conc (tracers.c:3047-3207),
 *       evaluated in FP64 from FP32 tracers and fluxes.
 *
 * Per step the FP64 fields are reduced once (k_prep32) to FP32 per-face
 * volume transports q = flux*dt (m3 per step) and Courant numbers, and the
 * column matrix is built and factored in FP64 (k_mat32) and stored
 * row-normalised in FP32, so the tracer kernels read 4-byte coefficients.
 * The update is the reference's tr - ((dtracer/area + Fz[zp1] - Fz[c])/dz)
 * in anomaly form, tr - (sum_f s_f (F_f - tr) q_f + tr D) / V with D the
 * net volume divergence of the cell (FP64 in the prep): the face values
 * are computed from the stencil's differences against the cell value, so
 * their rounding is relative to those differences.  The column solve is
 * in increment form, A (x - b) = dzdt b - A b.  Each new FP32 value is
 * formed in FP64 and rounded stochastically (a deterministic hash of cell,
 * tracer, step and stage), so sub-ulp increments do not bias the state or
 * the mass.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "emstr_dev.h"
#include "emstr_internal.h"
#include "emstr_kernels.h"
#include "emstr_prec.h"
#include "emstr_prof.h"

#define VANLEER 0x000080
#define QUICKEST 0x000020
#define IR_DMAX 1.0

typedef long long ll;

struct emstr_prec {
  float *tr[2]; /* FP32 tracer state [ntr][n3]                          */
  int cur;
  /* per-cell FP32 coefficients [n3] (W, S and bottom face of the cell) */
  float *qW, *cW, *qS, *cS, *qB, *cB;
  float *rvol;       /* 1/(area*dz)                                     */
  float *div;        /* net volume divergence per step (FP64 -> FP32)    */
  float *mg;         /* upper off-diagonal of the normalised matrix      */
  float *scr;        /* [ntr][n3] refinement scratch                     */
  unsigned nstep;    /* steps taken (stochastic rounding keys)           */
  float *w5, *w9;    /* QUICKEST vertical weights                       */
  float *mb, *mr, *mu; /* row-normalised factored column matrix         */
  int *stiff;        /* [plane] column gets the FP64 refinement pass    */
  int trasc, ultimate;
  emstr_prof *prof;
};

/*-------------------------------------------------------------------*/
/* deterministic stochastic rounding                                 */
/*-------------------------------------------------------------------*/
__device__ __forceinline__ uint32_t pmix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
/* key of (cell, tracer, step, stage) */
__device__ __forceinline__ uint32_t pkey(ll idx, int n, unsigned step, unsigned stage) {
  return pmix((uint32_t)idx ^ pmix((uint32_t)n * 0x85EBCA77U ^ step * 0xC2B2AE3DU ^
                                   stage * 0x27D4EB2FU));
}
/* the FP64 value x rounded to one of its two FP32 neighbours, the upper
 * with probability equal to the remainder's fraction of an ulp (a uniform
 * 29-bit number added to the 29 bits FP32 drops, then cleared) */
__device__ __forceinline__ float psround(double x, uint32_t r) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  u += (unsigned long long)(r >> 3);
  u &= ~0x1FFFFFFFULL;
  return (float)__longlong_as_double((long long)u);
}

/*-------------------------------------------------------------------*/
/* coefficient prep (FP64 in, FP32 out), one thread per wet cell      */
/*-------------------------------------------------------------------*/
__global__ void k_prep32(emstr_dev d, emstr_prec p, double dtu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const double area = d.cellarea[col];
  const double dz = d.dz[idx];
  const double dzb = (k > kbc) ? d.dz[idx - d.plane] : 0.0;
  const double w9 = 0.5 * (dz + dzb);
  p.qW[idx] = (float)(d.u1flux3d[idx] * dtu);
  p.cW[idx] = (float)(d.u1[idx] * dtu / d.h1au1[col]);
  p.qS[idx] = (float)(d.u2flux3d[idx] * dtu);
  p.cS[idx] = (float)(d.u2[idx] * dtu / d.h2au2[col]);
  p.qB[idx] = (float)(d.w[idx] * dtu * area);
  p.cB[idx] = (float)(d.w[idx] * dtu / w9);
  p.rvol[idx] = (dz != 0.0) ? (float)(1.0 / (area * dz)) : 0.0f;
  {
    /* E, N and top faces are the W, S and bottom faces of the neighbours */
    const int ksc = d.ks[col];
    const double qW = d.u1flux3d[idx], qS = d.u2flux3d[idx];
    const double qE = (k >= d.kb[col + 1]) ? d.u1flux3d[idx + 1] : 0.0;
    const double qN = (k >= d.kb[col + d.nx]) ? d.u2flux3d[idx + d.nx] : 0.0;
    const double qB = d.w[idx] * area;
    const double qT = (ksc >= 0 && k < ksc) ? d.w[idx + d.plane] * area : 0.0;
    p.div[idx] = (float)((qE - qW + qN - qS + qT - qB) * dtu);
  }
  if (p.w5) {
    p.w5[idx] = (float)(dzb / (dzb + dz));
    p.w9[idx] = (float)w9;
  }
}

/* column matrix, FP64 build and factor, stored row-normalised in FP32:
 * the solve of A x = dzdt*r becomes s_k = mr_k*(r_k - mb_k*s_{k-1}),
 * x_k = s_k - mu_{k+1}*x_{k+1} (tracers.c:3841-3889, 4185-4205); mb and mg
 * are the off-diagonals over dzdt. */
__global__ void k_mat32(emstr_dev d, emstr_prec p) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.ncol) return;
  const ll col = d.cols[q];
  const ll P = d.plane;
  const int kbc = d.kb[col];
  const int cs = d.knew[col];
  const double dt = d.dt;
  int stiff = 0;
  p.stiff[col] = 0;
  if (cs - 1 < kbc || d.dz[(ll)(cs - 1) * P + col] == 0.0) return;
  int c = cs;
  while (c >= kbc && d.dz[(ll)c * P + col] != 0.0) c--;
  const int cb = c + 1;
  if (cb >= cs) return;
  double dzk = d.dz[(ll)cb * P + col];
  double dzp = d.dz[(ll)(cb + 1) * P + col];
  double Kp = d.Kz[(ll)(cb + 1) * P + col];
  double dzf_p = 0.5 * (dzk + dzp);
  double dzdt = dzk / dt;
  double Cm1, Cp1 = -Kp / dzf_p, C = dzdt - Cp1;
  double div = C;
  ll ic = (ll)cb * P + col;
  p.mb[ic] = 0.0f;
  p.mr[ic] = (float)(dzdt / div);
  p.mu[ic] = 0.0f;
  p.mg[ic] = (float)(Cp1 / dzdt);
  double Cp1prev = Cp1;
  for (int kk = cb + 1; kk <= cs; kk++) {
    double dzc = dzp, dzf_c = dzf_p, Kc = Kp;
    dzdt = dzc / dt;
    Cm1 = -Kc / dzf_c;
    if (Kc * dt / (dzc * dzf_c) > IR_DMAX) stiff = 1;
    if (kk < cs) {
      dzp = d.dz[(ll)(kk + 1) * P + col];
      Kp = d.Kz[(ll)(kk + 1) * P + col];
      dzf_p = 0.5 * (dzc + dzp);
      Cp1 = -Kp / dzf_p;
      C = dzdt - Cm1 - Cp1;
    } else {
      Cp1 = 0.0;
      C = dzdt - Cm1;
    }
    double ud = Cp1prev / div;
    div = C - Cm1 * ud;
    ic = (ll)kk * P + col;
    p.mb[ic] = (float)(Cm1 / dzdt);
    p.mr[ic] = (float)(dzdt / div);
    p.mu[ic] = (float)ud;
    p.mg[ic] = (float)(Cp1 / dzdt);
    Cp1prev = Cp1;
  }
  p.stiff[col] = stiff;
}

/*-------------------------------------------------------------------*/
/* FP32 face values                                                  */
/*-------------------------------------------------------------------*/
__device__ __forceinline__ bool wetat(const int *__restrict__ kb, ll col, int k) {
  return k >= kb[col];
}

struct Fc {
  float q, c;          /* volume per step, Courant number            */
  float d3, hm, h0, hp; /* QUICKEST geometry                          */
};

/* van Leer (tracers.c:2097-2118); the sign of q is the sign of the
 * velocity.  The stencil is mirrored for q < 0: (bm2, bm1, c) for q > 0,
 * (fp1, c, bm1) for q < 0, with 1 - cn and 1 + cn.  The mirrored product is
 * the same product of negated differences and its quotient the negated
 * quotient of the reference's c - 0.5 (1 + cn) df, so the value is the same
 * to the last bit without a branch on the sign */
__device__ __forceinline__ float vlf(float bm2, float bm1, float c, float fp1,
                                     float q, float cn) {
  const bool pos = q > 0.0f;
  const float a2 = pos ? bm2 : fp1, a1 = pos ? bm1 : c, a0 = pos ? c : bm1;
  const float omc = 1.0f + (pos ? -cn : cn);
  float df = (a1 - a2) * (a0 - a1);
  df = (df > 0.0f) ? 2.0f * df / (a0 - a2) : 0.0f;
  const float F = a1 + 0.5f * omc * df;
  return (pos || q < 0.0f) ? F : 0.0f;
}

/* QUICKEST (tracers.c:1862-1884) */
__device__ __forceinline__ float qkf(float bm2, float bm1, float c, float fp1,
                                     const Fc &f) {
  const float cx = f.c;
  float face = f.d3 * (c - bm1) + bm1;
  float grad = c - bm1;
  float curv = 0.0f;
  if (cx > 0.0f)
    curv = 2.0f / (f.hm + f.h0) * (f.hm * c + f.h0 * bm2) - 2.0f * bm1;
  else if (cx < 0.0f)
    curv = 2.0f / (f.h0 + f.hp) * (f.h0 * fp1 + f.hp * bm1) - 2.0f * c;
  return face - 0.5f * cx * grad - (1.0f - cx * cx) * curv * (1.0f / 6.0f);
}

/* ULTIMATE (tracers.c:3517-3553) */
__device__ __forceinline__ float ultf(float cr, float F, float bm2, float bm1,
                                      float c, float fp1) {
  if (cr == 0.0f) return F;
  float tc, tu, td;
  if (cr > 0.0f) { tc = bm1; tu = bm2; td = c; }
  else { tc = c; tu = fp1; td = bm1; }
  float diff = td - tu, nc = 0.0f, nu = 0.0f;
  if (diff != 0.0f) {
    nc = (tc - tu) / diff;
    nu = (F - tu) / diff;
  }
  const float ac = fabsf(cr);
  if (nc >= 0.0f && nc <= 1.0f) {
    if (nu > 1.0f) F = td;
    else if (nu < nc) F = tu + nc * (td - tu);
    else if (nu > nc / ac) F = tu + nc * (td - tu) / ac;
  } else
    F = tu + nc * (td - tu);
  return F;
}

template <int SCH, int ULT>
__device__ __forceinline__ float flux32(float bm2, float bm1, float c, float fp1,
                                        const Fc &f) {
  float F = (SCH == QUICKEST) ? qkf(bm2, bm1, c, fp1, f)
                              : vlf(bm2, bm1, c, fp1, f.q, f.c);
  if (ULT) F = ultf(f.c, F, bm2, bm1, c, fp1);
  return F * f.q;
}

/* Stencil offsets with the ghost semantics of emstr_plain.hip. */
struct St32 {
  ll xm1, xm2, xp1, xp2, ym1, ym2, yp1, yp2, zm1, zm2, zp1, zp2;
  bool wp1, sp1;
};

__device__ __forceinline__ St32 sten32(const emstr_dev &d, ll col, int k, int kbc) {
  St32 s;
  const ll nx = d.nx, P = d.plane;
  const bool wm1 = wetat(d.kb, col - 1, k), sm1 = wetat(d.kb, col - nx, k);
  s.wp1 = wetat(d.kb, col + 1, k);
  s.sp1 = wetat(d.kb, col + nx, k);
  s.xm1 = wm1 ? -1 : 0;
  s.xp1 = s.wp1 ? 1 : 0;
  s.xm2 = wm1 ? (wetat(d.kb, col - 2, k) ? -2 : -1) : 0;
  s.xp2 = s.wp1 ? (wetat(d.kb, col + 2, k) ? 2 : 1) : 0;
  s.ym1 = sm1 ? -nx : 0;
  s.yp1 = s.sp1 ? nx : 0;
  s.ym2 = sm1 ? (wetat(d.kb, col - 2 * nx, k) ? -2 * nx : -nx) : 0;
  s.yp2 = s.sp1 ? (wetat(d.kb, col + 2 * nx, k) ? 2 * nx : nx) : 0;
  s.zm1 = (ll)(max(k - 1, kbc) - k) * P;
  s.zm2 = (ll)(max(k - 2, kbc) - k) * P;
  s.zp1 = (ll)(min(k + 1, d.nz - 1) - k) * P;
  s.zp2 = (ll)(min(k + 2, d.nz - 1) - k) * P;
  return s;
}

struct Co32 {
  Fc W, E, S, N, B, T;
  float rv;
};

template <int SCH>
__device__ __forceinline__ void co32(const emstr_dev &d, const emstr_prec &p,
                                     ll col, ll idx, int k, const St32 &s,
                                     bool top, Co32 &c) {
  const ll nx = d.nx, P = d.plane;
  c.W.q = p.qW[idx];
  c.W.c = p.cW[idx];
  c.S.q = p.qS[idx];
  c.S.c = p.cS[idx];
  c.B.q = p.qB[idx];
  c.B.c = p.cB[idx];
  c.E.q = s.wp1 ? p.qW[idx + 1] : 0.0f;
  c.E.c = s.wp1 ? p.cW[idx + 1] : 0.0f;
  c.N.q = s.sp1 ? p.qS[idx + nx] : 0.0f;
  c.N.c = s.sp1 ? p.cS[idx + nx] : 0.0f;
  c.T.q = top ? p.qB[idx + P] : 0.0f;
  c.T.c = top ? p.cB[idx + P] : 0.0f;
  c.rv = p.rvol[idx];
  if (SCH == QUICKEST) {
    const double *h1c = d.h1acell, *h2c = d.h2acell;
    c.W.d3 = (float)(h1c[col - 1] / (h1c[col - 1] + h1c[col]));
    c.S.d3 = (float)(h2c[col - nx] / (h2c[col - nx] + h2c[col]));
    c.W.hm = (float)d.h1au1[col - 1];
    c.W.h0 = (float)d.h1au1[col];
    c.W.hp = (float)d.h1au1[col + 1];
    c.S.hm = (float)d.h2au2[col - nx];
    c.S.h0 = (float)d.h2au2[col];
    c.S.hp = (float)d.h2au2[col + nx];
    if (s.wp1) {
      c.E.d3 = (float)(h1c[col] / (h1c[col] + h1c[col + 1]));
      c.E.hm = c.W.h0;
      c.E.h0 = c.W.hp;
      c.E.hp = (float)d.h1au1[col + 2];
    }
    if (s.sp1) {
      c.N.d3 = (float)(h2c[col] / (h2c[col] + h2c[col + nx]));
      c.N.hm = c.S.h0;
      c.N.h0 = c.S.hp;
      c.N.hp = (float)d.h2au2[col + 2 * nx];
    }
    /* vertical: w9 at zm1 (0 for the sediment ghost), c, zp1 */
    const float w9c = p.w9[idx];
    const float w9m = (s.zm1 != 0) ? p.w9[idx + s.zm1] : 0.0f;
    const float w9p = p.w9[idx + s.zp1];
    c.B.d3 = p.w5[idx];
    c.B.hm = w9m;
    c.B.h0 = w9c;
    c.B.hp = w9p;
    if (top) {
      c.T.d3 = p.w5[idx + P];
      c.T.hm = w9c;
      c.T.h0 = p.w9[idx + P];
      c.T.hp = p.w9[idx + P + ((k + 2 < d.nz) ? P : 0)];
    }
  }
}

/* Fluxes of one cell as anomalies: the stencil values less the cell's
 * value c0 (the schemes are translation invariant, so each face value
 * comes out as F - c0); returns the horizontal sum of (F - c0) q over
 * E - W + N - S, and the bottom and top terms in FB, FT. */
template <int SCH, int ULT>
__device__ __forceinline__ float fl32(const float *__restrict__ t, ll idx,
                                      const St32 &s, const Co32 &c, bool top,
                                      float &FB, float &FT) {
  const float c0 = t[idx];
  const float xm1 = t[idx + s.xm1] - c0, xm2 = t[idx + s.xm2] - c0,
              xp1 = t[idx + s.xp1] - c0, xp2 = t[idx + s.xp2] - c0;
  const float ym1 = t[idx + s.ym1] - c0, ym2 = t[idx + s.ym2] - c0,
              yp1 = t[idx + s.yp1] - c0, yp2 = t[idx + s.yp2] - c0;
  const float zm1 = t[idx + s.zm1] - c0, zm2 = t[idx + s.zm2] - c0,
              zp1 = t[idx + s.zp1] - c0, zp2 = t[idx + s.zp2] - c0;
  float FW = flux32<SCH, ULT>(xm2, xm1, 0.0f, xp1, c.W);
  float FE = s.wp1 ? flux32<SCH, ULT>(xm1, 0.0f, xp1, xp2, c.E) : 0.0f;
  float FS = flux32<SCH, ULT>(ym2, ym1, 0.0f, yp1, c.S);
  float FN = s.sp1 ? flux32<SCH, ULT>(ym1, 0.0f, yp1, yp2, c.N) : 0.0f;
  FB = flux32<SCH, ULT>(zm2, zm1, 0.0f, zp1, c.B);
  FT = top ? flux32<SCH, ULT>(zm1, 0.0f, zp1, zp2, c.T) : 0.0f;
  return FE - FW + FN - FS;
}

/* the horizontal volume divergence of the cell, in FP64 */
__device__ __forceinline__ double hvol(const Co32 &c) {
  return ((double)c.E.q - (double)c.W.q) + ((double)c.N.q - (double)c.S.q);
}

/* horizontal flux divergence (flux form, FP64) of the cell at layer k */
template <int SCH, int ULT>
__device__ double hdiv32(const emstr_dev &d, const emstr_prec &p,
                         const float *__restrict__ t, ll col, int k, int kbc) {
  const ll idx = (ll)k * d.plane + col;
  St32 s = sten32(d, col, k, kbc);
  Co32 c;
  co32<SCH>(d, p, col, idx, k, s, false, c);
  float FB, FT;
  const float h = fl32<SCH, ULT>(t, idx, s, c, false, FB, FT);
  return (double)h + (double)t[idx] * hvol(c);
}

/* unclipped sub-surface update of the cell at layer k, in FP64 */
template <int SCH, int ULT>
__device__ double upd32(const emstr_dev &d, const emstr_prec &p,
                        const float *__restrict__ t, ll col, int k, int kbc) {
  const ll idx = (ll)k * d.plane + col;
  St32 s = sten32(d, col, k, kbc);
  Co32 c;
  co32<SCH>(d, p, col, idx, k, s, true, c);
  float FB, FT;
  const float c0 = t[idx];
  const float h = fl32<SCH, ULT>(t, idx, s, c, true, FB, FT);
  return (double)c0 - (double)((h + (FT - FB) + c0 * p.div[idx]) * c.rv);
}

/*-------------------------------------------------------------------*/
/* advection: one thread per wet cell, tracer loop inside             */
/*-------------------------------------------------------------------*/
template <int SCH, int ULT>
__global__ void __launch_bounds__(256)
k_adv32(emstr_dev d, emstr_prec p, const float *__restrict__ tin,
        float *__restrict__ tout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3, P = d.plane;
  const int ksc = d.ks[col];
  if (ksc < 0) {
    for (int n = 0; n < d.ntr; n++) tout[n * N3 + idx] = tin[n * N3 + idx];
    return;
  }
  if (k > ksc) return;
  const bool sub = k < ksc;
  const St32 s = sten32(d, col, k, kbc);
  Co32 c;
  co32<SCH>(d, p, col, idx, k, s, sub, c);
  if (sub) {
    for (int n = 0; n < d.ntr; n++) {
      const float *t = tin + n * N3;
      if (!d.in_tbdy[n] || !d.advect[n]) {
        float r = t[idx];
        if (d.in_tbdy[n] && d.mintr[n] <= 1e-5 && r < (float)d.mintr[n]) r = (float)d.mintr[n];
        tout[n * N3 + idx] = r;
        continue;
      }
      float FB, FT;
      const float c0 = t[idx];
      const float h = fl32<SCH, ULT>(t, idx, s, c, true, FB, FT);
      double r = (double)c0 - (double)((h + (FT - FB) + c0 * p.div[idx]) * c.rv);
      if (d.mintr[n] <= 1e-5 && r < d.mintr[n]) r = d.mintr[n];
      tout[n * N3 + idx] = psround(r, pkey(idx, n, p.nstep, 0));
    }
    return;
  }
  /* processed surface cell: surf_conc volume balance in FP64 */
  const double area = d.cellarea[col];
  const int kold = d.kold[col];
  const double subeta = d.subeta[col], osubeta = d.osubeta[col];
  const double botz = d.botz[col];
  const double *lay = d.layers;
  const int ktop = d.nz - 1;
  for (int n = 0; n < d.ntr; n++) {
    const float *t = tin + n * N3;
    float *o = tout + n * N3;
    if (!d.in_tbdy[n]) {
      for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = t[(ll)kk * P + col];
      continue;
    }
    double trs;
    if (!d.advect[n]) {
      trs = (double)t[idx];
    } else {
      float FB, FT;
      const float c0 = t[idx];
      const float h0 = fl32<SCH, ULT>(t, idx, s, c, false, FB, FT);
      int cc = k;
      double waterbot = (cc == kbc) ? botz : lay[cc];
      double dzs = subeta - waterbot;
      double svol = dzs * area;
      /* -fcbot, the transport through the bottom, and the horizontal
       * divergence, in flux form */
      double surftr = (double)FB + (double)c0 * (double)c.B.q;
      double dtrc = (double)h0 + (double)c0 * hvol(c);
      while (cc != kold && cc < ktop) {
        surftr += (double)t[(ll)cc * P + col] * area * (lay[cc + 1] - waterbot) - dtrc;
        cc++;
        waterbot = lay[cc];
        dtrc = hdiv32<SCH, ULT>(d, p, t, col, cc, kbc);
      }
      surftr += (double)t[(ll)cc * P + col] * area * (osubeta - waterbot) - dtrc;
      while (cc != ktop) {
        cc++;
        surftr -= hdiv32<SCH, ULT>(d, p, t, col, cc, kbc);
      }
      if (svol == 0.0 || (surftr < 0.0 && dzs <= d.hmin))
        surftr = 0.0;
      else
        surftr /= svol;
      trs = (botz > 0.0 && d.dz[idx] < d.hmin) ? (double)t[idx] : surftr;
    }
    /* thin surface layer (tracers.c:3250-3272) */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * P + col] : 0.0;
    if (k > kbc && dzco < d.hmin && (trs < d.mintr[n] || trs > d.maxtr[n]))
      trs = d.advect[n] ? upd32<SCH, ULT>(d, p, t, col, k - 1, kbc) : (double)t[idx - P];
    if (d.mintr[n] <= 1e-5 && trs < d.mintr[n]) trs = d.mintr[n];
    const float r = psround(trs, pkey(idx, n, p.nstep, 0));
    for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = r;
  }
}

/*-------------------------------------------------------------------*/
/* vertical diffusion: FP32 sweeps, one FP64 residual pass of         */
/* iterative refinement in stiff columns                              */
/*-------------------------------------------------------------------*/
__global__ void __launch_bounds__(256)
k_vdiff32(emstr_dev d, emstr_prec p, const float *__restrict__ tin,
          float *__restrict__ tout) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  const int tl = threadIdx.y;
  if (q >= d.ncol) return;
  const ll col = d.cols[q];
  const ll P = d.plane, N3 = d.n3;
  const int kbc = d.kb[col];
  const int cs = d.knew[col];
  const int top = d.nz - 1;
  const bool single = (cs - 1 < kbc) || d.dz[(ll)(cs - 1) * P + col] == 0.0;
  int cb = cs;
  if (!single) {
    int c = cs;
    while (c >= kbc && d.dz[(ll)c * P + col] != 0.0) c--;
    cb = c + 1;
  }
  const bool solve = !single && cb < cs;
  const int stiff = solve ? p.stiff[col] : 0;
  const double dt = d.dt;
  for (int n = tl; n < d.ntr; n += 4) {
    const float *b = tin + n * N3;
    float *x = tout + n * N3;
    if (!d.in_tdif[n] || !solve) {
      for (int kk = kbc; kk <= top; kk++) x[(ll)kk * P + col] = b[(ll)kk * P + col];
      continue;
    }
    for (int kk = kbc; kk < cb; kk++) x[(ll)kk * P + col] = b[(ll)kk * P + col];
    for (int kk = cs + 1; kk <= top; kk++) x[(ll)kk * P + col] = b[(ll)kk * P + col];
    /* increment form: A d = dzdt b - A b, the right-hand side row k
     * -(gm (b_{k-1} - b_k) + gp (b_{k+1} - b_k)) over dzdt from the
     * differences of b; forward sweep into x, back sweep d into x */
    float s = 0.0f, bm = 0.0f;
    for (int kk = cb; kk <= cs; kk++) {
      const ll ic = (ll)kk * P + col;
      const float bc = b[ic];
      const float dm = (kk > cb) ? bm - bc : 0.0f;
      const float dp = (kk < cs) ? b[ic + P] - bc : 0.0f;
      const float r = -(p.mb[ic] * dm + p.mg[ic] * dp);
      s = p.mr[ic] * (r - p.mb[ic] * s);
      x[ic] = s;
      bm = bc;
    }
    for (int kk = cs - 1; kk >= cb; kk--) {
      const ll ic = (ll)kk * P + col;
      s = x[ic] - p.mu[ic + P] * s;
      x[ic] = s;
    }
    if (!stiff) {
      for (int kk = cb; kk <= cs; kk++) {
        const ll ic = (ll)kk * P + col;
        x[ic] = psround((double)b[ic] + (double)x[ic], pkey(ic, n, p.nstep, 1));
      }
      continue;
    }
    /* one FP64 residual pass: R = dzdt b - A (b + d) (A from Kz and dz as
     * tracers.c:3841-3889); its forward sweep in the scratch column, the
     * back sweep adds the correction, x = b + d + e rounded once */
    float *e = p.scr + n * N3;
    float es = 0.0f;
    double dzm = 0.0;
    for (int kk = cb; kk <= cs; kk++) {
      const ll ic = (ll)kk * P + col;
      const double dzc = d.dz[ic];
      const double dzdt = dzc / dt;
      const double xc = (double)b[ic] + (double)x[ic];
      double cm1 = 0.0, cp1 = 0.0, xm = 0.0, xp = 0.0;
      if (kk > cb) {
        cm1 = -d.Kz[ic] / (0.5 * (dzm + dzc));
        xm = (double)b[ic - P] + (double)x[ic - P];
      }
      if (kk < cs) {
        cp1 = -d.Kz[ic + P] / (0.5 * (dzc + d.dz[ic + P]));
        xp = (double)b[ic + P] + (double)x[ic + P];
      }
      const double C = dzdt - cm1 - cp1;
      const double r = dzdt * (double)b[ic] - (cm1 * xm + C * xc + cp1 * xp);
      es = p.mr[ic] * ((float)(r / dzdt) - p.mb[ic] * es);
      e[ic] = es;
      dzm = dzc;
    }
    float ec = 0.0f;
    for (int kk = cs; kk >= cb; kk--) {
      const ll ic = (ll)kk * P + col;
      ec = (kk == cs) ? e[ic] : e[ic] - p.mu[ic + P] * ec;
      x[ic] = psround((double)b[ic] + (double)x[ic] + (double)ec, pkey(ic, n, p.nstep, 1));
    }
  }
}

/*-------------------------------------------------------------------*/
/* sparse FP64 <-> dense FP32, mass                                   */
/*-------------------------------------------------------------------*/
__global__ void k_put32(const int *__restrict__ wsp, const ll *__restrict__ wdn,
                        int nwet, const double *__restrict__ sp,
                        float *__restrict__ dn) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nwet) dn[wdn[q]] = (float)sp[wsp[q]];
}
__global__ void k_get32(const int *__restrict__ wsp, const ll *__restrict__ wdn,
                        int nwet, const float *__restrict__ dn,
                        double *__restrict__ sp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nwet) sp[wsp[q]] = (double)dn[wdn[q]];
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

/* per-column (fixed order) then per-block TwoSum tree, FP64 */
__global__ void k_mass32(emstr_dev d, const float *__restrict__ tr,
                         double *__restrict__ part) {
  const int n = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double ss[256], se[256];
  double s = 0.0, e = 0.0;
  if (q < d.ncol) {
    const ll col = d.cols[q];
    const int kbc = d.kb[col], ksc = d.ks[col];
    const double a = d.cellarea[col];
    const float *t = tr + n * d.n3;
    for (int k = ksc; k >= kbc; k--) {
      const ll idx = (ll)k * d.plane + col;
      double m = (double)t[idx] * a * d.dz[idx];
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
extern "C" struct emstr_prec *emstr_prec_create(const emstr_dev *d, void *stream) {
  (void)stream;
  emstr_prec *p = (emstr_prec *)calloc(1, sizeof(emstr_prec));
  const size_t n3 = (size_t)d->n3;
  const size_t nt = (size_t)(d->ntr > 0 ? d->ntr : 1);
  bool ok = true;
  for (int b = 0; b < 2; b++) ok &= hipMalloc(&p->tr[b], nt * n3 * sizeof(float)) == hipSuccess;
  float **f[] = {&p->qW, &p->cW, &p->qS, &p->cS, &p->qB, &p->cB, &p->rvol,
                 &p->mb, &p->mr, &p->mu, &p->div, &p->mg};
  for (int q = 0; q < 12; q++) ok &= hipMalloc(f[q], n3 * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&p->scr, nt * n3 * sizeof(float)) == hipSuccess;
  if (d->trasc == QUICKEST) {
    ok &= hipMalloc(&p->w5, n3 * sizeof(float)) == hipSuccess;
    ok &= hipMalloc(&p->w9, n3 * sizeof(float)) == hipSuccess;
  }
  ok &= hipMalloc(&p->stiff, (size_t)d->plane * sizeof(int)) == hipSuccess;
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "precision form: device allocation failed");
    emstr_prec_destroy(p);
    return NULL;
  }
  for (int b = 0; b < 2; b++) hipMemset(p->tr[b], 0, nt * n3 * sizeof(float));
  for (int q = 0; q < 12; q++) hipMemset(*f[q], 0, n3 * sizeof(float));
  hipMemset(p->scr, 0, nt * n3 * sizeof(float));
  if (p->w5) hipMemset(p->w5, 0, n3 * sizeof(float)), hipMemset(p->w9, 0, n3 * sizeof(float));
  hipMemset(p->stiff, 0, (size_t)d->plane * sizeof(int));
  p->cur = 0;
  return p;
}

extern "C" void emstr_prec_destroy(struct emstr_prec *p) {
  if (!p) return;
  void *v[] = {p->tr[0], p->tr[1], p->qW, p->cW, p->qS, p->cS, p->qB, p->cB,
               p->rvol, p->w5, p->w9, p->mb, p->mr, p->mu, p->stiff, p->div, p->mg, p->scr};
  for (int q = 0; q < 18; q++)
    if (v[q]) hipFree(v[q]);
  free(p);
}

extern "C" void emstr_prec_set_prof(struct emstr_prec *p, struct emstr_prof *prof) {
  p->prof = prof;
}

extern "C" int emstr_prec_fields(struct emstr_prec *p, const emstr_dev *d, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (d->trasc != VANLEER && d->trasc != QUICKEST)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "precision form: van Leer or QUICKEST only");
  if (d->trasc == QUICKEST && !p->w5)
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                           "precision form: scheme changed after creation");
  p->trasc = d->trasc;
  p->ultimate = d->ultimate;
  dim3 b(64, 4, 1);
  dim3 g((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
  int pi = prof_begin(p->prof, KC_PREP, st);
  k_prep32<<<g, b, 0, st>>>(*d, *p, d->dt);
  if (d->ncol) k_mat32<<<(d->ncol + 255) / 256, 256, 0, st>>>(*d, *p);
  prof_end(p->prof, pi, st);
  return hipGetLastError() == hipSuccess ? 0 : emstr_set_error(EMSTR_EDEVICE, -1, "prep");
}

extern "C" int emstr_prec_put_tracer(struct emstr_prec *p, const emstr_dev *d, int n,
                                     const double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_put32<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      d->wsp, d->wdn, d->nwet, sparse, p->tr[p->cur] + (ll)n * d->n3);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_prec_get_tracer(struct emstr_prec *p, const emstr_dev *d, int n,
                                     double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_get32<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      d->wsp, d->wdn, d->nwet, p->tr[p->cur] + (ll)n * d->n3, sparse);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_prec_advect(struct emstr_prec *p, const emstr_dev *d, int stab,
                                 void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (stab & 22) /* SUB_STEP | SUB_STEP_NOSURF | SUB_STEP_TRACER */
    return emstr_set_error(EMSTR_EUNSUPPORTED, -1, "precision form: no sub-stepping");
  if (emstr_launch_subeta(d, d->dt, d->dt, 1, st)) return -1;
  dim3 b(64, 4, 1);
  dim3 g((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
  const float *in = p->tr[p->cur];
  float *out = p->tr[1 - p->cur];
  int pi = prof_begin(p->prof, KC_ADV, st);
  if (p->trasc == QUICKEST) {
    if (p->ultimate) k_adv32<QUICKEST, 1><<<g, b, 0, st>>>(*d, *p, in, out);
    else k_adv32<QUICKEST, 0><<<g, b, 0, st>>>(*d, *p, in, out);
  } else {
    if (p->ultimate) k_adv32<VANLEER, 1><<<g, b, 0, st>>>(*d, *p, in, out);
    else k_adv32<VANLEER, 0><<<g, b, 0, st>>>(*d, *p, in, out);
  }
  prof_end(p->prof, pi, st);
  p->cur = 1 - p->cur;
  if (emstr_launch_osub(d, st)) return -1;
  return hipGetLastError() == hipSuccess ? 0 : emstr_set_error(EMSTR_EDEVICE, -1, "adv32");
}

extern "C" int emstr_prec_vdiff(struct emstr_prec *p, const emstr_dev *d, int *err,
                                void *stream) {
  (void)err;
  hipStream_t st = (hipStream_t)stream;
  if (d->ncol == 0) return 0;
  dim3 b(64, 4, 1);
  int pi = prof_begin(p->prof, KC_VDIFF, st);
  k_vdiff32<<<(d->ncol + 63) / 64, b, 0, st>>>(*d, *p, p->tr[p->cur], p->tr[1 - p->cur]);
  prof_end(p->prof, pi, st);
  p->cur = 1 - p->cur;
  p->nstep++;
  return hipGetLastError() == hipSuccess ? 0 : emstr_set_error(EMSTR_EDEVICE, -1, "vdiff32");
}

extern "C" int emstr_prec_step(struct emstr_prec *p, const emstr_dev *d, int stab,
                               int *err, void *stream) {
  int r = emstr_prec_advect(p, d, stab, stream);
  if (r) return r;
  return emstr_prec_vdiff(p, d, err, stream);
}

extern "C" int emstr_prec_mass(struct emstr_prec *p, const emstr_dev *d, double *part,
                               void *stream) {
  int nb = (d->ncol + 255) / 256;
  if (nb == 0) return 0;
  dim3 g(nb, d->ntr, 1);
  k_mass32<<<g, 256, 0, (hipStream_t)stream>>>(*d, p->tr[p->cur], part);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
