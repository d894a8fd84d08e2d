/*
 * Plain form of the EMS tracer step on MI355X (gfx950): FP64 throughout,
 * the reference's schemes, limiters and operation order, so results agree
 * with the reference CPU path to the last bit.  Compiled with
 * -ffp-contract=off (no FMA contraction).
 *
 * Kernels (dense layout described in emstr_dev.h):
 *   k_adv      FFSL: one thread per wet cell, loop over tracers inside:
 *              fluxes from the trajectory integrals, flux divergence and
 *              sub-surface update (tracers.c:773-925); the processed
 *              surface cell also does surf_conc (tracers.c:3047-3207) and
 *              tr_bounds (tracers.c:3217-3324) for its column.
 *   k_face / k_upd  the other schemes, face-once: k_face computes each
 *              cell's west, south and bottom face values and fluxes per
 *              tracer (van Leer / QUICKEST / QUICKEST_CO / ORDER1 /
 *              ORDER2 / ORDER4 / ORDER2_UW, optional ULTIMATE and
 *              horizontal diffusion) into buffers; k_upd does the
 *              divergence, update, surf_conc and tr_bounds from them.
 *   k_vdiff    64 columns x 16 tracer lanes per block: the column's
 *              tridiagonal matrix is built and factored once
 *              (tracers.c:3836-3889, 4185-4195) into LDS and reused by
 *              every tracer's forward/back substitution
 *              (tracers.c:4135-4205), in place.
 *   k_setdz    per column: new surface, cells to process, dz
 *              (tracers.c:5682-5849).
 *   k_scatter / k_gather  sparse (reference numbering) <-> dense.
 *   k_mass     per-column and domain tracer mass, FP64 with error-free
 *              (TwoSum) accumulation (tracers.c:6608-6631).
 *   k_dtm      minimum sub-step (tracers.c:590-655).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emstr_dev.h"
#include "emstr_kernels.h"

#define ORDER1 0x000002
#define ORDER2 0x000004
#define ORDER4 0x000010
#define QUICKEST 0x000020
#define QUICKEST_CO 0x000040
#define VANLEER 0x000080
#define ORDER2_UW 0x002000
#define FFSL 0x080000

typedef long long ll;

/* the reference's max()/min() macros (ternaries, not fmax/fmin) */
__device__ __forceinline__ double fmx(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double fmn(double a, double b) { return (a < b) ? a : b; }

/*-------------------------------------------------------------------*/
/* Stencil geometry of one wet cell                                  */
/*-------------------------------------------------------------------*/
struct Sten {
  /* element offsets from the cell for the points reached by the maps
   * xm1, xm1(xm1), xp1, xp1(xp1) etc.; a point beyond land stops at the
   * last wet cell (ghost semantics).  gm: bit set if the point is a
   * lateral ghost (read from the step-start buffer). */
  ll xm1, xm2, xp1, xp2, ym1, ym2, yp1, yp2, zm1, zm2, zp1, zp2;
  unsigned gm;
  bool wetE, wetN;
};

enum { G_XM1 = 1, G_XM2 = 2, G_XP1 = 4, G_XP2 = 8, G_YM1 = 16, G_YM2 = 32,
       G_YP1 = 64, G_YP2 = 128 };

__device__ __forceinline__ bool wetat(const int *__restrict__ kb, ll col, int k) {
  return k >= kb[col];
}

/* Offsets along one horizontal direction with stride s (1 or nx). */
__device__ __forceinline__ void hdir(const int *__restrict__ kb, ll col, int k,
                                     ll s, ll &m1, ll &m2, ll &p1, ll &p2,
                                     unsigned &gm, unsigned gm1, unsigned gm2,
                                     unsigned gp1, unsigned gp2, bool &wetp) {
  bool wm1 = wetat(kb, col - s, k);
  bool wp1 = wetat(kb, col + s, k);
  wetp = wp1;
  if (wm1) {
    m1 = -s;
    if (wetat(kb, col - 2 * s, k)) m2 = -2 * s;
    else { m2 = -s; gm |= gm2; }
  } else {
    m1 = 0; m2 = 0; gm |= gm1 | gm2;
  }
  if (wp1) {
    p1 = s;
    if (wetat(kb, col + 2 * s, k)) p2 = 2 * s;
    else { p2 = s; gm |= gp2; }
  } else {
    p1 = 0; p2 = 0; gm |= gp1 | gp2;
  }
}

__device__ __forceinline__ Sten make_sten(const emstr_dev &d, ll col, int k, int kbc) {
  Sten s;
  s.gm = 0;
  hdir(d.kb, col, k, 1, s.xm1, s.xm2, s.xp1, s.xp2, s.gm, G_XM1, G_XM2, G_XP1,
       G_XP2, s.wetE);
  hdir(d.kb, col, k, d.nx, s.ym1, s.ym2, s.yp1, s.yp2, s.gm, G_YM1, G_YM2,
       G_YP1, G_YP2, s.wetN);
  const ll P = d.plane;
  /* vertical: the bed cell maps down to the sediment ghost (which holds
   * the bed value and maps to itself); the top layer maps to itself */
  s.zm1 = (ll)(max(k - 1, kbc) - k) * P;
  s.zm2 = (ll)(max(k - 2, kbc) - k) * P;
  s.zp1 = (ll)(min(k + 1, d.nz - 1) - k) * P;
  s.zp2 = (ll)(min(k + 2, d.nz - 1) - k) * P;
  return s;
}

/* Tracer values on the 13-point stencil of one cell for one tracer. */
struct Vals {
  double c, xm1, xm2, xp1, xp2, ym1, ym2, yp1, yp2, zm1, zm2, zp1, zp2;
};

__device__ __forceinline__ Vals load_vals(const double *__restrict__ t,
                                          const double *__restrict__ ts,
                                          ll idx, const Sten &s) {
  Vals v;
  v.c = t[idx];
  v.xm1 = ((s.gm & G_XM1) ? ts : t)[idx + s.xm1];
  v.xm2 = ((s.gm & G_XM2) ? ts : t)[idx + s.xm2];
  v.xp1 = ((s.gm & G_XP1) ? ts : t)[idx + s.xp1];
  v.xp2 = ((s.gm & G_XP2) ? ts : t)[idx + s.xp2];
  v.ym1 = ((s.gm & G_YM1) ? ts : t)[idx + s.ym1];
  v.ym2 = ((s.gm & G_YM2) ? ts : t)[idx + s.ym2];
  v.yp1 = ((s.gm & G_YP1) ? ts : t)[idx + s.yp1];
  v.yp2 = ((s.gm & G_YP2) ? ts : t)[idx + s.yp2];
  v.zm1 = t[idx + s.zm1];
  v.zm2 = t[idx + s.zm2];
  v.zp1 = t[idx + s.zp1];
  v.zp2 = t[idx + s.zp2];
  return v;
}

/*-------------------------------------------------------------------*/
/* Face values along one direction.  bm2, bm1: two cells upstream of   */
/* the face (for a positive velocity), c: the cell the face belongs    */
/* to, fp1: the cell beyond c.                                         */
/*-------------------------------------------------------------------*/
struct QkGeom { /* QUICKEST metric weights of one face */
  double d3;      /* interpolation weight (d3/d4/w5)                 */
  double hm, h0, hp; /* spacings at bm1, c, fp1 (h1au1 / w9)          */
};

/* tracers.c:2097-2118, with the stencil mirrored for a negative velocity:
 * (bm2, bm1, c) for vel > 0, (fp1, c, bm1) for vel < 0.  For vel < 0 the
 * reference forms df = (c - bm1) (fp1 - c), 2 df / (fp1 - bm1) and
 * c - 0.5 (1 + cn) df; the mirrored product is the same product of negated
 * differences and its quotient the negated quotient, so the face value is
 * the same to the last bit, and no branch on the sign is taken (in a wave
 * with both signs the quotient was computed twice) */
__device__ __forceinline__ double vl_face(double bm2, double bm1, double c,
                                          double fp1, double vel, double cn,
                                          bool &set) {
  set = vel > 0.0 || vel < 0.0;
  const bool pos = vel > 0.0;
  const double a2 = pos ? bm2 : fp1, a1 = pos ? bm1 : c, a0 = pos ? c : bm1;
  const double omc = 1.0 + (pos ? -cn : cn);
  double df = (a1 - a2) * (a0 - a1);
  df = (df > 0.0) ? 2.0 * df / (a0 - a2) : 0.0;
  const double F = a1 + 0.5 * omc * df;
  return set ? F : 0.0;
}

/* tracers.c:1862-1884 (non-uniform QUICKEST) */
__device__ __forceinline__ double qk_face(double bm2, double bm1, double c,
                                          double fp1, double cx,
                                          const QkGeom &q) {
  double db = 2.0 / (q.hm + q.h0);
  double df = 2.0 / (q.h0 + q.hp);
  double face = q.d3 * (c - bm1) + bm1;
  double grad = c - bm1;
  double curv = 0.0;
  if (cx > 0.0)
    curv = db * (q.hm * c + q.h0 * bm2) - 2.0 * bm1;
  else if (cx < 0.0)
    curv = df * (q.h0 * fp1 + q.hp * bm1) - 2.0 * c;
  return face - 0.5 * cx * grad - (1.0 - cx * cx) * curv / 6.0;
}

/* tracers.c:1988-1996 (uniform QUICKEST) */
__device__ __forceinline__ double qku_face(double bm2, double bm1, double c,
                                           double fp1, double cx) {
  double face = 0.5 * (c + bm1);
  double grad = c - bm1;
  double curv = 0.0;
  if (cx > 0.0)
    curv = c - 2.0 * bm1 + bm2;
  else if (cx < 0.0)
    curv = fp1 - 2.0 * c + bm1;
  return face - 0.5 * cx * grad - (1.0 - cx * cx) * curv / 6.0;
}

/* tracers.c:3517-3553 ULTIMATE on one face value F */
__device__ __forceinline__ double ultimate(double cr, double F, double bm2,
                                           double bm1, double c, double fp1) {
  double tc, tu, td, diff, nc, nu;
  if (cr == 0.0) return F;
  if (cr > 0.0) {
    tc = bm1;
    tu = bm2;
    td = c;
  } else {
    tc = c;
    tu = fp1;
    td = bm1;
  }
  diff = td - tu;
  nc = nu = 0.0;
  if (diff) {
    nc = (tc - tu) / diff;
    nu = (F - tu) / diff;
  }
  if (nc >= 0.0 && nc <= 1.0) {
    if (nu > 1.0)
      F = td;
    else if (nu < nc)
      F = tu + nc * (td - tu);
    else if (nu > nc / fabs(cr))
      F = tu + nc * (td - tu) / fabs(cr);
  } else
    F = tu + nc * (td - tu);
  return F;
}

/* Face inputs shared by all tracers. */
struct Face {
  double vel;  /* u1 / u2 / w at the face (sign for van Leer)     */
  double cn;   /* Courant number w6 / w7 / w8                     */
  double flux; /* u1flux3d / u2flux3d / w*dtu multiplier          */
  QkGeom q;
  double mgs;  /* ORDER2 weight (d3/d4/w5)                        */
};

/*
 * Advective flux through a face for one tracer (advect(),
 * tracers.c:1509-1574): face value by the scheme, ULTIMATE, then times
 * the volume flux.  ORDER1 returns the upwind flux directly
 * (tracers.c:1598-1616).  'open' is false for faces the reference never
 * visits (index not a wet cell), whose flux stays zero.
 */
template <int SCH>
__device__ __forceinline__ double face_flux(double bm2, double bm1, double c,
                                            double fp1, const Face &f,
                                            int ultimate_on) {
  double F = 0.0;
  if (SCH == ORDER1) {
    /* f.flux holds the signed volume flux; f.vel unused */
    double a = 0.5 * (f.flux + fabs(f.flux));
    double b = 0.5 * (f.flux - fabs(f.flux));
    return a * bm1 + b * c;
  }
  if (SCH == VANLEER) {
    bool set;
    F = vl_face(bm2, bm1, c, fp1, f.vel, f.cn, set);
  } else if (SCH == QUICKEST) {
    F = qk_face(bm2, bm1, c, fp1, f.cn, f.q);
  } else if (SCH == QUICKEST_CO) {
    F = qku_face(bm2, bm1, c, fp1, f.cn);
  } else if (SCH == ORDER2) {
    F = f.mgs * (c - bm1) + bm1;
  } else if (SCH == ORDER4) {
    F = (7.0 * (bm1 + c) - (fp1 + bm2)) / 12.0;
  } else if (SCH == ORDER2_UW) {
    /* tracers.c:3024-3030 */
    if (f.vel > 0.0)
      F = 0.5 * (3.0 - f.cn) * bm1 - 0.5 * (1.0 - f.cn) * bm2;
    else if (f.vel < 0.0)
      F = 0.5 * (3.0 - f.cn) * c - 0.5 * (1.0 - f.cn) * fp1;
  }
  if (ultimate_on) F = ultimate(f.cn, F, bm2, bm1, c, fp1);
  return F * f.flux;
}

/* ORDER1 vertical flux (tracers.c:1613-1616): already times dtu */
__device__ __forceinline__ double o1_fz(double w, double dtu, double bm1,
                                        double c) {
  return dtu * (0.5 * (w + fabs(w)) * bm1 + 0.5 * (w - fabs(w)) * c);
}

/*-------------------------------------------------------------------*/
/* Per-cell shared coefficients                                       */
/*-------------------------------------------------------------------*/
struct CellCo {
  Face W, E, S, N, B, T;
  double area, dz, dtu;
  /* horizontal diffusion coefficients (csx*u1kh etc.) and spacings    */
  double kW, kE, kS, kN, hW, hE, hS, hN;
};

/* dz of the cell k-1 below, 0 for the sediment ghost under the bed */
__device__ __forceinline__ double dz_below(const emstr_dev &d, ll idx, int k, int kbc) {
  return (k > kbc) ? d.dz[idx - d.plane] : 0.0;
}

template <int SCH>
__device__ void cell_coef(const emstr_dev &d, ll col, ll idx, int k, int kbc,
                          const Sten &s, double dtu, bool need_top,
                          int courf, CellCo &cc) {
  const ll nx = d.nx, P = d.plane;
  cc.area = d.cellarea[col];
  cc.dz = d.dz[idx];
  cc.dtu = dtu;
  /* x faces: W face of c, E face = W face of xp1 (only if wet) */
  cc.W.vel = d.u1[idx];
  cc.W.flux = d.u1flux3d[idx];
  if (courf) cc.W.cn = d.u1[idx] * dtu / d.h1au1[col];
  if (s.wetE) {
    cc.E.vel = d.u1[idx + 1];
    cc.E.flux = d.u1flux3d[idx + 1];
    if (courf) cc.E.cn = d.u1[idx + 1] * dtu / d.h1au1[col + 1];
  }
  cc.S.vel = d.u2[idx];
  cc.S.flux = d.u2flux3d[idx];
  if (courf) cc.S.cn = d.u2[idx] * dtu / d.h2au2[col];
  if (s.wetN) {
    cc.N.vel = d.u2[idx + nx];
    cc.N.flux = d.u2flux3d[idx + nx];
    if (courf) cc.N.cn = d.u2[idx + nx] * dtu / d.h2au2[col + nx];
  }
  /* vertical: bottom face of c, top face of c = bottom face of zp1 */
  double dzc = cc.dz;
  double dzb = dz_below(d, idx, k, kbc);
  double w9c = 0.5 * (dzc + dzb);
  cc.B.vel = d.w[idx];
  cc.B.flux = d.w[idx] * dtu;
  if (courf) cc.B.cn = d.w[idx] * dtu / (w9c * 1.0);
  double w9t = 0.0, dzt = 0.0;
  if (need_top) {
    dzt = d.dz[idx + P];
    w9t = 0.5 * (dzt + dzc);
    cc.T.vel = d.w[idx + P];
    cc.T.flux = d.w[idx + P] * dtu;
    if (courf) cc.T.cn = d.w[idx + P] * dtu / (w9t * 1.0);
  }
  if (SCH == QUICKEST || SCH == ORDER2) {
    /* d3/d4 (tracers.c:566-576) use the surface-layer 2D maps, whose
     * neighbours are the columns at i-1 / j-1 (wet or not) */
    const double *h1c = d.h1acell, *h2c = d.h2acell;
    double d3c = h1c[col - 1] / (h1c[col - 1] + h1c[col]);
    double d4c = h2c[col - nx] / (h2c[col - nx] + h2c[col]);
    cc.W.q.d3 = d3c;
    cc.W.mgs = d3c;
    cc.S.q.d3 = d4c;
    cc.S.mgs = d4c;
    if (s.wetE) {
      double d3e = h1c[col] / (h1c[col] + h1c[col + 1]);
      cc.E.q.d3 = d3e;
      cc.E.mgs = d3e;
    }
    if (s.wetN) {
      double d4n = h2c[col] / (h2c[col] + h2c[col + nx]);
      cc.N.q.d3 = d4n;
      cc.N.mgs = d4n;
    }
    cc.W.q.hm = d.h1au1[col - 1];
    cc.W.q.h0 = d.h1au1[col];
    cc.W.q.hp = d.h1au1[col + 1];
    cc.S.q.hm = d.h2au2[col - nx];
    cc.S.q.h0 = d.h2au2[col];
    cc.S.q.hp = d.h2au2[col + nx];
    if (s.wetE) {
      cc.E.q.hm = d.h1au1[col];
      cc.E.q.h0 = d.h1au1[col + 1];
      cc.E.q.hp = d.h1au1[col + 2];
    }
    if (s.wetN) {
      cc.N.q.hm = d.h2au2[col];
      cc.N.q.h0 = d.h2au2[col + nx];
      cc.N.q.hp = d.h2au2[col + 2 * nx];
    }
    /* vertical: w5 = dz(zm1)/(dz(zm1)+dz(c)), w9 at zm1, c, zp1
     * (tracers.c:562-565); w9 of the sediment ghost is 0 */
    double w5c = dzb / (dzb + dzc);
    cc.B.q.d3 = w5c;
    cc.B.mgs = w5c;
    double w9m = (k > kbc) ? 0.5 * (dzb + dz_below(d, idx - P, k - 1, kbc)) : 0.0;
    double w9p;
    if (k + 1 < d.nz) {
      double dzp = d.dz[idx + P];
      w9p = 0.5 * (dzp + dzc);
    } else {
      w9p = w9c;
    }
    cc.B.q.hm = w9m;
    cc.B.q.h0 = w9c;
    cc.B.q.hp = w9p;
    if (need_top) {
      double w5t = dzc / (dzc + dzt);
      cc.T.q.d3 = w5t;
      cc.T.mgs = w5t;
      cc.T.q.hm = w9c;
      cc.T.q.h0 = w9t;
      double w9pp;
      if (k + 2 < d.nz) w9pp = 0.5 * (d.dz[idx + 2 * P] + dzt);
      else w9pp = w9t;
      cc.T.q.hp = w9pp;
    }
  }
}

/*
 * Horizontal and vertical fluxes of one cell for one tracer.
 * Returns dtracer (tracers.c:888-892); Fz at the bottom and (if
 * need_top) the top face.
 */
template <int SCH>
__device__ __forceinline__ double cell_fluxes(const emstr_dev &d,
                                              const CellCo &cc, const Vals &v,
                                              const Sten &s, int n, ll idx,
                                              bool need_top, int adv, int dif,
                                              double &FzB, double &FzT) {
  const int ult = d.ultimate;
  double FxW = 0.0, FxE = 0.0, FyS = 0.0, FyN = 0.0;
  FzB = 0.0;
  FzT = 0.0;
  if constexpr (SCH == FFSL) {
    if (adv) {
      /* fluxes of the trajectory integrals (k_ffsl_flux) */
      FxW = d.ffx[idx];
      if (s.wetE) FxE = d.ffx[idx + 1];
      FyS = d.ffy[idx];
      if (s.wetN) FyN = d.ffy[idx + d.nx];
      FzB = d.ffz[idx];
      if (need_top) FzT = d.ffz[idx + d.plane];
    }
  } else if (adv) {
    /* W face: stencil of c;  E face: stencil of xp1 (bm1 = c) */
    FxW = face_flux<SCH>(v.xm2, v.xm1, v.c, v.xp1, cc.W, ult);
    if (s.wetE) FxE = face_flux<SCH>(v.xm1, v.c, v.xp1, v.xp2, cc.E, ult);
    FyS = face_flux<SCH>(v.ym2, v.ym1, v.c, v.yp1, cc.S, ult);
    if (s.wetN) FyN = face_flux<SCH>(v.ym1, v.c, v.yp1, v.yp2, cc.N, ult);
    if (SCH == ORDER1) {
      FzB = o1_fz(cc.B.vel, cc.dtu, v.zm1, v.c);
      if (need_top) FzT = o1_fz(cc.T.vel, cc.dtu, v.c, v.zp1);
    } else {
      FzB = face_flux<SCH>(v.zm2, v.zm1, v.c, v.zp1, cc.B, ult);
      if (need_top) FzT = face_flux<SCH>(v.zm1, v.c, v.zp1, v.zp2, cc.T, ult);
    }
  }
  if (dif) {
    /* tracers.c:3656-3664: Fx -= csx*u1kh*(tr[c]-tr[xm1])/h1au1 */
    FxW -= (cc.kW * (v.c - v.xm1) / cc.hW);
    if (s.wetE) FxE -= (cc.kE * (v.xp1 - v.c) / cc.hE);
    FyS -= (cc.kS * (v.c - v.ym1) / cc.hS);
    if (s.wetN) FyN -= (cc.kN * (v.yp1 - v.c) / cc.hN);
  }
  /* tracers.c:902-903: FFSL drops the horizontal divergence at nstep 0 */
  if (SCH == FFSL && d.nstep0) return 0.0;
  return (FxE - FxW + FyN - FyS) * cc.dtu;
}

__device__ __forceinline__ void diff_coef(const emstr_dev &d, ll col, ll idx,
                                          const Sten &s, CellCo &cc) {
  const ll nx = d.nx;
  /* csx = dzu1*h2au1*mdx (mdx = 1), then *u1kh */
  cc.kW = d.dzu1[idx] * d.h2au1[col] * 1.0 * d.u1kh[idx];
  cc.hW = d.h1au1[col];
  cc.kS = d.dzu2[idx] * d.h1au2[col] * 1.0 * d.u2kh[idx];
  cc.hS = d.h2au2[col];
  if (s.wetE) {
    cc.kE = d.dzu1[idx + 1] * d.h2au1[col + 1] * 1.0 * d.u1kh[idx + 1];
    cc.hE = d.h1au1[col + 1];
  }
  if (s.wetN) {
    cc.kN = d.dzu2[idx + nx] * d.h1au2[col + nx] * 1.0 * d.u2kh[idx + nx];
    cc.hN = d.h2au2[col + nx];
  }
}

/* Sub-surface update of one cell for one tracer (tracers.c:912-925,
 * z-levels: Ds = Hn1 = 1). */
__device__ __forceinline__ double subsurface(double trc, double dtr, double area,
                                             double FzT, double FzB, double dz) {
  return trc - ((dtr / area + (FzT - FzB)) / dz);
}

/* Recompute the full update of the cell at layer k of column col (used
 * for the layer below a thin surface in tr_bounds). */
template <int SCH>
__device__ double cell_update(const emstr_dev &d, const double *t,
                              const double *ts, ll col, int k, int kbc,
                              double dtu, int n, int courf, int adv, int dif) {
  ll idx = (ll)k * d.plane + col;
  Sten s = make_sten(d, col, k, kbc);
  CellCo cc;
  cell_coef<SCH>(d, col, idx, k, kbc, s, dtu, true, courf, cc);
  if (dif) diff_coef(d, col, idx, s, cc);
  Vals v = load_vals(t, ts, idx, s);
  double FzB, FzT;
  double dtr = cell_fluxes<SCH>(d, cc, v, s, n, idx, true, adv, dif, FzB, FzT);
  return subsurface(v.c, dtr, cc.area, FzT, FzB, cc.dz);
}

/* Horizontal divergence only (for cells above the processed surface). */
template <int SCH>
__device__ double cell_dtracer(const emstr_dev &d, const double *t,
                               const double *ts, ll col, int k, int kbc,
                               double dtu, int n, int courf, int adv, int dif) {
  ll idx = (ll)k * d.plane + col;
  Sten s = make_sten(d, col, k, kbc);
  CellCo cc;
  cell_coef<SCH>(d, col, idx, k, kbc, s, dtu, false, courf, cc);
  if (dif) diff_coef(d, col, idx, s, cc);
  Vals v = load_vals(t, ts, idx, s);
  double FzB, FzT;
  return cell_fluxes<SCH>(d, cc, v, s, n, idx, false, adv, dif, FzB, FzT);
}

/*-------------------------------------------------------------------*/
/* k_adv                                                             */
/*-------------------------------------------------------------------*/
template <int SCH>
__global__ void __launch_bounds__(256)
k_adv(emstr_dev d, const double *__restrict__ tin, double *__restrict__ tout,
      const double *__restrict__ tstart, double dtu, double trem, int anydif, int n0,
      int n1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  const int ksc = d.ks[col];
  const int courf = (d.ultimate || d.trasc >= QUICKEST) ? 1 : 0;

  if (ksc < 0) { /* column not processed: unchanged */
    for (int n = n0; n < n1; n++) tout[n * N3 + idx] = tin[n * N3 + idx];
    return;
  }
  if (k > ksc) return; /* above the processed surface: written below */

  const Sten s = make_sten(d, col, k, kbc);
  const bool sub = (k < ksc);
  CellCo cc;
  cell_coef<SCH>(d, col, idx, k, kbc, s, dtu, sub, courf, cc);
  if (anydif) diff_coef(d, col, idx, s, cc);

  if (sub) {
    for (int n = n0; n < n1; n++) {
      const double *t = tin + n * N3;
      if (!d.in_tbdy[n]) {
        tout[n * N3 + idx] = t[idx];
        continue;
      }
      const int adv = d.advect[n], dif = d.diffuse[n];
      Vals v = load_vals(t, tstart + n * N3, idx, s);
      double FzB, FzT;
      double dtr = cell_fluxes<SCH>(d, cc, v, s, n, idx, true, adv, dif, FzB, FzT);
      double r = subsurface(v.c, dtr, cc.area, FzT, FzB, cc.dz);
      /* tr_bounds clip (tracers.c:3279-3285) for s1 cells */
      if (d.mintr[n] <= 1e-5 && r < d.mintr[n]) r = d.mintr[n];
      tout[n * N3 + idx] = r;
    }
    return;
  }

  /* ---- processed surface cell of the column (k == ksc) ---- */
  const double area = cc.area;
  const int kold = d.kold[col];
  const double subeta = d.subeta[col];
  const double osubeta = d.osubeta[col];
  const double botz = d.botz[col];
  const double *lay = d.layers;
  const int ktop = d.nz - 1;
  for (int n = n0; n < n1; n++) {
    const double *t = tin + n * N3;
    const double *ts = tstart + n * N3;
    double *o = tout + n * N3;
    if (!d.in_tbdy[n]) {
      for (int kk = k; kk <= ktop; kk++) o[(ll)kk * d.plane + col] = t[(ll)kk * d.plane + col];
      continue;
    }
    const int adv = d.advect[n], dif = d.diffuse[n];
    double trs;
    if (!adv) {
      trs = t[idx];
    } else {
      Vals v = load_vals(t, ts, idx, s);
      double FzB, FzT;
      double dtr0 = cell_fluxes<SCH>(d, cc, v, s, n, idx, false, adv, dif, FzB, FzT);
      /* surf_conc (tracers.c:3047-3207) */
      double fcbot = -FzB * area;
      int c = k;
      double waterbot = (c == kbc) ? botz : lay[c];
      double dzs = subeta - waterbot;
      double svol = dzs * area;
      double surftr = -fcbot;
      double watertop;
      double dtrc = dtr0;
      while (c != kold && c < ktop) {
        watertop = lay[c + 1];
        surftr += (t[(ll)c * d.plane + col] * area * (watertop - waterbot) - dtrc);
        c = c + 1;
        waterbot = lay[c];
        dtrc = cell_dtracer<SCH>(d, t, ts, col, c, kbc, dtu, n, courf, adv, dif);
      }
      watertop = osubeta;
      waterbot = (c == kbc) ? botz : lay[c];
      surftr += (t[(ll)c * d.plane + col] * area * (watertop - waterbot) - dtrc);
      if (c != ktop) {
        do {
          c = c + 1;
          surftr -= (cell_dtracer<SCH>(d, t, ts, col, c, kbc, dtu, n, courf, adv, dif));
        } while (c != ktop);
      }
      if (svol == 0.0 || (surftr < 0.0 && dzs <= d.hmin))
        surftr = 0.0;
      else
        surftr /= svol;
      /* tracers.c:953-954 */
      trs = (botz > 0.0 && cc.dz < d.hmin) ? t[idx] : surftr;
    }
    /* tr_bounds, thin surface layer (tracers.c:3250-3272): the old
     * surface thickness dz[co] */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * d.plane + col] : 0.0;
    if (k > kbc && dzco < d.hmin) {
      if (trs < d.mintr[n] || trs > d.maxtr[n]) {
        trs = cell_update<SCH>(d, t, ts, col, k - 1, kbc, dtu, n, courf, adv, dif);
      }
    }
    /* clip (tracers.c:3279-3285) */
    if (d.mintr[n] <= 1e-5 && trs < d.mintr[n]) trs = d.mintr[n];
    /* the surface value fills every layer above (tracers.c:3306-3322) */
    for (int kk = k; kk <= ktop; kk++) o[(ll)kk * d.plane + col] = trs;
  }
}

/*-------------------------------------------------------------------*/
/* Face-once split of k_adv (schemes other than FFSL): k_face computes  */
/* each cell's own west, south and bottom face fluxes once per tracer   */
/* (advection times the transport, then horizontal diffusion), and      */
/* k_upd updates the cells from them.  The east face of a cell is the   */
/* west face of its east neighbour: the stencil walk from either side   */
/* reaches the same cells and ghosts, and the face coefficients are the */
/* same expressions, so the result equals k_adv's bit for bit.          */
/*-------------------------------------------------------------------*/
template <int SCH>
__global__ void __launch_bounds__(256)
k_face(emstr_dev d, const double *__restrict__ tin, const double *__restrict__ tstart,
       double dtu, int anydif, double *__restrict__ fw, double *__restrict__ fs,
       double *__restrict__ fb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  const int ksc = d.ks[col];
  const int courf = (d.ultimate || d.trasc >= QUICKEST) ? 1 : 0;
  const int ult = d.ultimate;
  /* vertical faces only of the cells to process */
  const bool vert = (ksc >= 0 && k <= ksc);
  const Sten s = make_sten(d, col, k, kbc);
  CellCo cc;
  cell_coef<SCH>(d, col, idx, k, kbc, s, dtu, false, courf, cc);
  if (anydif) diff_coef(d, col, idx, s, cc);
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n]) continue;
    const int adv = d.advect[n], dif = anydif ? d.diffuse[n] : 0;
    const Vals v = load_vals(tin + n * N3, tstart + n * N3, idx, s);
    double FxW = 0.0, FyS = 0.0, FzB = 0.0;
    if (adv) {
      FxW = face_flux<SCH>(v.xm2, v.xm1, v.c, v.xp1, cc.W, ult);
      FyS = face_flux<SCH>(v.ym2, v.ym1, v.c, v.yp1, cc.S, ult);
      if (vert) {
        if (SCH == ORDER1)
          FzB = o1_fz(cc.B.vel, cc.dtu, v.zm1, v.c);
        else
          FzB = face_flux<SCH>(v.zm2, v.zm1, v.c, v.zp1, cc.B, ult);
      }
    }
    if (dif) {
      FxW -= (cc.kW * (v.c - v.xm1) / cc.hW);
      FyS -= (cc.kS * (v.c - v.ym1) / cc.hS);
    }
    fw[n * N3 + idx] = FxW;
    fs[n * N3 + idx] = FyS;
    if (vert) fb[n * N3 + idx] = FzB;
  }
}

/* dtracer of the cell at layer k of column col from the face fluxes */
__device__ __forceinline__ double face_dtr(const emstr_dev &d, const double *__restrict__ fw,
                                           const double *__restrict__ fs, ll col, int k,
                                           double dtu) {
  const ll idx = (ll)k * d.plane + col, nx = d.nx;
  const double FxW = fw[idx], FyS = fs[idx];
  const double FxE = wetat(d.kb, col + 1, k) ? fw[idx + 1] : 0.0;
  const double FyN = wetat(d.kb, col + nx, k) ? fs[idx + nx] : 0.0;
  return (FxE - FxW + FyN - FyS) * dtu;
}

template <int SCH>
__global__ void __launch_bounds__(256)
k_upd(emstr_dev d, const double *__restrict__ tin, double *__restrict__ tout, double dtu,
      const double *__restrict__ fw, const double *__restrict__ fs,
      const double *__restrict__ fb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll P = d.plane;
  const ll idx = (ll)k * P + col;
  const ll N3 = d.n3;
  const int ksc = d.ks[col];
  if (ksc < 0) { /* column not processed: unchanged */
    for (int n = 0; n < d.ntr; n++) tout[n * N3 + idx] = tin[n * N3 + idx];
    return;
  }
  if (k > ksc) return; /* above the processed surface: written below */
  if (d.obcm && d.obcm[col]) {
    /* an open-boundary column: the update and surf_conc run over the
     * other columns (tracers.c:811-822, 912, 945: vc1 / vcs1), and so
     * does tr_bounds' thin-layer reset (tracers.c:3250, vcs1); its clip
     * runs over s1[1..vc1], which holds the boundary surface cells
     * (vcs1 < cc <= vcs) but not the cells below them (tracers.c:3277),
     * and its copy above the surface covers every processed column
     * (tracers.c:3306-3322) */
    for (int n = 0; n < d.ntr; n++) {
      const double *t = tin + n * N3;
      double *o = tout + n * N3;
      if (k < ksc) {
        o[idx] = t[idx];
      } else if (!d.in_tbdy[n]) {
        for (int kk = k; kk < d.nz; kk++) o[(ll)kk * P + col] = t[(ll)kk * P + col];
      } else {
        double v = t[idx];
        if (d.mintr[n] <= 1e-5 && v < d.mintr[n]) v = d.mintr[n];
        for (int kk = k; kk < d.nz; kk++) o[(ll)kk * P + col] = v;
      }
    }
    return;
  }
  const double area = d.cellarea[col], dzc = d.dz[idx];

  if (k < ksc) {
    /* face_dtr with the east and north wet tests taken once: a dry
     * neighbour's face reads the cell's own (in range) and is dropped */
    const bool wE = wetat(d.kb, col + 1, k), wN = wetat(d.kb, col + d.nx, k);
    const ll iE = idx + (wE ? 1 : 0), iN = idx + (wN ? (ll)d.nx : 0);
    for (int n = 0; n < d.ntr; n++) {
      const double *t = tin + n * N3;
      if (!d.in_tbdy[n]) {
        tout[n * N3 + idx] = t[idx];
        continue;
      }
      const double *fwn = fw + n * N3, *fsn = fs + n * N3;
      const double FxW = fwn[idx], FyS = fsn[idx], vE = fwn[iE], vN = fsn[iN];
      const double FxE = wE ? vE : 0.0, FyN = wN ? vN : 0.0;
      const double dtr = (FxE - FxW + FyN - FyS) * dtu;
      double FzB = 0.0, FzT = 0.0;
      if (d.advect[n]) {
        FzB = fb[n * N3 + idx];
        FzT = fb[n * N3 + idx + P];
      }
      double r = subsurface(t[idx], dtr, area, FzT, FzB, dzc);
      /* tr_bounds clip (tracers.c:3279-3285) for s1 cells */
      if (d.mintr[n] <= 1e-5 && r < d.mintr[n]) r = d.mintr[n];
      tout[n * N3 + idx] = r;
    }
    return;
  }

  /* ---- processed surface cell of the column (k == ksc) ---- */
  const int kold = d.kold[col];
  const double subeta = d.subeta[col];
  const double osubeta = d.osubeta[col];
  const double botz = d.botz[col];
  const double *lay = d.layers;
  const int ktop = d.nz - 1;
  for (int n = 0; n < d.ntr; n++) {
    const double *t = tin + n * N3;
    const double *fwn = fw + n * N3, *fsn = fs + n * N3, *fbn = fb + n * N3;
    double *o = tout + n * N3;
    if (!d.in_tbdy[n]) {
      for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = t[(ll)kk * P + col];
      continue;
    }
    const int adv = d.advect[n];
    double trs;
    if (!adv) {
      trs = t[idx];
    } else {
      const double dtr0 = face_dtr(d, fwn, fsn, col, k, dtu);
      /* surf_conc (tracers.c:3047-3207) */
      double fcbot = -fbn[idx] * area;
      int c = k;
      double waterbot = (c == kbc) ? botz : lay[c];
      double dzs = subeta - waterbot;
      double svol = dzs * area;
      double surftr = -fcbot;
      double watertop;
      double dtrc = dtr0;
      while (c != kold && c < ktop) {
        watertop = lay[c + 1];
        surftr += (t[(ll)c * P + col] * area * (watertop - waterbot) - dtrc);
        c = c + 1;
        waterbot = lay[c];
        dtrc = face_dtr(d, fwn, fsn, col, c, dtu);
      }
      watertop = osubeta;
      waterbot = (c == kbc) ? botz : lay[c];
      surftr += (t[(ll)c * P + col] * area * (watertop - waterbot) - dtrc);
      if (c != ktop) {
        do {
          c = c + 1;
          surftr -= (face_dtr(d, fwn, fsn, col, c, dtu));
        } while (c != ktop);
      }
      if (svol == 0.0 || (surftr < 0.0 && dzs <= d.hmin))
        surftr = 0.0;
      else
        surftr /= svol;
      /* tracers.c:953-954 */
      trs = (botz > 0.0 && dzc < d.hmin) ? t[idx] : surftr;
    }
    /* tr_bounds, thin surface layer (tracers.c:3250-3272) */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * P + col] : 0.0;
    if (k > kbc && dzco < d.hmin) {
      if (trs < d.mintr[n] || trs > d.maxtr[n]) {
        /* the full update of the layer below */
        const ll ib = idx - P;
        const double dtrb = face_dtr(d, fwn, fsn, col, k - 1, dtu);
        double FzB = 0.0, FzT = 0.0;
        if (adv) {
          FzB = fbn[ib];
          FzT = fbn[idx];
        }
        trs = subsurface(t[ib], dtrb, area, FzT, FzB, d.dz[ib]);
      }
    }
    /* clip (tracers.c:3279-3285) */
    if (d.mintr[n] <= 1e-5 && trs < d.mintr[n]) trs = d.mintr[n];
    /* the surface value fills every layer above (tracers.c:3306-3322) */
    for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = trs;
  }
}

static dim3 grid3(const emstr_dev *d, dim3 b);

/*-------------------------------------------------------------------*/
/* FFSL, structured grid (tracers.c:2132-2910)                        */
/*-------------------------------------------------------------------*/
/* A trajectory position is a wet cell or a ghost: p = idx * 8 + g with
 * idx a dense cell index and g = P_WET (the wet cell idx), P_GW..P_GN
 * (the lateral ghost west/east/south/north of wet cell idx) or P_SED
 * (the sediment ghost under bed cell idx).  The maps follow the grid's
 * ghost convention: a ghost maps back to its cell in the inward
 * direction and to itself in every other direction.  A lateral ghost
 * holds the step-start value of its cell (tracers.c:5574-5593), the
 * sediment ghost the bed value (tracers.c:827-841); a ghost's metrics
 * are those of the column it lies in. */
enum { P_WET = 0, P_GW = 1, P_GE = 2, P_GS = 3, P_GN = 4, P_SED = 5 };
#define FF_MAXWALK 4096 /* bound on the cells a trajectory crosses */

__device__ __forceinline__ ll ppack(ll idx, int g) { return idx * 8 + g; }
__device__ __forceinline__ ll pidx(ll p) { return p >> 3; }
__device__ __forceinline__ int pgh(ll p) { return (int)(p & 7); }

__device__ __forceinline__ ll p_xm1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  const int g = pgh(p);
  if (g == P_WET)
    return wetat(d.kb, idx % d.plane - 1, (int)(idx / d.plane)) ? ppack(idx - 1, P_WET)
                                                               : ppack(idx, P_GW);
  return (g == P_GE) ? ppack(idx, P_WET) : p;
}
__device__ __forceinline__ ll p_xp1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  const int g = pgh(p);
  if (g == P_WET)
    return wetat(d.kb, idx % d.plane + 1, (int)(idx / d.plane)) ? ppack(idx + 1, P_WET)
                                                               : ppack(idx, P_GE);
  return (g == P_GW) ? ppack(idx, P_WET) : p;
}
__device__ __forceinline__ ll p_ym1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  const int g = pgh(p);
  if (g == P_WET)
    return wetat(d.kb, idx % d.plane - d.nx, (int)(idx / d.plane)) ? ppack(idx - d.nx, P_WET)
                                                                  : ppack(idx, P_GS);
  return (g == P_GN) ? ppack(idx, P_WET) : p;
}
__device__ __forceinline__ ll p_yp1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  const int g = pgh(p);
  if (g == P_WET)
    return wetat(d.kb, idx % d.plane + d.nx, (int)(idx / d.plane)) ? ppack(idx + d.nx, P_WET)
                                                                  : ppack(idx, P_GN);
  return (g == P_GS) ? ppack(idx, P_WET) : p;
}
__device__ __forceinline__ ll p_zm1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  if (pgh(p) != P_WET) return p;
  return ((int)(idx / d.plane) > d.kb[idx % d.plane]) ? ppack(idx - d.plane, P_WET)
                                                      : ppack(idx, P_SED);
}
__device__ __forceinline__ ll p_zp1(const emstr_dev &d, ll p) {
  const ll idx = pidx(p);
  const int g = pgh(p);
  if (g == P_SED) return ppack(idx, P_WET);
  if (g != P_WET) return p;
  return ((int)(idx / d.plane) < d.nz - 1) ? ppack(idx + d.plane, P_WET) : p;
}
/* the column of a position (its 2D metrics) */
__device__ __forceinline__ ll p_col(const emstr_dev &d, ll p) {
  const ll col = pidx(p) % d.plane;
  switch (pgh(p)) {
    case P_GW: return col - 1;
    case P_GE: return col + 1;
    case P_GS: return col - d.nx;
    case P_GN: return col + d.nx;
    default: return col;
  }
}
/* a (the tracer or a modified copy) at p; a ghost holds the tracer */
__device__ __forceinline__ double p_val(const double *a, const double *t, const double *ts,
                                        ll p) {
  const int g = pgh(p);
  const ll idx = pidx(p);
  const double *b = (g == P_WET) ? a : ((g == P_SED) ? t : ts);
  return b[idx];
}
/* a 3-D cell field (dz, w9) at p: 0 at ghosts (never set there) */
__device__ __forceinline__ double p_f3(const double *f, ll p) {
  return (pgh(p) == P_WET) ? f[pidx(p)] : 0.0;
}

/* tracers.c:2771-2808 horizontal trajectory from head p (pcl, pcrf) */
template <int DIR>
__device__ __forceinline__ void ff_hor(const emstr_dev &d, double vel, double dt,
                                       const double *h, ll &cl, double &crf, int *err) {
  if (vel == 0.0) return;
  const double ltraj = fabs(vel) * dt;
  double lrem = ltraj;
  ll cp = cl;
  double dx = h[p_col(d, cp)];
  int it = 0;
  while (dx < lrem) {
    lrem -= dx;
    if (DIR == 0) cp = (vel < 0.0) ? p_xp1(d, cp) : p_xm1(d, cp);
    else cp = (vel < 0.0) ? p_yp1(d, cp) : p_ym1(d, cp);
    const double hh = h[p_col(d, cp)];
    if (hh > 0) dx = hh;
    if (++it > FF_MAXWALK) {
      atomicExch(err, 1);
      break;
    }
  }
  crf = fmx(fmn(lrem / dx, 1.0), 0.0);
  cl = cp;
}

/* tracers.c:2820-2859 vertical trajectory (h: w9 or dz) */
__device__ __forceinline__ void ff_vert(const emstr_dev &d, ll c, double vel, double dt,
                                        const double *h, ll &cl, double &crf, int *err) {
  if (vel == 0.0) return;
  const double ltraj = fabs(vel) * dt;
  double lrem = ltraj;
  ll cp = cl, cpb = cl;
  const double hcp = p_f3(h, cp);
  double dz = fabs((hcp > 0) ? hcp : p_f3(h, c));
  int it = 0;
  while (dz < lrem) {
    lrem -= dz;
    cp = (vel < 0.0) ? p_zp1(d, cp) : p_zm1(d, cp);
    if (cpb == cp) {
      lrem = fmod(lrem, dz);
      break;
    }
    const double hh = p_f3(h, cp);
    if (hh > 0) dz = fabs(hh);
    cpb = cp;
    if (++it > FF_MAXWALK) {
      atomicExch(err, 1);
      break;
    }
  }
  crf = fmx(fmn(lrem / dz, 1.0), 0.0);
  cl = cp;
}

/* cell-centred velocities and w9 (tracers.c:524-559), per wet cell */
__global__ void k_ffsl_vel(emstr_dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  d.fw9[idx] = 0.5 * (d.dz[idx] + ((k > kbc) ? d.dz[idx - d.plane] : 0.0));
  double nu = 0.0, nv = 0.0, nw = 0.0;
  const double u1c = d.u1[idx], u1p = d.u1[idx + 1];
  if ((u1c >= 0.0) && (u1p > 0.0))
    nu = u1c;
  else if ((u1c < 0.0) && (u1p <= 0.0))
    nu = u1p;
  const double u2c = d.u2[idx], u2p = d.u2[idx + d.nx];
  if ((u2c >= 0.0) && (u2p > 0.0))
    nv = u2c;
  else if ((u2c < 0.0) && (u2p <= 0.0))
    nv = u2p;
  const double wc = d.w[idx], wp = (k < d.nz - 1) ? d.w[idx + d.plane] : wc;
  if ((wc >= 0.0) && (wp > 0.0))
    nw = wc;
  else if ((wc < 0.0) && (wp <= 0.0))
    nw = wp;
  d.fnu[idx] = nu;
  d.fnv[idx] = nv;
  d.fnw[idx] = nw;
}

/* trajectories of one sub-step (tracers.c:2132-2251), per wet cell:
 * slots 0..2 the x, y, z faces of the cell, 3..5 its centre */
__global__ void k_ffsl_traj(emstr_dev d, double dt, int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  const ll c = ppack(idx, P_WET);
  ll cl;
  double crf;
  /* centres */
  const double nu = d.fnu[idx], nv = d.fnv[idx], nw = d.fnw[idx];
  cl = (nu < 0.0) ? p_xp1(d, c) : c;
  crf = 0.0;
  ff_hor<0>(d, nu, dt, d.h1au1, cl, crf, err);
  d.fcl[3 * N3 + idx] = cl;
  d.fcrf[3 * N3 + idx] = crf;
  cl = (nv < 0.0) ? p_yp1(d, c) : c;
  crf = 0.0;
  ff_hor<1>(d, nv, dt, d.h2au2, cl, crf, err);
  d.fcl[4 * N3 + idx] = cl;
  d.fcrf[4 * N3 + idx] = crf;
  cl = (nw < 0.0) ? p_zp1(d, c) : c;
  crf = 0.0;
  ff_vert(d, c, nw, dt, d.fw9, cl, crf, err);
  d.fcl[5 * N3 + idx] = cl;
  d.fcrf[5 * N3 + idx] = crf;
  /* faces: west and south faces of the cell, and its bottom face when it
   * is a cell to process */
  const double u1 = d.u1[idx], u2 = d.u2[idx];
  cl = (u1 > 0.0) ? p_xm1(d, c) : c;
  crf = 0.0;
  ff_hor<0>(d, u1, dt, d.h1acell, cl, crf, err);
  d.fcl[idx] = cl;
  d.fcrf[idx] = crf;
  cl = (u2 > 0.0) ? p_ym1(d, c) : c;
  crf = 0.0;
  ff_hor<1>(d, u2, dt, d.h2acell, cl, crf, err);
  d.fcl[N3 + idx] = cl;
  d.fcrf[N3 + idx] = crf;
  crf = 0.0;
  cl = c;
  const int ksc = d.ks[col];
  if (ksc >= 0 && k <= ksc) {
    const double w = d.w[idx];
    cl = (w > 0.0) ? p_zm1(d, c) : c;
    ff_vert(d, c, w, dt, d.dz, cl, crf, err);
  }
  d.fcl[2 * N3 + idx] = cl;
  d.fcrf[2 * N3 + idx] = crf;
}

/* interpolation of a along the trajectory (clc, crfc) of a cell in the
 * direction of velocity v (the transverse terms of tracers.c:2508-2565) */
template <int DIR>
__device__ __forceinline__ double ff_tv(const emstr_dev &d, const double *a, const double *t,
                                        const double *ts, double v, ll clc, double crfc) {
  ll up1, cp;
  if (DIR == 0) {
    up1 = (v < 0.0) ? clc : p_xm1(d, clc);
    cp = (v < 0.0) ? p_xm1(d, up1) : clc;
  } else if (DIR == 1) {
    up1 = (v < 0.0) ? clc : p_ym1(d, clc);
    cp = (v < 0.0) ? p_ym1(d, up1) : clc;
  } else {
    up1 = (v < 0.0) ? clc : p_zm1(d, clc);
    cp = (v < 0.0) ? p_zm1(d, up1) : clc;
  }
  const double acp = p_val(a, t, ts, cp);
  return acp + (crfc * (p_val(a, t, ts, up1) - acp));
}

/* one-dimensional transverse tracer along direction DIR with the
 * monotone clamp (tracers.c:2508-2546, monoco :2717-2726) */
template <int DIR>
__device__ __forceinline__ double ff_mod1(const emstr_dev &d, const double *t, const double *ts,
                                          double v, ll clc, double crfc) {
  ll up1, cp;
  if (DIR == 0) {
    up1 = (v < 0.0) ? clc : p_xm1(d, clc);
    cp = (v < 0.0) ? p_xm1(d, up1) : clc;
  } else if (DIR == 1) {
    up1 = (v < 0.0) ? clc : p_ym1(d, clc);
    cp = (v < 0.0) ? p_ym1(d, up1) : clc;
  } else {
    up1 = (v < 0.0) ? clc : p_zm1(d, clc);
    cp = (v < 0.0) ? p_zm1(d, up1) : clc;
  }
  const double tcp = p_val(t, t, ts, cp), tup = p_val(t, t, ts, up1);
  double r = tcp;
  if (crfc > 0.) r += (crfc * (tup - tcp));
  const double mn = fmn(tcp, tup), mx = fmx(tcp, tup);
  r = fmx(mn * 1.0, r);
  r = fmn(mx * 1.0, r);
  return r;
}

/* the transverse tracers of one tracer, per wet cell */
__global__ void k_ffsl_mod1(emstr_dev d, const double *__restrict__ t,
                            const double *__restrict__ ts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  if (k < d.kb[col]) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  d.ftmx[idx] = ff_mod1<0>(d, t, ts, d.fnu[idx], d.fcl[3 * N3 + idx], d.fcrf[3 * N3 + idx]);
  d.ftmy[idx] = ff_mod1<1>(d, t, ts, d.fnv[idx], d.fcl[4 * N3 + idx], d.fcrf[4 * N3 + idx]);
  d.ftmz[idx] = ff_mod1<2>(d, t, ts, d.fnw[idx], d.fcl[5 * N3 + idx], d.fcrf[5 * N3 + idx]);
}

/* the modified tracer for the fluxes along DIR (tracers.c:2568-2690) */
template <int DIR>
__global__ void k_ffsl_mod(emstr_dev d, const double *__restrict__ t,
                           const double *__restrict__ ts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  if (k < d.kb[col]) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  const double onesixth = 1.0 / 6.0;
  const double tc = t[idx];
  double tm = tc;
  if (DIR == 0) {
    tm += (tc + d.ftmy[idx] + d.ftmz[idx]);
    tm += ff_tv<2>(d, d.ftmy, t, ts, d.fnw[idx], d.fcl[5 * N3 + idx], d.fcrf[5 * N3 + idx]);
    tm += ff_tv<1>(d, d.ftmz, t, ts, d.fnv[idx], d.fcl[4 * N3 + idx], d.fcrf[4 * N3 + idx]);
  } else if (DIR == 1) {
    tm += (tc + d.ftmx[idx] + d.ftmz[idx]);
    tm += ff_tv<2>(d, d.ftmx, t, ts, d.fnw[idx], d.fcl[5 * N3 + idx], d.fcrf[5 * N3 + idx]);
    tm += ff_tv<0>(d, d.ftmz, t, ts, d.fnu[idx], d.fcl[3 * N3 + idx], d.fcrf[3 * N3 + idx]);
  } else {
    tm += (tc + d.ftmx[idx] + d.ftmy[idx]);
    tm += ff_tv<0>(d, d.ftmy, t, ts, d.fnu[idx], d.fcl[3 * N3 + idx], d.fcrf[3 * N3 + idx]);
    tm += ff_tv<1>(d, d.ftmx, t, ts, d.fnv[idx], d.fcl[4 * N3 + idx], d.fcrf[4 * N3 + idx]);
  }
  tm *= onesixth;
  d.ftm[idx] = tm;
}

/* van Leer face value on the modified tracer a (tracers.c:2868-2910) */
template <int DIR>
__device__ __forceinline__ double ff_vl(const emstr_dev &d, const double *a, const double *t,
                                        const double *ts, ll clc, double vel, double crf) {
  ll cp = clc, bm1, bm2, fp = 0;
  if (DIR == 0) {
    if (vel > 0.0) cp = p_xp1(d, cp);
    bm1 = p_xm1(d, cp);
    bm2 = p_xm1(d, bm1);
    if (vel < 0.0) fp = p_xp1(d, cp);
  } else if (DIR == 1) {
    if (vel > 0.0) cp = p_yp1(d, cp);
    bm1 = p_ym1(d, cp);
    bm2 = p_ym1(d, bm1);
    if (vel < 0.0) fp = p_yp1(d, cp);
  } else {
    if (vel > 0.0) cp = p_zp1(d, cp);
    bm1 = p_zm1(d, cp);
    bm2 = p_zm1(d, bm1);
    if (vel < 0.0) fp = p_zp1(d, cp);
  }
  /* the stencil mirrored for vel <= 0: (bm2, bm1, cp) for vel > 0,
   * (fp, cp, bm1) otherwise; the reference's acp - 0.5 (1 - crf) df of
   * df = (acp - abm1) (afp - acp), 2 df / (afp - abm1) is the same value
   * with the negated quotient, to the last bit, without a branch */
  const bool pos = vel > 0.0;
  const double acp = p_val(a, t, ts, cp), abm1 = p_val(a, t, ts, bm1);
  const double a2 = p_val(a, t, ts, pos ? bm2 : fp);
  const double a1 = pos ? abm1 : acp, a0 = pos ? acp : abm1;
  double df = (a1 - a2) * (a0 - a1);
  df = (df > 0.0) ? 2.0 * df / (a0 - a2) : 0.0;
  return a1 + 0.5 * (1.0 - crf) * df;
}

/* fluxes along DIR of one tracer from the modified tracer ftm
 * (tracers.c:2578-2705): the fractional part by van Leer, the whole
 * cells of the trajectory by their integrals, times the volume flux
 * (horizontal) or with the sign of w (vertical) */
template <int DIR>
__global__ void k_ffsl_flux(emstr_dev d, const double *__restrict__ t,
                            const double *__restrict__ ts, double dt, int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  const ll c = ppack(idx, P_WET);
  const double *a = d.ftm;
  const ll clc = d.fcl[DIR * N3 + idx];
  const double crf = d.fcrf[DIR * N3 + idx];
  double F = 0.0;
  if (DIR < 2) {
    const double vel = (DIR == 0) ? d.u1[idx] : d.u2[idx];
    const double *h = (DIR == 0) ? d.h1acell : d.h2acell;
    if (vel != 0.0) {
      if (crf != 0.0) F = ff_vl<DIR>(d, a, t, ts, clc, vel, crf);
      F *= (crf * h[p_col(d, clc)]);
      ll cp = (vel > 0.0) ? ((DIR == 0) ? p_xm1(d, c) : p_ym1(d, c)) : c;
      double dx = h[p_col(d, cp)];
      const double ltraj = fabs(vel * dt);
      double lrem = ltraj;
      int it = 0;
      while (dx < lrem) {
        F += (dx * p_val(a, t, ts, cp));
        lrem -= dx;
        if (DIR == 0) cp = (vel < 0) ? p_xp1(d, cp) : p_xm1(d, cp);
        else cp = (vel < 0) ? p_yp1(d, cp) : p_ym1(d, cp);
        const double hh = h[p_col(d, cp)];
        if (hh > 0) dx = hh;
        if (++it > FF_MAXWALK) {
          atomicExch(err, 1);
          break;
        }
      }
      F *= (((DIR == 0) ? d.u1flux3d[idx] : d.u2flux3d[idx]) / ltraj);
    }
    if (DIR == 0) d.ffx[idx] = F;
    else d.ffy[idx] = F;
    return;
  }
  const int ksc = d.ks[col];
  if (ksc >= 0 && k <= ksc) {
    const double w = d.w[idx];
    if (w != 0.0) {
      if (crf != 0.0) F = ff_vl<2>(d, a, t, ts, clc, w, crf);
      F *= (crf * p_f3(d.dz, clc));
      ll cp = (w > 0.0) ? p_zm1(d, c) : c;
      double dz = fabs(p_f3(d.dz, cp));
      double lrem = fabs(w * dt);
      int it = 0;
      while (dz < lrem) {
        F += (dz * p_val(a, t, ts, cp));
        lrem -= dz;
        cp = (w < 0.0) ? p_zp1(d, cp) : p_zm1(d, cp);
        const double hh = p_f3(d.dz, cp);
        if (hh > 0) dz = hh;
        if (++it > FF_MAXWALK) {
          atomicExch(err, 1);
          break;
        }
      }
      if (w < 0) F *= -1;
    }
  }
  d.ffz[idx] = F;
}

extern "C" int emstr_launch_ffsl_traj(const emstr_dev *d, double dtu, int *err,
                                      void *stream) {
  dim3 b(64, 4, 1);
  dim3 g = grid3(d, b);
  hipStream_t st = (hipStream_t)stream;
  k_ffsl_vel<<<g, b, 0, st>>>(*d);
  k_ffsl_traj<<<g, b, 0, st>>>(*d, dtu, err);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_launch_ffsl_fluxes(const emstr_dev *d, const double *t,
                                        const double *ts, double dtu, int *err,
                                        void *stream) {
  dim3 b(64, 4, 1);
  dim3 g = grid3(d, b);
  hipStream_t st = (hipStream_t)stream;
  k_ffsl_mod1<<<g, b, 0, st>>>(*d, t, ts);
  k_ffsl_mod<0><<<g, b, 0, st>>>(*d, t, ts);
  k_ffsl_flux<0><<<g, b, 0, st>>>(*d, t, ts, dtu, err);
  k_ffsl_mod<1><<<g, b, 0, st>>>(*d, t, ts);
  k_ffsl_flux<1><<<g, b, 0, st>>>(*d, t, ts, dtu, err);
  k_ffsl_mod<2><<<g, b, 0, st>>>(*d, t, ts);
  k_ffsl_flux<2><<<g, b, 0, st>>>(*d, t, ts, dtu, err);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

/*-------------------------------------------------------------------*/
/* k_vdiff: implicit vertical diffusion, in place                    */
/*-------------------------------------------------------------------*/
#define VD_COLS 64
#ifndef VD_TLANES
#define VD_TLANES 16 /* tracer lanes per block: 16 waves share one column matrix in LDS */
#endif

__global__ void __launch_bounds__(VD_COLS *VD_TLANES)
k_vdiff(emstr_dev d, double *__restrict__ tr, double *__restrict__ scratch,
        int *__restrict__ err) {
  extern __shared__ double lds[];
  const int nz = d.nz;
  double *Ldiv = lds;                       /* [nz][VD_COLS] */
  double *Lud = lds + nz * VD_COLS;         /* [nz][VD_COLS] */
  double *Lbm1 = lds + 2 * nz * VD_COLS;    /* [nz][VD_COLS] */
  __shared__ int Lcb[VD_COLS], Lcs[VD_COLS], Lmode[VD_COLS];
  const int lane = threadIdx.x, tl = threadIdx.y;
  const int cidx = blockIdx.x * VD_COLS + lane;
  const ll P = d.plane;
  const double dt = d.dt;
  ll col = (cidx < d.ncol) ? d.cols[cidx] : 0;

  if (tl == 0) {
    int mode = 0; /* 0 skip, 1 single layer, 2 tridiagonal */
    int cb = 0, cs = 0;
    if (cidx < d.ncol && d.knew[col] >= d.kb[col]) {
      const int kbc = d.kb[col];
      cs = d.knew[col];
      /* matrix coefficients (tracers.c:3836-3889) */
      int zm1 = cs - 1; /* layer below; kbc-1 is the sediment ghost */
      double dzm1 = (zm1 >= kbc) ? d.dz[(ll)zm1 * P + col] : 0.0;
      if (zm1 < kbc || dzm1 == 0.0) {
        mode = 1;
        cb = cs;
      } else {
        int c = cs;
        /* walk down while c is a cell (not the ghost) with dz != 0 */
        while (c >= kbc && d.dz[(ll)c * P + col] != 0.0) c--;
        cb = c + 1;
        mode = (cb < cs) ? 2 : 0; /* a zero-thickness surface: no solve */
      }
      if (mode == 2) {
        /* build and factor; dzface(k) = 0.5*(dz(k-1)+dz(k)) */
        double dzk = d.dz[(ll)cb * P + col];
        double dzp = (cb + 1 <= cs) ? d.dz[(ll)(cb + 1) * P + col] : 0.0;
        double Kp = d.Kz[(ll)(cb + 1) * P + col];
        double dzf_p = 0.5 * (dzk + dzp);
        double dzdt = dzk / dt;
        double Cm1 = 0.0;
        double Cp1 = -Kp / dzf_p;
        double C = dzdt - Cp1;
        double div = C;
        Ldiv[cb * VD_COLS + lane] = div;
        Lbm1[cb * VD_COLS + lane] = Cm1;
        double Cp1prev = Cp1;
        for (int kk = cb + 1; kk <= cs; kk++) {
          double dzc = dzp;
          double dzf_c = dzf_p;
          double Kc = Kp;
          dzdt = dzc / dt;
          Cm1 = -Kc / dzf_c;
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
          /* forward elimination, matrix part (tracers.c:4188-4193) */
          double ud = Cp1prev / div;
          div = C - Cm1 * ud;
          if (div == 0.0) atomicMax(err, (int)(kk * P + col) + 1);
          Lud[kk * VD_COLS + lane] = ud;
          Ldiv[kk * VD_COLS + lane] = div;
          Lbm1[kk * VD_COLS + lane] = Cm1;
          Cp1prev = Cp1;
        }
      }
    }
    Lcb[lane] = cb;
    Lcs[lane] = cs;
    Lmode[lane] = mode;
  }
  __syncthreads();
  const int mode = Lmode[lane];
  if (!mode) return;
  const int cb = Lcb[lane], cs = Lcs[lane];
  const ll N3 = d.n3;
  for (int n = tl; n < d.ntr; n += VD_TLANES) {
    if (!d.in_tdif[n]) continue;
    double *var = tr + n * N3;
    double *sol = scratch + n * N3;
    if (mode == 1) {
      /* tracers.c:4127-4133 with zero bottom and top fluxes */
      ll ic = (ll)cs * P + col;
      double dzc = d.dz[ic];
      var[ic] += dt * (0.0 - 0.0) / fmax(dzc, d.hmin);
      continue;
    }
    /* forward (tracers.c:4147-4195) */
    ll ic = (ll)cb * P + col;
    double dzdt = d.dz[ic] / dt;
    double rhs = dzdt * var[ic] + 0.0;
    double sp = rhs / Ldiv[cb * VD_COLS + lane];
    sol[ic] = sp;
    for (int kk = cb + 1; kk <= cs; kk++) {
      ic = (ll)kk * P + col;
      dzdt = d.dz[ic] / dt;
      rhs = dzdt * var[ic];
      if (kk == cs) rhs = rhs - 0.0;
      sp = (rhs - Lbm1[kk * VD_COLS + lane] * sp) / Ldiv[kk * VD_COLS + lane];
      sol[ic] = sp;
    }
    /* back substitution and update (tracers.c:4199-4205), scale = 1 */
    ic = (ll)cs * P + col;
    double sn = sp;
    var[ic] += (sn - var[ic]) * 1.0;
    for (int kk = cs - 1; kk >= cb; kk--) {
      ic = (ll)kk * P + col;
      double sk = sol[ic];
      sk -= Lud[(kk + 1) * VD_COLS + lane] * sn;
      var[ic] += (sk - var[ic]) * 1.0;
      sn = sk;
    }
  }
}

/*-------------------------------------------------------------------*/
/* k_setdz: set_dz for z-levels, per wet column (tracers.c:5722-5848) */
/*-------------------------------------------------------------------*/
__global__ void k_setdz(emstr_dev d, int *__restrict__ proc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (kbc >= d.nz) return;
  const ll P = d.plane;
  const double *lay = d.layers;
  const double eta = d.eta[col], oldeta = d.oldeta[col];
  /* old surface <- new surface; new surface: walk down while the lower
   * face is at or above eta; kbc-1 stands for the sediment ghost */
  const int kold = d.knew[col];
  int c = d.nz - 1;
  while (c >= kbc && lay[c] >= eta) c--;
  const int knew = c;
  d.kold[col] = kold;
  d.knew[col] = knew;
  /* processed? the lower of the two surfaces must be a cell */
  const int klow = min(kold, knew);
  double top = fmin(eta, oldeta);
  double bot = 0.05 * d.hmin + d.botz[col];
  /* tracers.c:5773 (SHOC) tests top > bot, hd-us tracers.c:7325 top >= bot */
  int p = ((d.us ? (top >= bot) : (top > bot)) && klow >= kbc) ? 1 : 0;
  d.ks[col] = p ? klow : -1;
  proc[col] = p;
  /* dz below the old surface, bed cell from botz, copy above */
  double tp = oldeta;
  int c3 = kold;
  while (c3 > kbc) {
    double b = lay[c3];
    d.dz[(ll)c3 * P + col] = tp - b;
    tp = b;
    c3--;
  }
  d.dz[(ll)kbc * P + col] = tp - d.botz[col];
  /* cells above the old surface take its thickness; a ghost surface
   * (kold < kbc) the ghost's */
  double dzs = (kold >= kbc) ? d.dz[(ll)kold * P + col] : (d.dzg ? d.dzg[col] : 0.0);
  for (int kk = max(kold, kbc - 1) + 1; kk < d.nz; kk++) d.dz[(ll)kk * P + col] = dzs;
}

/*-------------------------------------------------------------------*/
/* sub-step elevations (tracers.c:749-754, 1011-1016)                 */
/*-------------------------------------------------------------------*/
__global__ void k_subeta(emstr_dev d, double dtu, double trem, int init) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.ncol) return;
  const ll col = d.cols[q];
  if (init) d.osubeta[col] = d.oldeta[col];
  d.subeta[col] = d.oldeta[col] +
                  (1.0 + (dtu - trem) / d.dt) * (d.eta[col] - d.oldeta[col]);
}

__global__ void k_osub(emstr_dev d) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.ncol) return;
  const ll col = d.cols[q];
  d.osubeta[col] = d.subeta[col];
}

/*-------------------------------------------------------------------*/
/* minimum sub-step over the cells to process (tracers.c:595-655);    */
/* positive doubles compare as their bit patterns                     */
/*-------------------------------------------------------------------*/
__global__ void k_dtm(emstr_dev d, double dtu0, unsigned long long *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int ksc = d.ks[col];
  const int kbc = d.kb[col];
  if (ksc < 0 || k < kbc || k > ksc) return;
  const ll idx = (ll)k * d.plane + col;
  const double sf = 0.8, minval = 1e-10;
  double dtm = dtu0;
  const bool ffsl = (d.trasc == FFSL);
  double d4 = 0.5 * (d.u1[idx] + d.u1[idx + 1]);
  /* FFSL is limited by the velocity gradients (Lipschitz number) */
  if (ffsl) d4 = fmx(fabs(d.u1[idx + 1] - d.u1[idx]), d.u1kh[idx] / d.h1acell[col]);
  if (fabs(d4) < minval) d4 = minval;
  double d1 = sf * fabs(d.h1acell[col] / d4);
  if (d1 < dtm) dtm = d1;
  d4 = 0.5 * (d.u2[idx] + d.u2[idx + d.nx]);
  if (ffsl) d4 = fmx(fabs(d.u2[idx + d.nx] - d.u2[idx]), d.u2kh[idx] / d.h2acell[col]);
  if (fabs(d4) < minval) d4 = minval;
  double d2 = sf * fabs(d.h2acell[col] / d4);
  if (d2 < dtm) dtm = d2;
  const bool surf = (k == ksc);
  const ll zp = (k + 1 < d.nz) ? idx + d.plane : idx;
  d4 = surf ? 0.0 : 0.5 * (d.w[idx] + d.w[zp]);
  if (ffsl) d4 = surf ? 0.0 : fabs(d.w[zp] - d.w[idx]);
  if (fabs(d4) < minval) d4 = minval;
  double d3 = surf ? dtu0 : sf * fabs(d.dz[idx] * 1.0 / d4);
  if (d3 > 0.0 && d3 < dtm) dtm = d3;
  if (dtm < dtu0) atomicMin(out, (unsigned long long)__double_as_longlong(dtm));
}

/*-------------------------------------------------------------------*/
/* sparse <-> dense                                                  */
/*-------------------------------------------------------------------*/
__global__ void k_scatter(const int *__restrict__ wsp, const ll *__restrict__ wdn,
                          int nwet, const double *__restrict__ sp,
                          double *__restrict__ dn) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nwet) dn[wdn[q]] = sp[wsp[q]];
}

__global__ void k_gather(const int *__restrict__ wsp, const ll *__restrict__ wdn,
                         int nwet, const double *__restrict__ dn,
                         double *__restrict__ sp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nwet) sp[wsp[q]] = dn[wdn[q]];
}

/* ghost entries of a sparse array: dst[g[q]] = dst[src[q]] */
__global__ void k_ghosts(const int *__restrict__ g, const int *__restrict__ src,
                         int n, double *__restrict__ a) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) a[g[q]] = a[src[q]];
}

/*-------------------------------------------------------------------*/
/* tracer mass: per column (fixed k order, TwoSum-compensated) then   */
/* per block; partial (sum, err) pairs are added on the host in a    */
/* fixed order.                                                       */
/*-------------------------------------------------------------------*/
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__global__ void k_mass(emstr_dev d, const double *__restrict__ tr,
                       double *__restrict__ colmass, double *__restrict__ part) {
  const int n = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double ss[256], se[256];
  double s = 0.0, e = 0.0;
  if (q < d.ncol) {
    const ll col = d.cols[q];
    const int kbc = d.kb[col], ksc = d.ks[col];
    const double a = d.cellarea[col];
    const double *t = tr + n * d.n3;
    for (int k = ksc; k >= kbc; k--) {
      ll idx = (ll)k * d.plane + col;
      double m = t[idx] * a * d.dz[idx];
      double s2, e2;
      two_sum(s, m, s2, e2);
      s = s2;
      e += e2;
    }
    if (colmass) colmass[(ll)n * d.ncol + q] = s + e;
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
/* launchers                                                         */
/*-------------------------------------------------------------------*/
static dim3 grid3(const emstr_dev *d, dim3 b) {
  return dim3((d->nx + b.x - 1) / b.x, (d->ny + b.y - 1) / b.y, d->nz);
}

extern "C" int emstr_launch_adv(const emstr_dev *d, const double *tin,
                                double *tout, const double *tstart, double dtu,
                                double trem, int anydif, int n0, int n1, void *stream) {
  dim3 b(64, 4, 1);
  dim3 g = grid3(d, b);
  hipStream_t st = (hipStream_t)stream;
  /* FFSL only: the other schemes take the face-once path */
  if (d->trasc != FFSL) return -1;
  k_adv<FFSL><<<g, b, 0, st>>>(*d, tin, tout, tstart, dtu, trem, anydif, n0, n1);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <int SCH>
static void face_launch(const emstr_dev *d, const double *tin, double *tout,
                        const double *tstart, double dtu, int anydif, double *fw, double *fs,
                        double *fb, hipStream_t st) {
  dim3 b(64, 4, 1);
  dim3 g = grid3(d, b);
  k_face<SCH><<<g, b, 0, st>>>(*d, tin, tstart, dtu, anydif, fw, fs, fb);
  k_upd<SCH><<<g, b, 0, st>>>(*d, tin, tout, dtu, fw, fs, fb);
}

extern "C" int emstr_launch_adv_faces(const emstr_dev *d, const double *tin, double *tout,
                                      const double *tstart, double dtu, int anydif,
                                      double *fw, double *fs, double *fb, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  switch (d->trasc) {
    case VANLEER:
      face_launch<VANLEER>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case QUICKEST:
      face_launch<QUICKEST>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case QUICKEST_CO:
      face_launch<QUICKEST_CO>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case ORDER1:
      face_launch<ORDER1>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case ORDER2:
      face_launch<ORDER2>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case ORDER4:
      face_launch<ORDER4>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    case ORDER2_UW:
      face_launch<ORDER2_UW>(d, tin, tout, tstart, dtu, anydif, fw, fs, fb, st);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_vdiff(const emstr_dev *d, double *tr,
                                  double *scratch, int *err, void *stream) {
  if (d->ncol == 0) return 0;
  dim3 b(VD_COLS, VD_TLANES, 1);
  dim3 g((d->ncol + VD_COLS - 1) / VD_COLS);
  size_t lds = (size_t)3 * d->nz * VD_COLS * sizeof(double);
  if (lds > 48 * 1024)
    hipFuncSetAttribute((const void *)k_vdiff,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  k_vdiff<<<g, b, lds, (hipStream_t)stream>>>(*d, tr, scratch, err);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_setdz(const emstr_dev *d, int *proc, void *stream) {
  dim3 b(64, 4, 1);
  dim3 g((d->nx + 63) / 64, (d->ny + 3) / 4, 1);
  k_setdz<<<g, b, 0, (hipStream_t)stream>>>(*d, proc);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_subeta(const emstr_dev *d, double dtu, double trem,
                                   int init, void *stream) {
  if (d->ncol == 0) return 0;
  k_subeta<<<(d->ncol + 255) / 256, 256, 0, (hipStream_t)stream>>>(*d, dtu, trem, init);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_osub(const emstr_dev *d, void *stream) {
  if (d->ncol == 0) return 0;
  k_osub<<<(d->ncol + 255) / 256, 256, 0, (hipStream_t)stream>>>(*d);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_dtm(const emstr_dev *d, double dtu0,
                                unsigned long long *out, void *stream) {
  dim3 b(64, 4, 1);
  k_dtm<<<grid3(d, b), b, 0, (hipStream_t)stream>>>(*d, dtu0, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_scatter(const int *wsp, const long long *wdn,
                                    int nwet, const double *sp, double *dn,
                                    void *stream) {
  if (nwet == 0) return 0;
  k_scatter<<<(nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(wsp, wdn, nwet, sp, dn);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_gather(const int *wsp, const long long *wdn,
                                   int nwet, const double *dn, double *sp,
                                   void *stream) {
  if (nwet == 0) return 0;
  k_gather<<<(nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(wsp, wdn, nwet, dn, sp);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_ghosts(const int *g, const int *src, int n,
                                   double *a, void *stream) {
  if (n == 0) return 0;
  k_ghosts<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(g, src, n, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_mass_blocks(const emstr_dev *d) { return (d->ncol + 255) / 256; }

extern "C" int emstr_launch_mass(const emstr_dev *d, const double *tr,
                                 double *colmass, double *part, void *stream) {
  int nb = emstr_mass_blocks(d);
  if (nb == 0) return 0;
  dim3 g(nb, d->ntr, 1);
  k_mass<<<g, 256, 0, (hipStream_t)stream>>>(*d, tr, colmass, part);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/*-------------------------------------------------------------------*/
/* w from continuity, bottom up, per processed column                 */
/* (model/hd/momentum/vel3d.c:4680-4711; no sources or relaxation).   */
/* w must be zeroed before the launch.                                */
/*-------------------------------------------------------------------*/
__global__ void k_velw(emstr_dev d, const double *__restrict__ wbot) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.ncol) return;
  const ll col = d.cols[q];
  const int kbc = d.kb[col], kold = d.kold[col];
  const ll P = d.plane, nx = d.nx;
  const double area = d.cellarea[col];
  ll idx = (ll)kbc * P + col;
  d.w[idx] = wbot[col];
  double fcbot = d.w[idx] * area;
  for (int c = kbc; c < kold; c++) {
    idx = (ll)c * P + col;
    double fctop = fcbot + d.u1flux3d[idx] - d.u1flux3d[idx + 1] +
                   d.u2flux3d[idx] - d.u2flux3d[idx + nx];
    d.w[idx + P] = fctop / area;
    fcbot = fctop;
  }
}

/*-------------------------------------------------------------------*/
/* set_flux_3d (model/hd/momentum/vel3d.c:4041-4191, VEL3D mode), one  */
/* face direction, one thread per dense face: the window's wet faces   */
/* (fw) carry velocity x face thickness x face length x metric factor  */
/* in the reference's order, faces above the column's surface face     */
/* (ksur) and every other face zero; then k_fluxcb gives the faces one */
/* cell deep the 2D flux over the step / dt (the reference writes them */
/* before zeroing above the surface: a listed face above it stays      */
/* zero).  Under FFSL (uvm not NULL) the wet faces carry the mean      */
/* volume flux and nothing else is changed.                            */
/*-------------------------------------------------------------------*/
__global__ void __launch_bounds__(256)
k_flux3d(ll n3, ll plane, const unsigned char *__restrict__ fw, const int *__restrict__ ksur,
         const double *__restrict__ u, const double *__restrict__ dzu,
         const double *__restrict__ hf, const double *__restrict__ md,
         const double *__restrict__ uvm, double *__restrict__ out) {
  const ll idx = (ll)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n3) return;
  const ll col = idx % plane;
  const int k = (int)(idx / plane);
  double f = 0.0;
  if (uvm) {
    if (fw[idx]) f = uvm[idx];
  } else {
    if (fw[idx]) f = u[idx] * dzu[idx] * hf[col] * md[col];
    if (k > ksur[col]) f = 0.0;
  }
  out[idx] = f;
}

__global__ void __launch_bounds__(256)
k_fluxcb(int n, const ll *__restrict__ list, ll plane, const int *__restrict__ ksur,
         const double *__restrict__ uf, double dt, double *__restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const ll idx = list[q], col = idx % plane;
  if ((int)(idx / plane) <= ksur[col]) out[idx] = uf[col] / dt;
}

extern "C" int emstr_launch_flux3d(const emstr_dev *d, const unsigned char *fw, const int *ksur,
                                   const double *u, const double *dzu, const double *hf,
                                   const double *md, const double *uvm, int ncb,
                                   const long long *cb, const double *uf, double dt,
                                   double *out, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  if (d->n3)
    k_flux3d<<<(unsigned)((d->n3 + 255) / 256), 256, 0, st>>>(d->n3, d->plane, fw, ksur, u, dzu,
                                                              hf, md, uvm, out);
  if (!uvm && ncb > 0)
    k_fluxcb<<<(ncb + 255) / 256, 256, 0, st>>>(ncb, cb, d->plane, ksur, uf, dt, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/*-------------------------------------------------------------------*/
/* set_dz_at_u1 / set_dz_at_u2, z-levels (model/hd/momentum/vel3d.c:   */
/* 760-852), one thread per column: for a face column (kbf >= 0, the  */
/* layer of its bottom face) the surface face is the highest layer    */
/* whose lower face is not above the higher elevation either side (a  */
/* land neighbour reads the column's own elevation, its lateral       */
/* ghost's value) and not below the face bed; thicknesses from it     */
/* down to the bottom face (the bottom one to the face bed), the       */
/* surface face's above it, zero below.  ksur gets the surface face's  */
/* layer (nz for a column that is not a face column).  dir 1: e1       */
/* faces (neighbour at i-1), 2: e2 (j-1).                              */
/*-------------------------------------------------------------------*/
__global__ void __launch_bounds__(256)
k_dzu(emstr_dev d, int dir, const int *__restrict__ kbf, const double *__restrict__ botzu,
      double *__restrict__ dzu, int *__restrict__ ksur) {
  const ll col = (ll)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d.plane) return;
  const int nz = d.nz, kf = kbf[col];
  const ll P = d.plane;
  if (kf < 0) {
    for (int k = 0; k < nz; k++) dzu[(ll)k * P + col] = 0.0;
    ksur[col] = nz;
    return;
  }
  const ll i = col % d.nx, j = col / d.nx;
  const ll nb = (dir == 1) ? col - 1 : col - d.nx;
  const bool haven = (dir == 1) ? (i > 0) : (j > 0);
  const double en = (haven && d.kb[nb] < nz) ? d.eta[nb] : d.eta[col];
  const double ec = d.eta[col];
  double top = (ec > en) ? ec : en;
  top = (top > botzu[col]) ? top : botzu[col];
  const double ms = top;
  /* the walk down from the top layer stops at the sediment ghost */
  int ks = nz - 1;
  const int kb = d.kb[col];
  while (ks >= kb && d.layers[ks] > ms) ks--;
  ksur[col] = ks;
  for (int k = 0; k < kf; k++) dzu[(ll)k * P + col] = 0.0;
  double t = ms;
  for (int k = ks; k > kf; k--) {
    const double bt = d.layers[k];
    dzu[(ll)k * P + col] = t - bt;
    t = bt;
  }
  dzu[(ll)kf * P + col] = t - botzu[col];
  const double st = dzu[(ll)ks * P + col];
  for (int k = ks + 1; k < nz; k++) dzu[(ll)k * P + col] = st;
}

extern "C" int emstr_launch_dzu(const emstr_dev *d, int dir, const int *kbf,
                                const double *botzu, double *dzu, int *ksur, void *stream) {
  if (d->plane == 0) return 0;
  k_dzu<<<(unsigned)((d->plane + 255) / 256), 256, 0, (hipStream_t)stream>>>(*d, dir, kbf,
                                                                             botzu, dzu, ksur);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_launch_velw(const emstr_dev *d, const double *wbot,
                                 void *stream) {
  if (d->ncol == 0) return 0;
  k_velw<<<(d->ncol + 255) / 256, 256, 0, (hipStream_t)stream>>>(*d, wbot);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/*-------------------------------------------------------------------*/
/* auxiliary_routines (tracers.c:4296-4490): the parts that change   */
/* the window's state after the transport                            */
/*-------------------------------------------------------------------*/
/* reset_dz (tracers.c:5900-5932), per wet column: the cells from the
 * deeper of the old and new surfaces up to the new surface take the
 * full layer thickness from the deeper cell's lower face (or the bed),
 * the new surface cell the height of the updated elevation above its
 * lower face (kept if the elevation is not above it); layer kb - 1
 * stands for the sediment ghost, whose lower face is layers[kb - 1]
 * (its thickness goes to dzg, which set_dz copies up a column whose
 * old surface is the ghost).  p2: the column is a cell to process
 * for the updated elevation (tracers.c:5946-5952), its surface layer
 * the new surface. */
__global__ void k_resetdz(emstr_dev d, int *__restrict__ p2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (kbc >= d.nz) return;
  const ll P = d.plane;
  const double *lay = d.layers;
  const int knew = d.knew[col];
  int k = min(d.kold[col], knew);
  double bot = fmx(lay[max(k, 0)], d.botz[col]);
  while (k < knew) {
    const double top = lay[k + 1];
    if (k >= kbc) d.dz[(ll)k * P + col] = top - bot;
    else d.dzg[col] = top - bot;
    bot = top;
    k++;
  }
  const double top = d.eta[col];
  if (top > bot) {
    if (k >= kbc) d.dz[(ll)k * P + col] = top - bot;
    else d.dzg[col] = top - bot;
  }
  p2[col] = (d.eta[col] > 0.05 * d.hmin + d.botz[col] && knew >= kbc) ? 1 : 0;
}

/* bc_leastsq (boundary/bdrycond.c:180-214): the least-squares line
 * through (1, y1) .. (4, y4), evaluated at 5, in the reference's order */
__device__ double obc_leastsq(double y1, double y2, double y3, double y4) {
  const double x[5] = {0.0, 1.0, 2.0, 3.0, 4.0};
  const double y[5] = {0.0, y1, y2, y3, y4};
  double sxy = 0.0, sx = 0.0, sy = 0.0, sx2 = 0.0;
  const int nn = 4;
  for (int i = 1; i <= nn; i++) {
    sx += x[i];
    sy += y[i];
    sxy += (x[i] * y[i]);
    sx2 += (x[i] * x[i]);
  }
  const double a1 = (sxy - (sx * sy / nn)) / (sx2 - (sx * sx) / nn);
  const double a2 = sy / nn - a1 * sx / nn;
  return a2 + a1 * 5.0;
}

/* set_OBC_tr (tracers.c:4569-5165), the conditions built, one thread
 * per cell of one open boundary, every tracer of tbdy: the new value
 * from the condition bits (NOGRAD above trpc, NOTHIN, CLAMPD, LINEXT
 * through the inward chain, FILEIN below trpc, then UPSTRM with the
 * velocity the host picked by upmeth), written unless NOTHIN without
 * OP_OWRITE.  A boundary's cells read only cells inwards of it (and
 * themselves), so its threads do not read what another writes. */
__global__ void k_obc(emstr_dev d, ObcArgs a, double *__restrict__ tr,
                      const double *__restrict__ tsr) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const ll N3 = d.n3;
  const ll c = a.cell[q], c1 = a.in1[q];
  const ll e0 = a.chain[4 * q], e1 = a.chain[4 * q + 1], e2 = a.chain[4 * q + 2],
           e3 = a.chain[4 * q + 3];
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n]) continue;
    double *t = tr + n * N3;
    const double *ts = tsr + n * N3;
    /* a cell inwards, or a lateral ghost's starting value */
    auto rd = [&](ll e) { return (e >= 0) ? t[e] : ts[~e]; };
    const int bc = a.bcond[n];
    const double pc = a.trpc[n];
    double nv = 0.0;
    if (bc & OBC_NOGRAD) {
      if (pc == 0.0 || a.cz[q] > pc) nv = rd(c1);
    } else if (bc & OBC_NOTHIN) {
      nv = t[c];
    } else if (bc & OBC_CLAMPD) {
      nv = a.clampv[n];
    } else if (bc & OBC_LINEXT) {
      nv = obc_leastsq(rd(e3), rd(e2), rd(e1), rd(e0));
    }
    if (bc & OBC_FILEIN)
      if (pc == 0.0 || a.cz[q] <= pc) nv = a.fv[(ll)a.frow[n] * a.n + q];
    if (bc & OBC_UPSTRM) {
      const double v = a.vel[q], sc = a.sc, tc = t[c];
      nv = tc - a.dt / a.hat[q] *
                    (0.5 * (sc * v + fabs(v)) * (tc - rd(c1)) + 0.5 * (sc * v - fabs(v)) * (nv - tc));
    }
    if (!(bc & OBC_NOTHIN) || a.owrite) t[c] = nv;
  }
}

extern "C" int emstr_launch_obc(const emstr_dev *d, const ObcArgs *a, double *tr,
                                const double *ts, void *stream) {
  if (a->n > 0)
    k_obc<<<(a->n + 255) / 256, 256, 0, (hipStream_t)stream>>>(*d, *a, tr, ts);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* do_ts_relax's no-gradient relaxation field (tracers.c:6137-6148), per
 * wet column: below the first cell from the top whose centre is not
 * above depth (or the bed cell), every cell above it takes its value */
__global__ void k_relaxgrd(emstr_dev d, const double *__restrict__ cellz, double *__restrict__ rf,
                           double depth) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (kbc >= d.nz) return;
  const ll P = d.plane;
  int k = d.nz - 1;
  while (cellz[(ll)k * P + col] > depth && k > kbc) k--;
  const double v = rf[(ll)k * P + col];
  for (int kk = k + 1; kk < d.nz; kk++) rf[(ll)kk * P + col] = v;
}

/* Per wet cell, in the reference's order: calc_age (tracers.c:6966-6971),
 * tracer_decay over the cells to process (tracer_decay.c:55-61; the
 * tracers of tdec in order), do_ts_relax's relaxation of temperature then
 * salinity (tracers.c:6150-6190) */
__global__ void k_auxcell(emstr_dev d, AuxArgs a, double *__restrict__ tr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll idx = (ll)k * d.plane + col;
  const ll N3 = d.n3;
  if (a.age >= 0 && a.agemsk[idx] != 0.0) tr[a.age * N3 + idx] += a.dttr * (1.0 / 86400.0);
  const int ksc = d.ks[col];
  if (ksc >= 0 && k <= ksc) {
    for (int nn = 0; nn < a.ntdec; nn++) {
      const double decay = a.dec[nn];
      if (decay == 0.0) continue;
      double *t = tr + a.tdec[nn] * N3;
      const double dd = (a.dtr3[nn] >= 0) ? tr[a.dtr3[nn] * N3 + idx] : decay;
      t[idx] -= d.dt * t[idx] / (dd * a.dsc[nn]);
    }
  }
  for (int q = 0; q < 2; q++) {
    const int n = q ? a.sno : a.tno;
    const double *rf = q ? a.rsalt : a.rtemp;
    const double *tc = q ? a.salt_tc : a.temp_tc;
    const double rr = q ? a.rrS : a.rrT;
    const int tcf = q ? a.tcS : a.tcT;
    if (n < 0 || !rf) continue;
    double *t = tr + n * N3;
    if (rr != 0.0) t[idx] -= d.dt * rr * (t[idx] - rf[idx]);
    if (tc && tcf) {
      const double r2 = (tc[idx] != 0.0) ? (1.0 / 86400.0) / tc[idx] : 0.0;
      t[idx] -= d.dt * (t[idx] - rf[idx]) * r2;
    }
  }
}

/* do_ts_relax's no-gradient above the surface (tracers.c:6193-6204):
 * the cells above the new surface of every column to process for the
 * updated elevation take its temperature and salinity */
__global__ void k_tsabove(emstr_dev d, const int *__restrict__ p2, double *__restrict__ tr,
                          int tno, int sno) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  if (d.kb[col] >= d.nz || !p2[col]) return;
  const ll P = d.plane, N3 = d.n3;
  const int ks2 = d.knew[col];
  for (int q = 0; q < 2; q++) {
    const int n = q ? sno : tno;
    if (n < 0) continue;
    double *t = tr + n * N3;
    const double v = t[(ll)ks2 * P + col];
    for (int kk = ks2 + 1; kk < d.nz; kk++) t[(ll)kk * P + col] = v;
  }
}

extern "C" int emstr_launch_aux(const emstr_dev *d, const AuxArgs *a, double *tr, int *p2,
                                void *stream) {
  hipStream_t st = (hipStream_t)stream;
  dim3 b(64, 4, 1), g((d->nx + 63) / 64, (d->ny + 3) / 4, 1);
  k_resetdz<<<g, b, 0, st>>>(*d, p2);
  if (a->tracers) {
    if (a->tno >= 0 && a->rtemp && a->grdT)
      k_relaxgrd<<<g, b, 0, st>>>(*d, a->cellz, a->rtemp, a->depthT);
    if (a->sno >= 0 && a->rsalt && a->grdS)
      k_relaxgrd<<<g, b, 0, st>>>(*d, a->cellz, a->rsalt, a->depthS);
    k_auxcell<<<dim3(g.x, g.y, d->nz), b, 0, st>>>(*d, *a, tr);
    if (!d->sigma) k_tsabove<<<g, b, 0, st>>>(*d, p2, tr, a->tno, a->sno);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
