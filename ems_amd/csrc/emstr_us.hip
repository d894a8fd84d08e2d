/*
 * Plain form of the EMS COMPAS tracer step on MI355X (gfx950): FP64, the
 * reference's edge-based schemes, limiter and operation order, so results
 * agree with the reference CPU path to the last bit.  Compiled with
 * -ffp-contract=off (no FMA contraction).
 *
 * Kernels (layout in emstr_usdev.h):
 *   k_usflux<SCH> one thread per wet edge and layer, loop over tracers
 *                 inside: the edge flux (van Leer or
 *                 QUICKEST on edges with ULTIMATE, or first order,
 *                 tracers.c:1637-1743, 2055-2151, 2368-2573, 3785-3833;
 *                 horizontal diffusion, tracers.c:5383-5449) into a
 *                 [tracer][layer][edge] buffer;
 *   k_usfz<SCH>   one thread per wet centre: the vertical fluxes of the
 *                 cells to process (tracers.c:2405-2412, 2125-2150,
 *                 1777-1785) into a [tracer][layer][column] buffer;
 *   k_usadv<SCH>  one thread per wet centre, loop over tracers inside:
 *                 the edge-sign
 *                 divergence (tracers.c:949-959) and the one-stage
 *                 Runge-Kutta update (tracers.c:1000-1016, 1170-1214);
 *                 the processed surface centre also does surf_concrk
 *                 (tracers.c:3085-3260) and tr_bounds (tracers.c:3270-3370)
 *                 for its column.
 *   k_usdtm       minimum sub-step over the cells to process
 *                 (tracers.c:601-683), Courant or (FFSL) Lipschitz.
 * The FFSL fluxes are emstr_us_ffsl.hip's; k_usadv then updates.
 * The column kernels (set_dz, sub-step elevations, vertical diffusion,
 * mass, scatter/gather) are those of the plain SHOC form.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emstr_us_kernels.h"

#define ORDER1 0x000002
#define QUICKEST 0x000020
#define VANLEER 0x000080

typedef long long ll;

__device__ __forceinline__ bool wetk(const int *__restrict__ kb, int col, int k) {
  return col >= 0 && k >= kb[col];
}

/* ultimate_filter (tracers.c:3785-3833) on face value F; trc: tracer at
 * the face's cell c, trcm1 at its cell cm1 */
__device__ __forceinline__ double us_ult(double cr, double F, double trc, double trcm1,
                                         double trp, double trm) {
  double nc, nu, tc, tu, td, diff;
  if (cr == 0.0) return F;
  if (cr > 0.0) {
    tc = trcm1;
    tu = trm;
    td = trc;
  } else {
    tc = trc;
    tu = trp;
    td = trcm1;
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

/* van Leer face value (tracers.c:2544-2573) with the stencil mirrored
 * for a negative velocity: the stencil is (tm2, tm1, tp1) for vel > 0 and
 * (tp2, tp1, tm1) for vel < 0, omc is 1 - cn and 1 + cn.  For vel < 0 the
 * reference forms df = (tp1 - tm1) (tp2 - tp1), 2 df / (tp2 - tm1) and
 * tp1 - 0.5 (1 + cn) df; the mirrored product is the same product of
 * negated differences and its quotient the negated quotient, so the face
 * value is the same to the last bit, without a branch on the sign */
__device__ __forceinline__ double us_vlm(bool pos, double omc, double tp1, double tp2,
                                         double tm1, double tm2) {
  const double a2 = pos ? tm2 : tp2, a1 = pos ? tm1 : tp1, a0 = pos ? tp1 : tm1;
  double df = (a1 - a2) * (a0 - a1);
  df = (df > 0.0) ? 2.0 * df / (a0 - a2) : 0.0;
  return a1 + 0.5 * omc * df;
}

/*
 * Fluxes through wet surface edge e at layer k for every tracer: the face
 * value by the scheme, ULTIMATE, times u1flux3d (advect(),
 * tracers.c:1659-1728), then horizontal diffusion (tracers.c:5438-5446).
 * A far cell that is not wet at this layer is a ghost of the near cell:
 * its step-start value.  What depends on the edge only (transport,
 * Courant number, the wetness of the far cells, the diffusion
 * coefficient) is computed once, outside the loop over tracers.
 */
template <int SCH>
__device__ __forceinline__ void us_edge_fluxes(const emstr_usdev &D, int e, int k, ll ie,
                                               const double *__restrict__ tin,
                                               const double *__restrict__ tstart,
                                               double *__restrict__ fx, ll NE3, double dtu,
                                               int anydif, int ult) {
  const emstr_dev &d = D.d;
  const ll P = d.plane, N3 = d.n3;
  const int c1 = D.ec1[e], c2 = D.ec2[e];
  const ll i1 = (ll)k * P + c1, i2 = (ll)k * P + c2;
  const double fl = D.flux[ie];
  double vel = 0.0, cn = 0.0, omc = 0.0, h0 = 0.0, hm = 0.0, hp = 0.0, q3 = 0.0;
  bool wp = false, wm = false;
  unsigned sel = 0;
  ll jp = i1, jm = i2;
  if (SCH != ORDER1) {
    vel = D.u1[ie];
    h0 = D.h2e[e];
    cn = fmod(vel * dtu / h0, 1.0);
    omc = 1.0 + ((vel > 0.0) ? -cn : cn);
    const int cp = D.ecp[e], cm = D.ecm[e];
    wp = wetk(d.kb, cp, k);
    wm = wetk(d.kb, cm, k);
    if (wp) jp = (ll)k * P + cp;
    if (wm) jm = (ll)k * P + cm;
    if (SCH == QUICKEST) {
      hm = D.qhm[e];
      hp = D.qhp[e];
      q3 = D.qd3[e];
    }
    if (ult) sel = D.esel[e];
  }
  /* the far values: the neighbour's at this layer, or the ghost's */
  const double *__restrict__ bp = wp ? tin : tstart;
  const double *__restrict__ bm = wm ? tin : tstart;
  double cx = 0.0, h2 = 1.0;
  if (anydif) {
    cx = D.dzu[ie] * D.h1e[e] * 1.0 * D.kh[ie];
    h2 = D.h2e[e];
  }
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n]) continue;
    const int adv = d.advect[n], dif = anydif ? d.diffuse[n] : 0;
    const ll o = (ll)n * N3;
    const double tp1 = tin[o + i1], tm1 = tin[o + i2];
    double F = 0.0;
    if (adv) {
      if (SCH == ORDER1) {
        /* tracers.c:545-549, 1772-1775 */
        const double w6 = 0.5 * (fl + fabs(fl));
        const double w7 = 0.5 * (fl - fabs(fl));
        F = w6 * tm1 + w7 * tp1;
      } else {
        const double tp2 = bp[o + jp], tm2 = bm[o + jm];
        if (SCH == QUICKEST) {
          /* tracers.c:2098-2123 */
          const double db = 2.0 / (hm + h0);
          const double df = 2.0 / (hp + h0);
          const double face = q3 * (tm1 - tp1) + tp1;
          const double grad = tp1 - tm1;
          double curv = 0.0;
          if (cn > 0.0)
            curv = db * (hm * tp1 + h0 * tm2) - 2.0 * tm1;
          else if (cn < 0.0)
            curv = df * (h0 * tp2 + hp * tm1) - 2.0 * tp1;
          F = face - 0.5 * cn * grad - (1.0 - cn * cn) * curv / 6.0;
        } else {
          F = (vel > 0.0 || vel < 0.0) ? us_vlm(vel > 0.0, omc, tp1, tp2, tm1, tm2) : 0.0;
        }
        if (ult) {
          const double trp = (wp && (sel & 1)) ? tp2 : tp1;
          const double trm = wm ? ((sel & 2) ? tm2 : tm1) : tm2;
          F = us_ult(cn, F, tp1, tm1, trp, trm);
        }
        F *= fl;
      }
    }
    if (dif) F -= (cx * (tp1 - tm1) / h2);
    fx[n * NE3 + ie] = F;
  }
}

/* dtracer of the cell at layer k of column col (tracers.c:949-959); the
 * edge fluxes come from fxn ([nz][ne2], k_usflux) */
template <typename T>
__device__ double us_dtracer(const emstr_usdev &D, int col, int k,
                             const T *__restrict__ fxn, double dtu) {
  const emstr_dev &d = D.d;
  const int ncol = (int)d.plane;
  double dtr = 0.0;
  for (int j = 0; j < D.npe[col]; j++) {
    const int e = D.ce[j * ncol + col];
    const int other = (D.ec1[e] == col) ? D.ec2[e] : D.ec1[e];
    double F = 0.0;
    if (wetk(d.kb, other, k)) F = (double)fxn[(ll)k * D.ne2 + e];
    dtr += ((double)D.esc[j * ncol + col] * F * dtu);
  }
  return dtr;
}

/* the flux of every wet edge at every layer for every tracer (advect and
 * hor_diffuse, tracers.c:1637-1743, 5383-5449), one thread per wet edge */
template <int SCH>
__global__ void __launch_bounds__(256)
k_usflux(emstr_usdev D, const double *__restrict__ tin, const double *__restrict__ tstart,
         const ll *__restrict__ edn, int nesp, double *__restrict__ fx, double dtu,
         int anydif, int ult) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nesp) return;
  const ll ie = edn[q];
  const int k = (int)(ie / D.ne2), e = (int)(ie % D.ne2);
  us_edge_fluxes<SCH>(D, e, k, ie, tin, tstart, fx, (ll)d.nz * D.ne2, dtu, anydif, ult);
}

/* Runge-Kutta stage weights (rkweights of tracers.c:110-129 for one
 * stage): of the sub-step's starting value, of the stage's value, of its
 * tendency and volume */
struct UsRk {
  double w0, w1, w2;
  int last; /* the sub-step's last stage: tr_bounds follows */
  /* 5 stages: mode 1 stores the tendency (stage 3), mode 2 is the last
   * stage, w0 * ntr2 + w1 * ntr3 + w2 * (grad3 + grad) + w3 * ntr, volume
   * weight w4 (tracers.c:1199-1205) */
  double w3, w4;
  int mode;
};

/* runge_kutta (tracers.c:1170-1214) of one stage from the sub-step's
 * starting value ntr0 and the stage's value ntrc and tendency;
 * Ds = Hn1 = 1 at z-levels */
__device__ __forceinline__ double us_rk(const UsRk &w, double ovol, double nvol, double ntr0,
                                        double ntrc, double grad) {
  const double sf = 1.0;
  const double vol = ovol + w.w2 * (nvol - ovol);
  double tr = sf * w.w0 * ntr0 + sf * w.w1 * ntrc + w.w2 * grad;
  tr /= (1.0 * vol);
  return tr;
}

/* the last of 5 stages from the values of stages 2-4 and the tendencies
 * of stages 3 and 4 */
__device__ __forceinline__ double us_rk5(const UsRk &w, double ovol, double nvol, double ntr2,
                                         double ntr3, double grad3, double ntr4, double grad) {
  const double sf = 1.0;
  const double vol = ovol + w.w4 * (nvol - ovol);
  double tr = sf * w.w0 * ntr2 + sf * w.w1 * ntr3 + w.w2 * grad3 + sf * w.w3 * ntr4 +
              w.w2 * grad;
  tr /= (1.0 * vol);
  return tr;
}

/* the stage arrays of one tracer: the sub-step's starting values, the
 * inputs of stages 2 and 3 and the tendency of stage 3 (5 stages only) */
struct UsStage {
  const double *t0, *t2, *t3;
  double *g3;
};

/* the vertical flux through the lower face of every cell to process
 * (k <= ks of a processed column) for every advected tracer, one thread
 * per wet cell (dense indices wdn): van Leer (tracers.c:2405-2412), first
 * order (1777-1785), QUICKEST (2125-2150), ULTIMATE (1712-1720), times
 * w*dtu (1735-1738).  Tracer values at the layers below the bed are the
 * bed value (sediment ghost).  The Courant number and the layer
 * thicknesses are the cell's, computed once for every tracer */
template <int SCH>
__global__ void __launch_bounds__(256)
k_usfz(emstr_usdev D, const double *__restrict__ tin, const ll *__restrict__ wdn, int nwet,
       double *__restrict__ fz, double dtu, int ult) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nwet) return;
  const ll P = d.plane;
  const ll idx = wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int ksc = d.ks[col];
  if (ksc < 0 || k > ksc) return;
  const int kbc = d.kb[col];
  const ll N3 = d.n3;
  const double w = d.w[idx];
  const ll jm1 = (ll)max(k - 1, kbc) * P + col;
  const ll jm2 = (ll)max(k - 2, kbc) * P + col;
  const ll jp1 = (ll)min(k + 1, d.nz - 1) * P + col;
  const double wp = 0.5 * (w + fabs(w)), wm = 0.5 * (w - fabs(w));
  double cn = 0.0, omc = 0.0, w5 = 0.0, w9 = 0.0, w9m = 0.0, w9p = 0.0, db = 0.0, df = 0.0;
  if (SCH != ORDER1) {
    const double dzc = d.dz[idx];
    const double dzb = (k > kbc) ? d.dz[idx - P] : 0.0;
    w9 = 0.5 * (dzc + dzb);
    cn = fmod(w * dtu / (w9 * 1.0), 1.0);
    omc = 1.0 + ((w > 0.0) ? -cn : cn);
    if (SCH == QUICKEST) {
      /* w9 of the sediment ghost is 0 */
      w5 = dzb / (dzb + dzc);
      w9m = (k > kbc) ? 0.5 * (dzb + ((k - 1 > kbc) ? d.dz[idx - 2 * P] : 0.0)) : 0.0;
      w9p = (k + 1 < d.nz) ? 0.5 * (d.dz[idx + P] + dzc) : w9;
      db = 2.0 / (w9m + w9);
      df = 2.0 / (w9 + w9p);
    }
  }
  for (int n = 0; n < d.ntr; n++) {
    if (!(d.in_tbdy[n] && d.advect[n])) continue;
    const double *__restrict__ t = tin + n * N3;
    const double tc = t[idx], tzm1 = t[jm1];
    double F;
    if (SCH == ORDER1) {
      F = dtu * (wp * tzm1 + wm * tc);
    } else {
      const double tzm2 = t[jm2], tzp1 = t[jp1];
      if (SCH == QUICKEST) {
        const double face = w5 * (tc - tzm1) + tzm1;
        const double grad = tc - tzm1;
        double curv = 0.0;
        if (cn > 0.0)
          curv = db * (w9m * tc + w9 * tzm2) - 2.0 * tzm1;
        else if (cn < 0.0)
          curv = df * (w9 * tzp1 + w9p * tzm1) - 2.0 * tc;
        F = face - 0.5 * cn * grad - (1.0 - cn * cn) * curv / 6.0;
      } else {
        F = (w > 0.0 || w < 0.0) ? us_vlm(w > 0.0, omc, tc, tzp1, tzm1, tzm2) : 0.0;
      }
      if (ult) F = us_ult(cn, F, tc, tzm1, tzp1, tzm2);
      F = F * (w * dtu);
    }
    fz[n * N3 + idx] = F;
  }
}

/* new value of the sub-surface cell at layer k (tracers.c:1000-1016) */
template <typename T>
__device__ double us_subsurface(const emstr_usdev &D, const UsRk &w, int col, int k,
                                const UsStage &sa, const double *t, const T *fxn,
                                const T *fzn, double dtu, int adv) {
  const emstr_dev &d = D.d;
  const ll idx = (ll)k * d.plane + col;
  const double dtr = us_dtracer(D, col, k, fxn, dtu);
  double FzB = 0.0, FzT = 0.0;
  if (adv) {
    FzB = (double)fzn[idx];
    FzT = (double)fzn[idx + d.plane];
  }
  const double grad = -((dtr / d.cellarea[col] + (FzT - FzB)) / d.dz[idx]);
  if (w.mode == 2) return us_rk5(w, 1.0, 1.0, sa.t2[idx], sa.t3[idx], sa.g3[idx], t[idx], grad);
  return us_rk(w, 1.0, 1.0, sa.t0[idx], t[idx], grad);
}

#define US_MAXNPE 8

/* FFSL CLIP (clip_ffsl, transport.c:4107-4138): the range a centre's new
 * value is clamped to.  Its vertical streamline's (transport.c:2461,
 * 2480-2486): the tracer of the face's upstream centre, then of every
 * whole layer the trajectory crosses (the sediment ghost reads the bed
 * value; thicknesses and walk as in the vertical flux).  Then each edge's
 * (transport.c:4122-4127): a wet edge's is the range over the
 * least-squares stencil of its upstream centre (transport.c:2207,
 * 2309-2311; for a streamline of several segments the upstream value
 * and the stencil ranges of the cells it crosses into), a land edge's
 * is zero (its bounds are never written and
 * were allocated zero).  min and max are the reference's macros. */
__device__ __forceinline__ double mn_(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double mx_(double a, double b) { return (a > b) ? a : b; }
__device__ void us_clip_range(const emstr_usdev &D, int col, int k, int n, double dtu,
                              const double *__restrict__ tin, const double *__restrict__ bmn,
                              const double *__restrict__ bmx, double &lo, double &hi) {
  const emstr_dev &d = D.d;
  const ll P = d.plane, N3 = d.n3;
  const int kb = d.kb[col], nz = d.nz, ncol = (int)P;
  const double *t = tin + n * N3;
  const double w = d.w[(ll)k * P + col];
  auto zm1 = [&](int L) { return (L >= kb) ? L - 1 : L; };
  auto zp1 = [&](int L) { return (L < kb) ? kb : ((L < nz - 1) ? L + 1 : L); };
  auto tv = [&](int L) { return t[(ll)((L < kb) ? kb : L) * P + col]; };
  auto dzl = [&](int L) { return (L < kb) ? 0.0 : d.dz[(ll)L * P + col]; };
  lo = hi = tv((w > 0.0) ? zm1(k) : k);
  if (!(w == 0.0 || k == kb)) {
    int ci = (w > 0.0) ? zm1(k) : k, it = 0;
    double dz = dzl(ci), dist = fabs(w * dtu);
    while (dz < dist) {
      if (++it > 4 * nz + 4) break; /* reported by the vertical flux */
      const double v = tv(ci);
      lo = mn_(lo, v);
      hi = mx_(hi, v);
      dist -= dz;
      ci = (w < 0.0) ? zp1(ci) : zm1(ci);
      const double dn = dzl(ci);
      if (dn > 0) dz = dn;
    }
  }
  for (int j = 0; j < D.npe[col]; j++) {
    const int e = D.ce[j * ncol + col];
    const int other = (D.ec1[e] == col) ? D.ec2[e] : D.ec1[e];
    if (!wetk(d.kb, other, k)) {
      lo = mn_(lo, 0.0);
      hi = mx_(hi, 0.0);
    } else {
      const int up = (D.u1[(ll)k * D.ne2 + e] > 0.0) ? D.ec2[e] : D.ec1[e];
      const ll iu = n * N3 + (ll)k * P + up;
      const int r0 = D.lnx ? D.lnx[(ll)k * D.ne2 + e] : -1;
      if (r0 < 0) {
        lo = mn_(lo, bmn[iu]);
        hi = mx_(hi, bmx[iu]);
      } else {
        /* a streamline of several segments: the upstream value, then
         * the stencil range of each cell it crosses into, the last one
         * (where it ends) included */
        const double tu = t[(ll)k * P + up];
        lo = mn_(lo, tu);
        hi = mx_(hi, tu);
        for (int r = r0; r >= 0; r = D.sg_nx[r]) {
          const ll ir = n * N3 + (ll)k * P + D.sg_ci[r];
          lo = mn_(lo, bmn[ir]);
          hi = mx_(hi, bmx[ir]);
        }
      }
    }
  }
}
/* tr = max(min(mx, tr), mn) (transport.c:4137) */
__device__ __forceinline__ double us_clamp(double v, double lo, double hi) {
  return mx_(mn_(hi, v), lo);
}

template <int SCH, typename T, bool CLIP>
__global__ void __launch_bounds__(256)
k_usadv(emstr_usdev D, const double *__restrict__ t0in, const double *__restrict__ tin,
        double *__restrict__ tout, const T *__restrict__ fx,
        const T *__restrict__ fz, const ll *__restrict__ wdn, int nwet, double dtu,
        UsRk w, const double *__restrict__ t2in, const double *__restrict__ t3in,
        double *__restrict__ g3in, const double *__restrict__ bmn,
        const double *__restrict__ bmx) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nwet) return;
  const ll P = d.plane;
  const ll idx = wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int kbc = d.kb[col];
  const ll N3 = d.n3, NE3 = (ll)d.nz * D.ne2;
  const int ksc = d.ks[col];
  if (ksc < 0) { /* column not processed: unchanged */
    for (int n = 0; n < d.ntr; n++) tout[n * N3 + idx] = tin[n * N3 + idx];
    return;
  }
  if (k > ksc) return; /* above the processed surface: written below */

  if (k < ksc) {
    /* the cell's edges at this layer: surface edge, sign, wet */
    const int npe = D.npe[col], ncol = (int)P;
    /* eo: the flux's index, an in-range stand-in (edge 0 at this layer)
     * where the far cell is dry or the edge is missing, so the loads carry
     * no condition; ok selects */
    ll eo[US_MAXNPE];
    bool ok[US_MAXNPE];
    double sg[US_MAXNPE];
#pragma unroll
    for (int j = 0; j < US_MAXNPE; j++) {
      eo[j] = (ll)k * D.ne2;
      ok[j] = false;
      sg[j] = 0.0;
      if (j < npe) {
        const int e = D.ce[j * ncol + col];
        const int other = (D.ec1[e] == col) ? D.ec2[e] : D.ec1[e];
        ok[j] = wetk(d.kb, other, k);
        if (ok[j]) eo[j] += e;
        sg[j] = (double)D.esc[j * ncol + col];
      }
    }
    /* the most edges of a cell in this wave: a uniform bound on the loads */
    int npw = 0;
#pragma unroll
    for (int j = 0; j < US_MAXNPE; j++)
      if (__any(j < npe)) npw = j + 1;
    const double area = d.cellarea[col], dzc = d.dz[idx];
    for (int n = 0; n < d.ntr; n++) {
      const double *t = tin + n * N3;
      if (!d.in_tbdy[n]) {
        tout[n * N3 + idx] = t[idx];
        continue;
      }
      const T *fxn = fx + n * NE3;
      /* dtracer (tracers.c:949-959) */
      double fv[US_MAXNPE];
#pragma unroll
      for (int j = 0; j < US_MAXNPE; j++) fv[j] = (j < npw) ? (double)fxn[eo[j]] : 0.0;
      double dtr = 0.0;
#pragma unroll
      for (int j = 0; j < US_MAXNPE; j++) {
        const double a = dtr + (sg[j] * (ok[j] ? fv[j] : 0.0) * dtu);
        dtr = (j < npe) ? a : dtr;
      }
      double FzB = 0.0, FzT = 0.0;
      if (d.advect[n]) {
        FzB = (double)fz[n * N3 + idx];
        FzT = (double)fz[n * N3 + idx + P];
      }
      const double grad = -((dtr / area + (FzT - FzB)) / dzc);
      double r;
      if (w.mode == 2)
        r = us_rk5(w, 1.0, 1.0, t2in[n * N3 + idx], t3in[n * N3 + idx], g3in[n * N3 + idx],
                   t[idx], grad);
      else
        r = us_rk(w, 1.0, 1.0, t0in[n * N3 + idx], t[idx], grad);
      if (w.mode == 1) g3in[n * N3 + idx] = grad;
      if (CLIP) {
        double lo, hi;
        us_clip_range(D, col, k, n, dtu, tin, bmn, bmx, lo, hi);
        r = us_clamp(r, lo, hi);
      }
      /* tr_bounds clip (tracers.c:3328-3331), after the last stage */
      if (w.last && d.mintr[n] <= 1e-5 && r < d.mintr[n]) r = d.mintr[n];
      tout[n * N3 + idx] = r;
    }
    return;
  }

  /* ---- processed surface centre of the column (k == ksc) ---- */
  const double area = d.cellarea[col];
  const int kold = d.kold[col];
  const double subeta = d.subeta[col];
  const double osubeta = d.osubeta[col];
  const double botz = d.botz[col];
  const double *lay = d.layers;
  const int ktop = d.nz - 1;
  for (int n = 0; n < d.ntr; n++) {
    const double *t = tin + n * N3, *t0 = t0in + n * N3;
    const T *fxn = fx + n * NE3;
    double *o = tout + n * N3;
    UsStage sa = {t0, nullptr, nullptr, nullptr};
    if (w.mode) {
      sa.t2 = t2in + n * N3;
      sa.t3 = t3in + n * N3;
      sa.g3 = g3in + n * N3;
    }
    if (!d.in_tbdy[n]) {
      for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = t[(ll)kk * P + col];
      continue;
    }
    const int adv = d.advect[n];
    double trs;
    if (!adv) {
      trs = t[idx];
    } else {
      /* surf_concrk (tracers.c:3085-3260) and runge_kutta; the mass at the
       * start of the sub-step (stage 0's surf_concrk) from its buffer */
      const double FzB = (double)fz[n * N3 + idx];
      const double top = -FzB * area;
      int c = k;
      double waterbot = (c == kbc) ? botz : lay[c];
      const double dzs = subeta - waterbot;
      const double nvol = dzs * area;
      /* 5 stages, last: the masses of stages 2 and 3 in place of it */
      const double *ta = (w.mode == 2) ? sa.t2 : t0, *tb = (w.mode == 2) ? sa.t3 : t0;
      double surftr = 0.0, surftr0 = 0.0, surftr3 = 0.0, ovol = 0.0, vol, watertop;
      double grad = -top;
      while (c != kold && c < ktop) {
        watertop = lay[c + 1];
        vol = (watertop - waterbot) * area;
        ovol += vol;
        surftr += (t[(ll)c * P + col] * vol);
        surftr0 += (ta[(ll)c * P + col] * vol);
        surftr3 += (tb[(ll)c * P + col] * vol);
        grad -= us_dtracer(D, col, c, fxn, dtu);
        c = c + 1;
        waterbot = lay[c];
      }
      watertop = osubeta;
      waterbot = (c == kbc) ? botz : lay[c];
      vol = (watertop - waterbot) * area;
      ovol += vol;
      surftr += t[(ll)c * P + col] * vol;
      surftr0 += ta[(ll)c * P + col] * vol;
      surftr3 += tb[(ll)c * P + col] * vol;
      grad -= us_dtracer(D, col, c, fxn, dtu);
      if (c != ktop) {
        do {
          c = c + 1;
          grad -= us_dtracer(D, col, c, fxn, dtu);
        } while (c != ktop);
      }
      if (nvol == 0.0 || (surftr < 0.0 && dzs <= d.hmin)) surftr = 0.0;
      if (nvol == 0.0 || (surftr0 < 0.0 && dzs <= d.hmin)) surftr0 = 0.0;
      if (nvol == 0.0 || (surftr3 < 0.0 && dzs <= d.hmin)) surftr3 = 0.0;
      double r;
      if (w.mode == 2)
        r = us_rk5(w, ovol, nvol, surftr0, surftr3, sa.g3[idx], surftr, grad);
      else
        r = us_rk(w, ovol, nvol, surftr0, surftr, grad);
      if (w.mode == 1) sa.g3[idx] = grad;
      /* tracers.c:1056-1057 */
      trs = (botz > 0.0 && d.dz[idx] < d.hmin) ? t[idx] : r;
    }
    if (CLIP) {
      double lo, hi;
      us_clip_range(D, col, k, n, dtu, tin, bmn, bmx, lo, hi);
      trs = us_clamp(trs, lo, hi);
    }
    if (!w.last) {
      /* an intermediate stage: the layers above keep their values */
      o[idx] = trs;
      for (int kk = k + 1; kk <= ktop; kk++) o[(ll)kk * P + col] = t[(ll)kk * P + col];
      continue;
    }
    /* tr_bounds, thin surface layer (tracers.c:3303-3321): the old
     * surface thickness dz[co] */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * P + col] : 0.0;
    if (k > kbc && dzco < d.hmin) {
      if (trs < d.mintr[n] || trs > d.maxtr[n]) {
        trs = us_subsurface(D, w, col, k - 1, sa, t, fxn, fz + n * N3, dtu, adv);
        if (CLIP) {
          /* the layer below, as its own thread clamped it */
          double lo, hi;
          us_clip_range(D, col, k - 1, n, dtu, tin, bmn, bmx, lo, hi);
          trs = us_clamp(trs, lo, hi);
        }
      }
    }
    /* clip (tracers.c:3328-3331) */
    if (d.mintr[n] <= 1e-5 && trs < d.mintr[n]) trs = d.mintr[n];
    /* the surface value fills every layer above (tracers.c:3355-3366) */
    for (int kk = k; kk <= ktop; kk++) o[(ll)kk * P + col] = trs;
  }
}

/* minimum sub-step (tracers.c:601-683): edge velocities over the centre
 * distance, w over dz below the surface; positive doubles compare as
 * their bit patterns */
/* FFSL (tracers.c:624-631, 652-656): the one-sided Lipschitz number
 * of the centred velocities across each edge (a land edge's far side is
 * its ghost, gu/gv) and of w across the cell, scaled by trsf */
__global__ void k_usdtm(emstr_usdev D, double dtu0, unsigned long long *out, int ffsl,
                        double sff, const double *__restrict__ gu,
                        const double *__restrict__ gv) {
  const emstr_dev &d = D.d;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.z;
  if (col >= d.nx) return;
  const int ksc = d.ks[col];
  const int kbc = d.kb[col];
  if (ksc < 0 || k < kbc || k > ksc) return;
  const ll P = d.plane;
  const ll idx = (ll)k * P + col;
  const int ncol = (int)P;
  const double sf = ffsl ? sff : 0.8, minval = 1e-10;
  double dtm = dtu0, d4;
  for (int j = 0; j < D.npe[col]; j++) {
    const int e = D.ce[j * ncol + col];
    const ll ie = (ll)k * D.ne2 + e;
    if (ffsl) {
      const int c1 = D.ec1[e], c2 = D.ec2[e];
      double ua, va, ub, vb;
      if (wetk(d.kb, c1, k) && wetk(d.kb, c2, k)) {
        ua = D.uc[(ll)k * P + c1];
        va = D.vc[(ll)k * P + c1];
        ub = D.uc[(ll)k * P + c2];
        vb = D.vc[(ll)k * P + c2];
      } else {
        /* a land edge joins the wet centre (e2c[e][0]) to its ghost */
        ua = D.uc[idx];
        va = D.vc[idx];
        ub = gu[ie];
        vb = gv[ie];
      }
      d4 = (ua * D.n1x[e] + va * D.n1y[e]) - (ub * D.n1x[e] + vb * D.n1y[e]);
      d4 = (d4 < 0.0) ? -d4 : 0.0;
    } else
      d4 = D.u1[ie];
    if (fabs(d4) < minval) d4 = minval;
    const double d1 = sf * fabs(D.h2e[e] / d4);
    if (d1 < dtm) dtm = d1;
  }
  const bool surf = (k == ksc);
  const ll zp = (k + 1 < d.nz) ? idx + P : idx;
  if (ffsl) {
    d4 = surf ? 0.0 : d.w[zp] - d.w[idx];
    if (k == kbc) d4 = 0.0;
    d4 = (d4 < 0.0) ? -d4 : 0.0;
  } else
    d4 = surf ? 0.0 : 0.5 * (d.w[idx] + d.w[zp]);
  if (fabs(d4) < minval) d4 = minval;
  const double d3 = surf ? dtu0 : sf * fabs(d.dz[idx] * 1.0 / d4);
  if (d3 > 0.0 && d3 < dtm) dtm = d3;
  if (dtm < dtu0) atomicMin(out, (unsigned long long)__double_as_longlong(dtm));
}

template <int SCH>
static void us_adv_launch(const emstr_usdev *D, const double *const *ts, const double *tin,
                          double *tout, const double *tstart, const long long *edn, int nesp,
                          double *fx, double *fz, double dtu, int anydif, int ult,
                          const UsRk &w, double *g3, hipStream_t st) {
  const emstr_dev *d = &D->d;
  if (nesp)
    k_usflux<SCH><<<(nesp + 255) / 256, 256, 0, st>>>(*D, tin, tstart, edn, nesp, fx, dtu,
                                                       anydif, ult);
  if (d->nwet) {
    k_usfz<SCH><<<(d->nwet + 255) / 256, 256, 0, st>>>(*D, tin, d->wdn, d->nwet, fz, dtu, ult);
    k_usadv<SCH, double, false><<<(d->nwet + 255) / 256, 256, 0, st>>>(
        *D, ts[0], tin, tout, fx, fz, d->wdn, d->nwet, dtu, w, ts[1], ts[2], g3, nullptr,
        nullptr);
  }
}

static UsRk us_rkargs(const double *rkw, int last, int mode) {
  UsRk w;
  w.w0 = rkw[0];
  w.w1 = rkw[1];
  w.w2 = rkw[2];
  w.w3 = (mode == 2) ? rkw[3] : 0.0;
  w.w4 = (mode == 2) ? rkw[4] : 0.0;
  w.last = last;
  w.mode = mode;
  return w;
}

extern "C" int emstr_us_launch_adv(const emstr_usdev *D, const double *const *ts,
                                   const double *tin, double *tout, const double *tstart,
                                   const long long *edn, int nesp, double *fx, double *fz,
                                   double dtu, int anydif, int ult, const double *rkw, int last,
                                   int mode, double *g3, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  const UsRk w = us_rkargs(rkw, last, mode);
  switch (D->d.trasc) {
    case VANLEER:
      us_adv_launch<VANLEER>(D, ts, tin, tout, tstart, edn, nesp, fx, fz, dtu, anydif, ult, w,
                             g3, st);
      break;
    case ORDER1:
      us_adv_launch<ORDER1>(D, ts, tin, tout, tstart, edn, nesp, fx, fz, dtu, anydif, 0, w, g3,
                            st);
      break;
    case QUICKEST:
      us_adv_launch<QUICKEST>(D, ts, tin, tout, tstart, edn, nesp, fx, fz, dtu, anydif, ult, w,
                              g3, st);
      break;
    default:
      return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_us_launch_dtm(const emstr_usdev *D, double dtu0, unsigned long long *out,
                                   double sf, const double *gu, const double *gv,
                                   void *stream) {
  const emstr_dev *d = &D->d;
  dim3 b(256, 1, 1);
  dim3 g((d->nx + 255) / 256, 1, d->nz);
  k_usdtm<<<g, b, 0, (hipStream_t)stream>>>(*D, dtu0, out, D->ffsl, sf, gu, gv);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* the cell update from given edge and vertical fluxes (FFSL) */
extern "C" int emstr_us_launch_update(const emstr_usdev *D, const double *const *ts,
                                      const double *tin, double *tout, const double *fx,
                                      const double *fz, double dtu, const double *rkw,
                                      int last, int mode, double *g3, const double *bmn,
                                      const double *bmx, void *stream) {
  const emstr_dev *d = &D->d;
  const UsRk w = us_rkargs(rkw, last, mode);
  const int gw = (d->nwet + 255) / 256;
  hipStream_t st = (hipStream_t)stream;
  if (!d->nwet) return 0;
  if (bmn && bmx)
    k_usadv<VANLEER, double, true><<<gw, 256, 0, st>>>(*D, ts[0], tin, tout, fx, fz, d->wdn,
                                                       d->nwet, dtu, w, ts[1], ts[2], g3, bmn,
                                                       bmx);
  else
    k_usadv<VANLEER, double, false><<<gw, 256, 0, st>>>(*D, ts[0], tin, tout, fx, fz, d->wdn,
                                                        d->nwet, dtu, w, ts[1], ts[2], g3,
                                                        nullptr, nullptr);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* the same from the FP32 form's fluxes (FP64 state), one stage */
extern "C" int emstr_us_launch_update32(const emstr_usdev *D, const double *t0, const double *tin,
                                        double *tout, const float *fx, const float *fz,
                                        double dtu, const double *rkw, const double *bmn,
                                        const double *bmx, void *stream) {
  const emstr_dev *d = &D->d;
  const UsRk w = us_rkargs(rkw, 1, 0);
  const int gw = (d->nwet + 255) / 256;
  hipStream_t st = (hipStream_t)stream;
  if (!d->nwet) return 0;
  if (bmn && bmx)
    k_usadv<VANLEER, float, true><<<gw, 256, 0, st>>>(*D, t0, tin, tout, fx, fz, d->wdn,
                                                      d->nwet, dtu, w, nullptr, nullptr, nullptr,
                                                      bmn, bmx);
  else
    k_usadv<VANLEER, float, false><<<gw, 256, 0, st>>>(*D, t0, tin, tout, fx, fz, d->wdn,
                                                       d->nwet, dtu, w, nullptr, nullptr,
                                                       nullptr, nullptr, nullptr);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

/* FFSL CLIP: each wet centre's range over its least-squares stencil
 * (transport.c:2207, 2309-2311: the upstream value, then get_local_bounds
 * mode 2, :3824-3852, over lw->cells in their order), per tracer; a
 * lateral ghost in the stencil holds the step-start value of its wet
 * centre (tracers.c:185-200).  A column with water that is not processed
 * is reported (err 3): its cells would be clamped to bounds left from an
 * earlier call. */
__global__ void __launch_bounds__(256)
k_ffbnd(emstr_usdev D, const double *__restrict__ tin, const double *__restrict__ tst,
        double *__restrict__ bmn, double *__restrict__ bmx, int *__restrict__ err) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nwet) return;
  const ll P = d.plane, N3 = d.n3;
  const ll idx = d.wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  if (d.ks[col] < 0) atomicMax(err, 3);
  const int o0 = D.lsq_off[q], o1 = D.lsq_off[q + 1];
  const ll lay = (ll)k * P;
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n] || !d.advect[n]) continue;
    const double *t = tin + n * N3, *ts = tst + n * N3;
    double lo = t[idx], hi = lo;
    for (int o = o0; o < o1; o++) {
      const int gc = D.lsq_gc[o];
      const double v = (gc >= 0) ? t[lay + gc] : ts[lay + ~gc];
      lo = mn_(lo, v);
      hi = mx_(hi, v);
    }
    bmn[n * N3 + idx] = lo;
    bmx[n * N3 + idx] = hi;
  }
}

extern "C" int emstr_us_launch_ffsl_bounds(const emstr_usdev *D, const double *tin,
                                           const double *tstart, double *bmn, double *bmx,
                                           int *err, void *stream) {
  const emstr_dev *d = &D->d;
  if (d->nwet)
    k_ffbnd<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(*D, tin, tstart, bmn, bmx,
                                                                    err);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
