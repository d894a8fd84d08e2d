/*
 * COMPAS FFSL on the GPU, plain FP64 (model/hd-us/tracers/transport.c
 * ffsl_don :1961-2598, ffsl_init :2607-2723, get_posc :1476-1559,
 * intersect :1744-1826; tracers.c prep_ff_sl :2670-2721, ff_sl_do_vert
 * :2732-2770, get_linear_value :4682-4708, get_linear_limit :4718-4871).
 *
 * Cells are [layer][column] as in the rest of the COMPAS path; inside a
 * column a layer index stands for a centre, layer kb-1 for the sediment
 * ghost (which reads the bed value), and the top layer maps to itself
 * upwards.  Kernels, in launch order:
 *   k_ffnw     per step: the centred vertical velocity (Leonard et al.
 *              choice, or vel_center_w on the calls where ffsl_init runs)
 *   k_ffsub    per sub-step: vertical departure layers and fractions of
 *              the centres (nw, w9) and of the faces of the cells to
 *              process (w, dz), and their largest Courant number
 *   k_ffvfit   on the calls where ffsl_init runs: the velocity fit of
 *              every wet centre (delaunay_reinit / lsql_rebuild)
 *   k_ffline   per sub-step, one thread per wet edge: the streamline back
 *              from the edge midpoint through its upstream cell
 *              (get_posc / intersect), stored as offsets from the cell
 *              centre and a length; where it crosses into a wet cell it
 *              goes on from the crossing with the velocity fitted there
 *              (hd_trans_interp), each later segment a record chained
 *              from the edge
 *   k_ffmod    tr_mod: the vertical transverse tracer, no-gradient above
 *              the processed surface
 *   k_fflsq    the least-squares slopes of every centre and the vertex
 *              limiter beta
 *   k_ffedge   per wet edge: the streamline mean (trapezoid of the
 *              limited reconstruction), times u1flux3d, plus horizontal
 *              diffusion; the end value for the transverse terms
 *   k_ffmoc    tr_moc: the mean end value over the cell's wet edges, in
 *              the order the reference accumulates them
 *   k_fffz     the vertical flux of the cells to process (V_VC: van Leer
 *              on the fractional part, tr_moc over whole layers); run
 *              first without tracers to mark the tr_moc centres the
 *              fluxes read, so that k_ffedge stores end values and k_ffmoc
 *              computes tr_moc for those only
 * The cell update, surf_concrk and tr_bounds are k_usadv's.  k_fflsq
 * runs a wave per cell with a lane per tracer over tracer rows written by
 * k_ffmod; the other per-tracer kernels run one thread per cell or edge
 * over all tracers.  The per-tracer kernels are templates on the type T
 * of their intermediates (tr_mod, slopes and beta, end values, tr_moc,
 * fluxes) and of the arithmetic on tracer values: double for the plain
 * form (the reference's operations and order), float for the FP32 form,
 * whose tracer state stays FP64 and whose maps, streamlines and geometry
 * are the plain form's (FP64, the same kernels).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emstr_us_kernels.h"

typedef long long ll;

#define FF_TF 0.55     /* transport.c:2014 */
#define FF_SMALL 1e-10 /* transport.c:24 */

__device__ __forceinline__ double fmx(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double fmn(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ float fmx(float a, float b) { return (a > b) ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return (a < b) ? a : b; }

__device__ __forceinline__ bool ffwet(const int *__restrict__ kb, int col, int k) {
  return col >= 0 && k >= kb[col];
}
/* vertical maps by layer: the bed maps down to the sediment ghost (kb-1),
 * which maps to itself; the sediment maps up to the bed */
__device__ __forceinline__ int ffzm1(int L, int kb) { return (L >= kb) ? L - 1 : L; }
__device__ __forceinline__ int ffzp1(int L, int kb, int nz) {
  return (L < kb) ? kb : ((L < nz - 1) ? L + 1 : L);
}
/* layer thickness; 0 for the sediment ghost */
__device__ __forceinline__ double ffdz(const emstr_dev &d, int col, int L, int kb) {
  return (L < kb) ? 0.0 : d.dz[(ll)L * d.plane + col];
}
/* w9 = 0.5 (dz + dz below) of a centre (tracers.c:571-575); 0 for the
 * sediment ghost, whose entry the reference never writes */
__device__ __forceinline__ double ffw9(const emstr_dev &d, int col, int L, int kb) {
  return (L < kb) ? 0.0 : 0.5 * (ffdz(d, col, L, kb) + ffdz(d, col, L - 1, kb));
}

/*-------------------------------------------------------------------*/
/* per step: centred vertical velocity                               */
/*-------------------------------------------------------------------*/
__global__ void __launch_bounds__(256) k_ffnw(emstr_usdev D, int init, double *__restrict__ nw) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nwet) return;
  const ll P = d.plane;
  const ll idx = d.wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int kb = d.kb[col];
  const double *w = d.w;
  double v;
  if (!init) {
    /* tracers.c:579-586 */
    const double wc = w[idx], wp = w[(k < d.nz - 1) ? idx + P : idx];
    v = 0.0;
    if ((wc >= 0.0) && (wp > 0.0))
      v = wc;
    else if ((wc < 0.0) && (wp <= 0.0))
      v = wp;
  } else {
    /* vel_center_w (transport.c:6098-6141) from the old surface down */
    const int ko = d.kold[col];
    if (k > ko) {
      v = 0.0;
    } else {
      if (k == ko)
        v = 0.5 * (w[idx] + D.wtop[col]);
      else
        v = 0.5 * (w[idx + P] + ((k == kb) ? D.wbot[col] : w[idx]));
      if (fabs(v) < FF_SMALL) v = FF_SMALL;
    }
  }
  nw[idx] = v;
}

/*-------------------------------------------------------------------*/
/* per sub-step: vertical maps (prep_ff_sl) and Courant numbers      */
/*-------------------------------------------------------------------*/
/* ff_sl_do_vert for one centre or face: h is w9 (centres) or dz (faces) */
template <bool W9>
__device__ __forceinline__ void ffdovert(const emstr_dev &d, int col, int k, int kb,
                                         double vel, double dt, int &cl, double &crf) {
  crf = 0.0;
  if (vel == 0.0) return;
  const int nz = d.nz;
#define FFH(L) (W9 ? ffw9(d, col, (L), kb) : ffdz(d, col, (L), kb))
  const double lt = fabs(vel) * dt;
  double lrem = lt;
  int cp = cl, cpb = cl;
  double dz = fabs((FFH(cp) > 0) ? FFH(cp) : FFH(k));
  while (dz < lrem) {
    lrem -= dz;
    cp = (vel < 0.0) ? ffzp1(cp, kb, nz) : ffzm1(cp, kb);
    if (cpb == cp) {
      lrem = fmod(lrem, dz);
      break;
    }
    if (FFH(cp) > 0) dz = fabs(FFH(cp));
    cpb = cp;
  }
#undef FFH
  crf = fmx(fmn(lrem / dz, 1.0), 0.0);
  cl = cp;
}

__global__ void __launch_bounds__(256)
k_ffsub(emstr_usdev D, double dtu, const double *__restrict__ nw, int *__restrict__ clc,
        double *__restrict__ crc, int *__restrict__ clf, double *__restrict__ crf,
        double *__restrict__ cour) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  /* the streamline records of this sub-step start from zero (k_ffline
   * runs after this kernel on the stream) */
  if (q == 0 && D.sg_cnt) *D.sg_cnt = 0;
  if (q >= d.nwet) return;
  const ll P = d.plane;
  const ll idx = d.wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int kb = d.kb[col], ks = d.ks[col], nz = d.nz;
  /* centres (tracers.c:2697-2706) */
  {
    const double v = nw[idx];
    int cl = (v < 0.0) ? ffzp1(k, kb, nz) : k;
    double r;
    ffdovert<true>(d, col, k, kb, v, dtu, cl, r);
    clc[idx] = cl;
    crc[idx] = r;
  }
  /* faces of the cells to process (tracers.c:2711-2720) */
  if (ks >= 0 && k <= ks) {
    const double v = d.w[idx];
    int cl = (v > 0.0) ? ffzm1(k, kb) : k;
    double r;
    ffdovert<false>(d, col, k, kb, v, dtu, cl, r);
    clf[idx] = cl;
    crf[idx] = r;
    /* the largest Courant number (tracers.c:751-759, transport.c:2058-2068) */
    const double w8 = fmod(v * dtu / (ffw9(d, col, k, kb) * 1.0), 1.0);
    double cr = fabs(w8);
    const int ncol = (int)P;
    for (int j = 0; j < D.npe[col]; j++) {
      const int e = D.ce[j * ncol + col];
      const double w6 = fmod(D.u1[(ll)k * D.ne2 + e] * dtu / D.h2e[e], 1.0);
      cr = fmx(cr, fabs(w6));
    }
    cour[idx] = cr;
  }
}

/*-------------------------------------------------------------------*/
/* per sub-step: streamlines                                         */
/*-------------------------------------------------------------------*/
/* intersect (transport.c:1744-1826) with edge es of the surface layer */
__device__ __forceinline__ int ffcross(const emstr_usdev &D, int es, double xn, double yn,
                                       double xs, double ys, double &xi, double &yi) {
  const double v1x = D.ev[4 * es + 0], v1y = D.ev[4 * es + 1];
  const double v2x = D.ev[4 * es + 2], v2y = D.ev[4 * es + 3];
  double s1, s2;
  const double dx = fabs(v1x - v2x);
  if (dx) {
    s1 = (v2y - v1y) / (v2x - v1x);
    if (xs == xn)
      xi = xn;
    else {
      s2 = (ys - yn) / (xs - xn);
      xi = (s2 * xn - s1 * v1x + v1y - yn) / (s2 - s1);
    }
    yi = s1 * (xi - v1x) + v1y;
  } else {
    s1 = (v2x - v1x) / (v2y - v1y);
    if (ys == yn)
      yi = yn;
    else {
      s2 = (xs - xn) / (ys - yn);
      yi = (s1 * v1y - s2 * yn + xn - v1x) / (s1 - s2);
    }
    xi = s1 * (yi - v1y) + v1x;
  }
  if (xi >= fmn(v1x, v2x) && xi <= fmx(v1x, v2x) && yi >= fmn(v1y, v2y) &&
      yi <= fmx(v1y, v2y)) {
    if (xi >= fmn(xn, xs) && xi <= fmx(xn, xs) && yi >= fmn(yn, ys) && yi <= fmx(yn, ys))
      return 1;
  }
  return 0;
}

/* the velocity fit of every wet centre when ffsl_init runs
 * (delaunay_reinit of u and v, transport.c:4728-4929, lsql_rebuild,
 * lib/interp/lsq_lin/lsql.c:262-287): the slopes sum v_j bx_j, v_j by_j
 * in the table's order from 0, the value the centre's own; a ghost's v_j
 * is the edge velocity rotated to east / north, which ffsl_init stores
 * there (transport.c:2652-2666) */
__global__ void __launch_bounds__(256) k_ffvfit(emstr_usdev D) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nwet) return;
  const ll N3 = d.n3, idx = d.wdn[q];
  double w1u = 0.0, w2u = 0.0, w1v = 0.0, w2v = 0.0;
  for (int o = D.vls_off[q]; o < D.vls_off[q + 1]; o++) {
    const ll ent = D.vls_ent[o];
    double zu, zv;
    if (ent >= 0) {
      zu = D.uc[ent];
      zv = D.vc[ent];
    } else {
      const ll ie = ~ent;
      const int e = (int)(ie % D.ne2);
      zu = D.u1[ie] * D.n1x[e] + D.u2[ie] * D.t2x[e];
      zv = D.u1[ie] * D.n1y[e] + D.u2[ie] * D.t2y[e];
    }
    w1u += zu * D.vls_bx[o];
    w2u += zu * D.vls_by[o];
    w1v += zv * D.vls_bx[o];
    w2v += zv * D.vls_by[o];
  }
  D.vfit[idx] = D.uc[idx];
  D.vfit[N3 + idx] = w1u;
  D.vfit[2 * N3 + idx] = w2u;
  D.vfit[3 * N3 + idx] = D.vc[idx];
  D.vfit[4 * N3 + idx] = w1v;
  D.vfit[5 * N3 + idx] = w2v;
}

/* ffsl_don's streamline loop (transport.c:2192-2333) for one wet edge:
 * get_posc (transport.c:1476-1559, metres) from the edge midpoint through
 * the upstream centre; where the segment crosses into a wet cell, the
 * next one starts at the crossing, skips the edge crossed, runs for the
 * time left with the velocity fitted there, and so on (at most 100
 * segments, as the reference).  The velocity of a later segment is
 * hd_trans_interp's (transport.c:5171-5296): its depth is the centre's
 * own (pz = cellz[c]), so the vertical blend of the fits of the cell and
 * of the cell above gives the cell's fit exactly; that fit at (x, y) is
 * value + (x slope (x - xc) + y slope (y - yc)) (lsql_interpolate_point
 * with beta 1).  err: 1 a later segment before ffsl_init has run on this
 * context, 5 more than 100 segments, 6 the records full. */
__global__ void __launch_bounds__(256)
k_ffline(emstr_usdev D, double dtu, const ll *__restrict__ edn, int nesp,
         int *__restrict__ lci, double *__restrict__ ld, int *__restrict__ err) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nesp) return;
  const ll ie = edn[q];
  const ll NE3 = (ll)d.nz * D.ne2, N3 = d.n3;
  const int k = (int)(ie / D.ne2), e = (int)(ie % D.ne2);
  const int ncol = (int)d.plane;
  const double u1 = D.u1[ie], u2 = D.u2[ie];
  double u = u1 * D.n1x[e] + u2 * D.t2x[e];
  double v = u1 * D.n1y[e] + u2 * D.t2y[e];
  /* the upstream centre (transport.c:2162-2173) */
  int ci = (u1 > 0.0) ? D.ec2[e] : D.ec1[e];
  double px = D.ux[e], py = D.uy[e], tleft = dtu;
  int ei = e, prev = -1;
  lci[ie] = ci;
  if (D.lnx) D.lnx[ie] = -1;
  for (int nseg = 1;; nseg++) {
    if (nseg > 1) {
      if (!D.vfit_ok) {
        atomicMax(err, 1);
        break;
      }
      if (nseg > 100) {
        atomicMax(err, 5);
        break;
      }
      const ll ic = (ll)k * ncol + ci;
      const double xp = px - D.cx[ci], yp = py - D.cy[ci];
      u = D.vfit[ic] + (D.vfit[N3 + ic] * xp + D.vfit[2 * N3 + ic] * yp);
      v = D.vfit[3 * N3 + ic] + (D.vfit[4 * N3 + ic] * xp + D.vfit[5 * N3 + ic] * yp);
    }
    /* the direction's cosine and sine as u / s and v / s, the exact
     * values of cos(atan2(v, u)) and sin(atan2(v, u)) (correctly rounded
     * division and square root, so the end point does not depend on a
     * math library); at s = 0 cosine 1 and sine 0 as atan2(0, 0) = 0
     * gives */
    const double s = sqrt(u * u + v * v);
    const double dist = s * tleft * 1.0;
    const double costh = (s > 0.0) ? u / s : 1.0, sinth = (s > 0.0) ? v / s : 0.0;
    const double slon = px - dist * costh, slat = py - dist * sinth;
    double ex = px, ey = py, od = 0.0;
    int found = 0, cn = ci, ecross = ei;
    if (dist != 0.0) {
      double xi = 0.0, yi = 0.0;
      for (int j = 0; j < D.npe[ci]; j++) {
        const int ej = D.ce[j * ncol + ci];
        if (ej != ei && (found = ffcross(D, ej, px, py, slon, slat, xi, yi))) {
          cn = (D.ec1[ej] == ci) ? D.ec2[ej] : D.ec1[ej];
          ecross = ej;
          break;
        }
      }
      if (found) {
        ex = xi;
        ey = yi;
      } else {
        ex = slon;
        ey = slat;
      }
      const double dx = ex - px, dy = ey - py;
      od = sqrt(dx * dx + dy * dy) / 1.0;
    }
    const double cxc = D.cx[ci], cyc = D.cy[ci];
    if (nseg == 1) {
      ld[0 * NE3 + ie] = cxc - px;
      ld[1 * NE3 + ie] = cyc - py;
      ld[2 * NE3 + ie] = cxc - ex;
      ld[3 * NE3 + ie] = cyc - ey;
      ld[4 * NE3 + ie] = od;
    } else {
      const int r = atomicAdd(D.sg_cnt, 1);
      if (r >= D.sg_cap) {
        atomicMax(err, 6);
        break;
      }
      D.sg_ci[r] = ci;
      D.sg_nx[r] = -1;
      double *sd = D.sg_d + (ll)r * 5;
      sd[0] = cxc - px;
      sd[1] = cyc - py;
      sd[2] = cxc - ex;
      sd[3] = cyc - ey;
      sd[4] = od;
      if (prev < 0)
        D.lnx[ie] = r;
      else
        D.sg_nx[prev] = r;
      prev = r;
    }
    /* on into a wet cell: the time the segment took (transport.c:
     * 2286-2290) */
    if (!(found && ffwet(d.kb, cn, k))) break;
    tleft -= od / sqrt(u * u + v * v);
    ci = cn;
    ei = ecross;
    px = ex;
    py = ey;
  }
}

/*-------------------------------------------------------------------*/
/* per tracer                                                        */
/*-------------------------------------------------------------------*/
/* tr_mod (transport.c:2082-2106 with set_tr_nograd), written as rows of
 * the tracers of one wet centre, tq[q * tqs + n]: a block takes FM_C
 * consecutive wet centres, reads each tracer along the centres and writes
 * the rows through LDS.  Tracers that are not advected hold 0. */
#define FM_C 32 /* wet centres per block */
template <typename T>
__global__ void __launch_bounds__(256)
k_ffmod(emstr_usdev D, const double *__restrict__ tin, const double *__restrict__ nw,
        const int *__restrict__ clc, const double *__restrict__ crc, T *__restrict__ tq) {
  __shared__ T S[FM_C][65];
  const emstr_dev &d = D.d;
  const int t = threadIdx.x, c = t & (FM_C - 1), g = t / FM_C;
  const int q0 = blockIdx.x * FM_C, q = q0 + c;
  const int nc = min(FM_C, d.nwet - q0);
  const ll P = d.plane, N3 = d.n3, ts = D.tqs;
  ll ik = 0, iz = 0, ic = 0;
  T crf = 0;
  if (q < d.nwet) {
    const ll idx = d.wdn[q];
    const int col = (int)(idx % P), k = (int)(idx / P);
    const int kb = d.kb[col], ks = d.ks[col];
    const int kk = (ks >= 0 && k > ks) ? ks : k; /* above the surface: its value */
    ik = (ll)kk * P + col;
    const double vel = nw[ik];
    crf = (T)crc[ik];
    const int cl = clc[ik];
    const int zp = (vel < 0.0) ? cl : ffzm1(cl, kb);
    const int ci = (vel < 0.0) ? ffzm1(zp, kb) : cl;
    iz = (ll)((zp < kb) ? kb : zp) * P + col;
    ic = (ll)((ci < kb) ? kb : ci) * P + col;
  }
  for (int n0 = 0; n0 < d.ntr; n0 += 64) {
    const int tn = min(64, d.ntr - n0);
    if (q < d.nwet)
#pragma unroll 2
      for (int j = g; j < tn; j += 256 / FM_C) {
        const int n = n0 + j;
        T r = 0;
        if (d.in_tbdy[n] && d.advect[n]) {
          const double *tt = tin + n * N3;
          const T tc = (T)tt[ic], tk = (T)tt[ik], tz = (T)tt[iz]; /* iz in range */
          r = (T)FF_TF * tc + (T)(1.0 - FF_TF) * tk;
          const T rz = r + (T)FF_TF * (crf * (tz - tc));
          r = (crf > 0) ? rz : r;
        }
        S[c][j] = r;
      }
    __syncthreads();
    for (int f = t; f < nc * 64; f += 256) {
      const int cc = f >> 6, j = f & 63;
      if (j < tn) tq[(ll)(q0 + cc) * ts + n0 + j] = S[cc][j];
    }
    __syncthreads();
  }
}

/* block b's position when blocks are dealt to the 8 XCDs round robin and
 * each XCD should take one contiguous range of positions (its L2 then
 * holds the rows its neighbouring centres share) */
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int full = nb / 8, rem = nb % 8, x = b % 8;
  return x * full + min(x, rem) + b / 8;
}

/* lane l's value of a wave-uniform broadcast */
__device__ __forceinline__ int rdl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdl(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdl(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* get_linear_limit, df = 2: slopes w1, w2 from the stencil, beta from the
 * vertex ranges (lw->beta is 1 while the vertex values are taken).  One
 * wave per wet centre with a lane per tracer: the stencil's slots and
 * weights and the list of distinct centres are loaded an entry a lane and
 * broadcast, the vertex offsets and ranges come through the scalar unit,
 * the distinct centres' tr_mod rows (at most EMSTR_FF_ROWS) are loaded at
 * once, and the stencil and the vertex ranges take them by slot.  A block
 * takes LS_C consecutive centres and writes w1, w2, beta and w0 = tr_mod
 * as [4][tracer][cell] planes through LDS. */
#define LS_C 8 /* wet centres per block */
template <typename T>
__global__ void __launch_bounds__(256)
k_fflsq(emstr_usdev D, const T *__restrict__ tq, T *__restrict__ lq) {
  __shared__ T S[4][64][LS_C + 1];
  const emstr_dev &d = D.d;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q0 = xcd_block(blockIdx.x, gridDim.x) * LS_C;
  const int nc = min(LS_C, d.nwet - q0);
  const ll P = d.plane, N3 = d.n3, NT3 = (ll)d.ntr * N3, ts = D.tqs;
  /* the block's centres, a centre a lane */
  int m_col = 0, m_npe = 0, m_o0 = 0, m_o1 = 0, m_r0 = 0, m_r1 = 0;
  if (lane < nc) {
    const int q = q0 + lane;
    m_col = (int)(d.wdn[q] % P);
    m_o0 = D.lsq_off[q];
    m_o1 = D.lsq_off[q + 1];
    m_r0 = D.lr_off[q];
    m_r1 = D.lr_off[q + 1];
    m_npe = D.npe[m_col];
  }
  /* the write-out: thread (cell ccw, tracer jw + 32 h) of plane pl */
  const int ccw = threadIdx.x & (LS_C - 1), jw = threadIdx.x / LS_C;
  const ll iw = (ccw < nc) ? d.wdn[q0 + ccw] : 0;
  for (int n0 = 0; n0 < d.ntr; n0 += 64) {
    const int n = n0 + lane, nl = (n < d.ntr) ? n : 0;
    for (int cc = w; cc < nc; cc += 4) {
      const int q = q0 + cc;
      const int col = rdl(m_col, cc), npe = rdl(m_npe, cc);
      const int o0 = rdl(m_o0, cc), ns = rdl(m_o1, cc) - o0;
      const int r0 = rdl(m_r0, cc), nr = rdl(m_r1, cc) - r0;
      /* stencil and row list, an entry a lane (lanes past the end repeat
       * entry 0: no divergent loads; only entries are broadcast) */
      const int ls = (lane < ns) ? lane : 0, lr = (lane < nr) ? lane : 0;
      const int e_s = D.lsq_s[o0 + ls];
      const T e_b1 = (T)D.lsq_B[2 * (ll)(o0 + ls)], e_b2 = (T)D.lsq_B[2 * (ll)(o0 + ls) + 1];
      const int e_rq = D.lr_q[r0 + lr];
      /* vertex offsets and ranges (padded to EMSTR_FF_NPE) */
      const double *xy = D.vxy + (ll)col * (2 * EMSTR_FF_NPE);
      const unsigned *vw = D.vsl + (ll)q * EMSTR_FF_NPE;
      T vx[EMSTR_FF_NPE], vy[EMSTR_FF_NPE];
      unsigned wd[EMSTR_FF_NPE];
#pragma unroll
      for (int j = 0; j < EMSTR_FF_NPE; j++) {
        vx[j] = (T)xy[2 * j];
        vy[j] = (T)xy[2 * j + 1];
        wd[j] = vw[j];
      }
      /* rows of the distinct centres; lanes past the last tracer read
       * tracer 0 and are not written out */
      T t[EMSTR_FF_ROWS];
#pragma unroll
      for (int s = 0; s < EMSTR_FF_ROWS; s++) {
        const T *row = tq + (ll)rdl(e_rq, s) * ts;
        t[s] = (s < nr) ? row[nl] : (T)0;
      }
      const T w0 = t[0];
      T a1 = 0, a2 = 0, beta = 1;
      for (int i = 0; i < ns; i++) {
        const T tv = t[rdl(e_s, i)];
        a1 += (tv * rdl(e_b1, i));
        a2 += (tv * rdl(e_b2, i));
      }
      /* vertices in order; the wet centres around one in column order */
#pragma unroll
      for (int j = 0; j < EMSTR_FF_NPE; j++) {
        if (j >= npe) break;
        unsigned word = wd[j];
        T vmin = (T)__builtin_huge_val(), vmax = -(T)__builtin_huge_val();
        for (int i = 0; i < 4; i++, word >>= 8) {
          const int sl = (int)(word & 0xffu);
          if (sl == 0xff) break;
          const T tv = t[sl];
          vmin = fmn(vmin, tv);
          vmax = fmx(vmax, tv);
        }
        const T val = w0 + (T)1 * (a1 * vx[j] + a2 * vy[j]);
        const T dd = val - w0;
        if (dd != (T)0 && val > vmax) beta = fmx(fmn((vmax - w0) / dd, beta), (T)0);
        if (dd != (T)0 && val < vmin) beta = fmx(fmn((vmin - w0) / dd, beta), (T)0);
      }
      S[0][lane][cc] = a1;
      S[1][lane][cc] = a2;
      S[2][lane][cc] = fmn(fmx(beta, (T)0), (T)1);
      S[3][lane][cc] = w0;
    }
    __syncthreads();
    const int tn = min(64, d.ntr - n0);
    if (ccw < nc)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int j = jw + 32 * h, nn = n0 + j;
        if (j < tn && d.in_tbdy[nn] && d.advect[nn]) {
          T *o = lq + nn * N3 + iw;
#pragma unroll
          for (int pl = 0; pl < 4; pl++) o[pl * NT3] = S[pl][j][ccw];
        }
      }
    __syncthreads();
  }
}

/* edge fluxes (transport.c:2192-2386) and horizontal diffusion
 * (tracers.c:5438-5446) */
template <typename T>
__global__ void __launch_bounds__(256)
k_ffedge(emstr_usdev D, const double *__restrict__ tin, const ll *__restrict__ edn, int nesp,
         const int *__restrict__ lci, const double *__restrict__ ld,
         const T *__restrict__ lq, T *__restrict__ fx,
         T *__restrict__ te, int anydif, const int *__restrict__ need, int stamp) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nesp) return;
  const ll ie = edn[q];
  const ll P = d.plane, N3 = d.n3, NT3 = (ll)d.ntr * N3, NE3 = (ll)d.nz * D.ne2;
  const int k = (int)(ie / D.ne2), e = (int)(ie % D.ne2);
  const int ci = lci[ie];
  const ll cidx = (ll)k * P + ci;
  const T dxs = (T)ld[ie], dys = (T)ld[NE3 + ie], dxe = (T)ld[2 * NE3 + ie];
  const T dye = (T)ld[3 * NE3 + ie], dist = (T)ld[4 * NE3 + ie];
  const T f = (T)D.flux[ie];
  const int nx0 = D.lnx ? D.lnx[ie] : -1;
  const ll i1 = (ll)k * P + D.ec1[e], i2 = (ll)k * P + D.ec2[e];
  /* the end value is read only by the tr_moc of the centres either side */
  const bool wte = (D.ec1[e] >= 0 && need[i1] == stamp) || (D.ec2[e] >= 0 && need[i2] == stamp);
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n]) continue;
    T F = 0;
    if (d.advect[n]) {
      const T w0 = lq[3 * NT3 + n * N3 + cidx];
      const T w1 = lq[n * N3 + cidx], w2 = lq[NT3 + n * N3 + cidx];
      const T b = lq[2 * NT3 + n * N3 + cidx];
      T trs = w0 + b * (w1 * dxs + w2 * dys);
      T tre = w0 + b * (w1 * dxe + w2 * dye);
      T dint = (T)0 + dist;
      T ntr = (T)0 + (T)0.5 * dist * (trs + tre);
      /* later segments, each from its own centre's reconstruction */
      for (int r = nx0; r >= 0; r = D.sg_nx[r]) {
        const ll ri = (ll)k * P + D.sg_ci[r];
        const double *sd = D.sg_d + (ll)r * 5;
        const T r0 = lq[3 * NT3 + n * N3 + ri];
        const T r1 = lq[n * N3 + ri], r2 = lq[NT3 + n * N3 + ri];
        const T rb = lq[2 * NT3 + n * N3 + ri];
        const T rd = (T)sd[4];
        trs = r0 + rb * (r1 * (T)sd[0] + r2 * (T)sd[1]);
        tre = r0 + rb * (r1 * (T)sd[2] + r2 * (T)sd[3]);
        dint += rd;
        ntr += (T)0.5 * rd * (trs + tre);
      }
      F = (dint != (T)0) ? ntr / dint : (T)0;
      F *= f;
      if (wte) te[n * NE3 + ie] = tre;
    }
    if (anydif && d.diffuse[n]) {
      const double *t = tin + n * N3;
      const double csx = D.dzu[ie] * D.h1e[e] * 1.0;
      F -= (T)(csx * D.kh[ie] * (t[i1] - t[i2]) / D.h2e[e]);
    }
    fx[n * NE3 + ie] = F;
  }
}

/* tr_moc (transport.c:2323-2330, 2412-2420 with set_tr_nograd), of the
 * centres the vertical fluxes read (need[idx] == stamp; all of them at or
 * below the surface, where the no-gradient copy is the centre itself) */
template <typename T>
__global__ void __launch_bounds__(256)
k_ffmoc(emstr_usdev D, const double *__restrict__ tin, const T *__restrict__ te,
        T *__restrict__ moc, const int *__restrict__ need, int stamp) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nwet) return;
  const ll P = d.plane, N3 = d.n3, NE3 = (ll)d.nz * D.ne2;
  const ll idx = d.wdn[q];
  if (need[idx] != stamp) return;
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int ks = d.ks[col], ncol = (int)P;
  const int kk = (ks >= 0 && k > ks) ? ks : k;
  const ll ik = (ll)kk * P + col;
  ll el[8];
  int m = 0;
  for (int s = 0; s < D.npe[col] && s < 8; s++) {
    const int j = D.eord[s * ncol + col];
    const int e = D.ce[j * ncol + col];
    const int other = (D.ec1[e] == col) ? D.ec2[e] : D.ec1[e];
    if (ffwet(d.kb, other, kk)) el[m++] = (ll)kk * D.ne2 + e;
  }
  for (int n = 0; n < d.ntr; n++) {
    if (!d.in_tbdy[n] || !d.advect[n]) continue;
    const T *tn = te + n * NE3;
    T r = 0;
    for (int i = 0; i < m; i++) r += tn[el[i]];
    if (m) r /= (T)m;
    r = (T)FF_TF * r + (T)(1.0 - FF_TF) * (T)tin[n * N3 + ik];
    moc[n * N3 + idx] = r;
  }
}

template <typename T>
__device__ __forceinline__ T ffmark(int *need, ll i, int stamp) {
  need[i] = stamp;
  return (T)0;
}

/* vertical fluxes (transport.c:2441-2490, ffsl_van_leer :8216-8260).
 * MARK: one pass that follows the same stencil and whole-layer walk
 * without tracer values and sets need[] = stamp at every tr_moc centre
 * the fluxes read (tr_moc is read only there) */
template <bool MARK, typename T>
__global__ void __launch_bounds__(256)
k_fffz(emstr_usdev D, double dtu, const double *__restrict__ tin, const T *__restrict__ moc,
       const int *__restrict__ clf, const double *__restrict__ crf,
       const double *__restrict__ cour, T *__restrict__ fz, int *__restrict__ err,
       int *__restrict__ need, int stamp) {
  const emstr_dev &d = D.d;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.nwet) return;
  const ll P = d.plane, N3 = d.n3;
  const ll idx = d.wdn[q];
  const int col = (int)(idx % P), k = (int)(idx / P);
  const int kb = d.kb[col], ks = d.ks[col], nz = d.nz;
  if (ks < 0 || k > ks) return;
  const double w = d.w[idx];
  const int cl = clf[idx];
  const double cr = crf[idx];
  const bool usemoc = !(cour[idx] < 1.0);
  /* the van Leer stencil (fixed per cell) */
  int cp = cl;
  if (w > 0.0) cp = ffzp1(cp, kb, nz);
  const int bm1 = ffzm1(cp, kb), bm2 = ffzm1(bm1, kb), fp = ffzp1(cp, kb, nz);
  const double fr = (w == 0.0 || k == kb) ? 0.0 : cr * ffdz(d, col, cl, kb);
  const T crT = (T)cr, frT = (T)fr;
  for (int n = 0; n < (MARK ? 1 : d.ntr); n++) {
    if (!MARK && (!d.in_tbdy[n] || !d.advect[n])) continue;
    const double *t = tin + n * N3;
    const T *mc = moc + n * N3;
#define FFT(L) (MARK ? (T)0 : (T)t[(ll)(((L) < kb) ? kb : (L)) * P + col])
#define FFL(L) ((ll)(((L) < kb) ? kb : (((L) > ks) ? ks : (L))) * P + col)
#define FFM(L) (MARK ? ffmark<T>(need, FFL(L), stamp) : mc[FFL(L)])
#define FFV(L) (usemoc ? FFM(L) : FFT(L))
    /* the stencil mirrored for w < 0: (bm2, bm1, cp) for w > 0, (fp, cp,
     * bm1) for w < 0; the reference's cp - 0.5 (1 - cr) df of
     * df = (cp - bm1) (fp - cp), 2 df / (fp - bm1) is the same value with
     * the negated quotient, to the last bit, and the same three cells are
     * read (and marked) */
    T F = 0;
    if (cr != 0.0 && (w > 0.0 || w < 0.0)) {
      const bool pos = w > 0.0;
      const int l2 = pos ? bm2 : fp, l1 = pos ? bm1 : cp, l0 = pos ? cp : bm1;
      const T a2 = FFV(l2), a1 = FFV(l1), a0 = FFV(l0);
      T df = (a1 - a2) * (a0 - a1);
      df = (df > (T)0) ? (T)2 * df / (a0 - a2) : (T)0;
      F = a1 + (T)0.5 * ((T)1 - crT) * df;
    }
    if (w == 0.0 || k == kb) {
      F = 0;
    } else {
      F *= frT;
      /* whole layers of the trajectory (transport.c:2462-2486) */
      int ci = (w > 0.0) ? ffzm1(k, kb) : k;
      double dz = ffdz(d, col, ci, kb);
      double dist = fabs(w * dtu);
      int it = 0;
      while (dz < dist) {
        if (++it > 4 * nz + 4) {
          atomicMax(err, 2);
          break;
        }
        F += ((T)dz * FFM(ci));
        dist -= dz;
        ci = (w < 0.0) ? ffzp1(ci, kb, nz) : ffzm1(ci, kb);
        const double dn = ffdz(d, col, ci, kb);
        if (dn > 0) dz = dn;
      }
      if (w < 0) F *= (T)-1;
    }
#undef FFT
#undef FFL
#undef FFM
#undef FFV
    if (!MARK) fz[n * N3 + idx] = F;
  }
}

/*-------------------------------------------------------------------*/
/* launchers                                                         */
/*-------------------------------------------------------------------*/
extern "C" int emstr_us_launch_ffsl_step(const emstr_usdev *D, int init, double *nw,
                                         void *stream) {
  const emstr_dev *d = &D->d;
  if (d->nwet)
    k_ffnw<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(*D, init, nw);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_us_launch_ffsl_vfit(const emstr_usdev *D, void *stream) {
  const emstr_dev *d = &D->d;
  if (d->nwet && D->vls_off)
    k_ffvfit<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(*D);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_us_launch_ffsl_sub(const emstr_usdev *D, double dtu, const double *nw,
                                        int *clc, double *crc, int *clf, double *crf,
                                        double *cour, const long long *edn, int nesp, int *lci,
                                        double *ld, int *err, void *stream) {
  const emstr_dev *d = &D->d;
  hipStream_t st = (hipStream_t)stream;
  if (d->nwet)
    k_ffsub<<<(d->nwet + 255) / 256, 256, 0, st>>>(*D, dtu, nw, clc, crc, clf, crf, cour);
  if (nesp) k_ffline<<<(nesp + 255) / 256, 256, 0, st>>>(*D, dtu, edn, nesp, lci, ld, err);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename T>
static int ffsl_flux(const emstr_usdev *D, double dtu, const double *tin, const double *nw,
                     const int *clc, const double *crc, const int *clf, const double *crf,
                     const double *cour, const long long *edn, int nesp, const int *lci,
                     const double *ld, T *tm, T *lq, T *te, T *moc, T *fx, T *fz, int anydif,
                     int *need, int stamp, int *err, hipStream_t st) {
  const emstr_dev *d = &D->d;
  const int gw = (d->nwet + 255) / 256, ge = (nesp + 255) / 256;
  if (d->nwet) {
    k_fffz<true, T><<<gw, 256, 0, st>>>(*D, dtu, tin, moc, clf, crf, cour, fz, err, need, stamp);
    k_ffmod<T><<<(d->nwet + FM_C - 1) / FM_C, 256, 0, st>>>(*D, tin, nw, clc, crc, tm);
    k_fflsq<T><<<(d->nwet + LS_C - 1) / LS_C, 256, 0, st>>>(*D, tm, lq);
  }
  if (nesp)
    k_ffedge<T><<<ge, 256, 0, st>>>(*D, tin, edn, nesp, lci, ld, lq, fx, te, anydif, need,
                                    stamp);
  if (d->nwet) {
    k_ffmoc<T><<<gw, 256, 0, st>>>(*D, tin, te, moc, need, stamp);
    k_fffz<false, T><<<gw, 256, 0, st>>>(*D, dtu, tin, moc, clf, crf, cour, fz, err, need,
                                         stamp);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int emstr_us_launch_ffsl_flux(const emstr_usdev *D, double dtu, const double *tin,
                                         const double *nw, const int *clc, const double *crc,
                                         const int *clf, const double *crf, const double *cour,
                                         const long long *edn, int nesp, const int *lci,
                                         const double *ld, double *tm, double *lq, double *te,
                                         double *moc, double *fx, double *fz, int anydif,
                                         int *need, int stamp, int *err, void *stream) {
  return ffsl_flux<double>(D, dtu, tin, nw, clc, crc, clf, crf, cour, edn, nesp, lci, ld, tm,
                           lq, te, moc, fx, fz, anydif, need, stamp, err, (hipStream_t)stream);
}

/* the same with FP32 intermediates and arithmetic (the FP32 form) */
extern "C" int emstr_us_launch_ffsl_flux32(const emstr_usdev *D, double dtu, const double *tin,
                                           const double *nw, const int *clc, const double *crc,
                                           const int *clf, const double *crf,
                                           const double *cour, const long long *edn, int nesp,
                                           const int *lci, const double *ld, float *tm,
                                           float *lq, float *te, float *moc, float *fx,
                                           float *fz, int anydif, int *need, int stamp,
                                           int *err, void *stream) {
  return ffsl_flux<float>(D, dtu, tin, nw, clc, crc, clf, crf, cour, edn, nesp, lci, ld, tm, lq,
                          te, moc, fx, fz, anydif, need, stamp, err, (hipStream_t)stream);
}
