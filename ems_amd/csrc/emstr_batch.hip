/*
 * Batched form of the EMS tracer step on MI355X (gfx950): every tracer of
 * the step over every water column in one fused pass, FP32 arithmetic on a
 * 32-bit-significand state, FP64 budgets.
 *
 * Layout (HBM): tracers [ntr][nz][nyp][nxp] FP32 heads and, beside them,
 * one 8-bit tail per value (the state's next eight significand bits), the
 * tails of four tracers packed in one word [ntr/4][nz][nyp][nxp][4]; the
 * plane padded to
 * nxp = 64-multiple and nyp = 32-multiple columns (the padding is land), so
 * that a wave's block of columns (TX x 64/TX, TX consecutive x columns per
 * row) reads whole, aligned row segments.  Per cell, three float4 words of
 * tracer-independent coefficients [nz][nyp][nxp][3]: A = (qW, GW, qS, GS)
 * the transports of the west and south faces and the weights of their
 * limited van Leer terms, B = (qB, GB, 1/V, D) those of the bottom face,
 * the inverse volume and the net FP64 volume divergence, C = (gm, gp,
 * 1/div, ud) the normalised column matrix and its LU factors; the east,
 * north and top faces are the west, south and bottom faces of the
 * neighbouring cells.  Wetness is folded into the words: a face whose
 * near upwind cell is dry has a zero transport there, a face whose far
 * upwind cell is dry a zero limiter weight.  Per column a Col2 record
 * (surfaces, diffusion extent, flags, FP64 surface volume balance).
 *
 * Per step:
 *   k_cell2   one thread per cell: coefficients from the FP64 fields
 *             (tracers.c:496-578, 3841-3889 restated);
 *   k_col2    one thread per column: the column record and the LU factors
 *             of the column matrix;
 *   k_step2   one workgroup per (tile of 16 x 4 columns, 4 tracers): every
 *             lane walks its column from the bed to the top with a window
 *             of the tracer's values and vertical differences, the x
 *             neighbours from the lanes beside it, the y neighbours read
 *             from L1/L2, and computes
 *               - the sub-surface update (tracers.c:888-925) in anomaly
 *                 form, tr - (sum_f s_f q_f (F_f - tr) + tr D) / V, each
 *                 face's term the upwind value's difference against the
 *                 cell times max(q, 0) or min(q, 0), plus its limited van
 *                 Leer term G H (tracers.c:2097-2118), all from the
 *                 stencil's differences, each formed once per layer,
 *               - the processed surface cell by surf_conc
 *                 (tracers.c:3047-3207) as an FP64 volume balance, the
 *                 layers above it folded in on the way up,
 *             then tr_bounds (tracers.c:3217-3324) and the implicit
 *             vertical diffusion (tracers.c:3782-3957, 4067-4214) of the
 *             column in increment form, A (x - b) = dzdt b - A b, with one
 *             FP64 residual pass of iterative refinement where the
 *             diffusion number exceeds IR_DMAX, and writes the column once.
 * The blocks are dealt to the 8 XCDs in work-balanced chunks of vertical
 * strips, and the tracers of one block run back to back on one XCD, so
 * its coefficients and the halo rows of the block before it are L2 hits.
 * A cell's new value t0 + (advective increment) + (diffusive increment) is
 * formed in FP64 and rounded to nearest to the 32-bit state once per
 * step.  The vertical differences of the walk and of the column solve
 * carry the tails (head differences plus tail differences); the
 * horizontal stencil uses the heads, whose truncation is common to the
 * cell and its neighbours.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
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
#define SWAVES 4 /* rows of a tile: the four 16-lane rows of a wave */
/* tracers per lane (1, 2 or 4, per context): a tracer group's values of
 * a cell are that many consecutive floats (heads) and bytes (tails), one
 * wide load each */

typedef long long ll;

enum { F_PROC = 1, F_SINGLE = 2, F_STIFF = 4, F_THIN = 8, F_KEEPOLD = 16, F_MOVE = 32 };

/* per-column record (112 bytes) */
struct Col2 {
  int kb;    /* bed layer, nz for land                                 */
  int ks;    /* processed surface layer, -1: column not processed      */
  int kold;  /* old surface layer                                      */
  int cs;    /* new surface layer (top of the diffusion extent)        */
  int cb;    /* bottom of the diffusion extent                         */
  int flags;
  int nbw;   /* bed layers of the x neighbours: w1 | w2 << 8 | e1 << 16 | e2 << 24 */
  int nbs;   /* of the y neighbours: s1 | s2 << 8 | n1 << 16 | n2 << 24 */
  double vc;   /* consistent old volume of the surface cells + inflow */
  double svol; /* new surface volume                                   */
  double dzs;  /* subeta - waterbot of the processed surface cell      */
  double area, osub, botz;
  double pad0, pad1, pad2, pad3;
};

struct emstr_batch {
  int ntr, nz, nx, ny, nxp, nyp;
  int tpl;         /* tracers per lane (1, 2, 4)                         */
  ll P2;           /* nxp * nyp                                          */
  ll vol;          /* nz * P2: one tracer                                */
  float *tr[2];    /* [ntr][nz][nyp][nxp], the 24-bit heads              */
  uint8_t *rm[2];  /* [ntr/tpl][nz][nyp][nxp][tpl], the 8-bit tails      */
  int cur;
  float4 *co;      /* [nz][nyp][nxp][3]                                  */
  Col2 *ci;        /* [nyp][nxp]                                         */
  float *scr;      /* forward-sweep values, then the solve's increments, the
                      tracer layout                                      */
  float *scr2;     /* the refinement's corrections, the tracer layout    */
  float *mintr, *maxtr;
  /* block lists: per XCD a sequence of blocks (i0 | j0 << 16, kmin) */
  int tx;          /* block width (columns per row)                      */
  int strip;       /* strip width in columns                             */
  int nbins;       /* depth bins of the tile order in each XCD's chunk   */
  int *blk;        /* [16 + 2 nblk]: per XCD first block and count, then
                      (i0 | j0 << 16, kmin) per block                    */
  int nblk;
  int boff[8], bcnt[8], bmax;
  int kmin;        /* deepest bed of any block                           */
  emstr_prof *prof;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int graph_ok, graph_cur;
};

__device__ __forceinline__ float4 ld4(const float4 *p) { return *p; }

/*-------------------------------------------------------------------*/
/* prep: coefficients and the column record, one thread per column    */
/*-------------------------------------------------------------------*/
/* The weight of the limited term of a van Leer face (tracers.c:2097-2118)
 * in the face's flux, G = q (F - t_up) / H: with the upwind value t_up,
 * the limiter H(a, b) = a b / (a + b) for a b > 0 (else 0) of the
 * upwind difference a and the difference across the face b, and the
 * factor (1 - cn) of the face value for q > 0, (1 + cn) for q < 0 (the
 * limiter's factor 2 and the 1/2 of tracers.c:2108 cancel), G = |q| (1 -
 * |cn|).  G is zero where the stencil's far upwind cell is dry at the
 * face's layer: the reference reads that cell's ghost, which holds the
 * near upwind value (tracers.c:5574-5593 laterally, the sediment ghost
 * tracers.c:827 below the bed, the self-mapping top), so the upwind
 * difference and the limited term vanish. */
__device__ __forceinline__ float vlg(double q, double cn, bool farwet) {
  return farwet ? (float)(fabs(q) * ((q > 0.0) ? 1.0 - cn : 1.0 + cn)) : 0.0f;
}

struct PArgs {
  int nxp, nyp;
  ll P2;
};

/* per-cell coefficients, one thread per cell of the padded grid (cells
 * below the bed and land keep their zero words) */
__global__ void __launch_bounds__(256) k_cell2(emstr_dev d, PArgs a, float4 *__restrict__ co) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const ll col = (ll)j * d.nx + i;
  const int kbc = d.kb[col];
  if (k < kbc) return;
  const ll P = d.plane, nx = d.nx;
  const ll idx = (ll)k * P + col;
  const double dt = d.dt, area = d.cellarea[col];
  const int ksc = d.ks[col], cs = d.knew[col];
  const double dz = d.dz[idx];
  const double dzb = (k > kbc) ? d.dz[idx - P] : 0.0;
  const bool we = k >= d.kb[col + 1], wn = k >= d.kb[col + nx];
  /* the west and south neighbours (the faces' near upwind cells for q >
   * 0) wet at layer k; their own west and south neighbours are read only
   * then (a wet cell is not on the grid's edge) */
  const bool ww = k >= d.kb[col - 1], ws = k >= d.kb[col - nx];
  const bool tp = (ksc >= 0 && k < ksc);
  const double qW = d.u1flux3d[idx] * dt, qS = d.u2flux3d[idx] * dt;
  const double qE = we ? d.u1flux3d[idx + 1] * dt : 0.0;
  const double qN = wn ? d.u2flux3d[idx + nx] * dt : 0.0;
  const double qB = d.w[idx] * dt * area;
  /* the top face's transport is the bottom face's of the cell above */
  const double qT = tp ? d.w[idx + P] * dt * area : 0.0;
  const double cW = d.u1[idx] * dt / d.h1au1[col];
  const double cS = d.u2[idx] * dt / d.h2au2[col];
  const double cB = d.w[idx] * dt / (0.5 * (dz + dzb));
  /* far upwind cells: for q > 0 the second neighbour through the near
   * one, for q < 0 the neighbour beyond this cell */
  const bool fW = (qW > 0.0) ? (ww && k >= d.kb[col - 2]) : we;
  const bool fS = (qS > 0.0) ? (ws && k >= d.kb[col - 2 * nx]) : wn;
  const bool fB = (qB > 0.0) ? (k - 2 >= kbc) : (k + 1 < d.nz);
  float gm = 0.0f, gp = 0.0f;
  if (ksc >= 0 && k <= cs && cs - 1 >= kbc) {
    /* rows of the implicit vertical diffusion (tracers.c:3841-3889),
     * normalised by dzdt */
    const double dzdt = dz / dt;
    if (k > kbc) gm = (float)(-d.Kz[idx] / (0.5 * (dzb + dz)) / dzdt);
    if (k < cs) gp = (float)(-d.Kz[idx + P] / (0.5 * (dz + d.dz[idx + P])) / dzdt);
  }
  /* A face whose near upwind cell is dry carries the cell's own value
   * (its ghost's), so its anomaly term vanishes: its transport is stored
   * as zero in the word (the net divergence keeps it).  The bed cell's
   * bottom face reads the sediment ghost, the bed value. */
  const bool oW = ww, oS = ws, oB = k > kbc;
  float4 *o = co + ((ll)k * a.P2 + (ll)j * a.nxp + i) * 3;
  o[0] = make_float4(oW ? (float)qW : 0.0f, oW ? vlg(qW, cW, fW) : 0.0f, oS ? (float)qS : 0.0f,
                     oS ? vlg(qS, cS, fS) : 0.0f);
  o[1] = make_float4(oB ? (float)qB : 0.0f, oB ? vlg(qB, cB, fB) : 0.0f,
                     (dz != 0.0) ? (float)(1.0 / (area * dz)) : 0.0f,
                     (float)(qE - qW + qN - qS + qT - qB));
  o[2] = make_float4(gm, gp, 1.0f, 0.0f);
}

/* the column record and the LU factors of the column matrix, one thread
 * per column of the padded plane (after k_cell2) */
__global__ void __launch_bounds__(256) k_col2(emstr_dev d, PArgs a, float4 *__restrict__ co,
                                              Col2 *__restrict__ ci) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nxp || j >= a.nyp) return;
  const ll c2 = (ll)j * a.nxp + i;
  Col2 c;
  memset(&c, 0, sizeof(c));
  c.ks = -1;
  c.kb = d.nz;
  const bool in = i < d.nx && j < d.ny;
  const ll col = (ll)j * d.nx + i;
  const int kbc = in ? d.kb[col] : d.nz;
  if (kbc >= d.nz) {
    c.nbw = c.nbs = 0x01010101 * d.nz;
    ci[c2] = c;
    return;
  }
  const ll P = d.plane, nx = d.nx;
  const int top = d.nz - 1;
  const double dt = d.dt, area = d.cellarea[col];
  c.kb = kbc;
  /* neighbour beds; a second neighbour counts only if the first is wet */
  const int w1 = d.kb[col - 1], e1 = d.kb[col + 1], s1 = d.kb[col - nx], n1 = d.kb[col + nx];
  const int w2 = (w1 < d.nz) ? d.kb[col - 2] : d.nz, e2 = (e1 < d.nz) ? d.kb[col + 2] : d.nz;
  const int s2 = (s1 < d.nz) ? d.kb[col - 2 * nx] : d.nz;
  const int n2 = (n1 < d.nz) ? d.kb[col + 2 * nx] : d.nz;
  c.nbw = w1 | (w2 << 8) | (e1 << 16) | (e2 << 24);
  c.nbs = s1 | (s2 << 8) | (n1 << 16) | (n2 << 24);
  const int ksc = d.ks[col];
  const int cs = d.knew[col];
  c.ks = ksc;
  c.kold = d.kold[col];
  c.cs = cs;
  c.area = area;
  c.botz = d.botz[col];
  c.osub = d.osubeta[col];
  int flags = 0;
  if (ksc >= 0) {
    flags |= F_PROC;
    /* surface volume balance in FP64 (surf_conc): old volumes of the
     * cells from the processed surface to the old surface, minus the net
     * horizontal outflow of every cell from it to the top, plus the inflow
     * through its bottom face */
    const int kold = c.kold;
    const double waterbot = (ksc == kbc) ? c.botz : d.layers[ksc];
    c.dzs = d.subeta[col] - waterbot;
    c.svol = c.dzs * area;
    double vol = 0.0, wb = waterbot;
    int cc = ksc;
    while (cc != kold && cc < top) {
      vol += area * (d.layers[cc + 1] - wb);
      cc++;
      wb = d.layers[cc];
    }
    vol += area * (c.osub - ((cc == kbc) ? c.botz : d.layers[cc]));
    double out = 0.0;
    bool above = false;
    for (int k = ksc; k <= top; k++) {
      const ll idx = (ll)k * P + col;
      const bool we = k >= e1, wn = k >= n1;
      const double qE = we ? d.u1flux3d[idx + 1] * dt : 0.0;
      const double qN = wn ? d.u2flux3d[idx + nx] * dt : 0.0;
      out += qE - d.u1flux3d[idx] * dt + qN - d.u2flux3d[idx] * dt;
      if (k > ksc && (qE != 0.0 || qN != 0.0 || d.u1flux3d[idx] != 0.0 || d.u2flux3d[idx] != 0.0))
        above = true;
    }
    c.vc = vol - out + d.w[(ll)ksc * P + col] * dt * area;
    if (kold != ksc || above) flags |= F_MOVE;
    /* thin old surface layer (tracers.c:3250) */
    const double dzco = (kold >= kbc) ? d.dz[(ll)kold * P + col] : 0.0;
    if (ksc > kbc && dzco < d.hmin) flags |= F_THIN;
    if (c.botz > 0.0 && d.dz[(ll)ksc * P + col] < d.hmin) flags |= F_KEEPOLD;
    /* vertical diffusion extent (tracers.c:3836-3866) */
    if (cs < kbc || cs - 1 < kbc || d.dz[(ll)(cs - 1) * P + col] == 0.0) {
      flags |= F_SINGLE;
      c.cb = cs;
    } else {
      /* one above the highest zero-thickness layer at or below cs; the
       * loads of the scan are independent of each other */
      const double *__restrict__ DZ = d.dz, *__restrict__ KZ = d.Kz;
      int zq = kbc - 1;
#pragma unroll 8
      for (int k = kbc; k <= cs; k++)
        if (DZ[(ll)k * P + col] == 0.0) zq = k;
      c.cb = zq + 1;
      if (c.cb >= cs) flags |= F_SINGLE;
#pragma unroll 8
      for (int k = c.cb + 1; k <= cs; k++) {
        const double dzc = DZ[(ll)k * P + col];
        const double dzm = DZ[(ll)(k - 1) * P + col];
        if (KZ[(ll)k * P + col] * dt / (dzc * 0.5 * (dzm + dzc)) > IR_DMAX) flags |= F_STIFF;
      }
      /* LU factors of the normalised matrix (rows gm, 1 - gm - gp, gp):
       * ud_k = gp_{k-1} / div_{k-1}, div_k = 1 - gm - gp - gm ud_k; eight
       * rows loaded before their factors are stored */
      if (!(flags & F_SINGLE)) {
        float divp = 1.0f, gpp = 0.0f;
        float4 *oc = co + c2 * 3 + 2;
        for (int k0 = c.cb; k0 <= cs; k0 += 8) {
          float4 m[8];
#pragma unroll
          for (int q = 0; q < 8; q++) m[q] = oc[(ll)min(k0 + q, cs) * a.P2 * 3];
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const int k = k0 + q;
            if (k <= cs) {
              const float gm = (k > c.cb) ? m[q].x : 0.0f, gp = m[q].y;
              float ud = 0.0f, dv = 1.0f - gm - gp;
              if (k > c.cb) {
                ud = gpp / divp;
                dv = dv - gm * ud;
              }
              m[q].z = 1.0f / dv;
              m[q].w = ud;
              oc[(ll)k * a.P2 * 3] = m[q];
              divp = dv;
              gpp = gp;
            }
          }
        }
      }
    }
  }
  c.flags = flags;
  ci[c2] = c;
}

/*-------------------------------------------------------------------*/
/* van Leer face values as anomalies                                  */
/*-------------------------------------------------------------------*/
/* The van Leer limiter of a face (tracers.c:2097-2118) in difference
 * form.  For q > 0 the reference's df = (t[bm1] - t[bm2]) (t[c] -
 * t[bm1]) and t[c] - t[bm2] are a b and a + b, a = dl the upwind
 * difference and b = dm the difference across the face; for q < 0 the
 * mirrored stencil (fp1, c, bm1) gives -a b / (a + b) with a = dr, the
 * sign going into the face's weight.  H = a b / (a + b) for a b > 0,
 * else 0 (first order).  Every difference is formed once per layer and
 * shared by the faces that read it. */
__device__ __forceinline__ float vlh(bool pos, float dl, float dr, float dm) {
  const float up = pos ? dl : dr;
  const float df = up * dm;
  const float t = df * __builtin_amdgcn_rcpf(up + dm);
  return (df > 0.0f) ? t : 0.0f;
}

/* Compensated state: a value is held as an FP32 head h, the value
 * truncated to 24 significant bits, and an 8-bit tail r, its next eight
 * bits, so that the state keeps 32 significant bits:
 *   x = h + sign(h) r 2^(e - 31),  e the exponent of h.
 * enc rounds the FP64 value x to 32 bits, to nearest (half of the 21
 * dropped bits added, which are then cleared), and splits it; values
 * below the FP32 normal range keep r = 0.  One rounding per step of at
 * most 2^-32 of the value keeps 1000 steps within 2.4e-7 even if every
 * rounding went the same way. */
__device__ __forceinline__ void enc(double x, float &h, uint32_t &r) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  u += 0x100000ULL;
  u &= ~0x1FFFFFULL;
  const bool tiny = fabs(x) < 1.1754943508222875e-38;
  h = tiny ? (float)x : (float)__longlong_as_double((long long)(u & ~0x1FFFFFFFULL));
  r = tiny ? 0u : (uint32_t)(u >> 21) & 0xFFu;
}
/* the stored value in FP64 (exact) */
__device__ __forceinline__ double dec(float h, uint32_t r) {
  return __longlong_as_double(__double_as_longlong((double)h) | ((long long)r << 21));
}
/* the tail as an FP32 number: r 2^(e - 31) with the sign of h (exact
 * above 2^-95; h = m 2^f with m in [0.5, 1) gives f = e + 1) */
__device__ __forceinline__ float tail(float h, uint32_t r) {
  return __builtin_amdgcn_ldexpf(__builtin_copysignf((float)r, h),
                                 __builtin_amdgcn_frexp_expf(h) - 32);
}
/* the tail byte of h + t (t = tail(h, r)) */
__device__ __forceinline__ uint32_t tailbyte(float h, float t) {
  return (uint32_t)(__double_as_longlong((double)h + (double)t) >> 21) & 0xFFu;
}

/*-------------------------------------------------------------------*/
/* the fused step                                                    */
/*-------------------------------------------------------------------*/
struct SArgs {
  int nz, ntr, nxp;
  ll P2, vol;
  double hmin, dt;
  ll P;                 /* unpadded plane of the FP64 fields            */
  int nx;               /* unpadded row length of the FP64 fields       */
  int kl;               /* the lowest layer whose forward-sweep values
                           the wave keeps in LDS                        */
};

/* LDS per wave for the forward-sweep values of the top layers: 20 KiB, so
 * that eight waves fit a compute unit's 160 KiB */
constexpr int LDS_WAVE = 20480;

/* raw buffer access with 32-bit byte offsets from a wave-uniform base
 * (resource word 3 of the gfx9 family); a read past the end returns 0 */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
/* o: the lane's byte offset; so: a wave-uniform byte offset (a scalar
 * register), added to it (the range check covers o only: the buffers
 * carry slack beyond their last element for the reads past it) */
__device__ __forceinline__ float2 bl2(__amdgpu_buffer_rsrc_t r, uint32_t o, uint32_t so = 0) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
__device__ __forceinline__ float3 bl3(__amdgpu_buffer_rsrc_t r, uint32_t o, uint32_t so = 0) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, o, so, 0);
  return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}
__device__ __forceinline__ float4 bl4(__amdgpu_buffer_rsrc_t r, uint32_t o, uint32_t so = 0) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, o, so, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                     __uint_as_float(v[3]));
}
/* the T values of one cell of a tracer group (one wide load) */
template <int T>
__device__ __forceinline__ void blt(float (&v)[T], __amdgpu_buffer_rsrc_t r, uint32_t o,
                                    uint32_t so) {
  if constexpr (T == 4) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, o, so, 0);
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = __uint_as_float(x[u]);
  } else if constexpr (T == 2) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, o, so, 0);
#pragma unroll
    for (int u = 0; u < 2; u++) v[u] = __uint_as_float(x[u]);
  } else {
    v[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0));
  }
}
/* the T tail bytes of one cell of a tracer group, as one word */
template <int T>
__device__ __forceinline__ uint32_t blr(__amdgpu_buffer_rsrc_t r, uint32_t o, uint32_t so) {
  if constexpr (T == 4)
    return __builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0);
  else if constexpr (T == 2)
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, o, so, 0);
  else
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, o, so, 0);
}
template <int T> __device__ __forceinline__ void stt(float *p, const float (&v)[T]) {
  if constexpr (T == 4)
    *(float4 *)p = make_float4(v[0], v[1], v[2], v[3]);
  else if constexpr (T == 2)
    *(float2 *)p = make_float2(v[0], v[1]);
  else
    *p = v[0];
}
template <int T> __device__ __forceinline__ void str(uint8_t *p, uint32_t w) {
  if constexpr (T == 4)
    *(uint32_t *)p = w;
  else if constexpr (T == 2)
    *(uint16_t *)p = (uint16_t)w;
  else
    *p = (uint8_t)w;
}
__device__ __forceinline__ uint32_t byteof(uint32_t w, int u) { return (w >> (8 * u)) & 0xFFu; }

/* DPP moves inside rows of 16 lanes: lane l takes lane l - 1 (shr) or
 * l + 1, l + 2 (shl); a lane whose source is outside its row takes old */
__device__ __forceinline__ float shr1o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl1o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x101, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl2o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x102, 0xf, 0xf, false));
}

/* the loads of one layer k: the lane's column at layer k + 2 (the
 * window's new entry), its four y neighbours and the x neighbours beyond
 * its row segment (edge lanes), for the T tracers of the group; the
 * cell's west/south and bottom coefficient words, its top face (the
 * bottom word of the cell above), the matrix row of the layer below (the
 * forward row), the north neighbour's south face and, for the segment's
 * last lane, the east neighbour's west face */
template <int T> struct Lay {
  float z2[T], s1[T], s2[T], n1[T], n2[T], x1[T], x2[T];
  uint32_t z2r; /* the tails of z2 */
  float4 A, B;
  float2 U;
  float3 M;
  float2 N, E;
};

/* One wave per (tile of 16 x 4 columns, group of T tracers): lane l is
 * column i0 + l % 16 of row j0 + l / 16 and carries the T tracers of the
 * group, held tracer-innermost in memory so that one wide load brings
 * them all; the tracer-independent words (coefficients, wetness,
 * upwind directions) are loaded and evaluated once for the T tracers.  A
 * lane's x neighbours are the lanes beside it (DPP); every lane walks its
 * column from the tile's deepest bed to the top, the loads of layer k +
 * 1 issued before layer k is computed.  The advective increments go to
 * the output column and the forward-sweep values to the scratch column
 * until the back substitution forms each cell's value once. */
template <int T>
__global__ void __launch_bounds__(64)
k_step4(const SArgs a, const float *__restrict__ tin, float *__restrict__ tout,
        const uint8_t *__restrict__ rin, uint8_t *__restrict__ rout,
        const float4 *__restrict__ co, const Col2 *__restrict__ cinfo,
        const int *__restrict__ blk, const float *__restrict__ mintr,
        const float *__restrict__ maxtr, const double *__restrict__ layers,
        const double *__restrict__ Kz, const double *__restrict__ dzf,
        float *__restrict__ scr, float *__restrict__ scr2) {
  const int xcd = blockIdx.x & 7;
  const int q = blockIdx.x >> 3;
  const int NG = (a.ntr + T - 1) / T;
  const int bq = q / NG, g = q - bq * NG;
  /* blk: [0..7] first tile of each XCD, [8..15] tile counts, then per
   * tile (i0 | j0 << 16, the four rows' deepest beds as bytes) */
  if (bq >= blk[8 + xcd]) return;
  const int bi = blk[xcd] + bq;
  const int b0 = blk[16 + 2 * bi], kms = blk[17 + 2 * bi];
  const int kmin = min(min(kms & 255, (kms >> 8) & 255), min((kms >> 16) & 255, kms >> 24 & 255));
  if (kmin >= a.nz) return;
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15;
  const int i = (b0 & 0xffff) + cl, j = (b0 >> 16) + (lane >> 4);
  const int nz = a.nz, top = nz - 1;
  const uint32_t P2 = (uint32_t)a.P2, nxp = (uint32_t)a.nxp;
  const uint32_t col = (uint32_t)j * nxp + (uint32_t)i;
  const Col2 ci = cinfo[col];
  const int kb = ci.kb, ks = ci.ks, flags = ci.flags;
  /* the group's tracer bounds (wave-uniform; padding tracers take the
   * last tracer's) */
  float mn[T], mx[T];
  bool clip[T];
#pragma unroll
  for (int u = 0; u < T; u++) {
    const int n = min(g * T + u, a.ntr - 1);
    mn[u] = mintr[n];
    mx[u] = maxtr[n];
    clip[u] = mn[u] <= 1e-5f;
  }
  constexpr uint32_t VB = 4 * T; /* bytes of one cell's heads */
  const uint32_t vol = (uint32_t)a.vol;
  const __amdgpu_buffer_rsrc_t RT = rsrc(tin, vol * 4u * (uint32_t)(NG * T));
  const __amdgpu_buffer_rsrc_t RC = rsrc(co, vol * 48u);
  const __amdgpu_buffer_rsrc_t RR = rsrc(rin, vol * (uint32_t)(NG * T));
  /* element offsets of the lane's column in its group (heads: T floats,
   * tails: T bytes per cell) */
  const uint32_t ce = (uint32_t)g * vol + col;
  float *OB = tout + (size_t)ce * T;
  uint8_t *OR = rout + (size_t)ce * T;
  float *SB = scr + (size_t)ce * T;
  float *SE = scr2 + (size_t)ce * T;
  /* the forward-sweep values s_r of the lane's column: in the wave's LDS
   * for r >= kl (one row of 64 lanes x T per layer), in the scratch
   * column below */
  extern __shared__ float lsh[];
  auto sput = [&](int r, const float (&v)[T]) {
    if (r >= a.kl) {
      float *q = lsh + ((r - a.kl) * 64 + (threadIdx.x & 63)) * T;
#pragma unroll
      for (int u = 0; u < T; u++) q[u] = v[u];
    } else {
      stt<T>(SB + (size_t)r * P2 * T, v);
    }
  };
  auto sget = [&](int r, float (&v)[T]) {
    if (r >= a.kl) {
      const float *q = lsh + ((r - a.kl) * 64 + (threadIdx.x & 63)) * T;
#pragma unroll
      for (int u = 0; u < T; u++) v[u] = q[u];
    } else {
#pragma unroll
      for (int u = 0; u < T; u++) v[u] = SB[(size_t)r * P2 * T + u];
    }
  };
  /* byte offsets: of the lane's column (heads, tails), of a layer */
  const uint32_t c4 = ce * VB, L4 = P2 * VB;
  const uint32_t cr = ce * (uint32_t)T, LR = P2 * (uint32_t)T;
  const uint32_t c48 = col * 48u, L48 = P2 * 48u;
  const float4 *__restrict__ CC = co + col * 3 + 2; /* the column's matrix words */
  const float *__restrict__ CU = (const float *)CC + 3; /* their ud entries */
  const bool proc = (flags & F_PROC) != 0;
  const bool move = (flags & F_MOVE) != 0;
  const bool solve = proc && !(flags & F_SINGLE);
  const int cb = ci.cb, cs = ci.cs, kold = ci.kold;
  const uint32_t ry = nxp * VB;
  /* the lane's offsets of its y neighbours and of the x neighbours beyond
   * its row segment (x1: lane 0 its west, lane 15 its east neighbour; x2:
   * lane 0 its second west, lanes 14 and 15 their second east neighbours;
   * other lanes their own cell) */
  const uint32_t os1 = c4 - ry, os2 = c4 - 2 * ry, on1 = c4 + ry, on2 = c4 + 2 * ry;
  const uint32_t ox1 = (cl == 0) ? c4 - VB : (cl == 15) ? c4 + VB : c4;
  const uint32_t ox2 = (cl == 0) ? c4 - 2 * VB : (cl >= 14) ? c4 + 2 * VB : c4;
  const uint32_t oN = c48 + nxp * 48u + 8u;
  /* Issue the loads of layer k: every lane the same words at its own
   * offsets, the layer's offset uniform (below a lane's bed, above its
   * surface or at a neighbour that is land the words are read and not
   * used, or used with a zero weight: the unused cells hold zeros) */
  auto issue = [&](int k, Lay<T> &y) {
    const uint32_t kz = (uint32_t)min(k + 2, top);
    blt<T>(y.z2, RT, c4, kz * L4);
    y.z2r = blr<T>(RR, cr, kz * LR);
    const uint32_t kt = (uint32_t)k * L4, kc = (uint32_t)k * L48;
    y.A = bl4(RC, c48, kc);
    blt<T>(y.s1, RT, os1, kt);
    blt<T>(y.n1, RT, on1, kt);
    blt<T>(y.s2, RT, os2, kt);
    blt<T>(y.n2, RT, on2, kt);
    blt<T>(y.x1, RT, ox1, kt);
    blt<T>(y.x2, RT, ox2, kt);
    y.B = bl4(RC, c48 + 16u, kc);
    y.U = bl2(RC, c48 + 16u, k < top ? kc + L48 : kc);
    y.M = bl3(RC, c48 + 32u, (uint32_t)__builtin_amdgcn_readfirstlane(max(k - 1, 0)) * L48);
    y.N = bl2(RC, oN, kc);
    y.E = bl2(RC, c48 + 48u, kc);
  };
  /* Register window of the column, per tracer: the values of layers k and
   * k + 1 (heads and tails) and the vertical differences D(m) = x(m) -
   * x(m - 1) of layers k - 1 .. k + 1, clamped to the top (the top layer
   * maps to itself).  Below a lane's bed the window holds the zeros of the
   * unused cells: every term that would read them has a zero weight (the
   * bed cell's bottom transport, the limiter weights of faces whose far
   * upwind cell is the sediment ghost) and they are finite. */
  float t0[T], d00[T], zp1[T], dp1[T], Dzm1[T], Dz0[T], Dz1[T];
  {
    const int k0 = kmin;
    float h0[T], h1[T], hm1[T], hm2[T];
    blt<T>(h0, RT, c4, (uint32_t)min(k0, top) * L4);
    blt<T>(h1, RT, c4, (uint32_t)min(k0 + 1, top) * L4);
    blt<T>(hm1, RT, c4, (uint32_t)max(k0 - 1, 0) * L4);
    blt<T>(hm2, RT, c4, (uint32_t)max(k0 - 2, 0) * L4);
    const uint32_t r0 = blr<T>(RR, cr, (uint32_t)min(k0, top) * LR);
    const uint32_t r1 = blr<T>(RR, cr, (uint32_t)min(k0 + 1, top) * LR);
    const uint32_t rm1 = blr<T>(RR, cr, (uint32_t)max(k0 - 1, 0) * LR);
    const uint32_t rm2 = blr<T>(RR, cr, (uint32_t)max(k0 - 2, 0) * LR);
#pragma unroll
    for (int u = 0; u < T; u++) {
      t0[u] = h0[u];
      zp1[u] = h1[u];
      d00[u] = tail(h0[u], byteof(r0, u));
      dp1[u] = tail(h1[u], byteof(r1, u));
      const float zm1 = (k0 >= 1) ? hm1[u] : 0.0f, zm2 = (k0 >= 2) ? hm2[u] : 0.0f;
      const float em1 = (k0 >= 1) ? tail(hm1[u], byteof(rm1, u)) : 0.0f;
      const float em2 = (k0 >= 2) ? tail(hm2[u], byteof(rm2, u)) : 0.0f;
      Dzm1[u] = (zm1 - zm2) + (em1 - em2);
      Dz0[u] = (t0[u] - zm1) + (d00[u] - em1);
      Dz1[u] = (zp1[u] - t0[u]) + (dp1[u] - d00[u]);
    }
  }
  /* per tracer: the advective increments of the two layers below, the
   * last forward-sweep value, the bottom face's limited flux term (the top
   * face of the layer below) */
  float dd1[T], dd2[T], sp[T], GHp[T];
#pragma unroll
  for (int u = 0; u < T; u++) dd1[u] = dd2[u] = sp[u] = GHp[u] = 0.0f;
  /* surface cell (from the layer below it on): the anomaly terms of its
   * volume balance, its old value (head, tail), the unclipped increment of
   * the layer below it */
  double m2[T];
  float v0s[T], v0d[T], duk[T];
#pragma unroll
  for (int u = 0; u < T; u++) {
    m2[u] = 0.0;
    v0s[u] = v0d[u] = duk[u] = 0.0f;
  }
  /* the layers where no lane of the wave is within a layer of its
   * processed surface, walked without the surface terms */
  int kA = proc ? ks - 1 : nz;
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) kA = min(kA, __shfl_xor(kA, m));
  kA = __builtin_amdgcn_readfirstlane(kA);

  /* one layer; SURF: the surface terms may occur */
  auto layer = [&](auto surf_c, int k, const Lay<T> &y) {
    constexpr bool SURF = decltype(surf_c)::value;
    const bool has = k >= kb;
    /* the window's new entry and the differences of every tracer */
    float zp2[T], dp2[T], Dz2[T], Dx0[T], Dxm[T], Dxp[T], Dxp2[T];
#pragma unroll
    for (int u = 0; u < T; u++) {
      zp2[u] = y.z2[u];
      dp2[u] = tail(zp2[u], byteof(y.z2r, u));
      Dz2[u] = (zp2[u] - zp1[u]) + (dp2[u] - dp1[u]);
      /* x differences D(i - 1) .. D(i + 2) (D(i) = t(i) - t(i - 1)): the
       * lanes beside this one, at the row segment's ends the loads
       * beyond it */
      const float l1 = shr1o(y.x1[u], t0[u]), r1 = shl1o(y.x1[u], t0[u]);
      Dx0[u] = t0[u] - l1;
      Dxm[u] = shr1o(y.x1[u] - y.x2[u], Dx0[u]);
      Dxp[u] = shl1o(r1 - t0[u], Dx0[u]);
      Dxp2[u] = shl2o(y.x2[u] - r1, Dx0[u]);
    }
    /* the east face's transport and limiter weight: the west word of the
     * lane to the east, or of the east neighbour at the segment's end */
    const float qE = shl1o(y.E.x, y.A.x), gE = shl1o(y.E.y, y.A.y);
    const bool sub = proc && has && (!SURF || k < ks);
    const bool srf = SURF && proc && k == ks;
    const bool upp = SURF && proc && move && k > ks;
    if (!proc && has) {
      stt<T>(OB + (size_t)k * P2 * T, t0);
      uint32_t w = 0;
#pragma unroll
      for (int u = 0; u < T; u++) w |= tailbyte(t0[u], d00[u]) << (8 * u);
      str<T>(OR + (size_t)k * P2 * T, w);
    }
    if (sub || srf || upp) {
      const float qW = y.A.x, gW = y.A.y, qS = y.A.z, gS = y.A.w, qN = y.N.x, gN = y.N.y;
      const bool pW = qW > 0.0f, pE = qE > 0.0f, pS = qS > 0.0f, pN = qN > 0.0f;
      const float qWp = fmaxf(qW, 0.0f), qEm = fminf(qE, 0.0f), qSp = fmaxf(qS, 0.0f),
                  qNm = fminf(qN, 0.0f), qBp = fmaxf(y.B.x, 0.0f);
      /* per tracer: the sum over the lateral faces of s q (F - t0) (the
       * upwind value's difference against the cell, then the limited
       * terms) and the bottom face's term -qB (F - t0) */
      float hh[T], bot[T];
#pragma unroll
      for (int u = 0; u < T; u++) {
        const float Dy0 = t0[u] - y.s1[u], Dym = y.s1[u] - y.s2[u];
        const float Dyp = y.n1[u] - t0[u], Dyp2 = y.n2[u] - y.n1[u];
        float h = qWp * Dx0[u] + qEm * Dxp[u] + qSp * Dy0 + qNm * Dyp;
        h += gE * vlh(pE, Dx0[u], Dxp2[u], Dxp[u]) - gW * vlh(pW, Dxm[u], Dxp[u], Dx0[u]) +
             gN * vlh(pN, Dy0, Dyp2, Dyp) - gS * vlh(pS, Dym, Dyp, Dy0);
        hh[u] = h;
        bot[u] = qBp * Dz0[u] - GHp[u];
      }
      if (upp) {
        /* a layer above the processed surface cell (tracers.c:3117-3193):
         * the lateral anomalies against the surface cell's old value */
        const float dh = (qE - qW) + (qN - qS);
        const double vk = (k < kold) ? ci.area * (layers[k + 1] - layers[k])
                        : (k == kold) ? ci.area * (ci.osub - layers[k]) : 0.0;
#pragma unroll
        for (int u = 0; u < T; u++) {
          m2[u] -= (double)(hh[u] + (t0[u] - v0s[u]) * dh);
          const double an = ((double)t0[u] + (double)d00[u]) - ((double)v0s[u] + (double)v0d[u]);
          m2[u] += an * vk;
        }
      } else if (sub) {
        const float qT = y.U.x, gT = y.U.y, qTm = fminf(qT, 0.0f);
        const bool pT = qT > 0.0f;
        const int r = k - 1;
        const bool fwd = solve && r >= cb;
        const float gm = (r > cb) ? y.M.x : 0.0f;
        float dl[T], sv[T];
#pragma unroll
        for (int u = 0; u < T; u++) {
          /* the top face (the bottom face of the layer above) */
          const float GHT = gT * vlh(pT, Dz0[u], Dz2[u], Dz1[u]);
          const float tp = qTm * Dz1[u] + GHT;
          /* the advected value b = t0 + dlt: its increment is kept in
           * FP32 (the output column holds it until the column's values
           * are formed and rounded once, after the solve) */
          float dlt = -(((hh[u] + tp) + bot[u]) + (t0[u] + d00[u]) * y.B.w) * y.B.z;
          /* tr_bounds below the surface: the clip test in FP32 (a value
           * within an ulp of the minimum may go either way) */
          if (SURF) duk[u] = dlt;
          if (clip[u] && t0[u] + (d00[u] + dlt) < mn[u]) dlt = (mn[u] - t0[u]) - d00[u];
          dl[u] = dlt;
          /* the forward row r = k - 1: r_k = -(gm (b_{k-1} - b_k) + gp
           * (b_{k+1} - b_k)), s_k = (r_k - gm s_{k-1}) / div_k, the
           * differences of b from the window's and the increments' */
          const float dm = (r > cb) ? (dd2[u] - dd1[u]) - Dzm1[u] : 0.0f;
          const float dp = Dz0[u] + (dlt - dd1[u]);
          const float rr = -(gm * dm + y.M.y * dp);
          sv[u] = (rr - gm * sp[u]) * y.M.z;
          dd2[u] = dd1[u];
          dd1[u] = dlt;
          GHp[u] = GHT;
        }
        stt<T>(OB + (size_t)k * P2 * T, dl);
        if (fwd) {
#pragma unroll
          for (int u = 0; u < T; u++) sp[u] = sv[u];
          sput(r, sv);
        }
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) {
          m2[u] = -((double)hh[u] + (double)bot[u]);
          v0s[u] = t0[u];
          v0d[u] = d00[u];
        }
      }
    }
    /* advance the window */
#pragma unroll
    for (int u = 0; u < T; u++) {
      t0[u] = zp1[u];
      d00[u] = dp1[u];
      zp1[u] = zp2[u];
      dp1[u] = dp2[u];
      Dzm1[u] = Dz0[u];
      Dz0[u] = Dz1[u];
      Dz1[u] = Dz2[u];
    }
  };
  using NoSurf = std::integral_constant<bool, false>;
  using Surf = std::integral_constant<bool, true>;
  Lay<T> ya, yb;
  issue(kmin, ya);
  int k = kmin;
  for (; k + 1 < kA; k += 2) {
    issue(k + 1, yb);
    layer(NoSurf{}, k, ya);
    issue(k + 2, ya);
    layer(NoSurf{}, k + 1, yb);
  }
  for (; k + 1 <= top; k += 2) {
    issue(k + 1, yb);
    layer(Surf{}, k, ya);
    issue(k + 2, ya);
    layer(Surf{}, k + 1, yb);
  }
  if (k <= top) layer(Surf{}, k, ya);
  if (!proc) return;

  /* the stored value of the lane's cells at layer kk, and the increments
   * the walk left in the output column */
  auto told = [&](int kk, double (&v)[T]) {
    float h[T];
    blt<T>(h, RT, c4, (uint32_t)kk * L4);
    const uint32_t w = blr<T>(RR, cr, (uint32_t)kk * LR);
#pragma unroll
    for (int u = 0; u < T; u++) v[u] = dec(h[u], byteof(w, u));
  };
  /* a cell's new values, rounded once to the 32-bit state */
  auto put = [&](int kk, const double (&v)[T]) {
    float h[T];
    uint32_t w = 0;
#pragma unroll
    for (int u = 0; u < T; u++) {
      uint32_t r;
      enc(v[u], h[u], r);
      w |= r << (8 * u);
    }
    stt<T>(OB + (size_t)kk * P2 * T, h);
    str<T>(OR + (size_t)kk * P2 * T, w);
  };
  /* b of the layers below the surface that the forward rows read, and
   * the unclipped advected value below the surface */
  double bk1[T], bk2[T], bun[T];
#pragma unroll
  for (int u = 0; u < T; u++) bk1[u] = bk2[u] = bun[u] = 0.0;
  if (ks - 1 >= kb) {
    told(ks - 1, bk1);
#pragma unroll
    for (int u = 0; u < T; u++) {
      bun[u] = bk1[u] + (double)duk[u];
      bk1[u] += (double)OB[(size_t)(ks - 1) * P2 * T + u];
    }
  }
  /* surface value (surf_conc) in FP64, then tr_bounds at the surface */
  double surf[T];
#pragma unroll
  for (int u = 0; u < T; u++) {
    double s = ((double)v0s[u] + (double)v0d[u]) * ci.vc + m2[u];
    if (ci.svol == 0.0 || (s < 0.0 && ci.dzs <= a.hmin))
      s = 0.0;
    else
      s /= ci.svol;
    if (flags & F_KEEPOLD) s = (double)v0s[u] + (double)v0d[u];
    if ((flags & F_THIN) && (s < (double)mn[u] || s > (double)mx[u])) s = bun[u];
    if (clip[u] && s < (double)mn[u]) s = (double)mn[u];
    surf[u] = s;
  }
  /* the layers above the surface outside the solved range take the
   * surface value, rounded once */
  const int lo = solve ? cb : top + 1, hi = solve ? cs : top;
  for (int kk = ks; kk <= top; kk++)
    if (kk < lo || kk > hi) put(kk, surf);
  /* cells below the surface outside the solved range: t0 + dlt, rounded */
  for (int kk = kb; kk < min(ks, lo); kk++) {
    double v[T];
    told(kk, v);
#pragma unroll
    for (int u = 0; u < T; u++) v[u] += (double)OB[(size_t)kk * P2 * T + u];
    put(kk, v);
  }
  if (!solve) return;
  /* the remaining forward rows (their right-hand side reaches the
   * surface value), b below the surface from the layers below it */
  {
    const int rs = max(cb, ks - 1);
    if (rs == ks - 1 && rs > cb) {
      told(ks - 2, bk2);
#pragma unroll
      for (int u = 0; u < T; u++) bk2[u] += (double)OB[(size_t)(ks - 2) * P2 * T + u];
    }
    for (int r = rs; r <= cs; r++) {
      const float4 M = CC[(size_t)r * P2 * 3];
      const float gm = (r > cb) ? M.x : 0.0f;
      float sv[T];
#pragma unroll
      for (int u = 0; u < T; u++) {
        const double b0 = (r >= ks) ? surf[u] : bk1[u];
        const float dm = (r <= cb) ? 0.0f : (r == ks - 1) ? (float)(bk2[u] - bk1[u])
                       : (r == ks) ? (float)(bk1[u] - surf[u]) : 0.0f;
        const float dp = (r < cs) ? (float)(surf[u] - b0) : 0.0f;
        const float rr = -(gm * dm + M.y * dp);
        sp[u] = (rr - gm * sp[u]) * M.z;
        sv[u] = sp[u];
      }
      sput(r, sv);
    }
  }
  /* Back substitution d_k = s_k - ud_{k+1} d_{k+1}, x = t0 + dlt + d (or
   * the surface value + d) in FP64, rounded once; two rows at a time, the
   * next two rows' t0, dlt, s and ud loaded before this pair is stored
   * (d goes to the scratch column in a stiff column) */
  const bool stiff = (flags & F_STIFF) != 0;
  {
    float dn[T], udn = 0.0f;
#pragma unroll
    for (int u = 0; u < T; u++) dn[u] = 0.0f;
    struct Row {
      float t[T], d[T], s[T];
      uint32_t w;
      float ud;
    };
    auto fetch = [&](int r, Row &x) {
      const int rc = max(r, cb);
      blt<T>(x.t, RT, c4, (uint32_t)rc * L4);
      x.w = blr<T>(RR, cr, (uint32_t)rc * LR);
#pragma unroll
      for (int u = 0; u < T; u++) x.d[u] = OB[(size_t)rc * P2 * T + u];
      sget(rc, x.s);
      x.ud = CU[(size_t)rc * P2 * 12];
    };
    auto solve_row = [&](int r, const Row &x) {
      if (r < cb) return;
#pragma unroll
      for (int u = 0; u < T; u++) dn[u] = (r == cs) ? x.s[u] : x.s[u] - udn * dn[u];
      if (stiff) {
        stt<T>(SB + (size_t)r * P2 * T, dn);
      } else {
        double v[T];
#pragma unroll
        for (int u = 0; u < T; u++)
          v[u] = ((r < ks) ? dec(x.t[u], byteof(x.w, u)) + (double)x.d[u] : surf[u]) +
                 (double)dn[u];
        put(r, v);
      }
      udn = x.ud;
    };
    Row ra0, ra1, rb0, rb1;
    fetch(cs, ra0);
    fetch(cs - 1, ra1);
    for (int r0 = cs; r0 >= cb; r0 -= 4) {
      if (r0 - 2 >= cb) {
        fetch(r0 - 2, rb0);
        fetch(r0 - 3, rb1);
      }
      solve_row(r0, ra0);
      solve_row(r0 - 1, ra1);
      if (r0 - 2 < cb) break;
      if (r0 - 4 >= cb) {
        fetch(r0 - 4, ra0);
        fetch(r0 - 5, ra1);
      }
      solve_row(r0 - 2, rb0);
      solve_row(r0 - 3, rb1);
    }
  }
  if (!stiff) return;
  /* one FP64 residual pass of iterative refinement, the matrix and the
   * right-hand side in FP64 from Kz and dz (tracers.c:3841-3889); the
   * correction is solved with the FP32 factors; dlt is in the output
   * column, d in the scratch column, the correction in the second
   * scratch column */
  {
    const ll P = a.P;
    const ll fc = (ll)j * a.nx + i;
    const double dt = a.dt;
    auto bval = [&](int kk, double (&v)[T]) {
      if (kk < ks) {
        told(kk, v);
#pragma unroll
        for (int u = 0; u < T; u++) v[u] += (double)OB[(size_t)kk * P2 * T + u];
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) v[u] = surf[u];
      }
    };
    float e[T];
#pragma unroll
    for (int u = 0; u < T; u++) e[u] = 0.0f;
    for (int kk = cb; kk <= cs; kk++) {
      const ll ix = (ll)kk * P + fc;
      const size_t o = (size_t)kk * P2 * T;
      const double dzc = dzf[ix], dzdt = dzc / dt;
      double cm1 = 0.0, cp1 = 0.0;
      if (kk > cb) cm1 = -Kz[ix] / (0.5 * (dzf[ix - P] + dzc));
      if (kk < cs) cp1 = -Kz[ix + P] / (0.5 * (dzc + dzf[ix + P]));
      double bv[T], bvm[T], bvp[T];
      bval(kk, bv);
      if (kk > cb) bval(kk - 1, bvm);
      if (kk < cs) bval(kk + 1, bvp);
      const float4 M = CC[(size_t)kk * P2 * 3];
      const float gm = (kk > cb) ? M.x : 0.0f;
#pragma unroll
      for (int u = 0; u < T; u++) {
        const double bm = (kk > cb) ? bvm[u] : bv[u], bp = (kk < cs) ? bvp[u] : bv[u];
        const double dk = (double)SB[o + u];
        const double dkm = (kk > cb) ? (double)SB[o - (size_t)P2 * T + u] : 0.0;
        const double dkp = (kk < cs) ? (double)SB[o + (size_t)P2 * T + u] : 0.0;
        const double rhs = -(cm1 * (bm - bv[u]) + cp1 * (bp - bv[u]));
        const double Ad = cm1 * dkm + (dzdt - cm1 - cp1) * dk + cp1 * dkp;
        const float rho = (float)((rhs - Ad) / dzdt);
        e[u] = (rho - gm * e[u]) * M.z;
        SE[o + u] = e[u];
      }
    }
    float ec[T], udn = 0.0f;
#pragma unroll
    for (int u = 0; u < T; u++) ec[u] = 0.0f;
    for (int kk = cs; kk >= cb; kk--) {
      const size_t o = (size_t)kk * P2 * T;
#pragma unroll
      for (int u = 0; u < T; u++) {
        const float ek = SE[o + u];
        ec[u] = (kk == cs) ? ek : ek - udn * ec[u];
        SE[o + u] = ec[u];
      }
      udn = CC[(size_t)kk * P2 * 3].w;
    }
    /* every residual is formed before any value is replaced */
    for (int kk = cb; kk <= cs; kk++) {
      const size_t o = (size_t)kk * P2 * T;
      double v[T];
      bval(kk, v);
#pragma unroll
      for (int u = 0; u < T; u++) v[u] += (double)SB[o + u] + (double)SE[o + u];
      put(kk, v);
    }
  }
}

/*-------------------------------------------------------------------*/
/* layout conversion and mass                                        */
/*-------------------------------------------------------------------*/
struct LArgs {
  ll plane, P2, vol;
  int nx, nxp, nwet, tpl;
};

__device__ __forceinline__ ll bidx(const LArgs &a, ll idx) {
  const int k = (int)(idx / a.plane);
  const ll col = idx - (ll)k * a.plane;
  const ll jj = col / a.nx, ii = col - jj * a.nx;
  return (ll)k * a.P2 + jj * a.nxp + ii;
}

/* element offset of tracer n at cell index c of its volume, heads (in
 * floats) and tails (in bytes) alike: [n / tpl][vol][tpl] */
__device__ __forceinline__ ll toff(const LArgs &a, int n, ll c) {
  return ((ll)(n / a.tpl) * a.vol + c) * a.tpl + (n % a.tpl);
}

/* FP64 -> head and tail, rounded to nearest at 32 bits, into both buffers */
__global__ void k_put2(LArgs a, const int *__restrict__ wsp, const ll *__restrict__ wdn, int n,
                       const double *__restrict__ sp, float *__restrict__ t0,
                       float *__restrict__ t1, uint8_t *__restrict__ r0,
                       uint8_t *__restrict__ r1) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nwet) return;
  const ll c = bidx(a, wdn[q]), e = toff(a, n, c), o = e;
  float h;
  uint32_t r;
  enc(sp[wsp[q]], h, r);
  t0[e] = h;
  t1[e] = h;
  r0[o] = (uint8_t)r;
  r1[o] = (uint8_t)r;
}

__global__ void k_get2(LArgs a, const int *__restrict__ wsp, const ll *__restrict__ wdn, int n,
                       const float *__restrict__ t, const uint8_t *__restrict__ rm,
                       double *__restrict__ sp) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nwet) return;
  const ll c = bidx(a, wdn[q]);
  sp[wsp[q]] = dec(t[toff(a, n, c)], rm[toff(a, n, c)]);
}

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

/* per-block (sum, err) of sum(tr*area*dz) over the processed cells */
__global__ void k_mass2(emstr_dev d, LArgs la, int nz, const float *__restrict__ tr,
                        const uint8_t *__restrict__ rm, double *__restrict__ part) {
  const int n = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double ss[256], se[256];
  double s = 0.0, e = 0.0;
  if (q < d.ncol) {
    const ll col = d.cols[q];
    const int kbc = d.kb[col], ksc = d.ks[col];
    const double ar = d.cellarea[col];
    const ll jj = col / d.nx, ii = col - jj * d.nx;
    const ll c0 = jj * la.nxp + ii;
    for (int k = ksc; k >= kbc; k--) {
      const ll o = toff(la, n, (ll)k * la.P2 + c0);
      const double v = dec(tr[o], rm[o]);
      const double m = v * ar * d.dz[(ll)k * d.plane + col];
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
static int shape_ok(int tx) { return tx == 16; }

/* Tiles of 16 x 4 columns over the padded plane (a row of 16 columns per
 * wave of a workgroup); all-land tiles dropped.
 * Order: vertical strips of `strip` columns, in each strip the tiles row
 * by row (so a tile's south halo rows are those of a tile processed
 * shortly before it); the sequence is cut into 8 contiguous chunks of
 * equal work (wet cells), one per XCD.  With nbins > 1 the tiles of each
 * chunk are then binned by depth: the layers the tile's four waves walk
 * (each from its row's deepest bed to the top) split into nbins equal
 * ranges over the window, deepest bin first, the strip order kept
 * inside a bin, so the longest walks start first and short ones fill
 * the end of the chunk. */
static int build_blocks(emstr_batch *b, const int *kb) {
  const int tx = 16, ty = 1, nxp = b->nxp, nyp = b->nyp, nz = b->nz;
  const int tty = SWAVES * ty;
  const int strip = b->strip < tx ? tx : b->strip / tx * tx;
  const int nbx = nxp / tx, nby = (nyp + tty - 1) / tty;
  int *seq = (int *)malloc(sizeof(int) * (16 + 2 * (size_t)nbx * nby));
  ll *work = (ll *)malloc(sizeof(ll) * (size_t)nbx * nby + 1);
  int ns = 0, kmin_all = nz;
  ll total = 0;
  for (int s0 = 0; s0 < nxp; s0 += strip)
    for (int by = 0; by < nby; by++)
      for (int i0 = s0; i0 < s0 + strip && i0 < nxp; i0 += tx) {
        const int j0 = by * tty;
        int kms = 0, kt = nz;
        ll wk = 0;
        for (int w = 0; w < SWAVES; w++) {
          int km = nz;
          for (int jj = j0 + w * ty; jj < j0 + (w + 1) * ty; jj++)
            for (int ii = i0; ii < i0 + tx; ii++) {
              if (ii >= b->nx || jj >= b->ny) continue;
              const int v = kb[(ll)jj * b->nx + ii];
              if (v < nz) {
                km = v < km ? v : km;
                wk += nz - v;
              }
            }
          kms |= km << (8 * w);
          kt = km < kt ? km : kt;
        }
        if (kt >= nz) continue;
        seq[16 + 2 * ns] = i0 | (j0 << 16);
        seq[17 + 2 * ns] = kms;
        work[ns] = wk;
        total += wk;
        kmin_all = kt < kmin_all ? kt : kmin_all;
        ns++;
      }
  /* 8 chunks of equal work */
  int x = 0;
  ll acc = 0;
  for (int q = 0; q < 8; q++) b->bcnt[q] = 0;
  b->boff[0] = 0;
  for (int s = 0; s < ns; s++) {
    while (x < 7 && acc >= (total * (x + 1) + 7) / 8) {
      x++;
      b->boff[x] = s;
    }
    b->bcnt[x]++;
    acc += work[s];
  }
  for (int q = x + 1; q < 8; q++) b->boff[q] = ns;
  b->bmax = 0;
  for (int q = 0; q < 8; q++) b->bmax = b->bcnt[q] > b->bmax ? b->bcnt[q] : b->bmax;
  b->kmin = kmin_all;
  for (int q = 0; q < 8; q++) {
    seq[q] = b->boff[q];
    seq[8 + q] = b->bcnt[q];
  }
  if (b->nbins > 1 && ns > 1) {
    /* walk length of each tile; bins over the window's range */
    int *len = (int *)malloc(sizeof(int) * (size_t)ns);
    int *tmp = (int *)malloc(sizeof(int) * 2 * (size_t)ns);
    int lmin = 4 * nz, lmax = 0;
    for (int t = 0; t < ns; t++) {
      const int kms = seq[17 + 2 * t];
      int l = 0;
      for (int w = 0; w < SWAVES; w++) {
        const int km = (kms >> (8 * w)) & 255;
        if (km < nz) l += nz - km;
      }
      len[t] = l;
      lmin = l < lmin ? l : lmin;
      lmax = l > lmax ? l : lmax;
    }
    const int span = lmax - lmin + 1, nb = b->nbins;
    for (int q = 0; q < 8; q++) {
      const int o = b->boff[q], n = b->bcnt[q];
      int m = 0;
      for (int bin = 0; bin < nb; bin++)
        for (int t = o; t < o + n; t++) {
          const int bt = (int)((ll)(lmax - len[t]) * nb / span);
          if (bt != bin) continue;
          tmp[2 * m] = seq[16 + 2 * t];
          tmp[2 * m + 1] = seq[17 + 2 * t];
          m++;
        }
      memcpy(seq + 16 + 2 * o, tmp, sizeof(int) * 2 * (size_t)n);
    }
    free(len);
    free(tmp);
  }
  if (b->blk) (void)hipFree(b->blk);
  b->blk = NULL;
  b->nblk = ns;
  bool ok = hipMalloc(&b->blk, sizeof(int) * (16 + 2 * (size_t)ns)) == hipSuccess;
  if (ok)
    ok = hipMemcpy(b->blk, seq, sizeof(int) * (16 + 2 * (size_t)ns), hipMemcpyHostToDevice) ==
         hipSuccess;
  free(seq);
  free(work);
  if (!ok) return emstr_set_error(EMSTR_EDEVICE, -1, "batched form: block lists");
  return 0;
}

static void drop_graph(emstr_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec), b->gexec = NULL;
  if (b->graph) (void)hipGraphDestroy(b->graph), b->graph = NULL;
  b->graph_ok = 0;
}

/* tracers per lane of a new context: four for 32 tracers or more, two
 * below (measured on one MI355X: the 400 x 600 x 40 window with 40
 * tracers takes 1.19 / 1.68 / 2.59 ms a step with four / two / one, the
 * 64 windows of 5-20 tracers 3.48 / 3.25 / 3.42 ms).  EMSTR_BATCH_TPL
 * (1, 2 or 4) overrides it, for such measurements. */
static int batch_tpl(const emstr_dev *d) {
  const char *e = getenv("EMSTR_BATCH_TPL");
  if (e && (atoi(e) == 1 || atoi(e) == 2 || atoi(e) == 4)) return atoi(e);
  return (d->ntr >= 32) ? 4 : 2;
}

extern "C" struct emstr_batch *emstr_batch_create(const emstr_dev *d, const int *h_kb,
                                                  void *stream) {
  (void)stream;
  if (d->nz > NZMAX) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: at most %d layers", NZMAX);
    return NULL;
  }
  if (d->nx >= 65536 - 64 || d->ny >= 65536 - 8) {
    emstr_set_error(EMSTR_EUNSUPPORTED, -1, "batched form: grids up to 65472 x 65528");
    return NULL;
  }
  emstr_batch *b = (emstr_batch *)calloc(1, sizeof(emstr_batch));
  b->ntr = d->ntr;
  b->nz = d->nz;
  b->nx = d->nx;
  b->ny = d->ny;
  b->nxp = (d->nx + 63) / 64 * 64;
  b->nyp = (d->ny + 31) / 32 * 32; /* tiles up to 4 x 8 rows */
  b->P2 = (ll)b->nxp * b->nyp;
  b->vol = (ll)d->nz * b->P2;
  b->tpl = batch_tpl(d);
  const ll ntp = (ll)(d->ntr + b->tpl - 1) / b->tpl * b->tpl; /* tracers with the group padding */
  if (b->vol * 48 > 0x7fffffffLL || b->vol * 4 * ntp > 0xffffffffLL) {
    /* buffer offsets of the coefficients and tracers are 32-bit */
    emstr_set_error(EMSTR_EUNSUPPORTED, -1,
                    "batched form: at most %lld cells per window and 4 GiB of tracers",
                    0x7fffffffLL / 48);
    free(b);
    return NULL;
  }
  b->tx = 16;
  b->strip = 64;
  b->nbins = 1;
  const size_t nt = (size_t)b->vol * ntp;
  bool ok = true;
  const size_t nrm = (size_t)b->vol * ntp;
  /* slack for the neighbour reads of the last layer past the end: two
   * rows of tracers, one row of coefficient words */
  const size_t trb = nt * sizeof(float) + (size_t)b->nxp * 8 * b->tpl + 256;
  const size_t cob = (size_t)b->vol * 3 * sizeof(float4) + (size_t)b->nxp * 48 + 256;
  for (int q = 0; q < 2; q++) ok &= hipMalloc(&b->tr[q], trb) == hipSuccess;
  for (int q = 0; q < 2; q++) ok &= hipMalloc(&b->rm[q], nrm) == hipSuccess;
  ok &= hipMalloc(&b->scr, nt * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->scr2, nt * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->co, cob) == hipSuccess;
  ok &= hipMalloc(&b->ci, (size_t)b->P2 * sizeof(Col2)) == hipSuccess;
  ok &= hipMalloc(&b->mintr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->maxtr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  if (!ok) {
    emstr_set_error(EMSTR_EDEVICE, -1, "batched form: device allocation failed");
    emstr_batch_destroy(b);
    return NULL;
  }
  /* land and below-bed cells are never written: zero coefficients give
   * no transport through a face into land */
  hipMemset(b->co, 0, cob);
  for (int q = 0; q < 2; q++) hipMemset(b->tr[q], 0, trb);
  for (int q = 0; q < 2; q++) hipMemset(b->rm[q], 0, nrm);
  if (build_blocks(b, h_kb)) {
    emstr_batch_destroy(b);
    return NULL;
  }
  return b;
}

extern "C" void emstr_batch_destroy(struct emstr_batch *b) {
  if (!b) return;
  drop_graph(b);
  void *v[] = {b->tr[0], b->tr[1], b->rm[0], b->rm[1], b->scr, b->scr2, b->co, b->ci, b->mintr, b->maxtr, b->blk};
  for (unsigned q = 0; q < sizeof(v) / sizeof(v[0]); q++)
    if (v[q]) hipFree(v[q]);
  free(b);
}

extern "C" void emstr_batch_set_prof(struct emstr_batch *b, struct emstr_prof *p) { b->prof = p; }

/* tile_x: tile width (16 columns, the DPP row); tile_y: strip width in
 * columns of the XCD order; grid: depth bins of the tile order (the
 * launch grid follows from the tiles) */
extern "C" int emstr_batch_tune(struct emstr_batch *b, const int *h_kb, int tile_x, int tile_y,
                                int grid) {
  if (tile_x > 0) {
    if (!shape_ok(tile_x))
      return emstr_set_error(EMSTR_EARG, -1, "batched form: block width 16");
    b->tx = tile_x;
  }
  if (tile_y > 0) b->strip = tile_y;
  if (grid > 0) {
    if (grid > 64) return emstr_set_error(EMSTR_EARG, -1, "batched form: 1 to 64 depth bins");
    b->nbins = grid;
  }
  if (build_blocks(b, h_kb)) return -1;
  drop_graph(b);
  return 0;
}

extern "C" void emstr_batch_shape(const struct emstr_batch *b, int *tile_x, int *tile_y,
                                  int *grid, int *grid_full) {
  *tile_x = b->tx;
  *tile_y = b->strip;
  *grid = b->nbins;
  *grid_full = 8 * b->bmax;
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
  drop_graph(b);
  if (emstr_launch_subeta(d, d->dt, d->dt, 1, st)) return -1;
  return 0;
}

static void launch_step(emstr_batch *b, const SArgs &sa, const float *in, float *out,
                        const uint8_t *rin, uint8_t *rout, const emstr_dev *d,
                        hipStream_t st) {
  const int T = b->tpl, ng = (b->ntr + T - 1) / T;
  const size_t lds = (size_t)(b->nz - sa.kl) * 64 * T * sizeof(float);
  const unsigned nb = (unsigned)(8 * b->bmax * ng);
#define EMSTR_STEP(TT)                                                                       \
  k_step4<TT><<<nb, 64, lds, st>>>(sa, in, out, rin, rout, b->co, b->ci, b->blk, b->mintr,  \
                                   b->maxtr, d->layers, d->Kz, d->dz, b->scr, b->scr2)
  if (T == 4)
    EMSTR_STEP(4);
  else if (T == 2)
    EMSTR_STEP(2);
  else
    EMSTR_STEP(1);
#undef EMSTR_STEP
}

/* the per-step work: prep + fused step (captured in a graph) */
static int batch_launch(emstr_batch *b, const emstr_dev *d, hipStream_t st) {
  dim3 bp(64, 4, 1);
  dim3 gp((b->nxp + 63) / 64, (b->nyp + 3) / 4, 1);
  PArgs pa;
  pa.nxp = b->nxp;
  pa.nyp = b->nyp;
  pa.P2 = b->P2;
  int pi = prof_begin(b->prof, KC_PREP, st);
  k_cell2<<<dim3((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz), bp, 0, st>>>(*d, pa, b->co);
  k_col2<<<gp, bp, 0, st>>>(*d, pa, b->co, b->ci);
  prof_end(b->prof, pi, st);
  const float *in = b->tr[b->cur];
  float *out = b->tr[1 - b->cur];
  SArgs sa;
  sa.nz = b->nz;
  sa.ntr = b->ntr;
  sa.nx = b->nx;
  sa.nxp = b->nxp;
  sa.P2 = b->P2;
  sa.vol = b->vol;
  sa.hmin = d->hmin;
  sa.dt = d->dt;
  sa.P = d->plane;
  sa.kl = max(0, b->nz - LDS_WAVE / (64 * b->tpl * (int)sizeof(float)));
  if (b->bmax > 0) {
    pi = prof_begin(b->prof, KC_FUSED, st);
    launch_step(b, sa, in, out, b->rm[b->cur], b->rm[1 - b->cur], d, st);
    prof_end(b->prof, pi, st);
  }
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

static LArgs largs(const emstr_batch *b, const emstr_dev *d) {
  LArgs a;
  a.plane = d->plane;
  a.P2 = b->P2;
  a.vol = b->vol;
  a.nx = d->nx;
  a.nxp = b->nxp;
  a.nwet = d->nwet;
  a.tpl = b->tpl;
  return a;
}

extern "C" int emstr_batch_put(struct emstr_batch *b, const emstr_dev *d, int n,
                               const double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_put2<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      largs(b, d), d->wsp, d->wdn, n, sparse, b->tr[b->cur], b->tr[1 - b->cur], b->rm[b->cur],
      b->rm[1 - b->cur]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_batch_get(struct emstr_batch *b, const emstr_dev *d, int n,
                               double *sparse, void *stream) {
  if (d->nwet == 0) return 0;
  k_get2<<<(d->nwet + 255) / 256, 256, 0, (hipStream_t)stream>>>(largs(b, d), d->wsp, d->wdn,
                                                                  n, b->tr[b->cur],
                                                                  b->rm[b->cur], sparse);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int emstr_batch_mass(struct emstr_batch *b, const emstr_dev *d, double *part,
                                void *stream) {
  int nb = (d->ncol + 255) / 256;
  if (nb == 0) return 0;
  k_mass2<<<dim3(nb, d->ntr, 1), 256, 0, (hipStream_t)stream>>>(*d, largs(b, d), b->nz,
                                                                 b->tr[b->cur], b->rm[b->cur],
                                                                 part);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
