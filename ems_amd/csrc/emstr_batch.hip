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
 * tracer-independent coefficients [nz][nyp][nxp][3]: A = (qW, fW, qS, fS)
 * the transports and van Leer factors of the west and south faces, B =
 * (qB, fB, 1/V, D) those of the bottom face, the inverse volume and the net
 * FP64 volume divergence, C = (gm, gp, 1/div, ud) the normalised column
 * matrix and its LU factors; the east, north and top faces are the west,
 * south and bottom faces of the neighbouring cells.  Per column a Col2
 * record (beds of the eight stencil neighbours, surfaces, diffusion
 * extent, flags, FP64 surface volume balance).
 *
 * Per step:
 *   k_cell2   one thread per cell: coefficients from the FP64 fields
 *             (tracers.c:496-578, 3841-3889 restated);
 *   k_col2    one thread per column: the column record and the LU factors
 *             of the column matrix;
 *   k_step2   one wave per (block of 64 columns, tracer): every lane walks
 *             its column from the bed to the top with a five-layer register
 *             window of the tracer, the horizontal neighbours of the layer
 *             read from L1/L2, and computes
 *               - van Leer face values (tracers.c:2097-2118) as anomalies
 *                 against the cell value and the sub-surface update
 *                 (tracers.c:888-925) in anomaly form
 *                   tr - (sum_f s_f (F_f - tr) q_f + tr D) / V,
 *                 kept in the wave's LDS slice,
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
 * formed in FP64 and rounded to the 32-bit state once per step,
 * stochastically, with a deterministic hash of (cell, tracer, step), so
 * sub-ulp increments do not bias the state or the tracer mass, and results
 * do not depend on the launch shape.  The vertical differences of the walk
 * and of the column solve carry the tails (head differences plus tail
 * differences); the horizontal stencil uses the heads, whose truncation is
 * common to the cell and its neighbours.
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
#define SWAVES 4 /* waves per workgroup: four blocks of a tile, one tracer */

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
  ll P2;           /* nxp * nyp                                          */
  ll vol;          /* nz * P2: one tracer                                */
  float *tr[2];    /* [ntr][nz][nyp][nxp], the 24-bit heads              */
  uint8_t *rm[2];  /* [ntr/4][nz][nyp][nxp][4], the 8-bit tails          */
  int cur;
  float4 *co;      /* [nz][nyp][nxp][3]                                  */
  Col2 *ci;        /* [nyp][nxp]                                         */
  float *scr;      /* refinement scratch, the tracer layout              */
  float *mintr, *maxtr;
  int *stepctr;
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
/* the factor of the limited term of the van Leer face value for
 * transport q with Courant number cn: 1 - cn if q > 0, 1 + cn if q < 0
 * (the limiter's factor 2 and the 1/2 of tracers.c:2108 cancel) */
__device__ __forceinline__ float vlfac(double q, double cn) {
  return (q > 0.0) ? (float)(1.0 - cn) : (q < 0.0) ? (float)(1.0 + cn) : 1.0f;
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
  float gm = 0.0f, gp = 0.0f;
  if (ksc >= 0 && k <= cs && cs - 1 >= kbc) {
    /* rows of the implicit vertical diffusion (tracers.c:3841-3889),
     * normalised by dzdt */
    const double dzdt = dz / dt;
    if (k > kbc) gm = (float)(-d.Kz[idx] / (0.5 * (dzb + dz)) / dzdt);
    if (k < cs) gp = (float)(-d.Kz[idx + P] / (0.5 * (dz + d.dz[idx + P])) / dzdt);
  }
  float4 *o = co + ((ll)k * a.P2 + (ll)j * a.nxp + i) * 3;
  o[0] = make_float4((float)qW, vlfac(qW, cW), (float)qS, vlfac(qS, cS));
  o[1] = make_float4((float)qB, vlfac(qB, cB), (dz != 0.0) ? (float)(1.0 / (area * dz)) : 0.0f,
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
/* Van Leer face value (tracers.c:2097-2118) for transport q through the
 * face between bm1 and c, as its anomaly F - ref.  For q < 0 the stencil
 * is mirrored, (bm2, bm1, c) -> (fp1, c, bm1), which leaves df unchanged
 * and negates the reciprocal, so both signs share one expression; f is
 * 1 - cn for q > 0 and 1 + cn for q < 0.  With q == 0 the flux term
 * vanishes whatever the value.  The upwind value's anomaly is an exact
 * FP32 difference for neighbouring values within a factor of two, so the
 * rounding left is relative to the limited term and the anomaly itself. */
__device__ __forceinline__ float vla(float bm2, float bm1, float c, float fp1, float q, float f,
                                     float ref) {
  const bool pos = q > 0.0f;
  const float p2 = pos ? bm2 : fp1, p1 = pos ? bm1 : c, p0 = pos ? c : bm1;
  const float df = (p1 - p2) * (p0 - p1);
  const float t = df * __builtin_amdgcn_rcpf(p0 - p2);
  return __builtin_fmaf(f, (df > 0.0f) ? t : 0.0f, p1 - ref);
}

/* The same with the tails of the values (e*, of the four stencil values
 * and of ref): every difference is the heads' difference plus the
 * tails', so that it keeps the state's 32 bits. */
__device__ __forceinline__ float vlad(float bm2, float bm1, float c, float fp1, float ebm2,
                                      float ebm1, float ec, float efp1, float q, float f,
                                      float ref, float eref) {
  const bool pos = q > 0.0f;
  const float p2 = pos ? bm2 : fp1, p1 = pos ? bm1 : c, p0 = pos ? c : bm1;
  const float e2 = pos ? ebm2 : efp1, e1 = pos ? ebm1 : ec, e0 = pos ? ec : ebm1;
  const float d12 = (p1 - p2) + (e1 - e2), d01 = (p0 - p1) + (e0 - e1);
  const float df = d12 * d01;
  const float t = df * __builtin_amdgcn_rcpf(d01 + d12);
  return __builtin_fmaf(f, (df > 0.0f) ? t : 0.0f, (p1 - ref) + (e1 - eref));
}

/*-------------------------------------------------------------------*/
/* deterministic stochastic rounding                                 */
/*-------------------------------------------------------------------*/
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

/* per-cell key: one multiply round over the cell index and the wave's
 * (tracer, step) hash */
__device__ __forceinline__ uint32_t srkey(uint32_t cellkey, uint32_t hl) {
  uint32_t x = cellkey ^ hl;
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  return x;
}

/* Compensated state: a value is held as an FP32 head h, the value
 * truncated to 24 significant bits, and an 8-bit tail r, its next eight
 * bits, so that the state keeps 32 significant bits:
 *   x = h + sign(h) r 2^(e - 31),  e the exponent of h.
 * enc rounds the FP64 value x to 32 bits stochastically (a uniform 21-bit
 * number added to the 21 bits dropped, which are then cleared) and splits
 * it; values below the FP32 normal range keep r = 0. */
__device__ __forceinline__ void enc(double x, uint32_t rnd, float &h, uint32_t &r) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  u += (unsigned long long)(rnd >> 11);
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
  int nz, ntr, nx;      /* nx: unpadded, for the rounding keys          */
  int nxp;
  ll P2, vol;
  double hmin, dt;
  ll P;                 /* unpadded plane of the FP64 fields            */
  int nzl;              /* layers of a wave's LDS slice                 */
};

/* raw buffer access with 32-bit byte offsets from a wave-uniform base
 * (resource word 3 of the gfx9 family); a read past the end returns 0 */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
/* o: the lane's byte offset; so: a wave-uniform byte offset (a scalar
 * register), added to it (the range check covers o only: the buffers
 * carry slack beyond their last element for the reads past it) */
__device__ __forceinline__ float bl1(__amdgpu_buffer_rsrc_t r, uint32_t o, uint32_t so = 0) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, o, so, 0));
}
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
/* DPP moves inside rows of 16 lanes: lane l takes lane l - 1, l - 2
 * (shr) or l + 1, l + 2 (shl); a lane whose source is outside its row
 * takes old */
__device__ __forceinline__ float shr1o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float shr2o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x112, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl1o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x101, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl2o(float old, float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x102, 0xf, 0xf, false));
}

/* the loads of one layer: the column at layer k + 2 (the window's new
 * entry), the four y neighbours, the x neighbours beyond the lane's row
 * segment (edge lanes only), the cell's west/south coefficient word, the
 * north neighbour's south transport, the bottom-face word of the cell
 * above (this cell's top face) and, for the segment's last lane, the east
 * neighbour's west transport */
struct Lay {
  float z2, s1, s2, n1, n2, x1, x2;
  uint32_t z2r; /* the tail of z2 */
  float4 A, U;
  float3 M; /* gm, gp, 1/div: the walk's part of the matrix word */
  float2 N, E;
};

/* One workgroup per (tile of 16 x 4 columns, group of 4 tracers): wave w
 * is row j0 + w of the tile; lane l is column i0 + l % 16 of that row for
 * tracer 4 g + l / 16.  The four tracer rows of a wave read the same
 * coefficient words (one request per cell), and a lane's x neighbours are
 * the lanes beside it (DPP).  Every lane walks its column from the row's
 * deepest bed to the top; the loads of layer k + 1 are issued before
 * layer k is computed; the LDS slice holds the advected values b of the
 * lane's column, the scratch buffer its forward-sweep values. */
__global__ void __launch_bounds__(64 * SWAVES)
k_step2(const SArgs a, const float *__restrict__ tin, float *__restrict__ tout,
        const uint8_t *__restrict__ rin, uint8_t *__restrict__ rout,
        const float4 *__restrict__ co, const Col2 *__restrict__ cinfo,
        const int *__restrict__ blk, const float *__restrict__ mintr,
        const float *__restrict__ maxtr, const double *__restrict__ layers,
        const double *__restrict__ Kz, const double *__restrict__ dzf,
        float *__restrict__ scr, const int *__restrict__ stepctr) {
  extern __shared__ float lds[];
  const int xcd = blockIdx.x & 7;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = blockIdx.x >> 3;
  const int NG = (a.ntr + 3) >> 2;
  const int bq = q / NG, g = q - bq * NG;
  /* blk: [0..7] first tile of each XCD, [8..15] tile counts, then per
   * tile (i0 | j0 << 16, the four rows' deepest beds as bytes) */
  if (bq >= blk[8 + xcd]) return;
  const int bi = blk[xcd] + bq;
  const int b0 = blk[16 + 2 * bi], kmin = (blk[17 + 2 * bi] >> (8 * wv)) & 255;
  if (kmin >= a.nz) return;
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15;
  const int nr = 4 * g + (lane >> 4);
  const bool live = nr < a.ntr;
  const int n = live ? nr : a.ntr - 1;
  const int i = (b0 & 0xffff) + cl, j = (b0 >> 16) + wv;
  const int nz = a.nz, top = nz - 1;
  const uint32_t P2 = (uint32_t)a.P2, nxp = (uint32_t)a.nxp;
  const uint32_t col = (uint32_t)j * nxp + (uint32_t)i;
  const Col2 ci = cinfo[col];
  const int kb = ci.kb, ks = ci.ks, flags = ci.flags;
  const float mn = mintr[n], mx = maxtr[n];
  const bool clip = mn <= 1e-5f;
  const uint32_t hl = mix32((uint32_t)n * 0x85EBCA77U ^ (uint32_t)stepctr[0] * 0xC2B2AE3DU);
  const uint32_t cellkey0 = (uint32_t)(((ll)j * a.nx + i) * nz);
  const __amdgpu_buffer_rsrc_t RT = rsrc(tin, (uint32_t)(a.vol * 4 * a.ntr));
  const __amdgpu_buffer_rsrc_t RC = rsrc(co, (uint32_t)(a.vol * 48));
  const __amdgpu_buffer_rsrc_t RR = rsrc(rin, (uint32_t)(a.vol * 4 * NG));
  float *OB = tout + (ll)n * a.vol;
  /* the tails of the lane's tracer: byte 4 (k P2 + col) of the group's words */
  uint8_t *OR = rout + (ll)g * a.vol * 4 + (lane >> 4);
  float *SB = scr + (ll)n * a.vol;
  const float4 *__restrict__ CC = co + col * 3 + 2; /* the column's matrix words */
  const float *__restrict__ CU = (const float *)CC + 3; /* their ud entries */
  /* byte offsets: of the lane's column in its tracer, of a layer, of the
   * column's coefficients in a layer, of a coefficient layer */
  const uint32_t c4 = (uint32_t)(n * a.vol * 4) + col * 4u, L4 = P2 * 4u;
  const uint32_t cr = (uint32_t)(g * a.vol * 4) + col * 4u + (uint32_t)(lane >> 4);
  /* the tail of the lane's cell at layer k, as a byte */
  auto rb = [&](int k) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(RR, (uint32_t)k * L4 + cr, 0, 0);
  };
  const uint32_t c48 = col * 48u, L48 = P2 * 48u;
  const int nw1 = ci.nbw & 255, nw2 = (ci.nbw >> 8) & 255, ne1 = (ci.nbw >> 16) & 255,
            ne2 = (ci.nbw >> 24) & 255;
  const int ns1 = ci.nbs & 255, ns2 = (ci.nbs >> 8) & 255, nn1 = (ci.nbs >> 16) & 255,
            nn2 = (ci.nbs >> 24) & 255;
  const bool proc = (flags & F_PROC) != 0;
  const bool move = (flags & F_MOVE) != 0;
  float *L = lds + wv * a.nzl * 64 + lane;
  const uint32_t ry = nxp * 4u;
  auto clampk = [&](int k) { return (uint32_t)min(max(k, kb), top); };
  /* the lane's offsets of its y neighbours and of the x neighbours beyond
   * its row segment (x1: lane 0 its west, lane 1 its second west, lane 14
   * its second east, lane 15 its east neighbour; x2: lanes 0 and 15 their
   * second neighbours; other lanes their own cell) */
  const uint32_t os1 = c4 - ry, os2 = c4 - 2 * ry, on1 = c4 + ry, on2 = c4 + 2 * ry;
  const uint32_t ox1 = (cl == 0) ? c4 - 4 : (cl == 1) ? c4 - 8 : (cl == 14) ? c4 + 8
                     : (cl == 15) ? c4 + 4 : c4;
  const uint32_t ox2 = (cl == 0) ? c4 - 8 : (cl == 15) ? c4 + 8 : c4;
  const uint32_t oN = c48 + ry * 12u + 8u;
  /* Issue the loads of layer k: every lane the same words at its own
   * offsets, the layer's offset uniform (below a lane's bed, above its
   * surface or at a neighbour that is land the words are read and not
   * used; a neighbour that is land or below the bed at layer k is read as
   * the cell the map walk came from, tracers.c:5574-5593, when the layer
   * is computed). */
  auto issue = [&](int k, Lay &y) {
    const uint32_t kz = (uint32_t)min(k + 2, top) * L4;
    y.z2 = bl1(RT, c4, kz);
    y.z2r = (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(RR, cr, kz, 0);
    const uint32_t kt = (uint32_t)k * L4, kc = (uint32_t)k * L48;
    y.A = bl4(RC, c48, kc);
    y.s1 = bl1(RT, os1, kt);
    y.s2 = bl1(RT, os2, kt);
    y.n1 = bl1(RT, on1, kt);
    y.n2 = bl1(RT, on2, kt);
    y.x1 = bl1(RT, ox1, kt);
    y.x2 = bl1(RT, ox2, kt);
    y.U = bl4(RC, c48 + 16u, k < top ? kc + L48 : kc);
    y.M = bl3(RC, c48 + 32u, kc);
    y.N = bl2(RC, oN, kc);
    y.E = bl2(RC, c48 + 48u, kc);
  };
  /* register window of the column: layers k-2 .. k+2, clamped to the bed
   * (the sediment ghost holds the bed value) and to the top */
  float zm2 = bl1(RT, clampk(kmin - 2) * L4 + c4), zm1 = bl1(RT, clampk(kmin - 1) * L4 + c4);
  float t0 = bl1(RT, clampk(kmin) * L4 + c4), zp1 = bl1(RT, clampk(kmin + 1) * L4 + c4);
  float zp2 = 0.0f;
  /* and their tails (the vertical differences carry them) */
  float dm2 = tail(zm2, rb((int)clampk(kmin - 2))), dm1 = tail(zm1, rb((int)clampk(kmin - 1)));
  float d00 = tail(t0, rb((int)clampk(kmin))), dp1 = tail(zp1, rb((int)clampk(kmin + 1)));
  float dp2 = 0.0f;
  /* the bed cell (the window's value below the bed) */
  const float tbed = bl1(RT, (uint32_t)min(kb, top) * L4 + c4);
  const uint32_t rbed = rb(min(kb, top));
  float4 Bk = bl4(RC, clampk(kmin) * L48 + c48 + 16u); /* this layer's bottom face */
  double m2 = 0.0;    /* surface cell: anomaly terms of the volume balance */
  float v0s = 0.0f;   /* surface cell: its old value                       */
  float v0d = 0.0f;   /* and its tail                                      */
  double bunc = 0.0;  /* unclipped advected value below the surface        */
  float aT = 0.0f;    /* top-face anomaly of the layer below, against it   */
  const int kold = ci.kold;
  /* the column solve's forward sweep runs in the walk for the rows whose
   * right-hand side lies below the surface: row k - 1 at layer k */
  const bool solve = proc && !(flags & F_SINGLE);
  const int cb = ci.cb, cs = ci.cs;
  float dd1 = 0.0f, dd2 = 0.0f; /* advective increments of the two layers below */
  double bks1 = 0.0, bks2 = 0.0; /* b of the two layers below the surface      */
  float sp = 0.0f;              /* s of the last forward row            */
  float3 Mp = make_float3(0.0f, 0.0f, 1.0f); /* matrix row below      */
  float *Ls = L;                /* s_k of the lane's column in its LDS slice */
  Lay ya = {}, yb = {};
  issue(kmin, ya);
  auto layer = [&](int k, const Lay &y) {
    const bool has = k >= kb;
    {
      const bool bed = k + 2 < kb;
      zp2 = bed ? tbed : y.z2;
      dp2 = tail(zp2, bed ? rbed : y.z2r);
    }
    /* x neighbours at layer k from the lanes beside this one, at the
     * segment's ends from the loads beyond it */
    const float l1 = shr1o(y.x1, t0), l2 = shr2o((cl == 0) ? y.x2 : y.x1, t0);
    const float r1 = shl1o(y.x1, t0), r2 = shl2o((cl == 15) ? y.x2 : y.x1, t0);
    /* the east face's transport and factor: the west word of the lane to
     * the east, or of the east neighbour's word at the segment's end */
    const float qEl = shl1o(y.E.x, y.A.x), fEl = shl1o(y.E.y, y.A.y);
    const bool sub = proc && has && k < ks;
    const bool srf = proc && k == ks;
    const bool upp = proc && move && k > ks;
    if (!proc && has && live) {
      OB[(uint32_t)k * P2 + col] = t0;
      OR[((uint32_t)k * P2 + col) * 4u] = (uint8_t)tailbyte(t0, d00);
    }
    if (sub || srf || upp) {
      const bool bw1 = k >= nw1, be1 = k >= ne1;
      const bool bw2 = bw1 && k >= nw2, be2 = be1 && k >= ne2;
      const bool bs1 = k >= ns1, bn1 = k >= nn1;
      const bool bs2 = bs1 && k >= ns2, bn2 = bn1 && k >= nn2;
      const float w1 = bw1 ? l1 : t0, w2 = bw2 ? l2 : w1;
      const float e1 = be1 ? r1 : t0, e2 = be2 ? r2 : e1;
      const float s1 = bs1 ? y.s1 : t0, s2 = bs2 ? y.s2 : s1;
      const float n1 = bn1 ? y.n1 : t0, n2 = bn2 ? y.n2 : n1;
      const float qE = be1 ? qEl : 0.0f, qN = bn1 ? y.N.x : 0.0f;
      /* (a neighbour that is land at layer k has a zero word: its factor
       * must stay finite, the transport is zero) */
      const float fE = be1 ? fEl : 1.0f, fN = bn1 ? y.N.y : 1.0f;
      const float ref = upp ? v0s : t0;
      const float AW = vla(w2, w1, t0, e1, y.A.x, y.A.y, ref);
      const float AE = vla(w1, t0, e1, e2, qE, fE, ref);
      const float AS = vla(s2, s1, t0, n1, y.A.z, y.A.w, ref);
      const float AN = vla(s1, t0, n1, n2, qN, fN, ref);
      const float hh = AE * qE - AW * y.A.x + AN * qN - AS * y.A.z;
      if (upp) {
        /* a layer above the processed surface cell (tracers.c:3117-3193) */
        m2 -= (double)hh;
        const double an = ((double)t0 + (double)d00) - ((double)v0s + (double)v0d);
        if (k < kold)
          m2 += an * (ci.area * (layers[k + 1] - layers[k]));
        else if (k == kold)
          m2 += an * (ci.area * (ci.osub - layers[k]));
      } else {
        /* the bottom face is the top face of the layer below, against
         * this cell's value; zero at the bed (the ghost holds the bed
         * value) */
        const float fba = (k > kb) ? (aT + ((zm1 - t0) + (dm1 - d00))) * Bk.x : 0.0f;
        if (sub) {
          aT = vlad(zm1, t0, zp1, zp2, dm1, d00, dp1, dp2, y.U.x, y.U.y, t0, d00);
          /* the advected value b = t0 + dlt: its increment is kept in
           * FP32 (the output column holds it until the column's values
           * are formed and rounded once, after the solve) */
          float dlt = -((hh + aT * y.U.x - fba + (t0 + d00) * Bk.w) * Bk.z);
          /* tr_bounds below the surface: the clip test in FP32 (a value
           * within an ulp of the minimum may go either way) */
          const float du = dlt;
          if (clip && t0 + (d00 + dlt) < mn) dlt = (mn - t0) - d00;
          if (live) OB[(uint32_t)k * P2 + col] = dlt;
          const int r = k - 1;
          if (solve && r >= cb) {
            /* r_k = -(gm (b_{k-1} - b_k) + gp (b_{k+1} - b_k)),
             * s_k = (r_k - gm s_{k-1}) / div_k, the differences of b
             * from those of t0 and of the increments */
            const float gm = (r > cb) ? Mp.x : 0.0f;
            const float dm = (r > cb) ? ((zm2 - zm1) + (dm2 - dm1)) + (dd2 - dd1) : 0.0f;
            const float dp = ((t0 - zm1) + (d00 - dm1)) + (dlt - dd1);
            const float rr = -(gm * dm + Mp.y * dp);
            sp = (rr - gm * sp) * Mp.z;
            Ls[(r - kmin) * 64] = sp;
          }
          if (k == ks - 1) {
            const double T = (double)t0 + (double)d00;
            bunc = T + (double)du;
            bks1 = T + (double)dlt;
            bks2 = ((double)zm1 + (double)dm1) + (double)dd1;
          }
          dd2 = dd1;
          dd1 = dlt;
        } else {
          m2 = (double)fba - (double)hh;
          v0s = t0;
          v0d = d00;
        }
      }
    }
    /* the next layer's bottom face: loaded with this layer (below the
     * bed, the bed cell's word loaded at the start stays); this layer's
     * matrix row for the next forward row */
    if (has) {
      Bk = y.U;
      Mp = y.M;
    }
    /* advance the window */
    zm2 = zm1;
    zm1 = t0;
    t0 = zp1;
    zp1 = zp2;
    dm2 = dm1;
    dm1 = d00;
    d00 = dp1;
    dp1 = dp2;
  };
  int k = kmin;
  for (; k + 1 <= top; k += 2) {
    issue(k + 1, yb);
    layer(k, ya);
    issue(k + 2, ya);
    layer(k + 1, yb);
  }
  if (k <= top) layer(k, ya);
  if (!proc || !live) return;

  /* surface value (surf_conc) in FP64, then tr_bounds at the surface */
  double surf;
  {
    surf = ((double)v0s + (double)v0d) * ci.vc + m2;
    if (ci.svol == 0.0 || (surf < 0.0 && ci.dzs <= a.hmin))
      surf = 0.0;
    else
      surf /= ci.svol;
    if (flags & F_KEEPOLD) surf = (double)v0s + (double)v0d;
    if ((flags & F_THIN) && (surf < (double)mn || surf > (double)mx)) surf = bunc;
    if (clip && surf < (double)mn) surf = (double)mn;
  }
  const uint32_t hv = hl ^ 0x5bd1e995U;
  /* a cell's new value, rounded once to the 32-bit state */
  auto put = [&](uint32_t o, double v, uint32_t key) {
    float h;
    uint32_t r;
    enc(v, key, h, r);
    OB[o] = h;
    OR[o * 4u] = (uint8_t)r;
  };
  /* the stored value of the lane's cell at layer kk */
  auto told = [&](int kk) { return dec(bl1(RT, (uint32_t)kk * L4 + c4), rb(kk)); };
  /* the layers above the surface outside the solved range take the
   * surface value, rounded once */
  float trs;
  uint32_t trr;
  enc(surf, srkey(cellkey0 + (uint32_t)ks, hv), trs, trr);
  const int lo = solve ? cb : top + 1, hi = solve ? cs : top;
  for (int kk = ks; kk <= top; kk++)
    if (kk < lo || kk > hi) {
      OB[(uint32_t)kk * P2 + col] = trs;
      OR[((uint32_t)kk * P2 + col) * 4u] = (uint8_t)trr;
    }
  /* cells below the surface outside the solved range: t0 + dlt, rounded */
  for (int kk = kb; kk < min(ks, lo); kk++) {
    const uint32_t o = (uint32_t)kk * P2 + col;
    put(o, told(kk) + (double)OB[o], srkey(cellkey0 + (uint32_t)kk, hv));
  }
  if (!solve) return;
  /* the remaining forward rows (their right-hand side reaches the
   * surface value), b below the surface from the walk's last two layers */
  {
    const int rs = max(cb, ks - 1);
    for (int r = rs; r <= cs; r++) {
      const float4 M = CC[(ll)r * P2 * 3];
      const float gm = (r > cb) ? M.x : 0.0f;
      const double b0 = (r >= ks) ? surf : bks1;
      const double bm = (r - 1 >= ks) ? surf : (r - 1 == ks - 1 ? bks1 : bks2);
      const float dm = (r > cb) ? (float)(bm - b0) : 0.0f;
      const float dp = (r < cs) ? (float)(surf - b0) : 0.0f;
      const float rr = -(gm * dm + M.y * dp);
      sp = (rr - gm * sp) * M.z;
      Ls[(r - kmin) * 64] = sp;
    }
  }
  /* Back substitution d_k = s_k - ud_{k+1} d_{k+1}, x = t0 + dlt + d
   * (or the surface value + d) in FP64, rounded once; eight rows at a
   * time, the t0, dlt and ud of the next eight rows loaded before this
   * chunk's results are stored over its dlt (d goes to the scratch column
   * in a stiff column) */
  const bool stiff = (flags & F_STIFF) != 0;
  {
    float dn = 0.0f, udn = 0.0f;
    float ta[8], da[8], ua[8], tb[8], db[8], ub[8];
    uint32_t ra[8], rb2[8];
#define EMSTR_FETCH(R0, T, RR8, D, U)                               \
  _Pragma("unroll") for (int jj = 0; jj < 8; jj++) {                \
    const int rc = max((R0) - jj, cb);                              \
    T[jj] = bl1(RT, (uint32_t)rc * L4 + c4);                        \
    RR8[jj] = rb(rc);                                               \
    D[jj] = OB[(uint32_t)rc * P2 + col];                            \
    U[jj] = CU[(ll)rc * P2 * 12];                                   \
  }
#define EMSTR_SOLVE(R0, T, RR8, D, U)                                              \
  _Pragma("unroll") for (int jj = 0; jj < 8; jj++) {                               \
    const int r = (R0) - jj;                                                       \
    if (r >= cb) {                                                                 \
      const uint32_t o = (uint32_t)r * P2 + col;                                   \
      const float sk = Ls[(r - kmin) * 64];                                        \
      dn = (r == cs) ? sk : sk - udn * dn;                                         \
      if (stiff) {                                                                 \
        SB[o] = dn;                                                                \
      } else {                                                                     \
        const double bfull = (r < ks) ? dec(T[jj], RR8[jj]) + (double)D[jj] : surf; \
        put(o, bfull + (double)dn, srkey(cellkey0 + (uint32_t)r, hv));             \
      }                                                                            \
      udn = U[jj];                                                                 \
    }                                                                              \
  }
    EMSTR_FETCH(cs, ta, ra, da, ua)
    for (int r0 = cs; r0 >= cb; r0 -= 16) {
      if (r0 - 8 >= cb) {
        EMSTR_FETCH(r0 - 8, tb, rb2, db, ub)
      }
      EMSTR_SOLVE(r0, ta, ra, da, ua)
      if (r0 - 8 < cb) break;
      if (r0 - 16 >= cb) {
        EMSTR_FETCH(r0 - 16, ta, ra, da, ua)
      }
      EMSTR_SOLVE(r0 - 8, tb, rb2, db, ub)
    }
#undef EMSTR_FETCH
#undef EMSTR_SOLVE
  }
  if (!stiff) return;
  /* one FP64 residual pass of iterative refinement, the matrix and the
   * right-hand side in FP64 from Kz and dz (tracers.c:3841-3889); the
   * correction is solved with the FP32 factors; dlt is in the output
   * column, d in the scratch column, the correction in the LDS slice */
  {
    const ll P = a.P;
    const ll fc = (ll)(j) * a.nx + i;
    const double dt = a.dt;
    auto bval = [=](int kk) {
      return (kk < ks) ? dec(bl1(RT, (uint32_t)kk * L4 + c4),
                             (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(
                                 RR, (uint32_t)kk * L4 + cr, 0, 0)) +
                             (double)OB[(uint32_t)kk * P2 + col]
                       : surf;
    };
    float e = 0.0f;
    for (int kk = cb; kk <= cs; kk++) {
      const ll ix = (ll)kk * P + fc;
      const uint32_t o = (uint32_t)kk * P2 + col;
      const double dzc = dzf[ix], dzdt = dzc / dt;
      double cm1 = 0.0, cp1 = 0.0;
      if (kk > cb) cm1 = -Kz[ix] / (0.5 * (dzf[ix - P] + dzc));
      if (kk < cs) cp1 = -Kz[ix + P] / (0.5 * (dzc + dzf[ix + P]));
      const double bv = bval(kk);
      const double dk = (double)SB[o];
      const double bvm = (kk > cb) ? bval(kk - 1) : bv;
      const double bvp = (kk < cs) ? bval(kk + 1) : bv;
      const double dkm = (kk > cb) ? (double)SB[o - P2] : 0.0;
      const double dkp = (kk < cs) ? (double)SB[o + P2] : 0.0;
      const double rhs = -(cm1 * (bvm - bv) + cp1 * (bvp - bv));
      const double Ad = cm1 * dkm + (dzdt - cm1 - cp1) * dk + cp1 * dkp;
      const float rho = (float)((rhs - Ad) / dzdt);
      const float4 M = CC[(ll)kk * P2 * 3];
      const float gm = (kk > cb) ? M.x : 0.0f;
      e = (rho - gm * e) * M.z;
      Ls[(kk - kmin) * 64] = e;
    }
    float ec = 0.0f, udn = 0.0f;
    for (int kk = cs; kk >= cb; kk--) {
      const float ek = Ls[(kk - kmin) * 64];
      ec = (kk == cs) ? ek : ek - udn * ec;
      udn = CC[(ll)kk * P2 * 3].w;
      Ls[(kk - kmin) * 64] = ec;
    }
    /* every residual is formed before any value is replaced */
    for (int kk = cb; kk <= cs; kk++) {
      const uint32_t o = (uint32_t)kk * P2 + col;
      const double bv = bval(kk);
      put(o, bv + (double)SB[o] + (double)Ls[(kk - kmin) * 64],
          srkey(cellkey0 + (uint32_t)kk, hv));
    }
  }
}

/* advance the stochastic-rounding step counter (runs after the step) */
__global__ void k_tick2(int *ctr) { *ctr += 1; }

/*-------------------------------------------------------------------*/
/* layout conversion and mass                                        */
/*-------------------------------------------------------------------*/
struct LArgs {
  ll plane, P2, vol;
  int nx, nxp, nwet;
};

__device__ __forceinline__ ll bidx(const LArgs &a, ll idx) {
  const int k = (int)(idx / a.plane);
  const ll col = idx - (ll)k * a.plane;
  const ll jj = col / a.nx, ii = col - jj * a.nx;
  return (ll)k * a.P2 + jj * a.nxp + ii;
}

/* byte offset of the tail of tracer n at cell index c of its volume */
__device__ __forceinline__ ll toff(const LArgs &a, int n, ll c) {
  return (((ll)(n >> 2) * a.vol + c) << 2) + (n & 3);
}

/* FP64 -> head and tail, rounded to nearest at 32 bits, into both buffers */
__global__ void k_put2(LArgs a, const int *__restrict__ wsp, const ll *__restrict__ wdn, int n,
                       const double *__restrict__ sp, float *__restrict__ t0,
                       float *__restrict__ t1, uint8_t *__restrict__ r0,
                       uint8_t *__restrict__ r1) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nwet) return;
  const ll c = bidx(a, wdn[q]), e = (ll)n * a.vol + c, o = toff(a, n, c);
  float h;
  uint32_t r;
  enc(sp[wsp[q]], 0x80000000u, h, r);
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
  sp[wsp[q]] = dec(t[(ll)n * a.vol + c], rm[toff(a, n, c)]);
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
    const float *t = tr + (ll)n * la.vol + c0;
    for (int k = ksc; k >= kbc; k--) {
      const double v = dec(t[(ll)k * la.P2], rm[toff(la, n, (ll)k * la.P2 + c0)]);
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
  if (b->vol * 48 > 0x7fffffffLL || b->vol * 4 * ((d->ntr + 3) / 4 * 4) > 0xffffffffLL) {
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
  const size_t nt = (size_t)b->vol * d->ntr;
  bool ok = true;
  const size_t nrm = (size_t)b->vol * 4 * ((d->ntr + 3) / 4);
  /* slack for the neighbour reads of the last layer past the end: two
   * rows of tracers, one row of coefficient words */
  const size_t trb = nt * sizeof(float) + (size_t)b->nxp * 8 + 256;
  const size_t cob = (size_t)b->vol * 3 * sizeof(float4) + (size_t)b->nxp * 48 + 256;
  for (int q = 0; q < 2; q++) ok &= hipMalloc(&b->tr[q], trb) == hipSuccess;
  for (int q = 0; q < 2; q++) ok &= hipMalloc(&b->rm[q], nrm) == hipSuccess;
  ok &= hipMalloc(&b->scr, nt * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->co, cob) == hipSuccess;
  ok &= hipMalloc(&b->ci, (size_t)b->P2 * sizeof(Col2)) == hipSuccess;
  ok &= hipMalloc(&b->mintr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->maxtr, (d->ntr + 1) * sizeof(float)) == hipSuccess;
  ok &= hipMalloc(&b->stepctr, sizeof(int)) == hipSuccess;
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
  hipMemset(b->stepctr, 0, sizeof(int));
  if (build_blocks(b, h_kb)) {
    emstr_batch_destroy(b);
    return NULL;
  }
  return b;
}

extern "C" void emstr_batch_destroy(struct emstr_batch *b) {
  if (!b) return;
  drop_graph(b);
  void *v[] = {b->tr[0], b->tr[1], b->rm[0], b->rm[1], b->scr, b->co, b->ci, b->mintr, b->maxtr, b->stepctr, b->blk};
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
  const size_t lds = (size_t)SWAVES * (b->nz - b->kmin) * 64 * sizeof(float);
  const int ng = (b->ntr + 3) / 4;
  k_step2<<<(unsigned)(8 * b->bmax * ng), 64 * SWAVES, lds, st>>>(
      sa, in, out, rin, rout, b->co, b->ci, b->blk, b->mintr, b->maxtr, d->layers, d->Kz, d->dz, b->scr,
      b->stepctr);
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
  sa.nzl = b->nz - b->kmin;
  if (b->bmax > 0) {
    pi = prof_begin(b->prof, KC_FUSED, st);
    launch_step(b, sa, in, out, b->rm[b->cur], b->rm[1 - b->cur], d, st);
    prof_end(b->prof, pi, st);
  }
  k_tick2<<<1, 1, 0, st>>>(b->stepctr);
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
