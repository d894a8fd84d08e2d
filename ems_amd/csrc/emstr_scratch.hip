/*
 * Fixed device scratch for the entry point that allocates nothing
 * (emstr_tracer_mass_noalloc): the per-block (sum, error) partials of the
 * compensated mass sum and the per-tracer totals live in a file-scope
 * __device__ array instead of a buffer made with hipMalloc.  Each GPU the
 * library runs on holds its own copy of the array until the process ends.
 *
 * In host code the array's name is only a stand-in, not a device address;
 * its address on each device comes from hipGetSymbolAddress, taken once per
 * device with that device current, and only that pointer is handed to
 * kernels and copies.  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (two host threads, or two streams) race.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "emstr_internal.h"
#include "emstr_scratch.h"

static __device__ __align__(256) double scratch[EMSTR_SCRATCH_DOUBLES];

#define MAXDEV 64
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;
static double *scratch_dev[MAXDEV]; /* the array's address on each device */

extern "C" double *emstr_scratch_ptr(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAXDEV) {
    emstr_set_error(EMSTR_EDEVICE, -1, "scratch: no current device");
    return NULL;
  }
  pthread_mutex_lock(&scratch_lock);
  double *p = scratch_dev[dev];
  if (!p) {
    void *q = NULL;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q) {
      pthread_mutex_unlock(&scratch_lock);
      emstr_set_error(EMSTR_EDEVICE, -1, "scratch: hipGetSymbolAddress: %s",
                      hipGetErrorString(e));
      return NULL;
    }
    p = (double *)q;
    scratch_dev[dev] = p;
  }
  pthread_mutex_unlock(&scratch_lock);
  return p;
}

/* Per tracer, the compensated sum of the block (sum, err) pairs in block
 * order: the same operations, in the same order, as the host reduction of
 * emstr_tracer_mass, so both give the same bits. */
__global__ void k_mass_reduce(const double *__restrict__ part, int nb, int ntr,
                              double *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ntr) return;
  const double *p = part + 2 * (long long)n * nb;
  double s = 0.0, e = 0.0;
  for (int b = 0; b < nb; b++) {
    const double a = p[2 * b];
    const double t = s + a;
    const double bb = t - s;
    e += (s - (t - bb)) + (a - bb);
    s = t;
    e += p[2 * b + 1];
  }
  out[n] = s + e;
}

extern "C" int emstr_launch_mass_reduce(const double *part, int nb, int ntr, double *out,
                                        void *stream) {
  if (ntr <= 0) return 0;
  k_mass_reduce<<<(ntr + 63) / 64, 64, 0, (hipStream_t)stream>>>(part, nb, ntr, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
