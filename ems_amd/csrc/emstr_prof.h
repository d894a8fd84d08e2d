/* Per-kernel timing with HIP events recorded on the launching stream
 * around each launch (enabled with emstr_profile()); summed per kernel
 * class when read.  Shared by the host C layer and the .hip sources. */
#ifndef EMSTR_PROF_H
#define EMSTR_PROF_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMSTR_PROF_MAX 8192
/* kernel classes */
#define KC_ADV 0   /* advection (face fluxes + update + surface)     */
#define KC_VDIFF 1 /* vertical diffusion                             */
#define KC_PREP 2  /* per-step coefficient / matrix preparation       */
#define KC_OTHER 3 /* small per-column kernels                        */
#define KC_FUSED 4 /* fused advection + vertical diffusion            */
#define KC_N 5

typedef struct emstr_prof {
  int on;
  int n;
  int cls[EMSTR_PROF_MAX];
  hipEvent_t ev[2 * EMSTR_PROF_MAX];
  int created;
  double ms[KC_N];
  long long cnt[KC_N];
} emstr_prof;

static inline int prof_begin(emstr_prof *p, int cls, hipStream_t st) {
  if (!p || !p->on || p->n >= EMSTR_PROF_MAX) return -1;
  int i = p->n;
  if (i >= p->created) {
    hipEventCreate(&p->ev[2 * i]);
    hipEventCreate(&p->ev[2 * i + 1]);
    p->created = i + 1;
  }
  p->cls[i] = cls;
  hipEventRecord(p->ev[2 * i], st);
  return i;
}

static inline void prof_end(emstr_prof *p, int i, hipStream_t st) {
  if (i < 0) return;
  hipEventRecord(p->ev[2 * i + 1], st);
  p->n = i + 1;
}

/* Synchronise and fold the recorded pairs into ms[]/cnt[]. */
static inline void prof_collect(emstr_prof *p) {
  int i;
  if (!p) return;
  for (i = 0; i < p->n; i++) {
    float t = 0.0f;
    hipEventSynchronize(p->ev[2 * i + 1]);
    hipEventElapsedTime(&t, p->ev[2 * i], p->ev[2 * i + 1]);
    p->ms[p->cls[i]] += t;
    p->cnt[p->cls[i]]++;
  }
  p->n = 0;
}

static inline void prof_free(emstr_prof *p) {
  int i;
  if (!p) return;
  for (i = 0; i < p->created; i++) {
    hipEventDestroy(p->ev[2 * i]);
    hipEventDestroy(p->ev[2 * i + 1]);
  }
  p->created = 0;
}

#ifdef __cplusplus
}
#endif
#endif
