/* Fixed device scratch (emstr_scratch.hip) for the entry point that
 * allocates nothing. */
#ifndef EMSTR_SCRATCH_H
#define EMSTR_SCRATCH_H

/* 4 MiB of doubles on each device */
#define EMSTR_SCRATCH_DOUBLES (1 << 19)

#ifdef __cplusplus
extern "C" {
#endif

/* The scratch array's address on the current device (taken with
 * hipGetSymbolAddress on the first call on that device); NULL with the
 * error recorded on failure. */
double *emstr_scratch_ptr(void);
/* out[n] = compensated block-order sum of part's (sum, err) pairs. */
int emstr_launch_mass_reduce(const double *part, int nb, int ntr, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
