/* Internal helpers shared by the host-side C sources of libemstr. */
#ifndef EMSTR_INTERNAL_H
#define EMSTR_INTERNAL_H

#include "emstr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Record an error (thread-local); returns code. */
int emstr_set_error(int code, int cell, const char *fmt, ...);
void emstr_clear_error(void);

#ifdef __cplusplus
}
#endif
#endif
