/* Thread-local error record for the C-ABI (the reference exits through
 * hd_quit(); this library reports instead). */
#include <stdarg.h>
#include <stdio.h>

#include "emstr_internal.h"

static __thread char err_text[512];
static __thread int err_cell = -1;

int emstr_set_error(int code, int cell, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err_text, sizeof(err_text), fmt, ap);
  va_end(ap);
  err_cell = cell;
  return code;
}

void emstr_clear_error(void) {
  err_text[0] = 0;
  err_cell = -1;
}

const char *emstr_last_error(void) { return err_text; }

int emstr_last_error_cell(void) { return err_cell; }
