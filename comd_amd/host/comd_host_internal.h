/* comd_host_internal.h -- helpers shared by the host library sources. */
#ifndef COMD_HOST_INTERNAL_H
#define COMD_HOST_INTERNAL_H

#include "comd_host.h"

/* record an error for comd_host_status()/comd_host_error(); returns code */
int comd_host_fail(int code, const char *fmt, ...);
/* drop the device context bound to a simulation */
void comd_host_forget(simflat_t *s);

#endif
