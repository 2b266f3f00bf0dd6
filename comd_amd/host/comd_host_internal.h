/* comd_host_internal.h -- helpers shared by the host library sources. */
#ifndef COMD_HOST_INTERNAL_H
#define COMD_HOST_INTERNAL_H

#include "comd_host.h"

/* record an error for comd_host_status()/comd_host_error(); returns code */
int comd_host_fail(int code, const char *fmt, ...);
/* drop the device context bound to a simulation */
void comd_host_forget(simflat_t *s);
/* bind a simulation of potential kind (COMD_POT_*) to a device context,
 * created on first use; status */
int comd_dropin_bind(simflat_t *s, int kind);
/* device-resident mode (the device holds the authoritative state; the
 * adapters then skip their per-call upload and download); turning it on
 * uploads the host arrays once */
int comd_dropin_set_resident(simflat_t *s, int on);
int comd_dropin_resident(const simflat_t *s);
/* the LJ force slot of libcomd_host, so that the adapters recognise its
 * simulations */
void comd_dropin_register_lj(int (*slot)(void *));

#endif
