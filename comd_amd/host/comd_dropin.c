/*
 * comd_dropin.c -- the drop-in adapters (libcomd_dropin.so): the force slot
 * `int (*force)(void *s)` (src-flat/include/pmdTypes.h:77), nTimeSteps
 * (src-flat/timestep.c:114) and reBoxAll (src-flat/domains.c:347) of a
 * simflat_t run on the GPU through the C-ABI of comd_hip.h.
 *
 * This library defines none of the reference's own function names, so a
 * program built from the reference's sources can link it beside its own
 * objects and store comd_hip_eam_force / comd_hip_lj_force in its
 * potential's force slot.  The one global it shares is PERIODIC
 * (src-flat/mycommand.c:67), defined here for programs that do not define
 * it; a reference build's own definition takes its place at load time.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "comd_host.h"
#include "comd_host_internal.h"

int PERIODIC = 1;

static int g_status = 0;
static char g_error[512] = "";

/* Records (code, message) for comd_host_status / comd_host_error; code 0
 * clears the status. */
int comd_host_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    g_status = code;
    return code;
}

int comd_host_status(void) { return g_status; }
const char *comd_host_error(void) { return g_error; }

/* ---------------------------------------------------------------------
 * simulation <-> device context bindings.  The reference is single-threaded
 * and non-reentrant; a small table is enough. */

#define NBIND 64
typedef struct binding {
    const void *sim;
    comd_ctx *ctx;
    int kind;       /* COMD_POT_* */
    int precision;
    int eval;
    int resident;   /* device holds the authoritative state */
} binding;
static binding g_bind[NBIND];

static binding *find_binding(const void *s)
{
    for (int i = 0; i < NBIND; i++)
        if (g_bind[i].sim == s) return &g_bind[i];
    return NULL;
}

void comd_host_forget(simflat_t *s)
{
    binding *b = find_binding(s);
    if (!b) return;
    comd_destroy(b->ctx);
    memset(b, 0, sizeof *b);
}

/* The LJ force slot of a caller other than the drop-in adapter itself
 * (libcomd_host's LJ), registered by that library. */
static int (*g_lj_slot)(void *) = NULL;
void comd_dropin_register_lj(int (*slot)(void *)) { g_lj_slot = slot; }

static int pot_kind(const simflat_t *s)
{
    if (s->pot->force == comd_hip_lj_force || (g_lj_slot && s->pot->force == g_lj_slot)) return COMD_POT_LJ;
    return COMD_POT_EAM;
}

int comd_host_params(const simflat_t *s, comd_params *p)
{
    memset(p, 0, sizeof *p);
    p->potential = pot_kind(s);
    p->periodic = PERIODIC;
    p->cutoff = s->pot->cutoff;
    p->mass = s->pot->mass;
    for (int j = 0; j < 3; j++) {
        p->nbx[j] = s->nbx[j];
        p->bounds[j] = s->bounds[j];
        p->boxsize[j] = s->boxsize[j];
    }
    if (p->potential == COMD_POT_LJ) {
        const ljpotential_t *lj = (const ljpotential_t *)s->pot;
        p->lj_sigma = lj->sigma;
        p->lj_epsilon = lj->epsilon;
        return COMD_OK;
    }
    const eampotential_t *e = (const eampotential_t *)s->pot;
    const potentialarray_t *t[3] = {e->phi, e->rho, e->f};
    for (int k = 0; k < 3; k++) {
        if (!t[k]) return comd_host_fail(COMD_ERR_ARG, "EAM table missing");
        p->tab_n[k] = t[k]->n;
        p->tab_x0[k] = t[k]->x0;
        p->tab_xn[k] = t[k]->xn;
        p->tab_invdx[k] = t[k]->invDx;
        p->tab_values[k] = t[k]->values - 1;
    }
    return COMD_OK;
}

static double g_chebbuf[6 * 64];

static binding *bind_sim(simflat_t *s, int kind_hint)
{
    binding *b = find_binding(s);
    if (b) return b;
    b = find_binding(NULL);
    if (!b) {
        comd_host_fail(COMD_ERR_ARG, "too many simulations");
        return NULL;
    }
    comd_params p;
    if (comd_host_params(s, &p)) return NULL;
    if (kind_hint) p.potential = kind_hint;
    if (p.potential == COMD_POT_EAM && s->ch_pot && s->ch_pot->phi && s->ch_pot->phi->n <= 64) {
        const eam_cheby_t *ch = s->ch_pot;
        const potentialarray_t *c[6] = {ch->phi, ch->rho, ch->f, ch->dphi, ch->drho, ch->df};
        const int m = ch->phi->n;
        for (int k = 0; k < 6; k++) memcpy(g_chebbuf + k * m, c[k]->values, (size_t)m * sizeof(double));
        p.cheb_n = m;
        p.cheb_coef = g_chebbuf;
        p.cheb_x0[0] = ch->phi->x0; p.cheb_xn[0] = ch->phi->xn;
        p.cheb_x0[1] = ch->rho->x0; p.cheb_xn[1] = ch->rho->xn;
        p.cheb_x0[2] = ch->f->x0;   p.cheb_xn[2] = ch->f->xn;
    }
    comd_ctx *ctx = NULL;
    int rc = comd_create(&p, &ctx);
    if (rc) {
        comd_host_fail(rc, "%s", comd_last_error(NULL));
        return NULL;
    }
    b->sim = s;
    b->ctx = ctx;
    b->kind = p.potential;
    b->precision = COMD_PREC_FP64;
    b->eval = COMD_EVAL_TABLE;
    b->resident = 0;
    return b;
}

comd_ctx *comd_host_ctx(simflat_t *s)
{
    binding *b = bind_sim(s, 0);
    return b ? b->ctx : NULL;
}

int comd_host_set_precision(simflat_t *s, int precision)
{
    binding *b = bind_sim(s, 0);
    if (!b) return g_status;
    b->precision = precision;
    return comd_set_precision(b->ctx, precision);
}

int comd_host_set_eval(simflat_t *s, int cheby)
{
    binding *b = bind_sim(s, 0);
    if (!b) return g_status;
    b->eval = cheby ? COMD_EVAL_CHEB : COMD_EVAL_TABLE;
    int rc = comd_set_eval(b->ctx, b->eval);
    return rc ? comd_host_fail(rc, "%s", comd_last_error(b->ctx)) : COMD_OK;
}

static int upload_b(binding *b, simflat_t *s)
{
    int rc = comd_upload(b->ctx, s->natoms, s->id, s->iType, &s->r[0][0], &s->p[0][0]);
    if (rc) return comd_host_fail(rc, "%s", comd_last_error(b->ctx));
    return COMD_OK;
}

static int download_b(binding *b, simflat_t *s)
{
    int rc = comd_download(b->ctx, s->natoms, s->id, s->iType, &s->r[0][0], &s->p[0][0], &s->f[0][0], s->rho, s->fi);
    if (rc) return comd_host_fail(rc, "%s", comd_last_error(b->ctx));
    int tot = 0;
    for (int c = 0; c < s->nboxes; c++) {
        tot += s->natoms[c];
        for (int i = 0; i < s->natoms[c]; i++) s->mass[c * MAXATOMS + i] = s->pot->mass;
    }
    s->ntot = tot;
    return COMD_OK;
}

static int totals_b(binding *b, simflat_t *s)
{
    double pe = 0, st = 0;
    int rc = comd_energy(b->ctx, &pe, NULL, &st);
    if (rc) return comd_host_fail(rc, "%s", comd_last_error(b->ctx));
    s->e = pe;
    s->stress = st;
    return COMD_OK;
}

int comd_host_upload(simflat_t *s)
{
    binding *b = bind_sim(s, 0);
    return b ? upload_b(b, s) : g_status;
}

int comd_host_download(simflat_t *s)
{
    binding *b = bind_sim(s, 0);
    if (!b) return g_status;
    int rc = download_b(b, s);
    return rc ? rc : totals_b(b, s);
}

/* force slot: mirror the host arrays, compute on the device, write back */
static int force_adapter(void *sv, int kind)
{
    simflat_t *s = (simflat_t *)sv;
    binding *b = bind_sim(s, kind);
    if (!b) return g_status;
    int rc;
    if (!b->resident && (rc = upload_b(b, s))) return rc;
    if ((rc = comd_force(b->ctx))) return comd_host_fail(rc, "%s", comd_last_error(b->ctx));
    if ((rc = totals_b(b, s))) return rc;
    if (!b->resident) return download_b(b, s);
    return COMD_OK;
}

int comd_hip_eam_force(void *s) { return force_adapter(s, COMD_POT_EAM); }
int comd_hip_lj_force(void *s) { return force_adapter(s, COMD_POT_LJ); }

double comd_hip_nTimeSteps(int n, void *sv, double dt)
{
    simflat_t *s = (simflat_t *)sv;
    binding *b = bind_sim(s, 0);
    if (!b) return 0.0;
    if (!b->resident && upload_b(b, s)) return 0.0;
    double pe = 0.0;
    int rc = comd_nsteps(b->ctx, n, dt, &pe);
    if (rc) {
        comd_host_fail(rc, "%s", comd_last_error(b->ctx));
        return 0.0;
    }
    totals_b(b, s);
    if (!b->resident) download_b(b, s);
    return s->e;
}

int comd_hip_reBoxAll(void *sv)
{
    simflat_t *s = (simflat_t *)sv;
    binding *b = bind_sim(s, 0);
    if (!b) return g_status;
    int rc;
    if (!b->resident && (rc = upload_b(b, s))) return rc;
    if ((rc = comd_rebox(b->ctx))) return comd_host_fail(rc, "%s", comd_last_error(b->ctx));
    if (!b->resident) return download_b(b, s);
    return COMD_OK;
}

/* ---------------------------------------------------------------------
 * for libcomd_host's driver */

int comd_dropin_bind(simflat_t *s, int kind)
{
    return bind_sim(s, kind) ? COMD_OK : g_status;
}

int comd_dropin_set_resident(simflat_t *s, int on)
{
    binding *b = find_binding(s);
    if (!b) return comd_host_fail(COMD_ERR_STATE, "simulation not bound");
    if (on && !b->resident) {
        int rc = upload_b(b, s);
        if (rc) return rc;
    }
    b->resident = on ? 1 : 0;
    return COMD_OK;
}

int comd_dropin_resident(const simflat_t *s)
{
    binding *b = find_binding(s);
    return b ? b->resident : 0;
}
