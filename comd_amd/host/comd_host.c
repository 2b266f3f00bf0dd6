/*
 * comd_host.c -- the reference's hot-path entry points (force slots,
 * computeForce, nTimeSteps, reBoxAll) and driver (parseCommandLine,
 * initSimFromCmdLine, do_compute_work) on top of the GPU C-ABI.
 */
#include <getopt.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "comd_host.h"
#include "comd_host_internal.h"

static int g_status = 0;
static char g_error[512] = "";

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

static int pot_kind(const simflat_t *s)
{
    if (s->pot->force == LJ || s->pot->force == comd_hip_lj_force) return COMD_POT_LJ;
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
int eamForce(void *s) { return force_adapter(s, COMD_POT_EAM); }
int LJ(void *s) { return force_adapter(s, COMD_POT_LJ); }

int computeForce(simflat_t *s) { return s->pot->force(s); }

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

double nTimeSteps(int n, simflat_t *s, real_t dt) { return comd_hip_nTimeSteps(n, s, dt); }

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

void reBoxAll(simflat_t *s) { comd_hip_reBoxAll(s); }

/* ---------------------------------------------------------------------
 * driver */

static void usage(const char *exe)
{
    printf("usage: %s [options]\n"
           "  -f, --infile FILE      gz atom file (reference format)\n"
           "  -d, --potdir DIR       EAM potential directory\n"
           "  -p, --potname NAME     EAM potential name (e.g. cu500)\n"
           "  -e, --doeam            EAM instead of Lennard-Jones\n"
           "  -o, --noperiodic       no periodic boundaries\n"
           "  -x/-y/-z N             unit cells in x/y/z\n"
           "  -b, --bf F             box factor\n"
           "  -s, --defgrad F        deformation gradient (x stretch)\n"
           "  -t, --temp K           initial temperature (Maxwell-Boltzmann)\n"
           "      --disp A           Gaussian displacement sigma\n"
           "      --seed N           seed of the initial state\n"
           "      --precision fp64|mixed\n"
           "      --eval table|cheby\n"
           "      --niter N --nsteps N   driver loop (default 20 x 10)\n"
           "      --lj SIGMA,EPS,CUTOFF,MASS   Lennard-Jones parameters\n",
           exe);
}

int parseCommandLine(command_t *cmd, int argc, char **argv)
{
    memset(cmd, 0, sizeof *cmd);
    strcpy(cmd->potdir, "pots");
    strcpy(cmd->potname, "ag");
    cmd->periodic = 1;
    cmd->nx = cmd->ny = cmd->nz = 20;
    cmd->bf = 1.0;
    cmd->defgrad = 1.0;
    cmd->niter = 20;
    cmd->nsteps = 10;
    cmd->seed = 1;
    static struct option opts[] = {
        {"help", no_argument, 0, 'h'},      {"infile", required_argument, 0, 'f'},
        {"potdir", required_argument, 0, 'd'}, {"potname", required_argument, 0, 'p'},
        {"doeam", no_argument, 0, 'e'},     {"noperiodic", no_argument, 0, 'o'},
        {"usegpu", no_argument, 0, 'g'},    {"nx", required_argument, 0, 'x'},
        {"ny", required_argument, 0, 'y'},  {"nz", required_argument, 0, 'z'},
        {"bf", required_argument, 0, 'b'},  {"defgrad", required_argument, 0, 's'},
        {"temp", required_argument, 0, 't'}, {"precision", required_argument, 0, 1},
        {"eval", required_argument, 0, 2},  {"seed", required_argument, 0, 3},
        {"disp", required_argument, 0, 4},  {"niter", required_argument, 0, 5},
        {"nsteps", required_argument, 0, 6}, {"lj", required_argument, 0, 7},
        {0, 0, 0, 0}};
    optind = 1;
    int c;
    while ((c = getopt_long(argc, argv, "hf:d:p:eogx:y:z:b:s:t:", opts, NULL)) != -1) {
        switch (c) {
        case 'h': usage(argv[0]); return 2;
        case 'f': snprintf(cmd->filename, sizeof cmd->filename, "%s", optarg); break;
        case 'd': snprintf(cmd->potdir, sizeof cmd->potdir, "%s", optarg); break;
        case 'p': snprintf(cmd->potname, sizeof cmd->potname, "%s", optarg); break;
        case 'e': cmd->doeam = 1; break;
        case 'o': cmd->periodic = 0; break;
        case 'g': cmd->usegpu = 1; break;
        case 'x': cmd->nx = atoi(optarg); break;
        case 'y': cmd->ny = atoi(optarg); break;
        case 'z': cmd->nz = atoi(optarg); break;
        case 'b': cmd->bf = atof(optarg); break;
        case 's': cmd->defgrad = atof(optarg); break;
        case 't': cmd->temp = atof(optarg); break;
        case 1: cmd->precision = strcmp(optarg, "mixed") == 0 ? COMD_PREC_MIXED : COMD_PREC_FP64; break;
        case 2: cmd->cheby = strcmp(optarg, "cheby") == 0; break;
        case 3: cmd->seed = strtoull(optarg, NULL, 10); break;
        case 4: cmd->disp = atof(optarg); break;
        case 5: cmd->niter = atoi(optarg); break;
        case 6: cmd->nsteps = atoi(optarg); break;
        case 7:
            if (sscanf(optarg, "%lf,%lf,%lf,%lf", &cmd->lj_sigma, &cmd->lj_epsilon, &cmd->lj_cutoff, &cmd->lj_mass) != 4)
                return comd_host_fail(COMD_ERR_ARG, "--lj needs SIGMA,EPS,CUTOFF,MASS");
            break;
        default: usage(argv[0]); return COMD_ERR_ARG;
        }
    }
    PERIODIC = cmd->periodic;
    return COMD_OK;
}

simflat_t *initSimFromCmdLine(int argc, char **argv)
{
    command_t cmd;
    if (parseCommandLine(&cmd, argc, argv)) return NULL;
    pmd_base_potential_t *pot;
    if (cmd.doeam) {
        pot = setEamPot(cmd.potdir, cmd.potname);
        if (!pot) return NULL;
        cmd.lat = ((eampotential_t *)pot)->lat;  /* src-flat/docompute.c:101 */
    } else {
        ljpotential_t *lj = getLJPot();
        if (cmd.lj_sigma > 0) {
            lj->sigma = cmd.lj_sigma;
            lj->epsilon = cmd.lj_epsilon;
            lj->cutoff = cmd.lj_cutoff;
            lj->mass = cmd.lj_mass;
        }
        pot = (pmd_base_potential_t *)lj;
        cmd.lat = 1.122462048 * lj->cutoff;  /* src-flat/docompute.c:117 */
        if (cmd.lj_sigma > 0) cmd.lat = 1.4142135623730951 * 1.122462048309373 * lj->sigma;
    }
    simflat_t *sim = cmd.filename[0] ? fromFileGzip(cmd, pot) : create_fcc_lattice(cmd, pot);
    if (!sim) return NULL;
    if (cmd.doeam) sim->ch_pot = setChebPot((eampotential_t *)pot, 32);  /* docompute.c:133 */
    binding *b = bind_sim(sim, 0);
    if (!b) return NULL;
    if (comd_host_set_precision(sim, cmd.precision) || (cmd.cheby && comd_host_set_eval(sim, 1))) return NULL;
    if (upload_b(b, sim)) return NULL;
    b->resident = 1;
    reBoxAll(sim);
    printf("    total atoms is: %d\n", sim->ntot);
    return sim;
}

static double wall(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

/* src-flat/docompute.c:166-218: one force, then niter calls of
 * nTimeSteps(nsteps, 1 fs), printing energy and the timing per atom. */
void *do_compute_work(SimThreadData *data)
{
    simflat_t *s = data->sim;
    const int niter = data->niter > 0 ? data->niter : 20;
    const int nsteps = data->nsteps > 0 ? data->nsteps : 10;
    const double dt = 1.0e-15;
    binding *b = bind_sim(s, 0);
    double ts = wall();
    computeForce(s);
    if (b) comd_sync(b->ctx);
    double te = wall();
    double ke = 0;
    for (int iter = 0; iter < niter; iter++) {
        const double ns = iter == 0 ? 1.0 : (double)(1 + nsteps);
        if (b) comd_energy(b->ctx, NULL, &ke, NULL);
        printf(" %3d %30.20g computed in %.3fs (%8.4f us/atom for %d atoms)  ke %.12g\n", iter * nsteps, s->e, te - ts,
               1.0e6 * (te - ts) / ns / (double)s->ntot, s->ntot, ke);
        printf("Virial stress = %g\n", s->stress);
        ts = wall();
        nTimeSteps(nsteps, s, dt);
        te = wall();
    }
    printf(" %3d %30.20f computed in %.3fs (%8.4f us/atom for %d atoms)\n", niter * nsteps, s->e, te - ts,
           1.0e6 * (te - ts) / (double)(nsteps + 1) / (double)s->ntot, s->ntot);
    if (b && b->resident) comd_host_download(s);
    return NULL;
}
