/*
 * comd_host.c -- the reference's hot-path entry points (force slots,
 * computeForce, nTimeSteps, reBoxAll) and driver (parseCommandLine,
 * initSimFromCmdLine, do_compute_work) on top of the drop-in adapters
 * (comd_dropin.c) and the GPU C-ABI.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "comd_host.h"
#include "comd_host_internal.h"

int eamForce(void *s) { return comd_hip_eam_force(s); }
int LJ(void *s) { return comd_hip_lj_force(s); }

/* the drop-in library tells LJ simulations by their force slot */
__attribute__((constructor)) static void register_lj_slot(void) { comd_dropin_register_lj(LJ); }

int computeForce(simflat_t *s) { return s->pot->force(s); }

static int kind_of(const simflat_t *s)
{
    return (s->pot->force == LJ || s->pot->force == comd_hip_lj_force) ? COMD_POT_LJ : COMD_POT_EAM;
}

double nTimeSteps(int n, simflat_t *s, real_t dt)
{
    if (comd_dropin_bind(s, kind_of(s))) return 0.0;
    return comd_hip_nTimeSteps(n, s, dt);
}

void reBoxAll(simflat_t *s)
{
    if (comd_dropin_bind(s, kind_of(s)) == COMD_OK) comd_hip_reBoxAll(s);
}

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

static comd_run_options g_opts = {COMD_PREC_FP64, 0, 1, 0.0, 20, 10, 0.0, 0.0, 0.0, 0.0};

const comd_run_options *comd_host_run_options(void) { return &g_opts; }
void comd_host_set_run_options(const comd_run_options *o)
{
    if (o) g_opts = *o;
}

/* src-flat/mycommand.c:108-161: the reference's short options (-f -d -p -e
 * -o -g -x -y -z -b -s -t, defaults potdir "pots", potname "ag", periodic,
 * 20^3 cells, bf 1, defgrad 1) into cmd, and this library's long options
 * into the driver settings.  Errors go to comd_host_status() (the
 * reference's parser ignores unknown options; here they are reported). */
void parseCommandLine(command_t *cmd, int argc, char **argv)
{
    comd_host_fail(COMD_OK, "");
    memset(cmd, 0, sizeof *cmd);
    strcpy(cmd->potdir, "pots");
    strcpy(cmd->potname, "ag");
    cmd->periodic = 1;
    cmd->nx = cmd->ny = cmd->nz = 20;
    cmd->bf = 1.0;
    cmd->defgrad = 1.0;
    comd_run_options o = {COMD_PREC_FP64, 0, 1, 0.0, 20, 10, 0.0, 0.0, 0.0, 0.0};
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
    opterr = 0;
    int c;
    while ((c = getopt_long(argc, argv, "hf:d:p:eogx:y:z:b:s:t:", opts, NULL)) != -1) {
        switch (c) {
        case 'h': usage(argv[0]); comd_host_fail(COMD_ERR_ARG, "help requested"); break;
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
        case 1: o.precision = strcmp(optarg, "mixed") == 0 ? COMD_PREC_MIXED : COMD_PREC_FP64; break;
        case 2: o.cheby = strcmp(optarg, "cheby") == 0; break;
        case 3: o.seed = strtoull(optarg, NULL, 10); break;
        case 4: o.disp = atof(optarg); break;
        case 5: o.niter = atoi(optarg); break;
        case 6: o.nsteps = atoi(optarg); break;
        case 7:
            if (sscanf(optarg, "%lf,%lf,%lf,%lf", &o.lj_sigma, &o.lj_epsilon, &o.lj_cutoff, &o.lj_mass) != 4)
                comd_host_fail(COMD_ERR_ARG, "--lj needs SIGMA,EPS,CUTOFF,MASS");
            break;
        default: usage(argv[0]); comd_host_fail(COMD_ERR_ARG, "unknown option"); break;
        }
    }
    g_opts = o;
    PERIODIC = cmd->periodic;
}

simflat_t *initSimFromCmdLine(int argc, char **argv)
{
    command_t cmd;
    comd_host_fail(COMD_OK, "");
    parseCommandLine(&cmd, argc, argv);
    if (comd_host_status()) return NULL;
    const comd_run_options o = g_opts;
    pmd_base_potential_t *pot;
    if (cmd.doeam) {
        pot = setEamPot(cmd.potdir, cmd.potname);
        if (!pot) return NULL;
        cmd.lat = ((eampotential_t *)pot)->lat;  /* src-flat/docompute.c:101 */
    } else {
        ljpotential_t *lj = getLJPot();
        if (o.lj_sigma > 0) {
            lj->sigma = o.lj_sigma;
            lj->epsilon = o.lj_epsilon;
            lj->cutoff = o.lj_cutoff;
            lj->mass = o.lj_mass;
        }
        pot = (pmd_base_potential_t *)lj;
        cmd.lat = 1.122462048 * lj->cutoff;  /* src-flat/docompute.c:117 */
        if (o.lj_sigma > 0) cmd.lat = 1.4142135623730951 * 1.122462048309373 * lj->sigma;
    }
    /* the reference's lattice is at rest whatever -t says; --disp or -t
     * here give the seeded lattice */
    simflat_t *sim;
    if (cmd.filename[0]) sim = fromFileGzip(cmd, pot);
    else if (cmd.temp > 0.0 || o.disp > 0.0) sim = comd_host_seeded_lattice(cmd, pot, o.seed, o.disp);
    else sim = create_fcc_lattice(cmd, pot);
    if (!sim) return NULL;
    if (cmd.doeam) sim->ch_pot = setChebPot((eampotential_t *)pot, 32);  /* docompute.c:133 */
    if (comd_dropin_bind(sim, kind_of(sim))) return NULL;
    if (comd_host_set_precision(sim, o.precision) || (o.cheby && comd_host_set_eval(sim, 1))) return NULL;
    if (comd_dropin_set_resident(sim, 1)) return NULL;
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
    comd_ctx *ctx = comd_dropin_bind(s, kind_of(s)) ? NULL : comd_host_ctx(s);
    double ts = wall();
    computeForce(s);
    if (ctx) comd_sync(ctx);
    double te = wall();
    double ke = 0;
    for (int iter = 0; iter < niter; iter++) {
        const double ns = iter == 0 ? 1.0 : (double)(1 + nsteps);
        if (ctx) comd_energy(ctx, NULL, &ke, NULL);
        printf(" %3d %30.20g computed in %.3fs (%8.4f us/atom for %d atoms)  ke %.12g\n", iter * nsteps, s->e, te - ts,
               1.0e6 * (te - ts) / ns / (double)s->ntot, s->ntot, ke);
        printf("Virial stress = %g\n", s->stress);
        ts = wall();
        nTimeSteps(nsteps, s, dt);
        te = wall();
    }
    printf(" %3d %30.20f computed in %.3fs (%8.4f us/atom for %d atoms)\n", niter * nsteps, s->e, te - ts,
           1.0e6 * (te - ts) / (double)(nsteps + 1) / (double)s->ntot, s->ntot);
    if (ctx && comd_dropin_resident(s)) comd_host_download(s);
    return NULL;
}
