/*
 * comd_setup.c -- host-side setup of a simulation: EAM table reading,
 * Chebyshev fit, Lennard-Jones parameters, link-cell grid, atom insertion,
 * FCC lattice and seeded initial states.  Runs once per simulation on the
 * CPU, as the reference's setup does; the hot path is in libcomd_hip.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "comd_host.h"
#include "comd_host_internal.h"


/* ---------------------------------------------------------------------
 * EAM tables: Fortran-record binary (byte-swapped when the first record
 * marker exceeds 4096) or ASCII when the name ends in 't'; x0 shifted by one
 * interval; guard slots values[-1], values[n], values[n+1]
 * (src-flat/eam.c:130-288). */

static void bswap(void *p, int n, int size, int on)
{
    if (!on) return;
    unsigned char *b = (unsigned char *)p;
    for (int e = 0; e < n; e++, b += size)
        for (int i = 0, j = size - 1; i < j; i++, j--) {
            unsigned char t = b[i];
            b[i] = b[j];
            b[j] = t;
        }
}

static potentialarray_t *potarray_new(int n, double x0, double xn, double invDx)
{
    potentialarray_t *a = (potentialarray_t *)calloc(1, sizeof *a);
    double *store = (double *)calloc((size_t)n + 3, sizeof(double));
    if (!a || !store) {
        free(a);
        free(store);
        return NULL;
    }
    a->values = store + 1;
    a->n = n;
    a->invDx = invDx;
    a->xn = xn;
    a->x0 = x0 + (xn - x0) / (double)n;
    return a;
}

static void potarray_free(potentialarray_t **a)
{
    if (!a || !*a) return;
    if ((*a)->values) free((*a)->values - 1);
    free(*a);
    *a = NULL;
}

static potentialarray_t *read_table_binary(const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return NULL;
    int32_t rec, n, iflo, ifhi;
    double x0, xn, invDx;
    int ok = fread(&rec, 4, 1, fp) == 1;
    const int sw = rec > 4096;
    ok = ok && fread(&n, 4, 1, fp) == 1;
    ok = ok && fread(&x0, 8, 1, fp) == 1;
    ok = ok && fread(&xn, 8, 1, fp) == 1;
    ok = ok && fread(&invDx, 8, 1, fp) == 1;
    ok = ok && fread(&iflo, 4, 1, fp) == 1;
    ok = ok && fread(&ifhi, 4, 1, fp) == 1;
    ok = ok && fread(&rec, 4, 1, fp) == 1;
    ok = ok && fread(&rec, 4, 1, fp) == 1;
    bswap(&n, 1, 4, sw);
    bswap(&x0, 1, 8, sw);
    bswap(&xn, 1, 8, sw);
    bswap(&invDx, 1, 8, sw);
    bswap(&iflo, 1, 4, sw);
    bswap(&ifhi, 1, 4, sw);
    bswap(&rec, 1, 4, sw);
    if (!ok || n < 2 || n > (1 << 24) || rec != n * 8) {
        fclose(fp);
        comd_host_fail(COMD_ERR_ARG, "bad table header in %s", path);
        return NULL;
    }
    potentialarray_t *a = potarray_new(n, x0, xn, invDx);
    if (!a) {
        fclose(fp);
        return NULL;
    }
    ok = fread(a->values, 8, (size_t)n, fp) == (size_t)n;
    fclose(fp);
    if (!ok) {
        potarray_free(&a);
        comd_host_fail(COMD_ERR_ARG, "short table %s", path);
        return NULL;
    }
    bswap(a->values, n, 8, sw);
    const double lo = iflo ? a->values[0] : 0.0;
    const double hi = ifhi ? a->values[n - 1] : 0.0;
    a->values[-1] = lo;
    a->values[n] = hi;
    a->values[n + 1] = hi;
    return a;
}

static potentialarray_t *read_table_ascii(const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    char line[4096];
    int n;
    double x0, xn, invDx;
    if (!fgets(line, sizeof line, fp) || sscanf(line, "%10d %lg %lg %lg", &n, &x0, &xn, &invDx) != 4 || n < 2) {
        fclose(fp);
        return NULL;
    }
    potentialarray_t *a = potarray_new(n, x0, xn, invDx);
    for (int i = 0; a && i < n; i++)
        if (fscanf(fp, "%lg", &a->values[i]) != 1) potarray_free(&a);
    fclose(fp);
    if (!a) return NULL;
    a->values[-1] = a->values[0];
    a->values[n] = a->values[n - 1];
    a->values[n + 1] = a->values[n - 1];
    return a;
}

static potentialarray_t *read_table(const char *path)
{
    size_t len = strlen(path);
    return (len && path[len - 1] == 't') ? read_table_ascii(path) : read_table_binary(path);
}

/* line 2 of the .doc file holds the mass, line 3 the lattice constant
 * (src-flat/eam.c:290-329) */
static int read_doc(const char *path, double *mass, double *lat)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char line[4096];
    int ok = fgets(line, sizeof line, fp) != NULL;
    ok = ok && fgets(line, sizeof line, fp) && sscanf(line, "%lf", mass) == 1;
    ok = ok && fgets(line, sizeof line, fp) && sscanf(line, "%lf", lat) == 1;
    fclose(fp);
    return ok ? 0 : -1;
}

static void eam_destroy(void **pp)
{
    if (!pp || !*pp) return;
    eampotential_t *p = (eampotential_t *)*pp;
    potarray_free(&p->phi);
    potarray_free(&p->rho);
    potarray_free(&p->f);
    free(p);
    *pp = NULL;
}

pmd_base_potential_t *setEamPot(const char *dir, const char *name)
{
    char path[4096];
    eampotential_t *p = (eampotential_t *)calloc(1, sizeof *p);
    if (!p) return NULL;
    snprintf(path, sizeof path, "%s/%s.phi", dir, name);
    p->phi = read_table(path);
    snprintf(path, sizeof path, "%s/%s.rho", dir, name);
    p->rho = read_table(path);
    snprintf(path, sizeof path, "%s/%s.f", dir, name);
    p->f = read_table(path);
    snprintf(path, sizeof path, "%s/%s.doc", dir, name);
    double mass = -1, lat = 0;
    if (read_doc(path, &mass, &lat) || !p->phi || !p->rho || !p->f) {
        void *q = p;
        eam_destroy(&q);
        comd_host_fail(COMD_ERR_ARG, "unable to read potential %s/%s", dir, name);
        return NULL;
    }
    p->mass = mass;
    p->lat = lat;
    p->cutoff = p->phi->xn; /* src-flat/eam.c:392 */
    p->force = eamForce;
    p->destroy = eam_destroy;
    return (pmd_base_potential_t *)p;
}

static void lj_destroy(void **pp)
{
    if (pp && *pp) {
        free(*pp);
        *pp = NULL;
    }
}

/* src-flat/ljForce.c:75-86 */
ljpotential_t *getLJPot(void)
{
    ljpotential_t *p = (ljpotential_t *)calloc(1, sizeof *p);
    if (!p) return NULL;
    p->force = LJ;
    p->destroy = lj_destroy;
    p->sigma = 1.53;
    p->epsilon = 0.0085;
    p->cutoff = 3.0 * p->sigma;
    p->mass = 1.0;
    return p;
}

/* ---------------------------------------------------------------------
 * Chebyshev fit of a table: samples at the n Chebyshev nodes of [x0, xn]
 * (with sampled values above 1e10 replaced by 0), cosine transform, and the
 * derivative series (src-flat/cheby.c:15-270). */

static double sample_table(const potentialarray_t *a, double r)
{
    if (r < a->x0) r = a->x0;
    else if (r > a->xn) r = a->xn;
    r = (r - a->x0) * a->invDx;
    const int i = (int)floor(r);
    r = r - floor(r);
    const double *v = a->values;
    const double g0 = v[i + 1] - v[i - 1];
    double val = v[i] + 0.5 * r * (r * (v[i + 1] + v[i - 1] - 2.0 * v[i]) + g0);
    if (val > 1.0e10) val = 0.0;
    return val;
}

#define COMD_PI 3.141592653589793

static potentialarray_t *cheb_fit(const potentialarray_t *t, int n)
{
    potentialarray_t *c = (potentialarray_t *)calloc(1, sizeof *c);
    c->values = (double *)calloc((size_t)n + 1, sizeof(double)) + 1;
    c->n = n;
    c->x0 = t->x0;
    c->xn = t->xn;
    const double a = t->x0, b = t->xn;
    double *fk = (double *)malloc((size_t)n * sizeof(double));
    const double bma = 0.5 * (b - a), bpa = 0.5 * (b + a);
    for (int k = 0; k < n; k++) fk[k] = sample_table(t, cos(COMD_PI * (k + 0.5) / n) * bma + bpa);
    const double fac = 2.0 / n;
    for (int j = 0; j < n; j++) {
        double sum = 0.0;
        for (int k = 0; k < n; k++) sum += fk[k] * cos(COMD_PI * j * (k + 0.5) / n);
        c->values[j] = fac * sum;
    }
    free(fk);
    return c;
}

static potentialarray_t *cheb_deriv(const potentialarray_t *c)
{
    const int n = c->n;
    potentialarray_t *d = (potentialarray_t *)calloc(1, sizeof *d);
    d->values = (double *)calloc((size_t)n + 1, sizeof(double)) + 1;
    d->n = n;
    d->x0 = c->x0;
    d->xn = c->xn;
    d->invDx = 1.0;
    double *cd = d->values;
    cd[n - 1] = 0.0;
    cd[n - 2] = 2.0 * (n - 1) * c->values[n - 1];
    for (int j = n - 2; j > 0; j--) cd[j - 1] = cd[j + 1] + 2 * (j)*c->values[j];
    const double con = 2.0 / (c->xn - c->x0);
    for (int j = 0; j < n; j++) cd[j] = cd[j] * con;
    return d;
}

eam_cheby_t *setChebPot(eampotential_t *pot, int n)
{
    if (!pot || n < 2 || n > 64) return NULL;
    eam_cheby_t *ch = (eam_cheby_t *)calloc(1, sizeof *ch);
    ch->cutoff = pot->cutoff;
    ch->mass = pot->mass;
    ch->lat = pot->lat;
    ch->phi = cheb_fit(pot->phi, n);
    ch->dphi = cheb_deriv(ch->phi);
    ch->rho = cheb_fit(pot->rho, n);
    ch->drho = cheb_deriv(ch->rho);
    ch->f = cheb_fit(pot->f, n);
    ch->df = cheb_deriv(ch->f);
    return ch;
}

void destroyChebPot(eam_cheby_t **pc)
{
    if (!pc || !*pc) return;
    eam_cheby_t *c = *pc;
    potentialarray_t **a[6] = {&c->phi, &c->rho, &c->f, &c->dphi, &c->drho, &c->df};
    for (int k = 0; k < 6; k++) potarray_free(a[k]);
    free(c);
    *pc = NULL;
}

/* ---------------------------------------------------------------------
 * link-cell grid and atom insertion (src-flat/domains.c:155-291) */

simflat_t *blankSimulation(pmd_base_potential_t *pot)
{
    simflat_t *s = (simflat_t *)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->pot = pot ? pot : (pmd_base_potential_t *)getLJPot();
    return s;
}

int allocDomains(simflat_t *s)
{
    if (!s || !s->pot) return COMD_ERR_ARG;
    comd_params p;
    memset(&p, 0, sizeof p);
    p.cutoff = s->pot->cutoff;
    double b[3] = {s->bounds[0], s->bounds[1], s->bounds[2]};
    int rc = comd_grid_for(&p, b, s->defgrad, s->bf);
    if (rc) return comd_host_fail(rc, "need at least one cutoff-wide cell");
    s->nboxes = 1;
    for (int j = 0; j < 3; j++) {
        s->nbx[j] = p.nbx[j];
        s->boxsize[j] = p.boxsize[j];
        s->nboxes *= p.nbx[j];
    }
    const size_t nb = (size_t)s->nboxes, ns = nb * MAXATOMS;
    s->dcenter = (real3 *)calloc(nb, sizeof(real3));
    s->natoms = (int *)calloc(nb, sizeof(int));
    s->id = (int *)calloc(ns, sizeof(int));
    s->iType = (int *)calloc(ns, sizeof(int));
    s->mass = (real_t *)calloc(ns, sizeof(real_t));
    s->r = (real3 *)calloc(ns, sizeof(real3));
    s->p = (real3 *)calloc(ns, sizeof(real3));
    s->f = (real4 *)calloc(ns, sizeof(real4));
    s->rho = (real_t *)calloc(ns, sizeof(real_t));
    s->fi = (real_t *)calloc(ns, sizeof(real_t));
    if (!s->dcenter || !s->natoms || !s->id || !s->iType || !s->mass || !s->r || !s->p || !s->f || !s->rho || !s->fi)
        return comd_host_fail(COMD_ERR_NOMEM, "host allocation failed");
    for (int c = 0; c < s->nboxes; c++) {
        const int ix = c % s->nbx[0], iy = (c / s->nbx[0]) % s->nbx[1], iz = c / s->nbx[0] / s->nbx[1];
        s->dcenter[c][0] = s->boxsize[0] * (double)ix;
        s->dcenter[c][1] = s->boxsize[1] * (double)iy;
        s->dcenter[c][2] = s->boxsize[2] * (double)iz;
    }
    s->stateflag = 1;
    return COMD_OK;
}

int putAtomInBox(simflat_t *s, int id, char flagMove, int iType, real_t mass, double x, double y, double z,
                 double px, double py, double pz, double fx, double fy, double fz)
{
    (void)flagMove;
    double xyz[3] = {x, y, z};
    int ib[3];
    for (int m = 0; m < 3; m++) {
        if (xyz[m] < 0.0) xyz[m] += s->bounds[m];
        else if (xyz[m] >= s->bounds[m]) xyz[m] -= s->bounds[m];
        ib[m] = (int)floor(xyz[m] / s->boxsize[m]);
        if (ib[m] < 0 || ib[m] >= s->nbx[m]) return comd_host_fail(COMD_ERR_OUTSIDE, "atom %d outside the grid", id);
    }
    const int c = ib[0] + s->nbx[0] * ib[1] + s->nbx[0] * s->nbx[1] * ib[2];
    if (s->natoms[c] >= MAXATOMS) return comd_host_fail(COMD_ERR_OVERFLOW, "cell %d is full", c);
    const int o = c * MAXATOMS + s->natoms[c];
    s->ntot++;
    s->natoms[c]++;
    s->id[o] = id;
    s->iType[o] = iType;
    s->mass[o] = mass;
    for (int m = 0; m < 3; m++) s->r[o][m] = xyz[m] - s->dcenter[c][m];
    s->p[o][0] = px;
    s->p[o][1] = py;
    s->p[o][2] = pz;
    s->f[o][0] = fx;
    s->f[o][1] = fy;
    s->f[o][2] = fz;
    return COMD_OK;
}

void destroySimulation(simflat_t **ps)
{
    if (!ps || !*ps) return;
    simflat_t *s = *ps;
    comd_host_forget(s);
    if (s->pot && s->pot->destroy) {
        void *p = s->pot;
        ((void (*)(void **))s->pot->destroy)(&p);
    }
    destroyChebPot(&s->ch_pot);
    free(s->dcenter);
    free(s->natoms);
    free(s->id);
    free(s->iType);
    free(s->mass);
    free(s->r);
    free(s->p);
    free(s->f);
    free(s->rho);
    free(s->fi);
    free(s->comment);
    free(s);
    *ps = NULL;
}

/* ---------------------------------------------------------------------
 * initial conditions */

/* Walk of src-flat/ic_fcc.c:115-132: start at a/4, step a/2 by repeated
 * addition, keep points whose step parity (i+j+k) is odd; the counters are
 * never reset, as in the reference. */
int comd_fcc_positions(int nx, int ny, int nz, double lat, double defgrad, double *xyz, int max)
{
    const double b0 = nx * lat * defgrad, b1 = ny * lat, b2 = nz * lat;
    const double halflat = lat / 2.0;
    int i = 0, j = 0, k = 0, n = 0;
    double z = lat / 4.0;
    while (z < b2) {
        double y = lat / 4.0;
        while (y < b1) {
            double x = lat * defgrad / 4.0;
            while (x < b0) {
                if ((i + j + k) % 2) {
                    if (xyz && n < max) {
                        xyz[3 * n] = x;
                        xyz[3 * n + 1] = y;
                        xyz[3 * n + 2] = z;
                    }
                    n++;
                }
                x += halflat * defgrad;
                i++;
            }
            y += halflat;
            j++;
        }
        z += halflat;
        k++;
    }
    return n;
}

/* Lattice of create_fcc_lattice, optionally with seeded displacements and
 * momenta, binned in id order through putAtomInBox (ic_fcc.c:125). */
static simflat_t *fcc_lattice(command_t cmd, pmd_base_potential_t *pot, int seeded, unsigned long long seed,
                              double disp)
{
    simflat_t *s = blankSimulation(pot);
    if (!s) return NULL;
    s->comment = (char *)calloc(1024, 1);
    s->defgrad = cmd.defgrad;
    s->bounds[0] = cmd.nx * cmd.lat * cmd.defgrad;
    s->bounds[1] = cmd.ny * cmd.lat;
    s->bounds[2] = cmd.nz * cmd.lat;
    s->bf = cmd.bf;
    if (allocDomains(s)) {
        destroySimulation(&s);
        return NULL;
    }
    const int n = comd_fcc_positions(cmd.nx, cmd.ny, cmd.nz, cmd.lat, cmd.defgrad, NULL, 0);
    double *xyz = (double *)malloc((size_t)(n > 0 ? n : 1) * 3 * sizeof(double));
    double *p = (double *)calloc((size_t)(n > 0 ? n : 1) * 3, sizeof(double));
    comd_fcc_positions(cmd.nx, cmd.ny, cmd.nz, cmd.lat, cmd.defgrad, xyz, n);
    if (seeded && (cmd.temp > 0.0 || disp > 0.0)) comd_seed_state(seed, n, pot->mass, cmd.temp, disp, xyz, p);
    for (int a = 0; a < n; a++)
        if (putAtomInBox(s, a, 1, 1, pot->mass, xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], p[3 * a], p[3 * a + 1],
                         p[3 * a + 2], 0, 0, 0)) {
            free(xyz);
            free(p);
            destroySimulation(&s);
            return NULL;
        }
    free(xyz);
    free(p);
    return s;
}

/* src-flat/ic_fcc.c:63-141: a perfect lattice at rest (the reference does
 * not use cmd.temp), type 1, ids in creation order. */
simflat_t *create_fcc_lattice(command_t cmd, pmd_base_potential_t *pot) { return fcc_lattice(cmd, pot, 0, 0, 0.0); }

simflat_t *comd_host_seeded_lattice(command_t cmd, pmd_base_potential_t *pot, unsigned long long seed, double disp)
{
    return fcc_lattice(cmd, pot, 1, seed, disp);
}

/* SplitMix64 (Steele, Lea, Flood 2014) */
static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t *st)
{
    /* 53 random bits in (0, 1] */
    return ((double)(splitmix64(st) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

/* Box-Muller pair of standard normals */
static void normal2(uint64_t *st, double *g0, double *g1)
{
    const double u1 = uniform01(st), u2 = uniform01(st);
    const double rr = sqrt(-2.0 * log(u1));
    *g0 = rr * cos(2.0 * COMD_PI * u2);
    *g1 = rr * sin(2.0 * COMD_PI * u2);
}

void comd_seed_state(unsigned long long seed, int n, double mass_amu, double temp, double disp, double *xyz,
                     double *pxyz)
{
    /* Boltzmann constant in hartree per kelvin */
    const double kB = 3.166811563e-6;
    const double rmass = 1822.83 * mass_amu;
    const double sp = temp > 0.0 ? sqrt(rmass * kB * temp) : 0.0;
    double mean[3] = {0, 0, 0};
    for (int a = 0; a < n; a++) {
        uint64_t st = (uint64_t)seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(a + 1));
        splitmix64(&st);
        double g[6];
        normal2(&st, &g[0], &g[1]);
        normal2(&st, &g[2], &g[3]);
        normal2(&st, &g[4], &g[5]);
        if (xyz && disp > 0.0)
            for (int k = 0; k < 3; k++) xyz[3 * a + k] += disp * g[k];
        if (pxyz) {
            for (int k = 0; k < 3; k++) {
                pxyz[3 * a + k] = sp * g[3 + k];
                mean[k] += pxyz[3 * a + k];
            }
        }
    }
    if (pxyz && n > 0) {
        for (int k = 0; k < 3; k++) mean[k] /= n;
        for (int a = 0; a < n; a++)
            for (int k = 0; k < 3; k++) pxyz[3 * a + k] -= mean[k];
    }
}
