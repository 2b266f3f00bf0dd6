/*
 * orc_potential.c -- oracle restatement of the EAM table reader, the table
 * interpolation and the Chebyshev fit/evaluation of CoMD src-flat.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void swap_bytes(void *p, int nelem, int size, int doswap)
{
    if (!doswap) return;
    unsigned char *b = (unsigned char *)p;
    for (int e = 0; e < nelem; e++, b += size)
        for (int i = 0, j = size - 1; i < j; i++, j--) {
            unsigned char t = b[i];
            b[i] = b[j];
            b[j] = t;
        }
}

/* Storage layout and x0 shift of src-flat/eam.c:130-153. */
static int table_alloc(orc_table *t, int n, double x0, double xn, double invDx)
{
    t->store = (double *)calloc((size_t)n + 3, sizeof(double));
    if (!t->store) return ORC_ERR_ALLOC;
    t->vals = t->store + 1;
    t->n = n;
    t->invDx = invDx;
    t->xn = xn;
    t->x0 = x0 + (xn - x0) / (double)n;
    return ORC_OK;
}

/* Fortran-record binary table, src-flat/eam.c:156-249.  Record 1 holds
 * n, x0, xn, invDx, iflo, ifhi; record 2 holds n doubles.  The file is
 * byte-swapped when the first record marker reads above 4096 (eam.c:176). */
static int table_read_binary(orc_table *t, const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return ORC_ERR_NOFILE;
    int rec, n, iflo, ifhi;
    double x0, xn, invDx;
    int ok = 1;
    ok &= fread(&rec, 4, 1, fp) == 1;
    int sw = rec > 4096;
    ok &= fread(&n, 4, 1, fp) == 1;      swap_bytes(&n, 1, 4, sw);
    ok &= fread(&x0, 8, 1, fp) == 1;     swap_bytes(&x0, 1, 8, sw);
    ok &= fread(&xn, 8, 1, fp) == 1;     swap_bytes(&xn, 1, 8, sw);
    ok &= fread(&invDx, 8, 1, fp) == 1;  swap_bytes(&invDx, 1, 8, sw);
    ok &= fread(&iflo, 4, 1, fp) == 1;   swap_bytes(&iflo, 1, 4, sw);
    ok &= fread(&ifhi, 4, 1, fp) == 1;   swap_bytes(&ifhi, 1, 4, sw);
    ok &= fread(&rec, 4, 1, fp) == 1;    /* trailer of record 1 */
    ok &= fread(&rec, 4, 1, fp) == 1;    swap_bytes(&rec, 1, 4, sw);
    if (!ok || n <= 0 || n > (1 << 24) || rec != n * 8) { fclose(fp); return ORC_ERR_TABLE; }
    if (table_alloc(t, n, x0, xn, invDx)) { fclose(fp); return ORC_ERR_ALLOC; }
    double *buf = (double *)malloc((size_t)n * 8);
    if (!buf) { fclose(fp); return ORC_ERR_ALLOC; }
    if (fread(buf, 8, (size_t)n, fp) != (size_t)n) { free(buf); fclose(fp); return ORC_ERR_TABLE; }
    swap_bytes(buf, n, 8, sw);
    memcpy(t->vals, buf, (size_t)n * 8);
    free(buf);
    fclose(fp);
    /* guard slots, eam.c:232-246 */
    double lo = iflo ? t->vals[0] : 0.0;
    double hi = ifhi ? t->vals[n - 1] : 0.0;
    t->vals[-1] = lo;
    t->vals[n] = hi;
    t->vals[n + 1] = hi;
    return ORC_OK;
}

/* ASCII table (file name ending in 't'), src-flat/eam.c:252-288. */
static int table_read_ascii(orc_table *t, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return ORC_ERR_NOFILE;
    char line[4096];
    int n;
    double x0, xn, invDx;
    if (!fgets(line, sizeof line, fp) || sscanf(line, "%10d %lg %lg %lg", &n, &x0, &xn, &invDx) != 4 || n <= 0) {
        fclose(fp);
        return ORC_ERR_TABLE;
    }
    if (table_alloc(t, n, x0, xn, invDx)) { fclose(fp); return ORC_ERR_ALLOC; }
    for (int i = 0; i < n; i++)
        if (fscanf(fp, "%lg", &t->vals[i]) != 1) { fclose(fp); return ORC_ERR_TABLE; }
    fclose(fp);
    t->vals[-1] = t->vals[0];
    t->vals[n] = t->vals[n - 1];
    t->vals[n + 1] = t->vals[n - 1];
    return ORC_OK;
}

static int table_read(orc_table *t, const char *path)
{
    size_t len = strlen(path);
    if (len && path[len - 1] == 't') return table_read_ascii(t, path);
    return table_read_binary(t, path);
}

/* Mass from line 2, lattice constant from line 3 of the .doc file,
 * src-flat/eam.c:290-329. */
static int doc_read(const char *path, double *mass, double *lat)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return ORC_ERR_NOFILE;
    char line[4096];
    int ok = fgets(line, sizeof line, fp) != NULL;
    ok = ok && fgets(line, sizeof line, fp) && sscanf(line, "%lf", mass) == 1;
    ok = ok && fgets(line, sizeof line, fp) && sscanf(line, "%lf", lat) == 1;
    fclose(fp);
    return ok ? ORC_OK : ORC_ERR_TABLE;
}

/* src-flat/eam.c:347-398; cutoff = phi.xn (eam.c:392). */
int orc_pot_read_eam(orc_pot *pot, const char *dir, const char *name)
{
    char path[4096];
    memset(pot, 0, sizeof *pot);
    pot->kind = ORC_POT_EAM;
    int rc;
    snprintf(path, sizeof path, "%s/%s.phi", dir, name);
    if ((rc = table_read(&pot->phi, path))) return rc;
    snprintf(path, sizeof path, "%s/%s.rho", dir, name);
    if ((rc = table_read(&pot->rho, path))) return rc;
    snprintf(path, sizeof path, "%s/%s.f", dir, name);
    if ((rc = table_read(&pot->f, path))) return rc;
    snprintf(path, sizeof path, "%s/%s.doc", dir, name);
    if ((rc = doc_read(path, &pot->mass, &pot->lat))) return rc;
    pot->cutoff = pot->phi.xn;
    pot->eval = ORC_EVAL_TABLE;
    return ORC_OK;
}

/* Parameters of src-flat/ljForce.c:75-86, made settable. */
void orc_pot_set_lj(orc_pot *pot, double sigma, double epsilon, double cutoff, double mass)
{
    memset(pot, 0, sizeof *pot);
    pot->kind = ORC_POT_LJ;
    pot->sigma = sigma;
    pot->epsilon = epsilon;
    pot->cutoff = cutoff;
    pot->mass = mass;
    pot->lj_sign = 1.0;
}

void orc_pot_set_lj_sign(orc_pot *pot, double sign) { pot->lj_sign = sign < 0.0 ? -1.0 : 1.0; }

void orc_pot_free(orc_pot *pot)
{
    free(pot->phi.store);
    free(pot->rho.store);
    free(pot->f.store);
    pot->phi.store = pot->rho.store = pot->f.store = NULL;
}

void orc_pot_set_eval(orc_pot *pot, int eval) { pot->eval = eval; }

double orc_pot_get(const orc_pot *pot, int which)
{
    return which == 0 ? pot->cutoff : which == 1 ? pot->mass : pot->lat;
}

orc_pot *orc_pot_new(void) { return (orc_pot *)calloc(1, sizeof(orc_pot)); }
void orc_pot_delete(orc_pot *p) { if (p) { orc_pot_free(p); free(p); } }

/* 3-point quadratic value and interpolated central-difference derivative,
 * src-flat/eam.c:401-434. */
void orc_table_interp(const orc_table *a, double r, double *val, double *der)
{
    if (r < a->x0) r = a->x0;
    else if (r > a->xn) r = a->xn;
    r = (r - a->x0) * a->invDx;
    int i = (int)floor(r);
    r = r - floor(r);
    const double *v = a->vals;
    double g0 = v[i + 1] - v[i - 1];
    double g1 = v[i + 2] - v[i];
    *val = v[i] + 0.5 * r * (r * (v[i + 1] + v[i - 1] - 2.0 * v[i]) + g0);
    if (i <= 0) *der = 0.0;
    else *der = 0.5 * (g0 + r * (g1 - g0)) * a->invDx;
}

/* The sampling interpolator of the Chebyshev fit: same formula with values
 * above 1e10 replaced by zero, src-flat/cheby.c:15-60. */
static double table_sample(const orc_table *a, double r)
{
    double v, d;
    orc_table_interp(a, r, &v, &d);
    if (v > 1.0e10) v = 0.0;
    return v;
}

#define ORC_PI 3.141592653589793

/* Chebyshev coefficients from n samples at the Chebyshev nodes of
 * [x0, xn], src-flat/cheby.c:125-179. */
void orc_cheb_fit(const orc_table *t, int n, orc_cheb *out)
{
    double a = t->x0, b = t->xn;
    double fk[64];
    out->n = n;
    out->x0 = a;
    out->xn = b;
    double bma = 0.5 * (b - a);
    double bpa = 0.5 * (b + a);
    for (int k = 0; k < n; k++) {
        double y = cos(ORC_PI * (k + 0.5) / n);
        fk[k] = table_sample(t, y * bma + bpa);
    }
    double fac = 2.0 / n;
    for (int j = 0; j < n; j++) {
        double sum = 0.0;
        for (int k = 0; k < n; k++) sum += fk[k] * cos(ORC_PI * j * (k + 0.5) / n);
        out->c[j] = fac * sum;
    }
}

/* Coefficients of the derivative series, src-flat/cheby.c:219-270. */
void orc_cheb_deriv(const orc_cheb *c, orc_cheb *out)
{
    int n = c->n;
    out->n = n;
    out->x0 = c->x0;
    out->xn = c->xn;
    double *d = out->c;
    d[n - 1] = 0.0;
    d[n - 2] = 2.0 * (n - 1) * c->c[n - 1];
    for (int j = n - 2; j > 0; j--) d[j - 1] = d[j + 1] + 2 * (j)*c->c[j];
    double con = 2.0 / (c->xn - c->x0);
    for (int j = 0; j < n; j++) d[j] = d[j] * con;
}

/* Clenshaw evaluation with the lower clamp only (the upper clamp at
 * src-flat/cheby.c:293 assigns b to itself), src-flat/cheby.c:276-310. */
double orc_cheb_eval(const orc_cheb *cc, double x)
{
    double a = cc->x0, b = cc->xn;
    const double *c = cc->c;
    if (x < a) x = a;
    double d = 0.0, dd = 0.0, sv;
    double y = (2.0 * x - a - b) / (b - a);
    double y2 = 2.0 * y;
    for (int j = cc->n - 1; j > 0; j--) {
        sv = d;
        d = y2 * d - dd + c[j];
        dd = sv;
    }
    return y * d - dd + 0.5 * c[0];
}

/* src-flat/cheby.c:99-121 (called with n = 32 at docompute.c:56,133). */
void orc_pot_fit_cheb(orc_pot *pot, int n)
{
    orc_cheb_fit(&pot->phi, n, &pot->cphi);
    orc_cheb_deriv(&pot->cphi, &pot->cdphi);
    orc_cheb_fit(&pot->rho, n, &pot->crho);
    orc_cheb_deriv(&pot->crho, &pot->cdrho);
    orc_cheb_fit(&pot->f, n, &pot->cf);
    orc_cheb_deriv(&pot->cf, &pot->cdf);
}

void orc_pot_get_cheb(const orc_pot *p, double *out, int *n)
{
    const orc_cheb *cs[6] = {&p->cphi, &p->crho, &p->cf, &p->cdphi, &p->cdrho, &p->cdf};
    int m = p->cphi.n;
    *n = m;
    for (int k = 0; k < 6; k++) memcpy(out + k * m, cs[k]->c, (size_t)m * 8);
}

void orc_pot_get_table(const orc_pot *p, int which, int *n, double *hdr, double *vals)
{
    const orc_table *t = which == 0 ? &p->phi : which == 1 ? &p->rho : &p->f;
    *n = t->n;
    hdr[0] = t->x0;
    hdr[1] = t->xn;
    hdr[2] = t->invDx;
    if (vals) memcpy(vals, t->store, (size_t)(t->n + 3) * 8);
}

/* Value and derivative of function `which` (0 phi, 1 rho, 2 F) at x in the
 * potential's current evaluation mode: the table formula
 * (src-flat/eam.c:401-434) or the value and derivative series
 * (src-flat/eam.c:523-527,582-584,637-639). */
void orc_pot_eval(const orc_pot *p, int which, double x, double *val, double *der)
{
    if (p->eval == ORC_EVAL_CHEB) {
        const orc_cheb *cv[3] = {&p->cphi, &p->crho, &p->cf};
        const orc_cheb *cd[3] = {&p->cdphi, &p->cdrho, &p->cdf};
        *val = orc_cheb_eval(cv[which], x);
        *der = orc_cheb_eval(cd[which], x);
    } else {
        const orc_table *t = which == 0 ? &p->phi : which == 1 ? &p->rho : &p->f;
        orc_table_interp(t, x, val, der);
    }
}

/* Replaces the value series of phi, rho and F by coef[0..3n) (in that
 * order, on the fitted intervals) and rebuilds the derivative series from
 * them (src-flat/cheby.c:219-270). */
void orc_pot_set_cheb3(orc_pot *p, const double *coef, int n)
{
    orc_cheb *cs[3] = {&p->cphi, &p->crho, &p->cf};
    orc_cheb *ds[3] = {&p->cdphi, &p->cdrho, &p->cdf};
    for (int k = 0; k < 3; k++) {
        cs[k]->n = n;
        memcpy(cs[k]->c, coef + k * n, (size_t)n * 8);
        orc_cheb_deriv(cs[k], ds[k]);
    }
}
