/*
 * comd_io.c -- the reference's gz ASCII atom format (src-flat/read.c:155-241
 * reader; src-flat/write.c:175-220 writer): a line "natoms nmove", a comment
 * line, the periodic bounds, then one "x y z itype px py pz" line per atom
 * with momenta in the reference's internal units.  The writer prints %.17g
 * so that the doubles survive a round trip (the reference prints %.8lf).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "comd_host.h"
#include "comd_host_internal.h"

int comd_read_gz(const char *path, int *n, double bounds[3], double **xyz, double **pxyz)
{
    gzFile fp = gzopen(path, "rb");
    if (!fp) return comd_host_fail(COMD_ERR_ARG, "cannot open %s", path);
    char line[16384];
    int natoms = 0, nmove = 0;
    if (!gzgets(fp, line, sizeof line) || sscanf(line, "%d%d", &natoms, &nmove) < 1 || natoms < 0 ||
        !gzgets(fp, line, sizeof line) || !gzgets(fp, line, sizeof line) ||
        sscanf(line, "%lg %lg %lg", &bounds[0], &bounds[1], &bounds[2]) != 3) {
        gzclose(fp);
        return comd_host_fail(COMD_ERR_ARG, "bad header in %s", path);
    }
    double *x = (double *)malloc((size_t)(natoms > 0 ? natoms : 1) * 3 * sizeof(double));
    double *p = (double *)malloc((size_t)(natoms > 0 ? natoms : 1) * 3 * sizeof(double));
    if (!x || !p) {
        free(x);
        free(p);
        gzclose(fp);
        return comd_host_fail(COMD_ERR_NOMEM, "out of memory reading %s", path);
    }
    for (int i = 0; i < natoms; i++) {
        int it;
        if (!gzgets(fp, line, sizeof line) ||
            sscanf(line, "%lg %lg %lg %d %lg %lg %lg", &x[3 * i], &x[3 * i + 1], &x[3 * i + 2], &it, &p[3 * i],
                   &p[3 * i + 1], &p[3 * i + 2]) != 7) {
            free(x);
            free(p);
            gzclose(fp);
            return comd_host_fail(COMD_ERR_ARG, "bad atom line %d in %s", i, path);
        }
    }
    gzclose(fp);
    *n = natoms;
    *xyz = x;
    *pxyz = p;
    return COMD_OK;
}

int comd_write_gz(const char *path, int n, const double bounds[3], const double *xyz, const double *pxyz)
{
    gzFile fp = gzopen(path, "wb");
    if (!fp) return comd_host_fail(COMD_ERR_ARG, "cannot write %s", path);
    gzprintf(fp, "%d 0\n\n", n);
    gzprintf(fp, "%.17g %.17g %.17g\n", bounds[0], bounds[1], bounds[2]);
    for (int i = 0; i < n; i++)
        gzprintf(fp, "%.17g %.17g %.17g 1 %.17g %.17g %.17g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                 pxyz ? pxyz[3 * i] : 0.0, pxyz ? pxyz[3 * i + 1] : 0.0, pxyz ? pxyz[3 * i + 2] : 0.0);
    return gzclose(fp) == Z_OK ? COMD_OK : comd_host_fail(COMD_ERR_ARG, "write failed for %s", path);
}

/* src-flat/read.c:155-241: grid from the bounds with defgrad 0 (strain 1),
 * ids in file order, type 0, mass from the potential. */
simflat_t *fromFileGzip(command_t cmd, pmd_base_potential_t *pot)
{
    int n = 0;
    double b[3];
    double *x = NULL, *p = NULL;
    if (comd_read_gz(cmd.filename, &n, b, &x, &p)) return NULL;
    simflat_t *s = blankSimulation(pot);
    if (!s) {
        free(x);
        free(p);
        return NULL;
    }
    s->comment = (char *)calloc(1024, 1);
    s->bf = cmd.bf;
    for (int k = 0; k < 3; k++) s->bounds[k] = b[k];
    if (allocDomains(s)) {
        free(x);
        free(p);
        destroySimulation(&s);
        return NULL;
    }
    for (int i = 0; i < n; i++)
        if (putAtomInBox(s, i, 1, 0, pot->mass, x[3 * i], x[3 * i + 1], x[3 * i + 2], p[3 * i], p[3 * i + 1],
                         p[3 * i + 2], 0, 0, 0)) {
            destroySimulation(&s);
            break;
        }
    free(x);
    free(p);
    return s;
}
