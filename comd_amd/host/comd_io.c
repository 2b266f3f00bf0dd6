/*
 * comd_io.c -- the reference's atom file formats.
 *
 * gz ASCII (src-flat/read.c:155-241 reader, write.c:175-220 writer): a line
 * "natoms nmove", a comment line, the periodic bounds, then one
 * "x y z itype px py pz" line per atom, momenta in the reference's internal
 * units.  gzASCIIWrite prints %.8lf as the reference does; comd_write_gz
 * prints %.17g so that the doubles survive a round trip.
 *
 * clsman binary (src-flat/write.c:112-167, writeClsman): Fortran-style
 * records, each framed by its byte length before and after (4-byte ints):
 * (natoms, nmove = 0, 0), an empty comment record, the bounds (3 doubles),
 * then per atom in id order x y z (absolute, doubles), type 1 (int),
 * px py pz (doubles).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "comd_host.h"
#include "comd_host_internal.h"

/* The gz ASCII file at path; comment (1024 bytes, or NULL) receives the
 * comment line as the reference keeps it (with its newline, read.c:203). */
static int gz_read(const char *path, int *n, double bounds[3], double **xyz, double **pxyz, char *comment)
{
    gzFile fp = gzopen(path, "rb");
    if (!fp) return comd_host_fail(COMD_ERR_ARG, "cannot open %s", path);
    char line[16384];
    int natoms = 0, nmove = 0;
    if (!gzgets(fp, line, sizeof line) || sscanf(line, "%d%d", &natoms, &nmove) < 1 || natoms < 0 ||
        !gzgets(fp, line, sizeof line)) {
        gzclose(fp);
        return comd_host_fail(COMD_ERR_ARG, "bad header in %s", path);
    }
    if (comment) {
        strncpy(comment, line, 1023);
        comment[1023] = 0;
    }
    if (!gzgets(fp, line, sizeof line) || sscanf(line, "%lg %lg %lg", &bounds[0], &bounds[1], &bounds[2]) != 3) {
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

int comd_read_gz(const char *path, int *n, double bounds[3], double **xyz, double **pxyz)
{
    return gz_read(path, n, bounds, xyz, pxyz, NULL);
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
    char comment[1024];
    if (gz_read(cmd.filename, &n, b, &x, &p, comment)) return NULL;
    simflat_t *s = blankSimulation(pot);
    if (!s) {
        free(x);
        free(p);
        return NULL;
    }
    s->comment = (char *)calloc(1024, 1);
    memcpy(s->comment, comment, sizeof comment);
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

/* Slots of atoms 0 .. ntot-1 (the reference's getRL, write.c:57-87, which
 * reports missing or repeated ids on stderr; here they are an error). */
static int *slots_by_id(const simflat_t *s)
{
    int *slot = (int *)malloc((size_t)(s->ntot > 0 ? s->ntot : 1) * sizeof(int));
    if (!slot) return NULL;
    for (int i = 0; i < s->ntot; i++) slot[i] = -1;
    for (int b = 0; b < s->nboxes; b++)
        for (int i = 0; i < s->natoms[b]; i++) {
            const int o = b * MAXATOMS + i, a = s->id[o];
            if (a < 0 || a >= s->ntot || slot[a] >= 0) {
                free(slot);
                comd_host_fail(COMD_ERR_ARG, "atom ids are not 0 .. %d, each once", s->ntot - 1);
                return NULL;
            }
            slot[a] = o;
        }
    for (int i = 0; i < s->ntot; i++)
        if (slot[i] < 0) {
            free(slot);
            comd_host_fail(COMD_ERR_ARG, "atom id %d missing", i);
            return NULL;
        }
    return slot;
}

/* Absolute position (the cell's lower corner plus r) and momentum of the
 * atom in slot o. */
static void atom_xp(const simflat_t *s, int o, double td[6])
{
    const int b = o / MAXATOMS;
    for (int j = 0; j < 3; j++) {
        td[j] = s->dcenter[b][j] + s->r[o][j];
        td[j + 3] = s->p[o][j];
    }
}

/* src-flat/write.c:112-167 */
void writeClsman(simflat_t *s, char *fname)
{
    comd_host_fail(COMD_OK, "");  /* the status reports this call */
    int *slot = slots_by_id(s);
    if (!slot) return;
    FILE *fp = fopen(fname, "wb");
    if (!fp) {
        free(slot);
        comd_host_fail(COMD_ERR_ARG, "cannot write %s", fname);
        return;
    }
    const int head[7] = {3 * (int)sizeof(int), s->ntot, 0, 0, 3 * (int)sizeof(int), 0, 0};
    const int rec24 = 3 * (int)sizeof(double);
    const int recatoms = s->ntot * (6 * (int)sizeof(double) + (int)sizeof(int));
    const int type1 = 1;
    int ok = fwrite(head, sizeof(int), 7, fp) == 7;
    ok = ok && fwrite(&rec24, sizeof(int), 1, fp) == 1;
    ok = ok && fwrite(s->bounds, sizeof(double), 3, fp) == 3;
    ok = ok && fwrite(&rec24, sizeof(int), 1, fp) == 1;
    ok = ok && fwrite(&recatoms, sizeof(int), 1, fp) == 1;
    for (int i = 0; ok && i < s->ntot; i++) {
        double td[6];
        atom_xp(s, slot[i], td);
        ok = fwrite(td, sizeof(double), 3, fp) == 3 && fwrite(&type1, sizeof(int), 1, fp) == 1 &&
             fwrite(td + 3, sizeof(double), 3, fp) == 3;
    }
    ok = ok && fwrite(&recatoms, sizeof(int), 1, fp) == 1;
    ok = (fclose(fp) == 0) && ok;
    free(slot);
    if (!ok) comd_host_fail(COMD_ERR_ARG, "write failed for %s", fname);
}

/* src-flat/write.c:175-220: %.8lf, type 1, atoms in id order, the comment
 * line as the simulation keeps it.  The reference exits when a write
 * returns nothing, which an empty comment does (write.c:188); here an empty
 * comment is written as an empty line. */
void gzASCIIWrite(simflat_t *s, char *fname)
{
    comd_host_fail(COMD_OK, "");  /* the status reports this call */
    int *slot = slots_by_id(s);
    if (!slot) return;
    gzFile fp = gzopen(fname, "wb");
    if (!fp) {
        free(slot);
        comd_host_fail(COMD_ERR_ARG, "cannot write %s", fname);
        return;
    }
    int ok = gzprintf(fp, "%d 0\n", s->ntot) > 0;
    ok = ok && gzputs(fp, s->comment && s->comment[0] ? s->comment : "\n") > 0;
    ok = ok && gzprintf(fp, "%.8lf %.8lf %.8lf\n", s->bounds[0], s->bounds[1], s->bounds[2]) > 0;
    for (int i = 0; ok && i < s->ntot; i++) {
        double td[6];
        atom_xp(s, slot[i], td);
        ok = gzprintf(fp, "%.8lf %.8lf %.8lf %d %.8lf %.8lf %.8lf\n", td[0], td[1], td[2], 1, td[3], td[4], td[5]) > 0;
    }
    ok = (gzclose(fp) == Z_OK) && ok;
    free(slot);
    if (!ok) comd_host_fail(COMD_ERR_ARG, "write failed for %s", fname);
}

int comd_read_clsman(const char *path, int *n, double bounds[3], double **xyz, double **pxyz)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return comd_host_fail(COMD_ERR_ARG, "cannot open %s", path);
    int head[7], rec[2];
    int natoms = 0;
    if (fread(head, sizeof(int), 7, fp) != 7 || head[0] != 12 || head[4] != 12 || head[5] != 0 || head[6] != 0 ||
        (natoms = head[1]) < 0 || fread(rec, sizeof(int), 1, fp) != 1 || rec[0] != 24 ||
        fread(bounds, sizeof(double), 3, fp) != 3 || fread(rec, sizeof(int), 2, fp) != 2 || rec[0] != 24 ||
        rec[1] != natoms * 52) {
        fclose(fp);
        return comd_host_fail(COMD_ERR_ARG, "%s is not a clsman file", path);
    }
    double *x = (double *)malloc((size_t)(natoms > 0 ? natoms : 1) * 3 * sizeof(double));
    double *p = (double *)malloc((size_t)(natoms > 0 ? natoms : 1) * 3 * sizeof(double));
    int ok = x && p;
    for (int i = 0; ok && i < natoms; i++) {
        int t;
        ok = fread(&x[3 * i], sizeof(double), 3, fp) == 3 && fread(&t, sizeof(int), 1, fp) == 1 &&
             fread(&p[3 * i], sizeof(double), 3, fp) == 3;
    }
    ok = ok && fread(rec, sizeof(int), 1, fp) == 1 && rec[0] == natoms * 52;
    fclose(fp);
    if (!ok) {
        free(x);
        free(p);
        return comd_host_fail(COMD_ERR_ARG, "%s is truncated", path);
    }
    *n = natoms;
    *xyz = x;
    *pxyz = p;
    return COMD_OK;
}
