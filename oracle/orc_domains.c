/*
 * orc_domains.c -- oracle restatement of the link-cell grid of CoMD
 * src-flat/domains.c: grid sizing, atom insertion, rebinning with
 * swap-with-last removal, and the 27-cell stencil.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Grid sizing and cell corners, src-flat/domains.c:155-200.  The strain
 * is (defgrad, 1, 1) with defgrad 0 read as 1 (domains.c:165-167), the
 * box factor 0 read as 1 (171-172).  dcenter is the LOWER corner. */
orc_sim *orc_sim_create(const orc_pot *pot, const double bounds[3], double defgrad, double bf, int periodic)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    if (!s) return NULL;
    s->pot = pot;
    s->defgrad = defgrad;
    s->bf = bf;
    s->periodic = periodic;
    double strain[3] = {defgrad, 1.0, 1.0};
    if (strain[0] == 0.0) strain[0] = 1.0;
    double boxfactor = bf == 0.0 ? 1.0 : bf;
    s->nboxes = 1;
    for (int j = 0; j < 3; j++) {
        s->bounds[j] = bounds[j];
        s->nbx[j] = (int)floor(bounds[j] / (pot->cutoff * boxfactor * strain[j]));
        if (s->nbx[j] < 1) { s->err = ORC_ERR_GRID; return s; }
        s->nboxes *= s->nbx[j];
        s->boxsize[j] = bounds[j] / (double)s->nbx[j];
    }
    size_t nb = (size_t)s->nboxes, ns = nb * ORC_MAXATOMS;
    s->dcenter = (double *)calloc(nb * 3, sizeof(double));
    s->natoms = (int *)calloc(nb, sizeof(int));
    s->id = (int *)calloc(ns, sizeof(int));
    s->itype = (int *)calloc(ns, sizeof(int));
    s->mass = (double *)calloc(ns, sizeof(double));
    s->r = (double *)calloc(ns * 3, sizeof(double));
    s->p = (double *)calloc(ns * 3, sizeof(double));
    s->f = (double *)calloc(ns * 4, sizeof(double));
    s->rho = (double *)calloc(ns, sizeof(double));
    s->fi = (double *)calloc(ns, sizeof(double));
    if (!s->dcenter || !s->natoms || !s->id || !s->itype || !s->mass || !s->r || !s->p || !s->f || !s->rho || !s->fi) {
        s->err = ORC_ERR_ALLOC;
        return s;
    }
    for (int b = 0; b < s->nboxes; b++) {
        int ix = b % s->nbx[0];
        int iy = (b / s->nbx[0]) % s->nbx[1];
        int iz = b / s->nbx[0] / s->nbx[1];
        s->dcenter[3 * b + 0] = s->boxsize[0] * (double)ix;
        s->dcenter[3 * b + 1] = s->boxsize[1] * (double)iy;
        s->dcenter[3 * b + 2] = s->boxsize[2] * (double)iz;
    }
    return s;
}

void orc_sim_free(orc_sim *s)
{
    if (!s) return;
    free(s->dcenter); free(s->natoms); free(s->id); free(s->itype); free(s->mass);
    free(s->r); free(s->p); free(s->f); free(s->rho); free(s->fi);
    free(s);
}

/* src-flat/domains.c:237-291 with getBoxIDWorldCoords (71-82): wrap once
 * into [0,L), cell = floor(x/boxsize), append at the cell's end, store the
 * position relative to the cell corner.  The reference does not check the
 * slot count or the cell range here; the oracle reports both as errors. */
int orc_put_atom(orc_sim *s, int id, int itype, double mass,
                 double x, double y, double z, double px, double py, double pz)
{
    double xyz[3] = {x, y, z};
    for (int m = 0; m < 3; m++) {
        if (xyz[m] < 0.0) xyz[m] += s->bounds[m];
        else if (xyz[m] >= s->bounds[m]) xyz[m] -= s->bounds[m];
    }
    int ib[3];
    for (int m = 0; m < 3; m++) ib[m] = (int)floor(xyz[m] / s->boxsize[m]);
    for (int m = 0; m < 3; m++)
        if (ib[m] < 0 || ib[m] >= s->nbx[m]) { s->err = ORC_ERR_OUTSIDE; return s->err; }
    int ibox = ib[0] + s->nbx[0] * ib[1] + s->nbx[0] * s->nbx[1] * ib[2];
    if (s->natoms[ibox] >= ORC_MAXATOMS) { s->err = ORC_ERR_OVERFLOW; return s->err; }
    int off = ibox * ORC_MAXATOMS + s->natoms[ibox];
    s->ntot++;
    s->natoms[ibox]++;
    s->id[off] = id;
    s->itype[off] = itype;
    s->mass[off] = mass;
    for (int m = 0; m < 3; m++) s->r[3 * off + m] = xyz[m] - s->dcenter[3 * ibox + m];
    s->p[3 * off + 0] = px;
    s->p[3 * off + 1] = py;
    s->p[3 * off + 2] = pz;
    return ORC_OK;
}

/* Periodic wrap of a cell triple, or -1 outside a non-periodic box,
 * src-flat/domains.c:297-325. */
static int cell_of(const orc_sim *s, int ix[3])
{
    if (!s->periodic)
        for (int j = 0; j < 3; j++)
            if (ix[j] < 0 || ix[j] >= s->nbx[j]) return -1;
    for (int j = 0; j < 3; j++) {
        if (ix[j] < 0) ix[j] += s->nbx[j];
        else if (ix[j] >= s->nbx[j]) ix[j] -= s->nbx[j];
    }
    return ix[0] + ix[1] * s->nbx[0] + ix[2] * s->nbx[0] * s->nbx[1];
}

static void copy_slot(orc_sim *s, int from, int to)
{
    s->id[to] = s->id[from];
    s->itype[to] = s->itype[from];
    s->mass[to] = s->mass[from];
    memcpy(&s->r[3 * to], &s->r[3 * from], 3 * sizeof(double));
    memcpy(&s->f[4 * to], &s->f[4 * from], 4 * sizeof(double));
    memcpy(&s->p[3 * to], &s->p[3 * from], 3 * sizeof(double));
    s->fi[to] = s->fi[from];
    s->rho[to] = s->rho[from];
}

/* Append to the destination, then fill the hole with the source's last
 * atom, src-flat/domains.c:224-236.  The reference aborts the process when
 * the destination reaches 64 atoms (domains.c:229); here the error is
 * recorded and the caller stops at once, so no slot past a cell is ever
 * written. */
static int move_atom(orc_sim *s, int i, int ibox, int jbox)
{
    int nj = s->natoms[jbox];
    copy_slot(s, ibox * ORC_MAXATOMS + i, jbox * ORC_MAXATOMS + nj);
    s->natoms[jbox]++;
    if (s->natoms[jbox] >= ORC_MAXATOMS) {
        s->err = ORC_ERR_OVERFLOW;
        return 1;
    }
    s->natoms[ibox]--;
    int ni = s->natoms[ibox];
    if (ni) copy_slot(s, ibox * ORC_MAXATOMS + ni, ibox * ORC_MAXATOMS + i);
    return 0;
}

/* Ascending scan of cells; an atom outside its cell moves by one cell per
 * dimension with its relative coordinate shifted by one box size, and the
 * slot is re-examined after the swap, src-flat/domains.c:347-389. */
void orc_rebox_all(orc_sim *s)
{
    if (s->err) return;
    for (int ibox = 0; ibox < s->nboxes; ibox++) {
        int old[3] = {ibox % s->nbx[0], (ibox / s->nbx[0]) % s->nbx[1], (ibox / s->nbx[0] / s->nbx[1]) % s->nbx[2]};
        int i = 0;
        int off = ibox * ORC_MAXATOMS;
        while (i < s->natoms[ibox]) {
            int nw[3];
            double rn[3];
            memcpy(rn, &s->r[3 * off], sizeof rn);
            for (int k = 0; k < 3; k++) {
                if (s->r[3 * off + k] < 0.0) {
                    nw[k] = old[k] - 1;
                    rn[k] += s->boxsize[k];
                } else if (s->r[3 * off + k] >= s->boxsize[k]) {
                    nw[k] = old[k] + 1;
                    rn[k] -= s->boxsize[k];
                } else {
                    nw[k] = old[k];
                }
            }
            int jbox = cell_of(s, nw);
            if (jbox < 0 || jbox == ibox) {
                i++;
                off++;
            } else {
                memcpy(&s->r[3 * off], rn, sizeof rn);
                if (move_atom(s, i, ibox, jbox)) return;
            }
        }
    }
}

/* 27-cell stencil, x outer, y, z inner, with periodic wrap; a single-cell
 * grid returns just itself, src-flat/domains.c:391-430. */
int orc_neighbor_boxes(const orc_sim *s, int ibox, int *nbrs)
{
    if (s->nboxes == 1) {
        nbrs[0] = 0;
        return 1;
    }
    int ix = ibox % s->nbx[0];
    int iy = (ibox / s->nbx[0]) % s->nbx[1];
    int iz = (ibox / s->nbx[0] / s->nbx[1]) % s->nbx[2];
    int n = 0;
    for (int i = ix - 1; i <= ix + 1; i++)
        for (int j = iy - 1; j <= iy + 1; j++)
            for (int k = iz - 1; k <= iz + 1; k++) {
                int c[3] = {i, j, k};
                int b = cell_of(s, c);
                if (b >= 0) nbrs[n++] = b;
            }
    return n;
}

void orc_sim_info(const orc_sim *s, int *nbx3, int *nboxes, int *ntot,
                  double *bounds3, double *boxsize3, double *e, double *stress, int *err)
{
    for (int j = 0; j < 3; j++) {
        nbx3[j] = s->nbx[j];
        bounds3[j] = s->bounds[j];
        boxsize3[j] = s->boxsize[j];
    }
    *nboxes = s->nboxes;
    *ntot = s->ntot;
    *e = s->e;
    *stress = s->stress;
    *err = s->err;
}

void orc_sim_get(const orc_sim *s, int *natoms, int *id, int *itype, double *r, double *p,
                 double *f, double *rho, double *fi)
{
    size_t nb = (size_t)s->nboxes, ns = nb * ORC_MAXATOMS;
    if (natoms) memcpy(natoms, s->natoms, nb * sizeof(int));
    if (id) memcpy(id, s->id, ns * sizeof(int));
    if (itype) memcpy(itype, s->itype, ns * sizeof(int));
    if (r) memcpy(r, s->r, ns * 3 * sizeof(double));
    if (p) memcpy(p, s->p, ns * 3 * sizeof(double));
    if (f) memcpy(f, s->f, ns * 4 * sizeof(double));
    if (rho) memcpy(rho, s->rho, ns * sizeof(double));
    if (fi) memcpy(fi, s->fi, ns * sizeof(double));
}

void orc_sim_set_rp(orc_sim *s, const double *r, const double *p)
{
    size_t ns = (size_t)s->nboxes * ORC_MAXATOMS;
    if (r) memcpy(s->r, r, ns * 3 * sizeof(double));
    if (p) memcpy(s->p, p, ns * 3 * sizeof(double));
}
