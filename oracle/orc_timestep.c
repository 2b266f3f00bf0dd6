/*
 * orc_timestep.c -- oracle restatement of CoMD src-flat's position-Verlet
 * driver and of the two initial-condition paths (FCC lattice, gz input).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* src-flat/include/constants.h:61-65 */
static const double kBohrPerAtuToAPerS = (double)0.529 / (double)2.418e-17;
static const double kAmuToMe = (double)1822.83;

/* p -= dt f, src-flat/timestep.c:62-74 */
void orc_advance_velocity(orc_sim *s, double dt)
{
    for (int b = 0; b < s->nboxes; b++)
        for (int i = 0; i < s->natoms[b]; i++) {
            int o = b * ORC_MAXATOMS + i;
            for (int k = 0; k < 3; k++) s->p[3 * o + k] -= dt * s->f[4 * o + k];
        }
}

/* r += (dt p)/rmass then rebin, src-flat/timestep.c:76-108 */
void orc_advance_positions(orc_sim *s, double dt)
{
    if (s->err) return;
    double rmass = (double)(kAmuToMe * (double)s->pot->mass);
    for (int b = 0; b < s->nboxes; b++)
        for (int i = 0; i < s->natoms[b]; i++) {
            int o = b * ORC_MAXATOMS + i;
            for (int k = 0; k < 3; k++) s->r[3 * o + k] += dt * s->p[3 * o + k] / rmass;
        }
    orc_rebox_all(s);
}

/* src-flat/timestep.c:114-140: n DKD steps, then one more force call.
 * Stops after a rebin overflow (the reference exits there). */
double orc_n_time_steps(orc_sim *s, int n, double dt)
{
    dt = dt * kBohrPerAtuToAPerS;
    for (int i = 0; i < n && !s->err; i++) {
        orc_advance_positions(s, dt / 2.0);
        orc_compute_force(s);
        orc_advance_velocity(s, dt);
        orc_advance_positions(s, dt / 2.0);
    }
    if (!s->err) orc_compute_force(s);
    return s->e;
}

/* Sum |p|^2 / (2 rmass) in hartree; the reference has no kinetic energy,
 * this follows from its units (r in A, dt in A*atu/bohr, p = rmass v). */
double orc_kinetic_energy(const orc_sim *s)
{
    double rmass = kAmuToMe * s->pot->mass;
    double ke = 0.0;
    for (int b = 0; b < s->nboxes; b++)
        for (int i = 0; i < s->natoms[b]; i++) {
            int o = b * ORC_MAXATOMS + i;
            double p2 = s->p[3 * o] * s->p[3 * o] + s->p[3 * o + 1] * s->p[3 * o + 1] + s->p[3 * o + 2] * s->p[3 * o + 2];
            ke += p2 / (2.0 * rmass);
        }
    return ke;
}

/* Half-lattice walk from a/4 with parity (i+j+k) odd, accumulating the
 * coordinates by repeated addition; ids in creation order; zero momenta;
 * type 1.  src-flat/ic_fcc.c:63-141 */
orc_sim *orc_create_fcc(const orc_pot *pot, int nx, int ny, int nz, double lat,
                        double defgrad, double bf, int periodic)
{
    double bounds[3] = {nx * lat * defgrad, ny * lat, nz * lat};
    orc_sim *s = orc_sim_create(pot, bounds, defgrad, bf, periodic);
    if (!s || s->err) return s;
    double halflat = lat / 2.0;
    int i = 0, j = 0, k = 0, n = 0;
    double z = lat / 4.0;
    while (z < s->bounds[2]) {
        double y = lat / 4.0;
        while (y < s->bounds[1]) {
            double x = lat * defgrad / 4.0;
            while (x < s->bounds[0]) {
                if ((i + j + k) % 2) {
                    if (orc_put_atom(s, n++, 1, pot->mass, x, y, z, 0.0, 0.0, 0.0)) return s;
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
    return s;
}

/* clsman-style input: "natoms nmove", a comment line, the bounds, then one
 * "x y z itype px py pz" line per atom; id = line order, type 0, mass from
 * the potential; the grid ignores any stretch (defgrad 0).
 * src-flat/read.c:155-241 */
orc_sim *orc_read_gz(const orc_pot *pot, const char *path, double bf, int periodic)
{
    gzFile fp = gzopen(path, "rb");
    if (!fp) return NULL;
    char line[16384];
    int natoms = 0, nmove = 0;
    double bounds[3] = {0, 0, 0};
    if (!gzgets(fp, line, sizeof line) || sscanf(line, "%d%d", &natoms, &nmove) < 1) { gzclose(fp); return NULL; }
    if (!gzgets(fp, line, sizeof line)) { gzclose(fp); return NULL; }
    if (!gzgets(fp, line, sizeof line) || sscanf(line, "%lg %lg %lg", &bounds[0], &bounds[1], &bounds[2]) != 3) {
        gzclose(fp);
        return NULL;
    }
    orc_sim *s = orc_sim_create(pot, bounds, 0.0, bf, periodic);
    if (!s || s->err) { gzclose(fp); return s; }
    for (int i = 0; i < natoms; i++) {
        double x, y, z, px, py, pz;
        int it;
        if (!gzgets(fp, line, sizeof line) || sscanf(line, "%lg %lg %lg %d %lg %lg %lg", &x, &y, &z, &it, &px, &py, &pz) != 7) {
            s->err = ORC_ERR_NOFILE;
            break;
        }
        if (orc_put_atom(s, i, 0, pot->mass, x, y, z, px, py, pz)) break;
    }
    gzclose(fp);
    return s;
}

/* Replace the whole slot-layout state (natoms, id, itype, r, p). */
void orc_sim_set_state(orc_sim *s, const int *natoms, const int *id, const int *itype, const double *r, const double *p)
{
    size_t nb = (size_t)s->nboxes, ns = nb * ORC_MAXATOMS;
    memcpy(s->natoms, natoms, nb * sizeof(int));
    memcpy(s->id, id, ns * sizeof(int));
    memcpy(s->itype, itype, ns * sizeof(int));
    memcpy(s->r, r, ns * 3 * sizeof(double));
    memcpy(s->p, p, ns * 3 * sizeof(double));
    s->ntot = 0;
    for (size_t b = 0; b < nb; b++) {
        s->ntot += natoms[b];
        for (int i = 0; i < ORC_MAXATOMS; i++) s->mass[b * ORC_MAXATOMS + i] = s->pot->mass;
    }
}

/* The same insertion sequence as orc_read_gz, from arrays in memory. */
orc_sim *orc_from_atoms(const orc_pot *pot, const double bounds[3], double bf, int periodic,
                        int n, const double *xyz, const double *pxyz, int itype)
{
    orc_sim *s = orc_sim_create(pot, bounds, 0.0, bf, periodic);
    if (!s || s->err) return s;
    for (int i = 0; i < n; i++)
        if (orc_put_atom(s, i, itype, pot->mass, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2],
                         pxyz ? pxyz[3 * i] : 0.0, pxyz ? pxyz[3 * i + 1] : 0.0, pxyz ? pxyz[3 * i + 2] : 0.0))
            break;
    return s;
}
