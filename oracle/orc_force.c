/*
 * orc_force.c -- oracle restatement of the EAM (three passes, table or
 * Chebyshev evaluation) and Lennard-Jones forces of CoMD src-flat.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Conventions kept from the reference:
 *   EAM f holds +grad U (src-flat/eam.c:535-536, 645-646);
 *   LJ f holds -grad U (src-flat/ljForce.c:207-208);
 *   EAM pass-1 skips r2 > rc2, pass 3 skips r2 >= rc2, LJ skips r2 > rc2;
 *   per-atom LJ energy lacks the 4 eps sigma^6 factor (ljForce.c:200-201);
 *   "stress" is the scalar partial virial of eam.c:649 / ljForce.c:175,210.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* Minimum-imaged difference of two cell corners, src-flat/eam.c:495-502. */
static void cell_shift(const orc_sim *s, int ib, int jb, double d[3])
{
    for (int j = 0; j < 3; j++) {
        d[j] = s->dcenter[3 * ib + j] - s->dcenter[3 * jb + j];
        if (s->periodic) {
            if (d[j] < -0.5 * s->bounds[j]) d[j] += s->bounds[j];
            else if (d[j] > 0.5 * s->bounds[j]) d[j] -= s->bounds[j];
        }
    }
}

/* src-flat/eam.c:438-660. */
int orc_eam_force(orc_sim *s)
{
    const orc_pot *pot = s->pot;
    const int cheb = pot->eval == ORC_EVAL_CHEB;
    double rc2 = pot->cutoff * pot->cutoff;
    size_t ns = (size_t)s->nboxes * ORC_MAXATOMS;
    double etot = 0.0;
    memset(s->f, 0, ns * 4 * sizeof(double));
    memset(s->fi, 0, ns * sizeof(double));
    memset(s->rho, 0, ns * sizeof(double));
    s->stress = 0.0;
    int nbr[27];

    /* pass 1: pair energy, pair force, density (eam.c:484-549) */
    for (int ib = 0; ib < s->nboxes; ib++) {
        int ni = s->natoms[ib];
        int nn = orc_neighbor_boxes(s, ib, nbr);
        for (int t = 0; t < nn; t++) {
            int jb = nbr[t];
            if (jb < ib) continue;
            double drb[3];
            cell_shift(s, ib, jb, drb);
            int nj = s->natoms[jb];
            for (int ii = 0; ii < ni; ii++) {
                int io = ib * ORC_MAXATOMS + ii;
                for (int jj = 0; jj < nj; jj++) {
                    int jo = jb * ORC_MAXATOMS + jj;
                    if (ib == jb && jj <= ii) continue;
                    double dr[3], r2 = 0.0;
                    for (int k = 0; k < 3; k++) {
                        dr[k] = drb[k] + s->r[3 * io + k] - s->r[3 * jo + k];
                        r2 += dr[k] * dr[k];
                    }
                    if (r2 > rc2) continue;
                    double r = sqrt(r2);
                    double phiv, dphi, rhov, drho;
                    if (cheb) {
                        phiv = orc_cheb_eval(&pot->cphi, r);
                        dphi = orc_cheb_eval(&pot->cdphi, r);
                        rhov = orc_cheb_eval(&pot->crho, r);
                        drho = orc_cheb_eval(&pot->cdrho, r);
                    } else {
                        orc_table_interp(&pot->phi, r, &phiv, &dphi);
                        orc_table_interp(&pot->rho, r, &rhov, &drho);
                    }
                    (void)drho;
                    for (int k = 0; k < 3; k++) {
                        s->f[4 * io + k] += dphi * dr[k] / r;
                        s->f[4 * jo + k] -= dphi * dr[k] / r;
                    }
                    etot += phiv;
                    s->f[4 * io + 3] += phiv / 2.0;
                    s->f[4 * jo + 3] += phiv / 2.0;
                    s->rho[io] += rhov;
                    s->rho[jo] += rhov;
                }
            }
        }
    }

    /* pass 2: embedding energy and its derivative (eam.c:576-594) */
    for (int ib = 0; ib < s->nboxes; ib++) {
        for (int ii = 0; ii < s->natoms[ib]; ii++) {
            int io = ib * ORC_MAXATOMS + ii;
            double fv, fp;
            if (cheb) {
                fv = orc_cheb_eval(&pot->cf, s->rho[io]);
                fp = orc_cheb_eval(&pot->cdf, s->rho[io]);
            } else {
                orc_table_interp(&pot->f, s->rho[io], &fv, &fp);
            }
            s->fi[io] = fp;
            etot += fv;
            s->f[4 * io + 3] += fv;
        }
    }

    /* pass 3: embedding force and x-virial, atom-outer order (eam.c:595-653) */
    for (int ib = 0; ib < s->nboxes; ib++) {
        int ni = s->natoms[ib];
        int nn = orc_neighbor_boxes(s, ib, nbr);
        for (int ii = 0; ii < ni; ii++) {
            int io = ib * ORC_MAXATOMS + ii;
            for (int t = 0; t < nn; t++) {
                int jb = nbr[t];
                if (jb < ib) continue;
                double drb[3];
                cell_shift(s, ib, jb, drb);
                int nj = s->natoms[jb];
                for (int jj = 0; jj < nj; jj++) {
                    int jo = jb * ORC_MAXATOMS + jj;
                    if (ib == jb && jj <= ii) continue;
                    double dr[3], r2 = 0.0;
                    for (int k = 0; k < 3; k++) {
                        dr[k] = drb[k] + s->r[3 * io + k] - s->r[3 * jo + k];
                        r2 += dr[k] * dr[k];
                    }
                    if (r2 >= rc2) continue;
                    double r = sqrt(r2);
                    double rhov, drho;
                    if (cheb) drho = orc_cheb_eval(&pot->cdrho, r);
                    else orc_table_interp(&pot->rho, r, &rhov, &drho);
                    for (int k = 0; k < 3; k++) {
                        s->f[4 * io + k] += (s->fi[io] + s->fi[jo]) * drho * dr[k] / r;
                        s->f[4 * jo + k] -= (s->fi[io] + s->fi[jo]) * drho * dr[k] / r;
                    }
                    s->stress += 2.0 * (s->fi[io] + s->fi[jo]) * drho * dr[0] / r * dr[0];
                }
            }
        }
    }
    s->e = etot;
    s->stress = s->stress / (s->bounds[0] * s->bounds[1] * s->bounds[2]);
    return 0;
}

/* src-flat/ljForce.c:89-224: all 27 stencil cells, pair half-list by atom
 * id, kinetic stress term once per non-empty neighbour cell. */
int orc_lj_force(orc_sim *s)
{
    const orc_pot *pot = s->pot;
    double sigma = pot->sigma, epsilon = pot->epsilon;
    double rcut = pot->cutoff;
    double r2cut = rcut * rcut;
    double etot = 0.0;
    size_t ns = (size_t)s->nboxes * ORC_MAXATOMS;
    s->e = 0.0;
    memset(s->f, 0, ns * 4 * sizeof(double));
    double s6 = sigma * sigma * sigma * sigma * sigma * sigma;
    s->stress = 0.0;
    int nbr[27];
    for (int ib = 0; ib < s->nboxes; ib++) {
        int ni = s->natoms[ib];
        if (!ni) continue;
        int nn = orc_neighbor_boxes(s, ib, nbr);
        for (int t = 0; t < nn; t++) {
            int jb = nbr[t];
            double drb[3];
            cell_shift(s, ib, jb, drb);
            int nj = s->natoms[jb];
            if (!nj) continue;
            for (int ii = 0; ii < ni; ii++) {
                int io = ib * ORC_MAXATOMS + ii;
                s->stress -= s->p[3 * io] * s->p[3 * io] / s->mass[io];
                int idi = s->id[io];
                for (int jj = 0; jj < nj; jj++) {
                    int jo = jb * ORC_MAXATOMS + jj;
                    int idj = s->id[jo];
                    if (idj <= idi) continue;
                    double dr[3], r2 = 0.0;
                    for (int m = 0; m < 3; m++) {
                        dr[m] = drb[m] + s->r[3 * io + m] - s->r[3 * jo + m];
                        r2 += dr[m] * dr[m];
                    }
                    if (r2 > r2cut) continue;
                    r2 = 1.0 / r2;
                    double r6 = r2 * r2 * r2;
                    s->f[4 * io + 3] += 0.5 * r6 * (s6 * r6 - 1.0);
                    s->f[4 * jo + 3] += 0.5 * r6 * (s6 * r6 - 1.0);
                    etot += r6 * (s6 * r6 - 1.0);
                    double fr = 4.0 * epsilon * s6 * r6 * r2 * (12.0 * s6 * r6 - 6.0);
                    for (int m = 0; m < 3; m++) {
                        s->f[4 * io + m] += dr[m] * fr;
                        s->f[4 * jo + m] -= dr[m] * fr;
                    }
                    s->stress += 2.0 * fr * dr[0] * dr[0];
                }
            }
        }
    }
    etot = etot * 4.0 * epsilon * s6;
    s->e = etot;
    s->stress = s->stress / (s->bounds[0] * s->bounds[1] * s->bounds[2]);
    if (pot->lj_sign < 0.0)
        for (size_t o = 0; o < ns; o++)
            for (int m = 0; m < 3; m++) s->f[4 * o + m] = -s->f[4 * o + m];
    return 0;
}

/* src-flat/timestep.c:110-112 (the potential's force slot). */
int orc_compute_force(orc_sim *s)
{
    return s->pot->kind == ORC_POT_EAM ? orc_eam_force(s) : orc_lj_force(s);
}

/* Virial tensor of the current state, W_ab = sum_i p_ia p_ib / m +
 * sum_{i<j} r_ij,a f_ij,b with the physical pair force on i from j (EAM:
 * -(phi'(r) + (F'_i + F'_j) rho'(r)) r_ij / r with F' of the last force
 * call; LJ: -U'(r) r_ij / r), over the half-shell pairs of eamForce's loops
 * (src-flat/eam.c:484-519, 595-633: jBox >= iBox, same box jj > ii) with
 * the cutoff predicates of passes 1 (phi) and 3 (rho'), or of LJ
 * (ljForce.c:188).  The reference computes only its scalar "stress"; this
 * is the tensor it approximates.  out: W row-major (9), then the pair part
 * alone (9).  Needs at least 3 cells per dimension (no repeated stencil
 * cells). */
void orc_virial_tensor(const orc_sim *s, double *out)
{
    const orc_pot *pot = s->pot;
    const double rc2 = pot->cutoff * pot->cutoff;
    const double rmass = 1822.83 * pot->mass;  /* src-flat/timestep.c:90 */
    double w[3][3] = {{0}}, k[3][3] = {{0}};
    int nbr[27];
    for (int ib = 0; ib < s->nboxes; ib++) {
        int ni = s->natoms[ib];
        for (int ii = 0; ii < ni; ii++) {
            int io = ib * ORC_MAXATOMS + ii;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) k[a][b] += s->p[3 * io + a] * s->p[3 * io + b] / rmass;
        }
        int nn = orc_neighbor_boxes(s, ib, nbr);
        for (int t = 0; t < nn; t++) {
            int jb = nbr[t];
            if (jb < ib) continue;
            double drb[3];
            cell_shift(s, ib, jb, drb);
            int nj = s->natoms[jb];
            for (int ii = 0; ii < ni; ii++) {
                int io = ib * ORC_MAXATOMS + ii;
                for (int jj = (jb == ib ? ii + 1 : 0); jj < nj; jj++) {
                    int jo = jb * ORC_MAXATOMS + jj;
                    double dr[3], r2 = 0.0;
                    for (int m = 0; m < 3; m++) {
                        dr[m] = drb[m] + s->r[3 * io + m] - s->r[3 * jo + m];
                        r2 += dr[m] * dr[m];
                    }
                    if (r2 > rc2) continue;
                    double kk;
                    if (pot->kind == ORC_POT_LJ) {
                        double s6 = pot->sigma * pot->sigma * pot->sigma * pot->sigma * pot->sigma * pot->sigma;
                        double r2i = 1.0 / r2, r6 = r2i * r2i * r2i;
                        kk = 4.0 * pot->epsilon * s6 * r6 * r2i * (12.0 * s6 * r6 - 6.0);
                    } else {
                        double r = sqrt(r2), v, dphi, drho = 0.0;
                        orc_pot_eval(pot, 0, r, &v, &dphi);
                        if (r2 < rc2) orc_pot_eval(pot, 1, r, &v, &drho);
                        kk = -(dphi + (s->fi[io] + s->fi[jo]) * drho) / r;
                    }
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) w[a][b] += kk * dr[a] * dr[b];
                }
            }
        }
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            out[3 * a + b] = w[a][b] + k[a][b];
            out[9 + 3 * a + b] = w[a][b];
        }
}
