/*
 * orc_centro.c -- the centrosymmetry parameter of src-viz's
 * computeCentrosymmetry (src-viz/AtomVisualize.c:399-673), restated on the
 * oracle's slot layout.  TEST INFRASTRUCTURE ONLY: the checker of the GPU
 * kernel.
 *
 * Per atom, the 12 nearest neighbours: the atoms of its own cell first
 * (AtomVisualize.c:416-446, separations from r alone), then those of the
 * 26 surrounding cells in x-outer, y, z-inner order with periodic wrap
 * (448-503), where the image offset is the box length rounded to float
 * (465-472) and the separation is (offset + corner_j + r_j) - (corner_i +
 * r_i).  A candidate takes a free place while fewer than 12 are held, else
 * replaces the farthest held one (the first of equal ones) unless it is
 * farther (rsq > rmax).  With fewer than 12, the parameter is 1000
 * (508).  Otherwise it is the smallest, over all 10395 ways of pairing the
 * 12 neighbours, of sum over pairs |r_j + r_k|^2, each pairing's sum
 * accumulated in float, pairs taken in increasing order of their first
 * index (the exact branch, 540-651).  Cells are assumed to be at least the
 * 12-neighbour distance wide.
 */
#include <string.h>

#include "oracle.h"

typedef struct {
    double r[3], rsq;
} nbr_t;

static void offer(nbr_t *nl, int *n, const double d[3])
{
    const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    int i1;
    if (*n < 12) {
        i1 = (*n)++;
    } else {
        int kmax = 0;
        double rmax = nl[0].rsq;
        for (int k = 1; k < 12; k++)
            if (nl[k].rsq > rmax) {
                kmax = k;
                rmax = nl[k].rsq;
            }
        if (rsq > rmax) return;
        i1 = kmax;
    }
    nl[i1].rsq = rsq;
    memcpy(nl[i1].r, d, sizeof nl[i1].r);
}

/* the smallest pairing sum over the unpaired members of mask, the pairs
 * so far having summed to acc (float) */
static float best_pairing(const double s[12][12], int mask, float acc)
{
    if (mask == 0) return acc;
    const int n = __builtin_ctz(mask);
    const int rest = mask & ~(1 << n);
    float best = 3.0e38f;
    for (int m = rest; m; m &= m - 1) {
        const int p = __builtin_ctz(m);
        const float v = best_pairing(s, rest & ~(1 << p), (float)((double)acc + s[n][p]));
        if (v < best) best = v;
    }
    return best;
}

void orc_centrosymmetry(const orc_sim *s, double *out)
{
    const int nx = s->nbx[0], ny = s->nbx[1], nz = s->nbx[2];
    for (size_t q = 0; q < (size_t)s->nboxes * ORC_MAXATOMS; q++) out[q] = 0.0;
    for (int ib = 0; ib < s->nboxes; ib++) {
        const int bz = ib / (nx * ny), by = (ib - bz * nx * ny) / nx, bx = ib - bz * nx * ny - by * nx;
        for (int ii = 0; ii < s->natoms[ib]; ii++) {
            const int io = ib * ORC_MAXATOMS + ii;
            nbr_t nl[12];
            int n = 0;
            for (int jj = 0; jj < s->natoms[ib]; jj++) {
                const int jo = ib * ORC_MAXATOMS + jj;
                if (jo == io) continue;
                double d[3];
                for (int k = 0; k < 3; k++) d[k] = s->r[3 * jo + k] - s->r[3 * io + k];
                offer(nl, &n, d);
            }
            for (int cx = bx - 1; cx <= bx + 1; cx++)
                for (int cy = by - 1; cy <= by + 1; cy++)
                    for (int cz = bz - 1; cz <= bz + 1; cz++) {
                        if (cx == bx && cy == by && cz == bz) continue;
                        float off[3] = {0.0f, 0.0f, 0.0f};
                        int c[3] = {cx, cy, cz};
                        const int nb[3] = {nx, ny, nz};
                        for (int k = 0; k < 3; k++) {
                            if (c[k] < 0) { c[k] = nb[k] - 1; off[k] = (float)-s->bounds[k]; }
                            if (c[k] >= nb[k]) { c[k] = 0; off[k] = (float)s->bounds[k]; }
                        }
                        const int jb = c[2] * nx * ny + c[1] * nx + c[0];
                        for (int jj = 0; jj < s->natoms[jb]; jj++) {
                            const int jo = jb * ORC_MAXATOMS + jj;
                            double d[3];
                            for (int k = 0; k < 3; k++)
                                d[k] = ((double)off[k] + s->dcenter[3 * jb + k] + s->r[3 * jo + k]) -
                                       (s->dcenter[3 * ib + k] + s->r[3 * io + k]);
                            offer(nl, &n, d);
                        }
                    }
            if (n < 12) {
                out[io] = 1000.0;
                continue;
            }
            double sp[12][12];
            for (int j = 0; j < 12; j++)
                for (int k = 0; k < 12; k++) {
                    const double a = nl[j].r[0] + nl[k].r[0], b = nl[j].r[1] + nl[k].r[1], e = nl[j].r[2] + nl[k].r[2];
                    sp[j][k] = a * a + b * b + e * e;
                }
            const float best = best_pairing(sp, 0xfff, 0.0f);
            out[io] = best < 1000.0f ? (double)best : 1000.0;
        }
    }
}
