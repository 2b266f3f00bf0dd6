// comd_common.h -- device-state layout and shared helpers of the gfx950
// link-cell library (internal to comd_amd/csrc).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/comd_hip.h"

#define CM_SLOTS 64  // slots per link cell (src-flat/include/pmdTypes.h:57)

// Error bits raised by kernels into the context's device flag word.
enum {
    CM_FLAG_OVERFLOW = 1,  // a cell would reach 64 atoms (src-flat/domains.c:229)
    CM_FLAG_OUTSIDE = 2,   // atom outside the grid when binned
    CM_FLAG_MULTIHOP = 4,  // an arrival would move again in the same rebin pass
};

// Plain view of the device arrays passed to kernels by value.
struct DevState {
    int nbx0, nbx1, nbx2, ncells;
    int periodic;
    double bs0, bs1, bs2;      // boxsize
    double L0, L1, L2;         // bounds
    // moving state (slot layout, updated in place by the rebin)
    int *natoms;
    int *id;
    int *itype;
    double *rx, *ry, *rz;
    double *px, *py, *pz;
    // force outputs (slot layout)
    double *fx, *fy, *fz, *fe, *rho, *fi, *vir;
    int *nnb;           // per slot: neighbours inside the cutoff (last density/LJ pass)
    // rebin scratch
    int *dest;          // per slot: destination cell, or -1 to stay
    int *mvidx;         // per slot: index of its mover record (departing atoms)
    uint8_t *perm;      // per cell: arrangement after the cell's own scan
    int *meta;          // per cell: (len, nlow, nhigh, n0)
    int *counts;        // [0] movers, [1] touched cells of the current pass
    int *touched;       // per cell: 1 while listed in tc
    int *tc;            // touched cells of the current pass
    int *arr_cnt;       // per cell: arrivals of the current pass
    int *arr;           // per cell: up to 64 arriving mover indices
    // mover records: source slot, departure rank, and the atom with its
    // shifted coordinates
    int *mv_slot, *mv_rank, *mv_id, *mv_itype;
    double *mv_rx, *mv_ry, *mv_rz, *mv_px, *mv_py, *mv_pz;
    unsigned *flags;    // error flag word
};

__device__ __forceinline__ void cell_xyz(const DevState &d, int c, int &ix, int &iy, int &iz)
{
    ix = c % d.nbx0;
    iy = (c / d.nbx0) % d.nbx1;
    iz = (c / d.nbx0 / d.nbx1) % d.nbx2;
}

// Periodic wrap of a cell triple, or -1 outside a non-periodic box
// (src-flat/domains.c:297-325).
__device__ __forceinline__ int cell_wrap(const DevState &d, int ix, int iy, int iz)
{
    if (!d.periodic) {
        if (ix < 0 || ix >= d.nbx0 || iy < 0 || iy >= d.nbx1 || iz < 0 || iz >= d.nbx2) return -1;
    }
    if (ix < 0) ix += d.nbx0; else if (ix >= d.nbx0) ix -= d.nbx0;
    if (iy < 0) iy += d.nbx1; else if (iy >= d.nbx1) iy -= d.nbx1;
    if (iz < 0) iz += d.nbx2; else if (iz >= d.nbx2) iz -= d.nbx2;
    return ix + iy * d.nbx0 + iz * d.nbx0 * d.nbx1;
}

// Stencil entry e (0..26) of cell c in the reference's order: x outer, y,
// z inner (src-flat/domains.c:419-426); -1 outside a non-periodic box.
__device__ __forceinline__ int stencil_cell(const DevState &d, int c, int e)
{
    if (d.ncells == 1) return e == 0 ? 0 : -1;
    int ix, iy, iz;
    cell_xyz(d, c, ix, iy, iz);
    int ox = e / 9 - 1, oy = (e / 3) % 3 - 1, oz = e % 3 - 1;
    return cell_wrap(d, ix + ox, iy + oy, iz + oz);
}

// dcenter[c][k] = boxsize[k] * ix_k (src-flat/domains.c:191-197).  The
// corner arithmetic is never contracted into fma, even in files compiled
// with contraction: the shift of a cell against itself must be exactly 0
// (fma(b, i, -b*i) is the rounding error of b*i, not 0), or an atom would
// see itself as a neighbour at distance ~1e-16.
__device__ __forceinline__ void cell_corner(const DevState &d, int c, double &x, double &y, double &z)
{
#pragma clang fp contract(off)
    int ix, iy, iz;
    cell_xyz(d, c, ix, iy, iz);
    x = d.bs0 * (double)ix;
    y = d.bs1 * (double)iy;
    z = d.bs2 * (double)iz;
}

// Minimum-imaged corner difference dcenter[i]-dcenter[j]
// (src-flat/eam.c:495-502).
__device__ __forceinline__ double image1(double dd, double L, int periodic)
{
#pragma clang fp contract(off)
    if (periodic) {
        if (dd < -0.5 * L) dd += L;
        else if (dd > 0.5 * L) dd -= L;
    }
    return dd;
}

// dcenter[c] - dcenter[b], minimum-imaged, uncontracted (0 for b == c).
__device__ __forceinline__ void corner_shift(const DevState &d, int c, int b, double &sx, double &sy, double &sz)
{
#pragma clang fp contract(off)
    double cx, cy, cz, bx, by, bz;
    cell_corner(d, c, cx, cy, cz);
    cell_corner(d, b, bx, by, bz);
    sx = image1(cx - bx, d.L0, d.periodic);
    sy = image1(cy - by, d.L1, d.periodic);
    sz = image1(cz - bz, d.L2, d.periodic);
}
