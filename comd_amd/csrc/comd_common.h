// comd_common.h -- device-state layout and shared helpers of the gfx950
// link-cell library (internal to comd_amd/csrc).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/comd_hip.h"

#define CM_SLOTS 64  // slots per link cell (src-flat/include/pmdTypes.h:57)
#define CM_PLIST 24  // pair-list entries kept per lane of a cell (the force kernels' queue depth)

// Error bits raised by kernels into the context's device flag word.
enum {
    CM_FLAG_OVERFLOW = 1,  // a cell would reach 64 atoms (src-flat/domains.c:229)
    CM_FLAG_OUTSIDE = 2,   // atom outside the grid when binned
};

// Link-cell grid of one replica.  All replicas of a context share the slot
// arrays; replica r owns the global cells [cell_base, cell_base + ncells)
// and slot i of global cell c is c*64+i.
struct Grid {
    int nbx0, nbx1, nbx2, ncells, cell_base, periodic, pot;
    int nl_id_base;        // its atom ids i are neighbour-list ids nl_id_base + i
    double bs0, bs1, bs2;  // boxsize
    double L0, L1, L2;     // bounds
    double iL0, iL1, iL2;  // 1 / bounds
    double rmass;          // mass in the reference's internal unit (1822.83 m_amu)
};

// Plain view of the device arrays passed to kernels by value.
// Absolute position of an atom for the neighbour lists (angstrom).
struct CmPos {
    double x, y, z;
};

struct DevState {
    int ncells;              // cells of all replicas
    int ngrids;              // replicas
    Grid g0;                 // replica 0 (the only one of a single simulation)
    const Grid *grids;       // device copy of all replicas' grids
    const int *cell_grid;    // per global cell: its replica (used when ngrids > 1)
    // moving state (slot layout, updated in place by the rebin)
    int *natoms;
    int *id;
    int *itype;
    double *rx, *ry, *rz;
    double *px, *py, *pz;
    // force outputs (slot layout)
    double *fx, *fy, *fz, *fe, *rho, *fi, *vir;
    int *nnb;           // per slot: neighbours inside the cutoff (last density/LJ pass)
    float4 *nlr;        // per slot: the atom's FP32 position at the last list build (x, y, z)
                        // and its build position (w, int bits); moved with the atom
    // rebin scratch
    unsigned long long *leave;  // per cell: bit k set when slot k departs (cleared with touched)
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
    float4 *mv_nlr;
    double *mv_rx, *mv_ry, *mv_rz, *mv_px, *mv_py, *mv_pz;
    unsigned *flags;    // error flag word
    int *mh;            // per replica: an arrival of the current rebin pass would move again (the
                        // replica's pass is then replayed serially)
    // pair lists of the EAM density pass, replayed by the embedding-force
    // pass (null when not kept): per cell CM_PLIST x 64 halo indices
    // ([entry][lane]), per slot-lane the entry count, per cell 1 when no
    // lane's list overflowed
    uint16_t *plist;
    uint8_t *pcount;
    int *pvalid;
    // neighbour-list bookkeeping kept by the drift and rebin kernels (null
    // when lists are off): per build position the atom's absolute position;
    // per atom id its slot; a flag word whose bit CM_NL_REBUILD is raised
    // when an atom has moved more than half the skin
    CmPos *nl_pos;
    int *nl_slot;
    unsigned *nl_flags;
    double nl_half2;     // (skin / 2 - FP32 rounding of the reference positions)^2
    // when non-null, the link-cell force kernels run only if this word has
    // CM_NL_INVALID set (the lists could not be built)
    const unsigned *cell_gate;
};

enum { CM_NL_REBUILD = 1, CM_NL_INVALID = 2, CM_NL_OVERFLOW = 4 };

// Grid of the replica that owns global cell c.
__device__ __forceinline__ Grid grid_of(const DevState &d, int c)
{
    return d.ngrids == 1 ? d.g0 : d.grids[d.cell_grid[c]];
}

// Replica that owns global cell c.
__device__ __forceinline__ int rep_of(const DevState &d, int c) { return d.ngrids == 1 ? 0 : d.cell_grid[c]; }

__device__ __forceinline__ void cell_xyz(const Grid &g, int c, int &ix, int &iy, int &iz)
{
    const int l = c - g.cell_base;
    ix = l % g.nbx0;
    iy = (l / g.nbx0) % g.nbx1;
    iz = (l / g.nbx0 / g.nbx1) % g.nbx2;
}

// Periodic wrap of a cell triple to its global index, or -1 outside a
// non-periodic box (src-flat/domains.c:297-325).
__device__ __forceinline__ int cell_wrap(const Grid &g, int ix, int iy, int iz)
{
    if (!g.periodic) {
        if (ix < 0 || ix >= g.nbx0 || iy < 0 || iy >= g.nbx1 || iz < 0 || iz >= g.nbx2) return -1;
    }
    if (ix < 0) ix += g.nbx0; else if (ix >= g.nbx0) ix -= g.nbx0;
    if (iy < 0) iy += g.nbx1; else if (iy >= g.nbx1) iy -= g.nbx1;
    if (iz < 0) iz += g.nbx2; else if (iz >= g.nbx2) iz -= g.nbx2;
    return g.cell_base + ix + iy * g.nbx0 + iz * g.nbx0 * g.nbx1;
}

// Stencil entry e (0..26) of cell c in the reference's order: x outer, y,
// z inner (src-flat/domains.c:419-426); -1 outside a non-periodic box.  A
// one-cell grid sees only itself (domains.c:406-410).
__device__ __forceinline__ int stencil_cell(const Grid &g, int c, int e)
{
    if (g.ncells == 1) return e == 0 ? c : -1;
    int ix, iy, iz;
    cell_xyz(g, c, ix, iy, iz);
    const int ox = e / 9 - 1, oy = (e / 3) % 3 - 1, oz = e % 3 - 1;
    return cell_wrap(g, ix + ox, iy + oy, iz + oz);
}

// dcenter[c][k] = boxsize[k] * ix_k (src-flat/domains.c:191-197).  The
// corner arithmetic is never contracted into fma, even in files compiled
// with contraction: the shift of a cell against itself must be exactly 0
// (fma(b, i, -b*i) is the rounding error of b*i, not 0), or an atom would
// see itself as a neighbour at distance ~1e-16.
__device__ __forceinline__ void corner_of(const Grid &g, int ix, int iy, int iz, double &x, double &y, double &z)
{
#pragma clang fp contract(off)
    x = g.bs0 * (double)ix;
    y = g.bs1 * (double)iy;
    z = g.bs2 * (double)iz;
}

__device__ __forceinline__ void cell_corner(const Grid &g, int c, double &x, double &y, double &z)
{
    int ix, iy, iz;
    cell_xyz(g, c, ix, iy, iz);
    corner_of(g, ix, iy, iz, x, y, z);
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
__device__ __forceinline__ void corner_shift(const Grid &g, int c, int b, double &sx, double &sy, double &sz)
{
#pragma clang fp contract(off)
    double cx, cy, cz, bx, by, bz;
    cell_corner(g, c, cx, cy, cz);
    cell_corner(g, b, bx, by, bz);
    sx = image1(cx - bx, g.L0, g.periodic);
    sy = image1(cy - by, g.L1, g.periodic);
    sz = image1(cz - bz, g.L2, g.periodic);
}
