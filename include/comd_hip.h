/*
 * comd_hip.h -- C-ABI of the MI355X (gfx950) link-cell force and timestep
 * library (comd_amd/libcomd_hip.so).
 *
 * The library replaces the hot path of CoMD src-flat: the potential's force
 * slot `int (*force)(void *s)` (src-flat/include/pmdTypes.h:77, set at
 * src-flat/eam.c:394 and src-flat/ljForce.c:79, called from computeForce,
 * src-flat/timestep.c:110-112), the timestep driver `double nTimeSteps(int n,
 * simflat_t *s, real_t dt)` (src-flat/timestep.c:114-140) and the rebinning
 * `void reBoxAll(simflat_t *s)` (src-flat/domains.c:347-389).
 *
 * Only plain C types cross this boundary.  Host arrays use the reference's
 * slot layout: slot i of link cell c is index c*64+i (arrayAtBox,
 * src-flat/include/pmdTypes.h:100); r and p are 3 doubles per slot, f is 4
 * doubles per slot (fx, fy, fz, per-atom energy).  Every entry point returns
 * COMD_OK or a positive COMD_ERR_* code and never exits the process (the
 * reference calls exit() through simulationAbort, src-flat/utility.c:96-109).
 * A context owns one HIP stream; contexts are independent and reentrant.
 */
#ifndef COMD_HIP_H
#define COMD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden visibility; these entry points are its
 * only exported symbols. */
#pragma GCC visibility push(default)

#define COMD_MAXATOMS 64 /* slots per link cell, src-flat/include/pmdTypes.h:57 */

enum comd_status {
    COMD_OK = 0,
    COMD_ERR_ARG = 1,       /* bad argument or parameter */
    COMD_ERR_HIP = 2,       /* a HIP runtime call failed */
    COMD_ERR_NOMEM = 3,     /* device allocation failed */
    COMD_ERR_OVERFLOW = 4,  /* a link cell reached 64 atoms (src-flat/domains.c:229) */
    COMD_ERR_OUTSIDE = 5,   /* an atom fell outside the cell grid when binned */
    COMD_ERR_GRID = 6,      /* fewer than one cutoff-wide cell (src-flat/domains.c:180) */
    COMD_ERR_NODEVICE = 7,  /* no HIP device */
    COMD_ERR_STATE = 8,     /* call out of order (e.g. no atoms uploaded) */
};

enum comd_potential { COMD_POT_EAM = 1, COMD_POT_LJ = 2 };
enum comd_eval { COMD_EVAL_TABLE = 0, COMD_EVAL_CHEB = 1 };
enum comd_precision { COMD_PREC_FP64 = 0, COMD_PREC_MIXED = 1 };

/* Description of one simulation: potential and link-cell grid.
 * The grid fields are the values allocDomains computes
 * (src-flat/domains.c:155-200); comd_grid_for() fills them. */
typedef struct comd_params {
    int potential;          /* comd_potential */
    int eval;               /* comd_eval (EAM only) */
    int precision;          /* comd_precision */
    int periodic;           /* the reference's global PERIODIC (mycommand.c:67) */
    double cutoff;          /* A */
    double mass;            /* amu */
    /* Lennard-Jones (src-flat/ljForce.c:75-86) */
    double lj_sigma, lj_epsilon;
    /* EAM tables in the reference's potentialarray_t form
     * (src-flat/include/eamTypes.h:68-74): index 0 phi, 1 rho, 2 F.
     * x0 is the shifted origin (src-flat/eam.c:150); values points at n+3
     * doubles starting with the guard slot values[-1]. */
    int tab_n[3];
    double tab_x0[3], tab_xn[3], tab_invdx[3];
    const double *tab_values[3];
    /* Chebyshev series (src-flat/cheby.c:99-121): cheb_n coefficients each
     * for phi, rho, F, dphi, drho, dF (6*cheb_n doubles), with the domain
     * [cheb_x0[k], cheb_xn[k]] of phi, rho, F. */
    int cheb_n;
    const double *cheb_coef;
    double cheb_x0[3], cheb_xn[3];
    /* link-cell grid */
    int nbx[3];
    double bounds[3];
    double boxsize[3];
} comd_params;

typedef struct comd_ctx comd_ctx;

/* Grid of allocDomains (src-flat/domains.c:155-200): nbx = floor(bounds /
 * (cutoff*bf*strain)), strain = (defgrad,1,1), 0 read as 1 for defgrad and bf. */
int comd_grid_for(comd_params *p, const double bounds[3], double defgrad, double bf);

/* Create a context on the current HIP device; allocates device state for
 * nbx[0]*nbx[1]*nbx[2]*64 slots. */
int comd_create(const comd_params *p, comd_ctx **out);
int comd_destroy(comd_ctx *ctx);

/* Number of link cells and of atoms currently held. */
int comd_sizes(const comd_ctx *ctx, int *ncells, int *natoms_total);

/* Copy a slot-layout state to the device (the simflat_t arrays natoms, id,
 * iType, r, p of src-flat/include/pmdTypes.h:104-141). */
int comd_upload(comd_ctx *ctx, const int *natoms, const int *id, const int *itype,
                const double *r, const double *p);

/* Copy the device state back into slot-layout arrays; any pointer may be
 * NULL.  f, rho and fi are those of the last force evaluation. */
int comd_download(comd_ctx *ctx, int *natoms, int *id, int *itype, double *r, double *p,
                  double *f, double *rho, double *fi);

/* Bin n atoms given in world coordinates, in input order, with ids 0..n-1:
 * the sequence of putAtomInBox calls of src-flat/read.c:222 and
 * src-flat/ic_fcc.c:125 (wrap once into the box, cell = floor(x/boxsize),
 * append), done on the GPU as a stable counting sort.  xyz, pxyz: 3n doubles
 * in host memory.  itype is stored for every atom. */
int comd_bin_atoms(comd_ctx *ctx, int n, const double *xyz, const double *pxyz, int itype);

/* Same as comd_bin_atoms with the input already in device memory. */
int comd_bin_atoms_device(comd_ctx *ctx, int n, const double *d_xyz, const double *d_pxyz, int itype);

/* reBoxAll (src-flat/domains.c:347-389): same cell assignment and the same
 * in-cell order as the reference's sequential scan. */
int comd_rebox(comd_ctx *ctx);

/* One force evaluation, the potential's force slot (src-flat/eam.c:438-660,
 * src-flat/ljForce.c:89-224).  Results stay on the device. */
int comd_force(comd_ctx *ctx);

/* nTimeSteps (src-flat/timestep.c:114-140): n position-Verlet steps of dt
 * seconds, then one more force evaluation.  *pe_out (may be NULL) receives
 * the potential energy of that final force, as the reference returns s->e. */
int comd_nsteps(comd_ctx *ctx, int n, double dt_seconds, double *pe_out);

/* The loop of nTimeSteps without its trailing force call: n steps of
 * half drift + rebin, force, kick, half drift + rebin.  Nothing is copied
 * to the host and the stream is not synchronised; errors raised by the
 * kernels are reported by the next synchronising call (comd_sync,
 * comd_energy, comd_nsteps). */
int comd_steps(comd_ctx *ctx, int n, double dt_seconds);

/* The two halves of a step, for callers that drive the integrator
 * themselves: advancePositions (src-flat/timestep.c:76-108: r += (dt p)/m,
 * then reBoxAll) and advanceVelocity (timestep.c:62-74: p -= dt f), with dt
 * in the reference's internal unit (seconds * 0.529/2.418e-17). */
int comd_advance_positions(comd_ctx *ctx, double dt_internal);
int comd_advance_velocity(comd_ctx *ctx, double dt_internal);

/* Totals of the last force evaluation: potential energy (s->e), the
 * reference's scalar virial "stress" (src-flat/eam.c:649-657,
 * src-flat/ljForce.c:175-221), and the kinetic energy sum |p|^2/(2 m) in
 * hartree (which the reference never computes). */
int comd_energy(comd_ctx *ctx, double *pe, double *ke, double *stress);

int comd_set_precision(comd_ctx *ctx, int precision);
int comd_set_eval(comd_ctx *ctx, int eval);

/* Tuning of a context (no counterpart in the reference, whose force loops are
 * serial): launch shape of the link-cell force kernels (waves = home cells in
 * flight per block, persistent blocks per compute unit, tables in LDS, lane
 * balance), step replay, pair-list replay and the Verlet skin, none of which
 * changes results beyond the order of per-atom sums; and the Lennard-Jones
 * force convention (key 8), which does. */
enum comd_tune_key {
    COMD_TUNE_WAVES_PER_BLOCK = 1, /* 1..4, default 1 */
    COMD_TUNE_BLOCKS_PER_CU = 2,   /* 1..64, default 8 */
    COMD_TUNE_TABLES_IN_LDS = 3,   /* 0/1, default 1 */
    COMD_TUNE_BALANCED_DRAIN = 4,  /* 0/1: share an atom's accepted pairs among its lanes */
    COMD_TUNE_GRAPH = 5,           /* 0/1: comd_steps/comd_nsteps replay one captured step
                                      (default 1 for batches, 0 for single simulations) */
    COMD_TUNE_PAIR_LISTS = 6,      /* 0/1, default 1: the EAM embedding-force pass replays the pairs
                                      the density pass accepted instead of testing candidates again */
    COMD_TUNE_SKIN_MILLI = 7,      /* Verlet skin in 1e-3 A (default 400; 0 = link-cell kernels only).
                                      Periodic simulations, and batches whose replicas all are, with
                                      >= 3 cells of cutoff + skin per dimension (and one potential per
                                      kind in a batch) keep per-atom neighbour lists, rebuilt on the
                                      device when an atom has moved half the skin; same pairs and
                                      outputs, per-atom sums in another order */
    COMD_TUNE_LJ_PHYSICAL = 8,     /* 0/1, default 0: Lennard-Jones forces stored as +grad U, so the
                                      integrator moves atoms down the potential (physical motion,
                                      KE + PE conserved).  0 keeps src-flat's convention (ljForce.c
                                      stores -grad U and timestep.c:71 subtracts it: atoms climb the
                                      potential, KE - PE is conserved, and a 300 K copper lattice
                                      (epsilon 0.167 eV) flies apart within about 50 fs).  No counterpart in the
                                      reference. */
    COMD_TUNE_LIST_GROUPS = 9,     /* -1 (default), 0..4: auxiliary cells per group of the
                                      neighbour-list passes, 4x4x3, 4x4x2, 4x2x2, 2x2x2, 2x2x1 (from the
                                      option given, the first whose staged halo fits the LDS of a
                                      block; by default the largest fitting shape that still gives two
                                      groups per compute unit, else the one with the most groups) */
    COMD_TUNE_LIST_ROWS_LDS = 10,  /* -1 (default, = 1), 0, 1: table rows of the list passes staged in
                                      LDS when they fit beside the halo */
    COMD_TUNE_LIST_THREADS_DIV = 11, /* 1 (default) or 2..4: home atoms per thread of the list passes
                                      (threads per block = a group's atoms / this, in waves) */
    COMD_TUNE_ADJOINT_LISTS = 12,  /* the adjoint's pair passes over the neighbour lists when the
                                      context uses them: 1 (default) if it has a list group per CU at
                                      least, 2 always, 0 never (link cells) */
    COMD_TUNE_LIST_LANES = 13,     /* 0 (default): by the system's size; 1, 2 or 4: adjacent lanes per
                                      home atom in the list force passes, each taking every lanes-th
                                      chunk of the atom's list (more waves for small systems) */
};
int comd_tune(comd_ctx *ctx, int key, int value);

/* Pick the force kernels' launch shape by timing force calls, for the
 * kernels that run.  With neighbour lists (the default for periodic boxes
 * and batches): comd_autotune_lists, and waves, blocks_per_cu and balanced
 * return the link-cell settings unchanged.  Without: the link-cell kernels'
 * cells (waves) per workgroup, persistent workgroups per CU, and how an
 * atom's accepted pairs are split over its lanes (balanced or per lane).
 * The state does not change; the choice is applied and returned.  A new
 * shape changes the order of the per-atom sums, so forces may differ in
 * the last bits (within the parity bar). */
int comd_autotune(comd_ctx *ctx, int *waves, int *blocks_per_cu, int *balanced, double *best_ms);

/* The list passes' shape by timing: the group of auxiliary cells
 * (COMD_TUNE_LIST_GROUPS -1..4), the table rows in LDS or global memory
 * (COMD_TUNE_LIST_ROWS_LDS 0/-1, tried when table rows are used) and home
 * atoms per thread (COMD_TUNE_LIST_THREADS_DIV 1/2); the halo capacity and
 * threads per block follow from the group shape.  Each option rebuilds the
 * lists; ones whose lists overflow are skipped.  The fastest is applied (the
 * lists rebuilt and the forces of the state computed with it) and
 * returned; best_ms is the time of one force call with it.  Errors
 * COMD_ERR_STATE when the context has no lists. */
int comd_autotune_lists(comd_ctx *ctx, int *group_opt, int *rows_lds, int *threads_div, double *best_ms);

/* State of the Verlet neighbour lists (COMD_TUNE_SKIN_MILLI): whether the
 * context uses them, whether the last build fitted (otherwise the link-cell
 * kernels run), how many builds have run, and the list capacity per atom. */
int comd_list_info(comd_ctx *ctx, int *active, int *valid, int *builds, int *kmax);

/* Launch shape of the neighbour-list passes at the current precision: for
 * the density (or Lennard-Jones) pass and the embedding-force pass, the
 * blocks resident per compute unit (as the HIP occupancy calculator gives
 * it; -1 when it cannot) and the dynamic LDS bytes per block, and the
 * threads per block.  All zero while the lists are off.  A diagnostic for
 * tuning the group shape; no counterpart in the reference. */
int comd_list_occupancy(comd_ctx *ctx, int blocks_per_cu[2], int lds_bytes[2], int *threads);

/* Lanes per home atom of the neighbour-list passes (1, 2 or 4: as tuned with
 * COMD_TUNE_LIST_LANES, else the context's default for its size); 0 while
 * the lists are off.  A diagnostic beside comd_list_occupancy; no
 * counterpart in the reference. */
int comd_list_lanes(comd_ctx *ctx, int *lanes);

/* Number of atom pairs inside the cutoff found by the last EAM density or
 * LJ pass (each unordered pair once: the reference's half-pair count). */
int comd_pair_count(comd_ctx *ctx, long long *half_pairs);

/* Per-slot results beyond the simflat_t arrays: the number of neighbours
 * inside the cutoff of the last EAM density or LJ pass, and the per-atom
 * share of the virial "stress" sum (before division by the volume). */
int comd_download_aux(comd_ctx *ctx, int *nnb, double *vir);

/* Number of visible HIP devices and the name of the current one. */
int comd_device_info(int *count, char *name, int name_len);

/* Block until the context's stream is idle. */
int comd_sync(comd_ctx *ctx);

/* Human-readable text of the last error on this context (or of the last
 * failed comd_create when ctx is NULL). */
const char *comd_last_error(const comd_ctx *ctx);

/* Device-side timing of the kernels launched on the context's stream:
 * enables (on != 0) or disables per-kernel HIP-event timing, and reads the
 * accumulated milliseconds and launch counts per kernel class
 * (COMD_KCLASS_*). */
enum comd_kclass {
    COMD_K_REBIN = 0,   /* drift/kick + rebinning kernels */
    COMD_K_DENSITY = 1, /* EAM pass 1 (+ embedding) or LJ */
    COMD_K_EMBED_FORCE = 2, /* EAM pass 3 */
    COMD_K_REDUCE = 3,
    COMD_K_LIST_BUILD = 4, /* neighbour-list rebuild check (+ rebuild) */
    COMD_K_CELL_FALLBACK = 5, /* link-cell passes launched beside the list
                                 passes; they return at once while the lists
                                 are valid */
    COMD_K_NCLASS = 6
};
int comd_timing(comd_ctx *ctx, int on);
int comd_timing_read(comd_ctx *ctx, double *ms /* COMD_K_NCLASS */, int *launches /* COMD_K_NCLASS */);

/* ---- Observables beyond the reference's ----
 *
 * The reference reports only the potential energy and a partial scalar
 * "stress" (the xx component of the EAM embedding virial,
 * src-flat/eam.c:649-657; for LJ the pair xx term with a kinetic term
 * counted once per non-empty neighbour cell, ljForce.c:175-221).  These
 * add the kinetic energy and temperature and the full virial tensor
 *   W_ab = sum_i p_ia p_ib / m + sum_{i<j} r_ij,a f_ij,b
 * (f_ij the physical pair force on i from j, -grad_i U, whatever the sign
 * convention of the force array), from which the pressure tr(W) / (3V) and
 * the stress tensor -W / V follow.  No counterpart in the reference. */
typedef struct comd_observables {
    double pe;              /* potential energy, hartree (the reference's s->e) */
    double ke;              /* sum |p|^2 / (2 m), hartree */
    double etot;            /* pe + ke */
    double temperature;     /* 2 ke / (3 N k_B), kelvin, k_B = 3.166811563e-6 hartree/K */
    double virial[9];       /* W, hartree, row-major */
    double virial_pair[9];  /* its pair and embedding part */
    double pressure;        /* tr(W) / (3 V), hartree / A^3 (x 4359.74 for GPa) */
    double stress_ref;      /* the reference's scalar "stress" of the same force call */
    double volume;          /* A^3 */
    int natoms;
} comd_observables;

/* Runs one force evaluation at the current positions (so that F' and the
 * energies are current), then the observables of the simulation (of
 * replica r for comd_batch_observe).  The pair sums are FP64 over the
 * link-cell stencil with the force passes' cutoff predicates; F' is that
 * of the force call, in the context's precision.  Needs at least 3 link
 * cells per dimension (COMD_ERR_GRID otherwise). */
int comd_observe(comd_ctx *ctx, comd_observables *out);
int comd_batch_observe(comd_ctx *ctx, int r, comd_observables *out);

/* comd_observe that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The same force call
 * and reduction as comd_observe, stream-ordered on the context's stream
 * with no host synchronisation; the result is written to d_out, a
 * comd_observables in the caller's device memory.  The reduction's block
 * partials live in a static device array of 256 KiB (2048 blocks x 16
 * doubles; every simulation size fits, the reduction's grid being capped
 * at 2048 blocks), one per device, kept until the process ends: calls that
 * overlap on one device (contexts on different streams) race on it.
 * Errors of the force kernels surface at the next synchronising call
 * (comd_sync).  Single simulations (replica 0 of a batch). */
int comd_observe_noalloc(comd_ctx *ctx, comd_observables *d_out);

/* Energy-conservation monitor: the potential and kinetic energy (hartree)
 * after a force at the starting positions, then after each of nblocks
 * calls of nTimeSteps(block_steps, dt) -- src-flat/docompute.c:196-207's
 * loop, whose energies the reference prints -- recorded on the device
 * without synchronising with the host between blocks.  pe_ke receives
 * 2 (nblocks + 1) doubles (pe, ke) per record.  Single simulations only. */
int comd_run_monitored(comd_ctx *ctx, int nblocks, int block_steps, double dt_seconds, double *pe_ke);

/* Centrosymmetry parameter per atom, the defect measure of src-viz's
 * viewer (computeCentrosymmetry, src-viz/AtomVisualize.c:399-673): the
 * smallest, over the ways of pairing an atom's 12 nearest neighbours, of
 * the sum of |r_j + r_k|^2 (0 in a perfect FCC crystal, A^2), 1000 for
 * atoms with fewer than 12 neighbours in reach; the reference's neighbour
 * search and float sums.  out: one double per slot of the replica's cells
 * (slot layout, as comd_download), empty slots 0.  The reference's link
 * cells are at least the cutoff wide, which must exceed the 12th-neighbour
 * distance. */
int comd_centrosymmetry(comd_ctx *ctx, double *out);
int comd_batch_centrosymmetry(comd_ctx *ctx, int r, double *out);

/* ---- Snapshots ----
 *
 * The whole device state of a single simulation -- slot arrays (natoms,
 * id, iType, r, p, the last forces, rho, F', virial terms) and the
 * neighbour lists with their build state -- written to a file and read
 * back into a context created with the same parameters, so that the run
 * continues bit for bit as if it had not stopped.  Load applies the
 * snapshot's precision, evaluation, skin and LJ sign to the context.  The
 * reference has no restart (its clsman and gz writers are commented out of
 * its driver, src-flat/pmd.c:117,123, and the gz text is %.8lf). */
int comd_snapshot_save(comd_ctx *ctx, const char *path);
int comd_snapshot_load(comd_ctx *ctx, const char *path);

/* ---- Batches: many independent replicas in one context ----
 *
 * No counterpart in the reference (a src-flat process holds one simulation;
 * its globals make it non-reentrant).  Each replica has its own parameters
 * (potential, grid from its own bounds, periodicity) and its own slot layout
 * with the reference's semantics; their cells are numbered one after the
 * other, so every kernel of a step (drift, rebin, force passes) covers all
 * replicas in one launch, and comd_steps replays one captured step graph.
 * All replicas share the precision.  The single-simulation calls above act
 * on replica 0 (comd_upload, comd_download, comd_bin_atoms*) or on all
 * replicas (comd_force, comd_steps, comd_nsteps, comd_rebox; comd_energy
 * returns the sums over replicas). */
int comd_batch_create(int n, const comd_params *prms /* n */, comd_ctx **out);
/* Cells of replica r start at global cell *cell_base; *natoms atoms. */
int comd_batch_info(comd_ctx *ctx, int r, int *cell_base, int *ncells, int *natoms);
int comd_batch_bin_atoms(comd_ctx *ctx, int r, int n, const double *xyz, const double *pxyz, int itype);
int comd_batch_upload(comd_ctx *ctx, int r, const int *natoms, const int *id, const int *itype, const double *rr,
                      const double *p);
/* As comd_download, for replica r's own cells and slots. */
int comd_batch_download(comd_ctx *ctx, int r, int *natoms, int *id, int *itype, double *rr, double *p, double *f,
                        double *rho, double *fi);
int comd_batch_energy(comd_ctx *ctx, int r, double *pe, double *ke, double *stress);
/* n steps of every replica (comd_steps). */
int comd_batch_step(comd_ctx *ctx, int n, double dt_seconds);

/* ---- Chebyshev coefficients and their adjoint ---- */

/* Replace the value series of phi, rho and F (coef: 3*n doubles, in that
 * order, on the domains the context was created with) and rebuild the three
 * derivative series with the recurrence of src-flat/cheby.c:219-270 -- the
 * runtime form of setChebPot (src-flat/cheby.c:99-121).  n must equal the
 * context's cheb_n.  In a batch, every replica (all EAM with series) takes
 * the same coefficients. */
int comd_set_cheb(comd_ctx *ctx, const double *coef, int n);

/* The six series in use (phi rho F dphi drho dF, 6*cheb_n doubles; of
 * replica 0 in a batch). */
int comd_get_cheb(comd_ctx *ctx, double *coef6);

/* Force-matching loss and its gradient with respect to the 3*n value
 * coefficients (n <= 32).  Sets the coefficients (as comd_set_cheb), runs
 * the EAM force through the series in the context's precision, and returns
 *   loss = sum_i |f_i - f_target_i|^2 / (3 N) + w_e (E - e_target)^2
 * (f the force array of the reference's convention, +grad U, eam.c:529-533,
 * E the potential energy) and grad[3n] = dloss/dcoef from the adjoint
 * kernels.  f_target holds 3 doubles per slot in the context's current slot
 * layout (as comd_download returns f without its 4th component); empty
 * slots are ignored.  f_target NULL reuses the target of the previous call
 * on this context (it stays in device memory).  No counterpart in the
 * reference: it replaces fitting by finite differences over the force call. */
int comd_eam_cheb_grad(comd_ctx *ctx, const double *coef, int n, const double *f_target, double e_target,
                       double w_e, double *loss, double *grad);

/* The same over every replica of a batch of EAM replicas with the same
 * series domains and cutoff, in one launch per adjoint stage:
 *   loss = sum_r [ sum_{i in r} |f_i - f_target_i|^2 / (3 N_r)
 *                  + w_e (E_r - e_target[r])^2 ],
 * e_target[nrep]; f_target (or NULL, as above) covers the batch's slots.
 * The coefficients are set in every replica.  No counterpart in the
 * reference. */
int comd_batch_cheb_grad(comd_ctx *ctx, const double *coef, int n, const double *f_target, const double *e_target,
                         double w_e, double *loss, double *grad);

/* Gradient of the potential energy alone with respect to the value
 * coefficients of phi, rho and F (grad: 3 x n, as comd_eam_cheb_grad's),
 * at coef, through the same adjoint kernels with no force term: dE/dc_phi
 * = 1/2 sum_ij T_k(r_ij), dE/dc_rho = sum_ij F'(rho_i) T_k(r_ij), dE/dc_F =
 * sum_i T_k(rho_i) (T_0 halved as the series convention has it).  energy
 * receives E.  No counterpart in the reference (it serves potential fitting
 * on energies, beside the force-matching loss).  The batch form returns the
 * gradient of sum_r w_r E_r (weights NULL: all 1) and each E_r in
 * energies[nrep] (may be NULL). */
int comd_eam_cheb_energy_grad(comd_ctx *ctx, const double *coef, int n, double *energy, double *grad);
int comd_batch_cheb_energy_grad(comd_ctx *ctx, const double *coef, int n, const double *weights, double *energies,
                                double *grad);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif
#endif
