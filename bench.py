"""Benchmark of the link-cell EAM timestep on one MI355X.

Workload: EAM Cu (cu500 tables, 3-point table interpolation), FCC 80^3 =
2,048,000 atoms, a = 3.615 A, periodic, Maxwell-Boltzmann velocities at
300 K from a fixed seed, dt = 1 fs.  One step = half drift + rebin, force
(neighbour-list check and rebuild when an atom has moved half the 0.4 A
skin, density + embedding pass, embedding-force pass), kick, half drift +
rebin, all on the GPU with the state resident in HBM.  The headline value
is the mixed-precision path (FP32 table coefficients, fixed-point list
positions, FP64 sums and integration); the FP64 path is timed in the same
run.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs
on one GPU.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import comd_amd as CA  # noqa: E402

# MI355X published peaks.  The vector FP32 figure (157.3) counts packed
# FP32 instructions, which this library does not use; unpacked FP32 and
# FP64 vector instructions both peak at 78.6 TFLOP/s, the ceiling that
# applies to every kernel here.
PEAK_FP64_TFLOPS = 78.6
PEAK_VALU_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Algorithmic work per atom-step (EAM table; +,-,*,/,sqrt = 1 FLOP each,
# rejected candidate pairs not counted; P_half = half-pairs per atom):
#   density pass: 66 FLOP per half-pair + 20 per atom (embedding)
#   embedding-force pass: 49 FLOP per half-pair
#   drift/kick: 24 FLOP per atom
# Bytes per atom-step (FP64 state, compulsory): 232.
FLOP_PAIR_DENSITY = 66
FLOP_PAIR_EMBED_FORCE = 49
FLOP_ATOM_EMBED = 20
FLOP_ATOM_VERLET = 24
BYTES_ATOM_STEP = 232
# HIP-event timing classes of comd_timing_read
KCLASS = ("move+rebin", "density", "embed_force", "reduce", "list_build", "cell_fallback")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 1000 for eam, 500 for batch, 60 for adjoint)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=80, help="FCC unit cells per edge")
    ap.add_argument("--seed", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference timing")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-n", type=int, default=40, help="FCC cells per edge of the CPU reference's sample")
    ap.add_argument("--no-extra", action="store_true", help="skip the other configurations' timings")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default list-pass shapes (for profiles whose kernels should not vary)")
    ap.add_argument("--workload", choices=("eam", "batch", "adjoint"), default="eam",
                    help="eam: EAM Cu 2.05M atoms, FP64 and mixed (the headline line); batch: 64 LJ/EAM replicas "
                         "in one context; adjoint: Chebyshev-coefficient gradients of 16 Cu 20^3 configurations")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"eam": 1000, "batch": 500, "adjoint": 60}[args.workload]
    return args


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_precision(sim, precision, steps, warmup, n_atoms, timed_steps, autotune=True):
    """steps timed as the library runs them (the step replayed as a HIP
    graph), the energy drift over those steps, then timed_steps more with
    HIP events around every kernel class on the simulation's stream (the
    events break the graph into separate launches, so these steps are not
    the throughput figure)."""
    sim.set_precision(precision)
    tuned = sim.autotune_lists() if autotune and sim.list_info()["active"] else None  # setup, not timed
    log(f"{precision}: list passes autotuned: {tuned}")
    sim.steps(warmup)
    sim.sync()
    sim.force()
    pe0, ke0, _ = sim.energy()
    p0 = sim.pair_count() / n_atoms
    b0 = sim.list_info()["builds"]
    sim.sync()
    t0 = time.perf_counter()
    sim.steps(steps)
    sim.sync()
    t1 = time.perf_counter()
    info = sim.list_info()
    sim.force()
    pe1, ke1, _ = sim.energy()
    p1 = sim.pair_count() / n_atoms
    sim.timing(True)
    sim.sync()
    t2 = time.perf_counter()
    sim.steps(timed_steps)
    sim.sync()
    t3 = time.perf_counter()
    ms, launches = sim.timing_read()
    sim.timing(False)
    wall = t1 - t0
    return dict(
        wall=wall,
        atom_steps_per_s=n_atoms * steps / wall,
        ms_per_step=1e3 * wall / steps,
        ms_per_step_events=1e3 * (t3 - t2) / timed_steps,
        timed_steps=timed_steps,
        kernel_ms=ms,
        launches=launches,
        p_half=0.5 * (p0 + p1),
        e0=pe0 + ke0,
        e1=pe1 + ke1,
        lists=bool(info["active"] and info["valid"]),
        builds=info["builds"] - b0,
        kmax=info["kmax"],
        lanes=sim.list_lanes(),
        autotune=tuned,
    )


def mixed_force_error(sim):
    """Per-atom error of the mixed-precision force, rho and energy against
    FP64 on the simulation's current state: |df| / max(|f|, rms |f|)."""
    fe = {}
    for prec in ("fp64", "mixed"):
        sim.set_precision(prec)
        sim.force()
        st = sim.download()
        m = (np.arange(CA.MAXATOMS)[None, :] < st["natoms"][:, None]).reshape(-1)
        fe[prec] = (st["f"][m], st["rho"][m])
        del st
    f64, f32 = fe["fp64"][0], fe["mixed"][0]
    d = np.linalg.norm(f32[:, :3] - f64[:, :3], axis=1)
    mag = np.linalg.norm(f64[:, :3], axis=1)
    rel = d / np.maximum(mag, np.sqrt(np.mean(mag ** 2)))
    erel = np.abs(f32[:, 3] - f64[:, 3]) / np.abs(f64[:, 3])
    rrel = np.abs(fe["mixed"][1] - fe["fp64"][1]) / np.abs(fe["fp64"][1])
    sim.set_precision("fp64")
    return dict(force_max=float(rel.max()), force_median=float(np.median(rel)),
                energy_per_atom_max=float(erel.max()), rho_max=float(rrel.max()))


def measured_traffic(kname):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summaries (FETCH_SIZE and WRITE_SIZE runs, scaled by a copy kernel of
    known size); (None, reason) when none holds a line for the kernel."""
    pdir = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not name.endswith("hbm_traffic.txt"):
            continue
        path = os.path.join(pdir, name)
        for line in open(path):
            m = re.match(r"traffic_bytes_per_launch (.+?) = ([0-9]+)", line.strip())
            if m and m.group(1) == kname:
                return float(m.group(2)), os.path.relpath(path, ROOT)
    return None, "no counter summary for this kernel"


def host_cpu():
    """Model name of the host CPU and the number of CPUs this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, ncpu


def cpu_baseline(n, steps, seed):
    """The compiled reference (oracle/_ref/ref_table, built from the
    reference's own sources), one thread pinned to one core, on a bounded
    sample of the benchmark's workload: the same seeded 300 K Cu lattice at
    FCC n^3, read from a gz input file written at %.17g; the reference's
    extra end-of-call force is excluded.  Falls back to the oracle
    restatement when the compiled reference is absent."""
    import tempfile

    model, ncpu = host_cpu()
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_table")
    natoms = 4 * n ** 3
    sample = (f"EAM cu500 table, the seeded 300 K FCC {n}^3 lattice ({natoms} atoms) of the benchmark at a "
              f"smaller size, {steps} steps after one force, extra end-of-call force excluded, 1 thread")
    xyz = CA.fcc_positions(n, n, n, 3.615)
    xyz, p = CA.seed_state(xyz, 63.55, 300.0, 0.0, seed)
    bounds = (n * 3.615,) * 3
    extra = dict(cpu_model=model, nproc=ncpu)
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as tmp:
            inp = os.path.join(tmp, "seeded.inp.gz")
            CA.write_gz(inp, bounds, xyz, p)
            cmd = [exe, "--eam", CA.POT_DIR, "cu500", "--input", inp, "--steps", str(steps), "--dumps", "1", "--time"]
            cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else [0]
            pin = (lambda c: ["taskset", "-c", str(c)]) if shutil.which("taskset") else (lambda c: [])
            out = subprocess.run(pin(cpus[0]) + cmd, capture_output=True, text=True, check=True).stdout
            rate = lambda txt: float(re.search(r"atom_steps_per_s=([0-9.eE+-]+)", txt).group(1))
            one = rate(out)
            # the same sample in one independent process per available CPU
            # at once, at most 16 (the reference's path is single-threaded):
            # the host's aggregate throughput
            cpus = cpus[:16]
            procs = [subprocess.Popen(pin(c) + cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
                     for c in cpus]
            outs = [pr.communicate()[0] for pr in procs]
            if any(pr.returncode != 0 for pr in procs):
                raise RuntimeError("a reference process failed")
            extra["aggregate"] = dict(value=sum(rate(o) for o in outs), unit="atom-steps/s", processes=len(cpus),
                                      note="one independent reference process per available CPU (at most 16), the same sample")
        return dict(value=one, unit="atom-steps/s", cores=1, kind="reference", sample=sample, **extra)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    pot = O.OraclePot.eam(CA.POT_DIR, "cu500")
    s = O.OracleSim.from_atoms(pot, bounds, xyz, p)
    s.rebox()
    t0 = time.perf_counter()
    s.force()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.nsteps(steps)
    t = time.perf_counter() - t0 - tf
    return dict(value=natoms * steps / t, unit="atom-steps/s", cores=1, kind="port", sample=sample, **extra)


def timed_steps(sim, steps, warmup, natoms):
    """atom-steps/s of steps replayed as the library runs them, after warmup."""
    sim.steps(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.steps(steps)
    sim.sync()
    return natoms * steps / (time.perf_counter() - t0)


def other_configs(seed):
    """The other benchmark configurations, timed in the same run: plain
    Lennard-Jones copper (FCC 20^3, 600 K, 100 steps, FP64, physical force
    convention: in src-flat's the lattice flies apart within about 50
    steps), EAM Cu FCC 40^3 (256 000 atoms, 0.05 A displacements, 300 K, 100
    steps) by table and by 32-term Chebyshev series in FP64 and mixed
    precision, the 64-replica batch and the Chebyshev-coefficient gradient."""
    out = {}
    lj = CA.copper_lj()
    xyz = CA.fcc_positions(20, 20, 20, lj.lat)
    xyz, p = CA.seed_state(xyz, lj.mass, 600.0, 0.0, seed)
    s = CA.Simulation(lj, (20 * lj.lat,) * 3, defgrad=0.0)
    s.tune(lj_physical=1)
    s.bin_atoms(xyz, p)
    s.rebox()
    s.force()
    out["lj_cu_fcc20_600K_fp64"] = timed_steps(s, 100, 5, len(xyz))
    s.close()
    cu = CA.EAMPotential("cu500")
    xyz = CA.fcc_positions(40, 40, 40, 3.615)
    xyz, p = CA.seed_state(xyz, cu.mass, 300.0, 0.05, seed)
    for ev in ("table", "cheby"):
        s = CA.Simulation(cu, (40 * 3.615,) * 3, defgrad=0.0, evaluation=ev)
        s.bin_atoms(xyz, p)
        s.rebox()
        s.force()
        for prec in ("fp64", "mixed"):
            s.set_precision(prec)
            s.force()
            out[f"eam_cu_fcc40_{ev}_{prec}"] = timed_steps(s, 100, 5, len(xyz))
        s.close()
    log(f"other configs: {out}")
    return {k: {"value": v, "unit": "atom-steps/s"} for k, v in out.items()}

def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path; the sharding plan is a design "
              "note, not code", file=sys.stderr)
        sys.exit(2)
    ndev, devname = CA.device_info()
    if ndev < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batch":
        return bench_batch(args, devname)
    if args.workload == "adjoint":
        return bench_adjoint(args, devname)
    n = args.n
    pot = CA.EAMPotential("cu500")
    t_setup = time.perf_counter()
    sim, xyz, p = CA.fcc_simulation(pot, n, temp=300.0, disp=0.0, seed=args.seed)
    n_atoms = xyz.shape[0]
    del xyz, p
    sim.rebox()
    sim.force()
    t_setup = time.perf_counter() - t_setup

    # both precisions step from the same seeded state
    s0 = sim.download()
    res = {}
    force_err = None
    for prec in ("fp64", "mixed"):
        if prec != "fp64":
            sim.upload(s0["natoms"], s0["id"], s0["itype"], s0["r"], s0["p"])
            sim.rebox()
        res[prec] = run_precision(sim, prec, args.steps, args.warmup, n_atoms, min(args.steps, 200),
                                  autotune=not args.no_autotune)
        log(f"{prec}: {res[prec]['atom_steps_per_s'] / 1e9:.3f} G atom-steps/s, "
            f"{res[prec]['ms_per_step']:.3f} ms/step, kernel ms {['%.2f' % v for v in res[prec]['kernel_ms']]}")
        if prec == "fp64":
            force_err = mixed_force_error(sim)
            log(f"mixed vs fp64 at the thermalised end state of the FP64 run: {force_err}")
    del s0
    sim.close()

    mixed = res["mixed"]
    fp64 = res["fp64"]
    # dominant kernel of the mixed path (the force pass with the larger
    # time), timed with HIP events on the stream it runs on: the list pass
    # while the neighbour lists are in use, else the link-cell pass
    lists = mixed["lists"]
    rows = int(mixed["autotune"] is None or mixed["autotune"]["rows_lds"] != 0)  # table rows in LDS
    multi = int(mixed["lanes"] > 1)  # the passes' variant with several lanes per home atom
    kd = f"k_nl_pair<0, 0, 1, {rows}, {multi}>" if lists else "k_pair<0, 0, 1>"
    ke = f"k_nl_pair<1, 0, 1, {rows}, {multi}>" if lists else "k_pair<1, 0, 1>"
    passes = {
        CA.K_DENSITY: (kd, "density + embedding pass (src-flat/eam.c:484-594), table, mixed",
                       lambda ph: (FLOP_PAIR_DENSITY * ph + FLOP_ATOM_EMBED) * n_atoms),
        CA.K_EMBED_FORCE: (ke, "embedding-force pass (src-flat/eam.c:595-653), table, mixed",
                           lambda ph: FLOP_PAIR_EMBED_FORCE * ph * n_atoms),
    }
    k = max(passes, key=lambda c: mixed["kernel_ms"][c])
    kname, kdesc, kflop = passes[k]
    avg_s = mixed["kernel_ms"][k] / max(mixed["launches"][k], 1) * 1e-3
    flop_launch = kflop(mixed["p_half"])
    achieved = flop_launch / avg_s / 1e12
    traffic, traffic_src = measured_traffic(kname)
    k64 = fp64["kernel_ms"][k] / max(fp64["launches"][k], 1) * 1e-3
    flop64 = kflop(fp64["p_half"])

    f_alg = (FLOP_PAIR_DENSITY + FLOP_PAIR_EMBED_FORCE) * mixed["p_half"] + FLOP_ATOM_EMBED + FLOP_ATOM_VERLET
    ceil_mixed = min(PEAK_VALU_TFLOPS * 1e12 / f_alg, PEAK_HBM_GBS * 1e9 / BYTES_ATOM_STEP)
    ceil_fp64 = min(PEAK_FP64_TFLOPS * 1e12 / f_alg, PEAK_HBM_GBS * 1e9 / BYTES_ATOM_STEP)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_n, args.cpu_steps, args.seed)
        except Exception as exc:  # reported, not fatal
            print(f"[bench] cpu baseline failed: {exc}", file=sys.stderr)
    extra = None
    if not args.no_extra:
        try:
            extra = other_configs(args.seed)
            extra["batch_64_replicas_mixed"] = batch_throughput(args, 200)
            extra["adjoint_16x32000_cheb96"] = adjoint_throughput(args, 3)
        except Exception as exc:  # reported, not fatal
            print(f"[bench] other configurations failed: {exc}", file=sys.stderr)

    out = {
        "metric": "atom-steps/s, EAM Cu 2.05M atoms: mixed-prec vs FP64 on 1 GPU; % roofline",
        "value": mixed["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 potential values; f64 derivatives, accumulation and integration; 32-bit fixed-point staged positions",
        "data": "synthetic: FCC lattice, seeded Maxwell-Boltzmann velocities at 300 K",
        "config": {
            "workload": f"EAM Cu cu500 table, FCC {n}^3 = {n_atoms} atoms, a=3.615 A, 300 K, dt=1 fs, "
                        f"link cells {'x'.join(map(str, sim.nbx))}",
            "atoms": n_atoms,
            "precision": "mixed",
            "parallelism": "single GPU",
        },
        "fp64": {
            "value": fp64["atom_steps_per_s"],
            "ms_per_step": fp64["ms_per_step"],
            "dominant_kernel_ms": k64 * 1e3,
            "dominant_kernel_tflops": flop64 / k64 / 1e12,
            "dominant_kernel_frac_of_fp64_peak": flop64 / k64 / 1e12 / PEAK_FP64_TFLOPS,
            "kernel_ms_per_step": {name: fp64["kernel_ms"][i] / fp64["timed_steps"] for i, name in enumerate(KCLASS)},
            "ms_per_step_with_events": fp64["ms_per_step_events"],
            "step_roofline_frac": fp64["atom_steps_per_s"] / ceil_fp64,
            "energy_drift_per_atom_ha": (fp64["e1"] - fp64["e0"]) / n_atoms,
        },
        "mixed_over_fp64": mixed["atom_steps_per_s"] / fp64["atom_steps_per_s"],
        "roofline": {
            "bound": "valu",
            "kernel": f"{kname}: {kdesc}",
            "achieved": achieved,
            "peak": PEAK_VALU_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_VALU_TFLOPS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "hbm_frac": (traffic / avg_s / 1e9 / PEAK_HBM_GBS) if traffic else None,
            "avg_launch_ms": avg_s * 1e3,
            "flop_per_launch": flop_launch,
            "p_half": mixed["p_half"],
        },
        "step_roofline": {
            "f_alg_flop_per_atom_step": f_alg,
            "b_alg_bytes_per_atom_step": BYTES_ATOM_STEP,
            "ceiling_atom_steps_per_s": ceil_mixed,
            "frac": mixed["atom_steps_per_s"] / ceil_mixed,
        },
        "kernel_ms_per_step": {name: mixed["kernel_ms"][i] / mixed["timed_steps"] for i, name in enumerate(KCLASS)},
        "kernel_timing": {"steps": mixed["timed_steps"], "ms_per_step_with_events": mixed["ms_per_step_events"],
                          "note": "HIP events around each kernel class on the simulation's stream, over steps run "
                                  "after the timed ones (the events split the step graph into separate launches)"},
        "neighbour_lists": {"active": mixed["lists"], "skin_A": 0.4, "builds_in_timed_steps": mixed["builds"],
                            "list_capacity": mixed["kmax"], "lanes_per_atom": mixed["lanes"],
                            "autotuned_shape": {"mixed": mixed["autotune"], "fp64": fp64["autotune"]}},
        "energy_drift_per_atom_ha": (mixed["e1"] - mixed["e0"]) / n_atoms,
        "mixed_error_vs_fp64": dict(force_err, note="per-atom |df| / max(|f|, rms |f|) on the thermalised state "
                                    "at the end of the FP64 run; both paths run the timed steps from the same "
                                    "seeded state, their energy drifts are energy_drift_per_atom_ha"),
        "device": devname,
        "setup_s": t_setup,
        "other_configs": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_batch(args, steps):
    """64 independent replicas, FCC n^3 with n uniform in 8..24, x-stretch
    uniform in [1.00, 1.10] (file-input grid rule), even index Lennard-Jones
    copper, odd index EAM Cu (cu500 tables), 300 K, mixed precision, in one
    context: one launch per kernel phase, steps replayed from a HIP graph."""
    rng = np.random.default_rng(args.seed)
    lj = CA.copper_lj()
    cu = CA.EAMPotential("cu500")
    reps, atoms = [], []
    for r in range(64):
        n = int(rng.integers(8, 25))
        stretch = float(rng.uniform(1.0, 1.10))
        pot, lat = (lj, lj.lat) if r % 2 == 0 else (cu, 3.615)
        xyz = CA.fcc_positions(n, n, n, lat, stretch)
        xyz, p = CA.seed_state(xyz, pot.mass, 300.0, 0.0, args.seed * 1000 + r)
        reps.append(dict(pot=pot, bounds=(n * lat * stretch, n * lat, n * lat), defgrad=0.0))
        atoms.append((xyz, p, n, stretch, "lj" if r % 2 == 0 else "eam"))
    log(f"batch: {len(reps)} replicas, {sum(a[0].shape[0] for a in atoms)} atoms generated")
    t0 = time.perf_counter()
    b = CA.Batch(reps, precision="mixed")
    # Lennard-Jones replicas move down their potential: in src-flat's sign
    # convention a 300 K lattice flies apart within about 50 steps
    b.tune(lj_physical=1)
    for r, (xyz, p, *_rest) in enumerate(atoms):
        b.bin_atoms(r, xyz, p)
    b.rebox()
    b.force()
    b.sync()
    t_setup = time.perf_counter() - t0
    log(f"batch: binned and first force in {t_setup:.2f} s")
    sizes = np.array([a[0].shape[0] for a in atoms])
    total = int(sizes.sum())
    tuned = None
    if not getattr(args, "no_autotune", False):
        tuned = b.autotune_lists()  # the list passes' group shape, rows in LDS, atoms per thread
        tuned["occupancy"] = b.list_occupancy()
    log(f"batch: autotune {tuned}")
    b.steps(args.warmup)
    b.sync()
    pe0, ke0, _ = b.energy()
    t0 = time.perf_counter()
    b.steps(steps)
    b.sync()
    wall = time.perf_counter() - t0
    log(f"batch: {steps} graph steps in {wall:.3f} s")
    b.force()
    pe1, ke1, _ = b.energy()
    # the same steps launched one by one, and the kernel split, for reference
    b.tune(graph=0)
    ke = min(50, steps)
    b.timing(True)
    t1 = time.perf_counter()
    b.steps(ke)
    b.sync()
    wall_eager = time.perf_counter() - t1
    ms, launches = b.timing_read()
    b.timing(False)
    per_rep = sizes * steps / wall
    res = dict(value=total * steps / wall, unit="atom-steps/s", steps=steps, ms_per_step=1e3 * wall / steps,
               atoms=total, cells=b.ncells, atoms_list=atoms,
               per_replica_atom_steps_per_s={"min": float(per_rep.min()), "median": float(np.median(per_rep)),
                                             "max": float(per_rep.max())},
               eager_ms_per_step=1e3 * wall_eager / ke, graph_speedup=(wall_eager / ke) / (wall / steps),
               kernel_ms_per_step_eager={name: ms[i] / ke for i, name in enumerate(KCLASS)},
               autotune=tuned,
               energy_drift_per_atom_ha=((pe1 + ke1) - (pe0 + ke0)) / total, setup_s=t_setup)
    b.close()
    return res


def batch_throughput(args, steps):
    res = run_batch(args, steps)
    return {k: v for k, v in res.items() if k in ("value", "unit", "steps", "ms_per_step", "atoms", "graph_speedup",
                                                  "energy_drift_per_atom_ha")}


def bench_batch(args, devname):
    res = run_batch(args, args.steps)
    atoms = res.pop("atoms_list")
    out = {
        "metric": "atom-steps/s, 64-replica batch (FCC 8-24^3, x-stretch 1.00-1.10, half LJ half EAM Cu), "
                  "mixed precision, 1 GPU",
        "value": res["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 potential evaluation, f64 geometry, accumulation and integration",
        "data": "synthetic: FCC lattices, seeded Maxwell-Boltzmann velocities at 300 K",
        "config": {
            "workload": f"64 replicas, {res['atoms']} atoms, {res['cells']} link cells, one launch per kernel phase, "
                        f"HIP-graph step; Lennard-Jones replicas in the physical force convention",
            "replicas": [dict(kind=a[4], n=a[2], atoms=int(a[0].shape[0]), stretch=round(a[3], 4))
                         for a in atoms],
            "parallelism": "single GPU",
        },
        "per_replica_atom_steps_per_s": res["per_replica_atom_steps_per_s"],
        "eager_ms_per_step": res["eager_ms_per_step"],
        "graph_speedup": res["graph_speedup"],
        "kernel_ms_per_step_eager": res["kernel_ms_per_step_eager"],
        "autotune": res["autotune"],
        "energy_drift_per_atom_ha": res["energy_drift_per_atom_ha"],
        "device": devname,
        "setup_s": res["setup_s"],
    }
    print(json.dumps(out))


def run_adjoint(args, reps):
    """Gradient of the force-matching loss with respect to the 96 Chebyshev
    coefficients, over 16 configurations of Cu FCC 20^3 (32 000 atoms each,
    displacements sigma 0.1 A), targets from the FP64 forward at 1 %-perturbed
    coefficients; FP64 and mixed.  The 16 configurations are one batch
    (comd_batch_cheb_grad: one launch per adjoint stage for all of them)."""
    cu = CA.EAMPotential("cu500")
    c0 = cu.cheb[:3].copy()
    rng = np.random.default_rng(args.seed)
    ct = c0 * (1.0 + 0.01 * rng.standard_normal(c0.shape))
    n, nconf = 20, 16
    bounds = (n * 3.615,) * 3
    b = CA.Batch([dict(pot=cu, bounds=bounds, defgrad=0.0, evaluation="cheby") for _ in range(nconf)],
                 precision="fp64")
    for k in range(nconf):
        xyz = CA.fcc_positions(n, n, n, 3.615)
        xyz, p = CA.seed_state(xyz, cu.mass, 0.0, 0.1, args.seed * 100 + k)
        b.bin_atoms(k, xyz, p)
    b.rebox()
    b.set_cheb(ct)
    b.force()
    fts = [b.download(k)["f"][:, :3].copy() for k in range(nconf)]
    e_t = [b.energy_of(k)[0] for k in range(nconf)]
    w_e = 1.0 / (4 * n ** 3) ** 2
    b.cheb_grad(c0, fts, e_t, w_e)  # uploads the targets once
    del fts
    natoms = nconf * 4 * n ** 3
    res = {}
    for prec in ("fp64", "mixed"):
        b.set_precision(prec)
        b.cheb_grad(c0, None, e_t, w_e)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            loss, grad = b.cheb_grad(c0, None, e_t, w_e)
        wall = (time.perf_counter() - t0) / reps
        b.set_cheb(c0)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.force()
        b.sync()
        fwd = (time.perf_counter() - t0) / reps
        res[prec] = dict(seconds_per_gradient=wall, atom_gradients_per_s=natoms / wall, loss=loss,
                         forward_seconds=fwd, gradient_over_forward=wall / fwd, grad=grad)
    b.close()
    return res, natoms


def adjoint_throughput(args, reps):
    res, natoms = run_adjoint(args, reps)
    return {"value": res["mixed"]["atom_gradients_per_s"], "unit": "atom-gradients/s",
            "seconds_per_gradient_mixed": res["mixed"]["seconds_per_gradient"],
            "seconds_per_gradient_fp64": res["fp64"]["seconds_per_gradient"], "atoms": natoms}


def bench_adjoint(args, devname):
    reps = max(3, args.steps // 20)
    res, natoms = run_adjoint(args, reps)
    rel = np.abs(res["mixed"]["grad"] - res["fp64"]["grad"]).max(axis=1) / np.abs(res["fp64"]["grad"]).max(axis=1)
    out = {
        "metric": "Chebyshev-coefficient gradients: atom-gradients/s over 16 x Cu FCC 20^3, 1 GPU",
        "value": res["mixed"]["atom_gradients_per_s"],
        "unit": "atom-gradients/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": 1e3 * res["mixed"]["seconds_per_gradient"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 basis and potential evaluation, f64 geometry and accumulation",
        "data": "synthetic: FCC 20^3 Cu with seeded displacements sigma 0.1 A, targets at 1 %-perturbed coefficients",
        "config": {"workload": "16 configurations x 32 000 atoms in one batch, 96 coefficients",
                   "parallelism": "single GPU"},
        "fp64": {k: v for k, v in res["fp64"].items() if k != "grad"},
        "mixed": {k: v for k, v in res["mixed"].items() if k != "grad"},
        "mixed_gradient_rel_error_per_function": rel.tolist(),
        "device": devname,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
