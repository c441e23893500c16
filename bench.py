"""Throughput of the ECSIM cycle on one MI355X: particle updates per second.

A step is one full ECSIM cycle of the uniform magnetised plasma, 128^3 cells,
box 32^3 d_i, 2 species (q/m -64 and +1) x 64 particles per cell (268M
particles), B0 = 0.01 z plus a seeded k-spectrum of 1e-3 B0, dt = 0.1,
theta = 0.5, GMRES(30) to 1e-10: solve, move, field advance, sort, deposit
and assemble. Inputs are generated on the device (resident in HBM) before the
timed region. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle updates/s on 1 MI355X, plain FP64 vs mixed-prec at fixed energy drift"
HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP64_MFMA_PEAK = 78.6e12   # flop/s, dense FP64 matrix
FP32_MFMA_PEAK = 157.3e12  # flop/s, dense FP32 matrix
BF16_MFMA_PEAK = 2.5e15    # flop/s, dense BF16 matrix
MFMA_PEAK = {0: FP64_MFMA_PEAK, 1: FP32_MFMA_PEAK, 2: BF16_MFMA_PEAK}
MASS_NAMES = {0: "fp64", 1: "fp32-mfma", 2: "bf16x3-mfma"}
# The cycle's dominant kernel: the sort's record gather fused with the J + M
# deposition (one launch per cycle covers every species).
KERNEL_NAME = {0: "k_gather_deposit_fp64", 1: "k_gather_deposit_mixed<MomentsF32>",
               2: "k_gather_deposit_mixed<MomentsBf16x3>"}
# Mass-phase work per particle as the survey's roofline defines it: the dense
# 8 x 8 pair form, 9 tensor entries x 64 pairs x 2 = 1152 flop (648 of them
# are the 36 canonical pairs' useful products); BF16 hi/lo counts its three
# split products.
DEPOSIT_FLOPS_PER_PARTICLE = 1152
USEFUL_FLOPS_PER_PARTICLE = 648     # 27 pair moments x (9 alpha + 3 current) x 2
# Issued MFMA work per particle of each form: FP64/FP32 16x16x4, two per 4
# particles (2 x 2048 / 4); BF16 16x16x32, six per 32 particles (6 x 16384 / 32).
ISSUED_FLOPS_PER_PARTICLE = {0: 1024, 1: 1024, 2: 3072}
SPLIT_PRODUCTS = {0: 1, 1: 1, 2: 3}
# HBM bytes per launch of the dominant kernel at 128^3, 2 x 64 ppc: a
# profiled constant (FETCH_SIZE x 2 KiB + WRITE_SIZE x 1 KiB per dispatch from
# two rocprofv3 --pmc runs, calibrated on a plain device copy), not measured
# by this run.
DEPOSIT_TRAFFIC_128 = {0: 25.399e9 + 24.674e9, 1: 27.816e9 + 22.451e9, 2: 27.803e9 + 22.474e9}
DEPOSIT_TRAFFIC_SOURCE = "profiles/k_gather_deposit_hbm_traffic_128cube.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"

# Algorithmic bytes of the dominant kernel: the survey's J + M deposit (read
# r, v, q per particle, 56 B; per node write M, 1008 B FP64 or 504 B FP32, and
# J, 24 B, and read B at nodes, 24 B) plus the sort's record move it now
# carries (read the source index, 4 B, the key and id, 12 B, of a record
# whose r, v, q the deposit reads anyway; write all nine attributes, 68 B).
SORT_COPY_BYTES_PER_PARTICLE = 4 + 12 + 68
DEPOSIT_ALG_BYTES_PER_PARTICLE = 56 + SORT_COPY_BYTES_PER_PARTICLE
DEPOSIT_ALG_BYTES_PER_NODE = {0: 1008 + 48, 1: 504 + 48, 2: 504 + 48}
MOVER_BYTES_PER_PARTICLE = 112      # read r,v,q (56) + key (4); write r,v (48) + key (4)
DEPOSIT_BYTES_PER_PARTICLE = 140    # this build's kernel: source index, record read, record written
# Cell records written per cell: 243 mass moments (27 x 9) in FP64 (plain
# form) or FP32 (mixed forms, whose per-cell sums are FP32), and 24 FP64
# corner currents.
DEPOSIT_BYTES_PER_CELL = {0: 243 * 8 + 192, 1: 243 * 4 + 192, 2: 243 * 4 + 192}
SORT_PLAN_BYTES_PER_PARTICLE = 8    # the crosser plan: one key pass + the source index written
SPMV_BYTES_PER_NODE = 1008 + 24 * 2 + 48  # matrix + in/out vectors + curl cells


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=4, help="particles per cell per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cells", type=int, default=32)
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision runs (FP32 and BF16 hi/lo moment products)")
    ap.add_argument("--drift-cycles", type=int, default=3, help="untimed cycles measuring the per-cycle drift")
    ap.add_argument("--smoothing", default=None, metavar="OP:PASSES:ALPHA",
                    help="energy-conserving smoothing of every leg, e.g. face7:6:0.5 or binomial27:2:0 "
                         "(default: none, as in every reference configuration)")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="only time a plain 1 GiB device-to-device copy (calibrates the HBM byte counters)")
    ap.add_argument("--spmv-cells", type=int, default=64,
                    help="grid of the reference operator-stencil CPU baseline")
    ap.add_argument("--spmv-applies", type=int, default=16)
    ap.add_argument("--cpu-leg", choices=("cycle", "spmv"), default=None, help=argparse.SUPPRESS)
    return ap.parse_args()


def hbm_probe(nbytes=1 << 30, reps=5):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch
    scale the counters of the path's kernels to bytes."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"probe": "device copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps,
                      "GBps": 2 * nbytes * reps / dt / 1e9}), flush=True)


def host_threads():
    """Threads the all-cores CPU figure uses: OMP_NUM_THREADS when set (the
    host's share of cores), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_cycle_leg(cells, ppc_axis):
    """One CPU figure, in this process: the oracle (a C restatement of the
    reference's cycle) on a bounded sample of the workload; the thread count
    is OMP_NUM_THREADS of the OpenMP build, or 1 for the serial build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from tpic_amd import cases

    n = (cells, cells, cells)
    length = 32.0 * cells / 128.0
    lengths = (length,) * 3
    ppc = (ppc_axis,) * 3
    qom = [-64.0, 1.0]
    thermal = [0.045, 0.0056]
    g = O.make_grid(n, lengths)
    arrays, first = [], 0
    for s in range(2):
        loaded = O.load_uniform(g, ppc, 1.0, (0.0, 0.0, 0.0), (thermal[s],) * 3, False, O.load_seed(s), first,
                                qom[s])
        arrays.append(loaded.a)
        first += loaded.count
    counts = [len(a["cell"]) for a in arrays]
    sim = O.Sim(n, lengths, qom, counts, dt=0.1, theta=0.5, tol=1e-10)
    for s, a in enumerate(arrays):
        sim.set_particles(s, a)
    sim.field("B")[:] = cases.uniform_b(n, (0.0, 0.0, 0.01), amplitude=1e-3)
    sim.initialize()
    sim.cycle(1)  # warm-up cycle, untimed
    cycles = 2
    t0 = time.perf_counter()
    rc, _ = sim.cycle(cycles)
    dt = time.perf_counter() - t0
    sim.close()
    return {"rc": rc, "particles": sum(counts), "cycles": cycles, "seconds": dt,
            "value": sum(counts) * cycles / dt}


def ref_spmv_leg(cells, applies):
    """The reference's own operator stencils, compiled from its headers into
    oracle/_ref/libtpic_ref.so (apply_mass_matrix_node over every node and
    node_cell_curl both ways, mass_matrix_stencil.hpp:120-149,
    node_cell_curl.hpp:52-141), on random one-ghost arrays of a cells^3 grid:
    node updates per second of one operator apply (M v + curl_C2N curl_N2C v)."""
    import ctypes

    import numpy as np

    path = os.path.join(ROOT, "oracle", "_ref", "libtpic_ref.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    D = ctypes.POINTER(ctypes.c_double)
    I3 = ctypes.POINTER(ctypes.c_int)
    L.ref_mass_apply.argtypes = [I3, D, D, D, D, ctypes.c_double, D]
    L.ref_curl_n2c.argtypes = [I3, D, D, D, D, D]
    L.ref_curl_c2n.argtypes = [I3, D, D, D, D, D]
    n = (ctypes.c_int * 3)(cells, cells, cells)
    spacing = (ctypes.c_double * 3)(0.25, 0.25, 0.25)
    rng = np.random.default_rng(5)
    verts = (cells + 3) ** 3
    nodes = (cells + 1) ** 3
    mass = rng.standard_normal(verts * 126)
    vg = [rng.standard_normal(verts) for _ in range(3)]           # nodes with one ghost
    cg = [rng.standard_normal((cells + 2) ** 3) for _ in range(3)]  # cells with one ghost
    cells_out = np.zeros(3 * cells ** 3)
    nodes_out = np.zeros(3 * nodes)
    mass_out = np.zeros(3 * nodes)
    ptr = lambda a: a.ctypes.data_as(D)  # noqa: E731
    L.ref_mass_apply(n, ptr(mass), ptr(vg[0]), ptr(vg[1]), ptr(vg[2]), 1.0, ptr(mass_out))  # warm
    t0 = time.perf_counter()
    for _ in range(applies):
        L.ref_curl_n2c(n, spacing, ptr(vg[0]), ptr(vg[1]), ptr(vg[2]), ptr(cells_out))
        L.ref_curl_c2n(n, spacing, ptr(cg[0]), ptr(cg[1]), ptr(cg[2]), ptr(nodes_out))
        L.ref_mass_apply(n, ptr(mass), ptr(vg[0]), ptr(vg[1]), ptr(vg[2]), 1.0, ptr(mass_out))
    dt = time.perf_counter() - t0
    return {"nodes": nodes, "applies": applies, "seconds": dt, "value": nodes * applies / dt}


def run_leg(leg, threads, **kw):
    """A CPU leg in a child process with its own thread count."""
    import subprocess

    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env["TPIC_ORACLE_OPENMP"] = "1" if threads > 1 else "0"
    args = [sys.executable, os.path.join(ROOT, "bench.py"), "--cpu-leg", leg]
    for k, v in kw.items():
        args += ["--" + k.replace("_", "-"), str(v)]
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        progress("cpu leg", leg, "failed:", r.stderr[-400:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(cells, ppc_axis, spmv_cells):
    """The CPU baseline, on rank 0 at N=1 only: the oracle's cycle on a
    bounded sample of the workload on all host threads (the value) and on one
    thread, and the reference's own operator stencils (oracle/_ref) on all
    threads and on one."""
    threads = host_threads()
    progress("cpu baseline: oracle cycle, %d threads" % threads)
    many = run_leg("cycle", threads, cpu_cells=cells, ppc=ppc_axis)
    progress("cpu baseline: oracle cycle, 1 thread")
    one = run_leg("cycle", 1, cpu_cells=cells, ppc=ppc_axis)
    if not many or many["rc"] != 0 or not one or one["rc"] != 0:
        return None
    sample = ("oracle/tpic_oracle.c (the reference's cycle restated in C; OpenMP build), %d^3 cells x 2 species "
              "x %d ppc (%d particles), %d cycles after one warm-up cycle"
              % (cells, ppc_axis ** 3, many["particles"], many["cycles"]))
    out = {"value": many["value"], "unit": "particle updates/s", "cores": threads, "kind": "port",
           "sample": sample, "host_cpus": len(os.sched_getaffinity(0)),
           "one_thread": {"value": one["value"], "unit": "particle updates/s", "cores": 1}}
    progress("cpu baseline: reference operator stencils")
    ref_many = run_leg("spmv", threads, spmv_cells=spmv_cells)
    ref_one = run_leg("spmv", 1, spmv_cells=spmv_cells)
    if ref_many and ref_one:
        out["reference_spmv"] = {
            "kind": "reference", "unit": "node updates/s (operator applies x nodes / s)",
            "value": ref_many["value"], "cores": threads, "one_thread": ref_one["value"],
            "sample": "oracle/_ref/libtpic_ref.so: apply_mass_matrix_node + node_cell_curl (both directions) "
                      "of the reference's headers, %d^3 cells, %d applies" % (spmv_cells, ref_many["applies"])}
    return out


def progress(*what):
    print("[bench]", *what, file=sys.stderr, flush=True)


def smoothing_of(args):
    """(operator, passes, alpha) of --smoothing, or None."""
    if not args.smoothing:
        return None
    op, passes, alpha = args.smoothing.split(":")
    return ({"face7": 0, "binomial27": 1}[op.lower()], int(passes), float(alpha))


def run_case(args, mass_kernel):
    """Load, initialise and time the cycle with one mass kernel; returns the
    measured figures (wall time over the timed cycles, per-phase HIP-event
    times, GMRES iterations and the energy drift)."""
    import tpic_amd
    from tpic_amd import cases
    from tpic_amd import runtime as R

    c = args.cells
    cells = (c, c, c)
    lengths = (32.0 * c / 128.0,) * 3
    ppc = (args.ppc,) * 3
    per_species = c ** 3 * args.ppc ** 3
    qom = (-64.0, 1.0)
    thermal = (0.045, 0.0056)
    sim = tpic_amd.Simulation(cells, lengths, [(qom[0], per_species), (qom[1], per_species)], dt=0.1, theta=0.5,
                              gmres_tolerance=1e-10, mass_kernel=mass_kernel, smoothing=smoothing_of(args))
    first = 0
    for s in range(2):
        sim.load_uniform(s, ppc, rho_init=1.0, thermal=(thermal[s],) * 3, first_id=first)
        first += per_species
    sim.set_field(R.FIELD_B, cases.uniform_b(cells, (0.0, 0.0, 0.01), amplitude=1e-3))
    name = MASS_NAMES[mass_kernel]
    progress(name, "loaded", 2 * per_species, "particles")
    sim.initialize()
    progress(name, "initialized")
    e_start = sim.energy().sum()
    iters = []
    for _ in range(args.warmup):
        iters.append(sim.cycle(1)["iterations"])
    sim.synchronize()
    sim.enable_timing(True)
    sim.reset_timing()
    e0 = sim.energy().sum()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iters.append(sim.cycle(1)["iterations"])
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    progress(name, "timed %d steps in %.3f s" % (args.steps, elapsed))
    e1 = sim.energy().sum()
    phases = {p: sim.phase_time(p) for p in R.PHASES}
    sim.enable_timing(False)
    # Per-cycle drift on a few more (untimed) cycles.
    drifts = []
    e_prev = e1
    for _ in range(args.drift_cycles):
        iters.append(sim.cycle(1)["iterations"])
        e_next = sim.energy().sum()
        drifts.append(abs(e_next - e_prev) / abs(e_prev))
        e_prev = e_next
    flags = sim.validation()
    # The final state, for the mixed legs' errors against the FP64 leg (the
    # same load and the same number of cycles).
    final = {name: sim.get_field(which) for name, which in
             (("E", R.FIELD_E), ("B", R.FIELD_B), ("J", R.FIELD_J), ("M", R.FIELD_MASS))}
    sim.close()

    total = 2 * per_species
    out = {"value": total * args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps}
    ncell = c ** 3
    nodes = (c + 1) ** 3
    dep_ms, dep_calls = phases["deposit"]
    asm_ms, asm_calls = phases["assemble"]
    mov_ms, mov_calls = phases["move"]
    spmv_ms, spmv_calls = phases["spmv"]
    sort_ms, sort_calls = phases["sort"]
    # Dominant kernel: the gather and deposition (one launch covers both
    # species), timed with HIP events on the context stream it is launched on.
    dep_avg = dep_ms / max(dep_calls, 1) * 1e-3
    asm_avg = asm_ms / max(asm_calls, 1) * 1e-3
    dep_flops = DEPOSIT_FLOPS_PER_PARTICLE * SPLIT_PRODUCTS[mass_kernel] * total
    useful = USEFUL_FLOPS_PER_PARTICLE * SPLIT_PRODUCTS[mass_kernel] * total
    issued = ISSUED_FLOPS_PER_PARTICLE[mass_kernel] * total
    peak = MFMA_PEAK[mass_kernel]
    alg_bytes = DEPOSIT_ALG_BYTES_PER_PARTICLE * total + DEPOSIT_ALG_BYTES_PER_NODE[mass_kernel] * nodes
    traffic = DEPOSIT_TRAFFIC_128.get(mass_kernel) if (c, args.ppc) == (128, 4) and not args.smoothing else None
    mfma = {"achieved": dep_flops / dep_avg / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": dep_flops / dep_avg / peak, "flop_per_particle": DEPOSIT_FLOPS_PER_PARTICLE,
            "useful_frac": useful / dep_avg / peak, "issued_frac": issued / dep_avg / peak}
    hbm = {"achieved": alg_bytes / dep_avg / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": alg_bytes / dep_avg / HBM_PEAK, "algorithmic_bytes": alg_bytes,
           "bytes": "%d B per particle + %d B per node" % (DEPOSIT_ALG_BYTES_PER_PARTICLE,
                                                         DEPOSIT_ALG_BYTES_PER_NODE[mass_kernel])}
    # The bound reported is the roof the kernel comes closest to (the mixed
    # forms' MFMA peaks are far away; the FP64 form carries both).
    use_mfma = mfma["frac"] >= hbm["frac"]
    bound = mfma if use_mfma else hbm
    out["roofline"] = {"bound": "mfma" if use_mfma else "hbm", "achieved": bound["achieved"],
                       "peak": bound["peak"], "unit": bound["unit"], "frac": bound["frac"], "traffic": traffic,
                       "traffic_source": DEPOSIT_TRAFFIC_SOURCE if traffic else None,
                       "kernel": KERNEL_NAME[mass_kernel], "avg_ms": dep_avg * 1e3,
                       "traffic_GBps": traffic / dep_avg / 1e9 if traffic else None,
                       "mfma": mfma, "hbm": hbm,
                       "deposit_and_assembly": {"avg_ms": (dep_avg + asm_avg) * 1e3,
                                                "hbm_frac": alg_bytes / (dep_avg + asm_avg) / HBM_PEAK}}
    out["phases"] = {p: {"ms_per_step": ms / args.steps, "calls": n} for p, (ms, n) in phases.items()}
    mov_avg = mov_ms / max(mov_calls, 1) * 1e-3
    spmv_avg = spmv_ms / max(spmv_calls, 1) * 1e-3
    sort_avg = sort_ms / max(sort_calls, 1) * 1e-3
    out["kernels"] = {
        "mover_GBps": MOVER_BYTES_PER_PARTICLE * per_species / mov_avg / 1e9 if mov_calls else None,
        "sort_plan_GBps": SORT_PLAN_BYTES_PER_PARTICLE * per_species / sort_avg / 1e9 if sort_calls else None,
        "spmv_GBps": SPMV_BYTES_PER_NODE * nodes / spmv_avg / 1e9 if spmv_calls else None,
        "spmv_node_updates_per_s": nodes / spmv_avg if spmv_calls else None,
        "deposit_GBps": (DEPOSIT_BYTES_PER_PARTICLE * total + DEPOSIT_BYTES_PER_CELL[mass_kernel] * ncell) / dep_avg / 1e9,
    }
    out["gmres_iterations"] = iters
    out["energy_drift_per_cycle"] = abs(e1 - e0) / abs(e0) / args.steps
    out["energy_drift_max_per_cycle"] = max(drifts) if drifts else None
    out["energy_drift_total"] = abs(e_prev - e_start) / abs(e_start)
    out["validation"] = flags
    return out, final


def errors_against(final, reference):
    """Largest difference relative to the largest reference value, per field."""
    import numpy as np

    out = {}
    for name, ref in reference.items():
        scale = float(np.abs(ref).max())
        out[name] = float(np.abs(final[name] - ref).max()) / scale if scale > 0 else float(np.abs(final[name]).max())
    return out


def cpu_leg_main(args):
    """A CPU leg in a child process (no GPU use): one JSON line."""
    if args.cpu_leg == "cycle":
        print(json.dumps(cpu_cycle_leg(args.cpu_cells, args.ppc)), flush=True)
    else:
        print(json.dumps(ref_spmv_leg(args.spmv_cells, args.spmv_applies)), flush=True)


def main():
    args = parse()
    if args.cpu_leg:
        cpu_leg_main(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.hbm_probe:
        hbm_probe()
        return

    c = args.cells
    total = 2 * c ** 3 * args.ppc ** 3
    plain, plain_final = run_case(args, 0)
    mixed = {}
    if not args.no_mixed:
        for k in (1, 2):
            m, final = run_case(args, k)
            m["error_vs_fp64"] = errors_against(final, plain_final)
            del final
            mixed[MASS_NAMES[k]] = m
    del plain_final
    progress("cpu baseline" if not args.no_cpu_baseline else "done")
    base = None if args.no_cpu_baseline else cpu_baseline(args.cpu_cells, args.ppc, args.spmv_cells)
    line = {
        "metric": METRIC, "value": plain["value"], "unit": "particle updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (the reference's Uniform-case load: hipRAND XORWOW, seed 12345 + 100 s, on the device)",
        "config": {"workload": "uniform magnetised plasma %d^3 cells, 2 species x %d ppc (%d particles), "
                               "dt 0.1, theta 0.5, GMRES(30) 1e-10, FP64 mass matrix" % (c, args.ppc ** 3, total),
                   "cells": c, "particles": total, "mass": "fp64",
                   **({"smoothing": args.smoothing} if args.smoothing else {})},
        "roofline": plain.pop("roofline"),
        "cpu_baseline": base,
    }
    line.update({k: v for k, v in plain.items() if k not in ("value", "ms_per_step")})
    if mixed:
        ph = plain["phases"]
        plain_dep = ph["deposit"]["ms_per_step"]
        # Amdahl: the work that does not depend on the mass precision is the
        # move, the sort plan, the field advance and the sort's record copy
        # inside the fused gather and deposition (charged at its share of the
        # kernel's algorithmic bytes); with everything else free the cycle
        # could be at most this much faster than the FP64 cycle.
        nodes = (args.cells + 1) ** 3
        copy_share = SORT_COPY_BYTES_PER_PARTICLE * total / (
            DEPOSIT_ALG_BYTES_PER_PARTICLE * total + DEPOSIT_ALG_BYTES_PER_NODE[0] * nodes)
        independent = (sum(ph[p]["ms_per_step"] for p in ("move", "sort", "advance"))
                       + copy_share * plain_dep)
        bound = plain["ms_per_step"] / max(independent, 1e-9)
        for name, m in mixed.items():
            m["speedup_vs_fp64"] = m["value"] / plain["value"]
            m["speedup_bound"] = bound
            m["deposit_speedup_vs_fp64"] = plain_dep / m["phases"]["deposit"]["ms_per_step"]
        line["mixed"] = mixed
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
