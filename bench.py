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
KERNEL_NAME = {0: "k_deposit_fp64", 1: "k_deposit_mixed<MomentsF32>", 2: "k_deposit_mixed<MomentsBf16x3>"}
# Issued MFMA work per particle of each form (the roofline of that form):
# FP64/FP32 16x16x4: two per 4 particles (2 x 2048 / 4); BF16 16x16x32:
# six per 32 particles (6 x 16384 / 32).
ISSUED_FLOPS_PER_PARTICLE = {0: 1024, 1: 1024, 2: 3072}
# Products per moment term: the BF16 form computes hi*hi + hi*lo + lo*hi.
SPLIT_PRODUCTS = {0: 1, 1: 1, 2: 3}
# HBM bytes per deposition launch at 128^3, 2 x 64 ppc, from two rocprofv3
# --pmc runs (FETCH_SIZE x 2 KiB, WRITE_SIZE x 1 KiB, calibrated on a plain
# device copy); profiles/k_deposit_hbm_traffic_128cube.md.
DEPOSIT_TRAFFIC_128 = {0: 17.448e9 + 6.936e9, 1: 19.513e9 + 3.538e9, 2: 19.306e9 + 3.546e9}

# Algorithmic work per unit.
MOVER_BYTES_PER_PARTICLE = 112      # read r,v,q (56) + key (4); write r,v (48) + key (4)
DEPOSIT_FLOPS_PER_PARTICLE = 648    # 27 pair moments x (9 alpha + 3 current) x 2
DEPOSIT_BYTES_PER_PARTICLE = 60     # read r,v,q,key
# Cell records written per cell: 324 moment columns + 81 current moments, in
# FP64 (plain form) or FP32 (mixed forms, whose per-cell sums are FP32).
DEPOSIT_BYTES_PER_CELL = {0: 405 * 8, 1: 405 * 4, 2: 405 * 4}
SORT_BYTES_PER_PARTICLE = 144       # 2 key reads + 68 B read + 68 B written
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


def cpu_baseline(cells, ppc_axis):
    """The CPU oracle (one core) on a bounded sample of the same workload."""
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
    if rc != 0:
        return None
    return {"value": sum(counts) * cycles / dt, "unit": "particle updates/s", "cores": 1, "kind": "port",
            "sample": "oracle/tpic_oracle.c, %d^3 cells x 2 species x %d ppc (%d particles), %d cycles, "
                      "1 thread" % (cells, ppc_axis ** 3, sum(counts), cycles)}


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
    sim.close()

    total = 2 * per_species
    out = {"value": total * args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps}
    ncell = c ** 3
    nodes = (c + 1) ** 3
    dep_ms, dep_calls = phases["deposit"]
    mov_ms, mov_calls = phases["move"]
    spmv_ms, spmv_calls = phases["spmv"]
    sort_ms, sort_calls = phases["sort"]
    # Dominant kernel: the deposition (one launch covers both species).
    dep_avg = dep_ms / max(dep_calls, 1) * 1e-3
    dep_flops = DEPOSIT_FLOPS_PER_PARTICLE * SPLIT_PRODUCTS[mass_kernel] * total
    peak = MFMA_PEAK[mass_kernel]
    traffic = DEPOSIT_TRAFFIC_128.get(mass_kernel) if (c, args.ppc) == (128, 4) else None
    out["roofline"] = {"bound": "mfma", "achieved": dep_flops / dep_avg / 1e12, "peak": peak / 1e12,
                       "unit": "TFLOP/s", "frac": dep_flops / dep_avg / peak, "traffic": traffic,
                       "kernel": KERNEL_NAME[mass_kernel], "avg_ms": dep_avg * 1e3,
                       "traffic_GBps": traffic / dep_avg / 1e9 if traffic else None}
    out["phases"] = {p: {"ms_per_step": ms / args.steps, "calls": n} for p, (ms, n) in phases.items()}
    mov_avg = mov_ms / max(mov_calls, 1) * 1e-3
    spmv_avg = spmv_ms / max(spmv_calls, 1) * 1e-3
    sort_avg = sort_ms / max(sort_calls, 1) * 1e-3
    out["kernels"] = {
        "mover_GBps": MOVER_BYTES_PER_PARTICLE * per_species / mov_avg / 1e9 if mov_calls else None,
        "sort_GBps": SORT_BYTES_PER_PARTICLE * per_species / sort_avg / 1e9 if sort_calls else None,
        "spmv_GBps": SPMV_BYTES_PER_NODE * nodes / spmv_avg / 1e9 if spmv_calls else None,
        "deposit_GBps": (DEPOSIT_BYTES_PER_PARTICLE * total + DEPOSIT_BYTES_PER_CELL[mass_kernel] * ncell) / dep_avg / 1e9,
    }
    out["gmres_iterations"] = iters
    out["energy_drift_per_cycle"] = abs(e1 - e0) / abs(e0) / args.steps
    out["energy_drift_max_per_cycle"] = max(drifts) if drifts else None
    out["energy_drift_total"] = abs(e_prev - e_start) / abs(e_start)
    out["validation"] = flags
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.hbm_probe:
        hbm_probe()
        return

    c = args.cells
    total = 2 * c ** 3 * args.ppc ** 3
    plain = run_case(args, 0)
    mixed = {} if args.no_mixed else {MASS_NAMES[k]: run_case(args, k) for k in (1, 2)}
    progress("cpu baseline" if not args.no_cpu_baseline else "done")
    base = None if args.no_cpu_baseline else cpu_baseline(args.cpu_cells, args.ppc)
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
        plain_dep = plain["phases"]["deposit"]["ms_per_step"]
        for name, m in mixed.items():
            m["speedup_vs_fp64"] = m["value"] / plain["value"]
            m["deposit_speedup_vs_fp64"] = plain_dep / m["phases"]["deposit"]["ms_per_step"]
        line["mixed"] = mixed
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
