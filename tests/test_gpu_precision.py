"""The 2D Harris-sheet reconnection workload (GEMHarris: sheet B0 = 0.0195,
half-width 0.5, PEC walls in y, the X-point hump seed, currentFromAmpere,
a 10% background pair, mi/me = 25, Ti/Te = 5, dt = 0.125) with the three
moment-operand precisions of the mass matrix, through the C-ABI.

- At a reduced size (64 x 32 cells, 4 species x 16 ppc) every precision is
  compared with the FP64 oracle running the same load: J and M after
  startup, then E, B and the energies after four cycles, within the bounds
  stated per precision.
- At the full size of the workload (256 x 128 cells, 4 species x 256 ppc,
  33,554,432 particles, 100 cycles) through properties: energy change per
  cycle at most 1e-10 (FP64) or 1e-8 (FP32, BF16 hi/lo), every species'
  charge unchanged, keys sorted with bins their histogram and scan, ids a
  permutation, and a clear validation word."""
import numpy as np
import pytest

import oracle as O
import tpic_amd
from tpic_amd import runtime as R

pytestmark = pytest.mark.gpu

DELTA, B0X = 0.5, 0.0195
# Ti/Te = 5 at mi/me = 25: v_th,i = v_th,e sqrt(5 / 25); the sheet drifts
# keep w0_i / w0_e = -Ti/Te (input/GEMHarris.inp of the reference).
VTH_E = 0.045
VTH_I = 0.045 * (5.0 / 25.0) ** 0.5
SPECIES = [dict(qom=-25.0, rho=1.0, drift=(0.0, 0.0, 0.0065), th=VTH_E),
           dict(qom=1.0, rho=1.0, drift=(0.0, 0.0, -0.0325), th=VTH_I),
           dict(qom=-25.0, rho=0.1, drift=(0.0, 0.0, 0.0), th=VTH_E),
           dict(qom=1.0, rho=0.1, drift=(0.0, 0.0, 0.0), th=VTH_I)]
SEED = dict(pert_hump=0.12566370614359174, deltax_hump=1.0, deltay_hump=1.0, kx_hump=0.6283185307179586,
            ky_hump=0.6283185307179586)
PERIODIC = (1, 0, 1)
FACES = [(R.FACE_PERIODIC, R.FACE_PERIODIC), (R.FACE_PEC, R.FACE_PEC), (R.FACE_PERIODIC, R.FACE_PERIODIC)]


def say(*args):
    print(*args, flush=True)


def harris_case(cells, lengths):
    import harris as H
    sd = H.seed("GEMHarris", lengths[0], lengths[1], DELTA, **SEED)
    prof = H.profiles("GEMHarris", sd, B0X, DELTA, lengths[0], lengths[1],
                      [dict(qom=sp["qom"], rho=sp["rho"], drift=sp["drift"]) for sp in SPECIES],
                      current_from_ampere=True)
    b = H.cell_field("GEMHarris", sd, B0X, DELTA, cells, lengths)
    return prof, b


def device_profile(h):
    rp = R.HarrisProfile(harris_density=h.density, is_double=h.is_double, varying_thermal=h.varying_thermal,
                         ly=h.ly, delta=h.delta, ampere_drift=h.ampere_drift, curl_delta=h.curl_delta,
                         curl_sheet=h.sheet, curl_mode=h.mode, curl_hump=h.hump)
    rp.curl_center[0], rp.curl_center[1] = h.center_x, h.center_y
    rp.curl_mode_k[0], rp.curl_mode_k[1] = h.mode_kx, h.mode_ky
    rp.curl_hump_k[0], rp.curl_hump_k[1] = h.hump_kx, h.hump_ky
    rp.curl_hump_inverse_w2[0], rp.curl_hump_inverse_w2[1] = h.hump_inverse_wx2, h.hump_inverse_wy2
    return rp


def load_gpu(sim, ppc, prof, n):
    for s, (sp, h) in enumerate(zip(SPECIES, prof)):
        sim.load_harris(s, ppc, device_profile(h), rho_init=sp["rho"], drift=sp["drift"], thermal=(sp["th"],) * 3,
                        quiet_start=True, first_id=s * n)


# Bounds against the FP64 oracle, relative to the largest value of each
# quantity (energies: to the system energy). FP64: the reference's moment
# tolerance 3e-12 (sorted_mass_device.cpp:359-367), fields 1e-10 and energies
# 1e-12 (summation order only). FP32 operands round each product to 2^-24 and
# sum a cell in FP32; BF16 hi/lo drops lo*lo (about 2^-16 per product); the
# field errors follow the moment errors through the solve.
BOUNDS = {R.MASS_FP64: dict(moments=3e-12, fields=1e-10, energy=1e-12),
          R.MASS_FP32: dict(moments=2e-6, fields=1e-5, energy=1e-7),
          R.MASS_BF16X3: dict(moments=5e-5, fields=2e-4, energy=2e-6)}
KERNELS = [R.MASS_FP64, R.MASS_FP32, R.MASS_BF16X3]
IDS = ["fp64", "fp32", "bf16x3"]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_harris_precisions_against_the_fp64_oracle(kernel):
    cells, lengths = (64, 32, 1), (6.4, 3.2, 0.1)
    ppc = (4, 4, 1)
    n = int(np.prod(cells)) * 16
    prof, b = harris_case(cells, lengths)
    bound = BOUNDS[kernel]
    g = O.make_grid(cells, lengths, PERIODIC, FACES)
    say("harris reduced size, kernel", kernel)
    with tpic_amd.Simulation(cells, lengths, [(sp["qom"], n) for sp in SPECIES], dt=0.125, theta=0.5,
                             gmres_tolerance=1e-12, particle_periodic=PERIODIC, field_faces=FACES,
                             mass_kernel=kernel) as sim:
        ref = O.Sim(cells, lengths, [sp["qom"] for sp in SPECIES], [n] * 4, 0.125, tol=1e-12,
                    particle_periodic=PERIODIC, field_faces=FACES)
        load_gpu(sim, ppc, prof, n)
        for s, (sp, h) in enumerate(zip(SPECIES, prof)):
            p = O.load_uniform(g, ppc, sp["rho"], sp["drift"], (sp["th"],) * 3, True, O.load_seed(s), s * n,
                               sp["qom"], harris_profile=h)
            ref.set_particles(s, p.a)
        sim.set_field(R.FIELD_B, b)
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        err = {"J": rel(sim.get_field(R.FIELD_J), ref.field("J")),
               "M": rel(sim.get_field(R.FIELD_MASS), ref.field("mass"))}
        say("startup", err)
        assert err["J"] <= bound["moments"] and err["M"] <= bound["moments"]
        for cycle in range(4):
            rg = sim.cycle(1)
            rc, ro = ref.cycle(1)
            assert rc == 0 and rg["converged"]
            eg, eo = sim.energy(), ref.energy()
            e_err = float(np.abs(eg - eo).max() / abs(eo.sum()))
            errs = {name: rel(sim.get_field(w), ref.field(name)) for w, name in ((R.FIELD_E, "E"), (R.FIELD_B, "B"))}
            say("cycle", cycle, "iterations", rg["iterations"], ro.iterations, "energy err", e_err, errs)
            assert e_err <= bound["energy"]
            assert errs["E"] <= bound["fields"] and errs["B"] <= bound["fields"]
            if kernel == R.MASS_FP64:
                assert rg["iterations"] == ro.iterations
        assert sim.validation() == 0
        ref.close()


def check_species(sim, s, first, n):
    """Sortedness, bins, id permutation; returns the total charge."""
    a = sim.get_particles(s)
    keys = a["cell"]
    assert bool(np.all(keys[1:] >= keys[:-1])), "keys not sorted"
    counts, offsets = sim.get_bins(s)
    assert np.array_equal(np.bincount(keys, minlength=sim.bins), counts.astype(np.int64))
    assert offsets[0] == 0 and np.array_equal(np.cumsum(counts.astype(np.int64)), offsets[1:].astype(np.int64))
    seen = np.bincount((a["id"] - np.uint64(first)).astype(np.int64), minlength=n)
    assert seen.size == n and bool(np.all(seen == 1)), "ids are not a permutation"
    return float(np.sum(a["q"]))


@pytest.mark.parametrize("kernel", KERNELS, ids=IDS)
def test_harris_full_size_property_run(kernel):
    cells, lengths = (256, 128, 1), (25.6, 12.8, 0.1)
    ppc = (16, 16, 1)
    n = int(np.prod(cells)) * 256
    prof, b = harris_case(cells, lengths)
    say("harris full size, 4 species x", n, "kernel", kernel)
    drift_bound = 1e-10 if kernel == R.MASS_FP64 else 1e-8
    with tpic_amd.Simulation(cells, lengths, [(sp["qom"], n) for sp in SPECIES], dt=0.125, theta=0.5,
                             gmres_tolerance=1e-10, particle_periodic=PERIODIC, field_faces=FACES,
                             mass_kernel=kernel) as sim:
        load_gpu(sim, ppc, prof, n)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        charge0 = [check_species(sim, s, s * n, n) for s in range(4)]
        e_prev = sim.energy().sum()
        worst = 0.0
        for cycle in range(100):
            r = sim.cycle(1)
            assert r["converged"] == 1
            e = sim.energy().sum()
            worst = max(worst, abs(e - e_prev) / abs(e_prev))
            if cycle % 20 == 19:
                say("cycle", cycle, "iterations", r["iterations"], "worst dE/E per cycle so far", worst)
            e_prev = e
        assert worst <= drift_bound
        for s in range(4):
            charge = check_species(sim, s, s * n, n)
            assert abs(charge - charge0[s]) <= 1e-12 * abs(charge0[s])
        assert sim.validation() == 0
