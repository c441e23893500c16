"""The tpic_mi_run driver on the GPU: a deck run's reported energies against
the oracle running the same deck (the reference's Uniform-case load with
CenterInitialVelocity, B = B0, E = 0, then ECSIM cycles), and the report
lines of Pipeline::run_cycles (src/simulation/run_reporter.cpp)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "tpic_amd", "lib", "tpic_mi_run")
DECKS = os.path.join(ROOT, "tests", "decks")

# The decks under tests/decks, restated here so that a parser mistake
# cannot hide in both sides.
STARTUP = dict(
    deck="uniform_startup.inp", cells=(7, 3, 2), lengths=(7.0, 3.0, 2.0), periodic=(1, 1, 1), b0=(0.25, 0.125, -0.0625),
    dt=0.125, ncycles=3, theta=0.5, tol=1e-10, energy_cycle=1, quiet=False,
    species=[dict(rho=1.0, qom=-1.0, ppc=(2, 2, 2), th=(0.1,) * 3, drift=(0.125, 0.0, 0.0), center=True),
             dict(rho=0.5, qom=1.0, ppc=(2, 2, 2), th=(0.05,) * 3, drift=(-0.125, 0.0, 0.0), center=True),
             dict(rho=0.5, qom=0.5, ppc=(2, 2, 2), th=(0.025,) * 3, drift=(0.0, 0.0625, 0.0), center=True)])
MIRROR = dict(
    deck="uniform_mirror.inp", cells=(8, 6, 4), lengths=(8.0, 6.0, 4.0), periodic=(1, 0, 1), b0=(0.0, 0.0, 0.2),
    dt=0.1, ncycles=4, theta=0.5, tol=1e-12, energy_cycle=2, quiet=True,
    species=[dict(rho=1.0, qom=-16.0, ppc=(2, 2, 2), th=(0.08,) * 3, drift=(0.05, 0.02, 0.0), center=False),
             dict(rho=1.0, qom=1.0, ppc=(1, 2, 2), th=(0.02,) * 3, drift=(0.0, 0.0, 0.0), center=False)])

ANISOTROPY = dict(
    deck="anisotropy_2d.inp", cells=(32, 16, 1), lengths=(2.5, 1.25, 0.078125), periodic=(1, 1, 1),
    b0=(0.015, 0.0, 0.0), dt=0.015, ncycles=3, theta=0.5, tol=1e-14, energy_cycle=1, quiet=False,
    species=[dict(rho=1.0, qom=-1000.0, ppc=(4, 4, 2), th=(0.01, 0.03, 0.03), drift=(0.0, 0.0, 0.0), center=True),
             dict(rho=1.0, qom=1.0, ppc=(4, 4, 2), th=(0.0009486832980505138,) * 3, drift=(0.0, 0.0, 0.0),
                  center=True)])

TWO_PI = 6.283185307179586476925286766559
K2 = TWO_PI * 2 / 62.83185307179586


def two_stream_fields(cells, lengths):
    """E_x = -(rho_e A / k) sin(k x + phase) at the nodes x = i dx (two_stream.cpp:60-73)."""
    nx, ny, nz = cells
    x = np.arange(nx + 1) * (lengths[0] / nx)
    e = np.zeros((3, nz + 1, ny + 1, nx + 1))
    e[0] = (-(1.0 * 0.001 / K2)) * np.sin(K2 * x + 0.25)
    return e.ravel(), np.zeros(3 * nx * ny * nz)


def weibel_fields(cells, lengths):
    """B_z = A cos(k x) at the cell centres, E_y = -(growth / k) A sin(k x) at
    the nodes (weibel.cpp:9-40)."""
    nx, ny, nz = cells
    k = TWO_PI / lengths[0]
    dx = lengths[0] / nx
    e = np.zeros((3, nz + 1, ny + 1, nx + 1))
    e[1] = (-(0.1 / k) * 0.001) * np.sin(k * np.arange(nx + 1) * dx)
    b = np.zeros((3, nz, ny, nx))
    b[2] = 0.001 * np.cos(k * (np.arange(nx) + 0.5) * dx)
    return e.ravel(), b.ravel()


TWO_STREAM = dict(
    deck="two_stream_1d.inp", cells=(64, 1, 1), lengths=(62.83185307179586, 0.9817477042468103, 0.9817477042468103),
    periodic=(1, 1, 1), b0=(0.0, 0.0, 0.0), dt=0.1, ncycles=5, theta=0.5, tol=1e-12, energy_cycle=1, quiet=False,
    fields=two_stream_fields,
    species=[dict(rho=0.5, qom=-1.0, ppc=(16, 1, 1), th=(0.01,) * 3, drift=(0.2, 0.0, 0.0), center=False,
                  seed=(K2, 0.25, 0.001)),
             dict(rho=0.5, qom=-1.0, ppc=(16, 1, 1), th=(0.01,) * 3, drift=(-0.2, 0.0, 0.0), center=False,
                  seed=(K2, 0.25, 0.001)),
             dict(rho=1.0, qom=0.01, ppc=(16, 1, 1), th=(0.001,) * 3, drift=(0.0, 0.0, 0.0), center=False)])
WEIBEL = dict(
    deck="weibel_2d.inp", cells=(32, 32, 1), lengths=(3.2, 3.2, 0.1), periodic=(1, 1, 1), b0=(0.0, 0.0, 0.0),
    dt=0.05, ncycles=4, theta=0.5, tol=1e-12, energy_cycle=2, quiet=False, fields=weibel_fields,
    species=[dict(rho=1.0, qom=-1.0, ppc=(4, 4, 1), th=(0.3, 0.1, 0.3), drift=(0.0, 0.0, 0.0), center=True),
             dict(rho=1.0, qom=0.01, ppc=(4, 4, 1), th=(0.01,) * 3, drift=(0.0, 0.0, 0.0), center=True)])

PEC, PMC, PERIODIC = O.FACE_PEC, O.FACE_PMC, O.FACE_PERIODIC
PEC_SLAB = dict(
    deck="uniform_pec_slab.inp", cells=(8, 6, 4), lengths=(4.0, 3.0, 2.0), periodic=(1, 0, 1),
    faces=[(PERIODIC, PERIODIC), (PEC, PEC), (PERIODIC, PERIODIC)], b0=(0.2, 0.0, 0.1), dt=0.1, ncycles=3,
    theta=0.5, tol=1e-12, energy_cycle=1, quiet=False,
    species=[dict(rho=1.0, qom=-16.0, ppc=(2, 2, 1), th=(0.08,) * 3, drift=(0.0, 0.0, 0.0), center=True),
             dict(rho=1.0, qom=1.0, ppc=(2, 2, 1), th=(0.02,) * 3, drift=(0.0, 0.0, 0.0), center=True)])
PMC_BOX = dict(
    deck="uniform_pmc_box.inp", cells=(6, 4, 4), lengths=(3.0, 2.0, 2.0), periodic=(0, 1, 0),
    faces=[(PMC, PMC), (PERIODIC, PERIODIC), (PMC, PEC)], b0=(0.15, 0.0, 0.0), dt=0.1, ncycles=3,
    theta=0.5, tol=1e-12, energy_cycle=1, quiet=True,
    species=[dict(rho=1.0, qom=-9.0, ppc=(2, 1, 2), th=(0.07,) * 3, drift=(0.0, 0.0, 0.0), center=False),
             dict(rho=1.0, qom=1.0, ppc=(2, 1, 2), th=(0.02,) * 3, drift=(0.0, 0.0, 0.0), center=False)])
# The same runs with energy-conserving smoothing (tpic_mi_config.smooth_*).
PEC_SLAB_FACE7 = dict(PEC_SLAB, deck="uniform_pec_slab_face7.inp", smoothing=(O.SMOOTH_FACE7, 3, 0.5))
PMC_BOX_BINOMIAL27 = dict(PMC_BOX, deck="uniform_pmc_box_binomial27.inp", smoothing=(O.SMOOTH_BINOMIAL27, 2, 0.0))

def harris_fields(case, seed_kw, b0x, delta):
    def fields(cells, lengths):
        import harris as H
        sd = H.seed(case, lengths[0], lengths[1], delta, **seed_kw)
        nx, ny, nz = cells
        return np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1)), H.cell_field(case, sd, b0x, delta, cells, lengths)
    return fields


def harris_profiles(case, seed_kw, b0x, delta, ampere):
    def make(d):
        import harris as H
        sd = H.seed(case, d["lengths"][0], d["lengths"][1], delta, **seed_kw)
        species = [dict(qom=sp["qom"], rho=sp["rho"], drift=sp["drift"]) for sp in d["species"]]
        return H.profiles(case, sd, b0x, delta, d["lengths"][0], d["lengths"][1], species, current_from_ampere=ampere)
    return make


GEM_SEED = dict(pert_hump=0.12566370614359174, deltax_hump=1.0, deltay_hump=1.0, kx_hump=0.6283185307179586,
                ky_hump=0.6283185307179586)
HARRIS_SPECIES = [
    dict(rho=1.0, qom=-25.0, ppc=(2, 2, 1), th=(0.045,) * 3, drift=(0.0, 0.0, 0.0065), center=False),
    dict(rho=1.0, qom=1.0, ppc=(2, 2, 1), th=(0.0126,) * 3, drift=(0.0, 0.0, -0.0325), center=False),
    dict(rho=0.1, qom=-25.0, ppc=(2, 2, 1), th=(0.045,) * 3, drift=(0.0, 0.0, 0.0), center=False),
    dict(rho=0.1, qom=1.0, ppc=(2, 2, 1), th=(0.0126,) * 3, drift=(0.0, 0.0, 0.0), center=False)]
GEM_HARRIS = dict(
    deck="gem_harris_2d.inp", cells=(32, 16, 1), lengths=(12.8, 6.4, 0.4), periodic=(1, 0, 1),
    faces=[(PERIODIC, PERIODIC), (PEC, PEC), (PERIODIC, PERIODIC)], b0=(0.0, 0.0, 0.0), dt=0.125, ncycles=3,
    theta=0.5, tol=1e-12, energy_cycle=1, quiet=True, species=HARRIS_SPECIES,
    fields=harris_fields("GEMHarris", GEM_SEED, 0.0195, 0.5),
    profiles=harris_profiles("GEMHarris", GEM_SEED, 0.0195, 0.5, True))
DOUBLE_GEM = dict(
    deck="double_gem_2d.inp", cells=(32, 32, 1), lengths=(12.8, 12.8, 0.4), periodic=(1, 1, 1), b0=(0.0, 0.0, 0.0),
    dt=0.125, ncycles=3, theta=0.5, tol=1e-12, energy_cycle=1, quiet=False, species=HARRIS_SPECIES,
    fields=harris_fields("DoubleGEM", {}, 0.02, 0.5), profiles=harris_profiles("DoubleGEM", {}, 0.02, 0.5, False))

def spectrum_fields(amplitude):
    """B = B0 + the seeded k-spectrum (tpic_amd.cases.uniform_b, the same
    definition as the driver's MagneticSeedSpectrum* keys), E = 0."""
    def fields(cells, lengths):
        from tpic_amd import cases
        nx, ny, nz = cells
        return np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1)), cases.uniform_b(cells, (0.0, 0.0, 0.01), amplitude, 4, 7)
    return fields


# b0 is inside the fields image here (oracle_run adds d["b0"] on top).
UNIFORM_SPECTRUM = dict(
    deck="uniform_16_spectrum.inp", cells=(16, 16, 16), lengths=(4.0, 4.0, 4.0), periodic=(1, 1, 1),
    b0=(0.0, 0.0, 0.0), dt=0.1, ncycles=3, theta=0.5, tol=1e-10, energy_cycle=1, quiet=False,
    fields=spectrum_fields(0.1),
    species=[dict(rho=1.0, qom=-64.0, ppc=(4, 4, 4), th=(0.045,) * 3, drift=(0.0, 0.0, 0.0), center=False),
             dict(rho=1.0, qom=1.0, ppc=(4, 4, 4), th=(0.0056,) * 3, drift=(0.0, 0.0, 0.0), center=False)])

CYCLE = re.compile(r"ECSIM Cycle (\d+)  \|  t = (\S+)")
FINAL = re.compile(r"ECSIM Final State  \|  cycle = (\d+), t = (\S+)")
FIELD = re.compile(r"Field energy  : electric = (\S+)  magnetic = (\S+)")
TOTAL = re.compile(r"Total energy  : system   = (\S+)  kinetic  = (\S+)")
GMRES = re.compile(r"Maxwell GMRES : converged in (\d+) iterations; residual = (\S+)")


def say(*args):
    print(*args, flush=True)


def run_deck(deck, *args):
    say("tpic_mi_run --run", deck, *args)
    r = subprocess.run([RUN, "--run", os.path.join(DECKS, deck), *args], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    return r.stdout


def reported(text):
    """{cycle: (electric, magnetic, system, kinetic)} and the GMRES lines."""
    energies, gmres, current = {}, [], None
    for line in text.splitlines():
        m = CYCLE.search(line) or FINAL.search(line)
        if m:
            current = int(m.group(1))
            continue
        m = FIELD.search(line)
        if m:
            field = (float(m.group(1)), float(m.group(2)))
            continue
        m = TOTAL.search(line)
        if m:
            energies[current] = field + (float(m.group(1)), float(m.group(2)))
            continue
        m = GMRES.search(line)
        if m:
            gmres.append((int(m.group(1)), float(m.group(2))))
    return energies, gmres


def oracle_run(d):
    """The oracle's run of deck d: energies of the state entering each
    reported cycle and of the final state."""
    cells = d["cells"]
    g = O.make_grid(cells, d["lengths"], d["periodic"], d.get("faces"))
    loads, first = [], 0
    profiles = d["profiles"](d) if "profiles" in d else [None] * len(d["species"])
    for s, sp in enumerate(d["species"]):
        p = O.load_uniform(g, sp["ppc"], sp["rho"], sp["drift"], sp["th"], d["quiet"], O.load_seed(s), first,
                           sp["qom"], density_seed=sp.get("seed"), harris_profile=profiles[s])
        if sp["center"]:
            assert O.center_velocity(p, sp["drift"]) == 0
        first += p.count
        loads.append(p)
    sim = O.Sim(cells, d["lengths"], [sp["qom"] for sp in d["species"]], [p.count for p in loads], d["dt"],
                theta=d["theta"], tol=d["tol"], particle_periodic=d["periodic"], field_faces=d.get("faces"),
                smoothing=d.get("smoothing"))
    for s, p in enumerate(loads):
        sim.set_particles(s, p.a)
    b = sim.field("B")
    n = int(np.prod(cells))
    if "fields" in d:
        e_img, b_img = d["fields"](cells, d["lengths"])
        sim.field("E")[:] = e_img
        b[:] = b_img
    else:
        b[:] = 0.0
    for c in range(3):
        b[c * n:(c + 1) * n] += d["b0"][c]
    sim.initialize()

    def sample():
        e = sim.energy()
        kinetic = float(np.sum(e[2:]))
        return (e[0], e[1], e[0] + e[1] + kinetic, kinetic)

    out, iterations = {}, []
    for cycle in range(d["ncycles"]):
        if cycle % d["energy_cycle"] == 0:
            out[cycle] = sample()
        rc, r = sim.cycle(1)
        assert rc == 0
        iterations.append(r.iterations)
    out[d["ncycles"]] = sample()
    sim.close()
    return out, iterations


@pytest.mark.parametrize("d", [STARTUP, MIRROR, ANISOTROPY, TWO_STREAM, WEIBEL, PEC_SLAB, PMC_BOX, GEM_HARRIS,
                               DOUBLE_GEM, PEC_SLAB_FACE7, PMC_BOX_BINOMIAL27, UNIFORM_SPECTRUM],
                         ids=["uniform_startup", "uniform_mirror", "anisotropy_2d", "two_stream_1d", "weibel_2d",
                              "uniform_pec_slab", "uniform_pmc_box", "gem_harris_2d", "double_gem_2d",
                              "uniform_pec_slab_face7", "uniform_pmc_box_binomial27", "uniform_16_spectrum"])
def test_deck_run_energies_match_the_oracle(d):
    """FP64 run: every reported energy within 1e-10 of the oracle's (relative
    to the system energy; the report prints 13 significant digits and the two
    sides differ only in FP64 summation order), and the solver iteration
    counts within one of the oracle's."""
    energies, gmres = reported(run_deck(d["deck"]))
    expect, iterations = oracle_run(d)
    assert sorted(energies) == sorted(expect)
    for cycle, want in expect.items():
        got = energies[cycle]
        scale = abs(want[2])
        for k in range(4):
            assert abs(got[k] - want[k]) <= 1e-10 * scale, (cycle, k, got, want)
    assert len(gmres) == d["ncycles"]
    for (its, residual), want in zip(gmres, iterations):
        assert abs(its - want) <= 1 and residual <= d["tol"]


def test_deck_run_report_lines():
    text = run_deck("uniform_startup.inp", "--timing")
    assert "ECSIM Initialization complete" in text and "Starting ECSIM Main Simulation Loop" in text
    for cycle in range(3):
        assert " ======= ECSIM Cycle %d  |  t = %.6e =======" % (cycle, cycle * 0.125) in text
    assert " ======= ECSIM Final State  |  cycle = 3, t = 3.750000e-01 =======" in text
    assert "Energy change : dE/E0 = " in text
    assert "ECSIM Main Loop Complete" in text and "  Total cycles: 3" in text
    assert re.search(r"Ordinary after warm-up: 2 cycles; avg=\S+ std=\S+ min=\S+ max=\S+ ms", text)
    assert re.search(r"deposit  :\s+\S+ ms in \d+ calls", text)
    energies, _ = reported(text)
    # the first cycle's energy is the dE baseline: dE/E0 = 0 there
    assert "Energy change : dE/E0 = +0.000000e+00  dE = +0.000000e+00" in text
    assert energies[0][0] == 0.0  # E starts at zero


@pytest.mark.parametrize("operands", ["fp32", "bf16x3"])
def test_deck_run_with_mixed_moment_operands(operands):
    """The mixed moment operands change the run by rounding only: energies
    within 1e-6 of the FP64 run's (relative to the system energy)."""
    ref, _ = reported(run_deck("uniform_mirror.inp"))
    got, gmres = reported(run_deck("uniform_mirror.inp", "--mass-precision", operands))
    assert sorted(got) == sorted(ref) and len(gmres) == 4
    for cycle in ref:
        for k in range(4):
            assert abs(got[cycle][k] - ref[cycle][k]) <= 1e-6 * abs(ref[cycle][2]), (cycle, k)


def test_deck_initialize_only():
    say("tpic_mi_run --initialize uniform_startup.inp")
    r = subprocess.run([RUN, "--initialize", os.path.join(DECKS, "uniform_startup.inp")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Initialization complete" in r.stdout and "ECSIM Cycle" not in r.stdout


def read_vti(path):
    """(cycle, time, state, name, cell, values[n, 3]) of a field snapshot."""
    raw = open(path, "rb").read()
    marker = raw.index(b'<AppendedData encoding="raw">\n_') + len(b'<AppendedData encoding="raw">\n_')
    head = raw[:marker].decode()
    nbytes = int(np.frombuffer(raw[marker:marker + 8], dtype=np.uint64)[0])
    values = np.frombuffer(raw[marker + 8:marker + 8 + nbytes], dtype=np.float64).reshape(-1, 3)
    assert raw[marker + 8 + nbytes:] == b"\n  </AppendedData>\n</VTKFile>\n"
    cycle = int(re.search(r'Name="cycle"[^>]*>(-?\d+)<', head).group(1))
    time = float(re.search(r'Name="TimeValue"[^>]*>([^<]+)<', head).group(1))
    name = re.search(r'<DataArray type="Float64" Name="(\w+)" NumberOfComponents="3"', head).group(1)
    return cycle, time, name, "<CellData>" in head, values, head


def test_deck_run_writes_diagnostics_and_field_snapshots(tmp_path):
    """FieldOutputCycle / DiagnosticsOutputCycle: diagnostics.csv in the
    reference's layout (diagnostics_writer.cpp) with one pre_solve row per
    due cycle and a final row, and E/B snapshots as VTK ImageData with the
    appended raw Float64 array (vtk_field_engine.cpp), every value against
    the oracle's run of the same deck; existing files are refused unless
    --overwrite-output."""
    d = STARTUP
    out = tmp_path / "out"
    text = open(os.path.join(DECKS, d["deck"])).read()
    text += "SaveDirName = %s\nFieldOutputCycle = 2\nDiagnosticsOutputCycle = 1\nFieldOutputTag = B+E\n" % out
    deck = tmp_path / "deck.inp"
    deck.write_text(text)
    say("tpic_mi_run --run with output")
    r = subprocess.run([RUN, "--run", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Diagnostics rows" in r.stdout and "Field snapshots" in r.stdout
    # the oracle's states entering cycles 0..2 and the final state
    cells = d["cells"]
    g = O.make_grid(cells, d["lengths"], d["periodic"])
    sim = O.Sim(cells, d["lengths"], [sp["qom"] for sp in d["species"]],
                [int(np.prod(cells)) * 8] * 3, d["dt"], theta=d["theta"], tol=d["tol"])
    first = 0
    for s, sp in enumerate(d["species"]):
        p = O.load_uniform(g, sp["ppc"], sp["rho"], sp["drift"], sp["th"], False, O.load_seed(s), first, sp["qom"])
        O.center_velocity(p, sp["drift"])
        first += p.count
        sim.set_particles(s, p.a)
    nc = int(np.prod(cells))
    for c in range(3):
        sim.field("B")[c * nc:(c + 1) * nc] = d["b0"][c]
    sim.initialize()
    states = []
    for cycle in range(d["ncycles"] + 1):
        states.append((sim.scalars(), sim.field("E").copy(), sim.field("B").copy()))
        if cycle < d["ncycles"]:
            assert sim.cycle(1)[0] == 0
    sim.close()
    lines = (out / "diagnostics.csv").read_text().splitlines()
    assert lines[0] == "# precision_profile = double/double"
    assert lines[1].startswith("# precision_roles = storage: position=float64")
    header = [l for l in lines if l.startswith("cycle,")][0].split(",")
    assert header[:12] == ["cycle", "time", "state", "electric_energy", "magnetic_energy", "kinetic_energy",
                           "total_energy", "total_energy_change", "charge", "momentum_x", "momentum_y", "momentum_z"]
    assert len(header) == 12 + 5 * 3 and header[-1] == "momentum_z_2"
    rows = [l.split(",") for l in lines if l and l[0].isdigit()]
    assert [(int(r[0]), r[2]) for r in rows] == [(0, "pre_solve"), (1, "pre_solve"), (2, "pre_solve"), (3, "final")]
    e0 = states[0][0][0] + states[0][0][1] + sum(states[0][0][2::5])
    for r, (sc, _, _) in zip(rows, states):
        vals = np.array([float(v) for v in r[3:]])
        total = sc[0] + sc[1] + sum(sc[2::5])
        want = [sc[0], sc[1], sum(sc[2::5]), total, total - e0, sum(sc[3::5]), sum(sc[4::5]), sum(sc[5::5]),
                sum(sc[6::5])] + list(sc[2:])
        scale = np.maximum(np.abs(want), total)
        assert np.all(np.abs(vals - np.array(want)) <= 1e-10 * scale), r[0]
    nx, ny, nz = cells
    written = sorted(p.name for p in out.glob("*.vti"))
    assert written == ["B_t000000.vti", "B_t000002.vti", "B_t000003.vti", "E_t000000.vti", "E_t000002.vti",
                       "E_t000003.vti"]
    for cycle in (0, 2, 3):
        _, e_want, b_want = states[cycle]
        cyc, time, name, cell, values, head = read_vti(out / ("E_t%06d.vti" % cycle))
        assert (cyc, name, cell) == (cycle, "E", False) and abs(time - cycle * d["dt"]) <= 1e-15
        assert 'WholeExtent="0 %d 0 %d 0 %d"' % cells in head
        assert values.shape == ((nx + 1) * (ny + 1) * (nz + 1), 3)
        scale = max(np.abs(e_want).max(), 1e-30)
        assert np.abs(values.T.ravel() - e_want).max() <= 1e-10 * scale
        cyc, time, name, cell, values, head = read_vti(out / ("B_t%06d.vti" % cycle))
        assert (cyc, name, cell) == (cycle, "B", True)
        assert np.abs(values.T.ravel() - b_want).max() <= 1e-10 * np.abs(b_want).max()
        flag = re.search(r'Name="final_state" NumberOfTuples="1" format="ascii">(\d)<', head).group(1)
        assert flag == ("1" if cycle == 3 else "0")
    r = subprocess.run([RUN, "--run", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "File exists" in r.stderr
    r = subprocess.run([RUN, "--run", str(deck), "--overwrite-output"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def read_vtp(path):
    """(head, {name: values}) of a particle snapshot piece: the appended
    arrays in offset order, each a UInt64 byte count and its values."""
    raw = open(path, "rb").read()
    marker = raw.index(b'<AppendedData encoding="raw">\n_') + len(b'<AppendedData encoding="raw">\n_')
    head = raw[:marker].decode()
    arrays = re.findall(r'<DataArray type="(\w+)" Name="(\w+)" NumberOfComponents="(\d)" format="appended" '
                        r'offset="(\d+)"/>', head)
    types = {"Float64": np.float64, "UInt64": np.uint64, "Int64": np.int64}
    out, at = {}, marker
    for kind, name, comps, offset in sorted(arrays, key=lambda a: int(a[3])):
        assert at - marker == int(offset), name
        nbytes = int(np.frombuffer(raw[at:at + 8], dtype=np.uint64)[0])
        out[name] = np.frombuffer(raw[at + 8:at + 8 + nbytes], dtype=types[kind]).reshape(-1, int(comps))
        at += 8 + nbytes
    assert raw[at:] == b"\n  </AppendedData>\n</VTKFile>\n"
    return head, out


def test_deck_run_writes_particle_snapshots(tmp_path):
    """ParticlesOutputCycle: particles_t<cycle>_r000000.vtp (VTK PolyData:
    velocity, charge, id, species, positions (cell + fraction) * spacing, one
    vertex per particle; vtk_particle_engine.cpp) and its .pvtp manifest at
    the due cycles and the final state, against the oracle's particles of the
    same deck in the same stored order: ids and species exactly, positions
    and velocities within 1e-9 of their scale."""
    d = STARTUP
    out = tmp_path / "out"
    text = open(os.path.join(DECKS, d["deck"])).read()
    text += "SaveDirName = %s\nParticlesOutputCycle = 2\n" % out
    deck = tmp_path / "deck.inp"
    deck.write_text(text)
    say("tpic_mi_run --run with particle output")
    r = subprocess.run([RUN, "--run", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Particle snapshots             : 3" in r.stdout
    assert "Last particle snapshot         : %s" % (out / "particles_t000003.pvtp") in r.stdout
    cells = d["cells"]
    g = O.make_grid(cells, d["lengths"], d["periodic"])
    sim = O.Sim(cells, d["lengths"], [sp["qom"] for sp in d["species"]],
                [int(np.prod(cells)) * 8] * 3, d["dt"], theta=d["theta"], tol=d["tol"])
    first = 0
    for s, sp in enumerate(d["species"]):
        p = O.load_uniform(g, sp["ppc"], sp["rho"], sp["drift"], sp["th"], False, O.load_seed(s), first, sp["qom"])
        O.center_velocity(p, sp["drift"])
        first += p.count
        sim.set_particles(s, p.a)
    nc = int(np.prod(cells))
    for c in range(3):
        sim.field("B")[c * nc:(c + 1) * nc] = d["b0"][c]
    sim.initialize()
    spacing = [d["lengths"][a] / cells[a] for a in range(3)]
    ex, ey = cells[0] + 4, cells[1] + 4
    for cycle in range(d["ncycles"] + 1):
        if cycle in (0, 2, 3):
            parts = [sim.particles(s) for s in range(len(d["species"]))]
            key = np.concatenate([p["cell"] for p in parts]).astype(np.int64)
            ijk = [key % ex - 2, (key // ex) % ey - 2, key // (ex * ey) - 2]
            pos = np.stack([(0.0 + ijk[a] + np.concatenate([p[("rx", "ry", "rz")[a]] for p in parts])) * spacing[a]
                            for a in range(3)], axis=1)
            vel = np.stack([np.concatenate([p[k] for p in parts]) for k in ("u", "v", "w")], axis=1)
            head, got = read_vtp(out / ("particles_t%06d_r000000.vtp" % cycle))
            n = len(key)
            assert 'NumberOfPoints="%d" NumberOfVerts="%d"' % (n, n) in head
            assert np.array_equal(got["id"].ravel(), np.concatenate([p["id"] for p in parts]))
            assert np.array_equal(got["species"].ravel(),
                                  np.concatenate([np.full(len(p["id"]), s, np.uint64) for s, p in enumerate(parts)]))
            assert np.abs(got["position"] - pos).max() <= 1e-9 * max(d["lengths"])
            assert np.abs(got["velocity"] - vel).max() <= 1e-9 * np.abs(vel).max()
            assert np.abs(got["charge"].ravel() - np.concatenate([p["q"] for p in parts])).max() <= 1e-15
            assert np.array_equal(got["connectivity"].ravel(), np.arange(n))
            assert np.array_equal(got["offsets"].ravel(), np.arange(1, n + 1))
            manifest = (out / ("particles_t%06d.pvtp" % cycle)).read_text()
            assert '<Piece Source="particles_t%06d_r000000.vtp"/>' % cycle in manifest
            assert ('Name="final_state" NumberOfTuples="1" format="ascii">%d<' % (cycle == 3)) in manifest
        if cycle < d["ncycles"]:
            assert sim.cycle(1)[0] == 0
    sim.close()


def test_throughput_workload_deck_runs_the_bench_case(tmp_path):
    """tests/decks/uniform_128_spectrum.inp (the throughput workload, cut to
    two cycles) through the driver equals the same case set up through the
    Python binding as bench.py does (the reference's Uniform load, B0 plus
    the seeded spectrum): every reported energy within 1e-12 of the system
    energy (the report prints 13 significant digits)."""
    import tpic_amd
    from tpic_amd import cases
    from tpic_amd import runtime as R
    text = open(os.path.join(DECKS, "uniform_128_spectrum.inp")).read().replace("ncycles = 20", "ncycles = 2")
    deck = tmp_path / "c3.inp"
    deck.write_text(text)
    say("tpic_mi_run --run uniform_128_spectrum (2 cycles)")
    r = subprocess.run([RUN, "--run", str(deck)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    energies, gmres = reported(r.stdout)
    cells, n = (128, 128, 128), 128 ** 3 * 64
    with tpic_amd.Simulation(cells, (32.0,) * 3, [(-64.0, n), (1.0, n)], dt=0.1, theta=0.5,
                             gmres_tolerance=1e-10) as sim:
        for s, th in enumerate((0.045, 0.0056)):
            sim.load_uniform(s, (4, 4, 4), rho_init=1.0, thermal=(th,) * 3, first_id=s * n)
        sim.set_field(R.FIELD_B, cases.uniform_b(cells, (0.0, 0.0, 0.01), amplitude=1e-3))
        sim.initialize()
        want = {}
        for cycle in range(3):
            e = sim.energy()
            want[cycle] = (e[0], e[1], float(np.sum(e)), float(np.sum(e[2:])))
            if cycle < 2:
                sim.cycle(1)
    assert sorted(energies) == sorted(want)
    for cycle, w in want.items():
        for k in range(4):
            assert abs(energies[cycle][k] - w[k]) <= 1e-12 * abs(w[2]), (cycle, k, energies[cycle], w)



def test_deck_restart_continues_bit_for_bit(tmp_path):
    """RestartOutputCycle and --restart: a run of 2 cycles writes the restart
    file of its final state; --restart with ncycles = 4 continues it, and the
    report lines, diagnostics rows and final restart file equal those of a
    straight 4-cycle run (which writes a restart file before cycle 2, removed
    once the final one is complete)."""
    base = open(os.path.join(DECKS, STARTUP["deck"])).read()

    def deck(directory, ncycles):
        text = base + ("SaveDirName = {0}\nRestartDirName = {0}\nDiagnosticsOutputCycle = 1\n"
                       "RestartOutputCycle = 2\nncycles = {1}\n").format(directory, ncycles)
        path = tmp_path / ("deck_%s_%d.inp" % (directory.name, ncycles))
        path.write_text(text)
        return str(path)

    def go(path, *args):
        say("tpic_mi_run --run", path, *args)
        r = subprocess.run([RUN, "--run", path, *args], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (r.stdout, r.stderr)
        return r.stdout

    straight, parts = tmp_path / "straight", tmp_path / "parts"
    whole = go(deck(straight, 4))
    assert sorted(p.name for p in straight.glob("checkpoint_*")) == ["checkpoint_t4_n1.tpicmi"]
    first = go(deck(parts, 2))
    assert sorted(p.name for p in parts.glob("checkpoint_*")) == ["checkpoint_t2_n1.tpicmi"]
    second = go(deck(parts, 4), "--restart")
    assert "Restart from " in second and "(cycle 2)" in second
    assert sorted(p.name for p in parts.glob("checkpoint_*")) == ["checkpoint_t4_n1.tpicmi"]
    e_whole, g_whole = reported(whole)
    e_first, g_first = reported(first)
    e_second, g_second = reported(second)
    assert sorted(e_second) == [2, 3, 4]
    for cycle in (2, 3, 4):
        assert e_second[cycle] == e_whole[cycle], cycle
    assert e_first[0] == e_whole[0] and e_first[1] == e_whole[1]
    assert g_first + g_second == g_whole
    # dE/E0 keeps the first state's energy across the restart
    change = re.findall(r"dE/E0 = (\S+)", whole)
    assert re.findall(r"dE/E0 = (\S+)", second) == change[-3:]
    rows = lambda d: [l for l in (d / "diagnostics.csv").read_text().splitlines() if l and l[0].isdigit()]
    assert rows(parts) == rows(straight)
    assert [r.split(",")[0] for r in rows(parts)] == ["0", "1", "2", "3", "4"]
    assert ((parts / "checkpoint_t4_n1.tpicmi").read_bytes() == (straight / "checkpoint_t4_n1.tpicmi").read_bytes())


def test_load_deck_matches_the_driver():
    """tpic_mi_load_deck starts the same state as `tpic_mi_run --run`: the
    energies of every cycle of a context started from the deck and advanced
    with tpic_mi_cycle equal the driver's report lines (printed with 13
    significant digits), for a Uniform deck with three species and for the
    GEM Harris deck."""
    from tpic_amd import runtime as R
    for d in (STARTUP, GEM_HARRIS):
        path = os.path.join(DECKS, d["deck"])
        energies, _ = reported(run_deck(d["deck"]))
        say("tpic_mi_load_deck", d["deck"])
        sim, run = R.Simulation.from_deck(path)
        try:
            assert (run.ncycles, run.species, tuple(run.cells)) == (d["ncycles"], len(d["species"]), d["cells"])
            for cycle in range(run.ncycles + 1):
                sc = sim.scalars()
                kinetic = sum(sc[2::5])
                got = (sc[0], sc[1], sc[0] + sc[1] + kinetic, kinetic)
                if cycle in energies:
                    for a, b in zip(got, energies[cycle]):
                        assert float("%.12e" % a) == b, (d["deck"], cycle, got, energies[cycle])
                if cycle < run.ncycles:
                    sim.cycle(1)
            assert sim.validation() == 0
        finally:
            sim.close()
