"""GPU parity of energy-conserving smoothing (ECSimSmoothing=energy_conserving):
tpic_mi_smooth, the smoothed operator and smoothed cycles through the C-ABI
against the CPU oracle, whose smoothing is pinned in test_oracle (plane-wave
symbols, the mirror-doubled walled box, dual-volume self-adjointness, energy
conservation); and the input projection of the walled operator that the
smoothed operator relies on. The HIP smoothing may contract products into
FMAs, so values agree to the tolerances stated in each test rather than bit
for bit."""
import numpy as np
import pytest

import oracle as O
from tpic_amd import runtime as R
from test_gpu_parity import gpu, periodic_duplicates, say, walled_case

pytestmark = pytest.mark.gpu

SMOOTHINGS = [(R.SMOOTH_FACE7, 3, 0.5), (R.SMOOTH_BINOMIAL27, 2, 0.0)]
SMOOTHING_IDS = ["face7", "binomial27"]
FACES = {"periodic": (None, None), "PEC-y": (1, R.FACE_PEC), "PMC-x": (0, R.FACE_PMC)}


def smoothing_case(name):
    """A walled_case (5 x 4 x 3 cells, two species, 6 per cell); "periodic"
    reuses its particles and fields in the periodic box."""
    axis, kind = FACES[name]
    if axis is None:
        cells, lengths, _, _, species, b, e = walled_case(0, R.FACE_PEC)
        periodic = (1, 1, 1)
        faces = [(R.FACE_PERIODIC, R.FACE_PERIODIC)] * 3
        e = periodic_duplicates(cells, e, None)
    else:
        cells, lengths, periodic, faces, species, b, e = walled_case(axis, kind)
    return cells, lengths, periodic, faces, species, b, e, axis


@pytest.mark.parametrize("smoothing", SMOOTHINGS, ids=SMOOTHING_IDS)
@pytest.mark.parametrize("faces", list(FACES))
def test_smooth_matches_the_oracle(faces, smoothing):
    """tpic_mi_smooth of a random node vector (constrained rows not zero)
    against the oracle's S = P (pass P)^n: within 1e-14 of its size; in
    place gives the same bits as out of place."""
    cells, lengths, periodic, face, species, _, _, axis = smoothing_case(faces)
    rng = np.random.default_rng(5 + smoothing[0])
    nx, ny, nz = cells
    v = periodic_duplicates(cells, rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)), axis)
    g = O.make_grid(cells, lengths, periodic, face)
    want = O.smooth(g, smoothing, v)
    say("smooth", faces, smoothing)
    with gpu(cells, lengths, [(q, len(a["q"])) for q, a in species], particle_periodic=periodic,
             field_faces=face, smoothing=smoothing) as sim:
        sim.set_field(R.FIELD_E, v)
        sim.smooth(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
        out = sim.get_field(R.FIELD_RHS)
        assert np.abs(out - want).max() <= 1e-14 * np.abs(want).max()
        sim.smooth(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_E))
        assert np.array_equal(sim.get_field(R.FIELD_E), out)
        assert sim.validation() == 0


@pytest.mark.parametrize("smoothing", SMOOTHINGS, ids=SMOOTHING_IDS)
@pytest.mark.parametrize("faces", list(FACES))
def test_smoothed_operator_matches_the_oracle(faces, smoothing):
    """tpic_mi_apply_operator with smoothing (x = S(v), y = S((1/V) f_g M x),
    image = v + d^2 curl curl v + k y) on a random vector whose constrained
    rows are not zero (the entry projects it) against the oracle's smoothed
    operator with the oracle's mass matrix: within 1e-12 of its size."""
    cells, lengths, periodic, face, species, b, _, axis = smoothing_case(faces)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    rng = np.random.default_rng(17)
    nx, ny, nz = cells
    v = periodic_duplicates(cells, rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)), axis)
    g = O.make_grid(cells, lengths, periodic, face)
    say("smoothed operator", faces, smoothing)
    with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=face,
             smoothing=smoothing) as sim:
        ref = O.Sim(cells, lengths, qom, counts, 0.1, particle_periodic=periodic, field_faces=face,
                    smoothing=smoothing)
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
            ref.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        sim.set_field(R.FIELD_E, v)
        sim.apply_operator(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
        out = sim.get_field(R.FIELD_RHS)
        th, dt = 0.5, 0.1
        vol = g.spacing[0] * g.spacing[1] * g.spacing[2]
        want = O.maxwell_apply_smoothed(g, smoothing, ref.field("mass").copy(), th * dt, 4 * np.pi * th * dt,
                                        1 / vol, v)
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()
        ref.close()


@pytest.mark.parametrize("smoothing", SMOOTHINGS, ids=SMOOTHING_IDS)
@pytest.mark.parametrize("faces", list(FACES))
def test_smoothed_cycles_against_the_oracle(faces, smoothing):
    """Three smoothed cycles against the oracle: the same GMRES iteration
    counts, energies within 1e-12 of the total (and the GPU's total constant
    to 1e-11, the energy-conserving property), E, B and the movers' field
    S(E_theta) within 1e-10 of their size, keys bit for bit."""
    cells, lengths, periodic, face, species, b, e, _ = smoothing_case(faces)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    say("smoothed cycles", faces, smoothing)
    with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=face,
             gmres_tolerance=1e-12, smoothing=smoothing) as sim:
        ref = O.Sim(cells, lengths, qom, counts, 0.1, tol=1e-12, particle_periodic=periodic, field_faces=face,
                    smoothing=smoothing)
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
            ref.set_particles(s, a)
        sim.set_field(R.FIELD_E, e)
        sim.set_field(R.FIELD_B, b)
        ref.field("E")[:] = e
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        e0 = sim.energy().sum()
        for _ in range(3):
            rg = sim.cycle(1)
            rc, ro = ref.cycle(1)
            assert rc == 0 and rg["iterations"] == ro.iterations
            eg, eo = sim.energy(), ref.energy()
            assert np.all(np.abs(eg - eo) <= 1e-12 * eo.sum())
            assert abs(eg.sum() - e0) <= 1e-11 * e0
            for w, name in ((R.FIELD_E, "E"), (R.FIELD_B, "B"), (R.FIELD_EPARTICLE, "particle")):
                want = ref.field(name)
                assert np.abs(sim.get_field(w) - want).max() <= 1e-10 * np.abs(want).max(), name
        for s in range(len(species)):
            got, want = sim.get_particles(s), ref.particles(s)
            og, ow = np.argsort(got["id"]), np.argsort(want["id"])
            assert np.array_equal(got["cell"][og], want["cell"][ow])
        assert sim.validation() == 0
        ref.close()


def test_smoothing_settings_are_checked():
    """Creation rejects an unknown operator, SmoothNiter < 0 and Smooth
    outside [0, 1] (physics.cpp:65-71); Smooth = 1 or SmoothNiter = 0 leaves
    smoothing inactive, so tpic_mi_smooth is rejected."""
    cells, lengths = (4, 4, 4), (1.0, 1.0, 1.0)
    for bad in ((2, 1, 0.5), (R.SMOOTH_FACE7, -1, 0.5), (R.SMOOTH_FACE7, 1, 1.5), (R.SMOOTH_FACE7, 1, -0.1)):
        with pytest.raises(R.TpicError):
            gpu(cells, lengths, [(-1.0, 8)], smoothing=bad)
    for inactive in ((R.SMOOTH_FACE7, 4, 1.0), (R.SMOOTH_BINOMIAL27, 0, 0.0)):
        with gpu(cells, lengths, [(-1.0, 8)], smoothing=inactive) as sim:
            with pytest.raises(R.TpicError):
                sim.smooth(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))


@pytest.mark.parametrize("faces", ["PEC-y", "PMC-x"])
def test_walled_operator_projects_its_input(faces):
    """begin_apply_input (maxwell_device.hpp:270-272): on a walled box
    tpic_mi_apply_operator takes its input with the constrained rows zeroed,
    so an input with nonzero constrained rows has the image of its
    projection, bit for bit, and zero image rows there."""
    cells, lengths, periodic, face, species, b, _, axis = smoothing_case(faces)
    counts = [len(a["q"]) for _, a in species]
    rng = np.random.default_rng(23)
    nx, ny, nz = cells
    v = periodic_duplicates(cells, rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)), axis)
    g = O.make_grid(cells, lengths, periodic, face)
    projected = O.smooth(g, (O.SMOOTH_FACE7, 0, 0.0), v)  # zero passes: the projection alone
    assert not np.array_equal(projected, v)
    say("walled operator input projection", faces)
    with gpu(cells, lengths, [(q, n) for (q, _), n in zip(species, counts)], particle_periodic=periodic,
             field_faces=face) as sim:
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        images = []
        for vec in (v, projected):
            sim.set_field(R.FIELD_E, vec)
            sim.apply_operator(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
            images.append(sim.get_field(R.FIELD_RHS))
        assert np.array_equal(images[0], images[1])
        constrained = (projected == 0.0) & (v != 0.0)
        assert constrained.any() and np.all(images[0][constrained] == 0.0)


@pytest.mark.parametrize("faces", ["periodic", "PEC-y"])
def test_smoothed_cycles_with_the_fp32_matrix(faces):
    """The smoothed operator with the mixed (FP32-stored) mass matrix: three
    cycles track the FP64 smoothed run (energies within 1e-6 of the total, E
    within 1e-4 of its size, GMRES counts within one) and keep the energy
    constant to 1e-6 per cycle."""
    cells, lengths, periodic, face, species, b, e, _ = smoothing_case(faces)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    smoothing = (R.SMOOTH_FACE7, 3, 0.5)
    say("smoothed cycles fp32 matrix", faces)
    runs = []
    for kernel in (R.MASS_FP64, R.MASS_FP32):
        with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=face,
                 gmres_tolerance=1e-12, smoothing=smoothing, mass_kernel=kernel) as sim:
            for s, (_, a) in enumerate(species):
                sim.set_particles(s, a)
            sim.set_field(R.FIELD_E, e)
            sim.set_field(R.FIELD_B, b)
            sim.initialize()
            energies, iterations = [sim.energy()], []
            for _ in range(3):
                iterations.append(sim.cycle(1)["iterations"])
                energies.append(sim.energy())
            runs.append((energies, iterations, sim.get_field(R.FIELD_E), sim.validation()))
    (e64, it64, f64, v64), (e32, it32, f32, v32) = runs
    assert v64 == 0 and v32 == 0
    assert all(abs(a - b) <= 1 for a, b in zip(it64, it32))
    total = e64[0].sum()
    for a, b in zip(e64, e32):
        assert np.all(np.abs(a - b) <= 1e-6 * total)
    for n in range(3):
        assert abs(e32[n + 1].sum() - e32[n].sum()) <= 1e-6 * total
    assert np.abs(f32 - f64).max() <= 1e-4 * np.abs(f64).max()
