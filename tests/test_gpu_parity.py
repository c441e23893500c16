"""GPU parity: the HIP path through the C-ABI (libtpic_mi.so) against the CPU
oracle on the same seeded inputs. Integer and index results (sort
permutation, counts, offsets, keys) and the contraction-free mover are
bit-exact; floating-point moments and fields are within the tolerances
stated in each test (FP64)."""
import json
import os

import numpy as np
import pytest

import oracle as O
import tpic_amd
from tpic_amd import cases
from tpic_amd import runtime as R

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "reference_kats.json")))


def say(*args):
    print(*args, flush=True)


def gpu(cells, lengths, species, **kw):
    kw.setdefault("dt", 0.1)
    return tpic_amd.Simulation(cells, lengths, species, **kw)


def to_oracle(arrays):
    return O.ParticleArrays(source=arrays)


def periodic_nodes(cells, rng, scale=1.0):
    nx, ny, nz = cells
    v = rng.normal(size=(3, nz + 1, ny + 1, nx + 1)) * scale
    v[:, :, :, -1] = v[:, :, :, 0]
    v[:, :, -1] = v[:, :, 0]
    v[:, -1] = v[:, 0]
    return v.ravel().copy()


def pattern_particles(keys):
    n = len(keys)
    i = np.arange(n, dtype=np.float64)
    f = KATS["sort_pattern"]["factors"]
    out = {k: i * f[k] for k in ("rx", "ry", "rz", "u", "v", "w", "q")}
    out["cell"] = np.asarray(keys, np.int32)
    out["id"] = (np.arange(n) + 100).astype(np.uint64)
    return out


# ---- sort ------------------------------------------------------------------

SORT_CASES = ["pattern", "random", "near_sorted", "many_runs", "tiny", "one"]


@pytest.mark.parametrize("case", SORT_CASES)
def test_sort_bit_exact_against_stable_oracle(case):
    cells = (6, 5, 4)
    g = O.make_grid(cells, cells)
    bins = O.bin_count(g)
    rng = np.random.default_rng(11)
    if case == "pattern":
        keys = (np.arange(20000) * 7 + 3) % bins
    elif case == "random":
        keys = rng.integers(0, bins, size=50001)
    elif case == "near_sorted":
        keys = np.sort(rng.integers(0, bins, size=70000))
        flip = rng.random(keys.size) < 0.04
        keys[flip] = np.clip(keys[flip] + rng.integers(-30, 31, size=flip.sum()), 0, bins - 1)
    elif case == "many_runs":  # few bins, >24 runs per bin: the large-bin path
        keys = rng.integers(500, 506, size=30000)
    elif case == "tiny":
        keys = rng.integers(0, bins, size=37)
    else:
        keys = np.array([bins - 1])
    say("sort case", case, "n", len(keys))
    arrays = pattern_particles(keys)
    with gpu(cells, cells, [(-1.0, len(keys))]) as sim:
        sim.set_particles(0, arrays)
        sim.record_permutation(True)
        sim.sort(0)
        got = sim.get_particles(0)
        counts, offsets, perm = sim.get_bins(0, permutation=True)
        flags = sim.validation()
    out, operm, ocounts, ooffsets = O.sort(g, to_oracle(arrays))
    for k in got:
        assert np.array_equal(got[k], out.a[k]), k
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(offsets, ooffsets)
    assert np.array_equal(perm, operm)
    halo_used = any(not g_interior(g, k) for k in np.unique(keys))
    assert bool(flags & R.INVALID_HALO) == halo_used
    assert not flags & R.INVALID_SORT


def g_interior(g, key):
    ex, ey = g.n[0] + 4, g.n[1] + 4
    i, j, k = key % ex - 2, (key // ex) % ey - 2, key // (ex * ey) - 2
    return 0 <= i < g.n[0] and 0 <= j < g.n[1] and 0 <= k < g.n[2]


def test_sort_rejects_an_invalid_key_and_keeps_particles():
    cells = (3, 3, 3)
    g = O.make_grid(cells, cells)
    keys = np.array([5, 17, O.bin_count(g) + 3, 2], np.int32)
    arrays = pattern_particles(keys)
    with gpu(cells, cells, [(1.0, len(keys))]) as sim:
        sim.set_particles(0, arrays)
        sim.sort(0)
        got = sim.get_particles(0)
        assert sim.validation() & R.INVALID_SORT
    for k in got:
        assert np.array_equal(got[k], arrays[k]), k


# ---- mover -----------------------------------------------------------------

def random_state(cells, lengths, per_cell, rng, vscale=0.3):
    g = O.make_grid(cells, lengths)
    ncell = O.cell_count(g)
    n = ncell * per_cell
    c = np.repeat(np.arange(ncell), per_cell)
    i, j, k = c % cells[0], (c // cells[0]) % cells[1], c // (cells[0] * cells[1])
    a = {name: rng.random(n) for name in ("rx", "ry", "rz")}
    for name in ("u", "v", "w"):
        a[name] = rng.normal(size=n) * vscale
    a["q"] = rng.choice([-1.0, 0.5, 2.0], size=n)
    a["cell"] = O.encode(g, i, j, k).astype(np.int32)
    a["id"] = np.arange(n, dtype=np.uint64)
    return g, a


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1), (0, 0, 0)])
def test_mover_bit_exact_against_oracle(periodic):
    cells, lengths = (7, 5, 6), (1.4, 1.0, 1.5)
    rng = np.random.default_rng(12)
    g, a = random_state(cells, lengths, 9, rng, vscale=2.5)
    g.particle_periodic[0], g.particle_periodic[1], g.particle_periodic[2] = periodic
    # a few sub-ulp crossings and long flights
    a["rx"][:5] = 0.0
    a["u"][:5] = -1e-30
    a["v"][5:10] = 40.0
    e = periodic_nodes(cells, rng, 0.3)
    b = periodic_nodes(cells, rng, 0.8)
    qom, dt = -3.0, 0.17
    say("mover periodic", periodic, "n", len(a["cell"]))
    with gpu(cells, lengths, [(qom, len(a["cell"]))], dt=dt, particle_periodic=periodic) as sim:
        sim.set_particles(0, a)
        sim.set_field(R.FIELD_ETHETA, e)
        sim.set_field(R.FIELD_BNODE, b)
        sim.move(0)
        got = sim.get_particles(0)
        flags = sim.validation()
    p = to_oracle(a)
    oflags = O.move(g, O.mover_params(qom, dt, 1.0, g.spacing), p, e, b)
    assert flags == oflags == 0
    for k in got:
        assert np.array_equal(got[k].view(np.uint8), p.a[k].view(np.uint8)), k


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0)])
def test_sorted_mover_bit_exact_against_oracle(periodic):
    """The per-cell mover (bins of the last sort) against oracle sort + move."""
    cells, lengths = (9, 4, 5), (2.25, 1.0, 1.25)
    rng = np.random.default_rng(17)
    g, a = random_state(cells, lengths, 70, rng, vscale=1.5)  # > 64 per cell: two chunks
    g.particle_periodic[0], g.particle_periodic[1], g.particle_periodic[2] = periodic
    order = rng.permutation(len(a["cell"]))
    a = {k: v[order] for k, v in a.items()}
    e = periodic_nodes(cells, rng, 0.2)
    b = periodic_nodes(cells, rng, 0.6)
    qom, dt = 2.5, 0.11
    with gpu(cells, lengths, [(qom, len(a["cell"]))], dt=dt, particle_periodic=periodic) as sim:
        sim.set_particles(0, a)
        sim.sort(0)
        sim.set_field(R.FIELD_ETHETA, e)
        sim.set_field(R.FIELD_BNODE, b)
        sim.move(0)
        got = sim.get_particles(0)
        flags = sim.validation()
    sp, _, _, _ = O.sort(g, to_oracle(a))
    oflags = O.move(g, O.mover_params(qom, dt, 1.0, g.spacing), sp, e, b)
    assert flags == oflags == 0
    for k in got:
        assert np.array_equal(got[k].view(np.uint8), sp.a[k].view(np.uint8)), k


def test_sorted_mover_skips_halo_bins_and_a_failed_sort():
    """The sorted mover pushes the records of interior bins only: records
    sorted into halo bins stay as they are and raise no flag (the sort has
    flagged the halo), as in a walk over the interior bins; after a failed
    sort (every bin range empty) no record moves."""
    cells, lengths = (7, 3, 4), (1.75, 0.75, 1.0)
    rng = np.random.default_rng(29)
    g, a = random_state(cells, lengths, 40, rng, vscale=1.2)
    n = len(a["cell"])
    halo = [O.encode(g, -1, 1, 2), O.encode(g, 7, 0, 0), O.encode(g, 3, -2, 1), O.encode(g, 0, 0, 4)]
    for t, key in enumerate(halo):
        a["cell"][5 + 97 * t] = key
    order = rng.permutation(n)
    a = {k: v[order] for k, v in a.items()}
    e = periodic_nodes(cells, rng, 0.2)
    b = periodic_nodes(cells, rng, 0.6)
    qom, dt = 1.5, 0.1
    with gpu(cells, lengths, [(qom, n)], dt=dt) as sim:
        sim.set_particles(0, a)
        sim.sort(0)
        assert sim.validation() & R.INVALID_HALO
        sim.clear_validation()
        sim.set_field(R.FIELD_ETHETA, e)
        sim.set_field(R.FIELD_BNODE, b)
        sim.move(0)
        got = sim.get_particles(0)
        flags = sim.validation()
    sp, _, _, _ = O.sort(g, to_oracle(a))
    before = {k: v.copy() for k, v in sp.a.items()}
    oflags = O.move(g, O.mover_params(qom, dt, 1.0, g.spacing), sp, e, b)
    # The oracle's mover rejects the halo records with the key flag; the
    # sorted mover never visits them.
    assert oflags == R.INVALID_KEY and flags == 0
    in_halo = np.isin(before["cell"], np.array(halo, np.int32))
    assert in_halo.sum() == len(halo)
    for k in got:
        assert np.array_equal(got[k].view(np.uint8), sp.a[k].view(np.uint8)), k
        assert np.array_equal(got[k][in_halo], before[k][in_halo]), k
    # A failed sort leaves the records in input order and every bin empty.
    bad = dict(a)
    bad["cell"] = a["cell"].copy()
    bad["cell"][11] = O.bin_count(g) + 5
    with gpu(cells, lengths, [(qom, n)], dt=dt) as sim:
        sim.set_particles(0, bad)
        sim.sort(0)
        assert sim.validation() & R.INVALID_SORT
        sim.clear_validation()
        sim.set_field(R.FIELD_ETHETA, e)
        sim.set_field(R.FIELD_BNODE, b)
        sim.move(0)
        got = sim.get_particles(0)
        assert sim.validation() == 0
    for k in got:
        assert np.array_equal(got[k], bad[k]), k


@pytest.mark.parametrize("periodic", [(1, 1, 1), (0, 1, 0)])
def test_move_then_sort_bit_exact_against_oracle(periodic):
    """The cycle's particle half (per-cell mover, then the sort) against the
    oracle's sort + move + sort: same records, bins and permutation; a
    rejected record travels unchanged."""
    cells, lengths = (9, 4, 5), (2.25, 1.0, 1.25)
    rng = np.random.default_rng(23)
    g, a = random_state(cells, lengths, 70, rng, vscale=1.5)
    g.particle_periodic[0], g.particle_periodic[1], g.particle_periodic[2] = periodic
    order = rng.permutation(len(a["cell"]))
    a = {k: v[order] for k, v in a.items()}
    a["rx"][3] = 1.5  # fraction out of range: rejected by the mover
    e = periodic_nodes(cells, rng, 0.2)
    b = periodic_nodes(cells, rng, 0.6)
    qom, dt = 2.5, 0.11
    n = len(a["cell"])
    with gpu(cells, lengths, [(qom, n)], dt=dt, particle_periodic=periodic) as sim:
        sim.set_particles(0, a)
        sim.sort(0)
        sim.set_field(R.FIELD_ETHETA, e)
        sim.set_field(R.FIELD_BNODE, b)
        sim.record_permutation(True)
        sim.move(0)
        sim.sort(0)
        got = sim.get_particles(0)
        counts, offsets, perm = sim.get_bins(0, permutation=True)
        flags = sim.validation()
    sp, _, _, _ = O.sort(g, to_oracle(a))
    oflags = O.move(g, O.mover_params(qom, dt, 1.0, g.spacing), sp, e, b)
    sp2, operm, ocounts, ooffsets = O.sort(g, sp)
    assert flags == oflags == R.INVALID_FRACTION
    assert np.array_equal(counts, ocounts) and np.array_equal(offsets, ooffsets)
    assert np.array_equal(perm, operm)
    for k in got:
        assert np.array_equal(got[k].view(np.uint8), sp2.a[k].view(np.uint8)), k


@pytest.mark.parametrize("name", ["mover_long_flight", "mover_sub_ulp_crossing"])
def test_mover_known_answers_on_gpu(name):
    k = KATS[name]
    cells = tuple(k["cells"])
    g = O.make_grid(cells, cells, k["particle_periodic"])
    prm = k["params"]
    # beta = 0 through q/m = 0 is not admitted; a vanishing field gives the same push.
    a = {"rx": np.array([k["input"]["r"][0]]), "ry": np.array([k["input"]["r"][1]]),
         "rz": np.array([k["input"]["r"][2]]), "u": np.array([k["input"]["v"][0]]),
         "v": np.array([k["input"]["v"][1]]), "w": np.array([k["input"]["v"][2]]), "q": np.array([1.0]),
         "cell": np.array([O.encode(g, *k["input"]["cell"])], np.int32), "id": np.array([1], np.uint64)}
    with gpu(cells, cells, [(1.0, 1)], dt=prm["dt"], particle_periodic=k["particle_periodic"]) as sim:
        sim.set_particles(0, a)
        sim.move(0)
        got = sim.get_particles(0)
        assert sim.validation() == 0
    ex, ey = cells[0] + 4, cells[1] + 4
    key = int(got["cell"][0])
    assert [key % ex - 2, (key // ex) % ey - 2, key // (ex * ey) - 2] == k["expect"]["cell"]
    assert [got["rx"][0], got["ry"][0], got["rz"][0]] == k["expect"]["r"]
    assert [got["u"][0], got["v"][0], got["w"][0]] == k["expect"]["v"]


def test_mover_flags_invalid_records_and_keeps_them():
    cells = (4, 4, 4)
    rng = np.random.default_rng(13)
    g, a = random_state(cells, cells, 2, rng)
    a["u"][3] = np.nan
    a["rx"][7] = 1.0
    a["q"][9] = np.inf
    with gpu(cells, cells, [(1.0, len(a["cell"]))]) as sim:
        sim.set_particles(0, a)
        sim.move(0)
        got = sim.get_particles(0)
        flags = sim.validation()
    assert flags & R.INVALID_NONFINITE and flags & R.INVALID_FRACTION
    for idx in (3, 7, 9):
        for k in got:
            assert np.array_equal(got[k][idx:idx + 1], a[k][idx:idx + 1], equal_nan=True), (idx, k)


# ---- deposition and assembly ----------------------------------------------

def deposit_both(cells, lengths, species, rng, per_cell=11, **kw):
    """Deposit J and M for random sorted species on both sides."""
    g = O.make_grid(cells, lengths)
    arrays = []
    for s, qom in enumerate(species):
        _, a = random_state(cells, lengths, per_cell + s, rng)
        arrays.append(a)
    bnode = periodic_nodes(cells, rng, 0.5)
    dt = 0.13
    sim = gpu(cells, lengths, [(q, len(a["cell"])) for q, a in zip(species, arrays)], dt=dt, **kw)
    for s, a in enumerate(arrays):
        sim.set_particles(s, a)
        sim.sort(s)
    sim.set_field(R.FIELD_BNODE, bnode)
    sim.deposit()
    sim.assemble()
    vol = float(np.float64(np.longdouble(g.spacing[0]) * np.longdouble(g.spacing[1]) * np.longdouble(g.spacing[2])))
    cur = np.zeros(3 * O.node_count(g))
    mass = np.zeros(O.vertex_count(g) * 126)
    for s, a in enumerate(arrays):
        sp, _, counts, offsets = O.sort(g, to_oracle(a))
        boc = species[s] * dt / 2
        O.deposit_current(g, sp, counts, offsets, bnode, 1.0 / vol, boc, cur)
        O.deposit_mass(g, sp, counts, offsets, bnode, boc, mass)
    O.fold_current(g, cur)
    O.fold_mass(g, mass)
    return sim, g, cur, mass


def test_deposition_and_assembly_parity():
    rng = np.random.default_rng(14)
    cells, lengths = (6, 4, 5), (1.5, 1.0, 1.25)
    sim, g, cur, mass = deposit_both(cells, lengths, [-8.0, 1.0], rng)
    with sim:
        j = sim.get_field(R.FIELD_J)
        m = sim.get_field(R.FIELD_MASS)
        assert sim.validation() == 0
    # tests/ecsim/sorted_mass_device.cpp:359-367: 3e-12 max|K| + 1e-15 (FP64)
    assert np.abs(m - mass).max() <= 3e-12 * np.abs(mass).max() + 1e-15
    assert np.abs(j - cur).max() <= 3e-12 * np.abs(cur).max() + 1e-15


# Mixed-precision moment products: FP32 operands (rounding 2^-24 per operand,
# FP32 accumulation over ~25 particles per cell) stay near 1e-7 of the largest
# entry; BF16 hi/lo split products drop lo*lo and round lo (about 2^-17 per
# product), so near 1e-5.
MIXED_KERNELS = [("fp32", R.MASS_FP32, 2e-6), ("bf16x3", R.MASS_BF16X3, 5e-5)]


@pytest.mark.parametrize("name,kernel,tol", MIXED_KERNELS)
def test_mixed_moment_products_within_tolerance(name, kernel, tol):
    rng = np.random.default_rng(14)
    cells, lengths = (6, 4, 5), (1.5, 1.0, 1.25)
    sim, g, cur, mass = deposit_both(cells, lengths, [-8.0, 1.0], rng, mass_kernel=kernel)
    with sim:
        j = sim.get_field(R.FIELD_J)
        m = sim.get_field(R.FIELD_MASS)
        assert sim.validation() == 0
    err_m = np.abs(m - mass).max() / np.abs(mass).max()
    err_j = np.abs(j - cur).max() / np.abs(cur).max()
    say(name, "moments: mass rel err", err_m, "current rel err", err_j)
    assert err_m <= tol
    assert err_j <= tol
    # not the FP64 path: the rounding is visible
    assert err_m > 1e-12


# Per-cell populations: the reference's sorted-deposition edge cases
# (tests/golden/reference_kats.json, from tests/ecsim/sorted_current_mover_device.cpp:48:
# 0, 1, 3, 31, 32, 33, 129, 257) with the full 64-particle chunk and its
# neighbours (63, 64, 65), so that a partial first chunk, a full unrolled
# chunk, several chunks per cell and, for BF16 hi/lo, several 32-particle
# blocks per cell all run.
POPULATIONS = sorted(set(KATS["mover_parameters"]["populations"]) | {63, 64, 65})
# Bounds against the FP64 oracle, relative to the largest entry: FP64 the
# reference's 3e-12 (sorted_mass_device.cpp:359-367); FP32 operands and FP32
# cell sums over up to 257 products of mixed-sign charges 1e-5; BF16 hi/lo
# (lo*lo dropped, about 2^-16 per product) 1e-4.
POPULATION_KERNELS = [("fp64", R.MASS_FP64, 3e-12), ("fp32", R.MASS_FP32, 1e-5), ("bf16x3", R.MASS_BF16X3, 1e-4)]


def ragged_state(cells, lengths, populations, rng, shift):
    """Random particles with cell c holding populations[(c + shift) % len]."""
    g = O.make_grid(cells, lengths)
    ncell = O.cell_count(g)
    per = np.array([populations[(c + shift) % len(populations)] for c in range(ncell)])
    c = np.repeat(np.arange(ncell), per)
    n = len(c)
    i, j, k = c % cells[0], (c // cells[0]) % cells[1], c // (cells[0] * cells[1])
    a = {name: rng.random(n) for name in ("rx", "ry", "rz")}
    for name in ("u", "v", "w"):
        a[name] = rng.normal(size=n) * 0.3
    a["q"] = rng.choice([-1.0, 0.5, 2.0], size=n)
    a["cell"] = O.encode(g, i, j, k).astype(np.int32)
    a["id"] = np.arange(n, dtype=np.uint64)
    order = rng.permutation(n)  # unsorted input: the sort groups the cells
    return {name: v[order] for name, v in a.items()}


@pytest.mark.parametrize("name,kernel,tol", POPULATION_KERNELS)
def test_deposition_at_the_reference_populations(name, kernel, tol):
    """J and M of two species whose cells hold every population of
    POPULATIONS (each species rotated differently) against the FP64 oracle,
    for each moment-operand precision."""
    rng = np.random.default_rng(31)
    cells, lengths = (11, 3, 2), (2.75, 0.75, 0.5)
    species = [-8.0, 1.0]
    arrays = [ragged_state(cells, lengths, POPULATIONS, rng, shift) for shift in (0, 5)]
    g = O.make_grid(cells, lengths)
    bnode = periodic_nodes(cells, rng, 0.5)
    dt = 0.13
    say("populations", POPULATIONS, "kernel", name, "particles", [len(a["q"]) for a in arrays])
    with gpu(cells, lengths, [(q, len(a["cell"])) for q, a in zip(species, arrays)], dt=dt,
             mass_kernel=kernel) as sim:
        for s, a in enumerate(arrays):
            sim.set_particles(s, a)
            sim.sort(s)
        sim.set_field(R.FIELD_BNODE, bnode)
        sim.deposit()
        sim.assemble()
        j = sim.get_field(R.FIELD_J)
        m = sim.get_field(R.FIELD_MASS)
        assert sim.validation() == 0
    vol = float(np.float64(np.longdouble(g.spacing[0]) * np.longdouble(g.spacing[1]) * np.longdouble(g.spacing[2])))
    cur = np.zeros(3 * O.node_count(g))
    mass = np.zeros(O.vertex_count(g) * 126)
    for s, a in enumerate(arrays):
        sp, _, counts, offsets = O.sort(g, to_oracle(a))
        boc = species[s] * dt / 2
        O.deposit_current(g, sp, counts, offsets, bnode, 1.0 / vol, boc, cur)
        O.deposit_mass(g, sp, counts, offsets, bnode, boc, mass)
    O.fold_current(g, cur)
    O.fold_mass(g, mass)
    err_m = np.abs(m - mass).max() / np.abs(mass).max()
    err_j = np.abs(j - cur).max() / np.abs(cur).max()
    say(name, "populations: mass rel err", err_m, "current rel err", err_j)
    assert err_m <= tol and err_j <= tol


def test_deposition_thin_axes_and_empty_cells():
    rng = np.random.default_rng(15)
    cells, lengths = (9, 1, 1), (2.0, 0.2, 0.2)
    g = O.make_grid(cells, lengths)
    _, a = random_state(cells, lengths, 5, rng)
    keep = np.isin(a["cell"], [O.encode(g, i, 0, 0) for i in (0, 3, 4, 8)])
    a = {k: v[keep] for k, v in a.items()}
    bnode = periodic_nodes(cells, rng, 0.5)
    with gpu(cells, lengths, [(2.0, len(a["cell"]))], dt=0.2) as sim:
        sim.set_particles(0, a)
        sim.sort(0)
        sim.set_field(R.FIELD_BNODE, bnode)
        sim.deposit()
        sim.assemble()
        j = sim.get_field(R.FIELD_J)
        m = sim.get_field(R.FIELD_MASS)
    sp, _, counts, offsets = O.sort(g, to_oracle(a))
    vol = g.spacing[0] * g.spacing[1] * g.spacing[2]
    cur, _ = O.deposit_current(g, sp, counts, offsets, bnode, 1 / vol, 0.2)
    mass, _ = O.deposit_mass(g, sp, counts, offsets, bnode, 0.2)
    O.fold_current(g, cur)
    O.fold_mass(g, mass)
    assert np.abs(m - mass).max() <= 3e-12 * np.abs(mass).max() + 1e-15
    assert np.abs(j - cur).max() <= 3e-12 * np.abs(cur).max() + 1e-15


# ---- operator, solve and cycle ---------------------------------------------

def test_operator_parity():
    rng = np.random.default_rng(16)
    cells, lengths = (5, 6, 4), (1.0, 1.5, 1.0)
    sim, g, cur, mass = deposit_both(cells, lengths, [-8.0, 1.0], rng)
    v = periodic_nodes(cells, rng)
    with sim:
        sim.set_field(R.FIELD_E, v)
        sim.apply_operator(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
        out = sim.get_field(R.FIELD_RHS)
    th, dt = 0.5, 0.13
    vol = g.spacing[0] * g.spacing[1] * g.spacing[2]
    ref = O.maxwell_apply(g, mass, th * dt, 4 * np.pi * th * dt, 1 / vol, v)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def setup_plasma(cells, lengths, ppc, qom, thermal, b0, perturbation, seed_base=0):
    arrays = [cases.uniform_maxwellian(cells, lengths, ppc, qom[s], thermal=(thermal[s],) * 3,
                                       seed=cases.species_seed(s) + seed_base) for s in range(len(qom))]
    b = cases.uniform_b(cells, b0, amplitude=perturbation)
    return arrays, b


def run_both(cells, lengths, ppc, qom, thermal, b0, perturbation, dt, theta, tol, cycles):
    arrays, b = setup_plasma(cells, lengths, ppc, qom, thermal, b0, perturbation)
    counts = [len(a["cell"]) for a in arrays]
    sim = gpu(cells, lengths, list(zip(qom, counts)), dt=dt, theta=theta, gmres_tolerance=tol)
    for s, a in enumerate(arrays):
        sim.set_particles(s, a)
    sim.set_field(R.FIELD_B, b)
    sim.initialize()
    osim = O.Sim(cells, lengths, qom, counts, dt=dt, theta=theta, tol=tol)
    for s, a in enumerate(arrays):
        osim.set_particles(s, a)
    osim.field("B")[:] = b
    osim.initialize()
    history = []
    e0g, e0o = sim.energy(), osim.energy()
    history.append((e0g, e0o, None, None))
    for n in range(cycles):
        rg = sim.cycle(1)
        rc, ro = osim.cycle(1)
        assert rc == 0
        history.append((sim.energy(), osim.energy(), rg, ro))
        say("cycle", n, "gpu it", rg["iterations"], "res", rg["relative_residual"],
            "oracle it", ro.iterations, "res", ro.relative_residual)
    return sim, osim, history


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_solve_parity_after_bootstrap():
    cells, lengths = (8, 6, 5), (2.0, 1.5, 1.25)
    arrays, b = setup_plasma(cells, lengths, (2, 2, 2), [-16.0, 1.0], [0.05, 0.01], (0.01, 0.02, 0.05), 0.2)
    counts = [len(a["cell"]) for a in arrays]
    with gpu(cells, lengths, [(-16.0, counts[0]), (1.0, counts[1])], dt=0.1, gmres_tolerance=1e-12) as sim:
        for s, a in enumerate(arrays):
            sim.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        rc, res = sim.solve()
        eth = sim.get_field(R.FIELD_ETHETA)
        rhs = sim.get_field(R.FIELD_RHS)
    osim = O.Sim(cells, lengths, [-16.0, 1.0], counts, dt=0.1, tol=1e-12)
    for s, a in enumerate(arrays):
        osim.set_particles(s, a)
    osim.field("B")[:] = b
    osim.initialize()
    orc, ores = osim.solve()
    assert orc == 0 and res["converged"]
    assert res["iterations"] == ores.iterations
    assert rel(eth, osim.field("theta")) <= 1e-10


def test_cycle_parity_uniform_plasma_16():
    """The 16^3 uniform magnetised plasma (2 species, 64 ppc, dt 0.1, 3 cycles,
    GMRES 1e-10): fields within 1e-10 of the oracle, particles by id."""
    c = cases.UNIFORM_16
    sim, osim, hist = run_both(c["cells"], c["lengths"], c["ppc"], c["q_over_m"], c["thermal"], c["b0"], 1e-3,
                               c["dt"], c["theta"], c["gmres_tolerance"], c["cycles"])
    with sim:
        for which, name in ((R.FIELD_E, "E"), (R.FIELD_B, "B"), (R.FIELD_J, "J")):
            got = sim.get_field(which)
            assert rel(got, osim.field(name)) <= 1e-10, name
        for s in range(2):
            got = sim.get_particles(s)
            ref = osim.particles(s)
            assert np.array_equal(got["id"], ref["id"])  # same stable sort order
            assert np.array_equal(got["cell"], ref["cell"])
            for k in ("rx", "ry", "rz", "u", "v", "w"):
                assert np.abs(got[k] - ref[k]).max() <= 1e-12, k
        assert sim.validation() == 0
    for eg, eo, rg, ro in hist[1:]:
        assert rg["iterations"] == ro.iterations
        assert abs(eg.sum() - eo.sum()) <= 1e-12 * eo.sum()


def separate_cycle(sim, species):
    """One cycle through the separate entry points."""
    rc, res = sim.solve()
    for s in range(species):
        sim.move(s)
    sim.advance_fields()
    for s in range(species):
        sim.sort(s)
    sim.deposit(0, species)
    sim.assemble()
    return res


@pytest.mark.parametrize("walls,empty,dense", [(False, False, False), (True, False, False), (False, True, False),
                                               (False, False, True)],
                         ids=["periodic", "walls", "empty-species", "dense"])
@pytest.mark.parametrize("kernel", [R.MASS_FP64, R.MASS_FP32, R.MASS_BF16X3], ids=["fp64", "fp32", "bf16x3"])
def test_cycle_equals_the_separate_calls(kernel, walls, empty, dense):
    """tpic_mi_cycle fuses the sort's record copy into the deposition (and
    plans one species' sort on a second stream); after three cycles its state
    equals that of the separate calls (solve, move, advance, sort, deposit,
    assemble) bit for bit: E, B, B at nodes, J, the mass matrix, every
    species' bins and all nine particle attributes. Three species, the first
    with 80 particles in a cell (two chunks per cell), the last with a few;
    with walls, PEC faces in y and mirrored particles; a fourth species
    without particles between them; and a dense box whose cells hold 2048 and
    4608 particles of a species (runs longer than a wave's 1024-key span and
    a tile's 4096 keys)."""
    cells, lengths = (10, 7, 6), (2.5, 1.75, 1.5)
    qom = [-16.0, 1.0, 0.25]
    loads = [((4, 4, 5), 0.08), ((2, 2, 2), 0.02), ((1, 1, 3), 0.05)]
    if dense:
        cells, lengths = (4, 3, 2), (1.0, 0.75, 0.5)
        loads = [((16, 16, 8), 0.3), ((16, 16, 18), 0.05), ((2, 1, 1), 0.1)]
    arrays = [cases.uniform_maxwellian(cells, lengths, ppc, qom[s], thermal=(th,) * 3, seed=cases.species_seed(s))
              for s, (ppc, th) in enumerate(loads)]
    if empty:
        qom.insert(1, -2.0)
        arrays.insert(1, {k: np.zeros(0, v.dtype) for k, v in arrays[0].items()})
    b = cases.uniform_b(cells, (0.01, 0.02, 0.05), amplitude=0.2)
    counts = [len(a["cell"]) for a in arrays]
    kw = dict(dt=0.1, gmres_tolerance=1e-12, mass_kernel=kernel)
    if walls:
        kw.update(particle_periodic=(1, 0, 1),
                  field_faces=[(R.FACE_PERIODIC,) * 2, (R.FACE_PEC,) * 2, (R.FACE_PERIODIC,) * 2])
    say("fused against separate cycle, mass kernel", kernel, "walls", walls)
    fused = gpu(cells, lengths, list(zip(qom, counts)), **kw)
    separate = gpu(cells, lengths, list(zip(qom, counts)), **kw)
    try:
        for sim in (fused, separate):
            for s, a in enumerate(arrays):
                sim.set_particles(s, a)
            sim.set_field(R.FIELD_B, b)
            sim.initialize()
        for n in range(3):
            rf = fused.cycle(1)
            rs = separate_cycle(separate, len(qom))
            assert rf["iterations"] == rs["iterations"], n
        for which in (R.FIELD_E, R.FIELD_B, R.FIELD_BNODE, R.FIELD_J, R.FIELD_MASS):
            assert np.array_equal(fused.get_field(which), separate.get_field(which)), which
        for s in range(len(qom)):
            gf, gs = fused.get_particles(s), separate.get_particles(s)
            for k in R.ATTRIBUTES:
                assert np.array_equal(gf[k], gs[k]), (s, k)
            cf, of = fused.get_bins(s)
            cs, os_ = separate.get_bins(s)
            assert np.array_equal(cf, cs) and np.array_equal(of, os_), s
        assert fused.validation() == 0 and separate.validation() == 0
    finally:
        fused.close()
        separate.close()


def converging_slab(cells, per_cell, speed, rng):
    """Particles of the x column 2 of every (y, z) row moving +x by `speed`
    cells per step (fractions random), plus a sparse thermal background:
    every cell of column 3 receives the whole column-2 population in one
    step."""
    g = O.make_grid(cells, cells)
    ny, nz = cells[1], cells[2]
    rows = np.repeat(np.arange(ny * nz), per_cell)
    n = len(rows)
    j, k = rows % ny, rows // ny
    a = {name: rng.random(n) for name in ("rx", "ry", "rz")}
    a["u"] = np.full(n, speed)
    a["v"] = rng.normal(size=n) * 0.01
    a["w"] = rng.normal(size=n) * 0.01
    a["q"] = np.full(n, -0.01)
    a["cell"] = O.encode(g, np.full(n, 2), j, k).astype(np.int32)
    a["id"] = np.arange(n, dtype=np.uint64)
    bg = cases.uniform_maxwellian(cells, cells, (1, 1, 1), -1.0, thermal=(0.05,) * 3, seed=5, first_id=n)
    return {key: np.concatenate([a[key], bg[key]]) for key in a}


@pytest.mark.parametrize("case", ["uniform", "converging", "failed-sort"])
def test_cycle_incremental_plan_equals_the_run_based_plan(case):
    """In tpic_mi_cycle a species sorted before its move is planned from the
    particles whose cell changed (the crossers the block mover records);
    with a recorded permutation every species takes the run-based plan. Both
    give the stable counting sort, so after four cycles E, B, J, the mass
    matrix, the bins and every particle attribute are equal bit for bit:
    for a two-species uniform plasma (the crosser plan throughout), for a
    slab whose column-2 particles all cross into column 3 in the first step
    (about 400 arrivals per bin: the crosser plan hands the species to the
    run-based plan), and after a sort failed on an invalid key (the crosser plan finds
    the old bins inconsistent and hands over as well)."""
    rng = np.random.default_rng(43)
    if case == "converging":
        cells = (8, 4, 3)
        lengths = cells
        species = [(-1.0, None)]
        arrays = [converging_slab(cells, 400, 10.0, rng)]
    else:
        cells, lengths = (9, 6, 5), (2.25, 1.5, 1.25)
        species = [(-16.0, None), (1.0, None)]
        arrays = [cases.uniform_maxwellian(cells, lengths, (3, 3, 4), -16.0, thermal=(0.1,) * 3, seed=3),
                  cases.uniform_maxwellian(cells, lengths, (2, 2, 3), 1.0, thermal=(0.02,) * 3, seed=4,
                                           first_id=10 ** 6)]
        if case == "failed-sort":
            arrays[0]["cell"][17] = O.bin_count(O.make_grid(cells, lengths)) + 3
    spec = [(qom, len(a["cell"])) for (qom, _), a in zip(species, arrays)]
    b = cases.uniform_b(cells, (0.01, 0.02, 0.05), amplitude=0.2)
    say("crosser plan against run-based plan:", case, "particles", [n for _, n in spec])
    sims = [gpu(cells, lengths, spec, dt=0.1, gmres_tolerance=1e-12) for _ in range(2)]
    try:
        sims[1].record_permutation(True)
        for sim in sims:
            for s, a in enumerate(arrays):
                sim.set_particles(s, a)
            sim.set_field(R.FIELD_B, b)
            sim.initialize()
        for n in range(4):
            r0, r1 = sims[0].cycle(1), sims[1].cycle(1)
            assert r0["iterations"] == r1["iterations"], n
            if case == "converging" and n == 0:
                counts, _ = sims[0].get_bins(0)
                assert counts.max() >= 350, counts.max()
        for which in (R.FIELD_E, R.FIELD_B, R.FIELD_J, R.FIELD_MASS):
            assert np.array_equal(sims[0].get_field(which), sims[1].get_field(which)), which
        for s in range(len(spec)):
            g0, g1 = sims[0].get_particles(s), sims[1].get_particles(s)
            for k in R.ATTRIBUTES:
                assert np.array_equal(g0[k], g1[k]), (s, k)
            c0, o0 = sims[0].get_bins(s)
            c1, o1 = sims[1].get_bins(s)
            assert np.array_equal(c0, c1) and np.array_equal(o0, o1), s
        assert sims[0].validation() == sims[1].validation()
        if case == "failed-sort":
            assert sims[0].validation() & R.INVALID_SORT
        if case == "uniform":
            assert sims[0].validation() == 0
    finally:
        for sim in sims:
            sim.close()


def test_energy_conservation_on_gpu():
    cells, lengths = (12, 10, 8), (3.0, 2.5, 2.0)
    arrays, b = setup_plasma(cells, lengths, (2, 2, 2), [-25.0, 1.0], [0.08, 0.016], (0.0, 0.0, 0.02), 0.3)
    counts = [len(a["cell"]) for a in arrays]
    with gpu(cells, lengths, [(-25.0, counts[0]), (1.0, counts[1])], dt=0.1, gmres_tolerance=1e-12) as sim:
        for s, a in enumerate(arrays):
            sim.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        e0 = sim.energy().sum()
        for _ in range(4):
            sim.cycle(1)
            e1 = sim.energy().sum()
            assert abs(e1 - e0) / e0 <= 1e-11
        assert sim.validation() == 0


@pytest.mark.parametrize("name,kernel,tol", MIXED_KERNELS)
def test_mixed_moments_energy_drift(name, kernel, tol):
    # The mixed forms keep the ECSIM energy identity to the drift bar of
    # 1e-8 per cycle (the plain form holds it at the GMRES tolerance).
    cells, lengths = (12, 10, 8), (3.0, 2.5, 2.0)
    arrays, b = setup_plasma(cells, lengths, (2, 2, 2), [-25.0, 1.0], [0.08, 0.016], (0.0, 0.0, 0.02), 0.3)
    counts = [len(a["cell"]) for a in arrays]
    with gpu(cells, lengths, [(-25.0, counts[0]), (1.0, counts[1])], dt=0.1, gmres_tolerance=1e-12,
             mass_kernel=kernel) as sim:
        for s, a in enumerate(arrays):
            sim.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        e_prev = sim.energy().sum()
        for n in range(4):
            sim.cycle(1)
            e1 = sim.energy().sum()
            say(name, "moments cycle", n, "drift", abs(e1 - e_prev) / e_prev)
            assert abs(e1 - e_prev) / e_prev <= 1e-8
            e_prev = e1
        assert sim.validation() == 0


def test_device_loader_statistics():
    cells, lengths = (8, 8, 8), (2.0, 2.0, 2.0)
    ppc = (3, 2, 2)
    n = 8 * 8 * 8 * 12
    with gpu(cells, lengths, [(-4.0, n)]) as sim:
        sim.load_maxwellian(0, ppc, rho=1.0, drift=(0.1, 0.0, -0.2), thermal=(0.5, 0.25, 0.125), seed=99,
                            first_id=1000)
        p = sim.get_particles(0)
    g = O.make_grid(cells, lengths)
    for k in ("rx", "ry", "rz"):
        assert np.all((p[k] >= 0) & (p[k] < 1))
    assert np.array_equal(p["id"], 1000 + np.arange(n, dtype=np.uint64))
    vol = (0.25 * 0.25) * 0.25
    assert np.all(p["q"] == -vol / 12 * (1.0 / cases.FOUR_PI))
    counts = np.bincount(p["cell"], minlength=O.bin_count(g))
    assert set(np.unique(counts[counts > 0])) == {12}
    for k, mean, sd in (("u", 0.1, 0.5), ("v", 0.0, 0.25), ("w", -0.2, 0.125)):
        assert abs(p[k].mean() - mean) < 5 * sd / np.sqrt(n)
        assert abs(p[k].std() / sd - 1) < 0.03


@pytest.mark.parametrize("quiet,periodic", [(False, (1, 1, 1)), (True, (1, 1, 1)), (False, (0, 1, 0))])
def test_uniform_load_matches_the_reference_draws(quiet, periodic):
    """tpic_mi_load_uniform (hipRAND XORWOW, the reference's seed, sequence
    and draw order) against the oracle's restatement of the same generator:
    positions, keys, ids and charges bit for bit; velocities within
    1e-14 |v| + 4e-15 thermal (the device Box-Muller angle is sincospi, the
    host's sin(pi w) with the rounded product: about 4e-16 in the angle, so
    up to 2e-15 in a normal of radius 5)."""
    cells, lengths = (6, 5, 4), (1.5, 1.25, 1.0)
    ppc = (2, 3, 2)
    n = 6 * 5 * 4 * 12
    species = [(-64.0, 10), (1.0, n)]  # the loaded species is species 1 (seed 12445)
    say("uniform load quiet", quiet, "periodic", periodic)
    with gpu(cells, lengths, species, particle_periodic=periodic) as sim:
        sim.load_uniform(1, ppc, rho_init=1.0, drift=(0.1, 0.0, -0.2), thermal=(0.045, 0.02, 0.03),
                         quiet_start=quiet, first_id=500)
        got = sim.get_particles(1)
        flags = sim.validation()
    g = O.make_grid(cells, lengths, periodic)
    exp = O.load_uniform(g, ppc, 1.0, (0.1, 0.0, -0.2), (0.045, 0.02, 0.03), quiet, O.load_seed(1), 500, 1.0).a
    assert flags == 0
    for k in ("rx", "ry", "rz", "q", "cell", "id"):
        assert np.array_equal(got[k].view(np.uint8), exp[k].view(np.uint8)), k
    for k, th in (("u", 0.045), ("v", 0.02), ("w", 0.03)):
        assert np.all(np.abs(got[k] - exp[k]) <= 1e-14 * np.abs(exp[k]) + 4e-15 * th), k


def test_center_velocity_matches_the_oracle():
    """tpic_mi_center_velocity after the Uniform-case load against the oracle's
    CenterInitialVelocity: positions, keys, ids and charges untouched (bit for
    bit); velocities within 1e-13 |v| + 1e-15 (the FP64 sums differ only in
    summation order, so the removed mean differs by a few ulps of the sums)."""
    cells, lengths = (7, 3, 2), (7.0, 3.0, 2.0)
    ppc, drift, thermal = (2, 2, 2), (0.125, -0.0625, 0.0), (0.1, 0.05, 0.025)
    n = 7 * 3 * 2 * 8
    say("center velocity", cells, ppc)
    with gpu(cells, lengths, [(-1.0, n), (0.5, n)]) as sim:
        sim.load_uniform(1, ppc, rho_init=0.5, drift=drift, thermal=thermal, first_id=n)
        sim.center_velocity(1, drift)
        got = sim.get_particles(1)
        flags = sim.validation()
    g = O.make_grid(cells, lengths)
    exp = O.load_uniform(g, ppc, 0.5, drift, thermal, False, O.load_seed(1), n, 0.5)
    assert O.center_velocity(exp, drift) == 0
    exp = exp.a
    assert flags == 0
    for k in ("rx", "ry", "rz", "q", "cell", "id"):
        assert np.array_equal(got[k].view(np.uint8), exp[k].view(np.uint8)), k
    for k in ("u", "v", "w"):
        assert np.all(np.abs(got[k] - exp[k]) <= 1e-13 * np.abs(exp[k]) + 1e-15), k
    q = got["q"]
    for k, d in zip(("u", "v", "w"), drift):
        assert abs(np.sum(q * (got[k] - d))) <= 1e-14 * np.sum(np.abs(q) * (np.abs(got[k]) + abs(d)))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_one_cell_axis_against_the_oracle(axis):
    """A periodic axis of one cell: three cycles of the HIP path against the
    oracle's (which equals the doubled axis, test_oracle): energies within
    1e-12 of the system energy, E and B within 1e-12 of their largest value,
    particle keys bit for bit."""
    from test_oracle import one_cell_state
    cells, lengths, a = one_cell_state(axis, 1)
    n = len(a["q"])
    nc = int(np.prod(cells))
    b = np.empty(3 * nc)
    b[:nc], b[nc:2 * nc], b[2 * nc:] = 0.1, -0.05, 0.3
    say("one-cell axis", axis, cells)
    with gpu(cells, lengths, [(-1.0, n)], gmres_tolerance=1e-12) as sim:
        sim.set_particles(0, a)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        got_e = []
        for _ in range(3):
            got_e.append(sim.energy())
            sim.cycle(1)
        got_e.append(sim.energy())
        got_fields = {w: sim.get_field(w) for w in (R.FIELD_E, R.FIELD_B)}
        got_p = sim.get_particles(0)
        flags = sim.validation()
    ref = O.Sim(cells, lengths, [-1.0], [n], 0.1, tol=1e-12)
    ref.set_particles(0, a)
    ref.field("B")[:] = b
    ref.initialize()
    want_e = []
    for _ in range(3):
        want_e.append(ref.energy())
        assert ref.cycle(1)[0] == 0
    want_e.append(ref.energy())
    assert flags == 0
    got_e, want_e = np.array(got_e), np.array(want_e)
    assert np.all(np.abs(got_e - want_e) <= 1e-12 * np.abs(want_e).sum(axis=1).max())
    for w, name in ((R.FIELD_E, "E"), (R.FIELD_B, "B")):
        want = ref.field(name)
        assert np.max(np.abs(got_fields[w] - want)) <= 1e-12 * np.max(np.abs(want)), name
    want_p = ref.particles(0)
    order = np.argsort(got_p["id"])
    worder = np.argsort(want_p["id"])
    assert np.array_equal(got_p["cell"][order], want_p["cell"][worder])
    ref.close()


def test_density_seeded_load_matches_the_oracle():
    """tpic_mi_load_density_seeded (the TwoStream profile) against the oracle:
    positions, keys, ids bit for bit; charges within 4 ulp (the device and
    host cosines may differ in the last bit); velocities as in the uniform
    load."""
    cells, lengths = (32, 2, 2), (16.0, 1.0, 1.0)
    ppc, drift, thermal = (4, 1, 2), (0.2, 0.0, 0.0), (0.01, 0.01, 0.01)
    mode = (2 * np.pi * 3 / 16.0, 0.25, 0.05)
    n = 32 * 2 * 2 * 8
    say("density-seeded load", cells, ppc)
    with gpu(cells, lengths, [(-1.0, n)]) as sim:
        sim.load_density_seeded(0, ppc, *mode, rho_init=0.5, drift=drift, thermal=thermal, first_id=7)
        got = sim.get_particles(0)
        flags = sim.validation()
    g = O.make_grid(cells, lengths)
    exp = O.load_uniform(g, ppc, 0.5, drift, thermal, False, O.load_seed(0), 7, -1.0, density_seed=mode).a
    assert flags == 0
    for k in ("rx", "ry", "rz", "cell", "id"):
        assert np.array_equal(got[k].view(np.uint8), exp[k].view(np.uint8)), k
    assert np.all(np.abs(got["q"] - exp["q"]) <= 4 * np.spacing(np.abs(exp["q"])))
    for k, th in (("u", 0.01), ("v", 0.01), ("w", 0.01)):
        assert np.all(np.abs(got[k] - exp[k]) <= 1e-14 * np.abs(exp[k]) + 4e-15 * th), k


# ---- closed walls --------------------------------------------------------------

def walled_case(axis, kind, seed=21):
    """5 x 4 x 3 cells with PEC or PMC walls on `axis` (particles mirrored
    there), two species, 6 particles per cell, uniform B along the
    components the wall keeps even, and a small E with its constrained rows
    left for startup to clear."""
    rng = np.random.default_rng(seed + 7 * axis + kind)
    cells, lengths = (5, 4, 3), (2.5, 2.0, 1.5)
    faces = [(R.FACE_PERIODIC, R.FACE_PERIODIC)] * 3
    faces[axis] = (kind, kind)
    periodic = [1, 1, 1]
    periodic[axis] = 0
    g = O.make_grid(cells, lengths, periodic, faces)
    per = 6
    n = int(np.prod(cells)) * per
    slab = np.arange(n) // per
    species = []
    for s, (qom, vt) in enumerate(((-16.0, 0.09), (1.0, 0.03))):
        a = {k: rng.uniform(0.02, 0.98, n) for k in ("rx", "ry", "rz")}
        for k in ("u", "v", "w"):
            a[k] = rng.normal(size=n) * vt
        a["q"] = np.full(n, -0.02 if qom < 0 else 0.02)
        a["cell"] = O.encode(g, slab % cells[0], (slab // cells[0]) % cells[1],
                             slab // (cells[0] * cells[1])).astype(np.int32)
        a["id"] = (np.arange(n) + s * n).astype(np.uint64)
        species.append((qom, a))
    nc = int(np.prod(cells))
    b = np.zeros(3 * nc)
    even = [c for c in range(3) if (c != axis) == (kind == R.FACE_PEC)]
    for c in even:
        b[c * nc:(c + 1) * nc] = 0.1 + 0.05 * c
    nx, ny, nz = cells
    e = rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)) * 1e-3
    e = periodic_duplicates(cells, e, axis)
    return cells, lengths, periodic, faces, species, b, e


def periodic_duplicates(cells, e, wall_axis):
    """Copy node plane 0 to plane n on the periodic axes (the duplicates)."""
    nx, ny, nz = cells
    v = e.reshape(3, nz + 1, ny + 1, nx + 1).copy()
    for ax, n in enumerate(cells):
        if ax == wall_axis:
            continue
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[3 - ax] = 0
        hi[3 - ax] = n
        v[tuple(hi)] = v[tuple(lo)]
    return v.ravel()


@pytest.mark.parametrize("kind", [R.FACE_PEC, R.FACE_PMC], ids=["PEC", "PMC"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_walled_cycles_against_the_oracle(axis, kind):
    """Closed walls on one axis: after startup J and M against the oracle
    (3e-12 of their size), then three cycles: E and B within 1e-10 of their
    size, energies within 1e-12, the same GMRES iteration counts, keys bit
    for bit (the oracle's walls are pinned by the mirror-doubled periodic box
    and the curl adjointness, test_oracle)."""
    cells, lengths, periodic, faces, species, b, e = walled_case(axis, kind)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    say("walls axis", axis, "kind", kind)
    with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=faces,
             gmres_tolerance=1e-12) as sim:
        ref = O.Sim(cells, lengths, qom, counts, 0.1, tol=1e-12, particle_periodic=periodic, field_faces=faces)
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
            ref.set_particles(s, a)
        sim.set_field(R.FIELD_E, e)
        sim.set_field(R.FIELD_B, b)
        ref.field("E")[:] = e
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        j_ref = ref.field("J")
        assert np.abs(sim.get_field(R.FIELD_J) - j_ref).max() <= 3e-12 * np.abs(j_ref).max() + 1e-15
        m_ref = ref.field("mass")
        assert np.abs(sim.get_field(R.FIELD_MASS) - m_ref).max() <= 3e-12 * np.abs(m_ref).max() + 1e-15
        assert np.array_equal(sim.get_field(R.FIELD_E), ref.field("E"))  # constrained rows cleared alike
        for _ in range(3):
            rg = sim.cycle(1)
            rc, ro = ref.cycle(1)
            assert rc == 0 and rg["iterations"] == ro.iterations
            eg, eo = sim.energy(), ref.energy()
            assert np.all(np.abs(eg - eo) <= 1e-12 * eo.sum())
            for w, name in ((R.FIELD_E, "E"), (R.FIELD_B, "B")):
                want = ref.field(name)
                assert np.abs(sim.get_field(w) - want).max() <= 1e-10 * np.abs(want).max(), name
        for s in range(len(species)):
            got, want = sim.get_particles(s), ref.particles(s)
            og, ow = np.argsort(got["id"]), np.argsort(want["id"])
            assert np.array_equal(got["cell"][og], want["cell"][ow])
        assert sim.validation() == 0
        ref.close()


@pytest.mark.parametrize("kind", [R.FACE_PEC, R.FACE_PMC], ids=["PEC", "PMC"])
def test_walled_operator_against_the_oracle(kind):
    """tpic_mi_apply_operator with y walls against the oracle's operator on a
    random vector (constrained rows zero): within 1e-12 of its size."""
    cells, lengths, periodic, faces, species, b, e = walled_case(1, kind, seed=5)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    rng = np.random.default_rng(9)
    nx, ny, nz = cells
    v = periodic_duplicates(cells, rng.normal(size=3 * (nx + 1) * (ny + 1) * (nz + 1)), 1)
    g = O.make_grid(cells, lengths, periodic, faces)
    with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=faces) as sim:
        ref = O.Sim(cells, lengths, qom, counts, 0.1, particle_periodic=periodic, field_faces=faces)
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
            ref.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        # project v as the solver does: zero constrained rows (via E's startup clearing)
        ref.field("E")[:] = v
        ref.initialize()
        vp = ref.field("E").copy()
        sim.set_field(R.FIELD_E, vp)
        sim.apply_operator(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
        out = sim.get_field(R.FIELD_RHS)
        th, dt = 0.5, 0.1
        vol = g.spacing[0] * g.spacing[1] * g.spacing[2]
        want = O.maxwell_apply(g, ref.field("mass").copy(), th * dt, 4 * np.pi * th * dt, 1 / vol, vp)
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()
        ref.close()


@pytest.mark.parametrize("kind", [R.FACE_PEC, R.FACE_PMC], ids=["PEC", "PMC"])
def test_walled_solve_projects_a_field_set_after_startup(kind):
    """tpic_mi_set_field(E) on an initialised walled context with nonzero
    constrained rows, then a solve: the guess is projected as
    enqueue_rhs_fields does (maxwell_device.hpp:140-142), so E_theta has zero
    constrained rows, the iteration count equals the oracle's and E_theta is
    within 1e-10 of its size."""
    cells, lengths, periodic, faces, species, b, e = walled_case(1, kind, seed=11)
    counts = [len(a["q"]) for _, a in species]
    qom = [q for q, _ in species]
    say("walled solve after set_field, kind", kind)
    with gpu(cells, lengths, list(zip(qom, counts)), particle_periodic=periodic, field_faces=faces,
             gmres_tolerance=1e-12) as sim:
        ref = O.Sim(cells, lengths, qom, counts, 0.1, tol=1e-12, particle_periodic=periodic, field_faces=faces)
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
            ref.set_particles(s, a)
        sim.set_field(R.FIELD_B, b)
        ref.field("B")[:] = b
        sim.initialize()
        ref.initialize()
        # After startup: E with its constrained rows left nonzero.
        sim.set_field(R.FIELD_E, e)
        ref.field("E")[:] = e
        _, rg = sim.solve()
        rc, ro = ref.solve()
        assert rc == 0 and rg["converged"] and rg["iterations"] == ro.iterations
        got, want = sim.get_field(R.FIELD_ETHETA), ref.field("theta")
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
        nx, ny, nz = cells
        v = got.reshape(3, nz + 1, ny + 1, nx + 1)
        # y walls: PEC holds E_x, E_z at zero on the wall planes, PMC holds E_y.
        held = (0, 2) if kind == R.FACE_PEC else (1,)
        for c in held:
            assert np.all(v[c][:, 0, :] == 0) and np.all(v[c][:, ny, :] == 0), c
        ref.close()


@pytest.mark.parametrize("case", ["GEMHarris", "DoubleGEM"])
def test_harris_load_matches_the_oracle(case):
    """tpic_mi_load_harris against the oracle's Harris profile: positions,
    keys and ids bit for bit; charges and velocities within 1e-13 relative
    (the device's cosh, tanh, exp, cos and sin may differ from the host's in
    the last bits)."""
    import harris as H
    cells, lengths = (16, 12, 1), (8.0, 6.0, 0.5)
    delta, b0x = 0.5, 0.0195
    sd = H.seed(case, lengths[0], lengths[1], delta, pert_hump=0.12566370614359174, deltax_hump=1.0,
                deltay_hump=1.0, kx_hump=0.6283185307179586, ky_hump=0.6283185307179586)
    species = [dict(qom=-25.0, rho=1.0, drift=(0.0, 0.0, 0.0065)), dict(qom=1.0, rho=1.0, drift=(0.0, 0.0, -0.0325))]
    prof = H.profiles(case, sd, b0x, delta, lengths[0], lengths[1], species, current_from_ampere=case == "GEMHarris")
    periodic = (1, 0, 1)
    faces = [(0, 0), (1, 1), (0, 0)]
    g = O.make_grid(cells, lengths, periodic, faces)
    ppc = (2, 2, 1)
    n = int(np.prod(cells)) * 4
    say("harris load", case)
    with gpu(cells, lengths, [(sp["qom"], n) for sp in species], particle_periodic=periodic,
             field_faces=faces) as sim:
        for s, (sp, h) in enumerate(zip(species, prof)):
            rp = R.HarrisProfile(harris_density=h.density, is_double=h.is_double, varying_thermal=h.varying_thermal,
                                 ly=h.ly, delta=h.delta, ampere_drift=h.ampere_drift, curl_delta=h.curl_delta,
                                 curl_sheet=h.sheet, curl_mode=h.mode, curl_hump=h.hump)
            rp.curl_center[0], rp.curl_center[1] = h.center_x, h.center_y
            rp.curl_mode_k[0], rp.curl_mode_k[1] = h.mode_kx, h.mode_ky
            rp.curl_hump_k[0], rp.curl_hump_k[1] = h.hump_kx, h.hump_ky
            rp.curl_hump_inverse_w2[0], rp.curl_hump_inverse_w2[1] = h.hump_inverse_wx2, h.hump_inverse_wy2
            sim.load_harris(s, ppc, rp, rho_init=sp["rho"], drift=sp["drift"], thermal=(0.02,) * 3, first_id=s * n)
        got = [sim.get_particles(s) for s in range(2)]
        flags = sim.validation()
    assert flags == 0
    for s, (sp, h) in enumerate(zip(species, prof)):
        exp = O.load_uniform(g, ppc, sp["rho"], sp["drift"], (0.02,) * 3, False, O.load_seed(s), s * n, sp["qom"],
                             harris_profile=h).a
        for k in ("rx", "ry", "rz", "cell", "id"):
            assert np.array_equal(got[s][k].view(np.uint8), exp[k].view(np.uint8)), k
        assert np.all(np.abs(got[s]["q"] - exp["q"]) <= 1e-13 * np.abs(exp["q"])), "q"
        for k in ("u", "v", "w"):
            assert np.all(np.abs(got[s][k] - exp[k]) <= 1e-13 * np.abs(exp[k]) + 1e-16), k
