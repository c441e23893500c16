"""Output moments (tpic_mi_moments) against the oracle's or_deposit_moments
on the particles the device holds after startup and two cycles: every
component within 1e-12 of its largest value (the device sums each cell's
corners in particle order and then the 8 cells of a node; the oracle scatters
particle by particle), periodic and walled, all components and a selection."""

import numpy as np
import pytest

import oracle as O
from tpic_amd import runtime as R
from test_gpu_parity import gpu, periodic_nodes, say, walled_case

pytestmark = pytest.mark.gpu


def periodic_case(seed=5):
    rng = np.random.default_rng(seed)
    cells, lengths = (6, 5, 4), (3.0, 2.5, 2.0)
    g = O.make_grid(cells, lengths)
    per = 7
    n = int(np.prod(cells)) * per
    slab = np.arange(n) // per
    species = []
    for s, (qom, vt) in enumerate(((-16.0, 0.08), (1.0, 0.03))):
        a = {k: rng.uniform(0.0, 1.0, n) for k in ("rx", "ry", "rz")}
        for k in ("u", "v", "w"):
            a[k] = rng.normal(size=n) * vt + 0.01 * (s + 1)
        a["q"] = rng.uniform(0.5, 1.5, n) * (-0.02 if qom < 0 else 0.02)
        a["cell"] = O.encode(g, slab % cells[0], (slab // cells[0]) % cells[1],
                             slab // (cells[0] * cells[1])).astype(np.int32)
        a["id"] = (np.arange(n) + s * n).astype(np.uint64)
        species.append((qom, a))
    nc = int(np.prod(cells))
    b = np.concatenate([np.full(nc, 0.1), np.full(nc, -0.05), np.full(nc, 0.2)])
    e = periodic_nodes(cells, rng, 1e-3)
    return cells, lengths, [1, 1, 1], [(R.FACE_PERIODIC,) * 2] * 3, species, b, e


@pytest.mark.parametrize("case", ["periodic", "pec-y", "pmc-z"])
def test_moments_against_the_oracle(case):
    if case == "periodic":
        cells, lengths, periodic, faces, species, b, e = periodic_case()
    else:
        cells, lengths, periodic, faces, species, b, e = walled_case(
            1 if case == "pec-y" else 2, R.FACE_PEC if case == "pec-y" else R.FACE_PMC)
    say("moments case", case, "cells", cells)
    g = O.make_grid(cells, lengths, periodic, faces)
    inv_v = 1.0 / float(np.prod(np.array(lengths) / np.array(cells)))
    with gpu(cells, lengths, [(q, len(a["q"])) for q, a in species], particle_periodic=periodic,
             field_faces=faces) as sim:
        for s, (_, a) in enumerate(species):
            sim.set_particles(s, a)
        with pytest.raises(R.TpicError):  # not sorted before startup
            sim.moments(0)
        sim.set_field(R.FIELD_E, e)
        sim.set_field(R.FIELD_B, b)
        sim.initialize()
        sim.cycle(2)
        for s, (qom, _) in enumerate(species):
            got = sim.moments(s)
            want, flags = O.deposit_moments(g, O.ParticleArrays(source=sim.get_particles(s)), qom, inv_v)
            assert flags == 0
            for c in range(10):
                scale = np.abs(want[c]).max()
                assert scale > 0, c
                assert np.abs(got[c] - want[c]).max() <= 1e-12 * scale, (s, c)
            sel = 0b1010010001
            part = sim.moments(s, sel)
            for c in range(10):
                assert np.array_equal(part[c], got[c] if (sel >> c) & 1 else np.zeros_like(got[c])), (s, c)
        assert sim.validation() == 0
