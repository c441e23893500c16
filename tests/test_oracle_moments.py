"""The oracle's output moments (or_deposit_moments), pinned by a closed form
for one particle and by the CIC invariants on random particles: the weights of
a particle sum to one and reproduce linear functions, so over all nodes (wall
nodes divided by their dual-volume factor) rho sums to sum(q)/V and rho x_node
to sum(q x)/V; J and the second moments sum to the particles' own sums. The
component order, signs and the absence of the 4 pi output scale are those of
include/tpic/data/output/output_moment_quantities.hpp:26-34."""

import itertools

import numpy as np
import pytest

import oracle as O

PERIODIC, PEC = 0, 1


def arrays(g, ijk, frac, vel, q):
    n = len(q)
    return O.ParticleArrays(source={
        "rx": frac[:, 0], "ry": frac[:, 1], "rz": frac[:, 2],
        "u": vel[:, 0], "v": vel[:, 1], "w": vel[:, 2], "q": q,
        "cell": O.encode(g, *ijk).astype(np.int32), "id": np.arange(n, dtype=np.uint64)})


def node(cells, i, j, k):
    return i + (cells[0] + 1) * (j + (cells[1] + 1) * k)


def test_single_particle_closed_form():
    """One particle with dyadic fractions: every product is exact, so each of
    the 8 corner nodes (and its periodic copy at index 0 or n) holds
    amount x hat weight bit for bit, and every other node is zero."""
    cells, lengths = (4, 3, 5), (2.0, 1.5, 5.0)
    g = O.make_grid(cells, lengths)
    inv_v = 1.0 / (0.5 * 0.5 * 1.0)
    fr, vel, q, qom = (0.25, 0.5, 0.875), (1.0, -2.0, 3.0), 2.0, -0.5
    p = arrays(g, ([1], [2], [4]), np.array([fr]), np.array([vel]), np.array([q]))
    m, flags = O.deposit_moments(g, p, qom, inv_v)
    assert flags == 0
    charge = q * inv_v
    mass = charge / qom
    u, v, w = vel
    amount = [charge, charge * u, charge * v, charge * w,
              mass * u * u, mass * u * v, mass * u * w, mass * v * v, mass * v * w, mass * w * w]
    want = np.zeros_like(m)
    for a in range(8):
        bits = (a & 1, (a >> 1) & 1, (a >> 2) & 1)
        weight = np.prod([f if b else 1.0 - f for f, b in zip(fr, bits)])
        at = [c + b for c, b in zip((1, 2, 4), bits)]
        images = [sorted({x, 0, n}) if x in (0, n) else [x] for x, n in zip(at, cells)]
        for i, j, k in itertools.product(*images):
            want[:, node(cells, i, j, k)] += np.array(amount) * weight
    assert np.array_equal(m, want)
    assert m[0].max() == charge * (0.75 * 0.5 * 0.875)  # the largest weight: corner (0, y, 1)
    assert np.array_equal(m[5], m[1] * v / qom)  # PXY = J_x v / qom


def random_case(cells, lengths, n, seed, walls):
    rng = np.random.default_rng(seed)
    if walls:
        g = O.make_grid(cells, lengths, (0, 0, 0), [(PEC, PEC)] * 3)
    else:
        g = O.make_grid(cells, lengths)
    ijk = tuple(rng.integers(0, c, n) for c in cells)
    frac = rng.uniform(0.0, 1.0, (n, 3))
    vel = rng.normal(size=(n, 3))
    q = rng.uniform(0.5, 1.5, n) * np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    return g, arrays(g, ijk, frac, vel, q), np.stack(ijk, 1) + frac, vel, q


@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_cic_invariants(walls):
    cells, lengths = (5, 4, 6), (2.5, 1.0, 3.0)
    h = np.array(lengths) / np.array(cells)
    inv_v = 1.0 / np.prod(h)
    qom = 1.7
    g, p, pos, vel, q = random_case(cells, lengths, 400, 3 + walls, walls)
    m, flags = O.deposit_moments(g, p, qom, inv_v)
    assert flags == 0
    i, j, k = np.meshgrid(*[np.arange(c + 1) for c in cells], indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest, as the node index
    assert np.array_equal(node(cells, i, j, k), np.arange(len(i)))
    if walls:
        ends = [(x == 0) | (x == n) for x, n in zip((i, j, k), cells)]
        weight = 1.0 / 2.0 ** (ends[0].astype(int) + ends[1] + ends[2])
    else:  # one copy of each periodic node
        weight = ((i < cells[0]) & (j < cells[1]) & (k < cells[2])).astype(float)
    charge = q * inv_v
    mass = charge / qom
    u, v, w = vel.T
    sums = [charge, charge * u, charge * v, charge * w, mass * u * u, mass * u * v, mass * u * w,
            mass * v * v, mass * v * w, mass * w * w]
    for c in range(10):
        want = sums[c].sum()
        assert abs((weight * m[c]).sum() - want) <= 1e-12 * np.abs(sums[c]).sum(), c
    if walls:  # linear reproduction: corners mapped to the right nodes
        for axis, coord in enumerate((i, j, k)):
            got = (weight * m[0] * coord).sum()
            assert abs(got - (charge * pos[:, axis]).sum()) <= 1e-12 * np.abs(charge).sum() * cells[axis]


def test_selection_and_invalid_particles():
    cells, lengths = (3, 3, 3), (3.0, 3.0, 3.0)
    g, p, _, _, _ = random_case(cells, lengths, 60, 11, False)
    full, flags = O.deposit_moments(g, p, -2.0, 1.0)
    assert flags == 0
    sel = 0b1010010001  # rho, PXY, PYZ, PZZ
    part, flags = O.deposit_moments(g, p, -2.0, 1.0, sel)
    assert flags == 0
    for c in range(10):
        assert np.array_equal(part[c], full[c] if (sel >> c) & 1 else np.zeros_like(full[c]))
    bad = p.copy()
    bad.a["u"][0] = np.nan  # used by J_x and the x moments only
    rho, flags = O.deposit_moments(g, bad, -2.0, 1.0, 0b1)
    assert flags == 0 and np.array_equal(rho[0], full[0])
    _, flags = O.deposit_moments(g, bad, -2.0, 1.0)
    assert flags == 4  # non-finite amount: the particle is skipped
    bad = p.copy()
    bad.a["rx"][1] = 1.0
    bad.a["cell"][2] = O.encode(g, -1, 0, 0)
    _, flags = O.deposit_moments(g, bad, -2.0, 1.0)
    assert flags == 1 | 2
