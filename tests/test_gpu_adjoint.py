"""The adjoint path on the GPU through the C-ABI: the transposed operator and
solve, cycle checkpoints, the adjoint cycle against the oracle's adjoint
(oracle/oracle_adjoint.py, itself pinned by central differences of the
oracle's forward run in test_oracle), and the gradient harness
(tpic_amd.gradient.FieldEnergy) against central differences of the device's
forward run and of the oracle's."""
import numpy as np
import pytest

import oracle as O
import oracle_adjoint as AD
import tpic_amd
from tpic_amd import runtime as R

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


def case(rng, cells, lengths, ppc=5, qoms=(-8.0, 1.0), vscale=(0.3, 0.05)):
    """Random particles (fractions inside [0.15, 0.85]) with ids first + index."""
    g = O.make_grid(cells, lengths)
    nc = int(np.prod(cells))
    arrays = []
    for s, qom in enumerate(qoms):
        n = nc * ppc
        c = np.repeat(np.arange(nc), ppc)
        i, j, k = c % cells[0], (c // cells[0]) % cells[1], c // (cells[0] * cells[1])
        a = {name: rng.uniform(0.15, 0.85, n) for name in ("rx", "ry", "rz")}
        for name in ("u", "v", "w"):
            a[name] = rng.normal(size=n) * vscale[s]
        a["q"] = np.full(n, (-1.0 if qom < 0 else 1.0) * 0.05) * rng.uniform(0.8, 1.2, n)
        a["cell"] = O.encode(g, i, j, k).astype(np.int32)
        a["id"] = (np.arange(n) + s * 100000).astype(np.uint64)
        order = rng.permutation(n)
        arrays.append({key: val[order] for key, val in a.items()})
    return g, arrays


def started(cells, lengths, arrays, qoms, e, b, dt=0.2, tol=1e-13, **kw):
    sim = tpic_amd.Simulation(cells, lengths, [(q, len(a["q"])) for q, a in zip(qoms, arrays)], dt=dt,
                              gmres_tolerance=tol, gmres_max_iterations=400, **kw)
    for s, a in enumerate(arrays):
        sim.set_particles(s, a)
    sim.set_field(R.FIELD_E, e)
    sim.set_field(R.FIELD_B, b)
    sim.initialize()
    return sim


def test_transposed_operator_is_the_adjoint_of_the_operator():
    """<u, A v> = <A^T u, v> over the unique nodes for random periodic u, v
    (1e-12 of the products' scale), and A^T u equals the dense transpose of
    the oracle's operator on the same deposited matrix (1e-12)."""
    rng = np.random.default_rng(41)
    cells, lengths, qoms = (5, 4, 3), (2.5, 2.0, 1.5), (-8.0, 1.0)
    g, arrays = case(rng, cells, lengths, ppc=4, qoms=qoms)
    box = AD.Box(g)
    b0 = rng.normal(size=3 * O.cell_count(g)) * 0.3
    say("transposed operator")
    with started(cells, lengths, arrays, qoms, np.zeros(3 * box.nodes), b0) as sim:
        u = box.expand(rng.normal(size=3 * box.unique))
        v = box.expand(rng.normal(size=3 * box.unique))
        ein, rhs = sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS)
        sim.set_field(R.FIELD_E, v)
        sim.apply_operator(ein, rhs)
        av = sim.get_field(R.FIELD_RHS)
        sim.set_field(R.FIELD_E, u)
        sim.apply_operator_transposed(ein, rhs)
        atu = sim.get_field(R.FIELD_RHS)
        mass = sim.get_field(R.FIELD_MASS)
    left = np.dot(box.reduce(u), box.reduce(av))
    right = np.dot(box.reduce(atu), box.reduce(v))
    scale = np.linalg.norm(box.reduce(u)) * np.linalg.norm(box.reduce(av))
    assert abs(left - right) <= 1e-12 * scale, (left, right)
    dense = box.operator(mass, 0.5 * 0.2, 4 * np.pi * 0.5 * 0.2)
    want = dense.T @ box.reduce(u)
    assert np.abs(box.reduce(atu) - want).max() <= 1e-12 * np.abs(want).max()
    # Every periodic duplicate holds its unique node's value.
    assert np.abs(atu - box.expand(box.reduce(atu))).max() <= 1e-12 * np.abs(atu).max()


def test_transposed_solve_residual():
    """GMRES on A^T: |A^T x - b| <= tol |b| (tolerance 1e-12)."""
    rng = np.random.default_rng(43)
    cells, lengths, qoms = (6, 4, 2), (3.0, 2.0, 1.0), (-8.0, 1.0)
    g, arrays = case(rng, cells, lengths, ppc=4, qoms=qoms)
    box = AD.Box(g)
    with started(cells, lengths, arrays, qoms, np.zeros(3 * box.nodes), rng.normal(size=3 * O.cell_count(g)) * 0.3,
                 tol=1e-12) as sim:
        b = box.expand(rng.normal(size=3 * box.unique))
        sim.set_field(R.FIELD_RHS, b)
        sim.set_field(R.FIELD_ETHETA, np.zeros_like(b))
        r = sim.solve_transposed(sim.device_field(R.FIELD_RHS), sim.device_field(R.FIELD_ETHETA))
        x = sim.get_field(R.FIELD_ETHETA)
        sim.set_field(R.FIELD_E, x)
        sim.apply_operator_transposed(sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_RHS))
        atx = sim.get_field(R.FIELD_RHS)
    say("transposed solve", r)
    assert r["converged"] == 1
    assert np.linalg.norm(atx - b) <= 1e-11 * np.linalg.norm(b)


def test_checkpoint_restart_repeats_the_run_bit_for_bit():
    """Six cycles straight; three cycles, a checkpoint, three more; and a
    restore of the checkpoint followed by three cycles: E, B and every
    particle attribute equal bit for bit."""
    rng = np.random.default_rng(47)
    cells, lengths, qoms = (6, 4, 3), (3.0, 2.0, 1.5), (-8.0, 1.0)
    g, arrays = case(rng, cells, lengths, ppc=6, qoms=qoms)
    nn, nc = O.node_count(g), O.cell_count(g)
    e0 = np.zeros(3 * nn)
    b0 = rng.normal(size=3 * nc) * 0.2

    def state(sim):
        return [sim.get_field(R.FIELD_E), sim.get_field(R.FIELD_B)] + [sim.get_particles(s) for s in range(2)]

    def same(a, b):
        for x, y in zip(a[:2], b[:2]):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
        for pa, pb in zip(a[2:], b[2:]):
            for k in pa:
                assert np.array_equal(pa[k].view(np.uint8), pb[k].view(np.uint8)), k

    with started(cells, lengths, arrays, qoms, e0, b0, dt=0.25) as sim:
        sim.cycle(6)
        straight = state(sim)
    with started(cells, lengths, arrays, qoms, e0, b0, dt=0.25) as sim:
        sim.checkpoint_reserve(1)
        sim.cycle(3)
        sim.checkpoint_save(0)
        sim.cycle(3)
        same(state(sim), straight)
        sim.cycle(2)  # move on, then go back
        sim.checkpoint_restore(0)
        sim.cycle(3)
        same(state(sim), straight)
        assert sim.validation() == 0


def test_adjoint_cycles_match_the_oracle_adjoint():
    """Four adjoint cycles of W_E on the device against the oracle's adjoint
    of the same run: Ebar, Bbar and every particle's rbar, vbar, qbar within
    1e-8 of their largest value (the two forward runs agree to 1e-12 and the
    device solves A^T to 1e-13, the oracle exactly)."""
    rng = np.random.default_rng(3)
    cells, lengths, qoms, dt, cycles = (6, 4, 1), (3.0, 2.0, 0.5), (-8.0, 1.0), 0.2, 4
    g, arrays = case(rng, cells, lengths, qoms=qoms)
    box = AD.Box(g)
    nc = O.cell_count(g)
    e0 = box.expand(rng.normal(size=3 * box.unique) * 0.01)
    b0 = rng.normal(size=3 * nc) * 0.05 + np.repeat([0.1, 0.0, 0.2], nc)
    counts = [len(a["q"]) for a in arrays]
    say("adjoint cycles against the oracle")
    with started(cells, lengths, arrays, qoms, e0, b0, dt=dt) as sim:
        sim.checkpoint_reserve(cycles)
        for n in range(cycles):
            sim.checkpoint_save(n)
            sim.cycle(1)
        with sim.adjoint() as adj:
            adj.seed_field_energy(1.0, 0.0, [0, 100000])
            for n in range(cycles - 1, -1, -1):
                sim.checkpoint_restore(n)
                fwd, tr = adj.cycle()
                assert fwd["converged"] and tr["converged"]
            ebar, bbar = adj.field(R.FIELD_E), adj.field(R.FIELD_B)
            parts = [adj.particles(s) for s in range(2)]
    osim = O.Sim(cells, lengths, list(qoms), counts, dt, tol=1e-14, restart=60, max_iterations=400)
    for s, a in enumerate(arrays):
        osim.set_particles(s, a)
    osim.field("E")[:] = e0
    osim.field("B")[:] = b0
    osim.initialize()
    states = AD.run_forward(osim, cycles)
    want = AD.gradient_of_field_energy(osim, box, states, dt, 0.5, 1.0, 0.0)
    osim.close()
    assert np.abs(box.reduce(ebar) - want.ebar).max() <= 1e-8 * np.abs(want.ebar).max()
    assert np.abs(ebar - box.expand(box.reduce(ebar))).max() == 0.0
    assert np.abs(bbar - want.bbar).max() <= 1e-8 * np.abs(want.bbar).max()
    for s in range(2):
        got, exp = parts[s], want.particles[s]
        at = {int(x): n for n, x in enumerate(got["ids"])}
        idx = np.array([at[int(x)] for x in exp["ids"]])
        for key in ("rbar", "vbar", "qbar"):
            g_ = got[key][..., idx]
            assert np.abs(g_ - exp[key]).max() <= 1e-8 * np.abs(exp[key]).max(), (s, key)


def _oracle_two_stream_energy(model, v0, eps, cycles):
    """W_E after `cycles` of the oracle on the model's initial state at (v0, eps)."""
    v, q, e, b = (t.numpy() for t in model.state(v0, eps))
    sim = O.Sim(model.cells, model.lengths, [-1.0, -1.0], model.run.counts, model.dt, tol=1e-14, restart=30,
                max_iterations=400)
    off = 0
    for s, base in enumerate(model.base):
        n = len(base["q"])
        a = {k: base[k] for k in ("rx", "ry", "rz", "cell", "id")}
        a["u"], a["v"], a["w"] = v[off:off + n, 0], v[off:off + n, 1], v[off:off + n, 2]
        a["q"] = q[off:off + n]
        sim.set_particles(s, a)
        off += n
    sim.field("E")[:] = e
    sim.field("B")[:] = b
    sim.initialize()
    rc, _ = sim.cycle(cycles)
    assert rc == 0
    we = sim.energy()[0]
    sim.close()
    return we


def test_two_stream_gradient_small_against_finite_differences():
    """dW_E/dv0 and dW_E/deps of a 64-cell, 2 x 32 ppc two-stream run over 20
    cycles (checkpoints every 5): the adjoint through FieldEnergy against
    central differences (h = 1e-6) of the device's forward run and of the
    oracle's, within 1e-6 relative."""
    from tpic_amd import gradient as G
    model = G.TwoStream(cells=64, ppc=32, mode=2, cycles=20, every=5, gmres_tolerance=1e-13)
    try:
        value, grads = G.gradient(model, 0.2, 1e-3)
        fd_gpu = [G.central_difference(model, [0.2, 1e-3], i, 1e-6) for i in range(2)]
        fd_oracle = [(_oracle_two_stream_energy(model, 0.2 + h0, 1e-3 + h1, 20) -
                      _oracle_two_stream_energy(model, 0.2 - h0, 1e-3 - h1, 20)) / 2e-6
                     for h0, h1 in ((1e-6, 0.0), (0.0, 1e-6))]
    finally:
        model.close()
    say("two-stream small: W_E", value, "adjoint", grads, "fd gpu", fd_gpu, "fd oracle", fd_oracle)
    for got, a, b in zip(grads, fd_gpu, fd_oracle):
        assert abs(got - a) <= 1e-6 * abs(a)
        assert abs(got - b) <= 1e-6 * abs(b)


def test_restart_file_continues_the_run_bit_for_bit(tmp_path):
    """Three cycles, a restart file, and a fresh context that reads it,
    starts up and runs three more cycles: E, B and every particle attribute
    equal those of six cycles straight, bit for bit; the header carries the
    cycle, grid and species."""
    rng = np.random.default_rng(53)
    cells, lengths, qoms = (6, 4, 3), (3.0, 2.0, 1.5), (-8.0, 1.0)
    g, arrays = case(rng, cells, lengths, ppc=6, qoms=qoms)
    nn, nc = O.node_count(g), O.cell_count(g)
    e0 = np.zeros(3 * nn)
    b0 = rng.normal(size=3 * nc) * 0.2
    path = str(tmp_path / "checkpoint_t3.tpicmi")
    with started(cells, lengths, arrays, qoms, e0, b0, dt=0.25) as sim:
        sim.cycle(6)
        straight = [sim.get_field(R.FIELD_E), sim.get_field(R.FIELD_B)] + [sim.get_particles(s) for s in range(2)]
    with started(cells, lengths, arrays, qoms, e0, b0, dt=0.25) as sim:
        sim.cycle(3)
        sim.checkpoint_write(path, cycle=3, initial_energy=1.5)
    counts = [len(a["q"]) for a in arrays]
    with tpic_amd.Simulation(cells, lengths, list(zip(qoms, counts)), dt=0.25, gmres_tolerance=1e-13,
                             gmres_max_iterations=400) as sim:
        h = sim.checkpoint_read(path)
        assert (h.cycle, h.species, tuple(h.global_cells)) == (3, 2, cells)
        assert tuple(h.q_over_m[:2]) == qoms and tuple(h.global_count[:2]) == tuple(counts)
        assert h.initial_energy == 1.5 and h.dt == 0.25
        sim.initialize()
        sim.cycle(3)
        got = [sim.get_field(R.FIELD_E), sim.get_field(R.FIELD_B)] + [sim.get_particles(s) for s in range(2)]
    for x, y in zip(got[:2], straight[:2]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    for pa, pb in zip(got[2:], straight[2:]):
        for k in pa:
            assert np.array_equal(pa[k].view(np.uint8), pb[k].view(np.uint8)), k
