"""The adjoint of the oracle's ECSIM cycle, in numpy (test infrastructure).

The reference has no adjoint. This module differentiates the oracle's own
cycle (or_sim_cycle of tpic_oracle.c: solve, move, field advance, sort,
deposition, assembly on a periodic box) by reverse mode, with the field
operators taken as dense matrices built from the oracle's own kernels
(or_curl_n2c, or_curl_c2n, or_cell_to_node and the operator of the mass
matrix the oracle assembled), so that no transpose identity is assumed. It is
pinned by central finite differences of the oracle's forward run
(tests/test_oracle.py) and is the checker of the device adjoint
(tpic_mi_adjoint_cycle).

Variables are the unique nodes of the periodic box (the node vectors of the
oracle and of the device hold each periodic duplicate with its unique node's
value), the cells, and per particle its cell fractions r, velocity v and
charge q, in the sorted order of the state they belong to.
"""
import numpy as np

import oracle as O

FOUR_PI = 4.0 * 3.14159265358979323846


class Box:
    """Index maps and dense field operators of one periodic grid."""

    def __init__(self, g):
        self.g = g
        nx, ny, nz = g.n[0], g.n[1], g.n[2]
        self.n = (nx, ny, nz)
        self.nodes = O.node_count(g)
        self.cells = O.cell_count(g)
        self.unique = nx * ny * nz
        i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
        node = (i + (nx + 1) * (j + (ny + 1) * k)).ravel()
        uniq = ((i % nx) + nx * ((j % ny) + ny * (k % nz))).ravel()
        # P: unique -> all nodes (duplicates copied); R: all -> unique.
        self.copy_of = np.zeros(self.nodes, np.int64)
        self.copy_of[node] = uniq
        self.rep = np.zeros(self.unique, np.int64)
        base = (i < nx) & (j < ny) & (k < nz)
        self.rep[uniq[base.ravel()]] = node[base.ravel()]
        self.volume = float(np.float64(np.longdouble(g.spacing[0]) * np.longdouble(g.spacing[1]) *
                                       np.longdouble(g.spacing[2])))
        # Dense operators on unique nodes / cells, from the oracle's kernels.
        eye_u = np.eye(3 * self.unique)
        self.n2c = np.stack([O.curl_n2c(g, self.expand(col)) for col in eye_u.T], axis=1)  # 3C x 3U
        eye_c = np.eye(3 * self.cells)
        self.c2n = np.stack([self.reduce(O.curl_c2n(g, col)) for col in eye_c.T], axis=1)  # 3U x 3C
        self.c2n_avg = np.stack([self.reduce(O.cell_to_node(g, col)) for col in eye_c.T], axis=1)  # 3U x 3C

    def expand(self, u):
        """Unique-node vector (3U) -> node vector with duplicates (3N)."""
        u = np.asarray(u).reshape(3, self.unique)
        return u[:, self.copy_of].ravel()

    def reduce(self, v):
        """Node vector (3N) -> its unique nodes (3U)."""
        v = np.asarray(v).reshape(3, self.nodes)
        return v[:, self.rep].ravel()

    def operator(self, mass, delta, current_factor):
        """The oracle's Maxwell operator on unique nodes (3U x 3U)."""
        eye = np.eye(3 * self.unique)
        cols = [self.reduce(O.maxwell_apply(self.g, mass, delta, current_factor, 1.0 / self.volume,
                                            self.expand(col))) for col in eye.T]
        return np.stack(cols, axis=1)


def cic(r):
    """Weights w[p, a] and their fraction derivatives dw[p, a, axis]."""
    f = np.stack([r[0], r[1], r[2]], axis=1)
    ones = 1.0 - f
    w = np.zeros((f.shape[0], 8))
    dw = np.zeros((f.shape[0], 8, 3))
    for a in range(8):
        bits = (a & 1, (a >> 1) & 1, (a >> 2) & 1)
        fac = [f[:, x] if bits[x] else ones[:, x] for x in range(3)]
        sign = [1.0 if bits[x] else -1.0 for x in range(3)]
        w[:, a] = fac[0] * fac[1] * fac[2]
        dw[:, a, 0] = sign[0] * fac[1] * fac[2]
        dw[:, a, 1] = sign[1] * fac[0] * fac[2]
        dw[:, a, 2] = sign[2] * fac[0] * fac[1]
    return w, dw


def response(om):
    x, y, z = om[:, 0], om[:, 1], om[:, 2]
    d = 1.0 + (x * x + y * y + z * z)
    al = np.stack([1 + x * x, x * y + z, x * z - y, x * y - z, 1 + y * y, y * z + x, x * z + y, y * z - x,
                   1 + z * z], axis=1).reshape(-1, 3, 3) / d[:, None, None]
    return al, d


def omega_gradient(om, d, al, a, b):
    """grad_omega (a . alpha b) per particle."""
    ab = np.einsum("pij,pj->pi", al, b)
    form = np.sum(a * ab, axis=1)
    ao = np.sum(a * om, axis=1)
    bo = np.sum(b * om, axis=1)
    return (ao[:, None] * b + bo[:, None] * a + np.cross(a, b) - 2.0 * form[:, None] * om) / d[:, None]


def corners(box, cell_key):
    """Unique node of each particle's 8 cell corners [p, a]."""
    g = box.g
    nx, ny, nz = box.n
    ex, ey = nx + 4, ny + 4
    key = np.asarray(cell_key, np.int64)
    i = key % ex - 2
    j = (key // ex) % ey - 2
    k = key // (ex * ey) - 2
    out = np.zeros((len(key), 8), np.int64)
    for a in range(8):
        ax, ay, az = a & 1, (a >> 1) & 1, (a >> 2) & 1
        out[:, a] = ((i + ax) % nx) + nx * (((j + ay) % ny) + ny * ((k + az) % nz))
    assert g is not None
    return out


def gather(field_u, nodes, w, box):
    """Field (3U, unique nodes) at the particles: [p, 3]."""
    f = np.asarray(field_u).reshape(3, box.unique)
    return np.stack([np.sum(w * f[c][nodes], axis=1) for c in range(3)], axis=1)


def scatter(values, nodes, w, box):
    """sum_p w[p, a] values[p] into the unique nodes (3U)."""
    out = np.zeros((3, box.unique))
    for c in range(3):
        np.add.at(out[c], nodes.ravel(), (w * values[:, c:c + 1]).ravel())
    return out.ravel()


class Adjoint:
    """The adjoint of one state: unique-node E, cell B, and per species the
    particle adjoints rbar[3, n], vbar[3, n], qbar[n] in the order of `ids`."""

    def __init__(self, ebar, bbar, particles):
        self.ebar = ebar
        self.bbar = bbar
        self.particles = particles  # list of dicts rbar, vbar, qbar, ids


def seed_field_energy(box, e_nodes, b_cells, weight_e=1.0, weight_b=0.0, species_ids=()):
    scale = box.volume / FOUR_PI
    parts = [dict(rbar=np.zeros((3, len(ids))), vbar=np.zeros((3, len(ids))), qbar=np.zeros(len(ids)),
                  ids=np.asarray(ids).copy()) for ids in species_ids]
    return Adjoint(weight_e * scale * box.reduce(e_nodes), weight_b * scale * np.asarray(b_cells, float), parts)


def permute(adj_part, ids):
    """The particle adjoints moved to the order of `ids`."""
    order = {int(x): n for n, x in enumerate(adj_part["ids"])}
    idx = np.array([order[int(x)] for x in ids])
    return dict(rbar=adj_part["rbar"][:, idx], vbar=adj_part["vbar"][:, idx], qbar=adj_part["qbar"][idx],
                ids=np.asarray(ids).copy())


def adjoint_cycle(sim, box, adj, dt, theta, c=1.0):
    """One cycle backwards. `sim` (oracle Sim) holds the state the cycle
    starts from (sorted, deposited); `adj` is the adjoint of the state the
    cycle ends in. Returns the adjoint of the starting state."""
    g = box.g
    delta = c * theta * dt
    k = 4 * 3.14159265358979323846 * theta * dt
    inv_v = 1.0 / box.volume
    parts = []
    states = []
    for s in range(sim.species):
        p = sim.particles(s)
        parts.append(permute(adj.particles[s], p["id"]))
        states.append(p)
    rc, res = sim.solve()
    assert rc == 0
    etheta = box.reduce(sim.field("theta"))
    bn = box.reduce(sim.field("Bn"))
    mass = sim.field("mass").copy()
    A = box.operator(mass, delta, k)
    # Field advance: E' = E_theta / theta - (1 - theta)/theta E, B' = B - c dt curl_N2C(E_theta).
    if abs(theta - 1.0) < 1e-10:
        c0, c1 = 1.0, 0.0
    else:
        c0, c1 = 1.0 / theta, -(1.0 - theta) / theta
    ethetabar = c0 * adj.ebar + (-c * dt) * (box.n2c.T @ adj.bbar)
    ebar = c1 * adj.ebar
    bbar = adj.bbar.copy()
    spacing = np.array([g.spacing[0], g.spacing[1], g.spacing[2]])
    pre = []
    for s in range(sim.species):
        st, ad = states[s], parts[s]
        r = [st["rx"], st["ry"], st["rz"]]
        v = np.stack([st["u"], st["v"], st["w"]], axis=1)
        nodes = corners(box, st["cell"])
        w, dw = cic(r)
        beta = sim.q_over_m[s] * dt / 2
        beta_c = beta / c
        bp = gather(bn, nodes, w, box)
        ep = gather(etheta, nodes, w, box)
        om = beta_c * bp
        al, d = response(om)
        gv = ad["vbar"].T + (dt / spacing)[None, :] * ad["rbar"].T
        atg = np.einsum("pji,pj->pi", al, gv)
        vbar = 2.0 * atg - gv
        ebar_p = 2.0 * beta * atg
        b = v + beta * ep
        wbar = 2.0 * omega_gradient(om, d, al, gv, b)
        ea = np.stack([etheta.reshape(3, -1)[cc][nodes] for cc in range(3)], axis=2)  # [p, a, 3]
        rbar = ad["rbar"].T + np.einsum("pax,pa->px", dw, np.einsum("pac,pc->pa", ea, ebar_p))
        ethetabar = ethetabar + scatter(ebar_p, nodes, w, box)
        pre.append(dict(r=r, v=v, q=st["q"], nodes=nodes, w=w, dw=dw, beta_c=beta_c, om=om, al=al, d=d, ep=ep,
                        ea=ea, rbar=rbar, vbar=vbar, wbar=wbar, qbar=ad["qbar"].copy(), ids=ad["ids"]))
    mu = np.linalg.solve(A.T, ethetabar)
    ebar = ebar + mu
    bbar = bbar + delta * (box.c2n.T @ mu)
    bnbar = np.zeros(3 * box.unique)
    out_parts = []
    mu3 = mu.reshape(3, -1)
    bn3 = bn.reshape(3, -1)
    for s, x in enumerate(pre):
        nodes, w, dw, al, om, d = x["nodes"], x["w"], x["dw"], x["al"], x["om"], x["d"]
        q, v, ep = x["q"], x["v"], x["ep"]
        m = gather(mu, nodes, w, box)
        jb = -k * m
        sj = q * inv_v
        av = np.einsum("pij,pj->pi", al, v)
        ae = np.einsum("pij,pj->pi", al, ep)
        atj = np.einsum("pji,pj->pi", al, jb)
        atm = np.einsum("pji,pj->pi", al, m)
        vbar = x["vbar"] + sj[:, None] * atj
        qbar = x["qbar"] + inv_v * np.sum(jb * av, axis=1)
        wbar = x["wbar"] + sj[:, None] * omega_gradient(om, d, al, jb, v)
        t_mass = -(k * inv_v) * x["beta_c"]
        qbar = qbar + t_mass * np.sum(m * ae, axis=1)
        wbar = wbar + (t_mass * q)[:, None] * omega_gradient(om, d, al, m, ep)
        bbar_p = x["beta_c"] * wbar
        mua = np.stack([mu3[cc][nodes] for cc in range(3)], axis=2)
        ba = np.stack([bn3[cc][nodes] for cc in range(3)], axis=2)
        s_a = (sj[:, None] * (-k) * np.einsum("pac,pc->pa", mua, av) +
               (t_mass * q)[:, None] * (np.einsum("pac,pc->pa", mua, ae) + np.einsum("pac,pc->pa", x["ea"], atm)) +
               np.einsum("pac,pc->pa", ba, bbar_p))
        rbar = x["rbar"] + np.einsum("pax,pa->px", dw, s_a)
        bnbar = bnbar + scatter(bbar_p, nodes, w, box)
        out_parts.append(dict(rbar=rbar.T.copy(), vbar=vbar.T.copy(), qbar=qbar, ids=x["ids"]))
    bbar = bbar + box.c2n_avg.T @ bnbar
    return Adjoint(ebar, bbar, out_parts), res


def run_forward(sim, cycles):
    """Cycles with the start state of every cycle kept (particles, E, B)."""
    states = [snapshot(sim)]
    for _ in range(cycles):
        rc, _ = sim.cycle(1)
        assert rc == 0
        states.append(snapshot(sim))
    return states


def snapshot(sim):
    return dict(particles=[{k: v.copy() for k, v in sim.particles(s).items()} for s in range(sim.species)],
                E=sim.field("E").copy(), B=sim.field("B").copy())


def restore(sim, snap):
    for s, p in enumerate(snap["particles"]):
        sim.set_particles(s, p)
    sim.field("E")[:] = snap["E"]
    sim.field("B")[:] = snap["B"]
    sim.initialize()


def gradient_of_field_energy(sim, box, states, dt, theta, weight_e=1.0, weight_b=0.0, c=1.0):
    """The adjoint of L = weight_e W_E + weight_b W_B of the last state with
    respect to the first: walks the cycles backwards from the kept states."""
    last = states[-1]
    adj = seed_field_energy(box, last["E"], last["B"], weight_e, weight_b,
                            [p["id"] for p in last["particles"]])
    for n in range(len(states) - 2, -1, -1):
        restore(sim, states[n])
        adj, _ = adjoint_cycle(sim, box, adj, dt, theta, c)
    return adj
