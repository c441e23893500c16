"""Gradients of field energies through ECSIM cycles on the GPU.

`FieldEnergy` is a torch.autograd.Function whose forward runs the cycles of a
Simulation through the C-ABI and whose backward runs the adjoint kernels
(tpic_mi_adjoint_cycle) over cycle checkpoints with recomputation: the
forward keeps the state at the start of every segment of `every` cycles;
the backward restores each segment's start, recomputes its cycles keeping
each cycle's start, and walks them backwards. Torch holds only host tensors
here (the parameters and the initial state); every device buffer belongs to
the native context. Periodic fields and particles, no smoothing.

The two configurations with gradients:
- two_stream_1d: two counter-streaming electron beams (+-v0, thermal vth) on
  a periodic line, the TwoStream density seed (charge times
  1 + eps cos(k x + phase)) and its E_x seed; d(W_E at the end)/d(v0, eps).
- weibel_2d: bi-Maxwellian electrons (thermal speed along x and z = ratio x
  the one along y) with Maxwellian ions on a periodic plane;
  d(W_B at the end)/d(ratio).
"""
import numpy as np
import torch

from . import runtime as R

TWO_PI = 6.283185307179586476925286766559
FOUR_PI = 12.566370614359172953850573533118


def _slots(cycles, every):
    return (cycles + every - 1) // every + every


class Run:
    """A Simulation with its particle positions and ids fixed: the state a
    FieldEnergy forward starts from is (velocities, charges, E, B)."""

    def __init__(self, sim, base, first_ids, cycles, every, weight_e=1.0, weight_b=0.0):
        """base: per species the host arrays of the initial particles (their
        rx, ry, rz, cell and id are kept; u, v, w and q come from the
        forward's tensors); first_ids[s]: the species' ids fill
        [first, first + count)."""
        self.sim = sim
        self.base = base
        self.first_ids = list(first_ids)
        self.counts = [len(b["q"]) for b in base]
        self.cycles = int(cycles)
        self.every = max(1, int(every))
        self.weight_e, self.weight_b = float(weight_e), float(weight_b)
        sim.checkpoint_reserve(_slots(self.cycles, self.every))
        self.adjoint = sim.adjoint()
        self.iterations = []

    def load(self, v, q, e, b):
        offset = 0
        for s, base in enumerate(self.base):
            n = self.counts[s]
            a = {k: base[k] for k in ("rx", "ry", "rz", "cell", "id")}
            a["u"], a["v"], a["w"] = (np.ascontiguousarray(v[offset:offset + n, c]) for c in range(3))
            a["q"] = np.ascontiguousarray(q[offset:offset + n])
            self.sim.set_particles(s, a)
            offset += n
        self.sim.set_field(R.FIELD_E, e)
        self.sim.set_field(R.FIELD_B, b)
        self.sim.initialize()

    def forward(self, v, q, e, b):
        """Cycles with the start of every segment saved; returns the energy."""
        self.load(v, q, e, b)
        segments = range(0, self.cycles, self.every)
        for k, start in enumerate(segments):
            self.sim.checkpoint_save(k)
            for _ in range(min(self.every, self.cycles - start)):
                self.iterations.append(self.sim.cycle(1)["iterations"])
        energy = self.sim.energy()
        # The adjoint of L at the final state, taken now: backward() walks
        # back from it (the context may be run again before).
        self.adjoint.seed_field_energy(self.weight_e, self.weight_b, self.first_ids)
        return self.weight_e * energy[0] + self.weight_b * energy[1]

    def backward(self):
        """Adjoint of the last forward: (vbar [N, 3], qbar [N], Ebar [3N], Bbar [3C])
        in the order of the forward's inputs."""
        sim, adj = self.sim, self.adjoint
        outer = (self.cycles + self.every - 1) // self.every
        inner0 = outer
        for k in range(outer - 1, -1, -1):
            start = k * self.every
            length = min(self.every, self.cycles - start)
            sim.checkpoint_restore(k)
            for m in range(length):
                sim.checkpoint_save(inner0 + m)
                if m + 1 < length:
                    sim.cycle(1)
            for m in range(length - 1, -1, -1):
                sim.checkpoint_restore(inner0 + m)
                adj.cycle()
        vbar = np.zeros((sum(self.counts), 3))
        qbar = np.zeros(sum(self.counts))
        offset = 0
        for s, base in enumerate(self.base):
            p = adj.particles(s)
            # The adjoint comes back in the sorted order of the first state;
            # the inputs are in the order of base.
            where = np.empty(len(p["ids"]), np.int64)
            where[(p["ids"] - np.uint64(self.first_ids[s])).astype(np.int64)] = np.arange(len(p["ids"]))
            idx = where[(base["id"] - np.uint64(self.first_ids[s])).astype(np.int64)]
            n = self.counts[s]
            vbar[offset:offset + n] = p["vbar"][:, idx].T
            qbar[offset:offset + n] = p["qbar"][idx]
            offset += n
        ebar = adj.field(R.FIELD_E)
        # The adjoint is dL/dE of the unique nodes, held at each periodic
        # copy; the input vector's duplicate nodes carry none of it.
        nx, ny, nz = sim.cells
        dup = np.ones((nz + 1, ny + 1, nx + 1), bool)
        dup[:nz, :ny, :nx] = False
        ebar = ebar.reshape(3, -1)
        ebar[:, dup.ravel()] = 0.0
        return vbar, qbar, ebar.ravel(), adj.field(R.FIELD_B)


class FieldEnergy(torch.autograd.Function):
    """L = weight_e W_E + weight_b W_B after the run's cycles, as a function
    of the initial velocities v [N, 3], charges q [N], E [3N] and B [3C]
    (float64 host tensors)."""

    @staticmethod
    def forward(ctx, v, q, e, b, run):
        value = run.forward(v.detach().numpy(), q.detach().numpy(), e.detach().numpy(), b.detach().numpy())
        ctx.run = run
        return torch.tensor(value, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        vbar, qbar, ebar, bbar = ctx.run.backward()
        g = float(grad)
        as_t = lambda a: torch.from_numpy(np.ascontiguousarray(a * g))  # noqa: E731
        return as_t(vbar), as_t(qbar), as_t(ebar), as_t(bbar), None


# ---- configurations -------------------------------------------------------

def node_x(cells, lengths):
    """x of every node (x fastest, duplicates included)."""
    nx, ny, nz = cells
    x = np.arange(nx + 1) * (lengths[0] / nx)
    return np.broadcast_to(x, (nz + 1, ny + 1, nx + 1)).ravel()


def particle_x(base, cells, lengths):
    nx = cells[0]
    ex = nx + 4
    i = base["cell"].astype(np.int64) % ex - 2
    return (i + base["rx"]) * (lengths[0] / nx)


class TwoStream:
    """1D two-stream: beams +-v0 with thermal vth, ppc per beam, TwoStream
    density seed eps at `mode` (phase 0), ions omitted (an immobile uniform
    background does not enter the field equations of a periodic box)."""

    def __init__(self, cells=1024, ppc=1000, mode=8, v0=0.2, vth=0.01, eps=1e-3, dt=0.1, cycles=200, every=20,
                 mass_kernel=0, gmres_tolerance=1e-13):
        k = 0.2 * mode / 8
        length = TWO_PI * mode / k
        dx = length / cells
        self.cells = (cells, 1, 1)
        self.lengths = (length, dx, dx)
        self.k = TWO_PI * mode / length
        self.v0, self.vth, self.eps, self.dt = v0, vth, eps, dt
        self.rho = (0.5, 0.5)
        self.ppc = ppc
        n = cells * ppc
        self.sim = R.Simulation(self.cells, self.lengths, [(-1.0, n), (-1.0, n)], dt=dt, theta=0.5,
                                gmres_tolerance=gmres_tolerance, gmres_restart=30, gmres_max_iterations=400,
                                mass_kernel=mass_kernel)
        base = []
        for s, sign in enumerate((1.0, -1.0)):
            self.sim.load_uniform(s, (ppc, 1, 1), rho_init=self.rho[s], drift=(sign * v0, 0.0, 0.0),
                                  thermal=(vth, vth, vth), first_id=s * n)
            base.append(self.sim.get_particles(s))
        self.base = base
        self.run = Run(self.sim, base, [0, n], cycles, every, weight_e=1.0, weight_b=0.0)
        self.x = np.concatenate([particle_x(b, self.cells, self.lengths) for b in base])
        self.sign = np.concatenate([np.full(n, 1.0), np.full(n, -1.0)])
        # The loaded velocities and charges without drift and seed.
        self.thermal_v = np.concatenate([np.stack([b["u"] - sg * v0, b["v"], b["w"]], axis=1)
                                         for b, sg in zip(base, (1.0, -1.0))])
        self.q0 = np.concatenate([b["q"] for b in base])
        self.xn = node_x(self.cells, self.lengths)
        self.ncells = cells

    def state(self, v0, eps):
        """(v, q, E, B) as torch expressions of the parameters."""
        sign = torch.from_numpy(self.sign)
        thermal = torch.from_numpy(self.thermal_v)
        vx = thermal[:, 0] + sign * v0
        v = torch.stack([vx, thermal[:, 1], thermal[:, 2]], dim=1)
        q = torch.from_numpy(self.q0) * (1.0 + eps * torch.cos(torch.from_numpy(self.k * self.x)))
        ex = -(sum(self.rho) * eps / self.k) * torch.sin(torch.from_numpy(self.k * self.xn))
        nodes = ex.shape[0]
        e = torch.cat([ex, torch.zeros(2 * nodes, dtype=torch.float64)])
        b = torch.zeros(3 * self.ncells, dtype=torch.float64)
        return v, q, e, b

    def energy(self, v0, eps):
        return FieldEnergy.apply(*self.state(v0, eps), self.run)

    def close(self):
        self.run.adjoint.close()
        self.sim.close()


class Weibel:
    """2D Weibel: electrons with thermal speeds (ratio vth, vth, ratio vth),
    ions q/m = 1/mass_ratio at thermal vth_i, ppc each; d(W_B)/d(ratio)."""

    def __init__(self, cells=(512, 512), dx=0.1, ppc=(10, 10), vth_par=0.1, ratio=3.0, mass_ratio=100.0,
                 vth_i=0.01, dt=0.05, cycles=50, every=10, mass_kernel=0, gmres_tolerance=1e-13):
        self.cells = (cells[0], cells[1], 1)
        self.lengths = (cells[0] * dx, cells[1] * dx, dx)
        per = cells[0] * cells[1] * ppc[0] * ppc[1]
        self.sim = R.Simulation(self.cells, self.lengths, [(-1.0, per), (1.0 / mass_ratio, per)], dt=dt, theta=0.5,
                                gmres_tolerance=gmres_tolerance, gmres_restart=30, gmres_max_iterations=400,
                                mass_kernel=mass_kernel)
        self.vth_par, self.ratio = vth_par, ratio
        base = []
        thermals = [(ratio * vth_par, vth_par, ratio * vth_par), (vth_i, vth_i, vth_i)]
        for s in range(2):
            self.sim.load_uniform(s, (ppc[0], ppc[1], 1), rho_init=1.0, thermal=thermals[s], first_id=s * per)
            base.append(self.sim.get_particles(s))
        self.base = base
        self.run = Run(self.sim, base, [0, per], cycles, every, weight_e=0.0, weight_b=1.0)
        # Unit normals of the electrons' loaded velocities along x and z.
        self.z = np.stack([base[0]["u"] / (ratio * vth_par), base[0]["w"] / (ratio * vth_par)], axis=1)
        self.ions = np.stack([base[1]["u"], base[1]["v"], base[1]["w"]], axis=1)
        self.v_par = base[0]["v"]
        self.q = np.concatenate([b["q"] for b in base])
        nx, ny, _ = self.cells
        self.nodes = (nx + 1) * (ny + 1) * 2
        self.ncells = nx * ny

    def state(self, ratio):
        z = torch.from_numpy(self.z)
        u = z[:, 0] * (ratio * self.vth_par)
        w = z[:, 1] * (ratio * self.vth_par)
        ve = torch.stack([u, torch.from_numpy(self.v_par), w], dim=1)
        v = torch.cat([ve, torch.from_numpy(self.ions)], dim=0)
        q = torch.from_numpy(self.q)
        e = torch.zeros(3 * self.nodes, dtype=torch.float64)
        b = torch.zeros(3 * self.ncells, dtype=torch.float64)
        return v, q, e, b

    def energy(self, ratio):
        return FieldEnergy.apply(*self.state(ratio), self.run)

    def close(self):
        self.run.adjoint.close()
        self.sim.close()


def gradient(model, *params):
    """(value, [dL/dparam]) through FieldEnergy."""
    ts = [torch.tensor(float(p), dtype=torch.float64, requires_grad=True) for p in params]
    value = model.energy(*ts)
    value.backward()
    return float(value), [float(t.grad) for t in ts]


def central_difference(model, params, index, h):
    """(L(p + h e_i) - L(p - h e_i)) / 2h with the same forward."""
    up = list(params)
    down = list(params)
    up[index] += h
    down[index] -= h
    with torch.no_grad():
        lp = float(model.energy(*[torch.tensor(float(x), dtype=torch.float64) for x in up]))
        lm = float(model.energy(*[torch.tensor(float(x), dtype=torch.float64) for x in down]))
    return (lp - lm) / (2 * h)


__all__ = ["FieldEnergy", "Run", "TwoStream", "Weibel", "gradient", "central_difference"]
