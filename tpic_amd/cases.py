"""Seeded initial states of the uniform-plasma runs (host side, numpy).

The dense uniform path of the reference loader
(include/tpic/initialization/particles/particle_loading_device.hpp:69-159,
src/initialization/prepared_initialization.cpp:37-85): load index -> cell
index // ppc (x fastest) and subcell index % ppc; fraction (sub + U)/ppc;
velocity drift + thermal * N; charge (sign(q/m) V/ppc) * (rho/4pi); key
(i+2) + (nx+4)((j+2) + (ny+4)(k+2)); id = first_id + index. The random
numbers come from numpy's Philox bit generator with the reference's seed
formula 12345 + 100*species, so a state is reproducible; they are not the
hipRAND XORWOW draws of the reference.
"""
import math

import numpy as np

FOUR_PI = 12.566370614359172953850573533118


def encode(cells, i, j, k):
    nx, ny, _ = cells
    return (np.asarray(i) + 2) + (nx + 4) * ((np.asarray(j) + 2) + (ny + 4) * (np.asarray(k) + 2))


def species_seed(species):
    """prepared_initialization.cpp:46-47 on one rank."""
    return 12345 + 100 * species


def uniform_maxwellian(cells, lengths, ppc, q_over_m, rho=1.0, drift=(0.0, 0.0, 0.0),
                       thermal=(0.0, 0.0, 0.0), seed=12345, first_id=0):
    """SoA arrays of one species of a uniform Maxwellian plasma."""
    nx, ny, nz = (int(c) for c in cells)
    px, py, pz = (int(p) for p in ppc)
    per_cell = px * py * pz
    n = nx * ny * nz * per_cell
    spacing = [float(lengths[a]) / float(cells[a]) for a in range(3)]
    rng = np.random.Generator(np.random.Philox(seed))
    index = np.arange(n, dtype=np.int64)
    cell = index // per_cell
    sub = index % per_cell
    ijk = [cell % nx, (cell // nx) % ny, cell // (nx * ny)]
    sxyz = [sub % px, (sub // px) % py, sub // (px * py)]
    offsets = rng.random((3, n))
    frac = []
    for a, p in enumerate((px, py, pz)):
        f = (sxyz[a] + offsets[a]) * (1.0 / p)
        carry = f == 1.0
        f[carry] = 0.0
        ijk[a] = np.where(carry, (ijk[a] + 1) % cells[a], ijk[a])
        frac.append(f)
    normals = rng.standard_normal((3, n))
    volume = (spacing[0] * spacing[1]) * spacing[2]
    q_factor = (1.0 if q_over_m > 0.0 else -1.0) * volume / float(per_cell)
    charge = (q_factor * (rho / FOUR_PI)) * 1.0
    return {
        "rx": frac[0], "ry": frac[1], "rz": frac[2],
        "u": drift[0] + thermal[0] * normals[0],
        "v": drift[1] + thermal[1] * normals[1],
        "w": drift[2] + thermal[2] * normals[2],
        "q": np.full(n, charge),
        "cell": encode(cells, ijk[0], ijk[1], ijk[2]).astype(np.int32),
        "id": (first_id + index).astype(np.uint64),
    }


MASK64 = (1 << 64) - 1


def splitmix64(x):
    """One output of the splitmix64 generator (S. Vigna) at state x."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def spectrum_modes(modes, seed):
    """The seeded k-spectrum: per mode the integer wave vector (1..3 per
    axis, in box wavelengths), the phase and the unit polarisation
    perpendicular to k (divergence-free in the continuum). The same
    definition is in the deck driver (tpic_amd/csrc/driver/deck.cpp)."""
    out = []
    for m in range(int(modes)):
        h = splitmix64((int(seed) * 1000003 + m) & MASK64)
        k = [1.0 + float(h % 3), 1.0 + float((h >> 8) % 3), 1.0 + float((h >> 16) % 3)]
        phase = 2.0 * math.pi * float((h >> 32) & 0xFFFFFFFF) / 4294967296.0
        g = splitmix64(h)
        r = [float((g >> (21 * a)) & 0x1FFFFF) / 2097152.0 * 2.0 - 1.0 for a in range(3)]
        pol = [r[1] * k[2] - r[2] * k[1], r[2] * k[0] - r[0] * k[2], r[0] * k[1] - r[1] * k[0]]
        norm = math.sqrt(pol[0] * pol[0] + pol[1] * pol[1] + pol[2] * pol[2])
        if norm == 0.0:
            pol, norm = [0.0, -k[2], k[1]], math.sqrt(k[1] * k[1] + k[2] * k[2])
        out.append((k, phase, [p_ / norm for p_ in pol]))
    return out


def uniform_b(cells, b0, amplitude=0.0, modes=4, seed=7):
    """Cell-centred B = B0 plus a seeded k-spectrum of transverse
    perturbations of relative amplitude `amplitude` (of |B0|; compact cell
    layout): each mode adds amplitude |B0| pol cos(2 pi k . s + phase) at the
    cell centres s = ((i + 1/2)/nx, (j + 1/2)/ny, (k + 1/2)/nz)."""
    nx, ny, nz = (int(c) for c in cells)
    c = nx * ny * nz
    b = np.zeros(3 * c)
    for a in range(3):
        b[a * c:(a + 1) * c] = b0[a]
    if amplitude:
        i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        x = (i.transpose(2, 1, 0).ravel() + 0.5) / nx
        y = (j.transpose(2, 1, 0).ravel() + 0.5) / ny
        z = (k.transpose(2, 1, 0).ravel() + 0.5) / nz
        mag = math.sqrt(sum(v * v for v in b0)) or 1.0
        for kvec, phase, pol in spectrum_modes(modes, seed):
            arg = 2.0 * math.pi * (kvec[0] * x + kvec[1] * y + kvec[2] * z) + phase
            wave = np.cos(arg)
            for a in range(3):
                b[a * c:(a + 1) * c] += (amplitude * mag * pol[a]) * wave
    return b


# The uniform-plasma workloads (units c = 1, lengths in d_i, time 1/omega_pi).
UNIFORM_16 = dict(cells=(16, 16, 16), lengths=(4.0, 4.0, 4.0), ppc=(4, 4, 4),
                  q_over_m=(-64.0, 1.0), thermal=(0.045, 0.0056), b0=(0.0, 0.0, 0.01),
                  dt=0.1, theta=0.5, cycles=3, gmres_tolerance=1e-10)
UNIFORM_128 = dict(cells=(128, 128, 128), lengths=(32.0, 32.0, 32.0), ppc=(4, 4, 4),
                   q_over_m=(-64.0, 1.0), thermal=(0.045, 0.0056), b0=(0.0, 0.0, 0.01),
                   perturbation=1e-3, dt=0.1, theta=0.5, cycles=20, gmres_tolerance=1e-10)


def counts(case):
    nx, ny, nz = case["cells"]
    px, py, pz = case["ppc"]
    return [nx * ny * nz * px * py * pz] * len(case["q_over_m"])
