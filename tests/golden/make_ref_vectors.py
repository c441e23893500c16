"""Writes tests/golden/ref_vectors.json: outputs of the reference's own
host-callable functions (oracle/_ref/libtpic_ref.so, built from the reference
headers by `make -C oracle ref`) on fixed, seeded inputs. Values are stored as
float.hex() strings so the comparison is bit for bit.

Functions: make_cic_weights (include/tpic/ecsim/particles/cic.hpp:24-36),
ParticleResponse (particle_response.hpp:17-34), mass_pair_stencil
(mass_matrix_stencil.hpp:28-36), node_cell_curl (node_cell_curl.hpp:93-103),
cell_to_node_average (cell_to_node.hpp:19-32), apply_mass_matrix_node
(mass_matrix_stencil.hpp:120-149).
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

D = ctypes.POINTER(ctypes.c_double)


def hexes(a):
    return [float(x).hex() for x in np.asarray(a, dtype=np.float64).ravel()]


def ghost_nodes(n, compact):
    """Compact node component -> one-ghost array with periodic ghosts."""
    nx, ny, nz = n
    c = compact.reshape(nz + 1, ny + 1, nx + 1)
    out = np.zeros((nz + 3, ny + 3, nx + 3))
    out[1:-1, 1:-1, 1:-1] = c
    out[:, :, 0] = out[:, :, nx - 1 + 1]
    out[:, :, nx + 2] = out[:, :, 2]
    out[:, 0, :] = out[:, ny - 1 + 1, :]
    out[:, ny + 2, :] = out[:, 2, :]
    out[0, :, :] = out[nz - 1 + 1, :, :]
    out[nz + 2, :, :] = out[2, :, :]
    return np.ascontiguousarray(out.ravel())


def ghost_cells(n, compact):
    nx, ny, nz = n
    c = compact.reshape(nz, ny, nx)
    out = np.zeros((nz + 2, ny + 2, nx + 2))
    out[1:-1, 1:-1, 1:-1] = c
    out[:, :, 0] = out[:, :, nx]
    out[:, :, nx + 1] = out[:, :, 1]
    out[:, 0, :] = out[:, ny, :]
    out[:, ny + 1, :] = out[:, 1, :]
    out[0, :, :] = out[nz, :, :]
    out[nz + 1, :, :] = out[1, :, :]
    return np.ascontiguousarray(out.ravel())


def main():
    R = oracle.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libtpic_ref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20261202)
    out = {"source": "reference host functions compiled from the reference headers"}
    fractions = [0.0, -0.0, 5e-324, 2.2250738585072014e-308, 0.125, 0.5, 0.75, 0.987654,
                 np.nextafter(1.0, 0.0)]
    cic = []
    for x in fractions:
        for y in fractions[::3]:
            for z in fractions[::4]:
                w = np.zeros(8)
                R.ref_cic_weights(x, y, z, oracle.ptr(w))
                cic.append({"r": hexes([x, y, z]), "w": hexes(w)})
    out["cic"] = cic
    resp = []
    for _ in range(40):
        b = rng.normal(size=3) * rng.choice([1e-3, 1.0, 30.0])
        boc = rng.choice([-0.17, 0.03, -3.2, 0.5])
        a = np.zeros(9)
        n2 = np.zeros(1)
        R.ref_response(b[0], b[1], b[2], boc, oracle.ptr(a), oracle.ptr(n2))
        resp.append({"b": hexes(b), "beta_over_c": float(boc).hex(), "a": hexes(a), "norm2": float(n2[0]).hex()})
    out["response"] = resp
    out["pair_stencil"] = [[int(R.ref_mass_pair_stencil(a, b)) for b in range(8)] for a in range(8)]
    n = (4, 3, 5)
    spacing = np.array([0.25, 0.5, 0.2])
    nodes = (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
    cells = n[0] * n[1] * n[2]
    ni = (ctypes.c_int * 3)(*n)
    v = rng.normal(size=3 * nodes)
    # make the compact node vector periodic-consistent (duplicates equal)
    vv = v.reshape(3, n[2] + 1, n[1] + 1, n[0] + 1)
    vv[:, :, :, n[0]] = vv[:, :, :, 0]
    vv[:, :, n[1], :] = vv[:, :, 0, :]
    vv[:, n[2], :, :] = vv[:, 0, :, :]
    v = vv.ravel().copy()
    gv = [ghost_nodes(n, v[c * nodes:(c + 1) * nodes]) for c in range(3)]
    curl_cells = np.zeros(3 * cells)
    R.ref_curl_n2c(ni, oracle.ptr(spacing), oracle.ptr(gv[0]), oracle.ptr(gv[1]), oracle.ptr(gv[2]),
                   oracle.ptr(curl_cells))
    bcell = rng.normal(size=3 * cells)
    gb = [ghost_cells(n, bcell[c * cells:(c + 1) * cells]) for c in range(3)]
    curl_nodes = np.zeros(3 * nodes)
    R.ref_curl_c2n(ni, oracle.ptr(spacing), oracle.ptr(gb[0]), oracle.ptr(gb[1]), oracle.ptr(gb[2]),
                   oracle.ptr(curl_nodes))
    c2n = np.zeros(3 * nodes)
    for c in range(3):
        tmp = np.zeros(nodes)
        R.ref_cell_to_node(ni, oracle.ptr(gb[c]), oracle.ptr(tmp))
        c2n[c * nodes:(c + 1) * nodes] = tmp
    # Mass apply: a random folded matrix (periodic images in the ghost records).
    g = oracle.make_grid(n, n * spacing)
    mass = np.zeros(oracle.vertex_count(g) * 126)
    ex, ey, ez = n[0] + 3, n[1] + 3, n[2] + 3
    interior = rng.normal(size=(n[2] + 1, n[1] + 1, n[0] + 1, 126))
    interior[:, :, n[0]] = interior[:, :, 0]
    interior[:, n[1]] = interior[:, 0]
    interior[n[2]] = interior[0]
    m4 = mass.reshape(ez, ey, ex, 126)
    m4[1:-1, 1:-1, 1:-1] = interior
    oracle.fold_mass(g, mass)  # ghost phase fills ghost records (interior sums double, fine)
    inv_vol = 1.0 / (spacing[0] * spacing[1] * spacing[2])
    mv = np.zeros(3 * nodes)
    R.ref_mass_apply(ni, oracle.ptr(mass), oracle.ptr(gv[0]), oracle.ptr(gv[1]), oracle.ptr(gv[2]), inv_vol,
                     oracle.ptr(mv))
    out["fields"] = {
        "cells": list(n), "spacing": hexes(spacing), "node_vector": hexes(v), "cell_vector": hexes(bcell),
        "curl_n2c": hexes(curl_cells), "curl_c2n": hexes(curl_nodes), "cell_to_node": hexes(c2n),
        "mass_seed": 20261202, "inverse_volume": float(inv_vol).hex(), "mass_apply": hexes(mv),
    }
    np.save(os.path.join(HERE, "ref_mass_matrix.npy"), mass)
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
