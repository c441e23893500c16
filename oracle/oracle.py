"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own
host-callable functions (_ref/libtpic_ref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the tpic_amd package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

D = ctypes.POINTER(ctypes.c_double)
U32 = ctypes.POINTER(ctypes.c_uint32)
I32 = ctypes.POINTER(ctypes.c_int32)
U64 = ctypes.POINTER(ctypes.c_uint64)


class Grid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("spacing", ctypes.c_double * 3),
                ("particle_periodic", ctypes.c_int * 3), ("field_faces", (ctypes.c_int * 2) * 3)]


FACE_PERIODIC, FACE_PEC, FACE_PMC = 0, 1, 2


class Harris(ctypes.Structure):
    _fields_ = [("density", ctypes.c_int), ("is_double", ctypes.c_int), ("varying_thermal", ctypes.c_int),
                ("ly", ctypes.c_double), ("delta", ctypes.c_double), ("ampere_drift", ctypes.c_double),
                ("center_x", ctypes.c_double), ("center_y", ctypes.c_double), ("curl_delta", ctypes.c_double),
                ("sheet", ctypes.c_double), ("mode", ctypes.c_double), ("mode_kx", ctypes.c_double),
                ("mode_ky", ctypes.c_double), ("hump", ctypes.c_double), ("hump_kx", ctypes.c_double),
                ("hump_ky", ctypes.c_double), ("hump_inverse_wx2", ctypes.c_double),
                ("hump_inverse_wy2", ctypes.c_double)]


def harris(**kw):
    h = Harris()
    for k, v in kw.items():
        setattr(h, k, v)
    return h


class MoverParams(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_double), ("beta_over_c", ctypes.c_double),
                ("dt", ctypes.c_double), ("inverse_spacing", ctypes.c_double * 3)]


class Particles(ctypes.Structure):
    _fields_ = [("rx", D), ("ry", D), ("rz", D), ("u", D), ("v", D), ("w", D), ("q", D),
                ("cell", I32), ("id", U64), ("count", ctypes.c_uint32)]


class GmresResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("initial_residual", ctypes.c_double), ("residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double), ("relative_residual", ctypes.c_double)]


OPERATOR = ctypes.CFUNCTYPE(None, ctypes.c_void_p, D, D)


class Xorwow(ctypes.Structure):
    _fields_ = [("d", ctypes.c_uint32), ("x", ctypes.c_uint32 * 5), ("cached_normal", ctypes.c_double),
                ("has_cached", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        # TPIC_ORACLE_OPENMP=1 loads the OpenMP build (timing of the CPU
        # baseline only; its deposition and dot products round differently).
        name = "liboracle_omp.so" if os.environ.get("TPIC_ORACLE_OPENMP") == "1" else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built (run `make -C oracle`)" % name)
        _lib = ctypes.CDLL(path)
        L = _lib
        G = ctypes.POINTER(Grid)
        P = ctypes.POINTER(Particles)
        for name in ("or_node_count", "or_cell_count", "or_bin_count", "or_vertex_count"):
            getattr(L, name).argtypes = [G]
            getattr(L, name).restype = ctypes.c_int
        L.or_encode.argtypes = [G, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_encode.restype = ctypes.c_int32
        L.or_cic_weights.argtypes = [ctypes.c_double] * 3 + [D]
        L.or_response.argtypes = [ctypes.c_double] * 4 + [D, D]
        L.or_mass_pair_stencil.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_mass_pair_stencil.restype = ctypes.c_int
        L.or_move.argtypes = [G, ctypes.POINTER(MoverParams), P, D, D]
        L.or_move.restype = ctypes.c_uint
        L.or_sort.argtypes = [G, P, P, U32, U32, U32]
        L.or_sort.restype = ctypes.c_int
        L.or_deposit_current.argtypes = [G, P, U32, U32, D, ctypes.c_double, ctypes.c_double, D]
        L.or_deposit_current.restype = ctypes.c_uint
        L.or_deposit_mass.argtypes = [G, P, U32, U32, D, ctypes.c_double, D]
        L.or_deposit_mass.restype = ctypes.c_uint
        L.or_fold_current.argtypes = [G, D]
        L.or_deposit_moments.argtypes = [G, P, ctypes.c_double, ctypes.c_double, ctypes.c_uint, D]
        L.or_deposit_moments.restype = ctypes.c_uint
        L.or_fold_mass.argtypes = [G, D]
        L.or_curl_n2c.argtypes = [G, D, D]
        L.or_curl_c2n.argtypes = [G, D, D]
        L.or_cell_to_node.argtypes = [G, D, D]
        L.or_mass_apply.argtypes = [G, D, D, ctypes.c_double, D]
        L.or_maxwell_apply.argtypes = [G, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D]
        L.or_maxwell_rhs.argtypes = [G, D, D, D, ctypes.c_double, ctypes.c_double, D]
        SM = ctypes.POINTER(Smoothing)
        L.or_smoothing_active.argtypes = [SM]
        L.or_smooth.argtypes = [G, SM, D, D]
        L.or_maxwell_apply_smoothed.argtypes = [G, SM, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D]
        L.or_sim_set_smoothing.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.or_gmres.argtypes = [ctypes.c_int, OPERATOR, ctypes.c_void_p, D, D, ctypes.c_int,
                               ctypes.c_int, ctypes.c_double]
        L.or_gmres.restype = GmresResult
        L.or_sim_create.argtypes = [G, ctypes.c_int, D, U32, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_sim_create.restype = ctypes.c_void_p
        L.or_sim_destroy.argtypes = [ctypes.c_void_p]
        L.or_sim_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_sim_field.restype = D
        L.or_sim_particles.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_sim_particles.restype = P
        L.or_sim_bins.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.or_sim_bins.restype = U32
        for name in ("or_sim_initialize", "or_sim_sort"):
            getattr(L, name).restype = ctypes.c_int
        L.or_sim_initialize.argtypes = [ctypes.c_void_p]
        L.or_sim_sort.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_sim_solve.argtypes = [ctypes.c_void_p, ctypes.POINTER(GmresResult)]
        L.or_sim_solve.restype = ctypes.c_int
        L.or_sim_move.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_sim_move.restype = ctypes.c_uint
        L.or_sim_advance_fields.argtypes = [ctypes.c_void_p]
        L.or_sim_deposit.argtypes = [ctypes.c_void_p]
        L.or_sim_deposit.restype = ctypes.c_uint
        L.or_sim_cycle.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(GmresResult)]
        L.or_sim_cycle.restype = ctypes.c_int
        L.or_sim_energy.argtypes = [ctypes.c_void_p, D]
        L.or_sim_scalars.argtypes = [ctypes.c_void_p, D]
        L.or_sim_validation.argtypes = [ctypes.c_void_p]
        L.or_sim_validation.restype = ctypes.c_uint
        X = ctypes.POINTER(Xorwow)
        L.or_xorwow_init.argtypes = [X, ctypes.c_uint64, ctypes.c_uint64]
        L.or_xorwow_next.argtypes = [X]
        L.or_xorwow_next.restype = ctypes.c_uint32
        L.or_xorwow_uniform_double.argtypes = [X]
        L.or_xorwow_uniform_double.restype = ctypes.c_double
        L.or_xorwow_normal2_double.argtypes = [X, D]
        L.or_xorwow_normal_double.argtypes = [X]
        L.or_xorwow_normal_double.restype = ctypes.c_double
        L.or_load_uniform.argtypes = [G, ctypes.POINTER(ctypes.c_int * 3), ctypes.c_double, D, D, ctypes.c_int,
                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, P]
        L.or_load_seeded.argtypes = [G, ctypes.POINTER(ctypes.c_int * 3), ctypes.c_double, D, D, ctypes.c_int,
                                     ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, D, P]
        L.or_load_profile.argtypes = [G, ctypes.POINTER(ctypes.c_int * 3), ctypes.c_double, D, D, ctypes.c_int,
                                      ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, D,
                                      ctypes.POINTER(Harris), P]
        L.or_harris_curl.argtypes = [ctypes.POINTER(Harris), ctypes.c_double, ctypes.c_double]
        L.or_harris_curl.restype = ctypes.c_double
        L.or_center_velocity.argtypes = [P, D]
        L.or_center_velocity.restype = ctypes.c_int
    return _lib


def ref():
    """The reference's own functions, or None when they were not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libtpic_ref.so")
        if not os.path.exists(path):
            return None
        _ref = ctypes.CDLL(path)
        R = _ref
        R.ref_cic_weights.argtypes = [ctypes.c_double] * 3 + [D]
        R.ref_response.argtypes = [ctypes.c_double] * 4 + [D, D]
        R.ref_mass_pair_stencil.argtypes = [ctypes.c_int, ctypes.c_int]
        R.ref_mass_pair_stencil.restype = ctypes.c_int
        IP = ctypes.POINTER(ctypes.c_int)
        R.ref_mass_apply.argtypes = [IP, D, D, D, D, ctypes.c_double, D]
        R.ref_curl_n2c.argtypes = [IP, D, D, D, D, D]
        R.ref_curl_c2n.argtypes = [IP, D, D, D, D, D]
        R.ref_cell_to_node.argtypes = [IP, D, D]
    return _ref


def ptr(a, kind=D):
    return a.ctypes.data_as(kind)


def make_grid(n, lengths, particle_periodic=(1, 1, 1), field_faces=None):
    """field_faces: per axis (lower, upper) of FACE_PERIODIC, FACE_PEC, FACE_PMC
    (default: periodic fields)."""
    g = Grid()
    for a in range(3):
        g.n[a] = int(n[a])
        g.spacing[a] = float(lengths[a]) / float(n[a])
        g.particle_periodic[a] = int(particle_periodic[a])
        if field_faces is not None:
            g.field_faces[a][0] = int(field_faces[a][0])
            g.field_faces[a][1] = int(field_faces[a][1])
    return g


def node_count(g):
    return (g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1)


def cell_count(g):
    return g.n[0] * g.n[1] * g.n[2]


def bin_count(g):
    return (g.n[0] + 4) * (g.n[1] + 4) * (g.n[2] + 4)


def vertex_count(g):
    return (g.n[0] + 3) * (g.n[1] + 3) * (g.n[2] + 3)


def encode(g, i, j, k):
    ex, ey = g.n[0] + 4, g.n[1] + 4
    return (np.asarray(i) + 2) + ex * ((np.asarray(j) + 2) + ey * (np.asarray(k) + 2))


class ParticleArrays:
    """Owned numpy SoA arrays viewed as an or_particles struct."""

    NAMES = ("rx", "ry", "rz", "u", "v", "w", "q")

    def __init__(self, count=0, source=None):
        if source is not None:
            count = len(source["cell"])
        self.count = int(count)
        self.a = {n: np.zeros(self.count, np.float64) for n in self.NAMES}
        self.a["cell"] = np.zeros(self.count, np.int32)
        self.a["id"] = np.zeros(self.count, np.uint64)
        if source is not None:
            for k in self.a:
                self.a[k][:] = source[k]

    def struct(self):
        s = Particles()
        for n in self.NAMES:
            setattr(s, n, ptr(self.a[n]))
        s.cell = ptr(self.a["cell"], I32)
        s.id = ptr(self.a["id"], U64)
        s.count = self.count
        return s

    def copy(self):
        return ParticleArrays(source=self.a)


def move(g, params, particles, e_nodes, b_nodes):
    s = particles.struct()
    flags = lib().or_move(ctypes.byref(g), ctypes.byref(params), ctypes.byref(s),
                          ptr(np.ascontiguousarray(e_nodes)), ptr(np.ascontiguousarray(b_nodes)))
    return int(flags)


class XorwowStream:
    """rocRAND's XORWOW (hipRAND's default device generator) restated."""

    def __init__(self, seed, subsequence):
        self.s = Xorwow()
        lib().or_xorwow_init(ctypes.byref(self.s), int(seed), int(subsequence))

    def next(self):
        return int(lib().or_xorwow_next(ctypes.byref(self.s)))

    def uniform_double(self):
        return float(lib().or_xorwow_uniform_double(ctypes.byref(self.s)))

    def normal2_double(self):
        out = np.zeros(2)
        lib().or_xorwow_normal2_double(ctypes.byref(self.s), ptr(out))
        return out

    def normal_double(self):
        return float(lib().or_xorwow_normal_double(ctypes.byref(self.s)))


def load_seed(species, rank=0, max_species=10):
    """prepared_initialization.cpp:46-47."""
    return 12345 + 100 * (rank * max_species + species)


def load_uniform(g, ppc, rho_init, drift, thermal, quiet_start, seed, first_id, q_over_m, density_seed=None,
                 harris_profile=None):
    """The reference's Uniform-case load of one species (cells * ppc particles);
    density_seed = (k, phase, amplitude) applies the TwoStream density seed,
    harris_profile (a Harris structure) the Harris sheet profile."""
    n = cell_count(g) * int(ppc[0]) * int(ppc[1]) * int(ppc[2])
    out = ParticleArrays(n)
    s = out.struct()
    p3 = (ctypes.c_int * 3)(*[int(v) for v in ppc])
    mode = None if density_seed is None else np.asarray(density_seed, np.float64)
    lib().or_load_profile(ctypes.byref(g), ctypes.byref(p3), float(rho_init),
                          ptr(np.asarray(drift, np.float64)), ptr(np.asarray(thermal, np.float64)),
                          1 if quiet_start else 0, int(seed), int(first_id), float(q_over_m),
                          None if mode is None else ptr(mode),
                          None if harris_profile is None else ctypes.byref(harris_profile), ctypes.byref(s))
    return out


def harris_curl(h, x, y):
    return lib().or_harris_curl(ctypes.byref(h), float(x), float(y))


def center_velocity(particles, drift):
    """CenterInitialVelocity of one species with a uniform drift, in place;
    returns the oracle's non-finite flag."""
    s = particles.struct()
    return lib().or_center_velocity(ctypes.byref(s), ptr(np.asarray(drift, np.float64)))


def mover_params(q_over_m, dt, c, spacing):
    p = MoverParams()
    beta = q_over_m * dt / 2
    p.beta = beta
    p.beta_over_c = beta / c
    p.dt = dt
    for a in range(3):
        p.inverse_spacing[a] = 1.0 / spacing[a]
    return p


def sort(g, particles):
    out = ParticleArrays(particles.count)
    perm = np.zeros(max(particles.count, 1), np.uint32)
    counts = np.zeros(bin_count(g), np.uint32)
    offsets = np.zeros(bin_count(g) + 1, np.uint32)
    si, so = particles.struct(), out.struct()
    rc = lib().or_sort(ctypes.byref(g), ctypes.byref(si), ctypes.byref(so), ptr(perm, U32),
                       ptr(counts, U32), ptr(offsets, U32))
    if rc != 0:
        raise ValueError("oracle sort rejected the keys (rc=%d)" % rc)
    return out, perm[:particles.count], counts, offsets


def deposit_current(g, particles, counts, offsets, b_nodes, inverse_volume, beta_over_c, current=None):
    if current is None:
        current = np.zeros(3 * node_count(g))
    s = particles.struct()
    flags = lib().or_deposit_current(ctypes.byref(g), ctypes.byref(s), ptr(counts, U32), ptr(offsets, U32),
                                     ptr(np.ascontiguousarray(b_nodes)), inverse_volume, beta_over_c,
                                     ptr(current))
    return current, int(flags)


def deposit_mass(g, particles, counts, offsets, b_nodes, beta_over_c, mass=None):
    if mass is None:
        mass = np.zeros(vertex_count(g) * 126)
    s = particles.struct()
    flags = lib().or_deposit_mass(ctypes.byref(g), ctypes.byref(s), ptr(counts, U32), ptr(offsets, U32),
                                  ptr(np.ascontiguousarray(b_nodes)), beta_over_c, ptr(mass))
    return mass, int(flags)


def fold_current(g, current):
    lib().or_fold_current(ctypes.byref(g), ptr(current))
    return current


MOMENTS_ALL = 0x3FF


def deposit_moments(g, particles, q_over_m, inverse_volume, selected=MOMENTS_ALL):
    """Output moments of one species at nodes, shape (10, nodes): rho, Jx,
    Jy, Jz, PXX, PXY, PXZ, PYY, PYZ, PZZ (raw, rho without its 4 pi), folded
    and dual-volume scaled; returns (moments, invalid flags)."""
    out = np.zeros(10 * node_count(g))
    s = particles.struct()
    flags = lib().or_deposit_moments(ctypes.byref(g), ctypes.byref(s), float(q_over_m), float(inverse_volume),
                                     int(selected), ptr(out))
    return out.reshape(10, -1), int(flags)


def fold_mass(g, mass):
    lib().or_fold_mass(ctypes.byref(g), ptr(mass))
    return mass


def curl_n2c(g, nodes):
    out = np.zeros(3 * cell_count(g))
    lib().or_curl_n2c(ctypes.byref(g), ptr(np.ascontiguousarray(nodes)), ptr(out))
    return out


def curl_c2n(g, cells):
    out = np.zeros(3 * node_count(g))
    lib().or_curl_c2n(ctypes.byref(g), ptr(np.ascontiguousarray(cells)), ptr(out))
    return out


def cell_to_node(g, cells):
    out = np.zeros(3 * node_count(g))
    lib().or_cell_to_node(ctypes.byref(g), ptr(np.ascontiguousarray(cells)), ptr(out))
    return out


def mass_apply(g, mass, v, inverse_volume):
    out = np.zeros(3 * node_count(g))
    lib().or_mass_apply(ctypes.byref(g), ptr(mass), ptr(np.ascontiguousarray(v)), inverse_volume, ptr(out))
    return out


def maxwell_apply(g, mass, delta, current_factor, inverse_volume, v):
    out = np.zeros(3 * node_count(g))
    lib().or_maxwell_apply(ctypes.byref(g), ptr(mass), delta, current_factor, inverse_volume,
                           ptr(np.ascontiguousarray(v)), ptr(out))
    return out


SMOOTH_FACE7, SMOOTH_BINOMIAL27 = 0, 1


class Smoothing(ctypes.Structure):
    """or_smoothing: operator (SMOOTH_*), passes (SmoothNiter), alpha (Smooth)."""
    _fields_ = [("op", ctypes.c_int), ("iterations", ctypes.c_int), ("alpha", ctypes.c_double)]


def smooth(g, smoothing, v):
    """S(v) over the node vector v; smoothing = (op, iterations, alpha)."""
    sm = Smoothing(*smoothing)
    out = np.zeros(3 * node_count(g))
    lib().or_smooth(ctypes.byref(g), ctypes.byref(sm), ptr(np.ascontiguousarray(v, np.float64)), ptr(out))
    return out


def maxwell_apply_smoothed(g, smoothing, mass, delta, current_factor, inverse_volume, v):
    sm = Smoothing(*smoothing)
    out = np.zeros(3 * node_count(g))
    lib().or_maxwell_apply_smoothed(ctypes.byref(g), ctypes.byref(sm), ptr(mass), delta, current_factor,
                                    inverse_volume, ptr(np.ascontiguousarray(v)), ptr(out))
    return out


def gmres(apply, rhs, x0, restart=30, max_iterations=100, tol=1e-10):
    """Oracle GMRES on a Python operator apply(in_array) -> out_array."""
    n = len(rhs)

    def cb(_ctx, pin, pout):
        vin = np.ctypeslib.as_array(pin, shape=(n,))
        vout = np.ctypeslib.as_array(pout, shape=(n,))
        vout[:] = apply(vin.copy())

    fn = OPERATOR(cb)
    x = np.array(x0, dtype=np.float64, copy=True)
    r = lib().or_gmres(n, fn, None, ptr(np.ascontiguousarray(rhs, dtype=np.float64)), ptr(x),
                       restart, max_iterations, tol)
    return x, r


class Sim:
    """The oracle's whole-domain ECSIM cycle (periodic fields)."""

    FIELDS = {"E": 0, "B": 1, "Bn": 2, "J": 3, "theta": 4, "mass": 5, "particle": 6}

    def __init__(self, n, lengths, q_over_m, counts, dt, theta=0.5, c=1.0, tol=1e-10,
                 restart=30, max_iterations=100, particle_periodic=(1, 1, 1), field_faces=None, smoothing=None):
        self.g = make_grid(n, lengths, particle_periodic, field_faces)
        self.species = len(q_over_m)
        self.q_over_m = list(q_over_m)
        qom = np.asarray(q_over_m, np.float64)
        cnt = np.asarray(counts, np.uint32)
        self.h = lib().or_sim_create(ctypes.byref(self.g), self.species, ptr(qom), ptr(cnt, U32), dt, theta,
                                     c, tol, restart, max_iterations)
        if not self.h:
            raise RuntimeError("or_sim_create failed")
        self.counts = [int(c) for c in counts]
        if smoothing is not None and lib().or_sim_set_smoothing(self.h, int(smoothing[0]), int(smoothing[1]),
                                                                 float(smoothing[2])) != 0:
            raise ValueError("invalid smoothing %r" % (smoothing,))

    def close(self):
        if self.h:
            lib().or_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _size(self, which):
        g = self.g
        if which in ("E", "Bn", "J", "theta", "particle"):
            return 3 * node_count(g)
        if which == "B":
            return 3 * cell_count(g)
        return vertex_count(g) * 126

    def field(self, which):
        p = lib().or_sim_field(self.h, self.FIELDS[which])
        return np.ctypeslib.as_array(p, shape=(self._size(which),))

    def particles(self, s):
        p = lib().or_sim_particles(self.h, s).contents
        n = p.count
        out = {}
        for name in ParticleArrays.NAMES:
            out[name] = np.ctypeslib.as_array(getattr(p, name), shape=(n,))
        out["cell"] = np.ctypeslib.as_array(p.cell, shape=(n,))
        out["id"] = np.ctypeslib.as_array(p.id, shape=(n,))
        return out

    def set_particles(self, s, arrays):
        dst = self.particles(s)
        for k in dst:
            dst[k][:] = arrays[k]

    def bins(self, s):
        nb = bin_count(self.g)
        counts = np.ctypeslib.as_array(lib().or_sim_bins(self.h, s, 0), shape=(nb,))
        offsets = np.ctypeslib.as_array(lib().or_sim_bins(self.h, s, 1), shape=(nb + 1,))
        return counts, offsets

    def initialize(self):
        if lib().or_sim_initialize(self.h) != 0:
            raise RuntimeError("oracle initialize failed")

    def solve(self):
        r = GmresResult()
        rc = lib().or_sim_solve(self.h, ctypes.byref(r))
        return rc, r

    def move(self, s):
        return int(lib().or_sim_move(self.h, s))

    def advance_fields(self):
        lib().or_sim_advance_fields(self.h)

    def sort(self, s):
        return lib().or_sim_sort(self.h, s)

    def deposit(self):
        return int(lib().or_sim_deposit(self.h))

    def cycle(self, n=1):
        r = GmresResult()
        rc = lib().or_sim_cycle(self.h, n, ctypes.byref(r))
        return rc, r

    def energy(self):
        out = np.zeros(2 + self.species)
        lib().or_sim_energy(self.h, ptr(out))
        return out

    def scalars(self):
        out = np.zeros(2 + 5 * self.species)
        lib().or_sim_scalars(self.h, ptr(out))
        return out

    def validation(self):
        return int(lib().or_sim_validation(self.h))
