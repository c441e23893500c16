"""ctypes binding of libtpic_mi.so (the C-ABI of include/tpic_mi.h).

This is the host-side mirror of the reference's simulation-step interface for
the ECSIM hot path (simulation::Pipeline driving the ecsim enqueue_* functions,
include/tpic/simulation/driver/pipeline.hpp): a Simulation owns one native
context and exposes initialize / solve / move / advance_fields / sort /
deposit / assemble / cycle / energy with the reference's meaning. There is no
CPU fallback: a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtpic_mi.so")

OK = 0
INVALID_ARGUMENT = -1
INVALID_STATE = -2
OVERFLOW = -3
BACKEND = -4
NOT_CONVERGED = -5
UNSUPPORTED = -6

INVALID_KEY = 1
INVALID_FRACTION = 2
INVALID_NONFINITE = 4
INVALID_FLIGHT = 8
INVALID_HALO = 16
INVALID_SORT = 32
INVALID_FIELD = 64

FIELD_E, FIELD_B, FIELD_BNODE, FIELD_J, FIELD_ETHETA, FIELD_RHS, FIELD_MASS, FIELD_EPARTICLE = range(8)
PHASES = {"sort": 0, "move": 1, "deposit": 2, "assemble": 3, "solve": 4, "advance": 5, "spmv": 6}
MAX_SPECIES = 8


class TpicError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("tpic_mi status %d: %s" % (status, message))
        self.status = status


class _Species(ctypes.Structure):
    _fields_ = [("q_over_m", ctypes.c_double), ("count", ctypes.c_uint32)]


class Config(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int * 3), ("lengths", ctypes.c_double * 3),
                ("particle_periodic", ctypes.c_int * 3), ("dt", ctypes.c_double),
                ("theta", ctypes.c_double), ("c", ctypes.c_double), ("gmres_tolerance", ctypes.c_double),
                ("gmres_restart", ctypes.c_int), ("gmres_max_iterations", ctypes.c_int),
                ("mass_kernel", ctypes.c_int), ("species", ctypes.c_int),
                ("species_list", _Species * MAX_SPECIES), ("stream", ctypes.c_void_p),
                ("field_faces", (ctypes.c_int * 2) * 3), ("smooth_operator", ctypes.c_int),
                ("smooth_iterations", ctypes.c_int), ("smooth_alpha", ctypes.c_double)]


FACE_PERIODIC, FACE_PEC, FACE_PMC = 0, 1, 2
SMOOTH_FACE7, SMOOTH_BINOMIAL27 = 0, 1


D = ctypes.POINTER(ctypes.c_double)


CHECKPOINT_MAX_SPECIES = 10


class CheckpointHeader(ctypes.Structure):
    """tpic_mi_checkpoint_header (the reference's CheckpointHeader fields)."""
    _fields_ = [("cycle", ctypes.c_int32), ("species", ctypes.c_int32), ("processes", ctypes.c_int32 * 3),
                ("global_cells", ctypes.c_int32 * 3), ("lengths", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("c", ctypes.c_double), ("initial_energy", ctypes.c_double), ("faces", ctypes.c_int32 * 12),
                ("storage_bytes", ctypes.c_int32 * 5), ("q_over_m", ctypes.c_double * CHECKPOINT_MAX_SPECIES),
                ("global_count", ctypes.c_uint64 * CHECKPOINT_MAX_SPECIES)]


class DeckOptions(ctypes.Structure):
    _fields_ = [("mass_kernel", ctypes.c_int32), ("reserved", ctypes.c_int32 * 7)]


class DeckRun(ctypes.Structure):
    """tpic_mi_deck_run: what a started deck asks of its run."""
    _fields_ = [("ncycles", ctypes.c_int32), ("species", ctypes.c_int32), ("cells", ctypes.c_int32 * 3),
                ("energy_cycle", ctypes.c_int32), ("lengths", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("theta", ctypes.c_double), ("c", ctypes.c_double), ("count", ctypes.c_uint64 * MAX_SPECIES),
                ("q_over_m", ctypes.c_double * MAX_SPECIES)]


class ParticleArrays(ctypes.Structure):
    _fields_ = [("rx", D), ("ry", D), ("rz", D), ("u", D), ("v", D), ("w", D), ("q", D),
                ("cell", ctypes.POINTER(ctypes.c_int32)), ("id", ctypes.POINTER(ctypes.c_uint64))]


class SolveResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int),
                ("initial_residual", ctypes.c_double), ("residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double), ("relative_residual", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Load(ctypes.Structure):
    _fields_ = [("ppc", ctypes.c_int * 3), ("rho", ctypes.c_double), ("drift", ctypes.c_double * 3),
                ("thermal", ctypes.c_double * 3), ("seed", ctypes.c_uint64), ("first_id", ctypes.c_uint64)]


class UniformLoad(ctypes.Structure):
    _fields_ = [("ppc", ctypes.c_int * 3), ("rho_init", ctypes.c_double), ("drift", ctypes.c_double * 3),
                ("thermal", ctypes.c_double * 3), ("quiet_start", ctypes.c_int), ("rank", ctypes.c_int),
                ("max_species", ctypes.c_int), ("first_id", ctypes.c_uint64)]


class HarrisProfile(ctypes.Structure):
    _fields_ = [("harris_density", ctypes.c_int), ("is_double", ctypes.c_int), ("varying_thermal", ctypes.c_int),
                ("ly", ctypes.c_double), ("delta", ctypes.c_double), ("ampere_drift", ctypes.c_double),
                ("curl_center", ctypes.c_double * 2), ("curl_delta", ctypes.c_double),
                ("curl_sheet", ctypes.c_double), ("curl_mode", ctypes.c_double), ("curl_mode_k", ctypes.c_double * 2),
                ("curl_hump", ctypes.c_double), ("curl_hump_k", ctypes.c_double * 2),
                ("curl_hump_inverse_w2", ctypes.c_double * 2)]


EXPORTS = (
    "tpic_mi_create", "tpic_mi_destroy", "tpic_mi_last_error", "tpic_mi_synchronize", "tpic_mi_device_field",
    "tpic_mi_device_particles", "tpic_mi_set_particles", "tpic_mi_get_particles", "tpic_mi_load_maxwellian",
    "tpic_mi_load_uniform", "tpic_mi_load_density_seeded", "tpic_mi_load_harris", "tpic_mi_center_velocity",
    "tpic_mi_set_field",
    "tpic_mi_get_field", "tpic_mi_get_bins", "tpic_mi_record_permutation",
    "tpic_mi_initialize", "tpic_mi_sort", "tpic_mi_move", "tpic_mi_deposit",
    "tpic_mi_assemble",
    "tpic_mi_solve", "tpic_mi_apply_operator", "tpic_mi_smooth", "tpic_mi_advance_fields", "tpic_mi_cycle",
    "tpic_mi_energy",
    "tpic_mi_scalars", "tpic_mi_field_energy_static", "tpic_mi_moments",
    "tpic_mi_checkpoint_reserve", "tpic_mi_checkpoint_save", "tpic_mi_checkpoint_restore",
    "tpic_mi_checkpoint_write", "tpic_mi_checkpoint_read", "tpic_mi_checkpoint_read_header",
    "tpic_mi_apply_operator_transposed", "tpic_mi_solve_transposed", "tpic_mi_adjoint_create",
    "tpic_mi_adjoint_destroy", "tpic_mi_adjoint_seed", "tpic_mi_adjoint_cycle", "tpic_mi_adjoint_get_field",
    "tpic_mi_adjoint_get_particles",
    "tpic_mi_read_validation", "tpic_mi_clear_validation", "tpic_mi_enable_timing", "tpic_mi_phase_time",
    "tpic_mi_reset_timing", "tpic_mi_load_deck",
)

_lib = None


def library_path():
    return LIB_PATH


def load_library():
    """Load libtpic_mi.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtpic_mi.so is not built: run __graft_entry__.build() or "
                           "python tpic_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    V = ctypes.c_void_p
    I = ctypes.c_int
    L.tpic_mi_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(V)]
    L.tpic_mi_destroy.argtypes = [V]
    L.tpic_mi_last_error.argtypes = [V]
    L.tpic_mi_last_error.restype = ctypes.c_char_p
    L.tpic_mi_synchronize.argtypes = [V]
    L.tpic_mi_device_field.argtypes = [V, I, ctypes.POINTER(V)]
    L.tpic_mi_device_particles.argtypes = [V, I, I, ctypes.POINTER(V)]
    L.tpic_mi_set_particles.argtypes = [V, I, ctypes.POINTER(ParticleArrays)]
    L.tpic_mi_get_particles.argtypes = [V, I, ctypes.POINTER(ParticleArrays)]
    L.tpic_mi_load_maxwellian.argtypes = [V, I, ctypes.POINTER(Load)]
    L.tpic_mi_load_uniform.argtypes = [V, I, ctypes.POINTER(UniformLoad)]
    L.tpic_mi_load_density_seeded.argtypes = [V, I, ctypes.POINTER(UniformLoad), D]
    L.tpic_mi_center_velocity.argtypes = [V, I, D]
    L.tpic_mi_load_harris.argtypes = [V, I, ctypes.POINTER(UniformLoad), ctypes.POINTER(HarrisProfile)]
    L.tpic_mi_set_field.argtypes = [V, I, D]
    L.tpic_mi_get_field.argtypes = [V, I, D]
    U32 = ctypes.POINTER(ctypes.c_uint32)
    L.tpic_mi_get_bins.argtypes = [V, I, U32, U32, U32]
    L.tpic_mi_record_permutation.argtypes = [V, I]
    L.tpic_mi_initialize.argtypes = [V]
    L.tpic_mi_sort.argtypes = [V, I]
    L.tpic_mi_move.argtypes = [V, I]
    L.tpic_mi_deposit.argtypes = [V, I, I, I]
    L.tpic_mi_assemble.argtypes = [V]
    L.tpic_mi_solve.argtypes = [V, ctypes.POINTER(SolveResult)]
    L.tpic_mi_apply_operator.argtypes = [V, V, V]
    L.tpic_mi_smooth.argtypes = [V, V, V]
    L.tpic_mi_advance_fields.argtypes = [V]
    L.tpic_mi_cycle.argtypes = [V, I, ctypes.POINTER(SolveResult)]
    L.tpic_mi_energy.argtypes = [V, D]
    L.tpic_mi_scalars.argtypes = [V, D]
    L.tpic_mi_moments.argtypes = [V, I, ctypes.c_uint, V]
    L.tpic_mi_field_energy_static.argtypes = [ctypes.POINTER(Config), V, V, V, V]
    L.tpic_mi_checkpoint_reserve.argtypes = [V, I]
    L.tpic_mi_checkpoint_save.argtypes = [V, I]
    L.tpic_mi_checkpoint_restore.argtypes = [V, I]
    L.tpic_mi_checkpoint_write.argtypes = [V, ctypes.c_char_p, ctypes.POINTER(CheckpointHeader)]
    L.tpic_mi_checkpoint_read.argtypes = [V, ctypes.c_char_p, ctypes.POINTER(CheckpointHeader)]
    L.tpic_mi_checkpoint_read_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(CheckpointHeader)]
    L.tpic_mi_load_deck.argtypes = [ctypes.c_char_p, ctypes.POINTER(DeckOptions), V, ctypes.POINTER(V),
                                    ctypes.POINTER(DeckRun)]
    L.tpic_mi_apply_operator_transposed.argtypes = [V, V, V]
    L.tpic_mi_solve_transposed.argtypes = [V, V, V, ctypes.POINTER(SolveResult)]
    L.tpic_mi_adjoint_create.argtypes = [V, ctypes.POINTER(V)]
    L.tpic_mi_adjoint_destroy.argtypes = [V]
    L.tpic_mi_adjoint_seed.argtypes = [V, V, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
    L.tpic_mi_adjoint_cycle.argtypes = [V, V, ctypes.POINTER(SolveResult), ctypes.POINTER(SolveResult)]
    L.tpic_mi_adjoint_get_field.argtypes = [V, I, D]
    L.tpic_mi_adjoint_get_particles.argtypes = [V, I, D, D, D, ctypes.POINTER(ctypes.c_uint64)]
    L.tpic_mi_read_validation.argtypes = [V, ctypes.POINTER(ctypes.c_uint)]
    L.tpic_mi_clear_validation.argtypes = [V]
    L.tpic_mi_enable_timing.argtypes = [V, I]
    L.tpic_mi_phase_time.argtypes = [V, I, D, ctypes.POINTER(ctypes.c_uint64)]
    L.tpic_mi_reset_timing.argtypes = [V]
    for name in EXPORTS:
        fn = getattr(L, name)
        if name != "tpic_mi_last_error":
            fn.restype = ctypes.c_int
    _lib = L
    return L


def field_energy_static(cells, lengths, device_e, device_b, device_out, field_faces=None, stream=None):
    """tpic_mi_field_energy_static: W_E, W_B of device fields into device_out[0..1],
    enqueued on stream (None: the null stream); no context, no allocation.
    Returns the status; the error text is last_error()."""
    L = load_library()
    cfg = Config()
    for a in range(3):
        cfg.cells[a] = int(cells[a])
        cfg.lengths[a] = float(lengths[a])
        if field_faces is not None:
            cfg.field_faces[a][0] = int(field_faces[a][0])
            cfg.field_faces[a][1] = int(field_faces[a][1])
    return L.tpic_mi_field_energy_static(ctypes.byref(cfg), device_e, device_b, device_out, stream)


def checkpoint_header(path):
    """The header of a restart file, read without a context."""
    L = load_library()
    h = CheckpointHeader()
    status = L.tpic_mi_checkpoint_read_header(os.fsencode(path), ctypes.byref(h))
    if status:
        raise TpicError(status, last_error())
    return h


def last_error():
    """The error text of the last call that had no context."""
    return load_library().tpic_mi_last_error(None).decode()


def _ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


ATTRIBUTES = ("rx", "ry", "rz", "u", "v", "w", "q", "cell", "id")

# Mass-matrix deposition kernels (TPIC_MI_MASS_* of tpic_mi.h).
MASS_FP64, MASS_FP32, MASS_BF16X3 = 0, 1, 2
MASS_KERNELS = (MASS_FP64, MASS_FP32, MASS_BF16X3)


class Simulation:
    """One GPU context running the ECSIM cycle of a periodic box."""

    def __init__(self, cells, lengths, species, dt, theta=0.5, c=1.0, gmres_tolerance=1e-10,
                 gmres_restart=30, gmres_max_iterations=100, particle_periodic=(1, 1, 1), mass_kernel=0,
                 field_faces=None, smoothing=None):
        """species: list of (q_over_m, count); mass_kernel: MASS_FP64, MASS_FP32 or MASS_BF16X3;
        field_faces: per axis (lower, upper) of FACE_PERIODIC, FACE_PEC, FACE_PMC (default periodic);
        smoothing: (operator, iterations, alpha) with operator SMOOTH_FACE7 or SMOOTH_BINOMIAL27
        (default none)."""
        L = load_library()
        if not 0 < len(species) <= MAX_SPECIES:
            raise ValueError("1..%d species" % MAX_SPECIES)
        cfg = Config()
        for a in range(3):
            cfg.cells[a] = int(cells[a])
            cfg.lengths[a] = float(lengths[a])
            cfg.particle_periodic[a] = int(particle_periodic[a])
        cfg.dt, cfg.theta, cfg.c = float(dt), float(theta), float(c)
        cfg.gmres_tolerance = float(gmres_tolerance)
        cfg.gmres_restart = int(gmres_restart)
        cfg.gmres_max_iterations = int(gmres_max_iterations)
        if mass_kernel not in MASS_KERNELS:
            raise ValueError("mass_kernel must be one of %s" % (MASS_KERNELS,))
        cfg.mass_kernel = int(mass_kernel)
        cfg.species = len(species)
        for s, (qom, count) in enumerate(species):
            cfg.species_list[s].q_over_m = float(qom)
            cfg.species_list[s].count = int(count)
        cfg.stream = None
        if field_faces is not None:
            for a in range(3):
                cfg.field_faces[a][0] = int(field_faces[a][0])
                cfg.field_faces[a][1] = int(field_faces[a][1])
        if smoothing is not None:
            cfg.smooth_operator, cfg.smooth_iterations = int(smoothing[0]), int(smoothing[1])
            cfg.smooth_alpha = float(smoothing[2])
        h = ctypes.c_void_p()
        rc = L.tpic_mi_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != OK:
            raise TpicError(rc, L.tpic_mi_last_error(None).decode())
        self._adopt(L, h, cells, lengths, species)

    @classmethod
    def from_deck(cls, path, mass_kernel=MASS_FP64):
        """tpic_mi_load_deck: a context started from an input deck (parsed,
        admitted, loaded and initialized), ready for cycle(); returns it and
        the DeckRun of the deck."""
        L = load_library()
        opts = DeckOptions()
        opts.mass_kernel = int(mass_kernel)
        h, run = ctypes.c_void_p(), DeckRun()
        rc = L.tpic_mi_load_deck(os.fsencode(path), ctypes.byref(opts), None, ctypes.byref(h), ctypes.byref(run))
        if rc != OK:
            raise TpicError(rc, L.tpic_mi_last_error(None).decode())
        sim = cls.__new__(cls)
        sim._adopt(L, h, tuple(run.cells), tuple(run.lengths),
                   [(run.q_over_m[s], int(run.count[s])) for s in range(run.species)])
        return sim, run

    def _adopt(self, L, h, cells, lengths, species):
        self._L = L
        self.h = h
        self.cells = tuple(int(x) for x in cells)
        self.lengths = tuple(float(x) for x in lengths)
        self.species = [(float(q), int(n)) for q, n in species]
        nx, ny, nz = self.cells
        self.nodes = (nx + 1) * (ny + 1) * (nz + 1)
        self.ncells = nx * ny * nz
        self.bins = (nx + 4) * (ny + 4) * (nz + 4)
        self.vertices = (nx + 3) * (ny + 3) * (nz + 3)

    # ---- lifetime ------------------------------------------------------
    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._L.tpic_mi_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, allow=()):
        if rc != OK and rc not in allow:
            raise TpicError(rc, self._L.tpic_mi_last_error(self.h).decode())
        return rc

    # ---- state ---------------------------------------------------------
    def set_particles(self, s, arrays):
        n = self.species[s][1]
        keep = {}
        pa = ParticleArrays()
        for name in ATTRIBUTES:
            dtype = np.int32 if name == "cell" else (np.uint64 if name == "id" else np.float64)
            a = np.ascontiguousarray(arrays[name], dtype=dtype)
            if a.shape != (n,):
                raise ValueError("%s must have shape (%d,)" % (name, n))
            keep[name] = a
            ctype = ctypes.c_int32 if name == "cell" else (ctypes.c_uint64 if name == "id" else ctypes.c_double)
            setattr(pa, name, _ptr(a, ctype))
        self._check(self._L.tpic_mi_set_particles(self.h, s, ctypes.byref(pa)))

    def get_particles(self, s):
        n = self.species[s][1]
        out = {}
        pa = ParticleArrays()
        for name in ATTRIBUTES:
            dtype = np.int32 if name == "cell" else (np.uint64 if name == "id" else np.float64)
            out[name] = np.zeros(n, dtype)
            ctype = ctypes.c_int32 if name == "cell" else (ctypes.c_uint64 if name == "id" else ctypes.c_double)
            setattr(pa, name, _ptr(out[name], ctype))
        self._check(self._L.tpic_mi_get_particles(self.h, s, ctypes.byref(pa)))
        return out

    def load_maxwellian(self, s, ppc, rho=1.0, drift=(0, 0, 0), thermal=(0, 0, 0), seed=12345, first_id=0):
        ld = Load()
        for a in range(3):
            ld.ppc[a] = int(ppc[a])
            ld.drift[a] = float(drift[a])
            ld.thermal[a] = float(thermal[a])
        ld.rho = float(rho)
        ld.seed = int(seed)
        ld.first_id = int(first_id)
        self._check(self._L.tpic_mi_load_maxwellian(self.h, s, ctypes.byref(ld)))

    def load_uniform(self, s, ppc, rho_init=1.0, drift=(0, 0, 0), thermal=(0, 0, 0), quiet_start=False, rank=0,
                     max_species=10, first_id=0):
        """The reference's Uniform-case load of species s (hipRAND XORWOW,
        seed 12345 + 100 (rank max_species + s), subsequence = load index)."""
        ld = UniformLoad()
        for a in range(3):
            ld.ppc[a] = int(ppc[a])
            ld.drift[a] = float(drift[a])
            ld.thermal[a] = float(thermal[a])
        ld.rho_init = float(rho_init)
        ld.quiet_start = 1 if quiet_start else 0
        ld.rank = int(rank)
        ld.max_species = int(max_species)
        ld.first_id = int(first_id)
        self._check(self._L.tpic_mi_load_uniform(self.h, s, ctypes.byref(ld)))

    def load_density_seeded(self, s, ppc, wavenumber, phase, amplitude, rho_init=1.0, drift=(0, 0, 0),
                            thermal=(0, 0, 0), quiet_start=False, rank=0, max_species=10, first_id=0):
        """The TwoStream profile's load: the Uniform-case draws with each charge
        multiplied by 1 + amplitude cos(wavenumber x + phase)."""
        ld = UniformLoad()
        for a in range(3):
            ld.ppc[a] = int(ppc[a])
            ld.drift[a] = float(drift[a])
            ld.thermal[a] = float(thermal[a])
        ld.rho_init = float(rho_init)
        ld.quiet_start = 1 if quiet_start else 0
        ld.rank = int(rank)
        ld.max_species = int(max_species)
        ld.first_id = int(first_id)
        mode = (ctypes.c_double * 3)(float(wavenumber), float(phase), float(amplitude))
        self._check(self._L.tpic_mi_load_density_seeded(self.h, s, ctypes.byref(ld),
                                                        ctypes.cast(mode, D)))

    def load_harris(self, s, ppc, profile, rho_init=1.0, drift=(0, 0, 0), thermal=(0, 0, 0), quiet_start=False,
                    rank=0, max_species=10, first_id=0):
        """A Harris-sheet case's load (profile: a HarrisProfile structure)."""
        ld = UniformLoad()
        for a in range(3):
            ld.ppc[a] = int(ppc[a])
            ld.drift[a] = float(drift[a])
            ld.thermal[a] = float(thermal[a])
        ld.rho_init = float(rho_init)
        ld.quiet_start = 1 if quiet_start else 0
        ld.rank = int(rank)
        ld.max_species = int(max_species)
        ld.first_id = int(first_id)
        self._check(self._L.tpic_mi_load_harris(self.h, s, ctypes.byref(ld), ctypes.byref(profile)))

    def center_velocity(self, s, drift):
        """CenterInitialVelocity: the charge-weighted mean of v - drift is
        removed from species s (blocks the host)."""
        d = (ctypes.c_double * 3)(*[float(x) for x in drift])
        self._check(self._L.tpic_mi_center_velocity(self.h, s, ctypes.cast(d, D)))

    def _field_size(self, which):
        if which == FIELD_B:
            return 3 * self.ncells
        if which == FIELD_MASS:
            return self.vertices * 126
        return 3 * self.nodes

    def set_field(self, which, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        if a.size != self._field_size(which):
            raise ValueError("field size %d, expected %d" % (a.size, self._field_size(which)))
        self._check(self._L.tpic_mi_set_field(self.h, which, _ptr(a)))

    def get_field(self, which):
        out = np.zeros(self._field_size(which))
        self._check(self._L.tpic_mi_get_field(self.h, which, _ptr(out)))
        return out

    def record_permutation(self, enable=True):
        self._check(self._L.tpic_mi_record_permutation(self.h, 1 if enable else 0))

    def get_bins(self, s, permutation=False):
        counts = np.zeros(self.bins, np.uint32)
        offsets = np.zeros(self.bins + 1, np.uint32)
        perm = np.zeros(max(self.species[s][1], 1), np.uint32) if permutation else None
        self._check(self._L.tpic_mi_get_bins(self.h, s, _ptr(counts, ctypes.c_uint32),
                                             _ptr(offsets, ctypes.c_uint32),
                                             _ptr(perm, ctypes.c_uint32) if permutation else None))
        if permutation:
            return counts, offsets, perm[:self.species[s][1]]
        return counts, offsets

    def device_particles(self, s, attribute):
        """Device pointer of attribute (rx, ry, rz, u, v, w, q, cell, id = 0..8) of species s."""
        p = ctypes.c_void_p()
        self._check(self._L.tpic_mi_device_particles(self.h, s, attribute, ctypes.byref(p)))
        return p.value

    def device_field(self, which):
        p = ctypes.c_void_p()
        self._check(self._L.tpic_mi_device_field(self.h, which, ctypes.byref(p)))
        return p.value

    # ---- hot path ------------------------------------------------------
    def initialize(self):
        self._check(self._L.tpic_mi_initialize(self.h))

    def sort(self, s):
        self._check(self._L.tpic_mi_sort(self.h, s))

    def move(self, s):
        self._check(self._L.tpic_mi_move(self.h, s))

    def deposit(self, first=0, last=None, accumulate=False):
        last = len(self.species) if last is None else last
        self._check(self._L.tpic_mi_deposit(self.h, first, last, 1 if accumulate else 0))

    def assemble(self):
        self._check(self._L.tpic_mi_assemble(self.h))

    def solve(self, raise_on_failure=True):
        r = SolveResult()
        rc = self._L.tpic_mi_solve(self.h, ctypes.byref(r))
        if raise_on_failure:
            self._check(rc)
        return rc, r.as_dict()

    def apply_operator(self, device_in, device_out):
        self._check(self._L.tpic_mi_apply_operator(self.h, ctypes.c_void_p(device_in), ctypes.c_void_p(device_out)))

    def smooth(self, device_in, device_out):
        """device_out = S(device_in) (node vectors)."""
        self._check(self._L.tpic_mi_smooth(self.h, ctypes.c_void_p(device_in), ctypes.c_void_p(device_out)))

    def advance_fields(self):
        self._check(self._L.tpic_mi_advance_fields(self.h))

    def cycle(self, n=1):
        r = SolveResult()
        self._check(self._L.tpic_mi_cycle(self.h, int(n), ctypes.byref(r)))
        return r.as_dict()

    def scalars(self):
        """W_E, W_B, then per species kinetic energy, charge, momentum x, y, z."""
        out = np.zeros(2 + 5 * len(self.species))
        self._check(self._L.tpic_mi_scalars(self.h, _ptr(out)))
        return out

    def moments(self, species, selected=0x3FF):
        """Output moments of a sorted species at nodes (tpic_mi_moments) as a
        (10, nodes) float64 array, rows rho, Jx, Jy, Jz, PXX, PXY, PXZ, PYY,
        PYZ, PZZ; the device buffer is the HIP runtime's, freed on return."""
        # The runtime the library links (torch may have loaded its own copy
        # under the same name, whose allocations this context cannot use).
        hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        out = np.zeros((10, self.nodes))
        dev = ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(out.nbytes)) != 0:
            raise TpicError(BACKEND, "hipMalloc of the moment buffer failed")
        try:
            self._check(self._L.tpic_mi_moments(self.h, int(species), int(selected), dev))
            if hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, ctypes.c_size_t(out.nbytes), 2) != 0:
                raise TpicError(BACKEND, "moment readback failed")
        finally:
            hip.hipFree(dev)
        return out

    def energy(self):
        out = np.zeros(2 + len(self.species))
        self._check(self._L.tpic_mi_energy(self.h, _ptr(out)))
        return out

    def validation(self):
        v = ctypes.c_uint()
        self._check(self._L.tpic_mi_read_validation(self.h, ctypes.byref(v)))
        return int(v.value)

    def clear_validation(self):
        self._check(self._L.tpic_mi_clear_validation(self.h))

    def synchronize(self):
        self._check(self._L.tpic_mi_synchronize(self.h))

    # ---- cycle checkpoints and the adjoint ---------------------------------
    def checkpoint_reserve(self, slots):
        self._check(self._L.tpic_mi_checkpoint_reserve(self.h, int(slots)))

    def checkpoint_save(self, slot):
        self._check(self._L.tpic_mi_checkpoint_save(self.h, int(slot)))

    def checkpoint_restore(self, slot):
        self._check(self._L.tpic_mi_checkpoint_restore(self.h, int(slot)))

    def checkpoint_write(self, path, cycle=0, initial_energy=0.0):
        """Restart file of the current state (after a cycle)."""
        h = CheckpointHeader()
        h.cycle, h.initial_energy = int(cycle), float(initial_energy)
        self._check(self._L.tpic_mi_checkpoint_write(self.h, os.fsencode(path), ctypes.byref(h)))

    def checkpoint_read(self, path):
        """Loads a restart file (then call initialize); returns its header."""
        h = CheckpointHeader()
        self._check(self._L.tpic_mi_checkpoint_read(self.h, os.fsencode(path), ctypes.byref(h)))
        return h

    def apply_operator_transposed(self, device_in, device_out):
        self._check(self._L.tpic_mi_apply_operator_transposed(self.h, ctypes.c_void_p(device_in),
                                                               ctypes.c_void_p(device_out)))

    def solve_transposed(self, device_rhs, device_x):
        r = SolveResult()
        self._check(self._L.tpic_mi_solve_transposed(self.h, ctypes.c_void_p(device_rhs), ctypes.c_void_p(device_x),
                                                     ctypes.byref(r)))
        return r.as_dict()

    def adjoint(self):
        """A new adjoint state of this context (see Adjoint)."""
        return Adjoint(self)

    # ---- timing --------------------------------------------------------
    def enable_timing(self, enable=True):
        self._check(self._L.tpic_mi_enable_timing(self.h, 1 if enable else 0))

    def reset_timing(self):
        self._check(self._L.tpic_mi_reset_timing(self.h))

    def phase_time(self, phase):
        ms = ctypes.c_double()
        calls = ctypes.c_uint64()
        self._check(self._L.tpic_mi_phase_time(self.h, PHASES[phase], ctypes.byref(ms), ctypes.byref(calls)))
        return ms.value, int(calls.value)


class Adjoint:
    """The adjoint of one state of a Simulation (tpic_mi_adjoint_*): Ebar at
    nodes, Bbar at cells and per species rbar, vbar, qbar in the order of its
    ids. Periodic fields and particles, no smoothing."""

    def __init__(self, sim):
        self.sim = sim
        self._L = sim._L
        h = ctypes.c_void_p()
        sim._check(self._L.tpic_mi_adjoint_create(sim.h, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._L.tpic_mi_adjoint_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def seed_field_energy(self, weight_e=1.0, weight_b=0.0, first_ids=None):
        """Adjoint of weight_e W_E + weight_b W_B of the context's current state;
        first_ids[s]: the first id of species s (its ids fill [first, first + count))."""
        ns = len(self.sim.species)
        if first_ids is None:
            first_ids = [0] * ns
        ids = (ctypes.c_uint64 * ns)(*[int(x) for x in first_ids])
        self.sim._check(self._L.tpic_mi_adjoint_seed(self.sim.h, self.h, float(weight_e), float(weight_b),
                                                     ctypes.cast(ids, ctypes.POINTER(ctypes.c_uint64))))

    def cycle(self):
        """One cycle backwards from the context's restored starting state."""
        fwd, tr = SolveResult(), SolveResult()
        self.sim._check(self._L.tpic_mi_adjoint_cycle(self.sim.h, self.h, ctypes.byref(fwd), ctypes.byref(tr)))
        return fwd.as_dict(), tr.as_dict()

    def field(self, which):
        n = 3 * (self.sim.nodes if which == FIELD_E else self.sim.ncells)
        out = np.zeros(n)
        self.sim._check(self._L.tpic_mi_adjoint_get_field(self.h, which, _ptr(out)))
        return out

    def particles(self, s):
        n = self.sim.species[s][1]
        rbar, vbar = np.zeros((3, n)), np.zeros((3, n))
        qbar, ids = np.zeros(n), np.zeros(n, np.uint64)
        self.sim._check(self._L.tpic_mi_adjoint_get_particles(self.h, s, _ptr(rbar), _ptr(vbar), _ptr(qbar),
                                                              _ptr(ids, ctypes.c_uint64)))
        return dict(rbar=rbar, vbar=vbar, qbar=qbar, ids=ids)

