"""tpic_mi_field_energy_static, the entry point that allocates nothing,
against tpic_mi_energy on the same fields (through the C-ABI): two calls in a
row, the second on a smaller grid with walls, each within 1e-13 relative of
the context's W_E and W_B (the same kernel; only the number of block partials,
and so the summation order, differs)."""
import ctypes

import numpy as np
import pytest

from tpic_amd import runtime as R

pytestmark = pytest.mark.gpu


def _device_synchronize():
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libtpic_mi.so already loaded
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    assert hip.hipDeviceSynchronize() == 0


def _fields(sim, rng):
    e = rng.standard_normal(3 * sim.nodes)
    b = rng.standard_normal(3 * sim.ncells)
    sim.set_field(R.FIELD_E, e)
    sim.set_field(R.FIELD_B, b)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_static_field_energy_matches_the_context_energy(order):
    rng = np.random.default_rng(20261203)
    cases = [((16, 12, 10), (4.0, 3.0, 2.5), None),
             ((8, 6, 5), (2.0, 1.5, 1.25), ((R.FACE_PERIODIC, R.FACE_PERIODIC), (R.FACE_PEC, R.FACE_PEC),
                                            (R.FACE_PMC, R.FACE_PMC)))]
    for cells, lengths, faces in cases:
        print("no-alloc energy case", cells, faces, flush=True)
        with R.Simulation(cells, lengths, [(-1.0, 8)], dt=0.1, field_faces=faces) as sim:
            _fields(sim, rng)
            want = sim.energy()[:2]
            # The context's RHS owner serves as the caller's device output.
            out = sim.device_field(R.FIELD_RHS)
            rc = R.field_energy_static(cells, lengths, sim.device_field(R.FIELD_E), sim.device_field(R.FIELD_B),
                                       out, field_faces=faces)
            assert rc == R.OK, R.last_error()
            _device_synchronize()
            got = sim.get_field(R.FIELD_RHS)[:2]
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_static_field_energy_rejects_a_grid_beyond_its_scratch():
    # Validated before any device work: the pointers are never read.
    dummy = ctypes.c_void_p(16)
    rc = R.field_energy_static((512, 256, 256), (1.0, 1.0, 1.0), dummy, dummy, dummy)
    assert rc == R.INVALID_ARGUMENT
    assert "exceeds the scratch limit" in R.last_error()
