"""Gradients of field energies at the full sizes of the two differentiable
workloads, through tpic_amd.gradient.FieldEnergy (adjoint kernels over cycle
checkpoints with recomputation), against central differences of the device's
own forward run in FP64:

- two-stream, 1024 cells, 2 beams x 1000 ppc (v0 = 0.2, vth = 0.01), density
  seed eps = 1e-3 at mode 8, dt = 0.1, 200 cycles, checkpoints every 20:
  d W_E / d v0 and d W_E / d eps;
- Weibel, 512 x 512 cells (dx = 0.1), electrons with thermal speeds
  (0.3, 0.1, 0.3), ions q/m = 0.01 at 0.01, 100 ppc each, dt = 0.05, 50
  cycles, checkpoints every 10: d W_B / d ratio (ratio = 3), with the FP64
  mass matrix and again with BF16 hi/lo moment products, whose gradient is
  compared with the FP64 one.

The cycle map is piecewise smooth: a particle's CIC weights have kinks at cell
faces, so a central difference averages the derivative over the particles
whose trajectories cross a face inside the perturbation, an error that falls
with the step; at small steps the solver tolerance (1e-13) and rounding take
over. Each derivative is therefore checked at two steps, 1e-7 and 1e-8 of the
parameter: the closer difference must agree with the adjoint within 1e-5 and
both within 1e-3 (measured at 200 two-stream cycles: 5.9e-4 and 7.6e-6 for
v0, 6.5e-6 and 1.6e-4 for eps; at 1e-3 of the parameter the kinks alone give
1.8e-2)."""
import pytest

pytestmark = pytest.mark.gpu

FD_BOUND = 1e-5
FD_STEPS = (1e-7, 1e-8)   # relative to the parameter


def say(*args):
    print(*args, flush=True)


def check_against_differences(G, model, params, grads, label):
    for i, g in enumerate(grads):
        rel = []
        for step in FD_STEPS:
            fd = G.central_difference(model, params, i, abs(params[i]) * step)
            rel.append(abs(g - fd) / abs(fd))
            say(label, "parameter", i, "adjoint", g, "step", step, "central difference", fd, "relative", rel[-1])
        assert min(rel) <= FD_BOUND and max(rel) <= 1e-3, (label, i, rel)


def test_two_stream_gradient_full_size():
    from tpic_amd import gradient as G
    say("two-stream 1024 cells x 2 x 1000 ppc, 200 cycles")
    model = G.TwoStream(cells=1024, ppc=1000, mode=8, cycles=200, every=20, gmres_tolerance=1e-13)
    try:
        value, grads = G.gradient(model, 0.2, 1e-3)
        say("W_E", value, "adjoint", grads)
        check_against_differences(G, model, [0.2, 1e-3], grads, "two-stream")
    finally:
        model.close()


def test_weibel_gradient_full_size():
    from tpic_amd import gradient as G
    results = {}
    for kernel in (0, 2):
        say("Weibel 512 x 512, 2 x 100 ppc, 50 cycles, mass kernel", kernel)
        model = G.Weibel(cells=(512, 512), ppc=(10, 10), cycles=50, every=10, mass_kernel=kernel,
                         gmres_tolerance=1e-13)
        try:
            value, grads = G.gradient(model, 3.0)
            say("W_B", value, "adjoint", grads[0])
            if kernel == 0:
                check_against_differences(G, model, [3.0], grads, "Weibel")
        finally:
            model.close()
        results[kernel] = (value, grads[0])
    (v64, g64), (v16, g16) = results[0], results[2]
    say("BF16 hi/lo against FP64: W_B", abs(v16 - v64) / abs(v64), "gradient", abs(g16 - g64) / abs(g64))
    assert abs(g16 - g64) <= 1e-3 * abs(g64)
