"""GPU: adjoint gradients of singlet yields against central finite differences.

Bars: gradient relative error <= 1e-5 against central differences of the
yield (h = 1e-4 rad for angles, 1e-3 uT for field strength and pulse
amplitudes).  On the 576-dimensional random system the differenced yield is
the exact-dynamics Simpson yield of the dense oracle; on the full system it is
the GPU Krylov yield itself with the step schedule frozen (replayed) and a
tight error budget.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from oracle import core as ocore  # noqa: E402
from oracle import dense as odense  # noqa: E402
from rpm_amd import adjoint, batch, core, workloads  # noqa: E402

H_ANGLE = 1e-4      # rad
H_FIELD = 1e-3      # uT
TOL = 1e-5


def _small_prop(base, budget=1e-11):
    return batch.BatchedLanczosPropagator(base, minimum_dimension=24, maximum_dimension=64,
                                          state_error_budget=budget, initial_step_us=0.05, maximum_step_us=0.25)


def _dense_h(sys_, theta, field_uT, phi):
    o = sys_.hamiltonian(theta_deg=theta, field_mT=1e-3 * field_uT, phi_deg=phi, builder=ocore.build_hamiltonian)
    return o.explicit_matrix()


def _rel(g, fd, scale):
    return abs(g - fd) / max(abs(fd), scale)


def test_field_gradient_small_system_matches_dense_finite_differences():
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(241) * 0.00125
    tau = 1.0
    params = [(30.0, 20.0, 50.0), (75.0, -40.0, 35.0)]
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(len(params))])
    out = adjoint.yield_and_gradient(
        base, lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p), _small_prop(base),
        states, times, tau, [p[0] for p in params], [p[1] for p in params], [p[2] for p in params])
    print("recovered psi_0 error", out.recovered_initial_error, flush=True)
    assert out.recovered_initial_error < 1e-9
    for b, (th, ph, f) in enumerate(params):
        y0, _ = odense.dense_grid_yield([(times[-1], _dense_h(sys_, th, f, ph))], states[b], times, tau)
        assert abs(out.yields[b] - y0) < 1e-9
        fd = []
        for k, (dth, dph, df) in enumerate([(math.degrees(H_ANGLE), 0, 0), (0, math.degrees(H_ANGLE), 0),
                                            (0, 0, H_FIELD)]):
            yp, _ = odense.dense_grid_yield([(times[-1], _dense_h(sys_, th + dth, f + df, ph + dph))], states[b],
                                            times, tau)
            ym, _ = odense.dense_grid_yield([(times[-1], _dense_h(sys_, th - dth, f - df, ph - dph))], states[b],
                                            times, tau)
            fd.append((yp - ym) / (2 * (H_ANGLE if k < 2 else H_FIELD)))
        fd = np.array(fd)
        g = out.field_gradient[b]
        scale = 1e-3 * np.abs(fd).max()
        errs = [_rel(g[k], fd[k], scale) for k in range(3)]
        print(f"b={b} grad {g} fd {fd} rel {errs}", flush=True)
        assert max(errs) <= TOL


@pytest.mark.parametrize("window", [(0.0, 0.15), (0.05, 0.2)])
def test_pulse_gradient_small_system_matches_dense_finite_differences(window):
    print("case: pulse window", window, flush=True)
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(241) * 0.00125
    tau = 1.0
    rng = np.random.default_rng(7)
    amps = rng.uniform(-20.0, 20.0, size=6)
    th, ph, f = 40.0, 0.0, 50.0
    state = core.random_phase_initial_state(base.dimensions, 0, 3)[None]
    out = adjoint.yield_and_gradient(
        base, lambda t, fm, p: sys_.hamiltonian(theta_deg=t, field_mT=fm, phi_deg=p), _small_prop(base),
        state, times, tau, [th], [ph], [f], pulse_amplitudes_uT=amps, pulse_window_us=window)
    h0 = _dense_h(sys_, th, f, ph)
    mx = np.kron(adjoint.electron_derivative_operators()[0], np.eye(h0.shape[0] // 4))

    def dense_yield(a):
        # a window starting after 0 begins with a static-field piece
        edges = np.linspace(window[0], window[1], a.size + 1)
        lead = [(window[0], h0)] if window[0] > 0.0 else []
        pieces = lead + [(edges[s + 1], h0 + 1e-3 * a[s] * mx) for s in range(a.size)] + [(times[-1], h0)]
        return odense.dense_grid_yield(pieces, state[0], times, tau)[0]

    assert abs(out.yields[0] - dense_yield(amps)) < 1e-9
    fd = np.empty(amps.size)
    for s in range(amps.size):
        e = np.zeros(amps.size)
        e[s] = H_FIELD
        fd[s] = (dense_yield(amps + e) - dense_yield(amps - e)) / (2 * H_FIELD)
    g = out.pulse_gradient[0]
    scale = 1e-3 * np.abs(fd).max()
    errs = [_rel(g[s], fd[s], scale) for s in range(amps.size)]
    print("pulse grad", g, "fd", fd, "rel", errs, flush=True)
    assert max(errs) <= TOL
    # the static-field gradient is still available with the pulse on
    assert np.all(np.isfinite(out.field_gradient))


def test_field_gradient_full_system_matches_replayed_krylov_differences():
    base = core.build_hamiltonian(90.0, 0.05)
    times = np.arange(17) * 0.00125
    tau = 5.0
    params = [(90.0, 0.0, 50.0), (60.0, 30.0, 50.0)]
    states = np.stack([core.random_phase_initial_state(base.dimensions, core.MASTER_SEED, i) for i in range(2)])
    bp = batch.BatchedLanczosPropagator(base, state_error_budget=1e-12, initial_step_us=0.05)
    build = lambda t, fm, p: core.build_hamiltonian(t, fm, phi_deg=p)  # noqa: E731
    out = adjoint.yield_and_gradient(base, build, bp, states, times, tau, [p[0] for p in params],
                                     [p[1] for p in params], [p[2] for p in params])
    print("full system recovered psi_0 error", out.recovered_initial_error, flush=True)
    center = bp.propagate(states, times, pairs=batch.pair_matrices(base, [build(t, 1e-3 * f, p) for t, p, f in params]))
    sched = [r.schedule for r in center]

    def yields(ps):
        pr = batch.pair_matrices(base, [build(t, 1e-3 * f, p) for t, p, f in ps])
        res = bp.propagate(states, times, pairs=pr, schedules=sched)
        return np.array([core.integrate_yield(times, r.observable_values, tau) for r in res])

    fd = np.zeros((2, 3))
    for k, (dth, dph, df, step) in enumerate([(math.degrees(H_ANGLE), 0, 0, H_ANGLE),
                                              (0, math.degrees(H_ANGLE), 0, H_ANGLE), (0, 0, H_FIELD, H_FIELD)]):
        yp = yields([(t + dth, p + dph, f + df) for t, p, f in params])
        ym = yields([(t - dth, p - dph, f - df) for t, p, f in params])
        fd[:, k] = (yp - ym) / (2 * step)
    for b in range(2):
        scale = 1e-3 * np.abs(fd[b]).max()
        errs = [_rel(out.field_gradient[b, k], fd[b, k], scale) for k in range(3)]
        print(f"full b={b} grad {out.field_gradient[b]} fd {fd[b]} rel {errs}", flush=True)
        assert max(errs) <= TOL


def test_autograd_function_matches_direct_gradient():
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(81) * 0.00125
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(2)])
    build = lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p)  # noqa: E731
    prob = adjoint.FieldYield(base, build, _small_prop(base), states, times, 1.0, pulse_window_us=(0.0, 0.05))
    params = torch.tensor([[0.4, 0.1, 50.0], [1.2, -0.3, 40.0]], dtype=torch.float64, requires_grad=True)
    amps = torch.tensor([[5.0, -3.0], [1.0, 2.0]], dtype=torch.float64, requires_grad=True)
    y = adjoint.singlet_yield(prob, params, amps)
    (y * torch.tensor([1.0, -2.0], dtype=torch.float64)).sum().backward()
    direct = adjoint.yield_and_gradient(base, build, _small_prop(base), states, times, 1.0,
                                        np.degrees([0.4, 1.2]), np.degrees([0.1, -0.3]), [50.0, 40.0],
                                        pulse_amplitudes_uT=amps.detach().numpy(), pulse_window_us=(0.0, 0.05))
    np.testing.assert_allclose(y.detach().numpy(), direct.yields, rtol=0, atol=1e-15)
    w = np.array([1.0, -2.0])[:, None]
    np.testing.assert_allclose(params.grad.numpy(), w * direct.field_gradient, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(amps.grad.numpy(), w * direct.pulse_gradient, rtol=1e-12, atol=1e-18)


def test_autograd_two_forwards_then_one_backward():
    """A forward whose backward runs after a second forward keeps its own gradient."""
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    times = np.arange(81) * 0.00125
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(2)])
    build = lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p)  # noqa: E731
    prob = adjoint.FieldYield(base, build, _small_prop(base), states, times, 1.0)
    p1 = torch.tensor([[0.4, 0.1, 50.0], [1.2, -0.3, 40.0]], dtype=torch.float64, requires_grad=True)
    p2 = torch.tensor([[2.0, 0.7, 30.0], [0.2, 1.1, 60.0]], dtype=torch.float64, requires_grad=True)
    y1 = adjoint.singlet_yield(prob, p1)
    y2 = adjoint.singlet_yield(prob, p2)
    y1.sum().backward()
    assert p2.grad is None
    direct1 = adjoint.yield_and_gradient(base, build, _small_prop(base), states, times, 1.0,
                                         np.degrees([0.4, 1.2]), np.degrees([0.1, -0.3]), [50.0, 40.0])
    np.testing.assert_allclose(p1.grad.numpy(), direct1.field_gradient, rtol=1e-12, atol=1e-18)
    y2.sum().backward()
    direct2 = adjoint.yield_and_gradient(base, build, _small_prop(base), states, times, 1.0,
                                         np.degrees([2.0, 0.2]), np.degrees([0.7, 1.1]), [30.0, 60.0])
    np.testing.assert_allclose(p2.grad.numpy(), direct2.field_gradient, rtol=1e-12, atol=1e-18)
    assert np.abs(direct1.field_gradient - direct2.field_gradient).max() > 1e-6
