"""GPU: the run-time tuner (rpm_amd.autotune) picks among real variants and never changes results.

- hpsi_kernel times every H·psi kernel the production operator supports, keeps the fastest, caches
  it per batch bucket, and the tuned operator's H·psi equals the gather kernel's to 1e-13
  (relative L2);
- an operator without a tiled form has only its own kernel timed (or, in the generic local-term
  form, none);
- propagation_batch picks one of the candidate batch sizes, and directional yields with the tuned
  batch equal those with a fixed batch to 1e-12.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from rpm_amd import autotune, batch, core, engine, workloads  # noqa: E402


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def test_hpsi_kernel_tuning_production():
    print("case: tune H·psi kernel, production, batch 8", flush=True)
    autotune.clear_cache()
    h = core.build_hamiltonian(90.0, 0.05)
    op = h._device_op()
    tuned = autotune.hpsi_kernel(h, 8)
    print(tuned, flush=True)
    assert set(tuned["ms_per_launch"]) == {1, 2, 3, 4}
    assert tuned["kernel"] == min(tuned["ms_per_launch"], key=tuned["ms_per_launch"].get)
    assert op.kernel == tuned["kernel"]
    assert autotune.hpsi_kernel(h, 7) is tuned            # same power-of-two bucket: cached
    x = torch.from_numpy(workloads.normal_vectors(h.dimension, 8, seed=3)).cuda()
    y = torch.empty_like(x)
    h.apply_device(x, y, batch=8)
    op.set_kernel(1)
    ref = torch.empty_like(x)
    h.apply_device(x, ref, batch=8)
    op.set_kernel(0)
    assert _rel(y, ref) < 1e-13


def test_hpsi_kernel_tuning_operator_without_tiled_form():
    """Only the kernels an operator has are timed: a small system has no tiled form."""
    rng = np.random.default_rng(4)
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    h = engine.MatrixFreeHamiltonian((2, 2, 3), [engine.LocalTerm((2,), a + a.conj().T)])
    tuned = autotune.hpsi_kernel(h, 4)
    if tuned["name"] == "local-term":          # generic local-term form: one kernel, nothing to time
        assert tuned["ms_per_launch"] == {}
    else:
        assert tuned["kernel"] == 1 and set(tuned["ms_per_launch"]) == {1}


def test_propagation_batch_tuning_keeps_results():
    print("case: tune batch size, small system", flush=True)
    autotune.clear_cache()
    sys_ = workloads.small_system(0)
    base = sys_.hamiltonian()
    build = lambda t, f, p: sys_.hamiltonian(theta_deg=t, field_mT=f, phi_deg=p)  # noqa: E731
    states = np.stack([core.random_phase_initial_state(base.dimensions, 0, i) for i in range(3)])
    th, ph = workloads.fibonacci_sphere(3)
    fields = [25.0, 50.0]
    times = np.arange(81) * 0.00125
    bp = batch.BatchedLanczosPropagator(base, minimum_dimension=8, maximum_dimension=24, dimension_increment=8)
    auto = batch.directional_yields(base, build, bp, states, th, ph, fields, times, 1.0, batch_size="auto")
    fixed = batch.directional_yields(base, build, bp, states, th, ph, fields, times, 1.0, batch_size=5)
    assert np.abs(auto - fixed).max() < 1e-12
    hs = [build(float(th[0]), 0.05, float(ph[0]))] * 3
    tuned = autotune.propagation_batch(bp, states, times, batch.pair_matrices(base, hs), candidates=(1, 2, 3))
    print(tuned, flush=True)
    assert tuned["batch"] in (1, 2, 3) and set(tuned["trajectory_us_per_s"]) == {1, 2, 3}
