"""CPU pins of the test harness's FP64 adjoint (adjoint_ref.py): its gradient against central finite
differences of its own exact forward, and the wall invariant dE/dw0 = 0 on the edges."""
import numpy as np

from adjoint_ref import CavityAdjoint


def test_cpu_adjoint_matches_central_differences():
    n, steps, U = 12, 3, 1.1
    ref = CavityAdjoint(n, Re=400.0, dt=2e-3)
    rng = np.random.default_rng(3)
    w0 = rng.standard_normal((n, n))
    E, g, dU = ref.gradient(w0, U, steps)
    assert abs(E - ref.energy(w0, U, steps)) <= 1e-14 * abs(E)
    eps = 1e-6
    for _ in range(3):
        d = rng.standard_normal((n, n))
        fd = (ref.energy(w0 + eps * d, U, steps) - ref.energy(w0 - eps * d, U, steps)) / (2 * eps)
        assert abs(np.sum(g * d) - fd) <= 1e-6 * abs(fd), (np.sum(g * d), fd)
    fdU = (ref.energy(w0, U + eps, steps) - ref.energy(w0, U - eps, steps)) / (2 * eps)
    assert abs(dU - fdU) <= 1e-6 * abs(fdU), (dU, fdU)


def test_cpu_adjoint_wall_gradient_is_zero():
    n = 10
    ref = CavityAdjoint(n, Re=100.0, dt=1e-3)
    w0 = np.random.default_rng(5).standard_normal((n, n))
    _, g, _ = ref.gradient(w0, 1.0, 2)
    E = np.zeros((n, n), dtype=bool)
    E[0, :] = E[-1, :] = E[:, 0] = E[:, -1] = True
    assert np.all(g[E] == 0.0) and np.any(g[~E] != 0.0)
