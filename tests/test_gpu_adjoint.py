"""GPU tests of the adjoint (cnv_adj_*): the hand-written adjoint kernels' gradient of the final
kinetic energy against the FP64 CPU adjoint of the test harness (adjoint_ref.py, itself pinned by
central differences on the CPU), against central finite differences (eps = 1e-6) of the GPU
forward run, the wall invariant dE/dw0 = 0 (step 0 overwrites the wall vorticity before reading
it, src/main.c:311-320), independence of the checkpoint interval, and the gradient seen through a
torch.autograd.Function wrapper (test harness only)."""
import numpy as np
import pytest

import cnv
from adjoint_ref import CavityAdjoint

pytestmark = pytest.mark.gpu


def make(n, Re, dt, K, warm_start=0):
    p = cnv.default_params(n, Re=Re, dt=dt, beta=cnv.reference_beta(n), itmax=10_000_000, diagnostics=0,
                           warm_start=warm_start)
    return cnv.Adjoint(p, checkpoint_every=K, rtol_forward=1e-14, rtol_adjoint=1e-14)


def walls(n):
    E = np.zeros((n, n), dtype=bool)
    E[0, :] = E[-1, :] = E[:, 0] = E[:, -1] = True
    return E


def test_gpu_adjoint_matches_cpu_adjoint():
    n, steps, U, Re, dt = 24, 6, 1.1, 400.0, 4e-3
    w0 = cnv.gaussian_vorticity(n, count=8, seed=2)
    print("adjoint n=%d steps=%d" % (n, steps), flush=True)
    with make(n, Re, dt, K=2) as adj:
        E = adj.forward(w0, U, steps)
        g, dU = adj.backward()
        fs, bs = adj.sweeps()
    E_ref, g_ref, dU_ref = CavityAdjoint(n, Re, dt).gradient(w0, U, steps)
    print("E", E, E_ref, "dU", dU, dU_ref, "sweeps", fs, bs, flush=True)
    assert abs(E - E_ref) <= 1e-11 * abs(E_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-8 * np.max(np.abs(g_ref))
    assert abs(dU - dU_ref) <= 1e-8 * abs(dU_ref)
    assert np.all(g[walls(n)] == 0.0)
    assert fs > 0 and bs > 0


def test_gpu_adjoint_central_differences():
    n, steps, U, Re, dt = 32, 8, 0.9, 400.0, 3e-3
    rng = np.random.default_rng(7)
    w0 = cnv.gaussian_vorticity(n, count=12, seed=1)
    with make(n, Re, dt, K=3) as adj:
        adj.forward(w0, U, steps)
        g, dU = adj.backward()
        eps = 1e-6
        for trial in range(2):
            d = rng.standard_normal((n, n))
            fd = (adj.forward(w0 + eps * d, U, steps) - adj.forward(w0 - eps * d, U, steps)) / (2 * eps)
            print("direction", trial, float(np.sum(g * d)), fd, flush=True)
            assert abs(np.sum(g * d) - fd) <= 1e-5 * abs(fd)
        fdU = (adj.forward(w0, U + eps, steps) - adj.forward(w0, U - eps, steps)) / (2 * eps)
        print("lid", dU, fdU, flush=True)
        assert abs(dU - fdU) <= 1e-5 * abs(fdU)


def test_gpu_adjoint_checkpoint_interval_does_not_change_the_gradient():
    n, steps, U, Re, dt = 20, 7, 1.0, 250.0, 4e-3
    w0 = cnv.gaussian_vorticity(n, count=6, seed=4)
    out = []
    for K in (1, 3, 7, 100):
        with make(n, Re, dt, K) as adj:
            E = adj.forward(w0, U, steps)
            g, dU = adj.backward()
        out.append((E, g, dU))
    for E, g, dU in out[1:]:
        assert E == out[0][0] and dU == out[0][2] and np.array_equal(g, out[0][1])


def test_gpu_adjoint_warm_start():
    # warm-started solves (forward from the previous step's psi, adjoint from the previous reverse
    # step's solution) take fewer sweeps and give the gradient of the CPU adjoint within the same
    # bounds; the checkpoints carry the psi entering their step and the adjoint solution is carried
    # across segments, so the checkpoint interval and a second backward call change nothing
    n, steps, U, Re, dt = 24, 9, 1.1, 400.0, 4e-3
    w0 = cnv.gaussian_vorticity(n, count=8, seed=2)
    E_ref, g_ref, dU_ref = CavityAdjoint(n, Re, dt).gradient(w0, U, steps)
    out = []
    for K in (1, 4, 100):
        print("adjoint warm start K=%d" % K, flush=True)
        with make(n, Re, dt, K, warm_start=1) as adj:
            E = adj.forward(w0, U, steps)
            g, dU = adj.backward()
            warm = adj.sweeps()  # forward + recompute, adjoint
            g2, dU2 = adj.backward()
        assert dU2 == dU and np.array_equal(g2, g)
        out.append((E, g, dU))
    for E, g, dU in out[1:]:
        assert E == out[0][0] and dU == out[0][2] and np.array_equal(g, out[0][1])
    E, g, dU = out[0]
    assert abs(E - E_ref) <= 1e-11 * abs(E_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-8 * np.max(np.abs(g_ref))
    assert abs(dU - dU_ref) <= 1e-8 * abs(dU_ref)
    assert np.all(g[walls(n)] == 0.0)
    with make(n, Re, dt, 100) as adj:
        adj.forward(w0, U, steps)
        adj.backward()
        cold = adj.sweeps()
    print("sweeps warm", warm, "cold", cold, flush=True)
    assert warm[0] < cold[0] and warm[1] < cold[1]


def test_gpu_adjoint_backward_failure_ends_both_pipelines():
    # an adjoint-problem solve that hits itmax (a loose forward rule converges well within it, the
    # 1e-14 adjoint rule cannot) ends the backward pass with CNV_NOT_CONVERGED while the recompute
    # thread is still running or waiting: the call returns the error instead of hanging, and a
    # second call does the same
    n, steps, U, Re, dt = 24, 6, 1.0, 300.0, 4e-3
    p = cnv.default_params(n, Re=Re, dt=dt, beta=cnv.reference_beta(n), itmax=60, diagnostics=0)
    w0 = cnv.gaussian_vorticity(n, count=6, seed=3)
    with cnv.Adjoint(p, checkpoint_every=2, rtol_forward=1e-2, rtol_adjoint=1e-14) as adj:
        adj.forward(w0, U, steps)
        for _ in range(2):
            with pytest.raises(cnv.CnvError) as ei:
                adj.backward()
            assert ei.value.code == cnv.NOT_CONVERGED
            assert "adjoint-problem Poisson solve did not converge" in str(ei.value)


def test_gpu_adjoint_backward_refused_after_failed_forward():
    # a forward run whose solve hits itmax leaves no trajectory to differentiate
    n = 24
    p = cnv.default_params(n, Re=300.0, dt=4e-3, beta=cnv.reference_beta(n), itmax=60, diagnostics=0)
    with cnv.Adjoint(p, checkpoint_every=2, rtol_forward=1e-14, rtol_adjoint=1e-14) as adj:
        with pytest.raises(cnv.CnvError) as ei:
            adj.forward(cnv.gaussian_vorticity(n, count=6, seed=3), 1.0, 6)
        assert ei.value.code == cnv.NOT_CONVERGED
        with pytest.raises(cnv.CnvError) as ei:
            adj.backward()
        assert ei.value.code == cnv.ERR_ARG and "run cnv_adj_forward first" in str(ei.value)


def test_gpu_adjoint_through_torch_autograd():
    torch = pytest.importorskip("torch")
    n, steps, Re, dt = 16, 4, 300.0, 4e-3
    adj = make(n, Re, dt, K=2)

    class CavityEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, w0, U):
            return torch.tensor(adj.forward(w0.detach().numpy(), float(U), steps), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gE):
            g, dU = adj.backward()
            return gE * torch.from_numpy(g), gE * torch.tensor(dU, dtype=torch.float64)

    w0 = torch.from_numpy(cnv.gaussian_vorticity(n, count=5, seed=9)).requires_grad_(True)
    U = torch.tensor(1.2, dtype=torch.float64, requires_grad=True)
    E = CavityEnergy.apply(w0, U)
    E.backward()
    E_ref, g_ref, dU_ref = CavityAdjoint(n, Re, dt).gradient(w0.detach().numpy(), 1.2, steps)
    assert abs(float(E.detach()) - E_ref) <= 1e-11 * abs(E_ref)
    assert np.max(np.abs(w0.grad.numpy() - g_ref)) <= 1e-8 * np.max(np.abs(g_ref))
    assert abs(float(U.grad) - dU_ref) <= 1e-8 * abs(dU_ref)
    adj.close()


def test_gpu_adjoint_backward_twice():
    # the backward pass recomputes segments from the checkpoints, which moves the context's own
    # state; a second call after the same forward run must still seed from that run's final state
    n, steps, U, Re, dt = 20, 7, 1.0, 250.0, 4e-3
    w0 = cnv.gaussian_vorticity(n, count=6, seed=4)
    with make(n, Re, dt, K=3) as adj:
        adj.forward(w0, U, steps)
        g1, dU1 = adj.backward()
        g2, dU2 = adj.backward()
    assert dU1 == dU2 and np.array_equal(g1, g2)


def test_gpu_adjoint_workload_full_size():
    # the adjoint workload itself: 256^2, Re 400, dt = 20*2^-16, 2000 steps, w0 = the committed
    # 32 Gaussians (seed 1), lid U = 1, checkpoints every 100 steps; central differences
    # (eps = 1e-6) along one random direction and in U within 1e-5, and the same gradient with
    # checkpoints every 50 steps, and with warm-started solves within 1e-10
    n, steps, U, Re, dt = 256, 2000, 1.0, 400.0, 20 * 2.0 ** -16
    w0 = cnv.gaussian_vorticity(n)
    d = np.random.default_rng(11).standard_normal((n, n))
    eps = 1e-6
    with make(n, Re, dt, K=100) as adj:
        E = adj.forward(w0, U, steps)
        print("adjoint workload: E", E, flush=True)
        g, dU = adj.backward()
        print("adjoint workload: gradient done, dE/dU", dU, flush=True)
        fd = (adj.forward(w0 + eps * d, U, steps) - adj.forward(w0 - eps * d, U, steps)) / (2 * eps)
        print("adjoint workload: direction", float(np.sum(g * d)), fd, flush=True)
        fdU = (adj.forward(w0, U + eps, steps) - adj.forward(w0, U - eps, steps)) / (2 * eps)
        print("adjoint workload: lid", dU, fdU, flush=True)
    assert abs(np.sum(g * d) - fd) <= 1e-5 * abs(fd)
    assert abs(dU - fdU) <= 1e-5 * abs(fdU)
    assert np.all(g[walls(n)] == 0.0)
    with make(n, Re, dt, K=50) as adj:
        assert adj.forward(w0, U, steps) == E
        g50, dU50 = adj.backward()
    print("adjoint workload: K=50 done", flush=True)
    assert dU50 == dU and np.array_equal(g50, g)
    # warm-started solves: the same gradient within what the 1e-14 relative rule leaves
    with make(n, Re, dt, K=100, warm_start=1) as adj:
        Ew = adj.forward(w0, U, steps)
        gw, dUw = adj.backward()
    print("adjoint workload: warm start", Ew, dUw, flush=True)
    assert abs(Ew - E) <= 1e-11 * abs(E)
    assert np.max(np.abs(gw - g)) <= 1e-10 * np.max(np.abs(g)) and abs(dUw - dU) <= 1e-10 * abs(dU)
