"""GPU tests of the captured step (use_graph = 1): every step replays HIP graphs (pre-solve
launches, chunks of SOR launches, post-solve launches; in the refined form also the residual
and correction of each outer iteration) and must give exactly what the eager launches give,
step after step, with eager work and host synchronisation between the replays."""
import numpy as np
import pytest

import cnv

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,spl,ptype,poll", [(64, 8, 2, 32), (130, 4, 2, 5), (48, 8, 1, 8), (97, 2, 2, 16)])
def test_graph_step_equals_eager(n, spl, ptype, poll):
    kw = dict(Re=1000.0, dt=2e-4, tol=1e-8, itmax=200000, sweeps_per_launch=spl, poisson_type=ptype,
              launches_per_poll=poll, u=[0.0, 0.0, 0.0, 1.0])
    print("graph case n=%d spl=%d type=%d poll=%d" % (n, spl, ptype, poll), flush=True)
    with cnv.Cavity(cnv.default_params(n, **kw)) as eager, \
            cnv.Cavity(cnv.default_params(n, use_graph=1, **kw)) as graph:
        for t in range(4):
            se = eager.step(1)[0]
            sg = graph.step(1)[0]
            assert (sg.status, sg.k, sg.sweeps, sg.e) == (se.status, se.k, se.sweeps, se.e), t
            assert (sg.cont_max, sg.cont_min) == (se.cont_max, se.cont_min), t
            a, b = eager.get_state(), graph.get_state()
            for f in ("w", "psi", "u", "v"):
                assert np.array_equal(a[f], b[f]), (t, f)


def test_graph_multi_step_call():
    kw = dict(Re=400.0, dt=3e-4, tol=1e-9, itmax=200000)
    with cnv.Cavity(cnv.default_params(80, **kw)) as eager, \
            cnv.Cavity(cnv.default_params(80, use_graph=1, **kw)) as graph:
        se, sg = eager.step(5), graph.step(5)
        assert [s.k for s in se] == [s.k for s in sg]
        a, b = eager.get_state(), graph.get_state()
        assert all(np.array_equal(a[f], b[f]) for f in ("w", "psi", "u", "v"))


@pytest.mark.parametrize("warm,t32", [(0, 1), (1, 1), (0, 0)], ids=["cold", "warm", "stencils64"])
def test_graph_refined_step_equals_eager(warm, t32):
    # the refined form: per outer iteration a residual graph, replayed FP32 inner chunks (big and
    # tail) and a correction graph, the sweep budget on the device; transport and velocity
    # stencils in FP32 (default) or FP64 (transport_fp32 = 0), captured in the pre- and
    # post-solve graphs
    kw = dict(Re=1000.0, dt=1e-3, tol=1e-10, itmax=2_000_000, beta=1.9, precision=cnv.PREC_FP32_DC,
              warm_start=warm, transport_fp32=t32, ui=0.1, vi=-0.05, u=(0.2, -0.1, 0.05, 1.2),
              v=(-0.3, 0.1, 0.25, -0.15))
    print("refined graph case warm=%d transport_fp32=%d" % (warm, t32), flush=True)
    with cnv.Cavity(cnv.default_params(96, **kw)) as eager, \
            cnv.Cavity(cnv.default_params(96, use_graph=1, **kw)) as graph:
        for t in range(4):
            se = eager.step(1)[0]
            sg = graph.step(1)[0]
            assert (sg.status, sg.sweeps, sg.outer, sg.e) == (se.status, se.sweeps, se.outer, se.e), t
            a, b = eager.get_state(), graph.get_state()
            for f in ("w", "psi", "u", "v"):
                assert np.array_equal(a[f], b[f]), (t, f)


def test_graph_refined_budget():
    # a budget too small to converge: the device counter stops the inner solves, as the host does
    kw = dict(Re=1000.0, dt=2e-3, tol=1e-12, itmax=50, precision=cnv.PREC_FP32_DC)
    with cnv.Cavity(cnv.default_params(64, **kw)) as eager, \
            cnv.Cavity(cnv.default_params(64, use_graph=1, **kw)) as graph:
        se = eager.step(1, check=False)[0]
        sg = graph.step(1, check=False)[0]
    assert se.status == sg.status == cnv.NOT_CONVERGED
    assert sg.sweeps == se.sweeps
