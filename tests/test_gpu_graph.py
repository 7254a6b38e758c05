"""GPU tests of the captured step (use_graph = 1): every step replays HIP graphs (pre-solve
launches, chunks of SOR launches, post-solve launches) and must give exactly what the eager
launches give, step after step, with eager work and host synchronisation between the replays."""
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
