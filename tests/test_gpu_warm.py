"""GPU tests of the opt-in warm start (warm_start = 1: each Poisson solve starts from the previous
step's psi instead of 0, which the reference never does -- src/poisson.c:229): the first step is
the cold one bit for bit, later steps need fewer sweeps and agree with the cold run to far below
the precision form's bar (both solves stop at the same update norm), graph replay equals eager
launches, the refined form behaves alike, and a batch refuses the option."""
import numpy as np
import pytest

import cnv

pytestmark = pytest.mark.gpu

KW = dict(Re=1000.0, dt=2e-3, tol=1e-10, itmax=2_000_000, beta=1.9)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def run(n, steps, **kw):
    with cnv.Cavity(cnv.default_params(n, **KW, **kw)) as cav:
        st = cav.step(steps)
        return [s.sweeps for s in st], cav.get_state()


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_warm_start_fp64(graph):
    n, steps = 128, 6
    k_cold, s_cold = run(n, steps, use_graph=graph)
    k_warm, s_warm = run(n, steps, use_graph=graph, warm_start=1)
    print("sweeps cold", k_cold, "warm", k_warm, flush=True)
    assert k_warm[0] == k_cold[0]  # nothing to start from: the cold solve
    assert all(w < c for w, c in zip(k_warm[1:], k_cold[1:]))
    for f in ("psi", "w", "u", "v"):
        assert rel(s_warm[f], s_cold[f]) <= 1e-6, f


def test_warm_start_graph_equals_eager():
    _, a = run(96, 4, use_graph=0, warm_start=1)
    _, b = run(96, 4, use_graph=1, warm_start=1)
    for f in ("psi", "w", "u", "v"):
        assert np.array_equal(a[f], b[f]), f


def test_warm_start_refined():
    n, steps = 128, 4
    k_cold, s64 = run(n, steps)
    k_warm, s32 = run(n, steps, precision=cnv.PREC_FP32_DC, warm_start=1)
    print("sweeps fp64 cold", k_cold, "refined warm", k_warm, flush=True)
    assert all(w < c for w, c in zip(k_warm[1:], k_cold[1:]))
    for f in ("psi", "w", "u", "v"):
        assert rel(s32[f], s64[f]) <= 1e-6, f


def test_batch_refuses_warm_start():
    ps = [cnv.default_params(n, **KW) for n in (48, 64)]
    ps[1].warm_start = 1
    with pytest.raises(cnv.CnvError) as ei:
        cnv.Batch(ps)
    assert ei.value.code == cnv.ERR_ARG
