"""GPU tests of the batched form (cnv_batch_*): many cavities of different sizes, Reynolds numbers,
lid velocities, time steps and relaxation factors advanced together must give, cavity by
cavity, exactly what the same cavity gives alone (cnv_create / cnv_step) with the same launch
shape -- the batched kernels run the one-cavity kernels' per-point arithmetic and each cavity
keeps its own Poisson state.  The one-cavity path is itself pinned bit for bit to the oracle
(test_gpu_parity.py), so this chain pins every batched cavity to the oracle too; one cavity is
also checked against the oracle directly."""
import numpy as np
import pytest

import cnv
from oracle import orc

pytestmark = pytest.mark.gpu

# (n, Re, U, dt, beta): power-of-two and other sides, one taller than a tile, two sharing a side
MIX = [(24, 400.0, 1.0, 2e-4, 1.7), (33, 1000.0, 0.7, 1e-4, None), (64, 250.0, 1.3, 3e-4, None),
       (40, 3000.0, 0.9, 5e-5, 1.85), (130, 100.0, 1.1, 2e-4, None), (40, 150.0, 1.2, 2e-4, 1.6)]


def mix_params(precision=cnv.PREC_FP64, tol=1e-8, **kw):
    out = []
    for n, Re, U, dt, beta in MIX:
        out.append(cnv.default_params(n, Re=Re, dt=dt, beta=cnv.reference_beta(n) if beta is None else beta,
                                      tol=tol, itmax=200000, precision=precision, u=[0.0, 0.0, 0.0, U], **kw))
    return out


def same(a, b):
    return bool(np.array_equal(a, b))


def run_alone(p, steps):
    with cnv.Cavity(p) as cav:
        st = cav.step(steps)
        return st, cav.get_state()


@pytest.mark.parametrize("shape", [(5, 16, 4), (5, 4, 6)], ids=["s5_16x4", "s5_4x6"])
@pytest.mark.parametrize("precision,steps,every", [(cnv.PREC_FP64, 3, 0), (cnv.PREC_FP32_DC, 2, 1),
                                                   (cnv.PREC_FP32_DC, 2, 4)], ids=["fp64", "refined1", "refined4"])
def test_batch_equals_single(precision, steps, every, shape):
    # the same launch shape on both sides, named explicitly (a one-cavity default follows the
    # cavity's size): the reported L1 update is summed per tile; and the same stopping-test
    # cadence of the FP32 inner solves (the defaults differ: 4 for a context, 1 for a batch)
    ps = mix_params(precision, sweeps_per_launch=shape[0], tile_waves=shape[1], wave_rows=shape[2],
                    inner_test_every=every)
    print("batch of", [p.n for p in ps], "precision", precision, "shape", shape, flush=True)
    check_batch_equals_single(ps, steps)


def test_batch_fp64_stencils_equals_single():
    # the refined form with transport_fp32 = 0: the batched transport and velocity kernels' FP64
    # instantiations (the other batch tests run the default FP32 stencils)
    ps = mix_params(cnv.PREC_FP32_DC, transport_fp32=0, sweeps_per_launch=5, tile_waves=16, wave_rows=4,
                    inner_test_every=1)
    print("batch of", [p.n for p in ps], "refined, FP64 stencils", flush=True)
    check_batch_equals_single(ps, 2)


def check_batch_equals_single(ps, steps):
    with cnv.Batch(ps) as b:
        bst = b.step(steps)
        bstate = [b.get_state(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        st, state = run_alone(p, steps)
        for t in range(steps):
            assert bst[t][i].status == cnv.OK
            assert (bst[t][i].k, bst[t][i].sweeps, bst[t][i].outer) == (st[t].k, st[t].sweeps, st[t].outer), (i, t)
            assert bst[t][i].e == st[t].e
            assert (bst[t][i].cont_max, bst[t][i].cont_min) == (st[t].cont_max, st[t].cont_min)
        for f in ("w", "psi", "u", "v"):
            assert same(bstate[i][f], state[f]), (i, f)


def test_batch_cavity_matches_oracle():
    # cavity 1 of the mix (n = 33, reference beta), two steps, against the CPU restatement
    ps = mix_params()
    with cnv.Batch(ps) as b:
        bst = b.step(2)
        got = b.get_state(1)
    n, Re, U, dt, _ = MIX[1]
    op = orc.make_params(n, Re=Re, dt=dt, order=6, poisson_type=2, ordering=orc.REDBLACK, itmax=200000,
                         tol=1e-8, beta_=cnv.reference_beta(n), u=(0.0, 0.0, 0.0, U))
    st = orc.State(op)
    stats = st.run(2)
    assert [s.k for s in stats] == [bst[t][1].k for t in range(2)]
    for f in ("w", "psi", "u", "v"):
        assert same(got[f], getattr(st, f)), f


def test_batch_autotune_keeps_results():
    ps = mix_params()
    with cnv.Batch(ps) as b:
        shape = b.autotune(launches=6)
        print("autotuned", shape, flush=True)
        assert shape["sweeps_per_launch"] in (1, 2, 4, 5, 6, 8) and shape["ms_per_sweep"] > 0
        bst = b.step(2)
        got = [b.get_state(i) for i in range(len(ps))]
    for i, p in enumerate(ps):
        p.sweeps_per_launch = shape["sweeps_per_launch"]
        p.tile_waves = shape["tile_waves"]
        p.wave_rows = shape["wave_rows"]
        st, state = run_alone(p, 2)
        assert [bst[t][i].k for t in range(2)] == [s.k for s in st]
        for f in ("w", "psi", "u", "v"):
            assert same(got[i][f], state[f]), (i, f)


def test_batch_reports_a_failing_cavity():
    ps = mix_params()
    ps[2].itmax = 5  # cannot converge to 1e-8 in 5 sweeps
    with cnv.Batch(ps) as b:
        with pytest.raises(cnv.CnvError) as ei:
            b.step(1)
        assert ei.value.code == cnv.NOT_CONVERGED
        st = b.step(1, check=False)[0]
    assert st[2].status == cnv.NOT_CONVERGED
    assert all(st[i].status == cnv.OK for i in range(len(ps)) if i != 2)


def test_batch_workload_table_runs():
    # the 64-cavity workload itself (seed 0), one FP32-refined step
    rows = cnv.batch_table()
    assert len(rows) == 64 and sum(r["n"] ** 2 for r in rows) == 4916224
    ps = cnv.batch_params(rows, precision=cnv.PREC_FP32_DC)
    with cnv.Batch(ps) as b:
        st = b.step(1)[0]
    assert all(s.status == cnv.OK and s.outer >= 1 for s in st)
