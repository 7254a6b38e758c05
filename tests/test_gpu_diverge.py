"""GPU tests of divergence handling.  The reference has none: a solve whose iterate turns NaN
never meets its stopping rule and runs to itmax before exit(1) (src/poisson.c:272-284).  Here a
non-finite L1 update ends the solve at once with CNV_DIVERGED; a cavity whose largest |u| or |v|
exceeds its blowup_speed (the explicit scheme going unstable) is reported the same way; in a
batch such a cavity drops out and every other cavity goes on exactly as it would without it."""
import numpy as np
import pytest

import cnv

pytestmark = pytest.mark.gpu

SHAPE = dict(sweeps_per_launch=5, tile_waves=4, wave_rows=6)
S = 5
# launches a non-finite update can take to be seen: the launch that made it, decided by the next
# one; a one-cavity context's FP32 inner solves (refined form) test the last sweep of every fourth
# launch only (launches 3, 7, ...), so theirs is seen after the fourth launch; a batch tests every
# launch


def max_sweeps_to_diverged(precision, batch=False):
    return (4 if precision == cnv.PREC_FP32_DC and not batch else 2) * S


def params(n, Re, U, dt, precision=cnv.PREC_FP64, **kw):
    kw = {**SHAPE, **kw}
    return cnv.default_params(n, Re=Re, dt=dt, beta=cnv.reference_beta(n), tol=1e-8, itmax=2_000_000,
                              precision=precision, u=[0.0, 0.0, 0.0, U], **kw)


MIX = [(32, 400.0, 1.0, 2e-4), (48, 1000.0, 0.8, 1e-4), (40, 250.0, 1.2, 3e-4), (64, 100.0, 1.1, 2e-4)]


@pytest.mark.parametrize("precision", [cnv.PREC_FP64, cnv.PREC_FP32_DC], ids=["fp64", "refined"])
@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_nan_vorticity_ends_the_step_within_a_few_launches(precision, graph):
    p = params(48, 1000.0, 1.0, 1e-4, precision=precision, use_graph=graph)
    print("one cavity, NaN vorticity, precision", precision, "graph", graph, flush=True)
    with cnv.Cavity(p) as cav:
        cav.step(2)
        w = cav.get_state()["w"]
        w[20, 17] = np.nan
        cav.set_state(w=w)
        st = cav.step(1, check=False)[0]
        with pytest.raises(cnv.CnvError) as ei:
            cav.step(1)
    assert st.status == cnv.DIVERGED
    assert 0 < st.sweeps <= max_sweeps_to_diverged(precision)
    assert ei.value.code == cnv.DIVERGED


def test_blowup_speed_one_cavity():
    p = params(32, 400.0, 1.0, 2e-4, blowup_speed=1e-12)  # any motion at all exceeds it
    with cnv.Cavity(p) as cav:
        st = cav.step(1, check=False)[0]
    assert st.status == cnv.DIVERGED and st.sweeps > 0  # the solve converged; the speed check fired


def run_batch(ps, steps, inject=None, at=1):
    """Steps a batch `at` steps, optionally puts NaN into cavity `inject`'s vorticity, steps the
    rest; returns the stats and the final states."""
    with cnv.Batch(ps) as b:
        st = b.step(at)
        if inject is not None:
            w = b.get_state(inject)["w"]
            w[ps[inject].n // 2, ps[inject].n // 3] = np.nan
            b.set_state(inject, w=w)
        st += b.step(steps - at)
        status = b.status()
        states = [b.get_state(i) for i in range(len(ps))]
    return st, status, states


@pytest.mark.parametrize("precision", [cnv.PREC_FP64, cnv.PREC_FP32_DC], ids=["fp64", "refined"])
def test_batch_nan_cavity_drops_out_others_unchanged(precision):
    ps = [params(n, Re, U, dt, precision=precision) for n, Re, U, dt in MIX]
    print("batch", [p.n for p in ps], "precision", precision, flush=True)
    ref_st, ref_status, ref = run_batch(ps, 4)
    st, status, got = run_batch(ps, 4, inject=1, at=1)
    assert ref_status == [cnv.OK] * 4
    assert status == [cnv.OK, cnv.DIVERGED, cnv.OK, cnv.OK]
    assert st[1][1].status == cnv.DIVERGED and 0 < st[1][1].sweeps <= max_sweeps_to_diverged(precision, batch=True)
    for t in (2, 3):  # left out of the later steps
        assert st[t][1].status == cnv.DIVERGED and st[t][1].sweeps == 0
    for i in (0, 2, 3):
        for t in range(4):
            assert (st[t][i].status, st[t][i].sweeps, st[t][i].e) == (ref_st[t][i].status, ref_st[t][i].sweeps,
                                                                     ref_st[t][i].e), (i, t)
        for f in ("w", "psi", "u", "v"):
            assert np.array_equal(got[i][f], ref[i][f]), (i, f)
    assert np.isnan(got[1]["w"]).any()


def test_batch_blowup_speed_and_revival():
    # cavity 2's blow-up speed is below any motion: it is dropped after its first step, keeps that
    # step's fields, and set_state brings it back
    ps = [params(n, Re, U, dt) for n, Re, U, dt in MIX]
    ps[2].blowup_speed = 1e-12
    for i in (0, 1, 3):
        ps[i].blowup_speed = 100.0
    with cnv.Batch(ps) as b:
        st = b.step(3)
        assert b.status() == [cnv.OK, cnv.OK, cnv.DIVERGED, cnv.OK]
        assert st[0][2].status == cnv.DIVERGED and st[0][2].sweeps > 0
        assert st[1][2].sweeps == 0 and st[2][2].sweeps == 0
        s2 = b.get_state(2)
        with cnv.Cavity(params(*MIX[2])) as alone:
            alone.step(1)
            one = alone.get_state()
        for f in ("w", "psi", "u", "v"):
            assert np.array_equal(s2[f], one[f]), f
        b.set_state(2, w=s2["w"], u=s2["u"], v=s2["v"])
        assert b.status() == [cnv.OK] * 4
        st2 = b.step(1)
        assert st2[0][2].status == cnv.DIVERGED  # it moves again, above 1e-12
    with cnv.Batch([params(*m) for m in MIX]) as b:  # no threshold: all live
        b.step(2)
        assert b.status() == [cnv.OK] * 4
