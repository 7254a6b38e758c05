"""GPU tests at the workload's full size and beyond: 1024^2 steps bit-exact against the oracle
(the metric's first step, and two steps with the reference's relaxation rule), size-independent
properties of a 1024^2 step, and bit-exact Poisson iterates on 4096^2-class grids (thousands of
workgroups, several per CU, both access paths)."""
import numpy as np
import pytest

import cnv
from oracle import orc

pytestmark = pytest.mark.gpu

N, RE, DT, OMEGA, TOL = 1024, 1000.0, 50 * 2.0 ** -20, 1.95, 1e-10


def rb_sweep_update(psi, f, dx, dy, beta):
    """One red-black SOR sweep (src/poisson.c:268, red points first); returns the L1 update."""
    p = psi.copy()
    c1, c2, c3 = dy * dy, dx * dx, dx * dx * dy * dy
    B = 2 * (dx * dx + dy * dy)
    i, j = np.meshgrid(np.arange(1, p.shape[0] - 1), np.arange(1, p.shape[1] - 1), indexing="ij")
    total = 0.0
    for colour in (0, 1):
        m = ((i + j) & 1) == colour
        ii, jj = i[m], j[m]
        old = p[ii, jj].copy()
        new = beta * (c1 * (p[ii + 1, jj] + p[ii - 1, jj]) + c2 * (p[ii, jj + 1] + p[ii, jj - 1])
                      - c3 * f[ii, jj]) / B + (1 - beta) * old
        p[ii, jj] = new
        total += float(np.sum(np.abs(new - old)))
    return total


def test_fullsize_step_properties():
    kw = dict(Re=RE, dt=DT, order=6, beta=OMEGA, tol=TOL, itmax=50_000_000)
    print("1024^2 FP64 step", flush=True)
    with cnv.Cavity(cnv.default_params(N, **kw)) as cav:
        st = cav.step(1)[0]
        s64 = cav.get_state()
    # the returned iterate is the first with a sweep update below tol: one more sweep moves it
    # by a little less again (the asymptotic SOR contraction), never by more than tol
    nxt = rb_sweep_update(s64["psi"], -s64["w"], 1.0 / N, 1.0 / N, OMEGA)
    print("sweeps", st.sweeps, "last update", st.e, "next update", nxt, flush=True)
    assert st.status == cnv.OK and st.e < TOL
    assert 0.5 * st.e < nxt < TOL
    # walls of psi stay exactly 0; continuity stays at round-off
    assert not np.any(s64["psi"][[0, -1], :]) and not np.any(s64["psi"][:, [0, -1]])
    assert max(abs(st.cont_max), abs(st.cont_min)) < 1e-8
    print("1024^2 FP32-refined step", flush=True)
    with cnv.Cavity(cnv.default_params(N, precision=cnv.PREC_FP32_DC, **kw)) as cav:
        sd = cav.step(1)[0]
        s32 = cav.get_state()
    assert sd.status == cnv.OK and sd.outer >= 2
    for f in ("psi", "w", "u", "v"):
        err = np.max(np.abs(s32[f] - s64[f])) / np.max(np.abs(s64[f]))
        assert err <= 1e-6, (f, err)


def test_fullsize_step_bit_exact_reference_beta():
    # the metric's cavity (1024^2, Re 1000, dt = 50*2^-20, order 6, tol 1e-10) with the reference's
    # own relaxation rule (src/main.c:134, beta = 1.9939 here): about 5.5k sweeps per step, which
    # the oracle's OpenMP red-black sweeps finish in seconds.  Two steps, bit for bit: 240 tiles
    # per launch, the convergence decision over their partials, the rollback into the first sweep
    # with e < tol, all at full size.
    import os
    beta = cnv.reference_beta(N)
    kw = dict(Re=RE, dt=DT, order=6, tol=TOL, itmax=2_000_000)
    threads = min(16, os.cpu_count() or 1)
    print("1024^2 reference-beta steps, oracle on %d threads" % threads, flush=True)
    s_or = orc.State(orc.make_params(N, ordering=orc.REDBLACK, beta_=beta, threads=threads, **kw))
    per = s_or.run(2)
    with cnv.Cavity(cnv.default_params(N, beta=beta, **kw)) as cav:
        st = cav.step(2)
        s = cav.get_state()
    print("k", [x.k for x in st], [p.k for p in per], flush=True)
    assert [x.k for x in st] == [p.k for p in per]
    assert all(abs(x.e - p.e) <= 1e-12 * p.e for x, p in zip(st, per))
    for name in ("w", "psi", "u", "v"):
        assert np.array_equal(s[name], getattr(s_or, name)), name


def test_fullsize_metric_step_bit_exact():
    # the metric's configuration exactly (omega = 1.95): the first step from rest, 46352 sweeps
    # (about 9.3k launches of 5), against the oracle's OpenMP red-black step (about 35 s on 16
    # host threads), bit for bit
    import os
    kw = dict(Re=RE, dt=DT, order=6, tol=TOL, itmax=50_000_000)
    threads = min(16, os.cpu_count() or 1)
    print("1024^2 omega 1.95 step, oracle on %d threads" % threads, flush=True)
    s_or = orc.State(orc.make_params(N, ordering=orc.REDBLACK, beta_=OMEGA, threads=threads, **kw))
    per = s_or.run(1)
    with cnv.Cavity(cnv.default_params(N, beta=OMEGA, use_graph=1, **kw)) as cav:
        st = cav.step(1)
        s = cav.get_state()
    print("sweeps", st[0].sweeps, per[0].sweeps, flush=True)
    assert st[0].k == per[0].k and abs(st[0].e - per[0].e) <= 1e-12 * per[0].e
    for name in ("w", "psi", "u", "v"):
        assert np.array_equal(s[name], getattr(s_or, name)), name


@pytest.mark.parametrize("n", [4096, 4099])
def test_large_grid_poisson_iterates_bit_exact(n):
    # 7 sweeps on a random source: 4096 takes the paired-access, exact-reciprocal path, 4099 the
    # per-column, corrected-division path; both need thousands of workgroups
    f = np.random.default_rng(n).standard_normal((n, n))
    print("poisson n=%d" % n, flush=True)
    with cnv.Cavity(cnv.default_params(n, tol=1e-300, itmax=7, diagnostics=0)) as cav:
        rc, psi, k, e = cav.poisson(f)
    assert rc == cnv.NOT_CONVERGED
    psi_or, st_or = orc.poisson(orc.make_params(n, tol=1e-300, itmax=7, ordering=orc.REDBLACK, threads=8), f)
    assert st_or.sweeps == 7
    assert np.array_equal(psi, psi_or)
    # the last sweep's L1 update, summed over thousands of per-workgroup partials (several per
    # thread in the decision's reduction): equal up to the summation order
    assert abs(e - st_or.e) <= 1e-12 * st_or.e, (e, st_or.e)


def test_fullsize_benchmark_steps_match_oracle_digests():
    # bench.py's warm-up step and the four steps it times (1024^2, omega = 1.95, tol 1e-10,
    # graph-replayed: 46352, 48221, 49309, 50077 and 50668 sweeps), each equal to the oracle's
    # step in k, e (to 1e-12: another summation order), continuity and every field, through
    # SHA-256 digests (tests/golden/metric_steps_1024.npz, written by
    # tests/golden/make_metric_steps.py)
    import hashlib
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "metric_steps_1024.npz"))
    n, Re, dt, order, omega, tol, itmax, steps = g["params"]
    print("1024^2 benchmark steps against the oracle's digests", flush=True)

    def digest(a):
        return hashlib.sha256(np.ascontiguousarray(a + 0.0, dtype="<f8").tobytes()).digest()

    p = cnv.default_params(int(n), Re=float(Re), dt=float(dt), order=int(order), beta=float(omega),
                           tol=float(tol), itmax=int(itmax), use_graph=1)
    with cnv.Cavity(p) as cav:
        for t in range(int(steps)):
            st = cav.step(1)[0]
            s = cav.get_state()
            print("step %d: k %d (oracle %d)" % (t, st.k, g["k"][t]), flush=True)
            assert st.k == g["k"][t]
            assert abs(st.e - g["e"][t]) <= 1e-12 * g["e"][t]
            assert st.cont_max == g["cont_max"][t] and st.cont_min == g["cont_min"][t]
            for i, f in enumerate(("w", "psi", "u", "v")):
                assert digest(s[f]) == bytes(g["digests"][t][i]), (t, f)
