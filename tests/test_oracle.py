"""CPU tests that pin the oracle (oracle/cnv_oracle.c) before it is trusted as the checker.

Pins, per oracle function:
  - bit-equality with the compiled reference's outputs (tests/golden/*.npz, written by
    tests/golden/make_golden.py from oracle/_ref);
  - the per-step (k, e, continuity) known answers in the reference's own run logs;
  - mathematics that a plausible slip would break: polynomial exactness of the FD rows, the
    literal dense Kronecker matvec, a manufactured discrete solution, a dense direct solve, and
    the two SOR orderings meeting at one fixed point.
"""
import glob
import os

import numpy as np
import pytest

from oracle import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def same_bits(a, b):
    """Bit-exact up to the sign of zero (the only freedom a stencil sum has, see
    src/linearalg.c:269-280: adding +-0 products)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


# --------------------------------------------------------------------------------------
# finite-difference tables (src/finitediff.c)
# --------------------------------------------------------------------------------------
def test_diff_tables_match_reference_bits():
    g = load("diff_tables.npz")
    assert len(g.files) == 36
    for key in g.files:
        which, n, o, h = key.split("_")
        n, o, h = int(n[1:]), int(o[1:]), float.fromhex(h[1:])
        if n < o + 2:
            continue
        D = (orc.diff1 if which == "d1" else orc.diff2)(n, o, h)
        assert same_bits(D, g[key]), key


@pytest.mark.parametrize("o", [2, 4, 6])
def test_diff1_polynomial_exactness(o):
    n, h = 16, 0.1
    x = np.arange(n) * h
    D = orc.diff1(n, o, h)
    # interior rows: exact for degree <= o; every row: exact for linears
    for deg in range(0, o + 1):
        f, df = x ** deg, deg * x ** max(deg - 1, 0) if deg else np.zeros(n)
        got = D @ f
        rows = range(n) if deg <= 1 else range(o // 2, n - o // 2)
        for r in rows:
            assert abs(got[r] - df[r]) <= 1e-8 * max(1.0, abs(df[r])) * 10 ** deg, (deg, r)


@pytest.mark.parametrize("o", [2, 4, 6])
def test_diff2_polynomial_exactness_and_row_sums(o):
    n, h = 16, 0.1
    x = np.arange(n) * h
    D = orc.diff2(n, o, h)
    assert np.all(np.abs(D.sum(axis=1)) < 1e-9 / h ** 2)
    for deg in range(0, 4):  # every row (one-sided ones included) is exact for cubics
        f = x ** deg
        d2 = deg * (deg - 1) * x ** max(deg - 2, 0) if deg >= 2 else np.zeros(n)
        assert np.allclose(D @ f, d2, rtol=1e-8, atol=1e-6), deg
    for deg in range(4, o + 2):  # interior rows: exact to degree o + 1
        f = x ** deg
        d2 = deg * (deg - 1) * x ** (deg - 2)
        r = slice(o // 2, n - o // 2)
        assert np.allclose((D @ f)[r], d2[r], rtol=1e-7), deg


def test_bad_order_rejected():
    with pytest.raises(ValueError):
        orc.diff1(16, 3, 0.1)


# --------------------------------------------------------------------------------------
# Kronecker matvec == stencil sum, bit for bit
# --------------------------------------------------------------------------------------
@pytest.mark.parametrize("n,o", [(8, 2), (8, 4), (8, 6), (9, 6), (12, 4)])
def test_stencil_equals_dense_kron(n, o):
    rng = np.random.default_rng(n * 10 + o)
    f = rng.standard_normal((n, n))
    h = 1.0 / n
    for D in (orc.diff1(n, o, h), orc.diff2(n, o, h)):
        assert same_bits(orc.apply_x(D, f), orc.dense_kron_apply(D, f, True))
        assert same_bits(orc.apply_y(D, f), orc.dense_kron_apply(D, f, False))
        # and the operators act along the documented axes (j for DX, i for DY)
        assert np.allclose(orc.apply_x(D, f), f @ D.T)
        assert np.allclose(orc.apply_y(D, f), D @ f)


# --------------------------------------------------------------------------------------
# elementwise pieces
# --------------------------------------------------------------------------------------
def test_euler_matches_reference_bits():
    g = load("elementwise.npz")
    w = g["in_w"].copy()
    orc.lib().orc_euler(w.shape[0], w, g["in_dwdx"], g["in_dwdy"], g["in_d2x"], g["in_d2y"],
                        g["in_u"], g["in_v"], float(g["Re"]), float(g["dt"]))
    assert same_bits(w, g["euler_out"])
    assert same_bits(g["in_u"] + g["in_v"], g["continuity_out"])
    # vorticity() returns its 2nd argument minus its 1st (src/fluiddyn.c:194)
    assert same_bits(g["in_v"] - g["in_u"], g["vorticity_out"])


# --------------------------------------------------------------------------------------
# Poisson (src/poisson.c)
# --------------------------------------------------------------------------------------
def test_poisson_matches_reference_bits():
    g = load("poisson_cases.npz")
    idx = sorted({k.split("_")[0] for k in g.files})
    assert len(idx) == 9
    for key in idx:
        n, beta, omp, tol, k = g[key + "_meta"]
        n, k = int(n), int(k)
        p = orc.make_params(n, tol=tol, itmax=200000,
                            poisson_type=1 if beta < 0 else 2,
                            ordering=orc.REDBLACK if omp else orc.LEX,
                            beta_=None if beta < 0 else beta)
        psi, st = orc.poisson(p, g[key + "_f"])
        assert st.status == 0
        assert st.k == k, key
        assert "%E" % st.e == str(g[key + "_e"]), key
        assert same_bits(psi, g[key + "_psi"]), key


def _laplace_matrix(n, dx, dy):
    """The linear system whose solution is the SOR fixed point of src/poisson.c:268:
    2(dx^2+dy^2) psi - dy^2 (N+S) - dx^2 (E+W) = -dx^2 dy^2 f on the interior, psi=0 on edges."""
    m = n - 2
    A = np.zeros((m * m, m * m))
    for i in range(m):
        for j in range(m):
            r = i * m + j
            A[r, r] = 2 * (dx * dx + dy * dy)
            for di, dj, c in ((1, 0, dy * dy), (-1, 0, dy * dy), (0, 1, dx * dx), (0, -1, dx * dx)):
                ii, jj = i + di, j + dj
                if 0 <= ii < m and 0 <= jj < m:
                    A[r, ii * m + jj] = -c
    return A


@pytest.mark.parametrize("ordering", [orc.LEX, orc.REDBLACK])
@pytest.mark.parametrize("ptype", [1, 2])
def test_poisson_dense_direct_solve(ordering, ptype):
    n = 12
    rng = np.random.default_rng(7)
    f = rng.standard_normal((n, n))
    p = orc.make_params(n, tol=1e-15, itmax=100000, poisson_type=ptype, ordering=ordering)
    psi, st = orc.poisson(p, f)
    assert st.status == 0
    A = _laplace_matrix(n, p.dx, p.dy)
    x = np.linalg.solve(A, -(p.dx ** 2 * p.dy ** 2) * f[1:-1, 1:-1].ravel())
    assert np.max(np.abs(psi[1:-1, 1:-1].ravel() - x)) <= 1e-12 * np.max(np.abs(x))
    assert np.all(psi[0] == 0) and np.all(psi[-1] == 0) and np.all(psi[:, 0] == 0) and np.all(psi[:, -1] == 0)


def test_poisson_manufactured_solution_anisotropic_spacing():
    # dx != dy pins which spacing multiplies which neighbour pair (src/poisson.c:268)
    n = 14
    rng = np.random.default_rng(3)
    p = orc.make_params(n, tol=1e-12, itmax=200000, ordering=orc.REDBLACK, Lx=2, Ly=1)
    psi_star = np.zeros((n, n))
    psi_star[1:-1, 1:-1] = rng.standard_normal((n - 2, n - 2))
    dx, dy = p.dx, p.dy
    f = np.zeros((n, n))
    # f such that the SOR fixed point is psi_star: f = (dy^2(N+S) + dx^2(E+W) - 2(dx^2+dy^2) psi)/(dx^2 dy^2)
    f[1:-1, 1:-1] = (dy * dy * (psi_star[2:, 1:-1] + psi_star[:-2, 1:-1])
                     + dx * dx * (psi_star[1:-1, 2:] + psi_star[1:-1, :-2])
                     - 2 * (dx * dx + dy * dy) * psi_star[1:-1, 1:-1]) / (dx * dx * dy * dy)
    psi, st = orc.poisson(p, f)
    assert st.status == 0
    assert np.max(np.abs(psi - psi_star)) < 1e-10
    # the same f with the spacings swapped is a different problem
    q = orc.make_params(n, tol=1e-12, itmax=200000, ordering=orc.REDBLACK, Lx=1, Ly=2)
    psi2, _ = orc.poisson(q, f)
    assert np.max(np.abs(psi2 - psi_star)) > 1e-3


def test_orderings_share_the_fixed_point():
    n = 24
    f = np.random.default_rng(11).standard_normal((n, n)) * 30
    a, sa = orc.poisson(orc.make_params(n, tol=1e-13, itmax=100000, ordering=orc.LEX), f)
    b, sb = orc.poisson(orc.make_params(n, tol=1e-13, itmax=100000, ordering=orc.REDBLACK), f)
    assert sa.status == 0 and sb.status == 0
    assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a))
    assert sa.sweeps != sb.sweeps  # different orderings really ran


def test_poisson_stopping_rule_and_failures():
    n = 16
    f = np.random.default_rng(5).standard_normal((n, n))
    p = orc.make_params(n, tol=1e-9, itmax=100000)
    _, st = orc.poisson(p, f)
    assert st.status == 0 and st.k == st.sweeps - 1 and st.e < 1e-9
    # the sweep before the last one was not yet below tol
    p.itmax = st.sweeps - 1
    _, st2 = orc.poisson(p, f)
    assert st2.status == 1 and st2.e >= 1e-9
    # poisson_max_it = 1 reaches the exit path (src/poisson.c:280-284)
    p.itmax = 1
    assert orc.poisson(p, f)[1].status == 1
    # NaN never converges (NaN < tol is false)
    f2 = f.copy()
    f2[5, 5] = np.nan
    p.itmax = 50
    assert orc.poisson(p, f2)[1].status == 1
    # invalid poisson_type (src/main.c:357-361)
    p.poisson_type = 3
    assert orc.poisson(p, f)[1].status == 3


# --------------------------------------------------------------------------------------
# full steps (src/main.c:276-451)
# --------------------------------------------------------------------------------------
def step_cases():
    return sorted(glob.glob(os.path.join(GOLD, "step_*.npz")))


def params_from_fixture(g):
    meta = {k[5:]: g[k] for k in g.files if k.startswith("meta_")}
    beta = float(meta["beta"])
    return orc.make_params(
        int(meta["n"]), Re=float(meta["Re"]), dt=float(meta["dt"]), order=int(meta["order"]),
        poisson_type=int(meta["poisson_type"]),
        ordering=orc.REDBLACK if int(g["ordering"]) else orc.LEX,
        itmax=int(meta["itmax"]), tol=float(meta["tol"]), beta_=None if beta < 0 else beta,
        Lx=int(meta["Lx"]), Ly=int(meta["Ly"]), ui=float(meta["ui"]), vi=float(meta["vi"]),
        u=tuple(float(x) for x in meta["ub"]), v=tuple(float(x) for x in meta["vb"])), int(meta["steps"])


@pytest.mark.parametrize("path", step_cases(), ids=lambda s: os.path.basename(s)[5:-4])
def test_steps_match_reference_bits(path):
    g = load(path)
    p, steps = params_from_fixture(g)
    s = orc.State(p)
    per = s.run(steps)
    assert [st.k for st in per] == list(g["k"])
    assert ["%E" % st.e for st in per] == list(g["e"])
    assert same_bits([st.cont_max for st in per], g["cmax"])
    assert same_bits([st.cont_min for st in per], g["cmin"])
    for name in ("w", "psi", "u", "v"):
        assert same_bits(getattr(s, name), g[name]), name


@pytest.mark.parametrize("run", ["testRunNOOMP", "testRunOMP", "testRunNOOMPHIGHRES", "testRunOMPHIGHRES"])
def test_reference_run_log_known_answers(run):
    g = load("log_kat.npz")
    n, Re, dt, tol, itmax, ordering = g[run + "_meta"]
    p = orc.make_params(int(n), Re=Re, dt=dt, order=6, tol=tol, itmax=int(itmax),
                        ordering=int(ordering))
    ks = g[run + "_k"]
    s = orc.State(p)
    per = s.run(len(ks))
    assert [st.k for st in per] == list(ks)
    assert ["%E" % st.e for st in per] == list(g[run + "_e"])
    assert ["%E" % st.cont_max for st in per] == list(g[run + "_cmax"])
    assert ["%E" % st.cont_min for st in per] == list(g[run + "_cmin"])
