"""Every compiled SOR launch shape (sweeps per launch S, waves per workgroup, rows per wave) on the
single-cavity path, bit-exact against the oracle: steps with wall velocities on a grid with n
even and one with n odd (where the lane of column n-1 reads the pair before it and takes its
second element), and red-black Gauss-Seidel;
the refined (FP32 + FP64) form of each shape against the FP64 result; a shape that is not
compiled is refused."""
import numpy as np
import pytest

import cnv
from oracle import orc

pytestmark = pytest.mark.gpu

# S, waves, rows per wave: the shapes the library compiles
SHAPES = [(1, 4, 10), (2, 4, 10), (4, 8, 8), (6, 8, 8), (8, 8, 10), (4, 16, 4), (5, 16, 4), (6, 16, 4),
          (5, 12, 6), (5, 8, 4), (5, 8, 8), (4, 8, 4), (5, 4, 6)]
WALLS = dict(ui=0.1, vi=-0.05, u=(0.2, -0.1, 0.05, 1.2), v=(-0.3, 0.1, 0.25, -0.15))


def shape_kw(sh):
    return dict(sweeps_per_launch=sh[0], tile_waves=sh[1], wave_rows=sh[2])


def oracle_state(n, kw, steps):
    kw = dict(kw)
    u, v = kw.pop("u"), kw.pop("v")
    beta = kw.pop("beta", None)
    s = orc.State(orc.make_params(n, ordering=orc.REDBLACK, beta_=beta, u=tuple(u), v=tuple(v), **kw))
    return s, s.run(steps)


@pytest.mark.parametrize("n,ptype", [(150, 2), (151, 2), (77, 1)], ids=["n150_sor", "n151_sor", "n77_gs"])
def test_every_shape_bit_exact_vs_oracle(n, ptype):
    kw = dict(order=6, poisson_type=ptype, Re=800.0, dt=0.001, tol=1e-10, itmax=200000, **WALLS)
    if ptype == 2:
        kw["beta"] = 1.85
    s_or, per = oracle_state(n, kw, 2)
    for sh in SHAPES:
        print("shape", sh, "n", n, "poisson_type", ptype, flush=True)
        with cnv.Cavity(cnv.default_params(n, **kw, **shape_kw(sh))) as cav:
            assert cav.sweeps_per_launch == sh[0]
            st = cav.step(2)
            s = cav.get_state()
        assert [x.k for x in st] == [p.k for p in per], sh
        for name in ("w", "psi", "u", "v"):
            a, b = s[name], getattr(s_or, name)
            assert np.array_equal(a, b), (sh, name)


def test_every_shape_refined_form():
    n = 120
    kw = dict(order=6, Re=1000.0, dt=1e-3, tol=1e-10, itmax=2_000_000, beta=1.9, **WALLS)
    with cnv.Cavity(cnv.default_params(n, **kw)) as cav:
        cav.step(2)
        s64 = cav.get_state()
    for sh in SHAPES:
        print("refined shape", sh, flush=True)
        with cnv.Cavity(cnv.default_params(n, precision=cnv.PREC_FP32_DC, **kw, **shape_kw(sh))) as cav:
            st = cav.step(2)
            s32 = cav.get_state()
        assert all(x.status == cnv.OK for x in st), sh
        for name in ("psi", "w", "u", "v"):
            err = np.max(np.abs(s32[name] - s64[name])) / np.max(np.abs(s64[name]))
            assert err <= 1e-6, (sh, name, err)


def test_default_shape_follows_the_grid():
    # thin tiles on small grids (the GPU is otherwise idle), thick ones from 1024^2 on; the
    # results do not depend on the shape (bit-exact above), only the speed
    for n, want in ((64, 5), (256, 5), (768, 6), (1024, 5)):
        with cnv.Cavity(cnv.default_params(n)) as cav:
            assert cav.sweeps_per_launch == want


def test_uncompiled_shape_refused():
    with pytest.raises(cnv.CnvError) as ei:
        cnv.Cavity(cnv.default_params(64, sweeps_per_launch=3, tile_waves=16, wave_rows=4))
    assert ei.value.code == cnv.ERR_ARG
