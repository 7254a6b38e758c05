"""Cross-checks against the compiled reference itself (oracle/_ref, built by `make -C oracle ref`
in a container that has the reference sources).  Skipped where oracle/_ref is absent.

1. The fixture harness (tests/golden/refbind.py), which calls the reference library functions in
   src/main.c's order, reproduces the unmodified reference binary's own log line for line.
2. The oracle equals the reference library on fresh random cases not in the fixtures.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import refbind
from oracle import orc

pytestmark = pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")

CONT = re.compile(r"Continuity max: (\S+) \| Continuity min: (\S+) \|")


def _run_binary(tmp_path, binary, n, order, Re, dt, tol, steps, ptype, omp):
    cfg = tmp_path / "cfg.txt"
    tf = (steps + 0.5) * dt
    cfg.write_text(
        "Re = %r\nnx = %d\nny = %d\ndt = %r\ntf = %r\norder = %d\npoisson_tol = %r\n"
        "poisson_max_it = 50000\npoisson_type = %d\noutput_interval = 100000\nopenmp_enabled = %d\n"
        % (Re, n, n, dt, tf, order, tol, ptype, 1 if omp else 0))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([binary, str(cfg), "x"], cwd=tmp_path, env=env, check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=300)
    txt = (tmp_path / "output" / "logs" / "x.txt").read_text()
    return refbind.parse_solved(txt), CONT.findall(txt)


@pytest.mark.parametrize("omp", [False, True])
def test_harness_reproduces_reference_binary(tmp_path, omp):
    binary = os.path.join(refbind.REF_DIR, "cnavier_ref_omp" if omp else "cnavier_ref")
    if not os.path.exists(binary):
        pytest.skip("reference binary not built")
    n, order, Re, dt, tol, steps = 24, 6, 700.0, 0.004, 1e-7, 6
    solved, cont = _run_binary(tmp_path, binary, n, order, Re, dt, tol, steps, 2, omp)
    assert len(solved) == steps
    _, per = refbind.Ref(omp=omp).run_steps(n, order, Re, dt, tol, 50000, steps)
    assert [(k, e) for k, e, _, _ in per] == solved
    assert [("%E" % a, "%E" % b) for _, _, a, b in per] == cont


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_equals_reference_library_random(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(10, 30))
    order = int(rng.choice([2, 4, 6]))
    omp = bool(seed % 2)
    beta = float(rng.uniform(1.0, 1.9))
    ub = tuple(float(x) for x in rng.uniform(-1, 1, 4))
    vb = tuple(float(x) for x in rng.uniform(-1, 1, 4))
    os.environ["OMP_NUM_THREADS"] = "1"
    (w, psi, u, v), per = refbind.Ref(omp=omp).run_steps(
        n, order, 500.0, 0.002, 1e-9, 50000, 5, beta=beta, ub=ub, vb=vb)
    p = orc.make_params(n, Re=500.0, dt=0.002, order=order, tol=1e-9, itmax=50000, beta_=beta,
                        ordering=orc.REDBLACK if omp else orc.LEX, u=ub, v=vb)
    s = orc.State(p)
    got = s.run(5)
    assert [g.k for g in got] == [q[0] for q in per]
    for name, ref in zip(("w", "psi", "u", "v"), (w, psi, u, v)):
        assert np.array_equal(getattr(s, name), ref), name
