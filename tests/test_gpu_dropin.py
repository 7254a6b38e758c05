"""The drop-in route itself: the reference's unmodified src/main.c, compiled against include/ and
linked with libcnv.so in place of the reference's other src/*.c (oracle/_ref/cnavier_main_on_libcnv,
built by `make -C oracle ref` where the reference sources exist; it travels with the built
libraries), runs the reference's 64 x 64 sample configuration with its Poisson solves, Euler
updates and continuity diagnostics on the GPU, and logs the same per-step Poisson index k, L1
update e (printed with 6 digits) and continuity max/min as the reference's own OpenMP run
(output/logs/testRunOMP.txt of the reference, tests/golden/log_kat.npz) for its 40 steps.
Skipped where the binary is absent."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "oracle", "_ref", "cnavier_main_on_libcnv")
SOLVED = re.compile(r"solved with (\d+) iterations - root-sum-of-squares error: (\S+)")
CONT = re.compile(r"Continuity max: (\S+) \| Continuity min: (\S+) \|")

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref not built")]


def test_reference_main_on_libcnv_reproduces_openmp_log(tmp_path):
    # the reference's sample configuration (Re 1000, 64 x 64, dt 0.005, order 6, SOR with the
    # reference's relaxation rule to an L1 update below 1e-3), cut to 40 steps through tf:
    # it_max = (int)(tf/dt - 1) = 39 (src/main.c:162), steps 0..39
    cfg = ("Re = 1000.0\nnx = 64\nny = 64\ndt = 0.005\ntf = 0.2025\nmax_co = 1.0\norder = 6\n"
           "poisson_max_it = 10000\npoisson_tol = 1E-3\npoisson_type = 2\noutput_interval = 5\n")
    (tmp_path / "cfg.txt").write_text(cfg)
    print("reference main.c on libcnv: 64 x 64, 40 steps", flush=True)
    r = subprocess.run([BIN, "cfg.txt", "main"], cwd=tmp_path, capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = (tmp_path / "output" / "logs" / "main.txt").read_text()
    g = np.load(os.path.join(GOLD, "log_kat.npz"))
    solved, cont = SOLVED.findall(log), CONT.findall(log)
    assert len(solved) == 40 and len(cont) == 40, (len(solved), len(cont))
    assert [int(k) for k, _ in solved] == list(g["testRunOMP_k"][:40])
    for (_, e), er in zip(solved, g["testRunOMP_e"][:40]):
        assert abs(float(e) - float(er)) <= 2e-6 * float(er)
    assert [a for a, _ in cont] == list(g["testRunOMP_cmax"][:40])
    assert [b for _, b in cont] == list(g["testRunOMP_cmin"][:40])
    # VTK output every 5 steps, four fields each, through the library's printvtk
    vtk = sorted(os.listdir(tmp_path / "output" / "main"))
    assert len(vtk) == 32 and "stream-function-1-0.vtk" in vtk
