"""GPU tests of the reference-compatible C API (poisson.h, fluiddyn.h) and of the cnavier driver.

The driver test is the end-to-end drop-in check: run with the reference's own sample
configuration, the per-step Poisson iteration counts and the continuity diagnostics it logs
equal those in the reference's OpenMP run log (red-black ordering) line for line.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import legacy_bind as lb

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DRIVER = os.path.join(lb.ROOT, "fluid-dynamics1_amd", "cnavier")
SOLVED = re.compile(r"solved with (\d+) iterations - root-sum-of-squares error: (\S+)")
CONT = re.compile(r"Continuity max: (\S+) \| Continuity min: (\S+) \|")


def same(a, b):
    return bool(np.all((np.asarray(a) == np.asarray(b))))


def test_legacy_poisson_sor_log_bits():
    g = np.load(os.path.join(GOLD, "poisson_cases.npz"))
    for key in sorted({k.split("_")[0] for k in g.files}):
        n, beta, omp, tol, k = g[key + "_meta"]
        if not omp:
            continue
        n = int(n)
        print("case", key, flush=True)
        psi, txt = lb.poisson_log(g[key + "_f"], 1.0 / n, 1.0 / n, 200000, float(tol),
                                  None if beta < 0 else float(beta))
        (kk, e), = SOLVED.findall(txt)
        assert int(kk) == int(k), key
        assert abs(float(e) - float(str(g[key + "_e"]))) <= 1e-6 * float(e)
        assert same(psi, g[key + "_psi"]), key


def test_poisson_host_and_device_entry_points_match_reference():
    # cnv_poisson_host (host arrays in and out) and cnv_poisson_device (device buffers: here a
    # second context's w holds f and its u receives psi) give the reference's OpenMP iterate and
    # logged iteration count, bit for bit, on the same golden cases
    import ctypes
    import cnv
    g = np.load(os.path.join(GOLD, "poisson_cases.npz"))
    for key in sorted({k.split("_")[0] for k in g.files}):
        n, beta, omp, tol, k = g[key + "_meta"]
        if not omp or beta < 0:
            continue
        n = int(n)
        print("case", key, flush=True)
        f = np.ascontiguousarray(g[key + "_f"], dtype=np.float64)
        p = cnv.default_params(n, beta=float(beta), tol=float(tol), itmax=200000)
        with cnv.Cavity(p) as cav, cnv.Cavity(p) as aux:
            rc, psi, kh, eh = cav.poisson(f)
            assert rc == cnv.OK and kh == int(k), (key, rc, kh)
            assert same(psi, g[key + "_psi"]), key
            aux.set_state(w=f)
            dw, du = ctypes.c_void_p(), ctypes.c_void_p()
            assert cnv.lib().cnv_device_state(aux.h, ctypes.byref(dw), None, ctypes.byref(du), None) == cnv.OK
            kd, ed = ctypes.c_int(), ctypes.c_double()
            assert cnv.lib().cnv_poisson_device(cav.h, dw, du, ctypes.byref(kd), ctypes.byref(ed)) == cnv.OK
            assert (kd.value, ed.value) == (kh, eh), key
            assert same(aux.get_state()["u"], psi), key


def test_legacy_euler_continuity_vorticity_bits():
    g = np.load(os.path.join(GOLD, "elementwise.npz"))
    ms = [lb.to_m(g["in_" + k]) for k in ("w", "dwdx", "dwdy", "d2x", "d2y", "u", "v")]
    lb.lib().euler(*ms, float(g["Re"]), float(g["dt"]))
    assert same(lb.from_m(ms[0]), g["euler_out"])
    U, V = lb.to_m(g["in_u"]), lb.to_m(g["in_v"])
    Cm, Vm = lb.lib().continuity(U, V), lb.lib().vorticity(U, V)
    assert same(lb.from_m(Cm), g["continuity_out"])
    assert same(lb.from_m(Vm), g["vorticity_out"])
    lb.free(*ms, U, V, Cm, Vm)


def test_legacy_poisson_not_converged_exits(tmp_path):
    code = ("import sys, numpy as np; sys.path.insert(0, %r); import legacy_bind as lb; "
            "f = np.random.default_rng(1).standard_normal((24, 24)); "
            "lb.poisson_log(f, 1/24, 1/24, 3, 1e-12, 1.5); print('not reached')") % (
        os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not reached" not in r.stdout


def _run_driver(tmp_path, cfg_text, name):
    (tmp_path / "cfg.txt").write_text(cfg_text)
    r = subprocess.run([DRIVER, "cfg.txt", name], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return (tmp_path / "output" / "logs" / (name + ".txt")).read_text()


@pytest.mark.parametrize("run,cfg", [
    ("testRunOMP", "Re = 1000.0\nnx = 64\nny = 64\ndt = 0.005\ntf = 20.0\nmax_co = 1.0\norder = 6\n"
                   "poisson_max_it = 10000\npoisson_tol = 1E-3\npoisson_type = 2\noutput_interval = 5\n"
                   "nsteps = 40\n"),
    ("testRunOMPHIGHRES", "Re = 5000.0\nnx = 128\nny = 128\ndt = 0.001\ntf = 10.0\nmax_co = 0.5\n"
                          "order = 6\npoisson_max_it = 15000\npoisson_tol = 5E-4\npoisson_type = 2\n"
                          "output_interval = 100\nnsteps = 11\n"),
    ("testRunOMPHIGHRES", "Re = 5000.0\nnx = 128\nny = 128\ndt = 0.001\ntf = 10.0\nmax_co = 0.5\n"
                          "order = 6\npoisson_max_it = 15000\npoisson_tol = 5E-4\npoisson_type = 2\n"
                          "output_interval = 100\nnsteps = 11\nuse_graph = 1\n"),
])
def test_driver_reproduces_reference_openmp_log(tmp_path, run, cfg):
    g = np.load(os.path.join(GOLD, "log_kat.npz"))
    log = _run_driver(tmp_path, cfg, "gpu")
    solved, cont = SOLVED.findall(log), CONT.findall(log)
    steps = len(solved)
    assert steps >= 11
    assert [int(k) for k, _ in solved] == list(g[run + "_k"][:steps])
    for (_, e), er in zip(solved, g[run + "_e"][:steps]):
        assert abs(float(e) - float(er)) <= 2e-6 * float(er)
    assert [a for a, _ in cont] == list(g[run + "_cmax"][:steps])
    assert [b for _, b in cont] == list(g[run + "_cmin"][:steps])
    assert "Run name: gpu" in log and "Grid size: " in log and "=== Timing Summary ===" in log
    if run == "testRunOMP":
        vtk = sorted(os.listdir(tmp_path / "output" / "gpu"))
        assert "stream-function-1-0.vtk" in vtk and "y-velocity-1-3.vtk" in vtk and len(vtk) == 32
