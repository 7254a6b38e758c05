"""Times the reference's own red-black SOR on the host cores, for a bounded number of sweeps.

Test infrastructure only: used by bench.py's cpu_baseline leg (and its CPU test) as a measured
baseline, never by the product.  The library timed is oracle/_ref/libcnavier_ref_omp.so, which
oracle/Makefile ("make -C oracle ref") compiles from the reference's own sources with its
Makefile's flags ("-g -Wall -fopenmp", no optimisation level) plus -DOPENMP_ENABLED, the build its
OpenMP run logs come from.

poisson_SOR_log (src/poisson.c:224-285) returns only on convergence and otherwise ends the process
with exit(1) after itmax sweeps (src/poisson.c:280-284).  A sample therefore runs in a child
process with tol < 0 (never met): the child reads CLOCK_MONOTONIC just before the call and prints
it, and the parent reads the same clock once the child has exited.  The per-sweep time includes
what the reference does every sweep: the copy of the iterate (mtrxcpy), the two colour passes and
the serial L1 update norm (error, src/poisson.c:34-45).
"""
import ctypes as C
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libcnavier_ref_omp.so")


class Mtrx(C.Structure):  # the reference's mtrx (include/linearalg.h:6-11)
    _fields_ = [("M", C.POINTER(C.POINTER(C.c_double))), ("m", C.c_int), ("n", C.c_int)]


def available():
    return os.path.exists(LIB)


def _child(lib_path, f_path, sweeps, beta):
    L = C.CDLL(lib_path)
    L.initm.restype = Mtrx
    L.initm.argtypes = [C.c_int, C.c_int]
    L.set_poisson_openmp_config.argtypes = [C.c_int]
    L.poisson_SOR_log.restype = Mtrx
    L.poisson_SOR_log.argtypes = [Mtrx, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_void_p]
    f = np.load(f_path, allow_pickle=False)
    n = f.shape[0]
    F = L.initm(n, n)
    for i in range(n):
        np.ctypeslib.as_array(F.M[i], shape=(n,))[:] = f[i]
    L.set_poisson_openmp_config(1)
    h = 1.0 / n
    sys.stdout.write("start %r\n" % time.monotonic())
    sys.stdout.flush()
    L.poisson_SOR_log(F, h, h, sweeps, -1.0, beta, None)
    sys.stdout.write("returned\n")  # not reached: the reference exits at itmax
    sys.stdout.flush()


def time_sweeps(f, sweeps, beta, threads):
    """Seconds per sweep of the reference's poisson_SOR_log on source f (n x n, dx = dy = 1/n),
    measured over `sweeps` sweeps on `threads` OpenMP threads."""
    if not available():
        raise FileNotFoundError(LIB + " is not built (make -C oracle ref)")
    f = np.ascontiguousarray(f, dtype=np.float64)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "f.npy")
        np.save(path, f)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), LIB, path, str(int(sweeps)), repr(beta)],
                           capture_output=True, text=True, env=env)
        t_end = time.monotonic()
    starts = [ln for ln in r.stdout.splitlines() if ln.startswith("start ")]
    if r.returncode != 1 or len(starts) != 1 or "returned" in r.stdout:
        raise RuntimeError("reference SOR sample failed (exit %d): %s %s" % (r.returncode, r.stdout, r.stderr))
    return (t_end - float(starts[0].split()[1])) / sweeps


def sample(f, beta, threads, sample_seconds=10.0):
    """A short calibration run, then about sample_seconds of sweeps; returns (s/sweep, sweeps)."""
    per = time_sweeps(f, 10, beta, threads)
    sweeps = int(max(8, min(40000, sample_seconds / max(per, 1e-9))))
    return time_sweeps(f, sweeps, beta, threads), sweeps


if __name__ == "__main__":
    _child(sys.argv[1], sys.argv[2], int(sys.argv[3]), float(sys.argv[4]))
