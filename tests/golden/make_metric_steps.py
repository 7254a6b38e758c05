"""Writes tests/golden/metric_steps_1024.npz: the oracle's first five steps of the 1024 x 1024
cavity benchmark (Re 1000, lid u4 = 1, dt = 50*2^-20, order 6, red-black SOR with omega = 1.95 to
an L1 update below 1e-10, from rest; configs/cavity1024_re1000_fp64.txt): bench.py's warm-up step
and the four steps it times.

Stored, as data without pickles: per step the Poisson index k, the last L1 update e, the
continuity max/min, and the SHA-256 digests of the fields (w, psi, u, v) after the step, as
little-endian binary64 with -0.0 written as +0.0.  The oracle is the repository's CPU restatement
of the reference step (oracle/), checked bit for bit against the compiled reference; the
reference's own dense formulation cannot run at this size (35 TB of operators).

Run from the repository root (about 6 minutes on 8 threads):  python tests/golden/make_metric_steps.py
"""
import hashlib
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import orc  # noqa: E402

N, RE, DT, ORDER, OMEGA, TOL, ITMAX, STEPS = 1024, 1000.0, 50 * 2.0 ** -20, 6, 1.95, 1e-10, 50000000, 5
FIELDS = ("w", "psi", "u", "v")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a + 0.0, dtype="<f8").tobytes()).digest()


def main():
    p = orc.make_params(N, Re=RE, dt=DT, order=ORDER, poisson_type=2, ordering=orc.REDBLACK, itmax=ITMAX,
                        tol=TOL, beta_=OMEGA, threads=min(16, os.cpu_count() or 1))
    s = orc.State(p)
    k, e, cmax, cmin, dig = [], [], [], [], []
    for t in range(STEPS):
        t0 = time.time()
        st = s.step()
        print("oracle step %d: k %d, %.1f s" % (t, st.k, time.time() - t0), flush=True)
        k.append(st.k)
        e.append(st.e)
        cmax.append(st.cont_max)
        cmin.append(st.cont_min)
        dig.append(b"".join(digest(getattr(s, f)) for f in FIELDS))
    np.savez(os.path.join(HERE, "metric_steps_1024.npz"), k=np.array(k, dtype=np.int32), e=np.array(e),
             cont_max=np.array(cmax), cont_min=np.array(cmin),
             digests=np.frombuffer(b"".join(dig), dtype=np.uint8).reshape(STEPS, 4, 32),
             params=np.array([N, RE, DT, ORDER, OMEGA, TOL, ITMAX, STEPS], dtype=np.float64))


if __name__ == "__main__":
    main()
