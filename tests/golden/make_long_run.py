"""Writes tests/golden/long_run_128.npz: the oracle's whole 128 x 128 cavity run (Re 100, lid
u4 = 1, dt = 5*2^-14, order 6, red-black SOR with omega = 1.8 to an L1 update below 1e-8, 2000
steps from rest: the reference's CPU-runnable workload, configs/cavity128_re100_omega1.8.txt).

Stored, as data without pickles: the per-step Poisson index k, the last L1 update e and the
continuity max/min, and the SHA-256 digest of each final field (w, psi, u, v) as little-endian
binary64 with -0.0 written as +0.0 (the GPU and the reference may differ in the sign of a zero
sum).  The oracle is the repository's CPU restatement of the reference step (oracle/), checked
bit for bit against the compiled reference's multi-step fixtures; the reference's own dense
formulation would need about 4 hours for this run.

Run from the repository root (about 5 minutes on 8 threads):  python tests/golden/make_long_run.py
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

N, RE, DT, ORDER, OMEGA, TOL, ITMAX, STEPS = 128, 100.0, 3.0517578125e-4, 6, 1.8, 1e-8, 1000000, 2000


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a + 0.0, dtype="<f8").tobytes()).digest()


def main():
    p = orc.make_params(N, Re=RE, dt=DT, order=ORDER, poisson_type=2, ordering=orc.REDBLACK, itmax=ITMAX,
                        tol=TOL, beta_=OMEGA, threads=min(16, os.cpu_count() or 1))
    s = orc.State(p)
    t0 = time.time()
    st = s.run(STEPS)
    print("oracle: %d steps in %.1f s, k from %d to %d" % (STEPS, time.time() - t0, st[0].k, st[-1].k))
    np.savez(os.path.join(HERE, "long_run_128.npz"),
             k=np.array([x.k for x in st], dtype=np.int32),
             e=np.array([x.e for x in st]),
             cont_max=np.array([x.cont_max for x in st]),
             cont_min=np.array([x.cont_min for x in st]),
             digests=np.frombuffer(b"".join(digest(getattr(s, f)) for f in ("w", "psi", "u", "v")),
                                   dtype=np.uint8).reshape(4, 32),
             params=np.array([N, RE, DT, ORDER, OMEGA, TOL, ITMAX, STEPS], dtype=np.float64))


if __name__ == "__main__":
    main()
