"""The reference's CPU-runnable workload in full on the GPU: all 2000 steps of the 128 x 128
cavity (Re 100, lid u4 = 1, dt = 5*2^-14, order 6, SOR omega = 1.8 to an L1 update below 1e-8;
configs/cavity128_re100_omega1.8.txt) equal the oracle's red-black run step for step (Poisson
index k, continuity max/min; the L1 update e, summed in another order, to 1e-12) and in the final
fields bit for bit, through SHA-256 digests (tests/golden/long_run_128.npz, written by
tests/golden/make_long_run.py)."""
import hashlib
import os

import numpy as np
import pytest

import cnv

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def digest(a):
    # -0.0 as +0.0: the sign of a zero sum may differ
    return hashlib.sha256(np.ascontiguousarray(a + 0.0, dtype="<f8").tobytes()).digest()


@pytest.mark.parametrize("graph", [0, 1], ids=["eager", "graph"])
def test_whole_128_run_bit_exact_vs_oracle(graph):
    g = np.load(os.path.join(GOLD, "long_run_128.npz"))
    n, Re, dt, order, omega, tol, itmax, steps = g["params"]
    print("long run: %d steps of %dx%d, graph %d" % (steps, n, n, graph), flush=True)
    p = cnv.default_params(int(n), Re=float(Re), dt=float(dt), order=int(order), beta=float(omega),
                           tol=float(tol), itmax=int(itmax), use_graph=graph)
    with cnv.Cavity(p) as cav:
        st = cav.step(int(steps))
        s = cav.get_state()
    k = np.array([x.k for x in st], dtype=np.int64)
    first_diff = int(np.argmax(k != g["k"])) if not np.array_equal(k, g["k"]) else -1
    assert first_diff < 0, "k differs first at step %d: %d vs %d" % (first_diff, k[first_diff], g["k"][first_diff])
    e = np.array([x.e for x in st])
    assert float(np.max(np.abs(e - g["e"]) / np.abs(g["e"]))) <= 1e-12
    assert np.array_equal(np.array([x.cont_max for x in st]), g["cont_max"])
    assert np.array_equal(np.array([x.cont_min for x in st]), g["cont_min"])
    for i, f in enumerate(("w", "psi", "u", "v")):
        assert digest(s[f]) == bytes(g["digests"][i]), f
