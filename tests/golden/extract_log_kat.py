"""Extracts the known answers held in the reference's run logs into tests/golden/log_kat.npz.

The logs (output/logs/testRun*.txt of the reference) record, per time step, the Poisson
iteration index k, the L1 update e (src/poisson.c:275) and the continuity max/min
(src/main.c:407-408), each printed with %E.  The configs they ran are the reference's
config_default.txt (64^2, Re 1000) and config_high_re.txt (128^2, Re 5000); their parameters
are recorded next to the numbers.  Run in a container that has /root/reference.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LOGS = "/root/reference/output/logs"
RUNS = {
    # name: (n, Re, dt, tol, itmax, ordering 0=lexicographic 1=red-black)
    "testRunNOOMP": (64, 1000.0, 0.005, 1e-3, 10000, 0),
    "testRunOMP": (64, 1000.0, 0.005, 1e-3, 10000, 1),
    "testRunNOOMPHIGHRES": (128, 5000.0, 0.001, 5e-4, 15000, 0),
    "testRunOMPHIGHRES": (128, 5000.0, 0.001, 5e-4, 15000, 1),
}
SOLVED = re.compile(r"solved with (\d+) iterations - root-sum-of-squares error: (\S+)")
CONT = re.compile(r"Continuity max: (\S+) \| Continuity min: (\S+) \|")


def main():
    out = {}
    for name, meta in RUNS.items():
        txt = open(os.path.join(LOGS, name + ".txt")).read()
        solved = SOLVED.findall(txt)
        cont = CONT.findall(txt)
        m = min(len(solved), len(cont))  # the logs were interrupted mid-step
        out[name + "_k"] = np.array([int(k) for k, _ in solved[:m]])
        out[name + "_e"] = np.array([e for _, e in solved[:m]])
        out[name + "_cmax"] = np.array([a for a, _ in cont[:m]])
        out[name + "_cmin"] = np.array([b for _, b in cont[:m]])
        out[name + "_meta"] = np.array(meta, dtype=np.float64)
        print(name, m, "steps")
    np.savez_compressed(os.path.join(HERE, "log_kat.npz"), **out)


if __name__ == "__main__":
    main()
