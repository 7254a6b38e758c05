"""CPU test of fluid-dynamics1_amd/horizon.py's compare step (the error report of the full-horizon
runs): relative L-infinity and L2 errors of psi, omega and the centreline profiles u[:, n/2],
v[n/2, :], on synthetic states with known differences."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "fluid-dynamics1_amd", "horizon.py")


def test_compare_reports_relative_errors(tmp_path):
    n = 16
    rng = np.random.default_rng(0)
    b = {f: rng.standard_normal((n, n)) for f in ("w", "psi", "u", "v")}
    a = {f: b[f].copy() for f in b}
    a["psi"][3, 4] += 1e-6 * np.max(np.abs(b["psi"]))   # a known L-infinity error of 1e-6
    a["u"][5, n // 2] += 2e-7 * np.max(np.abs(b["u"][:, n // 2]))
    a["v"][n // 2 + 1, 3] += 1.0                          # off the centreline: not reported
    for name, d in (("a.npz", a), ("b.npz", b)):
        np.savez(tmp_path / name, step=np.array(10000), form=np.array("x"), **d)
    r = subprocess.run([sys.executable, SCRIPT, "compare", str(tmp_path / "a.npz"), str(tmp_path / "b.npz")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["steps"] == [10000, 10000]
    assert abs(out["linf_rel"]["psi"] - 1e-6) < 1e-12
    assert abs(out["linf_rel"]["u_centerline"] - 2e-7) < 1e-13
    assert out["linf_rel"]["omega"] == 0.0 and out["linf_rel"]["v_centerline"] == 0.0
    d = a["psi"] - b["psi"]
    assert abs(out["l2_rel"]["psi"] - np.linalg.norm(d) / np.linalg.norm(b["psi"])) < 1e-15
