"""Runs the workloads to their stated horizons, in resumable segments, and compares final states.

  horizon.py fp64    --steps 10000 --segment 3400 --state-in A.npz --state-out B.npz
  horizon.py refined --steps 10000 --segment 5000 --state-in A.npz --state-out B.npz
  horizon.py batch64 --steps 5000 [--precision refined|fp64] --out B.json
  horizon.py compare REFINED.npz FP64.npz

fp64 / refined: the 1024^2 cavity at Re 1000 (dt = 50*2^-20, order 6, SOR omega 1.95 to an L1 update
below 1e-10, lid velocity 1) in plain FP64 (captured-graph steps) or in the FP32 + FP64
defect-correction form (inner solves fold at 1e-5 of the correction's norm).  A segment runs at
most --segment steps from --state-in (or from rest) and writes the state reached (w, u, v; psi
too once the run is complete), with the per-step statistics of every segment so far, to
--state-out.  Each step's Poisson solve starts
from psi = 0 (src/poisson.c:229), so a step depends only on (w, u, v): a run in segments equals
one run, bit for bit.  Progress lines go to stderr every --print-every steps.

batch64: the 64-cavity workload (the committed table, numpy seed 0) in one batch: autotuned launch
shape, --steps steps, each cavity to its own L1 update 1e-8, with a blow-up speed of
--blowup-factor times its lid velocity: a cavity that exceeds it (or whose solve turns non-finite)
is reported diverged and dropped, the others go on.

compare: L-infinity and L2 relative errors of psi, omega and the centerline profiles
u(x = 0.5, y) = u[:, n/2] and v(x, y = 0.5) = v[n/2, :] (the node column / row at the centre,
src/main.c:138-139 puts node j at x = j/n) of the first state against the second.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import cnv  # noqa: E402

N, RE, DT, OMEGA, TOL = 1024, 1000.0, 50 * 2.0 ** -20, 1.95, 1e-10
FIELDS = ("w", "psi", "u", "v")
STAT_KEYS = ("sweeps", "outer", "e", "cont_max", "cont_min", "ms")


def cavity_params(form):
    prec = cnv.PREC_FP64 if form == "fp64" else cnv.PREC_FP32_DC
    return cnv.default_params(N, Re=RE, dt=DT, order=6, beta=OMEGA, tol=TOL, itmax=50_000_000, precision=prec,
                              diagnostics=1, use_graph=int(form == "fp64"))


def run_cavity(args):
    p = cavity_params(args.form)
    hist = {k: [] for k in STAT_KEYS}
    t0_step = 0
    with cnv.Cavity(p) as cav:
        if args.state_in:
            z = np.load(args.state_in)
            if str(z["form"]) != args.form:
                raise SystemExit("state %s is of form %s, not %s" % (args.state_in, z["form"], args.form))
            cav.set_state(w=z["w"], u=z["u"], v=z["v"])
            t0_step = int(z["step"])
            hist = {k: list(z[k]) for k in STAT_KEYS}
        todo = min(args.segment, args.steps - t0_step)
        print("horizon: %s from step %d, %d steps (of %d) on %s" % (args.form, t0_step, todo, args.steps,
                                                                    cnv.device_name(0)), file=sys.stderr, flush=True)
        wall0 = time.perf_counter()
        for t in range(todo):
            st, ms = cav.step_timed(1)
            s = st[0]
            for k, val in zip(STAT_KEYS, (s.sweeps, s.outer, s.e, s.cont_max, s.cont_min, ms)):
                hist[k].append(val)
            step = t0_step + t + 1
            if step % args.print_every == 0 or t == todo - 1:
                done = np.array(hist["ms"])
                print("horizon: %s step %d: %d sweeps, %.2f ms; mean %.2f ms/step over %d steps, %.0f s wall"
                      % (args.form, step, s.sweeps, ms, done.mean(), len(done), time.perf_counter() - wall0),
                      file=sys.stderr, flush=True)
        state = cav.get_state()
    # psi is kept only with the final state (each step recomputes it from zero, so a resumed
    # segment does not need it)
    final = t0_step + todo >= args.steps
    out = {f: state[f] for f in FIELDS if f != "psi" or final}
    out.update({k: np.array(v) for k, v in hist.items()})
    out["step"] = np.array(t0_step + todo)
    out["form"] = np.array(args.form)
    if args.state_out:
        np.savez(args.state_out, **out)
    ms = np.array(hist["ms"])
    summary = {"form": args.form, "steps_done": int(out["step"]), "steps_target": args.steps,
               "gpu_ms_total": float(ms.sum()),
               "Mcell_steps_per_s": N * N * len(ms) / (ms.sum() / 1e3) / 1e6,
               "sweeps_first_last": [int(hist["sweeps"][0]), int(hist["sweeps"][-1])],
               "sweeps_mean": float(np.mean(hist["sweeps"]))}
    print(json.dumps(summary))


def run_batch(args):
    prec = cnv.PREC_FP64 if args.precision == "fp64" else cnv.PREC_FP32_DC
    rows = cnv.batch_table()
    ps = cnv.batch_params(rows, precision=prec)
    for p, r in zip(ps, rows):
        p.blowup_speed = args.blowup_factor * r["U"]
    cells = [r["n"] ** 2 for r in rows]
    with cnv.Batch(ps) as b:
        tune = b.autotune(launches=20)
        print("horizon: batch64 %s autotuned %s" % (args.precision, tune), file=sys.stderr, flush=True)
        diverged_at = [None] * len(rows)
        ms_total, cell_steps = 0.0, 0
        sweeps = np.zeros(len(rows), dtype=np.int64)
        t = 0
        while t < args.steps:
            k = min(args.print_every, args.steps - t)
            st, ms = b.step_timed(k)
            ms_total += ms
            for tt in range(k):
                for i, s in enumerate(st[tt]):
                    if s.status == cnv.OK:
                        cell_steps += cells[i]
                        sweeps[i] += s.sweeps
                    elif s.status == cnv.DIVERGED and diverged_at[i] is None:
                        diverged_at[i] = t + tt
            t += k
            live = b.status().count(cnv.OK)
            print("horizon: batch64 step %d: %d live, %.2f ms/step over the last %d steps"
                  % (t, live, ms / k, k), file=sys.stderr, flush=True)
        status = b.status()
        finals = []
        for i in range(len(rows)):
            s = b.get_state(i)
            finals.append({"max_abs_w": float(np.nanmax(np.abs(s["w"]))) if np.isfinite(s["w"]).any() else None,
                           "max_speed": float(max(np.nanmax(np.abs(s["u"])), np.nanmax(np.abs(s["v"]))))
                           if np.isfinite(s["u"]).any() else None,
                           "finite": bool(np.isfinite(s["w"]).all() and np.isfinite(s["u"]).all())})
    per = [{"n": r["n"], "Re": r["Re"], "U": r["U"], "dt": r["dt"],
            "status": "ok" if status[i] == cnv.OK else "diverged", "diverged_at_step": diverged_at[i],
            "sweeps_total": int(sweeps[i]),
            # the explicit scheme's 1-D advection-diffusion bound U^2 dt <= 2/Re, at the lid speed
            "lid_U2dt_over_2nu": r["U"] ** 2 * r["dt"] * r["Re"] / 2.0, **finals[i]} for i, r in enumerate(rows)]
    out = {"workload": "64 lid-driven cavities (committed table, numpy seed 0), %s, SOR tol 1e-8 each, "
                       "blow-up speed %g x lid velocity" % (args.precision, args.blowup_factor),
           "steps": args.steps, "autotuned": tune, "gpu_ms_total": ms_total,
           "Mcell_steps_per_s": cell_steps / (ms_total / 1e3) / 1e6,
           "live_at_end": status.count(cnv.OK), "cavities": per}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "cavities"}))


def compare(args):
    a, b = np.load(args.a), np.load(args.b)
    mid = a["w"].shape[0] // 2
    pairs = {"psi": (a["psi"], b["psi"]), "omega": (a["w"], b["w"]),
             "u_centerline": (a["u"][:, mid], b["u"][:, mid]), "v_centerline": (a["v"][mid, :], b["v"][mid, :])}
    out = {"a": os.path.basename(args.a), "b": os.path.basename(args.b),
           "steps": [int(a["step"]), int(b["step"])], "linf_rel": {}, "l2_rel": {}}
    for k, (x, y) in pairs.items():
        out["linf_rel"][k] = float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
        out["l2_rel"][k] = float(np.linalg.norm(x - y) / np.linalg.norm(y))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("form", choices=("fp64", "refined", "batch64", "compare"))
    ap.add_argument("files", nargs="*")
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--segment", type=int, default=10000)
    ap.add_argument("--state-in")
    ap.add_argument("--state-out")
    ap.add_argument("--print-every", type=int, default=100)
    ap.add_argument("--precision", choices=("refined", "fp64"), default="refined")
    ap.add_argument("--blowup-factor", type=float, default=10.0)
    ap.add_argument("--out")
    args = ap.parse_args()
    if args.form == "compare":
        args.a, args.b = args.files
        compare(args)
    elif args.form == "batch64":
        run_batch(args)
    else:
        run_cavity(args)


if __name__ == "__main__":
    main()
