"""Benchmark: Mcell-steps/s and %HBM roofline of the 1024^2 Re=1000 lid-driven cavity step.

Workload (the metric's configuration): 1024 x 1024 cavity on the unit square, Re = 1000, lid
velocity 1, dt = 0.2*min(h^2 Re/4, h) = 50*2^-20, 6th-order stencils (the reference default),
plain FP64 red-black SOR with omega = 1.95 to an L1 update below 1e-10, psi = omega = 0 at t = 0.
A "step" is one full time step (wall vorticity, transport + Euler, Poisson solve, velocities,
continuity diagnostic).  The same steps are then run in the FP32 + FP64 defect-correction form
and compared with the FP64 run.

Prints one JSON line.  value = cells * steps / wall time of the timed steps (HIP events on the
solver stream), inputs resident in HBM.  roofline: the SOR kernel (the dominant one).  Its
achieved bandwidth is the bytes one launch must move -- read psi and omega, write psi, once per
launch: 24 B per cell in FP64, 12 B in FP32 (the S sweeps of a launch run from registers) --
over its mean launch duration, measured with HIP events on the stream it runs on; frac is that
over the 8 TB/s HBM peak.  Beside it: the HBM bytes the profiler's counters saw per launch
(traffic, read from the committed rocprofv3 --pmc summary named in traffic_source), the
per-sweep rate against the streaming bound of a solver without temporal blocking, and the
issue-side bound from the committed SQ counter summary.
cpu_baseline: the oracle (the repository's CPU restatement of the reference step), OpenMP
red-black, on a bounded sample of the same workload (see "sample"); serial_lexicographic: the
same on one thread in the reference's default (serial, lexicographic) order; reference_build
(when oracle/_ref is built): the reference's own OpenMP poisson_SOR_log, compiled from its
sources, timed on a bounded number of sweeps of the same source.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-dynamics1_amd"))
sys.path.insert(0, ROOT)

N = 1024
RE = 1000.0
DT = 50 * 2.0 ** -20          # 0.2 * min(h^2 Re / 4, h) with h = 1/1024
OMEGA = 1.95
TOL = 1e-10
HBM_PEAK = 8.0e12             # MI355X HBM3E, bytes/s
BYTES_PER_CELL_SWEEP = 24     # read psi, read omega, write psi (FP64)
BYTES_PER_CELL_SWEEP_FP32 = 12
C_CELLS = N * N
CLOCK_HZ = 2.4e9              # MI355X peak engine clock
SIMDS_PER_CU = 4


def workload_params(cnv, precision, args, graph=True, **kw):
    # steps replay captured HIP graphs unless --eager (or graph=False: the refined form with the
    # reference's inner relaxation runs eager launches, 4% faster there, since its short inner
    # solves leave each outer iteration's residual, chunks and correction graph launches with
    # their own submission cost; at omega 1.95 the two are within 1% of each other)
    return cnv.default_params(N, Re=RE, dt=DT, order=6, beta=OMEGA, tol=TOL, itmax=50_000_000,
                              precision=precision, sweeps_per_launch=args.spl, tile_waves=args.tile_waves,
                              wave_rows=args.wave_rows, diagnostics=1,
                              use_graph=0 if args.eager or not graph else 1, **kw)


def timed_steps(cav, k, tag):
    """k steps, each bracketed by HIP events on the solver stream; returns (stats, total ms)."""
    stats, total = [], 0.0
    for t in range(k):
        st, ms = cav.step_timed(1)
        stats += st
        total += ms
        print("bench: %s step %d: %.2f ms, %d sweeps" % (tag, t, ms, st[0].sweeps), file=sys.stderr, flush=True)
    return stats, total


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if one exists."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def read_sq(name):
    """Device totals per dispatch (SQ_WAVES, SQ_INSTS_VALU, SQ_WAVE_CYCLES, SQ_WAIT_ANY) from the
    committed rocprofv3 SQ counter summary, if one exists."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        f = line.split()
        if len(f) >= 2 and f[0] in ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY"):
            try:
                out.setdefault(f[0], float(f[1]))
            except ValueError:
                pass
    return out if len(out) == 4 else None


def roofline(cells, spl, launch_ms, bytes_per_cell, traffic_name, sq_name, waves_per_wg, kernel):
    """The SOR kernel against HBM: the bytes one launch must move (read psi and the source, write
    psi, once per launch) over the HIP-event launch time; the counter-measured bytes beside it;
    the per-sweep rate against a streaming solver's bound (bytes_per_cell per sweep at peak); and
    the issue bound of the VALU instructions per SIMD from the SQ counters (every VALU
    instruction issues in one quad-cycle, waves_per_wg / 4 waves per SIMD, one workgroup per CU)."""
    t = launch_ms / 1e3
    need = bytes_per_cell * cells
    traffic = read_traffic(traffic_name)
    out = {"bound": "hbm", "achieved": need / t / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": need / t / HBM_PEAK, "traffic": traffic,
           "kernel": kernel, "launch_ms": launch_ms, "bytes_per_launch": need,
           "bytes_model": "read psi and the source, write psi, once per %d-sweep launch: %d B per cell "
                          "x %d cells (the sweeps in between run from registers)" % (spl, bytes_per_cell, cells),
           "traffic_source": "profiles/%s (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each scaled by the same "
                             "counter on a 256 MiB copy kernel; median working dispatch)" % traffic_name,
           "counter_GBps": traffic / t / 1e9 if traffic else None,
           "counter_frac": traffic / t / HBM_PEAK if traffic else None,
           "sweep_rate": {"cell_sweeps_per_s": spl * cells / t,
                          "streaming_bound_cell_sweeps_per_s": HBM_PEAK / bytes_per_cell,
                          "ratio": spl * cells / t / (HBM_PEAK / bytes_per_cell),
                          "note": "a solver without temporal blocking moves %d B per cell-sweep; this "
                                  "kernel runs %d sweeps per tile load" % (bytes_per_cell, spl)}}
    sq = read_sq(sq_name)
    if sq:
        per_wave = sq["SQ_INSTS_VALU"] / sq["SQ_WAVES"]
        issue_s = per_wave * (waves_per_wg / SIMDS_PER_CU) * 4 / CLOCK_HZ
        out["issue"] = {"valu_per_wave": per_wave, "issue_bound_ms": issue_s * 1e3, "frac": issue_s / t,
                        "wave_wait_frac": sq["SQ_WAIT_ANY"] / sq["SQ_WAVE_CYCLES"],
                        "source": "profiles/%s (rocprofv3 --pmc SQ_*), %d waves per workgroup, one "
                                  "workgroup per CU, one quad-cycle per VALU instruction at %.1f GHz"
                                  % (sq_name, waves_per_wg, CLOCK_HZ / 1e9)}
    return out


def cpu_whole_step():
    """The oracle's first time step of the workload from rest, whole and timed (wall vorticity,
    transport + Euler, the Poisson solve to tol, velocities, continuity), OpenMP red-black on up
    to 16 host threads: the same arithmetic as the GPU's step, so the same 46352 sweeps."""
    from oracle import orc
    threads = min(16, os.cpu_count() or 1)
    p = orc.make_params(N, Re=RE, dt=DT, order=6, tol=TOL, itmax=50_000_000, beta_=OMEGA,
                        ordering=orc.REDBLACK, threads=threads)
    s = orc.State(p)
    t0 = time.perf_counter()
    st = s.step()
    return dict(seconds=time.perf_counter() - t0, sweeps=st.sweeps, threads=threads)


def cpu_baseline(sample_seconds=25.0, serial=False):
    """Oracle at the same workload on the host cores: the first time step's transport plus a
    bounded number of SOR sweeps, timed; scaled to cell-steps/s with the GPU run's mean sweeps per
    step.  Red-black on up to 16 threads (the reference's OpenMP build: identical arithmetic, so
    identical sweep counts), or serial=True: one thread in lexicographic order (the reference's
    default build; both orderings are consistent orderings of the 5-point operator, so SOR
    converges at the same rate and the sweep counts nearly agree)."""
    from oracle import orc
    threads = 1 if serial else min(16, os.cpu_count() or 1)
    p = orc.make_params(N, Re=RE, dt=DT, order=6, tol=-1.0, itmax=1, beta_=OMEGA,
                        ordering=orc.LEX if serial else orc.REDBLACK, threads=threads)
    s = orc.State(p)
    orc.lib().orc_apply_velocity_bc(orc.C.byref(p), s.u, s.v)
    orc.lib().orc_wall_vorticity(orc.C.byref(p), s.u, s.v, s.w)
    t0 = time.perf_counter()
    orc.lib().orc_transport(orc.C.byref(p), s.w, s.u, s.v)
    t_transport = time.perf_counter() - t0
    f = -s.w
    # calibrate, then run a sample of about sample_seconds
    p.itmax = 4
    t0 = time.perf_counter()
    orc.poisson(p, f)
    per = (time.perf_counter() - t0) / 4
    sweeps = int(max(8, min(40000, sample_seconds / max(per, 1e-9))))
    p.itmax = sweeps
    t0 = time.perf_counter()
    orc.poisson(p, f)
    t_sweep = (time.perf_counter() - t0) / sweeps
    return dict(t_transport=t_transport, t_sweep=t_sweep, sweeps=sweeps, threads=threads, f=f)


def host_buffer_leg(cnv, cav, state, params):
    """One Poisson solve of the workload's source (f = -omega after the timed steps) through the
    reference-compatible boundary, which hands over host arrays (cnv_poisson_host: f copied in, psi
    copied out, pageable memory), beside the same solve on device buffers (cnv_poisson_device; the
    buffers are a second context's w (holding f) and u (receiving psi)).  Both wall times include
    the host's launch and poll work; their difference is the PCIe part."""
    import ctypes
    f = np.ascontiguousarray(-state["w"])
    t0 = time.perf_counter()
    rc, psi, k, e = cav.poisson(f)
    t_host = time.perf_counter() - t0
    if rc != cnv.OK:
        raise RuntimeError("host-buffer Poisson solve failed (%d)" % rc)
    with cnv.Cavity(params) as aux:
        aux.set_state(w=f)
        dw, du = ctypes.c_void_p(), ctypes.c_void_p()
        cnv._check(cnv.lib().cnv_device_state(aux.h, ctypes.byref(dw), None, ctypes.byref(du), None), aux.h)
        kd, ed = ctypes.c_int(), ctypes.c_double()
        t0 = time.perf_counter()
        rc = cnv.lib().cnv_poisson_device(cav.h, dw, du, ctypes.byref(kd), ctypes.byref(ed))
        t_dev = time.perf_counter() - t0
        if rc != cnv.OK:
            raise RuntimeError("device-buffer Poisson solve failed (%d)" % rc)
        same = bool(np.array_equal(aux.get_state()["u"], psi)) and (k, e) == (kd.value, ed.value)
    print("bench: Poisson solve host buffers %.2f ms, device buffers %.2f ms, %d sweeps, identical %s"
          % (t_host * 1e3, t_dev * 1e3, k + 1, same), file=sys.stderr, flush=True)
    return {"solve_ms_host_buffers": t_host * 1e3, "solve_ms_device_buffers": t_dev * 1e3,
            "pcie_share": (t_host - t_dev) / t_host, "sweeps": k + 1, "identical_results": same,
            "Mcells_per_s_host_buffers": N * N / t_host / 1e6,
            "note": "one SOR solve of f = -omega (1024^2, omega 1.95, tol 1e-10) through cnv_poisson_host "
                    "(8 MiB in, 8 MiB out, pageable) and through cnv_poisson_device; not the metric"}


def batched_leg(cnv, steps, precision, tag):
    """The 64-cavity workload (seed 0) in one batch: autotuned launch shape, one warmup step, then
    `steps` timed steps; beside it the same cavities stepped one at a time through the one-cavity
    API (same warmup and steps), to show what batching buys."""
    rows = cnv.batch_table()
    cells = sum(r["n"] ** 2 for r in rows)
    ps = cnv.batch_params(rows, precision=precision)
    with cnv.Batch(ps) as b:
        tune = b.autotune(launches=20)
        print("bench: batch %s autotuned %s" % (tag, tune), file=sys.stderr, flush=True)
        b.step(1)
        st, ms = b.step_timed(steps)
    sweeps = [sum(s.sweeps for s in row) for row in st]
    print("bench: batch %s %.2f ms/step" % (tag, ms / steps), file=sys.stderr, flush=True)
    alone_ms = 0.0
    for p in ps:  # the same shape the autotuner chose, so the work per cavity is identical
        p.sweeps_per_launch, p.tile_waves, p.wave_rows = tune["sweeps_per_launch"], tune["tile_waves"], tune["wave_rows"]
        with cnv.Cavity(p) as cav:
            cav.step(1)
            _, t = cav.step_timed(steps)
            alone_ms += t
    print("bench: batch %s one at a time %.2f ms/step" % (tag, alone_ms / steps), file=sys.stderr, flush=True)
    return {
        "workload": "64 lid-driven cavities, seed 0: N in {64..512}, Re log-uniform [100, 5000], lid U in "
                    "[0.5, 1.5], dt = 0.2*min(h^2 Re/4, h), reference beta per N, SOR tol 1e-8, order 6",
        "cells": cells,
        "value": cells * steps / (ms / 1e3) / 1e6,
        "unit": "Mcell-steps/s",
        "ms_per_step": ms / steps,
        "steps": steps,
        "autotuned": tune,
        "sweeps_per_step_all_cavities": sweeps,
        "one_at_a_time_ms_per_step": alone_ms / steps,
        "speedup_vs_one_at_a_time": alone_ms / ms,
    }


def adjoint_leg(cnv, steps, K):
    """The adjoint workload: 256^2 cavity, Re 400, lid U = 1, dt = 20*2^-16, initial vorticity a sum
    of 32 Gaussians (seed 1); gradient of the final kinetic energy w.r.t. the initial vorticity and
    the lid velocity, checkpoints every K steps; checked by central differences (eps = 1e-6) along
    one random direction and in U.  Solves start from 0; then once more with warm-started solves,
    compared with the first gradient."""
    n, Re, U, dt = 256, 400.0, 1.0, 20 * 2.0 ** -16
    p = cnv.default_params(n, Re=Re, dt=dt, beta=cnv.reference_beta(n), itmax=50_000_000, diagnostics=0)
    w0 = cnv.gaussian_vorticity(n, count=32, seed=1)
    with cnv.Adjoint(p, checkpoint_every=K, rtol_forward=1e-14, rtol_adjoint=1e-14) as adj:
        t0 = time.perf_counter()
        E = adj.forward(w0, U, steps)
        t_fwd = time.perf_counter() - t0
        fs, _ = adj.sweeps()
        t0 = time.perf_counter()
        g, dU = adj.backward()
        t_bwd = time.perf_counter() - t0
        _, bs = adj.sweeps()
        print("bench: adjoint forward %.2f s, backward %.2f s" % (t_fwd, t_bwd), file=sys.stderr, flush=True)
        eps = 1e-6
        d = np.random.default_rng(11).standard_normal((n, n))
        fd = (adj.forward(w0 + eps * d, U, steps) - adj.forward(w0 - eps * d, U, steps)) / (2 * eps)
        fdU = (adj.forward(w0, U + eps, steps) - adj.forward(w0, U - eps, steps)) / (2 * eps)
    # the same with warm-started solves (forward from the previous step's psi, adjoint from the
    # previous reverse step's solution)
    p.warm_start = 1
    with cnv.Adjoint(p, checkpoint_every=K, rtol_forward=1e-14, rtol_adjoint=1e-14) as adj:
        t0 = time.perf_counter()
        Ew = adj.forward(w0, U, steps)
        tw_fwd = time.perf_counter() - t0
        fws, _ = adj.sweeps()
        t0 = time.perf_counter()
        gw, dUw = adj.backward()
        tw_bwd = time.perf_counter() - t0
        _, bws = adj.sweeps()
    print("bench: adjoint warm start forward %.2f s, backward %.2f s" % (tw_fwd, tw_bwd), file=sys.stderr,
          flush=True)
    walls = np.concatenate([g[0, :], g[-1, :], g[:, 0], g[:, -1]])
    return {
        "workload": "cavity 256x256 Re400 FP64, 2000 steps, w0 = 32 Gaussians (seed 1), E = final kinetic "
                    "energy, checkpoints every %d steps, Poisson solves to relative L1 update 1e-14" % K,
        "steps": steps, "energy": E, "dE_dU": dU,
        "forward_s": t_fwd, "backward_s": t_bwd,
        "gradient_over_forward": t_bwd / t_fwd,
        "sweeps_forward": fs, "sweeps_backward_adjoint_solves": bs,
        "fd_rel_err_direction": abs(float(np.sum(g * d)) - fd) / abs(fd),
        "fd_rel_err_lid": abs(dU - fdU) / abs(fdU),
        "wall_gradient_max_abs": float(np.max(np.abs(walls))),
        "warm_start": {
            "forward_s": tw_fwd, "backward_s": tw_bwd,
            "sweeps_forward": fws, "sweeps_backward_adjoint_solves": bws,
            "energy_rel_diff": abs(Ew - E) / abs(E),
            "gradient_max_rel_diff": float(np.max(np.abs(gw - g)) / np.max(np.abs(g))),
            "dE_dU_rel_diff": abs(dUw - dU) / abs(dU),
        },
    }


def dc_leg(cnv, args, state64, ms64, spl, inner_beta, tag, time_sor=True, transport_fp32=1):
    """The FP32 + FP64 defect-correction form of the same steps, compared with the FP64 run."""
    p = workload_params(cnv, cnv.PREC_FP32_DC, args, graph=inner_beta <= 0 and not args.refined_eager)
    p.inner_beta = inner_beta
    p.transport_fp32 = transport_fp32
    cd = cnv.Cavity(p)
    if args.warmup:
        timed_steps(cd, args.warmup, tag + " warmup")
    st32, ms32 = timed_steps(cd, args.steps, tag)
    s32 = cd.get_state()
    sor32_ms = cd.time_sor(args.sor_launches) if time_sor else None
    cd.close()
    mid = N // 2

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def rel2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    out = {
        "inner_beta": inner_beta if inner_beta > 0 else OMEGA,
        "step": "captured HIP graphs" if p.use_graph else "eager launches",
        "stencils": "fp32" if transport_fp32 else "fp64",  # transport and velocity stencils
        "value": C_CELLS * args.steps / (ms32 / 1e3) / 1e6,
        "ms_per_step": ms32 / args.steps,
        "speedup_vs_fp64": ms64 / ms32,
        "inner_sweeps_per_step": [s.sweeps for s in st32],
        "outer_per_step": [s.outer for s in st32],
        "err_linf_rel": {"psi": rel(s32["psi"], state64["psi"]), "omega": rel(s32["w"], state64["w"]),
                         "u_centerline": rel(s32["u"][:, mid], state64["u"][:, mid]),
                         "v_centerline": rel(s32["v"][mid, :], state64["v"][mid, :])},
        "err_l2_rel": {"psi": rel2(s32["psi"], state64["psi"]), "omega": rel2(s32["w"], state64["w"]),
                       "u_centerline": rel2(s32["u"][:, mid], state64["u"][:, mid]),
                       "v_centerline": rel2(s32["v"][mid, :], state64["v"][mid, :])},
    }
    if sor32_ms is not None:
        out["sor_fp32_launch_ms"] = sor32_ms
        out["roofline"] = roofline(C_CELLS, spl, sor32_ms, BYTES_PER_CELL_SWEEP_FP32,
                                   "sor_fp32_refined_n1024_traffic.txt", "sor_fp32_refined_n1024_sq_counters.txt",
                                   16, "sor_rb_kernel<float,...,S=%d>" % spl)
    print("bench: %s %.3f ms/step, inner sweeps %s" % (tag, ms32 / args.steps, out["inner_sweeps_per_step"]),
          file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spl", type=int, default=0, help="SOR sweeps per launch (0 = library default)")
    ap.add_argument("--tile-waves", type=int, default=0)
    ap.add_argument("--wave-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--eager", action="store_true", help="FP64 steps as eager launches instead of graph replays")
    ap.add_argument("--no-dc", action="store_true", help="skip the FP32-refined leg")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start FP64 leg")
    ap.add_argument("--sor-launches", type=int, default=400)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-cavity batched leg")
    ap.add_argument("--no-host-io", action="store_true",
                    help="skip the Poisson solve through host buffers (the PCIe-inclusive rate)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 256^2 adjoint leg")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--profile-sor", type=int, default=0, metavar="LAUNCHES",
                    help="profiling mode: one step, then LAUNCHES timed SOR launches and one 256 MiB "
                         "calibration copy (for counter scaling); prints no JSON")
    ap.add_argument("--profile-form", choices=("fp64", "refined"), default="fp64",
                    help="profiling mode: the FP64 step's SOR kernel, or the refined step's FP32 one")
    ap.add_argument("--refined-eager", action="store_true",
                    help="run the refined legs with eager launches (for rocprofv3 --kernel-trace, whose "
                         "interception of graph launches has faulted when a second context replayed its "
                         "graphs in one process)")
    ap.add_argument("--profile-adjoint", type=int, default=0, metavar="STEPS",
                    help="profiling mode: the adjoint workload's forward run (256^2, solves to the "
                         "relative rule) for STEPS steps; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the single cavity would shard "
              "by row strips with a halo exchange every few sweeps, the batch by whole cavities). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import cnv
    C = N * N
    dev = cnv.device_name(0)
    print("bench: %s, %dx%d Re=%g, %d warmup + %d timed steps" % (dev, N, N, RE, args.warmup, args.steps),
          file=sys.stderr, flush=True)

    if args.profile_adjoint:
        n = 256
        p = cnv.default_params(n, Re=400.0, dt=20 * 2.0 ** -16, beta=cnv.reference_beta(n), itmax=50_000_000,
                               diagnostics=0)
        with cnv.Adjoint(p, checkpoint_every=100, rtol_forward=1e-14, rtol_adjoint=1e-14) as adj:
            t0 = time.perf_counter()
            adj.forward(cnv.gaussian_vorticity(n, count=32, seed=1), 1.0, args.profile_adjoint)
            t = time.perf_counter() - t0
            fs, _ = adj.sweeps()
        print("bench: profile mode: adjoint forward %d steps %.3f s, %d sweeps" % (args.profile_adjoint, t, fs),
              file=sys.stderr, flush=True)
        return

    if args.profile_sor:
        refined = args.profile_form == "refined"
        cav = cnv.Cavity(workload_params(cnv, cnv.PREC_FP32_DC if refined else cnv.PREC_FP64, args,
                                         graph=not (refined and args.refined_eager)))
        timed_steps(cav, 1, "dc" if refined else "fp64")
        ms = cav.time_sor(args.profile_sor)
        cms = cav.calibration_copy(256 << 20)
        print("bench: profile mode: SOR launch %.4f ms, calibration copy of 256 MiB %.4f ms" % (ms, cms),
              file=sys.stderr, flush=True)
        cav.close()
        return

    # ---- FP64 plain ---------------------------------------------------------------------
    cav = cnv.Cavity(workload_params(cnv, cnv.PREC_FP64, args))
    if args.warmup:
        timed_steps(cav, args.warmup, "fp64 warmup")
    st, ms = timed_steps(cav, args.steps, "fp64")
    sweeps = [s.sweeps for s in st]
    T = ms / 1e3
    value = C * args.steps / T / 1e6
    alg_bytes = sum(24 + BYTES_PER_CELL_SWEEP * k for k in sweeps) * C
    state64 = cav.get_state()
    spl = cav.sweeps_per_launch
    print("bench: fp64 %.3f ms/step, sweeps/step %s" % (ms / args.steps, sweeps), file=sys.stderr, flush=True)
    sor_ms = cav.time_sor(args.sor_launches)
    host_io = None if args.no_host_io else host_buffer_leg(
        cnv, cav, state64, workload_params(cnv, cnv.PREC_FP64, args))
    cav.close()

    # ---- the same steps with the opt-in warm start (each solve from the previous psi) ---------
    warm = None
    if not args.no_warm:
        with cnv.Cavity(workload_params(cnv, cnv.PREC_FP64, args, warm_start=1)) as cw:
            if args.warmup:
                timed_steps(cw, args.warmup, "fp64 warm-start warmup")
            stw, msw = timed_steps(cw, args.steps, "fp64 warm-start")
            sw = cw.get_state()
        warm = {"ms_per_step": msw / args.steps, "value": C * args.steps / (msw / 1e3) / 1e6,
                "speedup_vs_cold": ms / msw, "sweeps_per_step": [s.sweeps for s in stw],
                "err_linf_rel_vs_cold": {f: float(np.max(np.abs(sw[f] - state64[f])) / np.max(np.abs(state64[f])))
                                         for f in ("psi", "w", "u", "v")},
                "note": "opt-in: each Poisson solve starts from the previous step's psi (the reference "
                        "starts from 0, src/poisson.c:229); not the metric's configuration"}
        print("bench: fp64 warm start %.3f ms/step, sweeps/step %s" % (warm["ms_per_step"], warm["sweeps_per_step"]),
              file=sys.stderr, flush=True)

    # ---- FP32 + FP64 defect correction ------------------------------------------------------
    dc = None
    if not args.no_dc:
        dc = dc_leg(cnv, args, state64, ms, spl, 0.0, "dc")
        # the same form with the inner FP32 solves relaxed by the reference's own rule
        # (src/main.c:134) instead of omega = 1.95: an algorithmic choice open to the inner solver
        dc["reference_inner_beta"] = dc_leg(cnv, args, state64, ms, spl, cnv.reference_beta(N), "dc-refbeta",
                                            time_sor=False)
        # the same form with the transport and velocity stencils in FP64 (transport_fp32 = 0): no
        # measurable difference in speed (the stencils are about 40 us of a step); the error
        # against FP64 drops from the FP32 stencils' rounding to the solver tolerance's own effect
        dc["stencils_fp64"] = dc_leg(cnv, args, state64, ms, spl, 0.0, "dc-stencils64", time_sor=False,
                                     transport_fp32=0)

    # ---- batched 64-cavity workload -----------------------------------------------------------
    batched = None
    if not args.no_batch:
        batched = batched_leg(cnv, args.batch_steps, cnv.PREC_FP32_DC, "fp32-refined")
        batched["fp64"] = batched_leg(cnv, args.batch_steps, cnv.PREC_FP64, "fp64")

    # ---- adjoint workload ---------------------------------------------------------------------
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(cnv, args.adjoint_steps, 100)

    # ---- CPU baseline -----------------------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        whole = cpu_whole_step()
        print("bench: cpu whole step %.1f s, %d sweeps" % (whole["seconds"], whole["sweeps"]),
              file=sys.stderr, flush=True)
        cpu = {"value": C / whole["seconds"] / 1e6, "unit": "Mcell-steps/s", "cores": whole["threads"],
               "kind": "port", "method": "timed",
               "sample": "oracle (C restatement of the reference step, OpenMP red-black, %d threads): the "
                         "workload's first time step from rest at 1024^2, whole and timed (%.1f s, %d SOR "
                         "sweeps; the GPU's timed steps take %d-%d sweeps)"
                         % (whole["threads"], whole["seconds"], whole["sweeps"], min(sweeps), max(sweeps))}
        b = cpu_baseline(sample_seconds=8.0)
        kmean = float(np.mean(sweeps))
        t_step = b["t_transport"] + kmean * b["t_sweep"]
        cpu["modelled_timed_window"] = {
            "value": C / t_step / 1e6, "unit": "Mcell-steps/s", "cores": b["threads"], "method": "model",
            "sample": "model, not a timed step: oracle (OpenMP red-black, %d threads) step-0 transport (%.3f s) "
                      "+ %d timed SOR sweeps (%.2f ms/sweep), scaled by the GPU's mean %.0f sweeps per timed step"
                      % (b["threads"], b["t_transport"], b["sweeps"], b["t_sweep"] * 1e3, kmean)}
        bs = cpu_baseline(sample_seconds=8.0, serial=True)
        t_step_s = bs["t_transport"] + kmean * bs["t_sweep"]
        cpu["serial_lexicographic"] = {
            "value": C / t_step_s / 1e6, "unit": "Mcell-steps/s", "cores": 1, "method": "model",
            "sample": "model, not a timed step: oracle, one thread, lexicographic SOR (the reference's default "
                      "build) at 1024^2: step-0 transport (%.3f s) + %d timed SOR sweeps (%.2f ms/sweep), scaled "
                      "by the GPU's mean %.0f sweeps per timed step"
                      % (bs["t_transport"], bs["sweeps"], bs["t_sweep"] * 1e3, kmean)}
        from oracle import ref_sor_sample
        if ref_sor_sample.available():
            # the reference's own poisson_SOR_log, compiled from its sources (oracle/_ref); its
            # transport is not timed: at 1024^2 its dense Kronecker operators would need 8 TB
            t_ref, n_ref = ref_sor_sample.sample(b["f"], OMEGA, b["threads"], sample_seconds=10.0)
            t_step_r = b["t_transport"] + kmean * t_ref
            cpu["reference_build"] = {
                "value": C / t_step_r / 1e6, "unit": "Mcell-steps/s", "cores": b["threads"], "kind": "reference",
                "method": "model",
                "sample": "model, not a timed step: the reference's own poisson_SOR_log (src/poisson.c, built by oracle/Makefile with its "
                          "Makefile's flags -g -fopenmp and -DOPENMP_ENABLED: red-black, %d threads) at 1024^2: "
                          "%d sweeps (%.2f ms/sweep) on the step-0 source, plus the oracle's step-0 transport "
                          "(%.3f s; the reference's dense Kronecker transport cannot be formed at 1024^2), "
                          "scaled by the GPU run's mean %.0f sweeps/step"
                          % (b["threads"], n_ref, t_ref * 1e3, b["t_transport"], kmean)}
        else:
            print("bench: oracle/_ref not built; reference_build baseline skipped", file=sys.stderr, flush=True)

    line = {
        "metric": "Mcell-steps/s & %HBM peak, 1024^2 Re1000 cavity; FP32-refined vs FP64 plain",
        "value": value,
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deterministic lid-driven cavity from rest)",
        "config": {"workload": "cavity 1024x1024 Re1000 FP64 red-black SOR omega=1.95 tol=1e-10, order 6",
                   "n": N, "Re": RE, "dt": DT, "sor_omega": OMEGA, "tol": TOL, "order": 6,
                   "step": "eager launches" if args.eager else "captured HIP graphs",
                   "sweeps_per_launch": spl, "sweeps_per_step": sweeps,
                   "algorithmic_GBps_step_loop": alg_bytes / T / 1e9},
        "roofline": roofline(C, spl, sor_ms, BYTES_PER_CELL_SWEEP, "sor_fp64_n1024_traffic.txt",
                             "sor_fp64_n1024_sq_counters.txt", 16, "sor_rb_kernel<double,...,S=%d>" % spl),
        "cpu_baseline": cpu,
        "fp64_warm_start": warm,
        "fp32_refined": dc,
        "host_buffers": host_io,
        "batched": batched,
        "adjoint": adjoint,
        "device": dev,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
