"""CPU tests of the drop-in boundary: the C-ABI library exports everything include/*.h declares,
and its host-side functions (no GPU involved) match the compiled reference's outputs."""
import os
import subprocess

import numpy as np
import pytest

import legacy_bind as lb

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Declared by the reference's header and never defined by it (include/fluiddyn.h:11); kept
# declared-only so that a caller's link behaves the same.
DECLARED_ONLY = {"pressure"}


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_every_declared_symbol_is_exported():
    decl = lb.declared_functions()
    assert len(decl) > 50, decl
    out = subprocess.run(["nm", "-D", "--defined-only", lb.LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = sorted(n for n in decl if n not in exported and n not in DECLARED_ONLY)
    assert not missing, missing
    assert "pressure" not in exported
    # and nothing else: the kernels' C++ symbols and the host helpers stay local
    every = {line.split()[-1] for line in out.splitlines() if len(line.split()) == 3}
    extra = sorted(every - set(decl))
    assert not extra, extra[:10]


def test_reference_headers_signatures_present():
    decl = lb.declared_functions()
    for name in ("Diff1", "Diff2", "error", "poisson", "poisson_SOR", "poisson_log", "poisson_SOR_log",
                 "set_poisson_openmp_config", "euler", "continuity", "vorticity",
                 "set_fluiddyn_openmp_config", "initm", "freem", "mtrxmul", "kronecker", "reshape",
                 "load_default_config", "load_config_from_file", "print_config", "printvtk",
                 "cnv_create", "cnv_step", "cnv_poisson_device"):
        assert name in decl, name


def test_diff_tables_bits():
    g = load("diff_tables.npz")
    for key in g.files:
        which, n, o, h = key.split("_")
        n, o, h = int(n[1:]), int(o[1:]), float.fromhex(h[1:])
        if n < o + 2:
            continue
        D = (lb.lib().Diff1 if which == "d1" else lb.lib().Diff2)(n, o, h)
        assert np.array_equal(lb.from_m(D), g[key]), key
        lb.free(D)


def test_diff_rows_sum_to_zero():
    # the FP32 transport stencils take each tap as a difference from the point's own value, which
    # equals the plain stencil only because every row of Diff1 and Diff2 sums to zero (the
    # derivative of a constant), boundary rows included; rounding leaves a few ulps of the row's
    # absolute sum
    for o in (2, 4, 6):
        for n in (12, 33, 64):
            for fn in (lb.lib().Diff1, lb.lib().Diff2):
                D = fn(n, o, 1.0 / (n - 1))
                A = lb.from_m(D)
                lb.free(D)
                tol = 16 * np.finfo(float).eps * np.abs(A).sum(axis=1)
                assert np.all(np.abs(A.sum(axis=1)) <= tol), (o, n, np.abs(A.sum(axis=1)).max())


def test_linearalg_matches_reference():
    g = load("host_linearalg.npz")
    A, B = lb.to_m(g["mul_A"]), lb.to_m(g["mul_B"])
    Cm = lb.lib().mtrxmul(A, B)
    assert np.array_equal(lb.from_m(Cm), g["mul_C"])
    K1, K2 = lb.to_m(g["kron_A"]), lb.to_m(g["kron_B"])
    Km = lb.lib().kronecker(K1, K2)
    assert np.array_equal(lb.from_m(Km), g["kron_C"])
    R = lb.lib().reshape(A, 3, 10)
    assert np.array_equal(lb.from_m(R), g["reshape_out"])
    assert lb.lib().maxel(A) == float(g["maxel"]) and lb.lib().minel(A) == float(g["minel"])
    E = lb.lib().eye(4)
    assert np.array_equal(lb.from_m(E), np.eye(4))
    lb.lib().invsig(E)
    assert np.array_equal(lb.from_m(E), -np.eye(4))
    lb.free(A, B, Cm, K1, K2, Km, R, E)


def test_legacy_error_is_row_major_l1():
    rng = np.random.default_rng(4)
    a, b = rng.standard_normal((9, 11)), rng.standard_normal((9, 11))
    A, B = lb.to_m(a), lb.to_m(b)
    e = lb.lib().error(A, B)
    ref = 0.0
    for x, y in zip(a.ravel(), b.ravel()):
        ref += abs(y - x)
    assert e == ref
    lb.free(A, B)


def test_config_parser_matches_reference():
    g = load("config_parser.npz")
    cfg = lb.lib().load_config_from_file(os.path.join(GOLD, "config_parser_input.txt").encode())
    for name, _ in lb.Config._fields_:
        assert getattr(cfg, name) == g[name].item(), name
    d = lb.lib().load_default_config()
    assert (d.Re, d.nx, d.dt, d.tf, d.order, d.poisson_max_it, d.poisson_tol, d.u4) == \
        (1000.0, 64, 0.005, 20.0, 6, 10000, 1e-3, 1.0)


def test_config_parser_warns_on_unknown_key(tmp_path):
    p = tmp_path / "c.txt"
    p.write_text("nx = 32\nsor_omega = 1.9\n")
    code = ("import sys; sys.path.insert(0, %r); import legacy_bind as lb; "
            "c = lb.lib().load_config_from_file(%r.encode()); print('NX', c.nx)") % (
        os.path.dirname(os.path.abspath(__file__)), str(p))
    out = subprocess.run(["python3", "-c", code], capture_output=True, text=True, check=True).stdout
    assert "Warning: Unknown configuration parameter 'sor_omega'" in out
    assert "NX 32" in out


def test_vtk_writer_matches_reference(tmp_path):
    g = load("vtk_writer.npz")
    V = lb.to_m(g["values"])
    code = ("import sys; sys.path.insert(0, %r); import numpy as np, legacy_bind as lb; "
            "g = np.load(%r); V = lb.to_m(g['values']); lb.lib().printvtk(V, b'field', %r.encode())") % (
        os.path.dirname(os.path.abspath(__file__)), os.path.join(GOLD, "vtk_writer.npz"), str(tmp_path))
    subprocess.run(["python3", "-c", code], capture_output=True, text=True, check=True)
    assert (tmp_path / "field-1-0.vtk").read_text() == str(g["text"])
    lb.free(V)


def test_driver_help_runs_without_gpu():
    exe = os.path.join(lb.ROOT, "fluid-dynamics1_amd", "cnavier")
    out = subprocess.run([exe, "--help-config"], capture_output=True, text=True, check=True).stdout
    assert "poisson_max_it" in out and "sor_omega" in out
    assert "output_async" in out and "transport_fp32" in out


@pytest.mark.parametrize("order", [3, 8])
def test_bad_order_exits_like_reference(order):
    code = ("import sys; sys.path.insert(0, %r); import legacy_bind as lb; lb.lib().Diff1(16, %d, 0.1)") % (
        os.path.dirname(os.path.abspath(__file__)), order)
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True)
    assert r.returncode == 1
    assert "** Error: valid orders are 2, 4 or 6 **" in r.stdout


def test_batch_rejects_mixed_precision_without_gpu():
    import cnv
    ps = [cnv.default_params(32), cnv.default_params(40, precision=cnv.PREC_FP32_DC)]
    with pytest.raises(cnv.CnvError) as ei:
        cnv.Batch(ps)
    assert ei.value.code == cnv.ERR_ARG and "must match item 0" in str(ei.value)


def _make_tables():
    import importlib.util
    import cnv
    spec = importlib.util.spec_from_file_location("make_tables", os.path.join(cnv.WORKLOADS, "make_tables.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_batch_workload_table_is_seed0_draw():
    # the committed table (what cnv.batch_table() and bench.py read) is numpy's seed-0 draw, value
    # for value, and follows the workload's rules
    import cnv
    rows = cnv.batch_table()
    assert len(rows) == 64 and {r["n"] for r in rows} <= {64, 96, 128, 192, 256, 384, 512}
    assert all(100.0 <= r["Re"] <= 5000.0 and 0.5 <= r["U"] <= 1.5 for r in rows)
    for r in rows:  # dt = 0.2*min(h*h*Re/4, h), h = 1/N; reference relaxation rule
        h = 1.0 / r["n"]
        assert r["dt"] == 0.2 * min(h * h * r["Re"] / 4.0, h)
        assert r["beta"] == cnv.reference_beta(r["n"])
    assert rows == _make_tables().draw_batch(64, 0)
    assert cnv.batch_table(65, 0)[:64] != rows  # another count is a fresh draw, not the table


def test_gaussian_table_is_seed1_draw():
    import numpy as np
    import cnv
    cx, cy, width, amp = cnv.gaussian_parameters()
    drawn = _make_tables().draw_gaussians(32, 1)
    for got, key in zip((cx, cy, width, amp), ("cx", "cy", "width", "amplitude")):
        assert np.array_equal(got, np.array(drawn[key]))
    assert np.all((0 <= cx) & (cx < 1) & (0 <= cy) & (cy < 1)) and np.all((0.02 <= width) & (width < 0.1))
    # the field at the nodes, from the table: a Gaussian's peak sits where its centre's node is
    w = cnv.gaussian_vorticity(64)
    x = np.arange(64) / 64
    k = int(np.argmax(np.abs(amp)))
    ref = sum(amp[m] * np.exp(-((x[None, :] - cx[m]) ** 2 + (x[:, None] - cy[m]) ** 2) / (2 * width[m] ** 2))
              for m in range(32))
    assert np.allclose(w, ref, rtol=0, atol=1e-12) and abs(w).max() > 0.5 * abs(amp[k])


def test_params_layout_matches_library():
    import ctypes
    import cnv
    L = cnv.lib()
    assert L.cnv_params_size() == ctypes.sizeof(cnv.Params)


def test_device_api_argument_errors_without_gpu():
    # argument checks come before any HIP call: each bad call returns CNV_ERR_ARG with a message
    import ctypes
    import threading
    import cnv
    L = cnv.lib()
    p = cnv.default_params(32)
    assert L.cnv_create(ctypes.byref(p), None) == cnv.ERR_ARG
    assert b"null output pointer" in L.cnv_last_error(None)
    h = ctypes.c_void_p()
    bad = cnv.default_params(32, order=3)
    assert L.cnv_create(ctypes.byref(bad), ctypes.byref(h)) == cnv.ERR_ARG and not h.value
    assert b"valid orders are 2, 4 or 6" in L.cnv_last_error(None)
    assert L.cnv_step(None, 1, None) == cnv.ERR_ARG
    assert L.cnv_set_state(None, None, None, None, None) == cnv.ERR_ARG
    assert L.cnv_get_state(None, None, None, None, None) == cnv.ERR_ARG
    assert L.cnv_device_state(None, None, None, None, None) == cnv.ERR_ARG
    assert L.cnv_poisson_device(None, None, None, None, None) == cnv.ERR_ARG
    assert L.cnv_sweeps_per_launch(None) == 0
    assert L.cnv_batch_create(ctypes.byref(p), 1, None) == cnv.ERR_ARG
    assert L.cnv_adj_create(ctypes.byref(p), 10, 1e-12, 1e-12, None) == cnv.ERR_ARG
    for every in (-1, 3, 128):
        bad_every = cnv.default_params(32, inner_test_every=every)
        assert L.cnv_create(ctypes.byref(bad_every), ctypes.byref(h)) == cnv.ERR_ARG and not h.value
        assert b"inner_test_every" in L.cnv_last_error(None)
    # the fallback message (no context) belongs to the calling thread
    assert L.cnv_step(None, -1, None) == cnv.ERR_ARG
    mine = L.cnv_last_error(None)
    seen = []

    def other():
        L.cnv_create(ctypes.byref(bad), ctypes.byref(ctypes.c_void_p()))
        seen.append(L.cnv_last_error(None))

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert b"valid orders" in seen[0]
    assert L.cnv_last_error(None) == mine == b"bad arguments to cnv_step"
