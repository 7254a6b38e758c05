"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has /root/reference:   python tests/golden/make_golden.py
It builds oracle/_ref (make -C oracle ref) and calls the reference library through refbind.py.
Fixtures are data only (inputs and the reference's outputs), stored as .npz without pickles.
The red-black (OpenMP-build) variant runs with OMP_NUM_THREADS=1 so that error() sums in
row-major order, as the serial build does.
"""
import os
import subprocess
import sys

os.environ["OMP_NUM_THREADS"] = "1"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import refbind  # noqa: E402

# Step scenarios: name -> kwargs of Ref.run_steps plus the build variant.
STEP_CASES = {
    "s16_o6_sor_lex": dict(omp=False, n=16, order=6, Re=1000.0, dt=0.005, tol=1e-6, itmax=20000, steps=20),
    "s16_o6_sor_rb": dict(omp=True, n=16, order=6, Re=1000.0, dt=0.005, tol=1e-6, itmax=20000, steps=20),
    "s24_o4_sor18_lex": dict(omp=False, n=24, order=4, Re=400.0, dt=0.002, tol=1e-8, itmax=20000, steps=10, beta=1.8),
    "s24_o4_sor18_rb": dict(omp=True, n=24, order=4, Re=400.0, dt=0.002, tol=1e-8, itmax=20000, steps=10, beta=1.8),
    "s32_o2_gs_lex": dict(omp=False, n=32, order=2, Re=100.0, dt=0.001, tol=1e-7, itmax=50000, steps=6, poisson_type=1),
    "s32_o2_sorb1_rb": dict(omp=True, n=32, order=2, Re=100.0, dt=0.001, tol=1e-7, itmax=50000, steps=6, beta=1.0),
    "s32_o6_bc_rb": dict(omp=True, n=32, order=6, Re=1000.0, dt=0.002, tol=1e-10, itmax=50000, steps=8,
                         ui=0.25, vi=-0.125, ub=(0.1, -0.2, 0.3, 1.5), vb=(-0.4, 0.05, 0.2, -0.3)),
    "s12_o6_sor_lex": dict(omp=False, n=12, order=6, Re=1000.0, dt=0.005, tol=1e-9, itmax=20000, steps=15),
    "s32_o6_re5000_lex": dict(omp=False, n=32, order=6, Re=5000.0, dt=0.001, tol=5e-4, itmax=15000, steps=10),
    "s20_o6_lx2_rb": dict(omp=True, n=20, order=6, Re=300.0, dt=0.002, tol=1e-9, itmax=20000, steps=8, Lx=2, Ly=1),
    "s40_o6_sor195_rb": dict(omp=True, n=40, order=6, Re=1000.0, dt=0.001, tol=1e-10, itmax=50000, steps=5, beta=1.95),
}

# Poisson-only scenarios on seeded random sources: (n, beta or None for Gauss-Seidel, omp, tol)
POISSON_CASES = [
    (8, None, False, 1e-12), (17, None, False, 1e-11), (24, 1.5, False, 1e-11),
    (17, 1.95, False, 1e-12), (24, 1.5, True, 1e-11), (17, 1.95, True, 1e-12),
    (31, "default", True, 1e-12), (31, "default", False, 1e-12), (16, 1.0, True, 1e-12),
]


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("needs /root/reference")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    ser, omp = refbind.Ref(omp=False), refbind.Ref(omp=True)

    # --- Diff1 / Diff2 tables --------------------------------------------------------
    diff = {}
    for n in (8, 9, 16):
        for o in (2, 4, 6):
            for h in (1.0 / n, 0.3):
                diff["d1_n%d_o%d_h%s" % (n, o, h.hex())] = ser.diff(1, n, o, h)
                diff["d2_n%d_o%d_h%s" % (n, o, h.hex())] = ser.diff(2, n, o, h)
    np.savez_compressed(os.path.join(HERE, "diff_tables.npz"), **diff)

    # --- euler / continuity / vorticity on random inputs ------------------------------
    rng = np.random.default_rng(20240601)
    n = 11
    ins = {k: rng.standard_normal((n, n)) for k in ("w", "dwdx", "dwdy", "d2x", "d2y", "u", "v")}
    out = ser.euler(ins["w"], ins["dwdx"], ins["dwdy"], ins["d2x"], ins["d2y"], ins["u"], ins["v"], 737.0, 0.0031)
    A, B = ser.to_m(ins["u"]), ser.to_m(ins["v"])
    Cm, Vm = ser.L.continuity(A, B), ser.L.vorticity(A, B)
    cont, vort = ser.from_m(Cm), ser.from_m(Vm)
    ser.free(A, B, Cm, Vm)
    np.savez_compressed(os.path.join(HERE, "elementwise.npz"), Re=737.0, dt=0.0031, euler_out=out,
                        continuity_out=cont, vorticity_out=vort, **{"in_" + k: v for k, v in ins.items()})

    # --- host helpers: linearalg, config parser, VTK writer ---------------------------------
    host = {}
    A = rng.standard_normal((6, 5))
    B = rng.standard_normal((5, 3))
    Am, Bm = ser.to_m(A), ser.to_m(B)
    Cm = ser.L.mtrxmul(Am, Bm)
    host["mul_A"], host["mul_B"], host["mul_C"] = A, B, ser.from_m(Cm)
    K1, K2 = rng.standard_normal((3, 3)), rng.standard_normal((3, 3))
    k1, k2 = ser.to_m(K1), ser.to_m(K2)
    Km = ser.L.kronecker(k1, k2)
    host["kron_A"], host["kron_B"], host["kron_C"] = K1, K2, ser.from_m(Km)
    Rm = ser.L.reshape(Am, 3, 10)
    host["reshape_in"], host["reshape_out"] = A, ser.from_m(Rm)
    host["maxel"] = np.array(ser.L.maxel(Am))
    host["minel"] = np.array(ser.L.minel(Am))
    ser.free(Am, Bm, Cm, k1, k2, Km, Rm)
    np.savez_compressed(os.path.join(HERE, "host_linearalg.npz"), **host)

    cfg_text = ("# comment line\n; another comment\n\nRe = 2500.5\nnx=48\n  ny   =   48   \n"
                "dt = 0.002 # trailing comment is dropped\ntf = 1.5\norder = 4\npoisson_tol = 2.5E-7\n"
                "poisson_type = 1\nbogus_key = 3\nu4 = 1.25\nv2=-0.5\nLx = 2\noutput_interval = 7\n"
                "poisson_max_it = 1234\nmax_co = 0.75\nopenmp_enabled = 1\nui = 0.125\n")
    cfg_path = os.path.join(HERE, "config_parser_input.txt")
    with open(cfg_path, "w") as fh:
        fh.write(cfg_text)
    cfg = ser.L.load_config_from_file(cfg_path.encode())
    np.savez_compressed(os.path.join(HERE, "config_parser.npz"),
                        **{name: np.array(getattr(cfg, name)) for name, _ in refbind.Config._fields_})

    import tempfile
    tmp = tempfile.mkdtemp()
    V = rng.standard_normal((5, 7)) * 3
    Vm = ser.to_m(V)
    ser.L.printvtk.argtypes = [refbind.Mtrx, refbind.C.c_char_p, refbind.C.c_char_p]
    ser.L.printvtk(Vm, b"field", tmp.encode())
    with open(os.path.join(tmp, "field-1-0.vtk")) as fh:
        vtk_text = fh.read()
    ser.free(Vm)
    np.savez_compressed(os.path.join(HERE, "vtk_writer.npz"), values=V, text=np.array(vtk_text))

    # --- Poisson --------------------------------------------------------------------------
    pois = {}
    for idx, (n, beta, use_omp, tol) in enumerate(POISSON_CASES):
        r = omp if use_omp else ser
        f = rng.standard_normal((n, n)) * 50.0
        dx = dy = 1.0 / n
        if beta == "default":
            s = np.sin(3.14159265359 / (n + 1))
            beta = 0.5 * (2 / (1 + s) + 2 / (1 + s))  # src/main.c:134 with nx == ny
        psi, k, e = r.poisson(f, dx, dy, 200000, tol, beta)
        key = "p%d" % idx
        pois[key + "_f"] = f
        pois[key + "_psi"] = psi
        pois[key + "_meta"] = np.array([n, -1.0 if beta is None else beta, 1.0 if use_omp else 0.0, tol, k])
        pois[key + "_e"] = np.array(e)
    np.savez_compressed(os.path.join(HERE, "poisson_cases.npz"), **pois)

    # --- full steps -----------------------------------------------------------------------
    for name, kw in STEP_CASES.items():
        kw = dict(kw)
        r = omp if kw.pop("omp") else ser
        (w, psi, u, v), per = r.run_steps(**kw)
        meta = dict(kw)
        meta.setdefault("poisson_type", 2)
        meta.setdefault("beta", -1.0)
        meta.setdefault("Lx", 1)
        meta.setdefault("Ly", 1)
        meta.setdefault("ui", 0.0)
        meta.setdefault("vi", 0.0)
        meta.setdefault("ub", (0.0, 0.0, 0.0, 1.0))
        meta.setdefault("vb", (0.0, 0.0, 0.0, 0.0))
        np.savez_compressed(
            os.path.join(HERE, "step_%s.npz" % name), w=w, psi=psi, u=u, v=v,
            k=np.array([p[0] for p in per]), e=np.array([p[1] for p in per]),
            cmax=np.array([p[2] for p in per]), cmin=np.array([p[3] for p in per]),
            ordering=np.array(1 if r.omp else 0),
            **{"meta_" + k: np.array(v) for k, v in meta.items()})
        print("wrote", name, "k:", [p[0] for p in per][:6], "...")


if __name__ == "__main__":
    main()
