"""ctypes binding of the CPU restatement (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product library never loads it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LEX, REDBLACK = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("order", C.c_int), ("poisson_type", C.c_int), ("ordering", C.c_int),
        ("itmax", C.c_int), ("threads", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("Re", C.c_double), ("dt", C.c_double), ("beta", C.c_double), ("tol", C.c_double),
        ("ui", C.c_double), ("vi", C.c_double),
        ("u1", C.c_double), ("u2", C.c_double), ("u3", C.c_double), ("u4", C.c_double),
        ("v1", C.c_double), ("v2", C.c_double), ("v3", C.c_double), ("v4", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [("status", C.c_int), ("k", C.c_int), ("sweeps", C.c_int), ("e", C.c_double),
                ("cont_max", C.c_double), ("cont_min", C.c_double)]


_DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_diff1.argtypes = [C.c_int, C.c_int, C.c_double, _DP]
        L.orc_diff2.argtypes = [C.c_int, C.c_int, C.c_double, _DP]
        L.orc_apply_x.argtypes = [C.c_int, _DP, _DP, _DP]
        L.orc_apply_y.argtypes = [C.c_int, _DP, _DP, _DP]
        L.orc_dense_kron_apply.argtypes = [C.c_int, _DP, C.c_int, _DP, _DP]
        L.orc_error.argtypes = [C.c_int, _DP, _DP]
        L.orc_error.restype = C.c_double
        L.orc_poisson.argtypes = [C.POINTER(Params), _DP, _DP, C.POINTER(Stats)]
        L.orc_init_state.argtypes = [C.POINTER(Params), _DP, _DP, _DP, _DP]
        L.orc_step.argtypes = [C.POINTER(Params), _DP, _DP, _DP, _DP, C.POINTER(Stats)]
        L.orc_run.argtypes = [C.POINTER(Params), C.c_int, _DP, _DP, _DP, _DP, C.POINTER(Stats)]
        L.orc_apply_velocity_bc.argtypes = [C.POINTER(Params), _DP, _DP]
        L.orc_wall_vorticity.argtypes = [C.POINTER(Params), _DP, _DP, _DP]
        L.orc_transport.argtypes = [C.POINTER(Params), _DP, _DP, _DP]
        L.orc_velocity.argtypes = [C.POINTER(Params), _DP, _DP, _DP,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_euler.argtypes = [C.c_int] + [_DP] * 7 + [C.c_double, C.c_double]
        L.orc_beta.argtypes = [C.c_int, C.c_int]
        L.orc_beta.restype = C.c_double
        _lib = L
    return _lib


def beta(n):
    return lib().orc_beta(n, n)


def make_params(n, Re=1000.0, dt=0.005, order=6, poisson_type=2, ordering=LEX, itmax=10000,
                tol=1e-3, beta_=None, Lx=1, Ly=1, threads=1, ui=0.0, vi=0.0,
                u=(0.0, 0.0, 0.0, 1.0), v=(0.0, 0.0, 0.0, 0.0)):
    """Parameters with the reference's conventions: dx = Lx/nx (src/main.c:138), beta from
    src/main.c:134 unless given."""
    p = Params()
    p.n, p.order, p.poisson_type, p.ordering, p.itmax, p.threads = n, order, poisson_type, ordering, itmax, threads
    p.dx, p.dy = float(Lx) / n, float(Ly) / n
    p.Re, p.dt, p.tol = Re, dt, tol
    p.beta = beta(n) if beta_ is None else beta_
    p.ui, p.vi = ui, vi
    p.u1, p.u2, p.u3, p.u4 = u
    p.v1, p.v2, p.v3, p.v4 = v
    return p


def diff1(n, o, h):
    D = np.zeros((n, n))
    rc = lib().orc_diff1(n, o, h, D)
    if rc:
        raise ValueError("orc_diff1 rc=%d" % rc)
    return D


def diff2(n, o, h):
    D = np.zeros((n, n))
    rc = lib().orc_diff2(n, o, h, D)
    if rc:
        raise ValueError("orc_diff2 rc=%d" % rc)
    return D


def apply_x(D, f):
    out = np.empty_like(f)
    lib().orc_apply_x(f.shape[0], np.ascontiguousarray(D), np.ascontiguousarray(f), out)
    return out


def apply_y(D, f):
    out = np.empty_like(f)
    lib().orc_apply_y(f.shape[0], np.ascontiguousarray(D), np.ascontiguousarray(f), out)
    return out


def dense_kron_apply(D, f, kron_x):
    out = np.empty_like(f)
    rc = lib().orc_dense_kron_apply(f.shape[0], np.ascontiguousarray(D), int(kron_x),
                                    np.ascontiguousarray(f), out)
    if rc:
        raise ValueError("grid too large for the dense Kronecker check")
    return out


def poisson(p, f):
    """Returns (psi, stats) for the Poisson solve of src/poisson.c (psi starts from zero)."""
    psi = np.zeros_like(f)
    st = Stats()
    lib().orc_poisson(C.byref(p), np.ascontiguousarray(f), psi, C.byref(st))
    return psi, st


class State:
    def __init__(self, p):
        n = p.n
        self.p = p
        self.w, self.psi = np.zeros((n, n)), np.zeros((n, n))
        self.u, self.v = np.zeros((n, n)), np.zeros((n, n))
        lib().orc_init_state(C.byref(p), self.w, self.psi, self.u, self.v)

    def step(self):
        st = Stats()
        rc = lib().orc_step(C.byref(self.p), self.w, self.psi, self.u, self.v, C.byref(st))
        if rc:
            raise RuntimeError("oracle step failed rc=%d (k=%d e=%g)" % (rc, st.k, st.e))
        return st

    def run(self, nsteps):
        arr = (Stats * nsteps)()
        rc = lib().orc_run(C.byref(self.p), nsteps, self.w, self.psi, self.u, self.v, arr)
        if rc:
            raise RuntimeError("oracle run failed rc=%d" % rc)
        return list(arr)
