"""ctypes binding of the compiled reference (oracle/_ref/libcnavier_ref*.so).

The reference is compiled from its own sources by oracle/Makefile (target "ref"); this module
only calls the functions it exports (include/linearalg.h, finitediff.h, poisson.h, fluiddyn.h,
config.h of the reference).  Used by make_golden.py to write fixtures and by the CPU tests that
cross-check the oracle against the reference when oracle/_ref is present.
"""
import ctypes as C  # noqa: N812
import os
import re
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


class Mtrx(C.Structure):
    # include/linearalg.h:6-11 of the reference: {double **M; int m; int n;}
    _fields_ = [("M", C.POINTER(C.POINTER(C.c_double))), ("m", C.c_int), ("n", C.c_int)]


class Config(C.Structure):
    # include/config.h:4-36 of the reference (field order and types)
    _fields_ = [("Re", C.c_double), ("Lx", C.c_int), ("Ly", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("dt", C.c_double), ("tf", C.c_double), ("max_co", C.c_double),
                ("order", C.c_int), ("poisson_max_it", C.c_int), ("poisson_tol", C.c_double),
                ("output_interval", C.c_int), ("poisson_type", C.c_int),
                ("openmp_enabled", C.c_int),
                ("ui", C.c_double), ("vi", C.c_double),
                ("u1", C.c_double), ("u2", C.c_double), ("u3", C.c_double), ("u4", C.c_double),
                ("v1", C.c_double), ("v2", C.c_double), ("v3", C.c_double), ("v4", C.c_double)]


def available():
    return os.path.exists(os.path.join(REF_DIR, "libcnavier_ref.so"))


_libc = C.CDLL(None)
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]


class Ref:
    """One compiled variant: omp=False is the stock Makefile build (serial, lexicographic SOR);
    omp=True is -DOPENMP_ENABLED (red-black SOR); run it with OMP_NUM_THREADS=1 for the
    deterministic serial error() summation."""

    def __init__(self, omp=False):
        path = os.path.join(REF_DIR, "libcnavier_ref_omp.so" if omp else "libcnavier_ref.so")
        L = C.CDLL(path)
        self.omp = omp
        M = Mtrx
        L.initm.restype = M
        L.initm.argtypes = [C.c_int, C.c_int]
        L.freem.restype = C.c_void_p
        L.freem.argtypes = [M]
        L.mtrxmul.restype = M
        L.mtrxmul.argtypes = [M, M]
        L.kronecker.restype = M
        L.kronecker.argtypes = [M, M]
        L.eye.restype = M
        L.eye.argtypes = [C.c_int]
        L.reshape.restype = M
        L.reshape.argtypes = [M, C.c_int, C.c_int]
        L.Diff1.restype = M
        L.Diff1.argtypes = [C.c_int, C.c_int, C.c_double]
        L.Diff2.restype = M
        L.Diff2.argtypes = [C.c_int, C.c_int, C.c_double]
        L.euler.argtypes = [M, M, M, M, M, M, M, C.c_double, C.c_double]
        L.continuity.restype = M
        L.continuity.argtypes = [M, M]
        L.vorticity.restype = M
        L.vorticity.argtypes = [M, M]
        L.maxel.restype = C.c_double
        L.maxel.argtypes = [M]
        L.minel.restype = C.c_double
        L.minel.argtypes = [M]
        L.invsig.argtypes = [M]
        L.error.restype = C.c_double
        L.error.argtypes = [M, M]
        L.poisson_log.restype = M
        L.poisson_log.argtypes = [M, C.c_double, C.c_double, C.c_int, C.c_double, C.c_void_p]
        L.poisson_SOR_log.restype = M
        L.poisson_SOR_log.argtypes = [M, C.c_double, C.c_double, C.c_int, C.c_double,
                                      C.c_double, C.c_void_p]
        L.set_openmp_config.argtypes = [C.c_int]
        L.set_poisson_openmp_config.argtypes = [C.c_int]
        L.set_fluiddyn_openmp_config.argtypes = [C.c_int]
        L.load_config_from_file.restype = Config
        L.load_config_from_file.argtypes = [C.c_char_p]
        L.load_default_config.restype = Config
        self.L = L
        on = 1 if omp else 0
        L.set_openmp_config(on)
        L.set_poisson_openmp_config(on)
        L.set_fluiddyn_openmp_config(on)

    # ---- containers -------------------------------------------------------------------
    def to_m(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        m, n = a.shape
        A = self.L.initm(m, n)
        for i in range(m):
            C.memmove(A.M[i], a[i].ctypes.data, 8 * n)
        return A

    @staticmethod
    def from_m(A):
        out = np.empty((A.m, A.n))
        for i in range(A.m):
            out[i] = np.ctypeslib.as_array(A.M[i], shape=(A.n,))
        return out

    def free(self, *ms):
        for A in ms:
            self.L.freem(A)

    # ---- single functions -------------------------------------------------------------
    def diff(self, which, n, o, h):
        D = (self.L.Diff1 if which == 1 else self.L.Diff2)(n, o, h)
        a = self.from_m(D)
        self.free(D)
        return a

    def euler(self, w, dwdx, dwdy, d2x, d2y, u, v, Re, dt):
        ms = [self.to_m(x) for x in (w, dwdx, dwdy, d2x, d2y, u, v)]
        self.L.euler(*ms, Re, dt)
        out = self.from_m(ms[0])
        self.free(*ms)
        return out

    def poisson(self, f, dx, dy, itmax, tol, beta=None):
        """poisson_log (beta None) or poisson_SOR_log; returns (psi, k, e_string)."""
        F = self.to_m(f)
        fd, path = tempfile.mkstemp(suffix=".log")
        os.close(fd)
        fp = _libc.fopen(path.encode(), b"w")
        try:
            if beta is None:
                P = self.L.poisson_log(F, dx, dy, itmax, tol, fp)
            else:
                P = self.L.poisson_SOR_log(F, dx, dy, itmax, tol, beta, fp)
        finally:
            _libc.fclose(fp)
        psi = self.from_m(P)
        self.free(F, P)
        with open(path) as fh:
            txt = fh.read()
        os.unlink(path)
        k, e = parse_solved(txt)[-1]
        return psi, k, e

    # ---- the time step in src/main.c order ----------------------------------------------
    def run_steps(self, n, order, Re, dt, tol, itmax, steps, poisson_type=2, beta=None, Lx=1,
                  Ly=1, ui=0.0, vi=0.0, ub=(0.0, 0.0, 0.0, 1.0), vb=(0.0, 0.0, 0.0, 0.0)):
        """Calls the reference library in the order of src/main.c:133-451 (the dense Kronecker
        operators included).  Returns final (w, psi, u, v) and per-step (k, e_str, cmax, cmin)."""
        L = self.L
        if beta is None:
            beta = 0.5 * (2 / (1 + np.sin(3.14159265359 / (n + 1))) +
                          2 / (1 + np.sin(3.14159265359 / (n + 1))))
        dx, dy = float(Lx) / n, float(Ly) / n
        d_x, d_y = L.Diff1(n, order, dx), L.Diff1(n, order, dy)
        d_x2, d_y2 = L.Diff2(n, order, dx), L.Diff2(n, order, dy)
        Ix, Iy = L.eye(n), L.eye(n)
        DX, DY = L.kronecker(Ix, d_x), L.kronecker(d_y, Iy)
        DX2, DY2 = L.kronecker(Ix, d_x2), L.kronecker(d_y2, Iy)
        self.free(d_x, d_y, d_x2, d_y2, Ix, Iy)

        def matvec(D, a):
            x = self.to_m(a.reshape(n * n, 1))
            y = L.mtrxmul(D, x)
            r = self.from_m(y).reshape(n, n)
            self.free(x, y)
            return r

        u, v, w, psi = (np.zeros((n, n)) for _ in range(4))
        u[1:-1, 1:-1] = ui
        v[1:-1, 1:-1] = vi
        u1, u2, u3, u4 = ub
        v1, v2, v3, v4 = vb
        stats = []
        fd, path = tempfile.mkstemp(suffix=".log")
        os.close(fd)
        fp = _libc.fopen(path.encode(), b"w")
        try:
            for t in range(steps):
                v[0, :], v[n - 1, :], u[0, :], u[n - 1, :] = v3, v4, u3, u4
                v[:, 0], v[:, n - 1], u[:, 0], u[:, n - 1] = v1, v2, u1, u2
                dudy, dvdx = matvec(DY, u), matvec(DX, v)
                w[0, :] = dvdx[0, :] - dudy[0, :]
                w[n - 1, :] = dvdx[n - 1, :] - dudy[n - 1, :]
                w[:, 0] = dvdx[:, 0] - dudy[:, 0]
                w[:, n - 1] = dvdx[:, n - 1] - dudy[:, n - 1]
                dwdx, dwdy = matvec(DX, w), matvec(DY, w)
                d2wdx2, d2wdy2 = matvec(DX2, w), matvec(DY2, w)
                w = self.euler(w, dwdx, dwdy, d2wdx2, d2wdy2, u, v, Re, dt)
                W = self.to_m(w)
                L.invsig(W)
                if poisson_type == 1:
                    P = L.poisson_log(W, dx, dy, itmax, tol, fp)
                else:
                    P = L.poisson_SOR_log(W, dx, dy, itmax, tol, beta, fp)
                L.invsig(W)
                w = self.from_m(W)
                psi = self.from_m(P)
                self.free(W, P)
                dpsidx, dpsidy = matvec(DX, psi), matvec(DY, psi)
                u = dpsidy.copy()
                v = -dpsidx
                dudx, dvdy = matvec(DX, u), matvec(DY, v)
                A, B = self.to_m(dudx), self.to_m(dvdy)
                Cm = L.continuity(A, B)
                stats.append((L.maxel(Cm), L.minel(Cm)))
                self.free(A, B, Cm)
        finally:
            _libc.fclose(fp)
            self.free(DX, DY, DX2, DY2)
        with open(path) as fh:
            solved = parse_solved(fh.read())
        os.unlink(path)
        assert len(solved) == steps, "reference Poisson did not converge every step"
        per_step = [(k, e, cmax, cmin) for (k, e), (cmax, cmin) in zip(solved, stats)]
        return (w, psi, u, v), per_step


_SOLVED = re.compile(r"solved with (\d+) iterations - root-sum-of-squares error: (\S+)")


def parse_solved(txt):
    return [(int(m.group(1)), m.group(2)) for m in _SOLVED.finditer(txt)]
