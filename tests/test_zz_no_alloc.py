"""GPU test of cnv_poisson_device_noalloc, the Poisson solve that allocates nothing: its scratch is
a static device array of the library.  It must give what cnv_poisson_device gives on a context of
the same grid and solve parameters -- the same iterate bit for bit, the same k and e -- on calls
in a row that reuse the array (a larger grid first, then a smaller one, whose solve must not see
the first one's leftovers), up to the largest grid it accepts, and refuse larger grids."""
import ctypes as C

import numpy as np
import pytest

import cnv

pytestmark = pytest.mark.gpu


def source(n, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(n) / n
    return np.sin(2 * np.pi * x)[None, :] * np.cos(3 * np.pi * x)[:, None] + 0.1 * rng.standard_normal((n, n))


def solve_both(n, beta, tol, ptype, f):
    """The same solve through cnv_poisson_device (a context) and cnv_poisson_device_noalloc, on device
    buffers of a helper context: its w holds f, its u receives the first psi, its v the second."""
    p = cnv.default_params(n, beta=beta, tol=tol, itmax=1_000_000, poisson_type=ptype)
    L = cnv.lib()
    with cnv.Cavity(p) as aux, cnv.Cavity(p) as ctx:
        aux.set_state(w=f, u=np.zeros((n, n)), v=np.full((n, n), 7.0))
        dw, du, dv = C.c_void_p(), C.c_void_p(), C.c_void_p()
        cnv._check(L.cnv_device_state(aux.h, C.byref(dw), None, C.byref(du), C.byref(dv)), aux.h)
        k1, e1, k2, e2 = C.c_int(), C.c_double(), C.c_int(), C.c_double()
        assert L.cnv_poisson_device(ctx.h, dw, du, C.byref(k1), C.byref(e1)) == cnv.OK
        rc = L.cnv_poisson_device_noalloc(n, p.dx, p.dy, beta, tol, p.itmax, ptype, dw, dv, C.byref(k2), C.byref(e2))
        assert rc == cnv.OK, L.cnv_last_error(None)
        s = aux.get_state()
    return s["u"], s["v"], (k1.value, e1.value), (k2.value, e2.value)


def test_noalloc_equals_context_solve_on_calls_in_a_row():
    cases = [(200, 1.9, 1e-9, 2, 1), (72, 1.0, 1e-9, 1, 2), (1024, 1.95, 1e-3, 2, 3), (48, 1.7, 1e-10, 2, 4)]
    for n, beta, tol, ptype, seed in cases:  # 72 and 48 reuse the array after larger grids
        print("noalloc: n=%d beta=%g tol=%g type=%d" % (n, beta, tol, ptype), flush=True)
        a, b, ka, kb = solve_both(n, beta, tol, ptype, source(n, seed))
        assert ka == kb, (n, ka, kb)
        assert np.array_equal(a, b), (n, np.max(np.abs(a - b)))
        assert np.all(a[0, :] == 0) and np.all(a[:, -1] == 0) and np.any(a != 0)


def test_noalloc_refuses_grids_beyond_its_scratch():
    L = cnv.lib()
    n = 1025  # one side more than the static array holds
    k, e = C.c_int(), C.c_double()
    rc = L.cnv_poisson_device_noalloc(n, 1.0 / n, 1.0 / n, 1.9, 1e-8, 100, 2, C.c_void_p(16), C.c_void_p(16),
                                      C.byref(k), C.byref(e))
    assert rc == cnv.ERR_ARG and b"static scratch" in L.cnv_last_error(None)
    assert L.cnv_poisson_device_noalloc(64, 1 / 64, 1 / 64, 1.9, 1e-8, 100, 3, C.c_void_p(16), C.c_void_p(16),
                                        None, None) == cnv.ERR_ARG
    assert L.cnv_poisson_device_noalloc(64, 1 / 64, 1 / 64, 1.9, 1e-8, 100, 2, None, None, None, None) == cnv.ERR_ARG
