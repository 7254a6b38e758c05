"""FP64 CPU adjoint of the cavity step (test harness only): the forward step of src/main.c:276-451
written with dense operators and an exact Poisson solve, and its reverse-mode gradient, for the
objective E = 1/2 dx dy sum(u^2 + v^2) after nsteps.  Pinned by its own central finite differences
(tests/test_adjoint_cpu.py) and then used to check the GPU adjoint kernels.

Operators (src/main.c:142-152): Dx f = f @ d1x.T (acts along j), Dy f = d1y @ f (along i), with
d1x = Diff1(n, o, dx), d1y = Diff1(n, o, dy) and the same for Diff2; tables from the oracle.
Poisson (src/poisson.c:268): the SOR fixed point on the interior, A psi = c3 w with
A psi = B psi - c1 (N + S) - c2 (E + W), c1 = dy^2, c2 = dx^2, c3 = dx^2 dy^2, B = 2(dx^2+dy^2),
psi = 0 on the walls; solved exactly here.
"""
import numpy as np

from oracle import orc


class CavityAdjoint:
    def __init__(self, n, Re, dt, order=6, walls=(0.0, 0.0, 0.0), ui=0.0, vi=0.0):
        self.n, self.Re, self.dt = n, Re, dt
        self.dx = self.dy = 1.0 / n
        self.d1x, self.d1y = orc.diff1(n, order, self.dx), orc.diff1(n, order, self.dy)
        self.d2x, self.d2y = orc.diff2(n, order, self.dx), orc.diff2(n, order, self.dy)
        self.u1, self.u2, self.u3 = walls  # u on walls j=0, j=n-1, i=0; v = 0 on every wall
        self.ui, self.vi = ui, vi
        dx, dy = self.dx, self.dy
        self.c1, self.c2, self.c3 = dy * dy, dx * dx, dx * dx * dy * dy
        B = 2 * (dx * dx + dy * dy)
        m = n - 2
        A = np.zeros((m * m, m * m))
        idx = lambda i, j: (i - 1) * m + (j - 1)
        for i in range(1, n - 1):
            for j in range(1, n - 1):
                r = idx(i, j)
                A[r, r] = B
                for (ii, jj, c) in ((i + 1, j, self.c1), (i - 1, j, self.c1), (i, j + 1, self.c2), (i, j - 1, self.c2)):
                    if 1 <= ii <= n - 2 and 1 <= jj <= n - 2:
                        A[r, idx(ii, jj)] -= c
        self.Ainv = np.linalg.inv(A)

    # ---- pieces of the forward step ----
    def bc(self, u, v, U):
        ub, vb = u.copy(), v.copy()
        ub[0, :], ub[-1, :] = self.u3, U          # rows first (src/main.c:283-296) ...
        ub[:, 0], ub[:, -1] = self.u1, self.u2    # ... columns win at the corners
        vb[0, :] = vb[-1, :] = 0.0
        vb[:, 0] = vb[:, -1] = 0.0
        return ub, vb

    def Dx(self, f):
        return f @ self.d1x.T

    def Dy(self, f):
        return self.d1y @ f

    def edge_mask(self):
        E = np.zeros((self.n, self.n), dtype=bool)
        E[0, :] = E[-1, :] = E[:, 0] = E[:, -1] = True
        return E

    def solve(self, rhs):
        """psi = embed(A^-1 rhs_interior) -- A is symmetric, so this is also the adjoint solve."""
        n = self.n
        out = np.zeros((n, n))
        out[1:-1, 1:-1] = (self.Ainv @ rhs[1:-1, 1:-1].ravel()).reshape(n - 2, n - 2)
        return out

    def step(self, w, u, v, U):
        ub, vb = self.bc(u, v, U)
        E = self.edge_mask()
        wt = w.copy()
        wt[E] = (self.Dx(vb) - self.Dy(ub))[E]
        lap = wt @ self.d2x.T + self.d2y @ wt
        w1 = (-ub * self.Dx(wt) - vb * self.Dy(wt) + (1.0 / self.Re) * lap) * self.dt + wt
        psi = self.solve(self.c3 * w1)
        return w1, self.Dy(psi), -self.Dx(psi), (ub, vb, wt)

    def initial(self, w0):
        n = self.n
        u = np.zeros((n, n))
        v = np.zeros((n, n))
        u[1:-1, 1:-1], v[1:-1, 1:-1] = self.ui, self.vi
        return w0.copy(), u, v

    def energy(self, w0, U, nsteps):
        w, u, v = self.initial(w0)
        for _ in range(nsteps):
            w, u, v, _ = self.step(w, u, v, U)
        return 0.5 * self.dx * self.dy * float(np.sum(u * u + v * v))

    def gradient(self, w0, U, nsteps):
        """(E, dE/dw0, dE/dU) by reverse mode over the stored trajectory."""
        states = []
        w, u, v = self.initial(w0)
        for _ in range(nsteps):
            states.append((w, u, v))
            w, u, v, _ = self.step(w, u, v, U)
        h2 = self.dx * self.dy
        Eval = 0.5 * h2 * float(np.sum(u * u + v * v))
        lw, lu, lv = np.zeros_like(w), h2 * u, h2 * v
        dU = 0.0
        Em = self.edge_mask()
        inner = ~Em
        for (w, u, v) in reversed(states):
            _, _, _, (ub, vb, wt) = self.step(w, u, v, U)
            lpsi = self.d1y.T @ lu - lv @ self.d1x           # Dy^T lu - Dx^T lv
            lw = lw + self.c3 * self.solve(lpsi)
            lwt = lw + self.dt * (-((ub * lw) @ self.d1x) - self.d1y.T @ (vb * lw)
                                  + (1.0 / self.Re) * (lw @ self.d2x + self.d2y.T @ lw))
            lub = -self.dt * self.Dx(wt) * lw
            lvb = -self.dt * self.Dy(wt) * lw
            g = np.where(Em, lwt, 0.0)
            lvb = lvb + g @ self.d1x
            lub = lub - self.d1y.T @ g
            dU += float(np.sum(lub[-1, 1:-1]))
            lw, lu, lv = np.where(inner, lwt, 0.0), np.where(inner, lub, 0.0), np.where(inner, lvb, 0.0)
        return Eval, lw, dU
