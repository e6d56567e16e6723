"""A differentiable FP64 restatement of the 2-D particle step in torch.

Test infrastructure: an independent implementation of the step the adjoint
engine differentiates (Yee B half steps, node-centred CIC gather, Boris,
zigzag deposit, periodic halos), written with torch tensor operations so
that torch.autograd gives the reference gradient of the objective
1/2 sum B^2 with respect to the initial momenta.  Its forward is pinned
against the C oracle in tests/test_torch_step.py.

Fields are (mx, my) tensors with the reference's ghost layers; element
[i-1, j-1] is cell/node (i, j).  Particles are described by their cell
(ci, cj, 0-based physical), in-cell offsets and momenta.
"""
import numpy as np
import torch


class Step2D:
    def __init__(self, nx, ny, c=0.45, c_ompe=10.0, epc=16, g0=1.0, mi_me=1.0):
        self.nx, self.ny = nx, ny
        self.mx, self.my = nx + 5, ny + 5
        self.c = c
        ompe = c / c_ompe
        self.qi = (g0 * ompe * ompe) / (epc * (1.0 + 1.0 / mi_me))
        self.qm = (0.5 / mi_me, -0.5)     # flvrqm / 2
        self.P = (slice(2, self.mx - 3), slice(2, self.my - 3))
        i = np.arange(self.mx) + 1
        si = i.copy(); si[i <= 2] += nx; si[i >= self.mx - 2] -= nx
        j = np.arange(self.my) + 1
        sj = j.copy(); sj[j <= 2] += ny; sj[j >= self.my - 2] -= ny
        self._gi, self._gj = np.ix_(si - 1, sj - 1)

    def ghost(self, F):
        return F[self._gi, self._gj]

    def fold(self, J):
        mx, my = self.mx, self.my
        J = J.clone()
        for a, b in ((3, mx - 2), (4, mx - 1), (5, mx), (mx - 4, 1), (mx - 3, 2)):
            J[a - 1, :] = J[a - 1, :] + J[b - 1, :]
        xs = slice(2, mx - 3)
        for a, b in ((3, my - 2), (4, my - 1), (5, my), (my - 4, 1), (my - 3, 2)):
            J[xs, a - 1] = J[xs, a - 1] + J[xs, b - 1]
        return J

    def _sh(self, F, di, dj):
        return F[2 + di:self.mx - 3 + di, 2 + dj:self.my - 3 + dj]

    def b_half(self, E, B):
        h = self.c * 0.5
        Ex, Ey, Ez = E
        Bx, By, Bz = [b.clone() for b in B]
        P, s = self.P, self._sh
        Bx[P] = B[0][P] - h * (s(Ez, 0, 1) - s(Ez, 0, 0))
        By[P] = B[1][P] + h * (s(Ez, 1, 0) - s(Ez, 0, 0))
        Bz[P] = B[2][P] - h * (s(Ey, 1, 0) - s(Ey, 0, 0)) + h * (s(Ex, 0, 1) - s(Ex, 0, 0))
        return [self.ghost(b) for b in (Bx, By, Bz)]

    def e_update(self, E, B, J):
        f = self.c
        Bx, By, Bz = B
        Ex, Ey, Ez = [e.clone() for e in E]
        P, s = self.P, self._sh
        Ex[P] = E[0][P] + f * (s(Bz, 0, 0) - s(Bz, 0, -1)) - J[0][P]
        Ey[P] = E[1][P] - f * (s(Bz, 0, 0) - s(Bz, -1, 0)) - J[1][P]
        Ez[P] = E[2][P] + f * (s(By, 0, 0) - s(By, -1, 0)) - f * (s(Bx, 0, 0) - s(Bx, 0, -1)) - J[2][P]
        return [self.ghost(e) for e in (Ex, Ey, Ez)]

    def nodes(self, E, B):
        mx, my = self.mx, self.my
        Ex, Ey, Ez = E
        Bx, By, Bz = B

        def s(F, di, dj):
            return F[2 + di:mx - 2 + di, 2 + dj:my - 2 + dj]

        vals = [(s(Ex, -1, 0) + s(Ex, 0, 0)) * 0.5, (s(Ey, 0, -1) + s(Ey, 0, 0)) * 0.5, s(Ez, 0, 0),
                (s(Bx, 0, -1) + s(Bx, 0, 0)) * 0.5, (s(By, -1, 0) + s(By, 0, 0)) * 0.5,
                (s(Bz, -1, -1) + s(Bz, -1, 0) + s(Bz, 0, -1) + s(Bz, 0, 0)) * 0.25]
        out = []
        for v in vals:
            t = torch.zeros(mx, my, dtype=torch.float64)
            t[2:mx - 2, 2:my - 2] = v
            out.append(t)
        return out

    def push(self, st, N):
        ci, cj, sp, ox, oy, oz, u, v, w = st
        c = self.c
        f64 = torch.float64
        qm = torch.where(sp == 0, torch.tensor(self.qm[0], dtype=f64), torch.tensor(self.qm[1], dtype=f64))
        qs = torch.where(sp == 0, torch.tensor(self.qi, dtype=f64), torch.tensor(-self.qi, dtype=f64))
        i, j = ci + 2, cj + 2
        gc = [torch.stack([N[k][i, j], N[k][i + 1, j], N[k][i, j + 1], N[k][i + 1, j + 1]]) for k in range(6)]
        w1 = (1 - ox) * (1 - oy); w2 = ox * (1 - oy); w3 = (1 - ox) * oy; w4 = ox * oy
        a = [(w1 * gc[k][0] + w2 * gc[k][1] + w3 * gc[k][2] + w4 * gc[k][3]) * qm for k in range(6)]
        pE, hB = a[:3], a[3:]
        U0 = [c * u + pE[0], c * v + pE[1], c * w + pE[2]]
        g = 1 / torch.sqrt(c * c + U0[0] ** 2 + U0[1] ** 2 + U0[2] ** 2)
        pB = [g * h for h in hB]
        f = 2 / (1 + pB[0] ** 2 + pB[1] ** 2 + pB[2] ** 2)
        V = [U0[0] + U0[1] * pB[2] - U0[2] * pB[1], U0[1] + U0[2] * pB[0] - U0[0] * pB[2],
             U0[2] + U0[0] * pB[1] - U0[1] * pB[0]]
        U1 = [x * f for x in V]
        U2 = [U0[0] + U1[1] * pB[2] - U1[2] * pB[1] + pE[0], U0[1] + U1[2] * pB[0] - U1[0] * pB[2] + pE[1],
              U0[2] + U1[0] * pB[1] - U1[1] * pB[0] + pE[2]]
        un = [x / c for x in U2]
        G = c * c / torch.sqrt(c * c + U2[0] ** 2 + U2[1] ** 2 + U2[2] ** 2)
        xs = [ox, oy, oz]
        x1 = [xs[k] + un[k] * G for k in range(3)]
        d = [torch.floor(x.detach()).long() for x in x1]
        xr = []
        for k in range(3):
            mid = 0.5 * (xs[k] + x1[k])
            df = d[k].double()
            xr.append(torch.where(d[k] == 0, mid, torch.where(d[k] > 0, df, df + 1)))
        F1 = [qs * (xr[k] - xs[k]) for k in range(3)]
        F2 = [qs * (x1[k] - xr[k]) for k in range(3)]
        W1 = [0.5 * (ox + xr[0]), 0.5 * (oy + xr[1])]
        W2 = [0.5 * (x1[0] + xr[0]) - d[0], 0.5 * (x1[1] + xr[1]) - d[1]]
        J = [torch.zeros(self.mx, self.my, dtype=torch.float64) for _ in range(3)]

        def dep(J, F, W, ii, jj):
            zl, zh = F[2] * (1 - W[0]), F[2] * W[0]
            J = [x.index_put((ii, jj), y, accumulate=True) for x, y in ((J[0], F[0] * (1 - W[1])),
                                                                          (J[1], F[1] * (1 - W[0])),
                                                                          (J[2], zl * (1 - W[1])))]
            J[0] = J[0].index_put((ii, jj + 1), F[0] * W[1], accumulate=True)
            J[1] = J[1].index_put((ii + 1, jj), F[1] * W[0], accumulate=True)
            J[2] = J[2].index_put((ii + 1, jj), zh * (1 - W[1]), accumulate=True)
            J[2] = J[2].index_put((ii, jj + 1), zl * W[1], accumulate=True)
            J[2] = J[2].index_put((ii + 1, jj + 1), zh * W[1], accumulate=True)
            return J

        J = dep(J, F1, W1, i, j)
        J = dep(J, F2, W2, i + d[0], j + d[1])
        nst = ((ci + d[0]) % self.nx, (cj + d[1]) % self.ny, sp, x1[0] - d[0], x1[1] - d[1], x1[2] - d[2],
               un[0], un[1], un[2])
        return nst, J

    def step(self, st, E, B):
        B = self.b_half(E, B)
        N = self.nodes(E, B)
        st, J = self.push(st, N)
        B = self.b_half(E, B)
        J = [self.fold(x) for x in J]
        E = self.e_update(E, B, J)
        return st, E, B

    def particles(self, P):
        """cells, offsets and species of the reference SoA P (x, y, z, flv)."""
        fx, fy, fz = np.floor(P["x"]), np.floor(P["y"]), np.floor(P["z"])
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
        return (t(fx.astype(np.int64) - 3), t(fy.astype(np.int64) - 3), t((P["flv"] == 2).astype(np.int64)),
                t(P["x"] - fx), t(P["y"] - fy), t(P["z"] - fz))

    def objective(self, P, u, v, w, nsteps):
        ci, cj, sp, ox, oy, oz = self.particles(P)
        E = [torch.zeros(self.mx, self.my, dtype=torch.float64) for _ in range(3)]
        B = [torch.zeros(self.mx, self.my, dtype=torch.float64) for _ in range(3)]
        st = (ci, cj, sp, ox, oy, oz, u, v, w)
        for _ in range(nsteps):
            st, E, B = self.step(st, E, B)
        return 0.5 * sum((b[self.P] ** 2).sum() for b in B)

    def gradient(self, P, nsteps):
        u, v, w = [torch.from_numpy(np.array(P[k], dtype=np.float64)).requires_grad_(True) for k in "uvw"]
        L = self.objective(P, u, v, w, nsteps)
        L.backward()
        return float(L), [u.grad.numpy(), v.grad.numpy(), w.grad.numpy()]
