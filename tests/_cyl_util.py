"""Helpers of the cylindrical tests: a random plasma filling a disk of the
(r, theta) grid and the node-by-node discrete Gauss residual

    G(i, j, k) = r_i dtheta div E - qi rho,
    r_i dtheta div E = dtheta (r_{i+1/2} Ex(i) - r_{i-1/2} Ex(i-1)) + Ey(j) - Ey(j-1)
                       + r_i dtheta (Ez(k) - Ez(k-1)),

with rho the CIC charge of the particles at the nodes in index space.  The
cylindrical E update and AddCurrent (cyl_em_update.F90:7-46, 114-140) keep
G constant wherever the deposit conserves charge.  `o` is a CylOracle whose
fields and particles hold the state to evaluate (from the oracle or a GPU
download).
"""
import numpy as np


def plasma(o, n, rng, rmax):
    r = np.sqrt(rng.uniform(0.0, rmax * rmax, n))
    th = rng.uniform(0.0, 2 * np.pi, n)
    x = 3.5 + r
    y = 3.0 + th / o.cy.dtheta
    z = rng.uniform(1.0, 2.0, n) if o.p.ndim == 2 else rng.uniform(3.0, o.p.mz - 2.0, n)
    u, v, w = rng.normal(0.0, 0.25, (3, n))
    flv = np.where(np.arange(n) % 2 == 0, 1, 2).astype(np.int32)
    q = np.where(flv == 1, 1.0, -1.0)
    return q, x, y, z, u, v, w, flv


def rho(o):
    """qi * sum q * CIC weight at the nodes (index space), periodic images folded."""
    p = o.p
    P = o.particles()
    rho = np.zeros((p.mx + 1, p.my + 1, p.mz + 1))
    i0 = np.floor(P["x"]).astype(int)
    j0 = np.floor(P["y"]).astype(int)
    dx, dy = P["x"] - i0, P["y"] - j0
    if p.ndim == 2:
        k0, dz, kw = np.ones_like(i0), np.zeros_like(dx), ((0, 1.0),)
    else:
        k0 = np.floor(P["z"]).astype(int)
        dz = P["z"] - k0
        kw = ((0, None), (1, None))
    for a in (0, 1):
        wa = (1 - dx) if a == 0 else dx
        for b in (0, 1):
            wb = (1 - dy) if b == 0 else dy
            for c, _ in kw:
                wc = 1.0 if p.ndim == 2 else ((1 - dz) if c == 0 else dz)
                np.add.at(rho, (i0 + a, j0 + b, k0 + c), P["q"] * wa * wb * wc)
    rho = rho[1:, 1:, 1:]        # 1-based -> 0-based (i-1)
    rho = rho[:p.mx, :p.my, :p.mz]
    # fold the periodic node images: theta node ny + 3 is node 3; z likewise (3-D)
    rho[:, 2, :] += rho[:, p.my - 3, :]
    rho[:, p.my - 3, :] = 0
    if p.ndim == 3:
        rho[:, :, 2] += rho[:, :, p.mz - 3]
        rho[:, :, p.mz - 3] = 0
    return p.qi * rho


def gauss(o):
    p, F, dth = o.p, o.F, o.cy.dtheta
    i = np.arange(p.mx)[:, None, None] + 1.0
    r = (i - 3.0) + o.cy.rmin
    G = np.zeros((p.mx, p.my, p.mz))
    Ex, Ey, Ez = F["Ex"], F["Ey"], F["Ez"]
    sl = (slice(1, None), slice(1, None), slice(None))
    G[sl] = (dth * ((r + 0.5)[1:] * Ex[1:, 1:, :] - (r - 0.5)[1:] * Ex[:-1, 1:, :]) + (Ey[1:, 1:, :] - Ey[1:, :-1, :]))
    if p.ndim == 3:
        G[1:, 1:, 1:] += (r[1:] * dth) * (Ez[1:, 1:, 1:] - Ez[1:, 1:, :-1])
    return G - rho(o)


def interior(o):
    p = o.p
    ks = slice(0, 1) if p.ndim == 2 else slice(2, p.mz - 3)
    return (slice(4, p.mx - 5), slice(2, p.my - 3), ks)   # i = 5 .. mx-5, j = 3 .. my-3
