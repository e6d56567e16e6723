"""Pins of the azimuthal-mode (m = 0, 1) cylindrical oracle (no GPU).

The reference has no modal solver, so the modal form is pinned against the
reference-derived full r-theta-z oracle and against what geometry fixes:
  * m = 0: theta-independent fields evolve exactly as on the full grid (the
    full grid's theta differences vanish; its AddCurrent per theta cell is the
    modal current over the circle), to rounding;
  * m = 1: a field sampled as F_c cos(theta) + F_s sin(theta) on the full grid
    (each component at its own theta stagger) follows the modal update with
    an error that falls as dtheta^2 (second order: the full grid's theta
    differences are centred);
  * a uniform Cartesian E_x (m = 1 slots) gives every particle, at any radius
    including the cells next to the axis, the momentum kick 2 (q/m) E_x / c
    along x exactly: gather, the axis parity fills, Boris and the local-frame
    rotation together;
  * without fields particles move on Cartesian straight lines, also across the
    axis;
  * the discrete Gauss law (flux of E out of every node cell minus the node
    charge, per mode) stays constant over a run with particles: deposit,
    AddCurrent and the modal divergence agree, including the angular current
    of m = 1.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import CylModalOracle, CylOracle

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def _full_and_modal(nr, nth, nz, c=0.3):
    full = CylOracle(3, nr, nth, nz, c=c, epc=1, cap=16)
    modal = CylModalOracle(nr, nz, dr=1.0, c=c, epc=1, cap=16)
    return full, modal


def _phys(a):
    # i = 4..mx-3, physical z
    return a[3:-3, ..., 2:-3]


def test_m0_matches_full_grid():
    full, modal = _full_and_modal(12, 8, 10)
    rng = np.random.default_rng(1)
    mx, mz = modal.shape[0], modal.shape[2]
    for n in FIELDS:
        f0 = rng.normal(size=(mx, mz))
        full.F[n][...] = f0[:, None, :]
        modal.F[n][...] = 0.0
        modal.F[n][:, 0, :] = f0
    # a theta-independent current: the modal m = 0 current is the full grid's summed
    # over the theta cells; the full grid's Jy counts theta cells (dy = dtheta / dtheta_cell),
    # the modal one radians, so it carries one more factor dtheta
    for n in ("Jx", "Jy", "Jz"):
        j0 = rng.normal(size=(mx, mz))
        full.F[n][...] = j0[:, None, :]
        modal.F[n][...] = 0.0
        modal.F[n][:, 0, :] = full.p.ny * j0 * (full.cy.dtheta if n == "Jy" else 1.0)
    for _ in range(3):
        full.update_b_half(); full.field_halo("B")
        full.update_e(); full.field_halo("E")
        full.add_current(); full.field_halo("E")
        modal.b_half()
        modal.e_phase(curl=True, add=False)
        modal.e_phase(curl=False, add=True)
    for n in FIELDS:
        a = _phys(modal.F[n][:, 0, :][:, None, :])
        for j in range(2, full.p.my - 3):
            b = _phys(full.F[n][:, j:j + 1, :])
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 1e-13 * scale, (n, j)
        assert np.abs(modal.F[n][:, 1:, :]).max() == 0.0, n


def _sample(full, comp, fc, fs):
    # full-grid sample of fc cos(theta) + fs sin(theta) at the component's theta
    # stagger: E_r, E_z, B_theta at theta_j = (j - 3) dtheta; E_theta, B_r, B_z at theta_{j + 1/2}
    dth = full.cy.dtheta
    half = comp in ("Ey", "Bx", "Bz")
    j = np.arange(full.p.my) + 1.0
    th = (j - 3.0 + (0.5 if half else 0.0)) * dth
    return fc[:, None, :] * np.cos(th)[None, :, None] + fs[:, None, :] * np.sin(th)[None, :, None], th


def _m1_update_error(nth, seed=3):
    nr, nz = 10, 8
    full, modal = _full_and_modal(nr, nth, nz, c=0.05)
    rng = np.random.default_rng(seed)
    mx, mz = modal.shape[0], modal.shape[2]
    # smooth radial and axial profiles, so that only the theta discretisation differs
    i = np.arange(mx)[:, None] + 1.0
    k = np.arange(mz)[None, :] + 1.0
    coef = {}
    for n in FIELDS:
        a, b, ph = rng.normal(size=3)
        fc = np.cos(0.3 * a * i + 0.2 * k + ph) * np.ones((mx, mz))
        fs = np.sin(0.25 * b * i - 0.3 * k + ph) * np.ones((mx, mz))
        coef[n] = (fc, fs)
        full.F[n][...] = _sample(full, n, fc, fs)[0]
        modal.F[n][...] = 0.0
        modal.F[n][:, 1, :] = fc
        modal.F[n][:, 2, :] = fs
    before = {n: full.F[n].copy() for n in FIELDS}
    full.update_b_half()
    full.update_e()
    modal.b_half()
    modal.e_phase(curl=True, add=False)
    err = 0.0
    for n in FIELDS:
        mod, _ = _sample(full, n, modal.F[n][:, 1, :], modal.F[n][:, 2, :])
        d_full = full.F[n] - before[n]
        d_mod = mod - before[n]
        sl = (slice(4, -4), slice(2, -3), slice(3, -4))   # away from the axis fills, wall and z ghosts
        scale = np.abs(d_mod[sl]).max()
        assert scale > 0, n
        err = max(err, np.abs(d_full[sl] - d_mod[sl]).max() / scale)
    return err


def test_m1_converges_to_full_grid_at_second_order():
    e16, e32, e64 = _m1_update_error(16), _m1_update_error(32), _m1_update_error(64)
    assert e16 < 0.05
    assert 3.3 < e16 / e32 < 4.8, (e16, e32)
    assert 3.3 < e32 / e64 < 4.8, (e32, e64)


def _cart(P, dr):
    r = (P["x"] - 3.5) * dr
    return r * np.cos(P["y"]), r * np.sin(P["y"])


def _cart_u(P):
    c, s = np.cos(P["y"]), np.sin(P["y"])
    return P["u"] * c - P["v"] * s, P["u"] * s + P["v"] * c


@pytest.mark.parametrize("dr", [1.0, 3.0])
def test_uniform_cartesian_ex_kicks_every_particle_along_x(dr):
    o = CylModalOracle(16, 12, dr=dr, c=0.4, c_ompe=10.0, epc=1, cap=2048)
    E0 = 1e-3
    o.F["Ex"][:, 1, :] = E0          # E_r = E0 cos(theta)
    o.F["Ey"][:, 2, :] = -E0         # E_theta = -E0 sin(theta)
    o.e_phase(curl=False, add=False)  # axis parity fills, halos, wall
    rng = np.random.default_rng(7)
    n = 500
    r = np.concatenate([rng.uniform(0.02, 1.5, n // 2) * dr, rng.uniform(0.0, 10.0, n - n // 2) * dr])
    q = -np.ones(n)
    x, th, z = 3.5 + r / dr, rng.uniform(0, 2 * np.pi, n), rng.uniform(3.0, 14.0, n)
    o.load_particles(q, x, th, z, np.zeros(n), np.zeros(n), np.zeros(n), np.full(n, 2, dtype=np.int32))
    o.move_deposit()
    P = o.particles()
    ux, uy = _cart_u(P)
    kick = 2.0 * (-1.0 * 0.5) * E0 / o.m.c     # 2 (q/m)/2 E / c for electrons
    np.testing.assert_allclose(ux, kick, rtol=1e-13, atol=0)
    assert np.abs(uy).max() <= 1e-13 * abs(kick)


@pytest.mark.parametrize("dr", [1.0, 2.5])
def test_free_particles_move_on_straight_lines(dr):
    o = CylModalOracle(20, 40, dr=dr, c=0.45, epc=1, cap=1024)
    rng = np.random.default_rng(11)
    n = 300
    r = rng.uniform(0.0, 4.0, n) * dr
    th = rng.uniform(0, 2 * np.pi, n)
    u, v, w = rng.normal(0, 0.6, (3, n))
    o.load_particles(-np.ones(n), 3.5 + r / dr, th, rng.uniform(10.0, 30.0, n), u, v, w,
                     np.full(n, 2, dtype=np.int32))
    P0 = o.particles()
    X0, Y0 = _cart(P0, dr)
    ux, uy = _cart_u(P0)
    g = np.sqrt(1.0 + ux ** 2 + uy ** 2 + P0["w"] ** 2)
    steps = 12
    for _ in range(steps):
        o.move_deposit()
    P = o.particles()
    X, Y = _cart(P, dr)
    # displacement per step c u / gamma, in axial cells
    np.testing.assert_allclose(X, X0 + steps * o.m.c * ux / g, rtol=0, atol=1e-11 * dr)
    np.testing.assert_allclose(Y, Y0 + steps * o.m.c * uy / g, rtol=0, atol=1e-11 * dr)
    np.testing.assert_allclose(P["z"], P0["z"] + steps * o.m.c * P0["w"] / g, rtol=0, atol=1e-11)


@pytest.mark.parametrize("dr", [1.0, 2.0])
def test_gauss_law_conserved_per_mode(dr):
    o = CylModalOracle(24, 32, dr=dr, c=0.4, c_ompe=10.0, epc=4)
    P = o.load_plasma(16, 3, 32, 8, seed=3)
    rng = np.random.default_rng(1)
    n = len(P["x"])
    for k in "uvw":
        o.P[k][:n] = rng.normal(0, 0.1, n)
    G0 = o.gauss()
    o.run(40)
    G1 = o.gauss()
    assert o.energy().sum() > 0
    for s in range(3):
        scale = np.abs(G0[:, s, :]).max()
        assert scale > 0
        assert np.abs(G1 - G0)[:, s, :].max() <= 1e-13 * max(scale, 1.0), s
