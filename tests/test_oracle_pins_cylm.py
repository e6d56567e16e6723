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


# ---- the Lehe B update per mode ----
# The modal Lehe keeps the reference's z averaging of the r and theta differences
# (beta = 1/8) and its wide z pair d_z, and leaves out the r pair (b_zx) and the theta
# pair (b_zy) of the z difference.  The reference's Lehe-cyl restated on the full grid
# with those two coefficients set to 0 is its reference.

def _lehe_full(nr, nth, nz, c):
    full = CylOracle(3, nr, nth, nz, c=c, epc=1, cap=16, fdtd="lehe")
    L = full.cy.lehe
    L[2] = L[2] + 2.0 * L[7] + 2.0 * L[8]
    L[7] = 0.0
    L[8] = 0.0
    return full


def test_lehe_m0_matches_full_grid():
    full = _lehe_full(12, 8, 12, 0.3)
    modal = CylModalOracle(12, 12, dr=1.0, c=0.3, epc=1, cap=16, fdtd=1)
    assert modal.m.l_zd != 0.0 and modal.m.l_rz == 0.125
    rng = np.random.default_rng(4)
    mx, mz = modal.shape[0], modal.shape[2]
    for n in FIELDS:
        f0 = rng.normal(size=(mx, mz))
        full.F[n][...] = f0[:, None, :]
        modal.F[n][...] = 0.0
        modal.F[n][:, 0, :] = f0
    for _ in range(3):
        full.update_b_half(); full.field_halo("B")
        full.update_e(); full.field_halo("E")
        modal.b_half()
        modal.e_phase(curl=True, add=False)
    for n in FIELDS:
        a = _phys(modal.F[n][:, 0, :][:, None, :])
        for j in range(2, full.p.my - 3):
            b = _phys(full.F[n][:, j:j + 1, :])
            assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), (n, j)
    # and the Lehe update differs from Yee's
    yee = CylModalOracle(12, 12, dr=1.0, c=0.3, epc=1, cap=16, fdtd=0)
    for n in FIELDS:
        yee.F[n][...] = 0.0
        yee.F[n][:, 0, :] = rng.normal(size=(mx, mz))
    lehe = CylModalOracle(12, 12, dr=1.0, c=0.3, epc=1, cap=16, fdtd=1)
    lehe.set_fields(**{n: yee.F[n].copy() for n in FIELDS})
    yee.b_half(); lehe.b_half()
    assert np.abs(_phys(yee.F["By"]) - _phys(lehe.F["By"])).max() > 1e-3


def _m1_lehe_error(nth, seed=3):
    nr, nz = 10, 8
    full = _lehe_full(nr, nth, nz, 0.05)
    modal = CylModalOracle(nr, nz, dr=1.0, c=0.05, epc=1, cap=16, fdtd=1)
    rng = np.random.default_rng(seed)
    mx, mz = modal.shape[0], modal.shape[2]
    i = np.arange(mx)[:, None] + 1.0
    k = np.arange(mz)[None, :] + 1.0
    for n in FIELDS:
        a, b, ph = rng.normal(size=3)
        fc = np.cos(0.3 * a * i + 0.2 * k + ph) * np.ones((mx, mz))
        fs = np.sin(0.25 * b * i - 0.3 * k + ph) * np.ones((mx, mz))
        full.F[n][...] = _sample(full, n, fc, fs)[0]
        modal.F[n][...] = 0.0
        modal.F[n][:, 1, :] = fc
        modal.F[n][:, 2, :] = fs
    before = {n: full.F[n].copy() for n in FIELDS}
    full.update_b_half()
    modal.b_half()
    err = 0.0
    for n in ("Bx", "By", "Bz"):
        mod, _ = _sample(full, n, modal.F[n][:, 1, :], modal.F[n][:, 2, :])
        sl = (slice(4, -4), slice(2, -3), slice(3, -4))
        d_full = (full.F[n] - before[n])[sl]
        d_mod = (mod - before[n])[sl]
        err = max(err, np.abs(d_full - d_mod).max() / np.abs(d_mod).max())
    return err


def test_lehe_m1_converges_to_full_grid_at_second_order():
    e16, e32, e64 = _m1_lehe_error(16), _m1_lehe_error(32), _m1_lehe_error(64)
    assert e16 < 0.05
    assert 3.3 < e16 / e32 < 4.8, (e16, e32)
    assert 3.3 < e32 / e64 < 4.8, (e32, e64)


def _vacuum_energy_halves(dr, c, slots, steps, restore_r_pair=False, seed=0):
    """largest E^2 + B^2 (over its start value) in the first and in the second half of a
    vacuum run from white-noise fields.  Near the Courant limit the start-up ratio is
    large (E^2 + B^2 at one time level is not the Lehe scheme's conserved norm); a
    bounded run keeps the second half at the first half's level."""
    o = CylModalOracle(12, 16, dr=dr, c=c, epc=1, cap=16, fdtd=1)
    if restore_r_pair:   # the reference's b_zx = (dz/dr)^2 / 8 with its share of the central term
        bzx = 0.125 / dr ** 2
        o.m.l_zr = bzx
        o.m.l_zz = o.m.l_zz - 2.0 * bzx
    rng = np.random.default_rng(seed)
    for n in FIELDS:
        for s in slots:
            o.F[n][4:-4, s, 3:-3] = rng.normal(size=(o.shape[0] - 8, o.shape[2] - 6))
    o.e_phase(curl=False, add=False)
    o.b_half(); o.b_half()
    e0 = o.energy().sum()
    halves = [0.0, 0.0]
    for t in range(steps):
        o.b_half(); o.b_half(); o.e_phase(curl=True, add=False)
        if t % 100 == 99:
            e = o.energy().sum()
            r = e / e0 if np.isfinite(e) else np.inf
            h = 0 if t < steps // 2 else 1
            halves[h] = max(halves[h], r)
    return halves


def test_reference_r_pair_makes_the_axis_unstable():
    # b_zx's r pair F(i+1) - 2 F(i) + F(i-1) reads the axis fills at i = 4
    a, b = _vacuum_energy_halves(1.0, 0.39, (0,), 2000, restore_r_pair=True)
    assert b > 1e4 * a
    a, b = _vacuum_energy_halves(1.0, 0.39, (0,), 2000)
    assert b <= 1.1 * a


def _courant_rule(dr):
    # sin^2(pi c / 2) + 1.25 c^2 / dr^2 < 1, and the axis limit of the Yee update
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if math.sin(0.5 * math.pi * mid) ** 2 + 1.25 * mid * mid / dr ** 2 < 1.0:
            lo = mid
        else:
            hi = mid
    return min(lo, 1.0 / math.sqrt(1.0 + 5.0 / dr ** 2))


@pytest.mark.parametrize("dr,c", [(4.0, 0.8), (4.0, None), (8.0, None), (2.0, None), (1.0, None)])
def test_modal_lehe_stays_bounded(dr, c):
    c = c if c is not None else 0.999 * _courant_rule(dr)
    for slots in ((0,), (1, 2)):
        a, b = _vacuum_energy_halves(dr, c, slots, 8000)
        assert np.isfinite(a) and b <= 1.1 * a, (dr, c, slots, a, b)


@pytest.mark.parametrize("zopen", [True, False])
def test_open_z_keeps_the_back_away_from_the_front(zopen):
    # the moving window's box has open z ends: fields in the back rows and electrons
    # leaving through the back must not reach the front.  Periodic z (zopen False) copies
    # the back rows into the front ghosts and wraps the electrons, which this test shows.
    nr, nz, c = 12, 32, 0.45
    o = CylModalOracle(nr, nz, dr=2.0, c=c, c_ompe=10.0, epc=4, cap=256)
    o.set_zopen(zopen)
    rng = np.random.default_rng(3)
    mx, _, mz = o.shape
    for n in FIELDS:
        a = np.zeros(o.shape, order="F")
        a[3:-3, :, 2:5] = 1e-2 * rng.normal(size=(mx - 6, 3, 3))   # Fortran rows 3..5
        o.F[n][...] = a
    n = 6
    o.load_particles(np.full(n, -1.0), rng.uniform(5.0, mx - 6.0, n), rng.uniform(0, 2 * np.pi, n),
                     rng.uniform(3.05, 3.3, n), np.zeros(n), np.zeros(n), np.full(n, -3.0),
                     np.full(n, 2, dtype=np.int32), cap=256)
    shifts = 0
    for t in range(1, 7):
        o.step()
        while t * c >= shifts + 1:
            shifts += 1
            o.shift(0, 0, 1, shifts)
    front = slice(mz - 8, mz)
    reached = max(np.abs(o.F[k][:, :, front]).max() for k in FIELDS)
    live = o.particles(live_only=True)
    if zopen:
        assert reached == 0.0
        assert len(live["x"]) == 0        # the electrons left through the back
    else:
        assert reached > 1e-6
        assert len(live["x"]) == n and np.all(live["z"] > mz // 2)   # wrapped to the front half
