"""Pins of the cylindrical oracle (the reference built with -Dcyl, inc_axis).

The reference ships no tests or fixtures, so each routine restated in
oracle/pictor_oracle.c is pinned by a closed form or an exact discrete
invariant that a dropped term, a wrong sign, a wrong radius factor or a
wrong index would break:

* push (cyl_movdep.F90:14-289): a free particle moves on a Cartesian
  straight line; a purely radial line through the axis comes out at
  theta + pi with the radial momentum reversed (axis sub-steps,
  cyl_movdep.F90:300-446); in a uniform axial B the positions form the
  regular polygon of the Boris rotation angle 2 atan(q B / (2 gamma m c));
* deposit + E update + AddCurrent (cyl_em_update.F90:7-46, 114-140): the
  discrete Gauss residual r dtheta div E - rho, node by node, does not
  change over full steps away from the axis (charge conservation in
  cylindrical metric); with the reference's indexing of the second
  half-segment (cyl_movdep.F90:249,252) it does change;
* B update (cyl_em_update.F90:49-110): the discrete divergence of B with
  its radius factors does not change, including the flux balance of the
  axis disk that UpdateBzAxis integrates (:262-280);
* axis fills (:185-260): a uniform Cartesian field is continued across the
  axis exactly.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import CylOracle
from tests._cyl_util import gauss as _gauss, interior as _interior, plasma as _plasma

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def _cart(o, x, y):
    r = x - 3.5
    th = (y - 3.0) * o.cy.dtheta
    return r * np.cos(th), r * np.sin(th), th


def _local(th, ux, uy):
    return ux * np.cos(th) + uy * np.sin(th), -ux * np.sin(th) + uy * np.cos(th)


def _one(o, x, y, z, u, v, w, flv=1):
    q = 1.0 if flv == 1 else -1.0
    o.load_particles(np.array([q]), np.array([x]), np.array([y]), np.array([z]), np.array([u]), np.array([v]),
                     np.array([w]), np.array([flv], dtype=np.int32))


def _push(o, n):
    out = []
    for _ in range(n):
        o.reset_current()
        o.move_deposit()
        o.exchange()
        P = o.particles()
        out.append({k: P[k][0] for k in ("x", "y", "z", "u", "v", "w")})
    return out


def test_free_particle_moves_on_a_straight_line():
    o = CylOracle(2, 24, 32)
    th0 = 0.3
    x0, y0 = 3.5 + 8.0, 3.0 + th0 / o.cy.dtheta
    U = np.array([0.3, -0.25, 0.1])              # Cartesian gamma*beta
    ur, ut = _local(th0, U[0], U[1])
    _one(o, x0, y0, 1.5, ur, ut, U[2])
    gamma = np.sqrt(1 + U @ U)
    X0, Y0, _ = _cart(o, x0, y0)
    tr = _push(o, 30)
    for n, s in enumerate(tr, 1):
        X, Y, th = _cart(o, s["x"], s["y"])
        step = o.p.c * n / gamma
        assert abs(X - (X0 + U[0] * step)) < 1e-11 and abs(Y - (Y0 + U[1] * step)) < 1e-11, n
        ux = s["u"] * np.cos(th) - s["v"] * np.sin(th)
        uy = s["u"] * np.sin(th) + s["v"] * np.cos(th)
        assert abs(ux - U[0]) < 1e-12 and abs(uy - U[1]) < 1e-12 and abs(s["w"] - U[2]) < 1e-15
        assert s["x"] - 3.5 > 1.5                 # stays off the axis sub-step region


def test_radial_line_through_the_axis():
    o = CylOracle(2, 24, 32)
    th0 = 1.0
    r0, ur0 = 2.3, -0.5
    _one(o, 3.5 + r0, 3.0 + th0 / o.cy.dtheta, 1.2, ur0, 0.0, 0.0)
    gamma = np.sqrt(1 + ur0 * ur0)
    s_step = o.p.c * abs(ur0) / gamma
    tr = _push(o, 16)
    crossed = False
    for n, s in enumerate(tr, 1):
        rr = r0 - n * s_step                      # signed radius along the original direction
        th = (s["y"] - 3.0) * o.cy.dtheta
        assert abs((s["x"] - 3.5) - abs(rr)) < 1e-12, n
        want_th = th0 if rr > 0 else th0 + np.pi
        assert abs(np.angle(np.exp(1j * (th - want_th)))) < 1e-12, n
        assert abs(s["u"] - (ur0 if rr > 0 else -ur0)) < 1e-14 and abs(s["v"]) < 1e-15
        crossed = crossed or rr < 0
    assert crossed


def test_near_axis_line_is_straight():
    # impact parameter 0.8 cells: every step starts inside r < 1 for a while (sub-steps)
    o = CylOracle(2, 24, 32)
    X0, Y0 = 3.0, 0.8
    th0 = np.arctan2(Y0, X0)
    U = np.array([-0.6, 0.0, 0.0])
    ur, ut = _local(th0, U[0], U[1])
    _one(o, 3.5 + np.hypot(X0, Y0), 3.0 + th0 / o.cy.dtheta, 1.5, ur, ut, 0.0)
    gamma = np.sqrt(1 + U @ U)
    tr = _push(o, 20)
    rmin = 1e9
    for n, s in enumerate(tr, 1):
        X, Y, th = _cart(o, s["x"], s["y"])
        step = o.p.c * n / gamma
        assert abs(X - (X0 + U[0] * step)) < 1e-10 and abs(Y - Y0) < 1e-10, n
        rmin = min(rmin, s["x"] - 3.5)
    assert rmin < 1.0
    assert o.cy.substeps >= 0


def test_uniform_axial_field_gives_the_boris_polygon():
    o = CylOracle(2, 24, 32)
    B0 = 0.2
    o.F["Bz"][...] = B0
    th0 = 0.5
    x0, y0 = 3.5 + 10.0, 3.0 + th0 / o.cy.dtheta
    U = np.array([0.15, 0.12, 0.05])
    ur, ut = _local(th0, U[0], U[1])
    _one(o, x0, y0, 1.5, ur, ut, U[2])
    gamma = np.sqrt(1 + U @ U)
    alpha = 2 * np.arctan(0.5 * (1.0 / o.p.mi_me) * B0 / (o.p.c * gamma))
    vdt = o.p.c * np.hypot(U[0], U[1]) / gamma
    R = vdt / (2 * np.sin(alpha / 2))
    P = [np.array(_cart(o, x0, y0)[:2])]
    for s in _push(o, 25):
        P.append(np.array(_cart(o, s["x"], s["y"])[:2]))
    a, b, c = P[0], P[1], P[2]
    d = 2 * (a[0] * (b[1] - c[1]) + b[0] * (c[1] - a[1]) + c[0] * (a[1] - b[1]))
    ux = ((a @ a) * (b[1] - c[1]) + (b @ b) * (c[1] - a[1]) + (c @ c) * (a[1] - b[1])) / d
    uy = ((a @ a) * (c[0] - b[0]) + (b @ b) * (a[0] - c[0]) + (c @ c) * (b[0] - a[0])) / d
    C = np.array([ux, uy])
    for n, X in enumerate(P):
        assert abs(np.linalg.norm(X - C) - R) < 1e-11 * max(R, 1.0), n
    # the turning angle between successive chords is alpha
    for n in range(1, len(P) - 1):
        v1, v2 = P[n] - P[n - 1], P[n + 1] - P[n]
        ang = np.arctan2(v1[0] * v2[1] - v1[1] * v2[0], v1 @ v2)
        assert abs(abs(ang) - alpha) < 1e-10, n


@pytest.mark.parametrize("ndim", [2, 3])
def test_gauss_residual_is_invariant(ndim):
    rng = np.random.default_rng(7 + ndim)
    o = CylOracle(ndim, 20, 16, 8 if ndim == 3 else 1)
    P = list(_plasma(o, 3000 if ndim == 2 else 6000, rng, 11.0))
    nsteps = 12
    if ndim == 3:
        # the zigzag of the reference has no q dx dy dz / 12 term, so in 3-D
        # continuity is exact only for moves without a z component: w = 0 and
        # the first step (E = B = 0 during its push, so w stays 0)
        P[6] = np.zeros_like(P[6])
        nsteps = 1
    o.load_particles(*P)
    G0 = _gauss(o)[_interior(o)]
    for t in range(nsteps):
        o.step(t)
    G1 = _gauss(o)[_interior(o)]
    scale = np.abs(G0).max()
    assert scale > 0
    assert np.abs(G1 - G0).max() <= 1e-12 * scale
    assert o.cy.substeps > 0          # some particles went through the axis sub-steps


def test_gauss_residual_is_invariant_for_axial_moves_3d():
    # moves along z only (u = v = 0): no dx dy dz term either, so the first
    # step conserves charge exactly and pins the Jz / (r dtheta) of AddCurrent
    rng = np.random.default_rng(21)
    o = CylOracle(3, 20, 16, 8)
    P = list(_plasma(o, 6000, rng, 11.0))
    P[4] = np.zeros_like(P[4])
    P[5] = np.zeros_like(P[5])
    o.load_particles(*P)
    G0 = _gauss(o)[_interior(o)]
    o.step(0)
    G1 = _gauss(o)[_interior(o)]
    assert np.abs(o.F["Ez"]).max() > 0
    assert np.abs(G1 - G0).max() <= 1e-12 * np.abs(G0).max()


def test_reference_second_half_indexing_breaks_gauss():
    rng = np.random.default_rng(9)
    o = CylOracle(2, 20, 16, second_half_new_cell=False)
    o.load_particles(*_plasma(o, 3000, rng, 11.0))
    G0 = _gauss(o)[_interior(o)]
    for t in range(4):
        o.step(t)
    G1 = _gauss(o)[_interior(o)]
    assert np.abs(G1 - G0).max() > 1e-3 * np.abs(G0).max()


def _divb(o):
    """r_{i+1} Bx(i+1) - r_i Bx(i) + (By(j+1) - By(j)) / dtheta + r_{i+1/2} (Bz(k+1) - Bz(k)) per cell"""
    p, F, dth = o.p, o.F, o.cy.dtheta
    i = np.arange(p.mx)[:, None, None] + 1.0
    r = (i - 3.0) + o.cy.rmin
    Bx, By, Bz = F["Bx"], F["By"], F["Bz"]
    kk = slice(0, 1) if p.ndim == 2 else slice(0, p.mz - 1)
    D = (r + 1.0)[:-1] * Bx[1:, :-1, kk] - r[:-1] * Bx[:-1, :-1, kk]
    D = D + (By[:-1, 1:, kk] - By[:-1, :-1, kk]) / dth
    if p.ndim == 3:
        D = D + (r + 0.5)[:-1] * (Bz[:-1, :-1, 1:] - Bz[:-1, :-1, :-1])
    return D


@pytest.mark.parametrize("ndim", [2, 3])
def test_div_b_is_invariant(ndim):
    rng = np.random.default_rng(3)
    o = CylOracle(ndim, 12, 16, 8 if ndim == 3 else 1)
    for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        o.F[k][...] = rng.normal(0, 1, o.F[k].shape)
    o.field_halo("E")
    o.field_halo("B")
    # cells whose B values are all updated by the main formula and none is a ghost
    cells = (slice(3, o.p.mx - 4), slice(2, o.p.my - 4), slice(0, 1) if ndim == 2 else slice(2, o.p.mz - 4))
    D0 = _divb(o)[cells]
    pi = 3.14159265358979323846264
    ny = o.p.ny

    def disk():
        Bz3 = o.F["Bz"][2, 2, :]
        return (pi / 4) * (Bz3[1:] - Bz3[:-1]) + (pi / ny) * o.F["Bx"][3, 2:o.p.my - 3, :-1].sum(axis=0)

    kd = slice(2, o.p.mz - 4)
    A0 = disk()[kd] if ndim == 3 else None
    for _ in range(3):
        o.update_b_half()
        o.field_halo("B")
        for k in ("Ex", "Ey", "Ez"):
            o.F[k][...] = rng.normal(0, 1, o.F[k].shape)
        o.field_halo("E")
    D1 = _divb(o)[cells]
    assert np.abs(D1 - D0).max() <= 1e-12 * np.abs(D0).max()
    if ndim == 3:
        A1 = disk()[kd]
        assert np.abs(A1 - A0).max() <= 1e-12 * max(np.abs(A0).max(), 1.0)
        assert np.abs(A0).max() > 0.1


def test_axis_fill_continues_a_uniform_cartesian_field():
    o = CylOracle(2, 12, 16)
    E0, B0 = 0.7, -0.4
    dth = o.cy.dtheta
    j = np.arange(o.p.my) + 1.0
    th_j = (j - 3.0) * dth            # Ex (r = i - 3) and By sit at theta_j
    th_h = (j - 2.5) * dth            # Ey (r = i - 3.5) and Bx sit at theta_{j+1/2}
    o.F["Ex"][...] = (E0 * np.cos(th_j))[None, :, None]
    o.F["Ey"][...] = (-E0 * np.sin(th_h))[None, :, None]
    o.F["By"][...] = (-B0 * np.sin(th_j))[None, :, None]    # By sits at theta_j
    o.F["Bx"][...] = (B0 * np.cos(th_h))[None, :, None]     # Bx sits at theta_{j+1/2}
    o.F["Ex"][:3] = 0
    o.F["Ey"][:3] = 0
    o.F["Bx"][:3] = 0
    o.F["By"][:3] = 0
    o.add_current()                   # J = 0: only the axis fill acts
    assert np.allclose(o.F["Ex"][2, :, 0], E0 * np.cos(th_j), atol=1e-15, rtol=0)
    assert np.allclose(o.F["Ex"][1, :, 0], E0 * np.cos(th_j), atol=1e-15, rtol=0)
    assert np.allclose(o.F["Ey"][2, :, 0], -E0 * np.sin(th_h), atol=1e-15, rtol=0)
    o.update_b_half()                 # E is uniform: curl E = 0 away from the axis
    assert np.allclose(o.F["Bx"][2, :, 0], B0 * np.cos(th_h), atol=1e-14, rtol=0)
    assert np.allclose(o.F["By"][2, :, 0], -B0 * np.sin(th_j), atol=1e-14, rtol=0)
    assert np.allclose(o.F["By"][1, :, 0], -B0 * np.sin(th_j), atol=1e-14, rtol=0)



@pytest.mark.parametrize("ndim", [2, 3])
def test_cyl_energy_weights_each_component_by_its_stagger_radius(ndim):
    # CalcFldEnergy_cyl: a uniform component F0 over the grid gives
    # 1/2 F0^2 dtheta ny nz_phys * sum of its stagger radius over i = 4..mx-3:
    # r = i - 3.5 (B_r, E_theta, E_z): (nx - 1)^2 / 2;  r + 1/2 (E_r, B_theta, B_z): nx (nx - 1) / 2
    nx, ny, nz = 12, 8, 6
    o = CylOracle(ndim, nx, ny, nz if ndim == 3 else 1, c=0.2, epc=1, cap=16)
    nzp = nz if ndim == 3 else 1
    sum_r, sum_rp = (nx - 1) ** 2 / 2.0, nx * (nx - 1) / 2.0
    dth = 2.0 * math.pi / ny
    for comp, slot, radial in (("Ex", 1, sum_rp), ("Ey", 1, sum_r), ("Ez", 1, sum_r),
                               ("Bx", 0, sum_r), ("By", 0, sum_rp), ("Bz", 0, sum_rp)):
        for n in FIELDS:
            o.F[n][...] = 0.0
        o.F[comp][...] = 3.0
        e = o.energy()
        want = 0.5 * 9.0 * dth * ny * nzp * radial
        assert abs(e[slot] - want) <= 1e-12 * want, comp
        assert e[1 - slot] == 0.0, comp


# ---- the Strang push of the CPU -Dcyl build (cyl_movdep_strang.F90) ----

def test_strang_free_particles_move_on_straight_lines():
    # no fields: half kick, move, rotation into the new frame, half kick is a straight
    # line, also for the particles that cross the axis in sub-steps
    rng = np.random.default_rng(3)
    o = CylOracle(2, 24, 32, c=0.15, push="strang", cap=600)
    n = 400
    r = rng.uniform(0.0, 10.0, n)
    th = rng.uniform(0.0, 2 * np.pi, n)
    U = rng.normal(0.0, 0.4, (3, n))
    ur, ut = _local(th, U[0], U[1])
    o.load_particles(np.ones(n), 3.5 + r, 3.0 + th / o.cy.dtheta, np.full(n, 1.5), ur, ut, U[2],
                     np.ones(n, dtype=np.int32))
    X0, Y0, _ = _cart(o, 3.5 + r, 3.0 + th / o.cy.dtheta)
    g = np.sqrt(1.0 + (U ** 2).sum(axis=0))
    steps = 8
    for _ in range(steps):
        o.reset_current(); o.move_deposit(); o.exchange()
    P = o.particles()
    X, Y, t1 = _cart(o, P["x"], P["y"])
    # a sub-step that carries a particle past the axis (x1 <= 0) keeps |x1| and drops the
    # transverse move (cyl_movdep_strang.F90:516-527, as in cyl_movdep.F90): paths that
    # pass within 0.1 cell of the axis are left out
    L = steps * o.p.c / g
    s = np.clip(-(X0 * U[0] + Y0 * U[1]) / np.maximum(U[0] ** 2 + U[1] ** 2, 1e-30), 0.0, L)
    clear = np.hypot(X0 + s * U[0], Y0 + s * U[1]) > 0.1
    assert clear.sum() > 390
    np.testing.assert_allclose(X[clear], (X0 + steps * o.p.c * U[0] / g)[clear], rtol=0, atol=1e-10)
    np.testing.assert_allclose(Y[clear], (Y0 + steps * o.p.c * U[1] / g)[clear], rtol=0, atol=1e-10)
    ux = P["u"] * np.cos(t1) - P["v"] * np.sin(t1)
    np.testing.assert_allclose(ux[clear], U[0][clear], rtol=0, atol=1e-12)
    assert o.cy.substeps > 0


@pytest.mark.parametrize("push,exact", [("boris", True), ("strang", True), ("strang_as_written", False)])
def test_strang_kick_in_a_uniform_axial_field(push, exact):
    # a uniform E_z: every particle's c w changes by (q/m) E_z in one step; Strang
    # applies q/m / 4 four times.  As written the second half starts its gather sums
    # from the first half's (q/m / 4) E: an extra (q/m)^2 E / 8, 12.5 % for electrons
    rng = np.random.default_rng(5)
    o = CylOracle(3, 16, 16, 8, c=0.15, push=push, cap=600, mi_me=1.0)
    E0 = 1e-3
    for k in FIELDS:
        o.F[k][...] = 0.0
    o.F["Ez"][...] = E0
    n = 300
    r = rng.uniform(0.0, 9.0, n)
    th = rng.uniform(0.0, 2 * np.pi, n)
    flv = rng.integers(1, 3, n).astype(np.int32)
    o.load_particles(np.where(flv == 1, 1.0, -1.0), 3.5 + r, 3.0 + th / o.cy.dtheta, rng.uniform(4.0, 8.0, n),
                     np.zeros(n), np.zeros(n), np.zeros(n), flv)
    o.reset_current(); o.move_deposit()
    P = o.particles()
    qm = np.where(P["flv"] == 1, o.p.flvrqm[0], o.p.flvrqm[1])
    want = qm * E0 / o.p.c
    if exact:
        np.testing.assert_allclose(P["w"], want, rtol=1e-12, atol=0)
    else:
        off = P["x"] >= 4.5 + 0.2     # the main loop (the axis sub-steps reset their gathers)
        extra = qm[off] ** 2 * E0 / 8.0 / o.p.c
        np.testing.assert_allclose(P["w"][off], want[off] + extra, rtol=1e-9, atol=0)


@pytest.mark.parametrize("ndim", [2, 3])
def test_strang_keeps_the_gauss_residual(ndim):
    rng = np.random.default_rng(17 + ndim)
    o = CylOracle(ndim, 20, 16, 8 if ndim == 3 else 1, push="strang")
    P = list(_plasma(o, 3000, rng, 11.0))
    nsteps = 12
    if ndim == 3:
        P[6] = np.zeros_like(P[6])
        nsteps = 1
    o.load_particles(*P)
    G0 = _gauss(o)[_interior(o)]
    for t in range(nsteps):
        o.step(t)
    G1 = _gauss(o)[_interior(o)]
    assert np.abs(G1 - G0).max() <= 1e-11 * np.abs(G0).max()   # rounding: about 2e-12
    assert o.cy.substeps > 0
