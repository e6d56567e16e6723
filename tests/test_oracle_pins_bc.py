"""Pins of the oracle's boundary faces (no GPU).

The reference has no tests for src/bc, so each face type is pinned by what
the mathematics or the physics fixes:
  * non-periodic halos: the MPI face emulation of comm_fld.F90 with the
    faces of a non-periodic axis left out (null neighbours) gives the same
    arrays;
  * exchange: a particle leaving through a non-periodic face is gone, one
    leaving through a periodic face wraps;
  * reflecting faces: the specular image 2 x_w - x of the free flight with
    the normal momentum reversed and |u| kept; with a wall on a grid node
    the normal current folded back across the wall keeps div E - rho
    unchanged node by node, and the tangential fold keeps the summed
    tangential current;
  * removing faces: exactly the particles beyond the face become holes;
  * conducting faces: a vacuum pulse is reflected with its energy kept and
    nothing passes the wall; the as-written top face (bc.F90:318) wipes the
    field above the bottom face;
  * PML: the coefficients at depth d are a = exp(-sigma - alpha), b =
    sigma/(sigma + alpha) (a - 1), sigma = (d/L)^4, alpha = 9e-4, c = 0; a
    pulse entering the layer leaves less than 1e-5 of its energy, where a
    conducting face keeps it;
  * open faces: a pulse leaving through them keeps only a few percent;
  * moving faces advance by speed * c per step.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle

from tests.test_oracle_pins import one_particle


def _emulate_np(F, add, m, ndim, periodic):
    """The face emulation with the exchanges of non-periodic axes removed."""
    import tests.test_oracle_pins as base
    if all(periodic[: ndim]):
        return base._emulate_mpi_faces(F, add, m, ndim)
    mx, my, mz = m
    full = {0: (1, mx), 1: (1, my), 2: (1, mz)}

    def layers(axis, off, nl):
        mm = m[axis]
        o = off + mm if off < 0 else off
        return (o + 1, o + nl)

    def box(axis, rng):
        b = dict(full)
        if axis != 0:
            b[0] = (3, mx - 3)
        b[axis] = rng
        return b

    def sl(b):
        return tuple(slice(b[a][0] - 1, b[a][1]) for a in range(3))

    if ndim == 3:
        order = [0, 1, 2] if add else [1, 2, 0]
    else:
        order = [0, 1] if add else [1, 0]
    for axis in order:
        if not periodic[axis]:
            continue
        send = {1: box(axis, layers(axis, 2, 3)), 2: box(axis, layers(axis, -5, 2))}
        recv = {1: box(axis, layers(axis, 0, 2)), 2: box(axis, layers(axis, -3, 3))}
        if not add:
            bufs = {n: F[sl(send[n])].copy() for n in (1, 2)}
            F[sl(recv[1])] = bufs[2]
            F[sl(recv[2])] = bufs[1]
        else:
            bufs = {n: F[sl(recv[n])].copy() for n in (1, 2)}
            F[sl(send[1])] = F[sl(send[1])] + bufs[2]
            F[sl(send[2])] = F[sl(send[2])] + bufs[1]
    return F


@pytest.mark.parametrize("add", [False, True])
@pytest.mark.parametrize("ndim,periodic", [(2, (0, 1, 1)), (2, (1, 0, 1)), (2, (0, 0, 1)),
                                           (3, (0, 1, 1)), (3, (1, 1, 0)), (3, (1, 0, 0))])
def test_nonperiodic_halo_matches_face_emulation(add, ndim, periodic):
    o = Oracle(ndim, 10, 12, 8, epc=2)
    for a in range(3):
        o.p.periodic[a] = periodic[a]
    rng = np.random.default_rng(5)
    names = ("Jx", "Jy", "Jz") if add else ("Bx", "By", "Bz")
    A = [rng.standard_normal(o.shape) for _ in range(3)]
    for n, a in zip(names, A):
        o.F[n][...] = a
    if add:
        o.current_halo_add()
    else:
        o.field_halo("B")
    for n, a in zip(names, A):
        want = _emulate_np(np.asfortranarray(a.copy()), add, o.shape, ndim, periodic)
        assert np.array_equal(o.F[n], want), n
    # a periodic axis changes the arrays; with none left nothing moves at all
    assert np.array_equal(o.F[names[0]], A[0]) == (not any(periodic[:ndim]))


def test_exchange_drops_particles_leaving_a_nonperiodic_axis():
    o = Oracle(2, 16, 16, epc=2)
    o.p.periodic[0] = 0          # x faces lead nowhere, y wraps
    xlen, ylen = o.p.xlen, o.p.ylen
    xs = np.array([2.9, o.p.xmax + 0.2, 10.0, 10.0, 10.0])
    ys = np.array([10.0, 10.0, 2.7, o.p.ymax + 0.1, 10.0])
    n = len(xs)
    o.load_particles(np.ones(n), xs, ys, np.full(n, 1.5), np.zeros(n), np.zeros(n), np.zeros(n),
                     np.ones(n, np.int32))
    o.exchange()
    P = o.particles()
    got = sorted(zip(P["x"], P["y"]))
    want = sorted([(10.0, 2.7 + ylen), (10.0, o.p.ymax + 0.1 - ylen), (10.0, 10.0)])
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    assert xlen > 0


# ---------------------------------------------------------------- reflection
FACES = [("x", "lower"), ("x", "upper"), ("y", "lower"), ("y", "upper")]


@pytest.mark.parametrize("axis,side", FACES)
def test_reflection_is_specular_image_of_free_flight(axis, side):
    o = Oracle(2, 24, 24, epc=2)
    wall = 7.37 if side == "lower" else 16.61      # global position, not on a node
    o.set_bc_prtl(axis, wall, "refl", side)
    a = 0 if axis == "x" else 1
    wl = wall + 3.0
    rng = np.random.default_rng(3)
    for _ in range(20):
        u = rng.uniform(-2.0, 2.0, 3)
        u[a] = -abs(u[a]) if side == "lower" else abs(u[a])
        r0 = np.array([rng.uniform(12.0, 14.0), rng.uniform(12.0, 14.0), 1.5])
        gam = math.sqrt(1.0 + u @ u)
        step = u * o.p.c / gam
        r0[a] = wl - 0.6 * step[a]          # crosses the wall in this step
        one_particle(o, *r0, *u)
        o.reset_current()
        o.move_deposit()
        o.bc_post_movdep()
        P = o.particles()
        free = r0 + step
        got = np.array([P["x"][0], P["y"][0]])
        want = free[:2].copy()
        want[a] = 2.0 * wl - free[a]
        assert np.allclose(got, want, rtol=0, atol=1e-13)
        uu = np.array([P["u"][0], P["v"][0], P["w"][0]])
        assert abs(uu[a] + u[a]) < 1e-14 * max(1.0, abs(u[a]))
        assert abs(np.linalg.norm(uu) - np.linalg.norm(u)) < 1e-14 * np.linalg.norm(u)
        assert (got[a] > wl) if side == "lower" else (got[a] < wl)


def _residual(o):
    d = o.div_e() - o.rho()
    return d[3:-3, 3:-3, :] if o.p.ndim == 2 else d[3:-3, 3:-3, 3:-3]


@pytest.mark.parametrize("axis,side", FACES)
def test_reflection_keeps_gauss_law_for_normal_motion(axis, side):
    # a wall on a grid node; every particle moves normal to it and half of them cross it
    nx = 24
    o = Oracle(2, nx, nx, epc=2)
    wall = 6.0 if side == "lower" else 17.0
    o.set_bc_prtl(axis, wall, "refl", side)
    a = 0 if axis == "x" else 1
    wl = wall + 3.0
    rng = np.random.default_rng(11)
    n = 64
    pos = rng.uniform(8.0, 18.0, (n, 2))
    sgn = -1.0 if side == "lower" else 1.0
    un = sgn * rng.uniform(0.1, 3.0, n)
    gam = np.sqrt(1.0 + un * un)
    d = un * o.p.c / gam
    pos[:, a] = wl - rng.uniform(-0.9, 0.9, n) * np.abs(d)    # inside, some crossing
    pos[:, a] = np.where(sgn * (pos[:, a] - wl) > 0, 2 * wl - pos[:, a], pos[:, a])
    uvw = np.zeros((n, 3))
    uvw[:, a] = un
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    flv = np.where(q > 0, 1, 2).astype(np.int32)
    o.load_particles(q, pos[:, 0], pos[:, 1], np.full(n, 1.5), uvw[:, 0], uvw[:, 1], uvw[:, 2], flv)
    r0 = _residual(o).copy()
    o.step(1)
    P = o.particles()
    r1 = _residual(o)
    assert np.max(np.abs(r1 - r0)) < 1e-16, np.max(np.abs(r1 - r0))
    assert np.max(np.abs(r0)) > 1e-4
    # some particles were reflected
    uu = P["u"] if a == 0 else P["v"]
    assert np.any(np.sign(uu) != sgn)


@pytest.mark.parametrize("axis,side", FACES)
def test_reflection_tangential_fold_keeps_summed_current(axis, side):
    nx = 24
    o = Oracle(2, nx, nx, epc=2)
    wall = 6.0 if side == "lower" else 17.0
    o.set_bc_prtl(axis, wall, "refl", side)
    a = 0 if axis == "x" else 1
    t = 1 - a
    wl = wall + 3.0
    rng = np.random.default_rng(13)
    n = 40
    pos = rng.uniform(8.0, 18.0, (n, 2))
    sgn = -1.0 if side == "lower" else 1.0
    pos[:, a] = wl - sgn * rng.uniform(0.05, 0.5, n)    # just inside the wall
    u = np.zeros((n, 3))
    u[:, a] = sgn * rng.uniform(0.5, 2.0, n)            # crossing
    u[:, t] = rng.uniform(-1.0, 1.0, n)                 # and moving along it
    u[:, 2] = rng.uniform(-1.0, 1.0, n)
    o.load_particles(np.ones(n), pos[:, 0], pos[:, 1], np.full(n, 1.25), u[:, 0], u[:, 1], u[:, 2],
                     np.ones(n, np.int32))
    o.reset_current()
    o.move_deposit()
    names = ("Jx", "Jy", "Jz")
    before = {k: o.F[k].copy() for k in names}
    o.bc_post_movdep()
    tang = names[t]
    assert abs(o.F[tang].sum() - before[tang].sum()) < 1e-15
    assert abs(o.F["Jz"].sum() - before["Jz"].sum()) < 1e-15
    # no current is left beyond the wall
    beyond = (slice(int(wl), None) if side == "upper" else slice(0, int(wl) - 1))
    idx = [slice(None)] * 3
    idx[a] = beyond
    assert np.all(o.F[tang][tuple(idx)] == 0.0)
    assert np.any(before[tang][tuple(idx)] != 0.0)


def test_reflection_applied_twice_is_the_same():
    # EnforceBC_PostMovDep calls each reflecting face twice (bc.F90:245-266); the
    # second call changes nothing
    o = Oracle(2, 24, 24, epc=2)
    for axis, side, pos in (("x", "lower", 5.0), ("x", "upper", 18.3), ("y", "lower", 4.6), ("y", "upper", 19.0)):
        o.set_bc_prtl(axis, pos, "refl", side)
    rng = np.random.default_rng(21)
    n = 400
    x = rng.uniform(7.0, 22.5, n)
    y = rng.uniform(7.0, 22.5, n)
    u = rng.uniform(-3, 3, (n, 3))
    o.load_particles(np.ones(n), x, y, np.full(n, 1.5), u[:, 0], u[:, 1], u[:, 2], np.ones(n, np.int32))
    o.reset_current()
    o.move_deposit()
    o.bc_post_movdep()
    P1 = o.particles()
    J1 = {k: o.F[k].copy() for k in ("Jx", "Jy", "Jz")}
    o.bc_post_movdep()
    P2 = o.particles()
    for k in P1:
        assert np.array_equal(P1[k], P2[k])
    for k in J1:
        assert np.array_equal(J1[k], o.F[k])


# ---------------------------------------------------------------- removal
def test_removing_faces_make_holes_of_exactly_the_particles_beyond():
    o = Oracle(3, 16, 16, 16, epc=2)
    pos = {("x", "lower"): 3.2, ("x", "upper"): 12.7, ("y", "lower"): 2.5, ("y", "upper"): 13.1,
           ("z", "lower"): 4.0, ("z", "upper"): 11.9}
    for (axis, side), p in pos.items():
        o.set_bc_prtl(axis, p, "remv", side)
    rng = np.random.default_rng(2)
    n = 3000
    r = rng.uniform(3.0, 19.0, (n, 3))
    o.load_particles(np.ones(n), r[:, 0], r[:, 1], r[:, 2], np.zeros(n), np.zeros(n), np.zeros(n),
                     np.ones(n, np.int32))
    o.bc_post_movdep()
    lo = np.array([pos[("x", "lower")], pos[("y", "lower")], pos[("z", "lower")]]) + 3.0
    hi = np.array([pos[("x", "upper")], pos[("y", "upper")], pos[("z", "upper")]]) + 3.0
    keep = np.all((r >= lo) & (r <= hi), axis=1)
    flv = o.P["flv"][:n]
    assert np.array_equal(flv != 0, keep)
    gone = ~keep
    assert np.all(o.P["x"][:n][gone] == 3.8) and np.all(o.P["z"][:n][gone] == 3.8)
    assert np.all(o.P["q"][:n][gone] == 0) and np.all(o.P["u"][:n][gone] == 0)
    assert np.array_equal(o.P["x"][:n][keep], r[keep, 0])
    assert 0 < keep.sum() < n


# ---------------------------------------------------------------- field faces
def _pulse(o, x0, k=0.8, w=6.0):
    i = np.arange(o.p.mx) - 3.0
    c = o.p.c
    prof = np.exp(-((i - x0) / w) ** 2) * np.cos(k * (i - x0))
    xb = i + 0.5 + 0.5 * c           # B half a step behind, at the staggered x
    profB = np.exp(-((xb - x0) / w) ** 2) * np.cos(k * (xb - x0))
    o.F["Ey"][...] = prof[:, None, None]
    o.F["Bz"][...] = profB[:, None, None]
    o.field_halo("E")
    o.field_halo("B")


def _run_energy(o, steps):
    e0 = o.energy()[:2].sum()
    for t in range(1, steps + 1):
        o.step(t)
    return o.energy()[:2].sum() / e0


def test_conducting_face_reflects_the_pulse_and_stops_it():
    nx = 128
    o = Oracle(2, nx, 8, epc=1, cap=16)
    o.set_bc_fld("x", 10.0, "cond", "lower")
    o.set_bc_fld("x", nx - 10.0, "cond", "upper")
    _pulse(o, nx / 2)
    ratio = _run_energy(o, 200)
    assert abs(ratio - 1.0) < 5e-3, ratio
    right = int(math.ceil(nx - 10.0 + 3.0))
    assert np.all(o.F["Ey"][right - 1:, :, :] == 0.0) and np.all(o.F["Ez"][right - 1:, :, :] == 0.0)
    # the reflected pulse travels back at the Yee group speed for k = 0.8 per cell
    c, k = o.p.c, 0.8
    vg = c * math.cos(k / 2) / math.sqrt(1.0 - (c * math.sin(k / 2)) ** 2)
    wall = nx - 10.0
    xpk = 2 * wall - (nx / 2 + vg * 200)
    ey = np.abs(o.F["Ey"][:, 3, 0])
    assert abs(np.argmax(ey) - 3 - xpk) < 4, (np.argmax(ey) - 3, xpk)


def test_conducting_top_as_written_wipes_the_box():
    o = Oracle(2, 16, 32, epc=1, cap=16)
    o.set_bc_fld("y", 4.0, "cond", "lower")
    o.set_bc_fld("y", 28.0, "cond", "upper")
    rng = np.random.default_rng(0)
    for n in ("Ex", "Ey", "Ez"):
        o.F[n][...] = rng.standard_normal(o.shape)
    E0 = o.F["Ex"].copy()
    o.bc_final()
    assert np.any(o.F["Ex"][:, 8:28, :] != 0.0)
    assert np.array_equal(o.F["Ex"][:, 8:30, :], E0[:, 8:30, :])
    o.p.cond_top_as_written = 1
    o.p.bc_pos_fld[3] = 28.0
    o.F["Ex"][...] = E0
    o.bc_final()
    assert np.all(o.F["Ex"][:, 7:, :] == 0.0)   # everything above the bottom face


def test_pml_coefficients_closed_form():
    o = Oracle(2, 64, 8, epc=1, cap=16)
    L = 6.0
    o.set_bc_fld("x", 50.0, "pml", "upper", thickness=L)
    o.set_bc_fld("x", 9.0, "pml", "lower", thickness=L)
    for d in (0.5, 1.5, 3.0, 5.25, 6.0, 7.0):
        sigma = min(1.0, (d / L) ** 4)
        a = math.exp(-sigma - 9e-4)
        b = sigma / (sigma + 9e-4) * (a - 1.0)
        for x in (50.0 + d, 9.0 - d):
            got = o.pml_coeffs(x, 0)
            assert abs(got[0] - a) < 1e-15 and abs(got[1] - b) < 1e-15 and got[2] == 0.0
    assert np.array_equal(o.pml_coeffs(30.0, 0), np.zeros(3))     # interior
    assert np.array_equal(o.pml_coeffs(60.0, 1), np.zeros(3))     # another axis


def test_pml_absorbs_the_pulse():
    nx = 128
    o = Oracle(2, nx, 8, epc=1, cap=16)
    o.set_bc_fld("x", 10.0, "pml", "lower", thickness=8.0)
    o.set_bc_fld("x", nx - 10.0, "pml", "upper", thickness=8.0)
    _pulse(o, nx / 2)
    ratio = _run_energy(o, 220)
    assert ratio < 1e-5, ratio
    assert np.any(o.pml_psi[8] != 0.0) and not np.any(o.pml_psi[11] != 0.0)   # B psi1 z (x layer) only


def test_pml_absorbs_in_3d():
    n = 48
    o = Oracle(3, n, 8, 8, epc=1, cap=16)
    o.set_bc_fld("x", 6.0, "pml", "lower", thickness=6.0)
    o.set_bc_fld("x", n - 6.0, "pml", "upper", thickness=6.0)
    _pulse(o, n / 2, k=0.9, w=4.0)
    ratio = _run_energy(o, 110)
    assert ratio < 1e-3, ratio


def test_open_face_lets_the_pulse_out():
    nx = 128
    o = Oracle(2, nx, 8, epc=1, cap=16)
    o.set_bc_fld("x", 10.0, "open", "lower")
    o.set_bc_fld("x", nx - 10.0, "open", "upper")
    _pulse(o, nx / 2)
    ratio = _run_energy(o, 220)
    assert ratio < 0.1, ratio


def test_faces_move_at_speed_times_c():
    o = Oracle(2, 32, 16, epc=1, cap=16)
    o.set_bc_prtl("x", 20.0, "refl", "upper")
    o.set_bc_fld("x", 24.0, "cond", "upper")
    o.set_boundary_speed("x", "upper", 0.3)
    pos_p, pos_f = 20.0, 24.0
    for t in range(1, 11):
        o.step(t)
        pos_p = pos_p + 0.3 * o.p.c
        pos_f = pos_f + 0.3 * o.p.c
    assert o.p.bc_pos_prtl[1] == pos_p and o.p.bc_pos_fld[1] == pos_f
    assert o.p.bc_pos_prtl[0] == 0.0


@pytest.mark.parametrize("axis,side", FACES)
def test_reflection_fold_moves_current_onto_the_wall_node(axis, side):
    # wall on the node row w (local index): the tangential current deposited on the
    # row beyond the wall lands on the wall's own row; the normal current of the cell
    # beyond the wall comes back into the cell inside it with its sign reversed
    nx = 24
    wall = 6.0 if side == "lower" else 17.0
    a = 0 if axis == "x" else 1
    w = int(wall) + 3
    rng = np.random.default_rng(17)
    n = 60
    pos = rng.uniform(8.0, 18.0, (n, 2))
    sgn = -1.0 if side == "lower" else 1.0
    pos[:, a] = w - sgn * rng.uniform(0.02, 0.4, n)
    u = rng.uniform(-1.0, 1.0, (n, 3))
    u[:, a] = sgn * rng.uniform(0.5, 2.0, n)
    J = []
    for bc in (False, True):
        o = Oracle(2, nx, nx, epc=2)
        if bc:
            o.set_bc_prtl(axis, wall, "refl", side)
        o.load_particles(np.ones(n), pos[:, 0], pos[:, 1], np.full(n, 1.5), u[:, 0], u[:, 1], u[:, 2],
                         np.ones(n, np.int32))
        o.reset_current()
        o.move_deposit()
        o.bc_post_movdep()
        J.append({k: o.F[k][:, :, 0].copy() if a == 0 else o.F[k][:, :, 0].T.copy() for k in ("Jx", "Jy", "Jz")})
    raw, got = J
    nrm = "Jx" if a == 0 else "Jy"
    tng = "Jy" if a == 0 else "Jx"
    W = w - 1                      # 0-based row of the wall node
    out = W + 1 if side == "upper" else W - 1      # node row beyond the wall
    want = {k: raw[k].copy() for k in raw}
    for k in (tng, "Jz"):
        want[k][W] += raw[k][out]
        want[k][out] = 0.0
    cin, cout = (W - 1, W) if side == "upper" else (W, W - 1)   # normal-flux cells inside / beyond
    want[nrm][cin] -= raw[nrm][cout]
    want[nrm][cout] = 0.0
    for k in raw:
        assert np.array_equal(got[k], want[k]), k
    assert np.any(raw[tng][out] != 0.0) and np.any(raw[nrm][cout] != 0.0)
