"""Pins of the oracle's diagnostics and of the 2-D B gather (no GPU).

What each check rests on, independently of the oracle's own formulas:
  * 2-D B gather (movdep_routines.F90:115-223): B linear in space, on the
    Yee staggering (Bx at (i, j+1/2), By at (i+1/2, j), Bz at (i+1/2, j+1/2)),
    is reproduced exactly at the particle.  B is read back from the Boris
    rotation, which is exactly the implicit-midpoint (Cayley) rotation
    u' - u = (u' + u) x t with t = (q/2m) B / (c gamma): three particles of
    equal |u| at one point give t by least squares.
  * CalcFldEnergy (savedata_routines.F90:310-327): a uniform value F0 in one
    component gives 1/2 F0^2 nx ny (nz) whatever the ghost cells hold.
  * CalcPrtlEnergy (savedata_routines.F90:283-308): particles of known gamma
    give (gamma - 1) m c^2 per flavour, with massi = qi mi and masse = qi
    from InitScale's closed form (initialise.F90:44-53), qi computed here
    from the configuration; holes count for nothing.
  * momentum sum: a uniform E kicks every particle by (q/m) E / c, so the
    total m u changes by (N_ion - N_electron) qi E / c, which fixes the mass
    of each flavour against its charge; hand-built particles give sum m u
    with the masses above.
  * modal energies: the energy of a field sampled as F0 (m = 0) or
    F1c cos(theta) + F1s sin(theta) (m = 1) on the full r-theta-z grid (whose
    radius weights are pinned in test_oracle_pins_cyl.py) equals the modal
    energy of the slots, because sum_j cos^2(theta_j) dtheta = pi and
    sum_j dtheta = 2 pi on a uniform theta grid: a wrong mode weight (2 pi
    for m = 1) fails.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import CylModalOracle, CylOracle, Oracle

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def _qi(c=0.45, c_ompe=10.0, g0=1.0, mi_me=1.0, epc=16):
    # InitScale, initialise.F90:44-53: ompe = c / c_ompe, qi = g0 ompe^2 / (epc (1 + me/mi))
    ompe = c / c_ompe
    return g0 * ompe * ompe / (epc * (1.0 + 1.0 / mi_me))


# ---------------------------------------------------------------- 2-D B gather
def _b_from_rotation(o, x, y, umag=0.1):
    """B at (x, y) read from the Boris rotation of three particles (E = 0)."""
    U = umag * np.eye(3)
    n = 3
    o.load_particles(np.ones(n), np.full(n, x), np.full(n, y), np.full(n, 1.5), U[:, 0].copy(), U[:, 1].copy(),
                     U[:, 2].copy(), np.ones(n, dtype=np.int32), cap=64)
    o.move_deposit()
    P = o.particles()
    V = np.stack([P["u"], P["v"], P["w"]], axis=1)
    # order of the particles after the push is kept (no exchange): row k is particle k
    A = []
    d = []
    for k in range(n):
        s = V[k] + U[k]
        # s x t as a matrix acting on t
        A.append(np.array([[0.0, -s[2], s[1]], [s[2], 0.0, -s[0]], [-s[1], s[0], 0.0]]))
        d.append(V[k] - U[k])
    t, *_ = np.linalg.lstsq(np.concatenate(A), np.concatenate(d), rcond=None)
    qm_half = 0.5 * o.p.flvrqm[0]
    g = 1.0 / (o.p.c * math.sqrt(1.0 + umag * umag))   # 1 / (c gamma)
    return t / (g * qm_half)


def test_boris_is_the_cayley_rotation():
    # the relation the B read-back rests on: u' - u = (u' + u) x t for a known uniform B
    o = Oracle(2, 16, 16, epc=4)
    B = np.array([0.013, -0.021, 0.034])
    for n, v in zip(("Bx", "By", "Bz"), B):
        o.F[n][...] = v
    got = _b_from_rotation(o, 9.3, 8.6)
    assert np.allclose(got, B, rtol=1e-12, atol=0), (got, B)


def test_gather_reproduces_linear_b_fields_2d():
    o = Oracle(2, 16, 16, epc=4)
    p = o.p
    i = np.arange(p.mx)[:, None, None] + 1.0   # Fortran index of element [i-1]
    j = np.arange(p.my)[None, :, None] + 1.0
    b0 = np.array([0.02, -0.015, 0.03])
    gx = np.array([1e-3, -2e-3, 1.5e-3])
    gy = np.array([-1.2e-3, 0.8e-3, 2e-3])
    # Yee staggering of the reference 2-D grid
    pos = {"Bx": (i, j + 0.5), "By": (i + 0.5, j), "Bz": (i + 0.5, j + 0.5)}
    for a, n in enumerate(("Bx", "By", "Bz")):
        X, Y = pos[n]
        o.F[n][...] = b0[a] + gx[a] * X + gy[a] * Y
    rng = np.random.default_rng(4)
    for _ in range(12):
        x, y = rng.uniform(5, 14), rng.uniform(5, 14)
        got = _b_from_rotation(o, x, y)
        want = b0 + gx * x + gy * y
        assert np.allclose(got, want, rtol=0, atol=5e-14), (x, y, got, want)


# ---------------------------------------------------------------- field energy
@pytest.mark.parametrize("ndim,n", [(2, (16, 12, 1)), (3, (8, 10, 6))])
@pytest.mark.parametrize("comp", FIELDS)
def test_field_energy_of_a_uniform_component(ndim, n, comp):
    o = Oracle(ndim, *n, epc=2)
    F0 = 0.37
    for m in FIELDS:
        o.F[m][...] = 1e3   # ghost cells: never counted
    for m in FIELDS:
        sl = (slice(2, -2), slice(2, -2), slice(2, -2) if ndim == 3 else slice(None))
        o.F[m][sl] = F0 if m == comp else 0.0
    e = o.energy()
    cells = n[0] * n[1] * (n[2] if ndim == 3 else 1)
    want = 0.5 * F0 * F0 * cells
    got = e[0] if comp[0] == "B" else e[1]
    other = e[1] if comp[0] == "B" else e[0]
    assert abs(got - want) <= 1e-14 * want, (got, want)
    assert other == 0.0


# ---------------------------------------------------------------- particle energy
def test_particle_energy_per_flavour_closed_form():
    mi_me, c, epc = 25.0, 0.45, 4
    o = Oracle(2, 16, 16, epc=epc, c=c, mi_me=mi_me)
    qi = _qi(c=c, mi_me=mi_me, epc=epc)
    massi, masse = qi * mi_me, qi
    # gamma exact in binary for the first rows: u = 3/4 -> 5/4, (u, v) = (0, 4/3) -> 5/3
    U = np.array([[0.75, 0, 0], [0, 4.0 / 3.0, 0], [0.3, -0.2, 0.5], [0, 0, 0], [1.0, 2.0, 2.0], [0.1, 0.1, 0.1]])
    flv = np.array([1, 1, 2, 2, 2, 0], dtype=np.int32)    # the last one is a hole
    q = np.where(flv == 1, 1.0, -1.0)
    n = len(flv)
    o.load_particles(q, np.full(n, 8.5), np.full(n, 8.5), np.full(n, 1.5), U[:, 0].copy(), U[:, 1].copy(),
                     U[:, 2].copy(), flv, cap=64)
    e = o.energy()
    gam = np.sqrt(1.0 + (U * U).sum(axis=1))
    want1 = ((gam - 1.0) * massi * c * c)[flv == 1].sum()
    want2 = ((gam - 1.0) * masse * c * c)[flv == 2].sum()
    assert abs(e[2] - want1) <= 1e-14 * want1, (e[2], want1)
    assert abs(e[3] - want2) <= 1e-14 * want2, (e[3], want2)
    # the exact rows on their own
    assert abs(((gam[0] - 1.0) + (gam[1] - 1.0)) - (0.25 + 2.0 / 3.0)) < 1e-15


# ---------------------------------------------------------------- momentum
def test_momentum_sum_closed_form():
    mi_me, c, epc = 25.0, 0.45, 4
    o = Oracle(2, 16, 16, epc=epc, c=c, mi_me=mi_me)
    qi = _qi(c=c, mi_me=mi_me, epc=epc)
    rng = np.random.default_rng(2)
    n = 40
    flv = np.where(np.arange(n) % 3 == 0, 1, 2).astype(np.int32)
    q = np.where(flv == 1, 1.0, -1.0)
    U = rng.normal(0, 0.2, (n, 3))
    o.load_particles(q, rng.uniform(5, 14, n), rng.uniform(5, 14, n), np.full(n, 1.5), U[:, 0].copy(),
                     U[:, 1].copy(), U[:, 2].copy(), flv, cap=128)
    m = np.where(flv == 1, qi * mi_me, qi)
    want = (m[:, None] * U).sum(axis=0)
    got = o.momentum()
    assert np.allclose(got, want, rtol=1e-14, atol=1e-18), (got, want)


def test_momentum_change_in_uniform_e_is_net_charge_times_e():
    mi_me, c, epc = 25.0, 0.45, 4
    o = Oracle(2, 16, 16, epc=epc, c=c, mi_me=mi_me)
    qi = _qi(c=c, mi_me=mi_me, epc=epc)
    E = np.array([2e-3, -1e-3, 3e-3])
    for nm, v in zip(("Ex", "Ey", "Ez"), E):
        o.F[nm][...] = v
    rng = np.random.default_rng(6)
    n = 30
    flv = np.where(np.arange(n) < 19, 1, 2).astype(np.int32)   # 19 ions, 11 electrons
    q = np.where(flv == 1, 1.0, -1.0)
    U = rng.normal(0, 0.1, (n, 3))
    o.load_particles(q, rng.uniform(5, 14, n), rng.uniform(5, 14, n), np.full(n, 1.5), U[:, 0].copy(),
                     U[:, 1].copy(), U[:, 2].copy(), flv, cap=128)
    p0 = o.momentum()
    o.move_deposit()
    p1 = o.momentum()
    want = (19 - 11) * qi * E / c
    assert np.allclose(p1 - p0, want, rtol=1e-12, atol=0), (p1 - p0, want)


# ---------------------------------------------------------------- modal energy weights
def _sample(full, comp, fc, fs):
    # E_r, E_z, B_theta at theta_j; E_theta, B_r, B_z at theta_{j+1/2}
    dth = full.cy.dtheta
    half = comp in ("Ey", "Bx", "Bz")
    j = np.arange(full.p.my) + 1.0
    th = (j - 3.0 + (0.5 if half else 0.0)) * dth
    return fc[:, None, :] * np.cos(th)[None, :, None] + fs[:, None, :] * np.sin(th)[None, :, None]


@pytest.mark.parametrize("slot", [0, 1, 2])
@pytest.mark.parametrize("comp", FIELDS)
def test_modal_energy_equals_full_grid_energy(slot, comp):
    nr, nth, nz = 9, 16, 8
    full = CylOracle(3, nr, nth, nz, c=0.1, epc=1, cap=16)
    modal = CylModalOracle(nr, nz, dr=1.0, c=0.1, epc=1, cap=16)
    mx, mz = modal.shape[0], modal.shape[2]
    rng = np.random.default_rng(10 * slot + FIELDS.index(comp))
    prof = rng.uniform(0.5, 1.5, (mx, mz))   # any (r, z) profile
    for n in FIELDS:
        full.F[n][...] = 0.0
        modal.F[n][...] = 0.0
    modal.F[comp][:, slot, :] = prof
    if slot == 0:
        full.F[comp][...] = prof[:, None, :]
    elif slot == 1:
        full.F[comp][...] = _sample(full, comp, prof, 0.0 * prof)
    else:
        full.F[comp][...] = _sample(full, comp, 0.0 * prof, prof)
    ef = full.energy()[:2]
    em = modal.energy()
    k = 0 if comp[0] == "B" else 1
    assert ef[k] > 0
    assert abs(em[k] - ef[k]) <= 1e-13 * ef[k], (em, ef)
    assert em[1 - k] == 0.0


def test_modal_uniform_m0_energy_closed_form():
    # uniform E_z = F0 in m = 0 on i = 4..mx-3 (radii (i - 3.5) dr, the node stagger of E_z):
    # 1/2 F0^2 2 pi dr sum_i r_i nz = 1/2 F0^2 2 pi dr^2 nz (0 + 1 + ... + nr-1) ... at r = i - 3.5
    nr, nz, dr = 8, 9, 2.0
    modal = CylModalOracle(nr, nz, dr=dr, c=0.1, epc=1, cap=16)
    F0 = 0.8
    modal.F["Ez"][:, 0, :] = F0
    em = modal.energy()
    r = (np.arange(4, modal.shape[0] - 2) - 3.5) * dr      # i = 4..mx-3 (1-based)
    want = 0.5 * F0 * F0 * 2.0 * math.pi * dr * r.sum() * nz
    assert abs(em[1] - want) <= 1e-14 * want, (em, want)
