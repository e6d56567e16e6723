"""Pins of the oracle's Lehe B half step (em_update_lehe.F90) -- no GPU.

The Lehe scheme replaces the forward differences of the B update by
  gradX F = (a_x + 2 b_xy cos k_y + 2 b_xz cos k_z) sin(k_x/2) + d_x sin(3 k_x/2)
(in Fourier form, times 2i e^{i k_x/2}), likewise gradY, gradZ, with the E
update left Yee.  A plane wave exp(i(k.r - w n)) is then an exact discrete
eigenmode with
  sin^2(w/2) = c^2 sum_a sin(k_a/2) S_a,
where S_a is the bracket above.  The tests evolve such eigenmodes (on-axis
and oblique, 2-D and 3-D) and check them after many steps, which pins every
coefficient and stencil offset independently of the oracle's code.  The
k = pi mode along x travels at exactly c: the coefficient d_x of the
reference is chosen for that, and the test uses w = pi c directly.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.test_oracle_pins_3d import _grid_coords


def _coef(o):
    L = list(o.p.lehe)
    return dict(ax=L[0], ay=L[1], az=L[2], bxy=L[3], bxz=L[4], byx=L[5], byz=L[6], bzx=L[7], bzy=L[8],
                dx=L[9], dy=L[10], dz=L[11])


def _run(o, n):
    for _ in range(n):
        o.update_b_half(); o.field_halo("B")
        o.update_b_half(); o.field_halo("B")
        o.update_e(); o.field_halo("E")


def test_lehe_coefficients_follow_the_reference_definition():
    # em_update_lehe.F90:8-25 with unit cell size: the identities the scheme is built on
    for ndim in (2, 3):
        o = Oracle(ndim, 8, 8, 8, fdtd="lehe")
        k = _coef(o)
        c = o.p.c
        assert abs((1 - 4 * k["dx"]) - math.sin(math.pi * c / 2) ** 2 / c ** 2) < 1e-15
        s_x = k["ax"] + 2 * k["bxy"] + (2 * k["bxz"] if ndim == 3 else 0) + 3 * k["dx"]
        s_y = k["ay"] + 2 * k["byx"] + (2 * k["byz"] if ndim == 3 else 0) + 3 * k["dy"]
        assert abs(s_x - 1) < 1e-15 and abs(s_y - 1) < 1e-15   # consistency: gradX -> d/dx as k -> 0


@pytest.mark.parametrize("m", [1, 3])
def test_lehe_plane_wave_along_x_2d(m):
    nx, ny = 32, 8
    o = Oracle(2, nx, ny, fdtd="lehe")
    k_ = _coef(o)
    c = o.p.c
    k = 2 * np.pi * m / nx
    S = (k_["ax"] + 2 * k_["bxy"]) * math.sin(k / 2) + k_["dx"] * math.sin(1.5 * k)
    w = 2 * math.asin(c * math.sqrt(math.sin(k / 2) * S))
    A = -c * S / math.sin(w / 2)
    i, j, kk = _grid_coords(o)
    X = (i - 3.0) * np.ones(o.shape)
    o.F["Ez"][...] = np.cos(k * X)
    o.F["By"][...] = A * np.cos(k * (X + 0.5) + 0.5 * w)
    n = 50
    _run(o, n)
    assert np.max(np.abs(o.F["Ez"] - np.cos(k * X - w * n))) < 1e-11
    assert np.max(np.abs(o.F["By"] - A * np.cos(k * (X + 0.5) - w * (n - 0.5)))) < 1e-11


def test_lehe_nyquist_mode_travels_at_c():
    # k = pi along x: w = pi c exactly (the reference's choice of d_x)
    nx, ny = 16, 8
    o = Oracle(2, nx, ny, fdtd="lehe")
    c = o.p.c
    k, w = math.pi, math.pi * c
    k_ = _coef(o)
    S = (k_["ax"] + 2 * k_["bxy"]) * 1.0 + k_["dx"] * (-1.0)   # sin(pi/2) = 1, sin(3pi/2) = -1
    A = -c * S / math.sin(w / 2)
    i, j, kk = _grid_coords(o)
    X = (i - 3.0) * np.ones(o.shape)
    o.F["Ez"][...] = np.cos(k * X)
    o.F["By"][...] = A * np.cos(k * (X + 0.5) + 0.5 * w)
    n = 40
    _run(o, n)
    assert np.max(np.abs(o.F["Ez"] - np.cos(k * X - w * n))) < 1e-11


def test_lehe_oblique_tm_wave_2d():
    nx, ny = 24, 16
    o = Oracle(2, nx, ny, fdtd="lehe")
    q = _coef(o)
    c = o.p.c
    kx, ky = 2 * np.pi * 2 / nx, 2 * np.pi / ny
    Sx = (q["ax"] + 2 * q["bxy"] * math.cos(ky)) * math.sin(kx / 2) + q["dx"] * math.sin(1.5 * kx)
    Sy = (q["ay"] + 2 * q["byx"] * math.cos(kx)) * math.sin(ky / 2) + q["dy"] * math.sin(1.5 * ky)
    w = 2 * math.asin(c * math.sqrt(math.sin(kx / 2) * Sx + math.sin(ky / 2) * Sy))
    sw = math.sin(w / 2)
    Ay, Ax = -c * Sx / sw, c * Sy / sw
    i, j, kk = _grid_coords(o)
    X, Y = (i - 3.0) * np.ones(o.shape), (j - 3.0) * np.ones(o.shape)

    def ez(n):
        return np.cos(kx * X + ky * Y - w * n)

    def by(n):
        return Ay * np.cos(kx * (X + 0.5) + ky * Y - w * (n - 0.5))

    def bx(n):
        return Ax * np.cos(kx * X + ky * (Y + 0.5) - w * (n - 0.5))

    o.F["Ez"][...] = ez(0); o.F["By"][...] = by(0); o.F["Bx"][...] = bx(0)
    n = 40
    _run(o, n)
    assert np.max(np.abs(o.F["Ez"] - ez(n))) < 1e-11
    assert np.max(np.abs(o.F["By"] - by(n))) < 1e-11
    assert np.max(np.abs(o.F["Bx"] - bx(n))) < 1e-11


def test_lehe_oblique_wave_3d():
    nx, nz = 16, 12
    o = Oracle(3, nx, 6, nz, fdtd="lehe")
    q = _coef(o)
    c = o.p.c
    kx, kz = 2 * np.pi / nx, 2 * np.pi / nz
    Sx = (q["ax"] + 2 * q["bxy"] + 2 * q["bxz"] * math.cos(kz)) * math.sin(kx / 2) + q["dx"] * math.sin(1.5 * kx)
    Sz = (q["az"] + 2 * q["bzx"] * math.cos(kx) + 2 * q["bzy"]) * math.sin(kz / 2) + q["dz"] * math.sin(1.5 * kz)
    w = 2 * math.asin(c * math.sqrt(math.sin(kx / 2) * Sx + math.sin(kz / 2) * Sz))
    sw = math.sin(w / 2)
    Ax, Az = -c * Sz / sw, c * Sx / sw
    i, j, k = _grid_coords(o)
    X, Z = (i - 3.0) * np.ones(o.shape), (k - 3.0) * np.ones(o.shape)

    def ey(n):
        return np.cos(kx * X + kz * Z - w * n)

    def bx(n):
        return Ax * np.cos(kx * X + kz * (Z + 0.5) - w * (n - 0.5))

    def bz(n):
        return Az * np.cos(kx * (X + 0.5) + kz * Z - w * (n - 0.5))

    o.F["Ey"][...] = ey(0); o.F["Bx"][...] = bx(0); o.F["Bz"][...] = bz(0)
    n = 30
    _run(o, n)
    assert np.max(np.abs(o.F["Ey"] - ey(n))) < 1e-11
    assert np.max(np.abs(o.F["Bx"] - bx(n))) < 1e-11
    assert np.max(np.abs(o.F["Bz"] - bz(n))) < 1e-11


def test_lehe_gauss_law_over_run():
    # the E update is Yee, so div E - rho stays at rounding level with the Lehe B step
    from pictor_amd import setup_weibel
    o = Oracle(2, 32, 32, epc=8, fdtd="lehe")
    P = setup_weibel(32, 32, epc=8, seed=11)
    o.load_particles(**P)
    for t in range(1, 41):
        o.step(t)
    r, rmax = o.gauss()
    assert r < 1e-13 * rmax, (r, rmax)
