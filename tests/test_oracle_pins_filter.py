"""Pins of the oracle's current filter (smoothen_current, current.F90:19-30), no GPU.

What the mathematics fixes, independently of the oracle's loops:
  * on a periodic box one pass along an axis is the circular convolution with
    (1, 2, 1) / 4, so a Fourier mode cos(2 pi m i / n) is an eigenvector with
    eigenvalue cos^2(pi m / n) per axis pass; curr_filter passes raise it to
    that power (weights, the periodic images and the pass count all enter);
  * a unit spike becomes the outer product of (1, 2, 1) / 4 along every axis,
    wrapped across the periodic boundary (exact: powers of two);
  * the sum over the physical cells is unchanged (the weights sum to one);
  * curr_filter = 0 leaves J untouched, and E and B are never touched.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle

J = ("Jx", "Jy", "Jz")


def interior(o):
    if o.p.ndim == 2:
        return (slice(2, o.p.mx - 3), slice(2, o.p.my - 3), slice(0, 1))
    return (slice(2, o.p.mx - 3), slice(2, o.p.my - 3), slice(2, o.p.mz - 3))


def periodic_fill(o, fn):
    """J(i,j,k) = fn(i-3, j-3, k-3) on every cell, ghosts as periodic images."""
    nx, ny, nz = o.p.nx, o.p.ny, o.p.nz
    i = (np.arange(o.p.mx) - 2) % nx
    j = (np.arange(o.p.my) - 2) % ny
    k = (np.arange(o.p.mz) - 2) % nz if o.p.ndim == 3 else np.zeros(1, dtype=int)
    I, Jj, K = np.meshgrid(i, j, k, indexing="ij")
    return fn(I, Jj, K)


@pytest.mark.parametrize("ndim,passes,m", [(2, 1, (1, 0, 0)), (2, 2, (3, 2, 0)), (2, 3, (0, 5, 0)),
                                           (3, 1, (1, 2, 3)), (3, 2, (2, 0, 1))])
def test_fourier_mode_eigenvalue(ndim, passes, m):
    n = (16, 12, 10) if ndim == 3 else (16, 12, 1)
    o = Oracle(ndim, *n, epc=1, curr_filter=passes)
    kk = [2 * math.pi * m[a] / n[a] for a in range(3)]
    mode = lambda I, Jj, K: np.cos(kk[0] * I + 0.3) * np.cos(kk[1] * Jj - 0.7) * np.cos(kk[2] * K + 0.1)
    for c, name in enumerate(J):
        o.F[name][...] = (c + 1.0) * periodic_fill(o, mode)
    before = {name: o.F[name][interior(o)].copy() for name in J}
    o.smoothen_current()
    lam = 1.0
    for a in range(ndim):
        lam *= math.cos(math.pi * m[a] / n[a]) ** 2
    lam **= passes
    for name in J:
        np.testing.assert_allclose(o.F[name][interior(o)], lam * before[name], rtol=0, atol=1e-13)


@pytest.mark.parametrize("ndim,at", [(2, (3, 4)), (2, (0, 0)), (2, (15, 11)), (3, (0, 5, 9)), (3, (7, 0, 4))])
def test_spike_becomes_121_stencil_with_wrap(ndim, at):
    n = (16, 12, 10) if ndim == 3 else (16, 12, 1)
    o = Oracle(ndim, *n, epc=1, curr_filter=1)
    at3 = tuple(at) + ((0,) if ndim == 2 else ())
    spike = np.zeros(n if ndim == 3 else n[:2] + (1,))
    spike[at3] = 64.0
    for name in J:
        o.F[name][...] = periodic_fill(o, lambda I, Jj, K: spike[I, Jj, K])
    o.smoothen_current()
    want = np.zeros_like(spike)
    w = {-1: 0.25, 0: 0.5, 1: 0.25}
    shifts = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in ((-1, 0, 1) if ndim == 3 else (0,))]
    for s in shifts:
        idx = tuple((at3[d] + s[d]) % (n[d] if d < ndim else 1) for d in range(3))
        want[idx] += 64.0 * w[s[0]] * w[s[1]] * (w[s[2]] if ndim == 3 else 1.0)
    for name in J:
        assert np.array_equal(o.F[name][interior(o)], want), name


def test_sum_preserved_and_other_fields_untouched():
    o = Oracle(2, 20, 14, epc=1, curr_filter=2)
    rng = np.random.default_rng(3)
    for name in o.FIELD_NAMES:
        o.F[name][...] = rng.normal(size=o.shape)
    others = {n: o.F[n].copy() for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")}
    # the filter reads the physical cells through their periodic images
    for name in J:
        o.F[name][...] = periodic_fill(o, lambda I, Jj, K, a=o.F[name][2:-3, 2:-3, 0].copy(): a[I, Jj])
    s0 = {name: o.F[name][interior(o)].sum() for name in J}
    o.smoothen_current()
    for name in J:
        assert abs(o.F[name][interior(o)].sum() - s0[name]) <= 1e-12 * np.abs(o.F[name]).sum()
    for n, a in others.items():
        assert np.array_equal(o.F[n], a), n


def test_zero_passes_is_identity():
    o = Oracle(3, 10, 8, 6, epc=1, curr_filter=0)
    rng = np.random.default_rng(5)
    for name in J:
        o.F[name][...] = rng.normal(size=o.shape)
    before = {name: o.F[name].copy() for name in J}
    o.smoothen_current()
    for name in J:
        assert np.array_equal(o.F[name], before[name])


def test_step_applies_filter_between_fold_and_add_current():
    # one step with curr_filter = 1 equals the phases run by hand in MAIN.F90 order
    rng = np.random.default_rng(11)
    n = 2000
    args = (np.ones(n), 3 + 16 * rng.random(n), 3 + 12 * rng.random(n), 1 + rng.random(n),
            rng.normal(0, 0.2, n), rng.normal(0, 0.2, n), rng.normal(0, 0.2, n), np.where(np.arange(n) % 2, 1, 2))
    a = Oracle(2, 16, 12, epc=4, curr_filter=1, cap=4096)
    b = Oracle(2, 16, 12, epc=4, curr_filter=1, cap=4096)
    for o in (a, b):
        o.load_particles(*[np.asarray(x) for x in args[:7]], np.asarray(args[7], dtype=np.int32))
    a.run(0, 1)
    b.reset_current(); b.update_b_half(); b.field_halo("B"); b.move_deposit(); b.exchange()
    b.update_b_half(); b.field_halo("B"); b.update_e(); b.field_halo("E"); b.current_halo_add()
    b.smoothen_current(); b.add_current(); b.field_halo("E")
    for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz"):
        assert np.array_equal(a.F[name], b.F[name]), name
    c = Oracle(2, 16, 12, epc=4, curr_filter=0, cap=4096)
    c.load_particles(*[np.asarray(x) for x in args[:7]], np.asarray(args[7], dtype=np.int32))
    c.run(0, 1)
    assert not np.array_equal(a.F["Ex"], c.F["Ex"])
