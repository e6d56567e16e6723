"""pictor_energy_noalloc against pictor_diagnostics, through the C ABI (needs an MI355X).

The allocation-free entry point keeps its partial sums in a static device
array shared by every call on the GPU.  Two calls in a row, the second on a
smaller context, must each equal out[0..6] of pictor_diagnostics on the same
state bit for bit (same kernels, same summation order); a grid whose tiles
need more partials than the array holds is refused with PICTOR_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from pictor_amd import Simulation, setup_weibel
from pictor_amd._lib import PICTOR_ERR_ARG, lib
from tests._util import say

pytestmark = pytest.mark.gpu


def _noalloc(sim):
    L = lib()
    out = np.full(7, np.nan)
    rc = L.pictor_energy_noalloc(sim._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, out


def _diag(sim):
    L = lib()
    out = np.zeros(9)
    assert L.pictor_diagnostics(sim._ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return out


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
def test_noalloc_equals_diagnostics_twice(form):
    for n, tile, steps in ((128, 16, 4), (32, 8, 2)):   # the second call on a smaller grid
        say(f"noalloc {form} {n}^2 tile {tile}")
        P = setup_weibel(n, n, epc=4, drift=0.5, uth=0.01, seed=7 + n)
        with Simulation(n, n, epc=4, form=form, tile=tile) as sim:
            sim.upload_particles(**P)
            sim.step(steps)
            ref = _diag(sim)
            rc, got = _noalloc(sim)
            assert rc == 0, lib().pictor_last_error()
            assert np.array_equal(got, ref[:7]), (got, ref[:7])
            assert ref[0] > 0 and ref[2] > 0 and ref[3] > 0
            # a second call on the same state (the array keeps the last call's contents)
            rc, again = _noalloc(sim)
            assert rc == 0 and np.array_equal(again, got)


def test_noalloc_refuses_too_many_tiles():
    say("noalloc refusal 2048^2 tile 8")
    with Simulation(2048, 2048, epc=1, form="fp32_split", tile=8) as sim:
        rc, _ = _noalloc(sim)
        assert rc == PICTOR_ERR_ARG
        assert b"scratch" in lib().pictor_last_error()
