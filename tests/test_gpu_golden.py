"""The HIP path from the golden fixtures' inputs against their stored outputs
(tests/golden/*.npz, 10 oracle steps each), through the C ABI (needs an MI355X).

Bars: the FP64 form within 1e-9 of every stored field (E, B and J; the GPU
sums the current in another order), energies within 1e-9 (fields) and 1e-12
(particles), the same particle count and every particle column within 1e-9
after sorting; the FP32 split form within 1e-3 of the stored field and
particle energies, with the same particle count.
"""
import os

import numpy as np
import pytest

from pictor_amd import Simulation
from tests._util import FIELDS, rel_err, say, sort_particles
from tests.test_golden import load, mg

pytestmark = pytest.mark.gpu


def _run(name, form):
    ref = load(name)
    ndim, nx, ny, nz, epc, fdtd, nsteps, seed, pfaces, ffaces = mg.CASES[name]
    say(f"golden {name} form={form}")
    sim = Simulation(nx, ny, nz, ndim=ndim, epc=epc, form=form, fdtd=fdtd, tile=8 if ndim == 2 else 4)
    with sim:
        for axis, side, kind, pos in ffaces:
            sim.SetBC_Fld(axis, pos, kind, side)
        for axis, side, kind, pos in pfaces:
            sim.SetBC_Prtl(axis, pos, kind, side)
        sim.upload_fields(*[ref["in_" + f] for f in FIELDS])
        sim.upload_particles(**{k: ref["in_" + k] for k in ("q", "x", "y", "z", "u", "v", "w", "flv")})
        sim.step(nsteps)
        return ref, sim.download_fields(), sim.download_particles(), sim.diagnostics()


@pytest.mark.parametrize("name", list(mg.CASES))
def test_golden_fp64(name):
    ref, G, P, d = _run(name, "fp64")
    for f in FIELDS + ("Jx", "Jy", "Jz"):
        assert rel_err(G[f], ref["out_" + f]) < 1e-9, f
    e = ref["out_energy"]
    assert abs(d["B_energy"] - e[0]) <= 1e-9 * e[0]
    assert abs(d["E_energy"] - e[1]) <= 1e-9 * e[1]
    assert abs(d["prtl_energy_1"] - e[2]) <= 1e-12 * e[2]
    assert abs(d["prtl_energy_2"] - e[3]) <= 1e-12 * e[3]
    assert len(P["x"]) == len(ref["out_x"])
    S = sort_particles(P)
    assert np.array_equal(S["flv"], ref["out_flv"])
    for c in ("q", "x", "y", "z", "u", "v", "w"):
        assert np.max(np.abs(S[c] - ref["out_" + c])) < 1e-9, c


@pytest.mark.parametrize("name", list(mg.CASES))
def test_golden_split(name):
    ref, G, P, d = _run(name, "fp32_split")
    e = ref["out_energy"]
    got = (d["B_energy"], d["E_energy"], d["prtl_energy_1"], d["prtl_energy_2"])
    for a, b in zip(got, e):
        assert abs(a - b) <= 1e-3 * b, (got, e)
    assert len(P["x"]) == len(ref["out_x"])
