"""Energy and field files (save_total_energy, save_field_collective) on the GPU.

The energy file holds CalcEnergy of the device state: it must equal
pictor_diagnostics bit for bit and the oracle's energies within 1e-9 (FP64
form).  The field file must equal the CollectFld restatement (tests/_collect.py,
pinned in test_output_collect.py) applied to the downloaded fields, bit for bit,
for both forms, 2-D and 3-D, several sampling ratios, with the current.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import Simulation, setup_weibel
from tests._collect import collect
from tests._util import FIELDS, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
def test_energy_file(form, tmp_path):
    nx = ny = 32
    P = setup_weibel(nx, ny, epc=8, drift=0.5, uth=0.05, seed=4)
    o = Oracle(2, nx, ny, epc=8, cap=2 * len(P["x"]))
    o.load_particles(**P)
    for t in range(1, 11):
        o.step(t)
    with Simulation(nx, ny, epc=8, form=form, tile=8) as sim:
        sim.upload_particles(**P)
        sim.step(10)
        path = str(tmp_path / "energy_10")
        sim.save_total_energy(path)
        d = sim.diagnostics()
    z = np.load(path)
    assert sorted(z.files) == ["Benergy", "Eenergy", "PrtlEnergy"]
    assert z["Benergy"].shape == (1,) and z["PrtlEnergy"].shape == (2,)
    assert z["Benergy"][0] == d["B_energy"] and z["Eenergy"][0] == d["E_energy"]
    assert z["PrtlEnergy"][0] == d["prtl_energy_1"] and z["PrtlEnergy"][1] == d["prtl_energy_2"]
    e = o.energy()
    tol = 1e-9 if form == "fp64" else 1e-3
    for got, want in ((z["Benergy"][0], e[0]), (z["Eenergy"][0], e[1]), (z["PrtlEnergy"][0], e[2]),
                      (z["PrtlEnergy"][1], e[3])):
        assert abs(got - want) <= tol * abs(want), (got, want)


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
@pytest.mark.parametrize("ndim,res", [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_field_file_equals_collectfld(form, ndim, res, tmp_path):
    n = 32 if ndim == 2 else 16
    say(f"field file {form} {ndim}-D n={n} res={res}")
    P = setup_weibel(n, n, n, ndim=ndim, epc=4, drift=0.5, uth=0.05, seed=6)
    with Simulation(n, n, n, ndim=ndim, epc=4, form=form, tile=8 if ndim == 2 else 4) as sim:
        sim.upload_particles(**P)
        sim.step(6)
        path = str(tmp_path / "fld_6")
        sim.save_field_collective(path, res=res, save_current=True)
        F = sim.download_fields()
    z = np.load(path)
    assert sorted(z.files) == sorted(FIELDS + ("Jx", "Jy", "Jz"))
    want = collect(F, ndim, res)
    for name in want:
        assert z[name].dtype == np.float32 and np.isfortran(z[name]), name
        assert np.array_equal(z[name], want[name]), (name, np.max(np.abs(z[name] - want[name])))
    assert np.max(np.abs(z["Jx"])) > 0 and np.max(np.abs(z["Bz"])) > 0
