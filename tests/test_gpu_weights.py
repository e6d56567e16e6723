"""Weighted particles in the FP32 split form (the reference's q = wtp * FlvrCharge,
init_prtl.F90:161), through the C ABI (needs an MI355X).

Bars: a split run whose particles carry charge weights follows the FP64 oracle run with
the same per-particle charges (field and particle energies within 1e-3 after 20 steps,
the Gauss residual within 1e-6 of max |rho| of the oracle's, the same particle count); downloads return
each particle's charge (weights exact in FP32 come back exactly); the weight travels
with tags and restart files; a charge of the wrong sign for its flavour is refused.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import PictorError, Simulation, setup_weibel
from tests._util import say

pytestmark = pytest.mark.gpu


def _weighted(nx, ny, epc, seed):
    P = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.05, seed=seed)
    rng = np.random.default_rng(seed)
    wt = rng.integers(4, 17, len(P["x"])) / 8.0          # 0.5 .. 2, exact in FP32
    P["q"] = P["q"] * wt
    return P, wt


def test_weighted_split_follows_the_oracle():
    nx = ny = 32
    P, wt = _weighted(nx, ny, 8, 3)
    o = Oracle(2, nx, ny, epc=8, cap=2 * len(P["x"]) + 1024)
    o.load_particles(**P)
    for t in range(1, 21):
        o.step(t)
    say("weighted split, 20 steps")
    with Simulation(nx, ny, epc=8, form="fp32_split", tile=8) as sim:
        sim.upload_particles(**P)
        G0 = sim.download_particles()
        assert np.array_equal(np.sort(G0["q"]), np.sort(P["q"]))   # weights exact in FP32
        sim.step(20)
        d = sim.diagnostics()
        n = sim.particle_count()
    eo = o.energy()
    go = o.gauss()
    say("split", d, "oracle", eo, go)
    assert abs(d["B_energy"] - eo[0]) <= 1e-3 * eo[0]
    assert abs(d["E_energy"] - eo[1]) <= 1e-3 * eo[1]
    assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-3 * eo[2]
    assert abs(d["prtl_energy_2"] - eo[3]) <= 1e-3 * eo[3]
    # unequal weights at one position leave a net charge with E = 0 at t = 0, a Gauss
    # residual that the charge-conserving deposit then keeps: the split form's equals the
    # oracle's to 1e-6 of max |rho|
    assert abs(d["gauss_residual"] - go[0]) <= 1e-6 * d["rho_max"], (d, go)
    assert n == len(P["x"])


def test_weights_with_tags_and_restart(tmp_path):
    nx = ny = 32
    P, wt = _weighted(nx, ny, 4, 8)
    tag = np.arange(len(P["x"]), dtype=np.int32)
    path = str(tmp_path / "weighted_state")
    with Simulation(nx, ny, epc=4, form="fp32_split", tile=8) as sim:
        sim.upload_particles(**P, tag=tag)
        sim.step(5)
        sim.save_state(path, 5)
        sim.step(5)
        A = sim.download_particles(with_tag=True)
    with Simulation(nx, ny, epc=4, form="fp32_split", tile=8) as sim:
        assert sim.load_state(path) == 5
        sim.step(5)
        B = sim.download_particles(with_tag=True)
    # every tag still carries the charge it was uploaded with
    assert np.array_equal(A["q"], P["q"][A["tag"]])
    oa, ob = np.argsort(A["tag"]), np.argsort(B["tag"])
    for c in ("q", "x", "y", "u", "v", "w"):
        assert np.array_equal(A[c][oa], B[c][ob]), c


def test_wrong_sign_charge_refused():
    P = setup_weibel(16, 16, epc=2, seed=1)
    P["q"][3] = -P["q"][3]
    with Simulation(16, 16, epc=2, form="fp32_split", tile=8) as sim:
        with pytest.raises(PictorError):
            sim.upload_particles(**P)
