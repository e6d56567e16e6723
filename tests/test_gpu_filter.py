"""Current filter (smoothen_current, current.F90:19-30) on the GPU vs the oracle.

Bars:
  * one pictor_smoothen_current from the same J: the physical cells equal the
    oracle's bit for bit in the FP64 form, 2-D and 3-D, 1 and 3 passes
    (the FP32 split form filters its fixed-point counts, so 1e-15 relative);
  * 40 steps with curr_filter = 1, FP64: fields within 1e-9 relative of the
    oracle, particle energies within 1e-12 (the J summation order aside, the
    same bar as the unfiltered run);
  * the fused step equals the phase-by-phase calls (bit for bit in the split
    form, whose current is fixed-point; to 1e-12 in the FP64 form);
  * the FP64 adjoint with curr_filter = 1 matches central differences of the
    objective (the filter is its own transpose on the periodic box);
  * the split adjoint engine's objective equals the production forward's.
"""
import numpy as np
import pytest
import torch

from oracle.pyoracle import Oracle
from pictor_amd import Simulation, setup_weibel
from tests._util import FIELDS, rel_err, say

pytestmark = pytest.mark.gpu

J = ("Jx", "Jy", "Jz")


def _phys(a, ndim):
    return a[2:-3, 2:-3, :] if ndim == 2 else a[2:-3, 2:-3, 2:-3]


def _deposit(sim):
    sim.ResetCurrent(); sim.UpdateBfieldHalf(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
    sim.UpdateCurrentAllFaces()


@pytest.mark.parametrize("ndim,passes,form", [(2, 1, "fp64"), (2, 3, "fp64"), (3, 1, "fp64"), (3, 2, "fp64"),
                                              (2, 2, "fp32_split")])
def test_filter_matches_oracle(ndim, passes, form):
    n = (32, 24, 1) if ndim == 2 else (16, 16, 16)
    say(f"[case] filter ndim={ndim} passes={passes} form={form}")
    P = setup_weibel(n[0], n[1], n[2], ndim=ndim, epc=4, drift=0.5, uth=0.05, seed=21)
    with Simulation(n[0], n[1], n[2], ndim=ndim, epc=4, form=form, tile=8 if ndim == 2 else 4,
                    curr_filter=passes) as sim:
        sim.upload_particles(**P)
        _deposit(sim)
        J0 = {k: v.copy() for k, v in sim.download_fields().items() if k in J}
        sim.smoothen_current()
        J1 = {k: v for k, v in sim.download_fields().items() if k in J}
    o = Oracle(ndim, *n, epc=4, curr_filter=passes)
    o.set_fields(**J0)
    o.smoothen_current()
    for k in J:
        a, b = _phys(J1[k], ndim), _phys(o.F[k], ndim)
        assert np.abs(b).max() > 0, k
        if form == "fp64":
            assert np.array_equal(a, b), f"{k}: max diff {np.abs(a - b).max()}"
        else:
            assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max(), k


def test_steps_with_filter_match_oracle():
    nx, ny, epc, nsteps = 32, 32, 16, 40
    say(f"[case] filtered steps nx={nx} ny={ny} epc={epc}")
    P = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=4)
    o = Oracle(2, nx, ny, epc=epc, cap=2 * len(P["x"]) + 4096, curr_filter=1)
    o.load_particles(**P)
    with Simulation(nx, ny, epc=epc, form="fp64", tile=8, curr_filter=1) as sim:
        sim.upload_particles(**P)
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        G = sim.download_fields()
        d = sim.diagnostics()
    eo = o.energy()
    for k in FIELDS:
        assert rel_err(G[k], o.F[k]) < 1e-9, k
    assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-12 * eo[2]
    assert abs(d["B_energy"] - eo[0]) <= 1e-9 * eo[0]
    # the filtered run differs from the unfiltered one
    o2 = Oracle(2, nx, ny, epc=epc, cap=2 * len(P["x"]) + 4096, curr_filter=0)
    o2.load_particles(**P)
    o2.run(1, nsteps)
    assert abs(o2.energy()[0] - eo[0]) > 1e-6 * eo[0]


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
def test_fused_step_with_filter_equals_phases(form):
    say(f"[case] fused vs phases with filter form={form}")
    P = setup_weibel(32, 32, epc=8, drift=0.5, uth=0.01, seed=8)
    out = []
    for fused in (True, False):
        with Simulation(32, 32, epc=8, form=form, tile=8, curr_filter=2) as sim:
            sim.upload_particles(**P)
            for _ in range(6):
                if fused:
                    sim.step(1)
                else:
                    sim.step_phases()
            out.append(sim.download_fields())
    for k in FIELDS:
        if form == "fp32_split":   # fixed-point current: the two runs agree bit for bit
            assert np.array_equal(out[0][k], out[1][k]), k
        else:                      # FP64 atomics: two runs agree to rounding
            assert rel_err(out[0][k], out[1][k]) < 1e-12, k


def _objective(sim, P0, u, v, w, nsteps, seg):
    return sim.adjoint_forward(P0["q"], P0["x"], P0["y"], P0["z"], u, v, w, P0["flv"], nsteps, seg)


def test_fp64_adjoint_with_filter_matches_finite_differences():
    from pictor_amd.adjoint import ParticleStepObjective
    nx = ny = 16
    epc, nsteps, seg = 4, 6, 3
    say(f"[case] adjoint with filter nx={nx} epc={epc} nsteps={nsteps}")
    P0 = setup_weibel(nx, ny, epc=epc, drift=0.3, uth=0.05, seed=31)
    with Simulation(nx, ny, epc=epc, form="fp64", tile=8, curr_filter=2) as sim:
        u = torch.tensor(P0["u"], requires_grad=True)
        v = torch.tensor(P0["v"], requires_grad=True)
        w = torch.tensor(P0["w"], requires_grad=True)
        L = ParticleStepObjective.apply(u, v, w, sim, P0, nsteps, seg)
        L.backward()
        rng = np.random.default_rng(2)
        dirs = [rng.normal(size=len(P0["u"])) for _ in range(3)]
        eps = 1e-7
        for comp, name in enumerate(("u", "v", "w")):
            g = (u, v, w)[comp].grad.numpy()
            d = dirs[comp]
            mom = [P0["u"].copy(), P0["v"].copy(), P0["w"].copy()]
            mom[comp] = P0[name] + eps * d
            lp = _objective(sim, P0, *mom, nsteps, seg)
            mom[comp] = P0[name] - eps * d
            lm = _objective(sim, P0, *mom, nsteps, seg)
            fd = (lp - lm) / (2 * eps)
            ad = float(np.dot(g, d))
            say(f"{name}: fd {fd:.10e} adjoint {ad:.10e}")
            assert abs(fd - ad) <= 1e-3 * abs(fd), name


def test_split_engine_objective_with_filter_equals_production():
    nx = ny = 32
    epc, nsteps, seg = 8, 8, 4
    say("[case] split engine with filter")
    P0 = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=5)
    with Simulation(nx, ny, epc=epc, form="fp32_split", tile=8, curr_filter=1) as sim:
        L = _objective(sim, P0, P0["u"], P0["v"], P0["w"], nsteps, seg)
        sim.upload_particles(**P0)
        sim.step(nsteps)
        assert L == sim.diagnostics()["B_energy"]


@pytest.mark.parametrize("ndim,passes,form", [(2, 1, "fp64"), (2, 3, "fp64"), (3, 2, "fp64"), (2, 2, "fp32_split")])
def test_filtered_efield_matches_oracle(ndim, passes, form):
    # SetFilteredEfldGPU (fields_gpu.F90:70-90): E through `passes` rounds of the same
    # 1-2-1 filter, then the halo copy; the oracle filters the same values as its J
    n3 = (32, 24, 1) if ndim == 2 else (12, 12, 12)
    say(f"filtered E ndim={ndim} passes={passes} form={form}")
    o = Oracle(ndim, *n3, epc=1, cap=16, curr_filter=passes)
    rng = np.random.default_rng(passes + 10 * ndim)
    F = {n: np.asfortranarray(rng.standard_normal(o.shape)) for n in FIELDS}
    o.set_fields(**F)
    o.field_halo("E"); o.field_halo("B")
    with Simulation(*n3, ndim=ndim, epc=1, form=form, tile=8 if ndim == 2 else 4) as sim:
        sim.upload_fields(*[o.F[n] for n in FIELDS])
        G = sim.filtered_efield(passes)
        E0 = sim.download_fields()
    for a, b in (("Ex", "Jx"), ("Ey", "Jy"), ("Ez", "Jz")):
        o.F[b][...] = E0[a]                 # the context's E as the GPU holds it (FP32 pair in split)
    o.smoothen_current()
    o.field_halo("J")
    for a, b in (("Ex", "Jx"), ("Ey", "Jy"), ("Ez", "Jz")):
        assert rel_err(_phys(G[a], ndim), _phys(o.F[b], ndim)) <= 1e-15, a
        assert not np.array_equal(_phys(G[a], ndim), _phys(E0[a], ndim))   # the filter did act
