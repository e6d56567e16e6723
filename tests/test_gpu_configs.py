"""The five benchmark configurations at full size on the GPU, through the C ABI.

C1 runs in full against the oracle; the others are too large for the oracle
and are checked through what does not depend on size.

Bars:
  * C1, 2-D Weibel 256^2, 16 ppc per species, 1000 steps, against the oracle
    on the host's threads: the FP64 form's fields within 1e-9 and energies
    within 1e-9 at every 100th step, Gauss residual within 1e-12 of max |rho|;
    the split form's field energy within 1e-4 of the oracle's at every 100th
    step and its Gauss residual within 1e-6 of max |rho|.
  * C2, 2-D Weibel 2048^2, 32 ppc per species (268,435,456 particles), and C3,
    3-D Weibel 256^3, 8 ppc per species (268,435,456 particles), 20 steps of
    each form: no error, the particle count kept, two split runs bitwise equal,
    split and FP64 field energies within 1e-5; Gauss residual within 1e-6
    (split) and 1e-12 (FP64) of max |rho| in 2-D; in 3-D the reference's
    zigzag drops its q dx dy dz / 12 term, so the split residual is held to
    twice the FP64 form's.
  * C4, the laser wakefield in the modal cylinder (4096 x 512, 31.5 M
    particles, Lehe, moving window, sorted by cell every 20 steps), 300 steps
    (240 window shifts): two split runs bitwise equal; per mode the discrete
    Gauss residual of the interior rows equals the initial one translated by
    the window's shifts, to 1e-11 (FP64) and 1e-8 (m = 0) / 1e-6 (m = 1) of its
    largest value for the split form, whose current is exact for its own FP32
    phases while the oracle's charge uses FP64 phases.
  * C5, the adjoint on 2-D Weibel 512^2, 16 ppc per species, 200 steps,
    checkpoints every 20: the engine's split objective equals the production
    split form's bit for bit; the split drift gradient within 3e-3 of the FP64
    adjoint's, the momentum gradients within 2e-2 (relative L2).
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import CylModalOracle, Oracle
from pictor_amd import Simulation, setup_weibel
from tests._util import FIELDS, rel_err, say

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))   # the host's share of cores per GPU
C, C_OMPE = 0.45, 10.0


def _field_energy(d):
    return d["B_energy"] + d["E_energy"]


def test_c1_weibel2d_256_against_the_oracle():
    nx = ny = 256
    epc, steps, every = 16, 1000, 100
    P = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=12345)
    n = len(P["x"])
    say(f"C1 oracle: {n} particles, {steps} steps on {THREADS} threads")
    o = Oracle(2, nx, ny, epc=epc, c=C, c_ompe=C_OMPE, nthreads=THREADS, cap=int(n * 1.05) + 4096)
    o.load_particles(**P)
    ref = []
    for t in range(1, steps + 1):
        o.step(t)
        if t % every == 0:
            e, g = o.energy(), o.gauss()
            ref.append({"B_energy": e[0], "E_energy": e[1], "prtl_energy_1": e[2], "prtl_energy_2": e[3],
                        "gauss_residual": g[0], "rho_max": g[1]})
    F_ref = {k: o.F[k].copy() for k in FIELDS}
    for form in ("fp64", "fp32_split"):
        say(f"C1 {form}")
        with Simulation(nx, ny, epc=epc, c=C, c_ompe=C_OMPE, form=form) as sim:
            sim.upload_particles(**P)
            worst = 0.0
            for i, r in enumerate(ref):
                sim.step(every)
                d = sim.diagnostics()
                err = abs(_field_energy(d) - _field_energy(r)) / _field_energy(r)
                worst = max(worst, err)
                if form == "fp64":
                    for k in ("B_energy", "E_energy", "prtl_energy_1", "prtl_energy_2"):
                        assert abs(d[k] - r[k]) <= 1e-9 * abs(r[k]), (i, k, d[k], r[k])
                    assert d["gauss_residual"] <= 1e-12 * d["rho_max"], (i, d)
                else:
                    assert err <= 1e-4, (i, err)
                    assert d["gauss_residual"] <= 1e-6 * d["rho_max"], (i, d)
            assert sim.particle_count() == n
            if form == "fp64":
                G = sim.download_fields()
                for k in FIELDS:
                    assert rel_err(G[k], F_ref[k]) <= 1e-9, k
            say(f"C1 {form}: max field-energy error vs the oracle {worst:.3e}")


def _weibel_big(ndim, nx, ny, nz, epc, drift, uth, steps=20):
    P = setup_weibel(nx, ny, nz, ndim=ndim, epc=epc, drift=drift, uth=uth, seed=12345)
    n = len(P["x"])
    out = {}
    for name, form in (("split_a", "fp32_split"), ("split_b", "fp32_split"), ("fp64", "fp64")):
        say(f"{ndim}-D {nx}x{ny}x{nz} {name}: {n} particles, {steps} steps")
        with Simulation(nx, ny, nz, ndim=ndim, epc=epc, c=C, c_ompe=C_OMPE, form=form) as sim:
            sim.upload_particles(**P)
            sim.step(steps)
            d = sim.diagnostics()
            assert sim.particle_count() == n, name
            G = sim.download_fields()
        out[name] = (d, G)
        say(f"{name}: {d}")
    # the split step is deterministic: fields, current and field energies repeat bit for
    # bit; the particle sums and the diagnostic's charge density add the particles in
    # their slot order within a bin (which varies from run to run) and with FP64
    # atomics, so they repeat to rounding
    for k in FIELDS + ("Jx", "Jy", "Jz"):
        assert np.array_equal(out["split_a"][1][k], out["split_b"][1][k]), k
    for k in ("B_energy", "E_energy"):
        assert out["split_a"][0][k] == out["split_b"][0][k], k
    for k in ("prtl_energy_1", "prtl_energy_2", "rho_max"):
        a, b = out["split_a"][0][k], out["split_b"][0][k]
        assert abs(a - b) <= 1e-11 * abs(b), (k, a, b)
    ga, gb = out["split_a"][0]["gauss_residual"], out["split_b"][0]["gauss_residual"]
    assert abs(ga - gb) <= 1e-12 * out["split_b"][0]["rho_max"], (ga, gb)
    ds, d64 = out["split_a"][0], out["fp64"][0]
    err = abs(_field_energy(ds) - _field_energy(d64)) / _field_energy(d64)
    say(f"split vs FP64 field energy after {steps} steps: {err:.3e}")
    assert err <= 1e-5
    return ds, d64


def test_c2_weibel2d_2048_full_size():
    ds, d64 = _weibel_big(2, 2048, 2048, 1, 32, 0.5, 0.01)
    assert ds["gauss_residual"] <= 1e-6 * ds["rho_max"], ds
    assert d64["gauss_residual"] <= 1e-12 * d64["rho_max"], d64


def test_c3_weibel3d_256_full_size():
    ds, d64 = _weibel_big(3, 256, 256, 256, 8, 0.3, 0.02)
    assert ds["gauss_residual"] <= 2.0 * d64["gauss_residual"] + 1e-11 * d64["rho_max"], (ds, d64)


# the laser wakefield: bench.py's lwfa_4096x512
LWFA = dict(nr=512, nz=4096, dr=4.0, cps=20.0, nrp=480, ppc=16, c=0.8)


def _lwfa(form, steps):
    p = LWFA
    a0, omega0, tau, w0 = 2.0, (1.0 / 0.005) ** 0.5, 10.0, 15.0
    zc = p["nz"] - 3.0 * tau * p["cps"]
    with Simulation(p["nr"], 2, p["nz"], ndim=3, epc=p["ppc"], c=p["c"], c_ompe=p["cps"], mi_me=1e30, form=form,
                    geometry="cyl_modal", dr=p["dr"], fdtd="lehe") as sim:
        sim.cyl_plasma(p["nrp"], 3, p["nz"], p["ppc"], seed=2024)
        sim.cyl_laser(a0, omega0, tau, w0, zc)
        F0, P0 = sim.download_fields(), sim.download_particles()
        sim.moving_window(True, nr_plasma=p["nrp"], ppc=p["ppc"], seed=2024)
        sim.cyl_sort_period(20)
        sim.step(steps)
        d = sim.diagnostics()
        F1, P1 = sim.download_fields(), sim.download_particles()
    return F0, P0, F1, P1, d


def _modal_gauss(F, P):
    p = LWFA
    cap = len(P["x"]) + 16
    o = CylModalOracle(p["nr"], p["nz"], dr=p["dr"], c=p["c"], c_ompe=p["cps"], epc=p["ppc"], cap=16)
    o.set_zopen(True)
    o.set_fields(**{k: F[k] for k in FIELDS})
    o.load_particles(P["q"], P["x"], P["y"], P["z"], P["u"], P["v"], P["w"], P["flv"], cap=cap)
    return o.gauss()


def test_c4_lwfa_4096x512_window_and_gauss():
    steps = 300
    shifts = int(steps * LWFA["c"] + 1e-9)
    runs = {}
    for name, form in (("split_a", "fp32_split"), ("split_b", "fp32_split"), ("fp64", "fp64")):
        say(f"C4 {name}: {steps} steps, {shifts} window shifts")
        runs[name] = _lwfa(form, steps)
        say(f"C4 {name}: {runs[name][4]}")
    a, b = runs["split_a"], runs["split_b"]
    for k in FIELDS:
        assert np.array_equal(a[2][k], b[2][k]), k
    for k in ("x", "y", "z", "u", "v", "w", "flv"):
        assert np.array_equal(a[3][k], b[3][k]), k
    for name, bars in (("fp64", (1e-11, 1e-11, 1e-11)), ("split_a", (1e-8, 1e-6, 1e-6))):
        F0, P0, F1, P1, _ = runs[name]
        g0, g1 = _modal_gauss(F0, P0), _modal_gauss(F1, P1)
        mz = g0.shape[2]
        # interior rows (1-based k = 7 .. mz - 9 - shifts) against the rows they came from
        lo, hi = 6, mz - 9 - shifts
        for s in range(3):
            scale = np.abs(g0[:, s, :]).max()
            d = np.abs(g1[:, s, lo:hi] - g0[:, s, lo + shifts:hi + shifts]).max() / scale
            say(f"C4 {name} mode slot {s}: Gauss change {d:.3e} of max {scale:.3e}")
            assert d <= bars[s], (name, s, d)


def test_c5_adjoint_weibel2d_512():
    import torch

    from pictor_amd.adjoint import ParticleStepObjective, drift_boost, drift_signs

    nx = ny = 512
    epc, nsteps, seg = 16, 200, 20
    P0 = setup_weibel(nx, ny, epc=epc, drift=0.0, uth=0.01, seed=12345)
    sgn = torch.from_numpy(drift_signs(len(P0["x"]), epc))
    th = [torch.from_numpy(P0[k]) for k in ("u", "v", "w")]
    res = {}
    for form in ("fp32_split", "fp64"):
        say(f"C5 {form}: {len(P0['x'])} particles, {nsteps} steps, segment {seg}")
        d = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
        tl = [t.clone().requires_grad_(True) for t in th]
        u, v, w = drift_boost(tl[0], tl[1], tl[2], sgn, d)
        for t in (u, v, w):
            t.retain_grad()
        with Simulation(nx, ny, epc=epc, c=C, c_ompe=C_OMPE, form=form) as sim:
            L = ParticleStepObjective.apply(u, v, w, sim, P0, nsteps, seg)
            L.backward()
            st = sim.adjoint_stats()
            if form == "fp32_split":
                P = dict(P0)
                P["u"], P["v"], P["w"] = u.detach().numpy(), v.detach().numpy(), w.detach().numpy()
                sim.upload_particles(**P)
                sim.step(nsteps)
                L_prod = sim.diagnostics()["B_energy"]
                say(f"C5 split objective {float(L)!r}, production forward {L_prod!r}")
                assert float(L) == L_prod
        res[form] = (float(d.grad), [t.grad.numpy().copy() for t in (u, v, w)])
        say(f"C5 {form}: L={float(L):.9e} dL/ddrift={float(d.grad):.6e} stats={st}")
    g32, g64 = res["fp32_split"], res["fp64"]
    rd = abs(g32[0] - g64[0]) / abs(g64[0])
    num = np.sqrt(sum(np.sum((g32[1][k] - g64[1][k]) ** 2) for k in range(3)))
    den = np.sqrt(sum(np.sum(g64[1][k] ** 2) for k in range(3)))
    say(f"C5 split vs FP64: drift {rd:.3e}, momenta (relative L2) {num / den:.3e}")
    assert rd <= 3e-3
    assert num / den <= 2e-2
