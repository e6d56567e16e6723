"""Boundary faces on the GPU against the oracle, through the C ABI (needs an MI355X).

Bars:
  * one push with reflecting / removing particle faces and non-periodic axes
    (FP64 form): the particle multiset bit for bit, the current within 1e-13
    before and after the halo (the GPU sums it in another order);
  * vacuum fields with conducting, open and PML faces: bit for bit over 30
    steps (no current, so no summation order);
  * whole steps with plasma and faces: fields within 1e-9, energies within
    1e-9, the same particle count (FP64); the split form within 2e-3 of the
    FP64 field energy and within 0.5 % of its particle count;
  * the phase calls equal the fused step to 1e-12;
  * invalid faces are refused.
"""
import ctypes

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import PictorError, Simulation, setup_weibel
from pictor_amd._lib import PICTOR_ERR_ARG, lib
from tests._util import FIELDS, assert_same_multiset, rel_err, say

pytestmark = pytest.mark.gpu

# (name, ndim, particle faces, field faces): face = (axis, side, type, position[, thickness])
SCEN = {
    "refl_x": (2, [("x", "lower", "refl", 4.0), ("x", "upper", "refl", 27.5)], [("x", "lower", "cond", 3.0),
                                                                                ("x", "upper", "cond", 28.0)]),
    "refl_y": (2, [("y", "lower", "refl", 3.25), ("y", "upper", "refl", 20.0)], [("y", "lower", "cond", 2.0),
                                                                                 ("y", "upper", "cond", 21.0)]),
    "remv_pml": (2, [("y", "lower", "remv", 2.5), ("y", "upper", "remv", 21.5)], [("y", "lower", "pml", 6.0, 4.0),
                                                                                  ("y", "upper", "pml", 18.0, 4.0)]),
    "open_x": (2, [("x", "lower", "remv", 1.0), ("x", "upper", "remv", 31.0)], [("x", "lower", "open", 5.0),
                                                                                ("x", "upper", "open", 27.0)]),
    "mixed": (2, [("x", "lower", "refl", 4.5), ("x", "upper", "remv", 28.0), ("y", "lower", "remv", 2.0),
                  ("y", "upper", "refl", 21.0)], [("x", "upper", "pml", 24.0, 6.0)]),
    "walls3d": (3, [("x", "lower", "refl", 3.0), ("x", "upper", "refl", 13.0), ("z", "lower", "remv", 1.5),
                    ("z", "upper", "remv", 14.5)], [("z", "lower", "pml", 5.0, 3.0), ("y", "lower", "cond", 2.0)]),
}
SIZE = {2: (32, 24, 1), 3: (16, 16, 16)}


def _inside(P, faces):
    keep = np.ones(len(P["x"]), bool)
    for axis, side, kind, pos, *_ in faces:
        c = P[axis] - 3.0
        keep &= (c > pos + 0.02) if side == "lower" else (c < pos - 0.02)
    return {k: v[keep] for k, v in P.items()}


def _build(name, form="fp64", epc=8, seed=3, drift=0.5, uth=0.3, moving=False, with_fields=True):
    ndim, pfaces, ffaces = SCEN[name]
    nx, ny, nz = SIZE[ndim]
    say(f"bc scenario {name} form={form} nx={nx} ny={ny} nz={nz}")
    P = setup_weibel(nx, ny, nz, ndim=ndim, epc=epc, drift=drift, uth=uth, seed=seed)
    P = _inside(P, pfaces)
    o = Oracle(ndim, nx, ny, nz, epc=epc, cap=2 * len(P["x"]) + 4096)
    sim = Simulation(nx, ny, nz, ndim=ndim, epc=epc, form=form, tile=8 if ndim == 2 else 4)
    for axis, side, kind, pos, *rest in ffaces:
        o.set_bc_fld(axis, pos, kind, side, *rest)
        sim.SetBC_Fld(axis, pos, kind, side, *rest)
    for axis, side, kind, pos in pfaces:
        o.set_bc_prtl(axis, pos, kind, side)
        sim.SetBC_Prtl(axis, pos, kind, side)
    if moving:
        axis, side = pfaces[1][0], pfaces[1][1]
        o.set_boundary_speed(axis, side, -0.2)
        sim.SetBoundarySpeed(axis, side, -0.2)
    o.load_particles(**P)
    if with_fields:
        rng = np.random.default_rng(seed)
        F = {n: np.asfortranarray(0.02 * rng.standard_normal(o.shape)) for n in FIELDS}
        o.set_fields(**F)
        o.field_halo("E"); o.field_halo("B")
        F = {n: o.F[n].copy(order="F") for n in FIELDS}
        sim.upload_fields(*[F[n] for n in FIELDS])
    sim.upload_particles(**P)
    return o, sim, P


@pytest.mark.parametrize("name", list(SCEN))
def test_bc_one_push_bitwise_fp64(name):
    o, sim, P = _build(name)
    with sim:
        sim.ResetCurrent(); sim.MoveDepositPrtl()
        o.reset_current(); o.move_deposit(); o.bc_post_movdep()
        sim.ExchangePrtl(); o.exchange()
        G = sim.download_particles()
        R = o.particles()
        assert_same_multiset(G, R)
        if any(f[2] == "remv" for f in SCEN[name][1]):
            assert len(G["x"]) < len(P["x"])      # some particles left through a removing face
        J = sim.download_fields()
        for n in ("Jx", "Jy", "Jz"):
            assert rel_err(J[n], o.F[n]) < 1e-13, n
        sim.UpdateCurrentAllFaces(); o.current_halo_add()
        J = sim.download_fields()
        for n in ("Jx", "Jy", "Jz"):
            assert rel_err(J[n], o.F[n]) < 1e-13, n
        assert len(G["x"]) > 0


@pytest.mark.parametrize("name", ["refl_x", "refl_y", "remv_pml", "open_x", "mixed", "walls3d"])
def test_bc_vacuum_fields_bitwise(name):
    ndim, _, ffaces = SCEN[name]
    nx, ny, nz = SIZE[ndim]
    o = Oracle(ndim, nx, ny, nz, epc=1, cap=16)
    with Simulation(nx, ny, nz, ndim=ndim, epc=1, form="fp64", tile=8 if ndim == 2 else 4) as sim:
        for axis, side, kind, pos, *rest in ffaces:
            o.set_bc_fld(axis, pos, kind, side, *rest)
            sim.SetBC_Fld(axis, pos, kind, side, *rest)
        rng = np.random.default_rng(9)
        F = {n: np.asfortranarray(rng.standard_normal(o.shape)) for n in FIELDS}
        o.set_fields(**F)
        sim.upload_fields(*[F[n] for n in FIELDS])
        say(f"vacuum {name}: 30 steps")
        for t in range(1, 31):
            o.step(t)
        sim.step(30)
        G = sim.download_fields()
        for n in FIELDS:
            assert np.array_equal(G[n], o.F[n]), (n, np.max(np.abs(G[n] - o.F[n])))
        if any(f[2] == "pml" for f in ffaces):
            assert np.max(np.abs(G["Bz"])) > 0


@pytest.mark.parametrize("name,moving", [("refl_x", False), ("refl_x", True), ("remv_pml", False),
                                         ("open_x", False), ("mixed", False), ("walls3d", False)])
def test_bc_steps_fp64_vs_oracle(name, moving):
    o, sim, P = _build(name, moving=moving)
    nsteps = 30
    with sim:
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        G = sim.download_fields()
        for n in FIELDS:
            assert rel_err(G[n], o.F[n]) < 1e-9, n
        d = sim.diagnostics()
        eo = o.energy()
        assert abs(d["B_energy"] - eo[0]) <= 1e-9 * eo[0]
        assert abs(d["E_energy"] - eo[1]) <= 1e-9 * eo[1]
        assert sim.particle_count() == len(o.particles()["x"])


@pytest.mark.parametrize("name", ["refl_x", "remv_pml", "mixed", "walls3d"])
def test_bc_split_form_against_fp64(name):
    res = []
    for form in ("fp64", "fp32_split"):
        o, sim, P = _build(name, form=form)
        with sim:
            sim.step(30)
            d = sim.diagnostics()
            res.append((d["B_energy"] + d["E_energy"], sim.particle_count()))
    (e64, n64), (e32, n32) = res
    assert abs(e32 - e64) <= 2e-3 * e64, (e32, e64)
    assert abs(n32 - n64) <= 0.005 * n64, (n32, n64)


def test_bc_phases_equal_fused():
    res = []
    for fused in (False, True):
        o, sim, P = _build("mixed")
        with sim:
            for _ in range(4):
                if fused:
                    sim.step(1)
                else:
                    sim.step_phases()
            res.append((sim.download_fields(), sim.particle_count()))
    for n in FIELDS:
        assert rel_err(res[0][0][n], res[1][0][n]) < 1e-12, n
    assert res[0][1] == res[1][1]


def test_bc_refusals():
    with Simulation(32, 32, epc=2, form="fp64", tile=8) as sim:
        for args in (("z", 3.0, "refl", "lower"),):
            with pytest.raises(PictorError):
                sim.SetBC_Prtl(*args)
        with pytest.raises(PictorError):
            sim.SetBC_Fld("y", 2.0, "open", "lower")            # open faces are x only
        with pytest.raises(PictorError):
            sim.SetBC_Fld("x", 2.0, "pml", "lower", 4.0)        # the layer leaves the box
        with pytest.raises(PictorError):
            sim.SetBC_Fld("x", 3.0, "prdc", "lower")
        sim.SetBC_Prtl("x", 3.0, "refl", "lower")
        P = setup_weibel(32, 32, epc=2, seed=1)
        L = lib()
        obj = ctypes.c_double()
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))   # noqa: E731
        rc = L.pictor_adjoint_forward(sim._ctx, len(P["x"]), *[dp(P[k]) for k in ("q", "x", "y", "z", "u", "v", "w")],
                                      P["flv"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 2, 1,
                                      ctypes.byref(obj))
        assert rc == PICTOR_ERR_ARG and b"boundary" in L.pictor_last_error()
    with Simulation(32, 32, epc=2, form="fp64", tile=8, curr_filter=1) as sim:
        with pytest.raises(PictorError):
            sim.SetBC_Prtl("x", 3.0, "refl", "lower")
    with Simulation(16, 16, 16, ndim=3, epc=1, form="fp64", tile=4) as sim:
        with pytest.raises(PictorError):
            sim.SetBC_Fld("z", 3.0, "cond", "lower")


@pytest.mark.parametrize("form", ["fp32_split", "fp64"])
def test_bc_restart_continues_the_run(form, tmp_path):
    # faces (with a moving one and a PML) travel in the restart file; the split form
    # continues bit for bit, the FP64 form to the rounding of its current sums
    ndim, pfaces, ffaces = SCEN["mixed"]
    runs = []
    for split in (False, True):
        o, sim, P = _build("mixed", form=form, moving=True)
        with sim:
            sim.step(6)
            if split:
                path = str(tmp_path / f"state_{form}")
                sim.save_state(path, 6)
                nx, ny, nz = SIZE[ndim]
                with Simulation(nx, ny, nz, ndim=ndim, epc=8, form=form, tile=8) as sim2:
                    assert sim2.load_state(path) == 6
                    sim2.step(6)
                    runs.append((sim2.download_fields(), sim2.download_particles()))
            else:
                sim.step(6)
                runs.append((sim.download_fields(), sim.download_particles()))
    (Fa, Pa), (Fb, Pb) = runs
    if form == "fp32_split":
        for n in FIELDS:
            assert np.array_equal(Fa[n], Fb[n]), n
        assert_same_multiset(Pa, Pb)
    else:
        for n in FIELDS:
            assert rel_err(Fa[n], Fb[n]) < 1e-12, n
        assert len(Pa["x"]) == len(Pb["x"])
