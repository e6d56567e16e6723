"""3-D HIP path vs the FP64 CPU oracle, through the C ABI (needs an MI355X).

Same bars as the 2-D parity tests: FP64 form bit-exact for Yee/halo sweeps
and for particles (positions, momenta, cells, tiles) after a push from
identical state, including particles that sit in a ghost layer for a step
after a corner crossing; current and multi-step fields within the stated
tolerances.  The reference's 3-D deposit is not exactly charge conserving
(it omits the q dx dy dz / 12 term of the exact current), so the Gauss
residual is compared with the oracle's own residual rather than with zero.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import Simulation, setup_weibel
from tests._util import FIELDS, assert_same_multiset, rel_err, say, smooth_fields, sort_particles, tiles_of3

pytestmark = pytest.mark.gpu


def _pair3(n, epc, form="fp64", tile=4, seed=3, drift=0.3, uth=0.02, fields=True):
    nx, ny, nz = n
    say(f"3d n={n} epc={epc} form={form} tile={tile} seed={seed}")
    P = setup_weibel(nx, ny, nz, ndim=3, epc=epc, drift=drift, uth=uth, seed=seed)
    o = Oracle(3, nx, ny, nz, epc=epc, cap=2 * len(P["x"]) + 4096)
    o.load_particles(**P)
    sim = Simulation(nx, ny, nz, ndim=3, epc=epc, form=form, tile=tile)
    if fields:
        F = smooth_fields(o, seed=seed)
        sim.upload_fields(*[F[k] for k in FIELDS])
    sim.upload_particles(**P)
    return o, sim, P


def test_yee3d_and_halo_bitwise():
    nx, ny, nz = 16, 12, 8
    say("3d yee", nx, ny, nz)
    o = Oracle(3, nx, ny, nz, epc=2)
    F = smooth_fields(o, amp_e=1.0, amp_b=1.0, seed=11, modes=6)
    rng = np.random.default_rng(5)
    for k in FIELDS:
        F[k][3:-2, 3:-2, 3:-2] += 1e-3 * rng.standard_normal(F[k][3:-2, 3:-2, 3:-2].shape)
    o.set_fields(**F)
    o.field_halo("E"); o.field_halo("B")
    F = {k: o.F[k].copy(order="F") for k in FIELDS}
    with Simulation(nx, ny, nz, ndim=3, epc=2, form="fp64", tile=4) as sim:
        sim.upload_fields(*[F[k] for k in FIELDS])
        for _ in range(5):
            sim.UpdateBfieldHalf()
            o.update_b_half(); o.field_halo("B")
            sim.UpdateEfield()
            o.update_e(); o.field_halo("E")
        G = sim.download_fields()
    for k in FIELDS:
        assert np.array_equal(G[k], o.F[k]), k


@pytest.mark.parametrize("tile", [4, 8])
def test_push3d_one_step_bitwise_fp64(tile):
    n = (16, 16, 16)
    o, sim, P = _pair3(n, epc=2, tile=tile, seed=7, drift=0.5, uth=0.05)
    with sim:
        sim.ResetCurrent(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
        o.reset_current(); o.move_deposit(); o.exchange()
        G = sim.download_particles(with_tile=True)
        assert_same_multiset(G, o.particles())
        assert np.array_equal(G["tile"], tiles_of3(G, *n, tile))
        J = sim.download_fields()
        for k in ("Jx", "Jy", "Jz"):
            assert rel_err(J[k], o.F[k]) < 1e-13, k
        sim.UpdateCurrentAllFaces(); o.current_halo_add()
        J = sim.download_fields()
        for k in ("Jx", "Jy", "Jz"):
            assert rel_err(J[k], o.F[k]) < 1e-13, k


def test_push3d_corner_ghost_sitters_bitwise():
    # particles that leave through an edge or a corner of the periodic box:
    # the diagonal rules of the exchange wrap the face they left furthest
    # through and leave them one step in a ghost layer for the other faces
    nx = ny = nz = 8
    o = Oracle(3, nx, ny, nz, epc=2, cap=4096)
    lo, hi = o.p.xmin, o.p.xmax
    rng = np.random.default_rng(3)
    m = 400
    corner = rng.integers(0, 2, (m, 3))
    pos = np.where(corner == 0, lo + rng.uniform(0.0, 0.6, (m, 3)), hi - rng.uniform(0.0, 0.6, (m, 3)))
    vel = np.where(corner == 0, -1.0, 1.0) * rng.uniform(0.3, 3.0, (m, 3))
    P = {"q": np.where(np.arange(m) % 2 == 0, 1.0, -1.0), "x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
         "u": vel[:, 0], "v": vel[:, 1], "w": vel[:, 2], "flv": np.where(np.arange(m) % 2 == 0, 1, 2).astype(np.int32)}
    o.load_particles(**P)
    with Simulation(nx, ny, nz, ndim=3, epc=2, form="fp64", tile=4) as sim:
        sim.upload_particles(**P)
        ghosts = 0
        for step in range(4):
            say(f"corner step {step}")
            sim.ResetCurrent(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
            o.reset_current(); o.move_deposit(); o.exchange()
            R = o.particles()
            ghosts += int(np.sum((R["x"] < lo) | (R["x"] > hi) | (R["y"] < lo) | (R["y"] > hi)
                                 | (R["z"] < lo) | (R["z"] > hi)))
            G = sim.download_particles()
            assert_same_multiset(G, R)
            J = sim.download_fields()
            for k in ("Jx", "Jy", "Jz"):
                assert rel_err(J[k], o.F[k]) < 1e-13, k
        assert ghosts > 0


def test_steps3d_fp64_vs_oracle():
    n = (16, 16, 16)
    o, sim, P = _pair3(n, epc=4, seed=4, fields=False, drift=0.3, uth=0.05)
    nsteps = 20
    with sim:
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        G = sim.download_fields()
        d = sim.diagnostics()
        eo = o.energy()
        go = o.gauss()
        say("3d fp64:", d, eo, go)
        for k in FIELDS:
            assert rel_err(G[k], o.F[k]) < 1e-9, k
        assert abs(d["B_energy"] - eo[0]) <= 1e-9 * eo[0]
        assert abs(d["E_energy"] - eo[1]) <= 1e-9 * eo[1]
        assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-12 * eo[2]
        assert abs(d["prtl_energy_2"] - eo[3]) <= 1e-12 * eo[3]
        # the same residual as the oracle's (the scheme's own, not rounding)
        assert abs(d["gauss_residual"] - go[0]) <= 1e-6 * go[0] + 1e-13 * go[1], (d, go)
        assert abs(d["rho_max"] - go[1]) <= 1e-12 * go[1]
        assert sim.particle_count() == len(P["x"])


def test_split3d_vs_fp64_oracle():
    n = (16, 16, 16)
    o, sim, P = _pair3(n, epc=4, form="fp32_split", seed=4, fields=False, drift=0.3, uth=0.05)
    nsteps = 20
    with sim:
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        d = sim.diagnostics()
        eo = o.energy()
        go = o.gauss()
        say("3d split vs fp64:", d, eo, go)
        # FP32 arithmetic: field energy within 1e-3, particle energy within 1e-5;
        # Gauss residual within 1e-6 of max |rho| of the oracle's own residual
        assert abs(d["B_energy"] - eo[0]) <= 1e-3 * eo[0]
        assert abs(d["E_energy"] - eo[1]) <= 1e-3 * eo[1]
        assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-5 * eo[2]
        assert abs(d["prtl_energy_2"] - eo[3]) <= 1e-5 * eo[3]
        assert abs(d["gauss_residual"] - go[0]) <= 1e-6 * go[1], (d, go)
        assert sim.particle_count() == len(P["x"])


def test_split3d_deterministic():
    n = (16, 16, 8)
    P = setup_weibel(*n, ndim=3, epc=4, drift=0.3, uth=0.05, seed=9)
    res = []
    for rep in range(2):
        say(f"3d rep={rep}")
        with Simulation(*n, ndim=3, epc=4, form="fp32_split", tile=4) as sim:
            sim.upload_particles(**P)
            sim.step(15)
            res.append((sim.download_fields(), sim.diagnostics(), sort_particles(sim.download_particles())))
    for k in FIELDS + ("Jx", "Jy", "Jz"):
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
    for c in ("x", "y", "z", "u", "v", "w"):
        assert np.array_equal(res[0][2][c], res[1][2][c]), c
