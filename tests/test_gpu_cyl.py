"""Cylindrical geometry (the reference's -Dcyl build with the axis) on the GPU
vs the cylindrical oracle, through the C ABI (needs an MI355X).

Bars (the Boris push of cyl_movdep.F90 and the Strang push of
cyl_movdep_strang.F90 alike):
  * field updates, axis fills and halos (cyl_em_update.F90, Yee 2-D/3-D and
    Lehe-cyl 3-D) bit-exact against the oracle from identical fields;
  * one push from identical state: positions and momenta within 1e-13
    relative (the GPU's atan may differ from the host's in the last bit),
    current within 1e-12 of its maximum (atomic summation order), the same
    particles going through the axis sub-steps;
  * 20 full steps: fields and particles within 1e-9 relative;
  * the GPU run keeps the discrete Gauss residual r dtheta div E - rho
    constant node by node away from the axis (a size-independent property,
    checked on a larger grid than the oracle comparison);
  * a Courant number at or above the limit of the cells next to the axis
    (0.5 dtheta wide in theta), where the reference's scheme diverges, is
    rejected at pictor_init.
"""
import numpy as np
import pytest

from oracle.pyoracle import CylOracle
from pictor_amd import Simulation
from tests._cyl_util import gauss, interior, plasma
from tests._util import say

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


# c below the Courant limit of the cells next to the axis (0.193 for 16 theta cells in 2-D, 0.189 in 3-D)
C = 0.15


def _pair(ndim, nr, nth, nz=1, fdtd="yee", seed=1, n=0, rmax=None, epc=4, push="boris"):
    say(f"cyl ndim={ndim} nr={nr} ntheta={nth} nz={nz} fdtd={fdtd} seed={seed} n={n} push={push}")
    o = CylOracle(ndim, nr, nth, nz, c=C, epc=epc, fdtd=fdtd, cap=max(n, 1) + 16, push=push)
    sim = Simulation(nr, nth, nz, ndim=ndim, c=C, epc=epc, form="fp64", fdtd=fdtd, geometry="cyl")
    sim.cyl_push_scheme(push)
    rng = np.random.default_rng(seed)
    P = None
    if n:
        P = plasma(o, n, rng, rmax if rmax else nr - 8.0)
        o.load_particles(*P)
        sim.upload_particles(*P)
    return o, sim, rng, P


def _random_fields(o, sim, rng, amp=1e-3):
    F = {k: amp * rng.normal(0, 1, o.shape) for k in FIELDS}
    o.set_fields(**F)
    sim.upload_fields(*[np.asfortranarray(F[k]) for k in FIELDS])


def _same(a, b, what):
    d = np.abs(a - b).max()
    assert d == 0.0, f"{what}: max difference {d}"


@pytest.mark.parametrize("ndim,fdtd", [(2, "yee"), (3, "yee"), (3, "lehe")])
def test_cyl_field_updates_bitwise(ndim, fdtd):
    o, sim, rng, _ = _pair(ndim, 16, 16, 10 if ndim == 3 else 1, fdtd)
    _random_fields(o, sim, rng)
    for _ in range(3):
        sim.UpdateBfieldHalf()
        o.update_b_half()
        o.field_halo("B")
        sim.UpdateEfield()
        o.update_e()
        o.field_halo("E")
    G = sim.download_fields()
    for k in FIELDS:
        _same(G[k], o.F[k], k)
    sim.close()


def _push_once(o, sim):
    sim.ResetCurrent()
    sim.MoveDepositPrtl()
    sim.ExchangePrtl()
    o.reset_current()
    ns = o.move_deposit()
    o.exchange()
    return ns


@pytest.mark.parametrize("ndim,push", [(2, "boris"), (3, "boris"), (2, "strang"), (3, "strang")])
def test_cyl_push_matches_oracle(ndim, push):
    o, sim, rng, P = _pair(ndim, 24, 16, 12 if ndim == 3 else 1, n=6000, rmax=16.0, push=push)
    _random_fields(o, sim, rng, amp=2e-2)
    # fields with consistent ghosts and axis values, as after a step
    sim.UpdateBfieldHalf(); o.update_b_half(); o.field_halo("B")
    sim.UpdateEfield(); o.update_e(); o.field_halo("E")
    ns = _push_once(o, sim)
    assert ns > 0, "no particle went through the axis sub-steps"
    A = sim.download_particles()
    B = o.particles()
    assert len(A["x"]) == len(B["x"])
    for k in ("x", "y", "z", "u", "v", "w"):
        d = np.abs(A[k] - B[k]) / np.maximum(np.abs(B[k]), 1.0)
        say(f"{k}: max rel diff {d.max():.3e}")
        assert d.max() <= 1e-13, k
    G = sim.download_fields()
    for k in ("Jx", "Jy", "Jz"):
        scale = np.abs(o.F[k]).max()
        assert scale > 0
        d = np.abs(G[k] - o.F[k]).max() / scale
        say(f"{k}: max diff / max {d:.3e}")
        assert d <= 1e-12, k
    sim.close()


@pytest.mark.parametrize("ndim,push", [(2, "boris"), (3, "boris"), (2, "strang"), (3, "strang")])
def test_cyl_steps_match_oracle(ndim, push):
    o, sim, rng, P = _pair(ndim, 20, 16, 10 if ndim == 3 else 1, n=3000, rmax=11.0, push=push)
    nsteps = 20
    sim.step(nsteps)
    o.run(0, nsteps)
    G = sim.download_fields()
    for k in FIELDS:
        scale = np.abs(o.F[k]).max()
        d = np.abs(G[k] - o.F[k]).max() / scale
        say(f"{k}: max diff / max {d:.3e} (max {scale:.3e})")
        assert scale > 0 and d <= 1e-9, k
    A = sim.download_particles()
    B = o.particles()
    for k in ("x", "y", "z", "u", "v", "w"):
        d = np.abs(A[k] - B[k]).max() / max(np.abs(B[k]).max(), 1.0)
        assert d <= 1e-9, k
    # CalcEnergy with -Dcyl on the GPU state equals the oracle's on the same state
    dg = sim.diagnostics()
    o.set_fields(**{k: G[k] for k in FIELDS})
    o.load_particles(A["q"], A["x"], A["y"], A["z"], A["u"], A["v"], A["w"], A["flv"])
    eo = o.energy()
    for i, key in enumerate(("B_energy", "E_energy", "prtl_energy_1", "prtl_energy_2")):
        say(f"{key}: GPU {dg[key]:.12e} oracle {eo[i]:.12e}")
        assert eo[i] > 0 and abs(dg[key] - eo[i]) <= 1e-12 * eo[i], key
    sim.close()


def test_cyl_gpu_run_conserves_charge():
    # 2-D r-theta, 96 x 16 cells, 60 steps: G constant at nodes 5 <= i <= mx-5
    o = CylOracle(2, 96, 16, c=C, epc=8, cap=16)
    rng = np.random.default_rng(4)
    P = plasma(o, 100000, rng, 80.0)
    sim = Simulation(96, 16, 1, ndim=2, c=C, epc=8, form="fp64", geometry="cyl")
    sim.upload_particles(*P)
    o.load_particles(*P)
    G0 = gauss(o)[interior(o)]
    sim.step(60)
    F = sim.download_fields()
    o.set_fields(**{k: F[k] for k in FIELDS})
    A = sim.download_particles()
    o.load_particles(A["q"], A["x"], A["y"], A["z"], A["u"], A["v"], A["w"], A["flv"])
    G1 = gauss(o)[interior(o)]
    d = np.abs(G1 - G0).max() / np.abs(G0).max()
    say(f"Gauss residual change / max {d:.3e}")
    assert d <= 1e-11
    sim.close()


def test_cyl_rejects_a_particle_at_the_wall():
    sim = Simulation(16, 16, 1, ndim=2, c=C, epc=4, form="fp64", geometry="cyl")
    x = np.array([17.95])               # mx - 3 = 18: the first push moves it past the wall limit
    P = dict(q=np.ones(1), x=x, y=np.array([5.0]), z=np.array([1.5]), u=np.array([2.0]), v=np.zeros(1),
             w=np.zeros(1), flv=np.ones(1, dtype=np.int32))
    sim.upload_particles(**P)
    with pytest.raises(RuntimeError):
        sim.step(1)
    sim.close()


@pytest.mark.parametrize("ndim,ny,c", [(2, 16, 0.2), (3, 16, 0.19), (2, 8, 0.37)])
def test_cyl_rejects_a_courant_number_above_the_axis_limit(ndim, ny, c):
    # the cells next to the axis are 0.5 dtheta wide in theta: the reference's scheme diverges there
    with pytest.raises(RuntimeError):
        Simulation(16, ny, 8 if ndim == 3 else 1, ndim=ndim, c=c, epc=4, form="fp64", geometry="cyl")
