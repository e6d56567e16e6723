"""HIP path vs the FP64 CPU oracle, through the C ABI (needs an MI355X).

Bars: FP64 form bit-exact for fields after Yee/halo sweeps and for particle
positions, momenta, cell indices and tile assignment after one push from
identical state; current and multi-step fields within a stated tolerance
(the current is summed in a different order on the GPU); Gauss's law at
rounding level.  The FP32 split form is checked against the FP64 oracle with
the tolerances written in each test, and for bitwise run-to-run determinism.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import PictorError, Simulation, setup_weibel
from tests._util import FIELDS, assert_same_multiset, rel_err, say, smooth_fields, sort_particles, tiles_of

pytestmark = pytest.mark.gpu


def _pair(nx, ny, epc, form="fp64", tile=8, seed=3, drift=0.5, uth=0.01, fields=True, c=0.45):
    say(f"nx={nx} ny={ny} epc={epc} form={form} tile={tile} seed={seed}")
    P = setup_weibel(nx, ny, epc=epc, drift=drift, uth=uth, seed=seed)
    o = Oracle(2, nx, ny, epc=epc, c=c, cap=2 * len(P["x"]) + 4096)
    o.load_particles(**P)
    sim = Simulation(nx, ny, epc=epc, c=c, form=form, tile=tile)
    if fields:
        F = smooth_fields(o, seed=seed)
        sim.upload_fields(*[F[n] for n in FIELDS])
    sim.upload_particles(**P)
    return o, sim, P


def test_yee_and_halo_bitwise():
    nx, ny = 40, 24
    o = Oracle(2, nx, ny, epc=2)
    F = smooth_fields(o, amp_e=1.0, amp_b=1.0, seed=11, modes=6)
    rng = np.random.default_rng(5)
    for n in FIELDS:   # add small-scale noise so that every stencil term matters
        F[n][3:-2, 3:-2, :] += 1e-3 * rng.standard_normal(F[n][3:-2, 3:-2, :].shape)
    o.set_fields(**F)
    o.field_halo("E"); o.field_halo("B")
    F = {n: o.F[n].copy(order="F") for n in FIELDS}
    with Simulation(nx, ny, epc=2, form="fp64", tile=8) as sim:
        sim.upload_fields(*[F[n] for n in FIELDS])
        for _ in range(7):
            sim.UpdateBfieldHalf()
            o.update_b_half(); o.field_halo("B")
            sim.UpdateEfield()
            o.update_e(); o.field_halo("E")
        G = sim.download_fields()
    for n in FIELDS:
        assert np.array_equal(G[n], o.F[n]), n


# every tile-size instantiation of the push kernels (tile 8: 256-lane
# workgroups; 16: 512 lanes; 32: 1024 lanes), with and without the XCD tile
# remap (number of tiles divisible by 8)
TILE_CASES = [(32, 24, 8), (64, 32, 8), (64, 64, 16), (48, 32, 16), (128, 64, 32), (64, 64, 32)]


@pytest.mark.parametrize("nx,ny,tile", TILE_CASES)
def test_push_one_step_bitwise_fp64(nx, ny, tile):
    o, sim, P = _pair(nx, ny, epc=8, seed=7, tile=tile)
    with sim:
        sim.ResetCurrent(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
        o.reset_current(); o.move_deposit(); o.exchange()
        G = sim.download_particles(with_tile=True)
        assert_same_multiset(G, o.particles())
        # particle -> tile assignment is a function of the cell index
        assert np.array_equal(G["tile"], tiles_of(G, nx, ny, tile))
        J = sim.download_fields()
        for n in ("Jx", "Jy", "Jz"):
            assert rel_err(J[n], o.F[n]) < 1e-13, n
        sim.UpdateCurrentAllFaces(); o.current_halo_add()
        J = sim.download_fields()
        for n in ("Jx", "Jy", "Jz"):
            assert rel_err(J[n], o.F[n]) < 1e-13, n


def test_fused_step_equals_phases():
    nx, ny = 32, 32
    P = setup_weibel(nx, ny, epc=4, seed=2)
    o = Oracle(2, nx, ny, epc=4)
    F = smooth_fields(o, seed=2)
    res = []
    for fused in (False, True):
        say(f"fused={fused}")
        with Simulation(nx, ny, epc=4, form="fp64", tile=8) as sim:
            sim.upload_fields(*[F[n] for n in FIELDS])
            sim.upload_particles(**P)
            for _ in range(3):
                if fused:
                    sim.step(1)
                else:
                    sim.step_phases()
            res.append((sim.download_fields(), sim.download_particles()))
    # the FP64 form sums the current with FP64 atomics, so two runs agree to rounding only
    for n in FIELDS:
        assert rel_err(res[0][0][n], res[1][0][n]) < 1e-12, n
    a, b = sort_particles(res[0][1]), sort_particles(res[1][1])
    for c in ("x", "y", "z", "u", "v", "w"):
        assert np.max(np.abs(a[c] - b[c])) < 1e-12, c
    assert np.array_equal(a["flv"], b["flv"])


@pytest.mark.parametrize("nx,ny,tile", [(32, 32, 8), (64, 64, 32)])
def test_steps_fp64_vs_oracle(nx, ny, tile):
    o, sim, P = _pair(nx, ny, epc=16, seed=4, fields=False, tile=tile)
    nsteps = 40
    with sim:
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        G = sim.download_fields()
        d = sim.diagnostics()
        eo = o.energy()
        go = o.gauss()
        for n in FIELDS:
            assert rel_err(G[n], o.F[n]) < 1e-9, n
        assert abs(d["B_energy"] - eo[0]) <= 1e-9 * eo[0]
        assert abs(d["E_energy"] - eo[1]) <= 1e-9 * eo[1]
        assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-12 * eo[2]
        assert abs(d["prtl_energy_2"] - eo[3]) <= 1e-12 * eo[3]
        # Gauss's law at rounding level (relative to the largest node charge)
        assert d["gauss_residual"] <= 1e-12 * d["rho_max"], d
        assert go[0] <= 1e-12 * go[1]
        assert sim.particle_count() == len(P["x"])


@pytest.mark.parametrize("nx,ny,tile", [(32, 32, 8), (64, 64, 16), (128, 64, 32), (64, 64, 32)])
def test_split_vs_fp64_oracle(nx, ny, tile):
    o, sim, P = _pair(nx, ny, epc=16, form="fp32_split", seed=4, fields=False, tile=tile)
    nsteps = 40
    with sim:
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        d = sim.diagnostics()
        eo = o.energy()
        say("split vs fp64:", d, eo)
        # FP32 particle arithmetic: field energy within 1e-3 relative after 40 steps,
        # particle kinetic energy within 1e-5, Gauss residual within 1e-6 of max |rho|
        assert abs(d["B_energy"] - eo[0]) <= 1e-3 * eo[0]
        assert abs(d["E_energy"] - eo[1]) <= 1e-3 * eo[1]
        assert abs(d["prtl_energy_1"] - eo[2]) <= 1e-5 * eo[2]
        assert d["gauss_residual"] <= 1e-6 * d["rho_max"], d
        assert sim.particle_count() == len(P["x"])


@pytest.mark.parametrize("ndim", [2, 3])
def test_split_gauss_residual_stays_at_rounding(ndim):
    # the split form deposits the exact charge flux of its stored positions (2-D) and
    # accumulates E in FP64 as an FP32 (hi, lo) pair, so in 2-D div E - rho only drifts by
    # the rounding of each deposit to the 2^-30 q fixed-point grid of the current arrays
    # (a random walk of about 1e-9 q per deposit: about 2e-8 of max |rho| after 300 steps
    # here, where an FP32 E accumulates 1e-6 over 2000 steps at 2048^2); in 3-D the
    # reference zigzag itself drops the q dx dy dz / 12 term, so there the bar is the FP64
    # form's own residual
    from pictor_amd import setup_weibel
    n = (64, 64, 1) if ndim == 2 else (16, 16, 16)
    P = setup_weibel(n[0], n[1], n[2], ndim=ndim, epc=8, drift=0.5 if ndim == 2 else 0.3,
                     uth=0.01 if ndim == 2 else 0.02, seed=9)
    nsteps = 300 if ndim == 2 else 60
    res = {}
    for form in ("fp32_split", "fp64"):
        with Simulation(n[0], n[1], n[2], ndim=ndim, epc=8, form=form, tile=8 if ndim == 2 else 4) as sim:
            sim.upload_particles(**P)
            sim.step(nsteps)
            d = sim.diagnostics()
            res[form] = d["gauss_residual"] / d["rho_max"]
    say(f"{ndim}-D Gauss residual / max rho after {nsteps} steps: {res}")
    if ndim == 2:
        assert res["fp32_split"] <= 1e-7
    else:
        assert res["fp32_split"] <= 2 * res["fp64"] + 1e-11


@pytest.mark.parametrize("nx,ny,tile", [(64, 32, 8), (64, 64, 16), (128, 64, 32)])
def test_split_deterministic(nx, ny, tile):
    P = setup_weibel(nx, ny, epc=8, seed=9)
    res = []
    for rep in range(2):
        say(f"rep={rep} nx={nx} ny={ny} tile={tile}")
        with Simulation(nx, ny, epc=8, form="fp32_split", tile=tile) as sim:
            sim.upload_particles(**P)
            sim.step(25)
            res.append((sim.download_fields(), sim.diagnostics()))
    for n in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jx", "Jy", "Jz"):
        assert np.array_equal(res[0][0][n], res[1][0][n]), n
    for k in ("B_energy", "E_energy"):
        assert res[0][1][k] == res[1][1][k]


def test_spill_relayout_keeps_particles():
    # every particle starts in one 8x8 tile and streams into empty tiles: the
    # bins of the empty tiles overflow, particles spill and the layout is rebuilt
    nx, ny = 32, 32
    n = 4000
    rng = np.random.default_rng(1)
    P = {"q": np.where(np.arange(n) % 2 == 0, 1.0, -1.0), "x": 3 + rng.uniform(0, 8, n), "y": 3 + rng.uniform(0, 8, n),
         "z": 1 + rng.uniform(0, 1, n), "u": np.full(n, 0.8), "v": np.full(n, 0.6), "w": np.zeros(n),
         "flv": np.where(np.arange(n) % 2 == 0, 1, 2).astype(np.int32)}
    o = Oracle(2, nx, ny, epc=4, cap=3 * n)
    o.load_particles(**P)
    with Simulation(nx, ny, epc=4, form="fp64", tile=8) as sim:
        sim.upload_particles(**P)
        for t in range(1, 31):
            say(f"step {t}")
            sim.step(1)
            o.step(t)
            assert sim.particle_count() == n
        G = sim.download_particles()
        assert np.allclose(np.sort(G["x"]), np.sort(o.particles()["x"]), rtol=0, atol=1e-9)


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
def test_sort_tiles_against_oracle_reorder(form):
    # pictor_sort_tiles (ReorderPrtlArr) after a push that moved particles
    # across tiles (the push writes each into an arrival bin of its new tile),
    # against orc_reorder on the same particles: the same multiset (FP64; the
    # split form's offsets differ from the oracle's positions at FP32 rounding),
    # every particle in its own tile's bin, the bins in tile order and, per tile,
    # the particles that the oracle's cell-ordered array holds in that tile
    nx, ny, T = 64, 32, 8
    say(f"sort_tiles form={form} nx={nx} ny={ny} tile={T}")
    rng = np.random.default_rng(21)
    P = setup_weibel(nx, ny, epc=4, drift=0.9, uth=0.3, seed=21)
    perm = rng.permutation(len(P["x"]))
    P = {k: v[perm] for k, v in P.items()}
    o = Oracle(2, nx, ny, epc=4, c=0.45, cap=2 * len(P["x"]) + 4096)
    F = smooth_fields(o, seed=21)
    o.load_particles(**P)
    with Simulation(nx, ny, epc=4, c=0.45, form=form, tile=T) as sim:
        sim.upload_fields(*[F[n] for n in FIELDS])
        sim.upload_particles(**P)
        sim.ResetCurrent(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
        o.reset_current(); o.move_deposit(); o.exchange()
        moved = np.sum(tiles_of(P, nx, ny, T) != tiles_of(o.particles(), nx, ny, T)) if form == "fp64" else 1
        sim.ReorderPrtl()
        o.reorder()
        G = sim.download_particles(with_tile=True)
    R = o.particles()
    assert moved > 0
    cells = (np.floor(R["y"]).astype(np.int64) - 3) * nx + np.floor(R["x"]).astype(np.int64) - 3
    assert np.all(np.diff(cells) >= 0)                    # the oracle: ordered by cell
    assert len(G["x"]) == len(R["x"]) == len(P["x"])
    assert np.array_equal(G["tile"], tiles_of(G, nx, ny, T))
    assert np.all(np.diff(G["tile"]) >= 0)
    rt = tiles_of(R, nx, ny, T)
    for t in np.unique(rt):
        a = {k: v[G["tile"] == t] for k, v in G.items() if k != "tile"}
        b = {k: v[rt == t] for k, v in R.items()}
        if form == "fp64":
            assert_same_multiset(a, b)
        else:
            assert len(a["x"]) == len(b["x"]), t
            assert np.max(np.abs(np.sort(a["x"]) - np.sort(b["x"]))) < 1e-5, t


def test_upload_rejects_out_of_domain():
    with Simulation(16, 16, epc=2, form="fp64", tile=8) as sim:
        P = setup_weibel(16, 16, epc=2, seed=1)
        P["x"][5] = 0.5
        with pytest.raises(PictorError) as e:
            sim.upload_particles(**P)
        assert e.value.code == 1


@pytest.mark.parametrize("ndim", [2, 3])
def test_lehe_fields_bitwise(ndim):
    n = (24, 16, 8) if ndim == 3 else (32, 24, 1)
    say(f"lehe ndim={ndim} n={n}")
    o = Oracle(ndim, *n, epc=2, fdtd="lehe")
    F = smooth_fields(o, amp_e=1.0, amp_b=1.0, seed=13, modes=6)
    rng = np.random.default_rng(8)
    sl = (slice(3, -2), slice(3, -2), slice(3, -2) if ndim == 3 else slice(None))
    for k in FIELDS:
        F[k][sl] += 1e-3 * rng.standard_normal(F[k][sl].shape)
    o.set_fields(**F)
    o.field_halo("E"); o.field_halo("B")
    F = {k: o.F[k].copy(order="F") for k in FIELDS}
    with Simulation(*n, ndim=ndim, epc=2, form="fp64", tile=8 if ndim == 2 else 4, fdtd="lehe") as sim:
        sim.upload_fields(*[F[k] for k in FIELDS])
        for _ in range(6):
            sim.UpdateBfieldHalf()
            o.update_b_half(); o.field_halo("B")
            sim.UpdateEfield()
            o.update_e(); o.field_halo("E")
        G = sim.download_fields()
    for k in FIELDS:
        assert np.array_equal(G[k], o.F[k]), k


@pytest.mark.parametrize("form", ["fp64", "fp32_split"])
def test_lehe_steps_vs_oracle(form):
    nx, ny = 32, 32
    say(f"lehe steps form={form}")
    P = setup_weibel(nx, ny, epc=16, drift=0.5, uth=0.01, seed=4)
    o = Oracle(2, nx, ny, epc=16, cap=2 * len(P["x"]) + 4096, fdtd="lehe")
    o.load_particles(**P)
    nsteps = 40
    with Simulation(nx, ny, epc=16, form=form, tile=8, fdtd="lehe") as sim:
        sim.upload_particles(**P)
        for t in range(1, nsteps + 1):
            o.step(t)
        sim.step(nsteps)
        d = sim.diagnostics()
        eo = o.energy()
        if form == "fp64":
            G = sim.download_fields()
            for k in FIELDS:
                assert rel_err(G[k], o.F[k]) < 1e-9, k
            assert abs(d["B_energy"] - eo[0]) <= 1e-9 * eo[0]
            assert d["gauss_residual"] <= 1e-12 * d["rho_max"], d
        else:
            assert abs(d["B_energy"] - eo[0]) <= 1e-3 * eo[0]
            assert abs(d["E_energy"] - eo[1]) <= 1e-3 * eo[1]
            assert d["gauss_residual"] <= 1e-6 * d["rho_max"], d
