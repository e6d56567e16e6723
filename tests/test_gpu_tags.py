"""Per-particle tags (the reference's tagp) through the C ABI (needs an MI355X).

Bars: after one push (FP64, 2-D and 3-D) every downloaded particle, found by its tag,
equals the oracle's update of the particle uploaded with that tag (positions within
1e-12 after the periodic wrap, momenta bit for bit); tags survive re-layouts, spills,
boundary faces (removed particles take their tags with them) and restart files; a
tagged run follows the same physics as an untagged one (split form: the same particle
multiset bit for bit); an untagged context refuses to download tags.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from pictor_amd import PictorError, Simulation, setup_weibel
from tests._util import FIELDS, assert_same_multiset, say, smooth_fields

pytestmark = pytest.mark.gpu


def _tags(n, seed=5):
    # distinct, non-trivial tags (a permutation of 0..n-1 scaled and offset)
    return (np.random.default_rng(seed).permutation(n) * 3 + 7).astype(np.int32)


@pytest.mark.parametrize("ndim", [2, 3])
def test_tags_follow_the_oracle_over_one_push(ndim):
    n3 = (32, 24, 1) if ndim == 2 else (16, 16, 16)
    tile = 8 if ndim == 2 else 4
    say(f"tags one push ndim={ndim}")
    P = setup_weibel(*n3, ndim=ndim, epc=4, drift=0.5, uth=0.1, seed=11)
    tag = _tags(len(P["x"]))
    o = Oracle(ndim, *n3, epc=4, cap=2 * len(P["x"]) + 1024)
    F = smooth_fields(o, seed=11)
    o.load_particles(**P)
    with Simulation(*n3, ndim=ndim, epc=4, form="fp64", tile=tile) as sim:
        sim.upload_fields(*[F[k] for k in FIELDS])
        sim.upload_particles(**P, tag=tag)
        sim.ResetCurrent(); sim.MoveDepositPrtl(); sim.ExchangePrtl()
        G = sim.download_particles(with_tag=True)
    o.reset_current(); o.move_deposit()          # in place: slot i is uploaded particle i
    R = o.particles(live_only=False)
    assert np.array_equal(np.sort(G["tag"]), np.sort(tag))
    where = {t: i for i, t in enumerate(tag)}
    idx = np.array([where[t] for t in G["tag"]])
    for c in ("u", "v", "w", "q"):
        assert np.array_equal(G[c], R[c][idx]), c
    assert np.array_equal(G["flv"], R["flv"][idx])
    axes = (("x", n3[0]), ("y", n3[1])) + ((("z", n3[2]),) if ndim == 3 else ())
    for c, m in axes:
        r = R[c][idx]
        r = np.where(r >= 3 + m, r - m, np.where(r < 3, r + m, r))
        assert np.max(np.abs(G[c] - r)) < 1e-12, c


@pytest.mark.parametrize("form", ["fp32_split", "fp64"])
def test_tags_survive_relayouts(form):
    # every particle starts in one tile and streams into empty ones: bins overflow,
    # particles spill and the layout is rebuilt several times
    nx = ny = 32
    n = 4000
    rng = np.random.default_rng(1)
    P = {"q": np.where(np.arange(n) % 2 == 0, 1.0, -1.0), "x": 3 + rng.uniform(0, 8, n), "y": 3 + rng.uniform(0, 8, n),
         "z": 1 + rng.uniform(0, 1, n), "u": np.full(n, 0.8), "v": np.full(n, 0.6), "w": np.zeros(n),
         "flv": np.where(np.arange(n) % 2 == 0, 1, 2).astype(np.int32)}
    tag = _tags(n, 2)
    res = {}
    for tagged in (True, False):
        say(f"relayout form={form} tagged={tagged}")
        with Simulation(nx, ny, epc=4, form=form, tile=8) as sim:
            sim.upload_particles(**P, tag=tag if tagged else None)
            sim.step(30)
            assert sim.relayout_stats()[0] > 0
            res[tagged] = sim.download_particles(with_tag=tagged)
    G = res[True]
    assert np.array_equal(np.sort(G["tag"]), np.sort(tag))
    # the tag follows its particle: the flavour of each tag is the one uploaded with it
    where = {t: i for i, t in enumerate(tag)}
    idx = np.array([where[t] for t in G["tag"]])
    assert np.array_equal(G["flv"], P["flv"][idx])
    if form == "fp32_split":
        # deterministic current: tags change nothing in the physics
        assert_same_multiset({k: v for k, v in G.items() if k != "tag"}, res[False])


def test_tags_with_faces_and_restart(tmp_path):
    nx, ny = 32, 24
    P = setup_weibel(nx, ny, epc=4, drift=0.5, uth=0.3, seed=4)
    keep = (P["x"] - 3 > 4.02) & (P["x"] - 3 < 27.98)
    P = {k: v[keep] for k, v in P.items()}
    tag = _tags(len(P["x"]), 9)
    path = str(tmp_path / "tagged_state")
    runs = []
    for restart in (False, True):
        with Simulation(nx, ny, epc=4, form="fp32_split", tile=8) as sim:
            sim.SetBC_Prtl("x", 4.0, "refl", "lower")
            sim.SetBC_Prtl("x", 28.0, "remv", "upper")
            sim.SetBC_Fld("x", 3.0, "cond", "lower")
            sim.upload_particles(**P, tag=tag)
            sim.step(10)
            if restart:
                sim.save_state(path, 10)
                with Simulation(nx, ny, epc=4, form="fp32_split", tile=8) as sim2:
                    assert sim2.load_state(path) == 10
                    sim2.step(10)
                    runs.append(sim2.download_particles(with_tag=True))
            else:
                sim.step(10)
                runs.append(sim.download_particles(with_tag=True))
    a, b = runs
    assert len(a["x"]) < len(P["x"])                       # the removing face took particles
    assert len(np.unique(a["tag"])) == len(a["tag"])
    assert set(a["tag"]) <= set(tag)
    oa, ob = np.argsort(a["tag"]), np.argsort(b["tag"])
    assert np.array_equal(a["tag"][oa], b["tag"][ob])
    for c in ("x", "y", "u", "v", "w", "flv"):
        assert np.array_equal(a[c][oa], b[c][ob]), c


def test_untagged_context_refuses_tag_download():
    P = setup_weibel(16, 16, epc=2, seed=1)
    with Simulation(16, 16, epc=2, form="fp64", tile=8) as sim:
        sim.upload_particles(**P)
        with pytest.raises(PictorError):
            sim.download_particles(with_tag=True)
        sim.upload_particles(**P, tag=_tags(len(P["x"])))   # a tagged upload switches the record
        assert len(sim.download_particles(with_tag=True)["tag"]) == len(P["x"])
        sim.upload_particles(**P)                           # and an untagged one back
        with pytest.raises(PictorError):
            sim.download_particles(with_tag=True)
