"""C ABI boundary and host-side logic (no GPU needed).

* libpictor_hip.so loads and exports every function include/pictor_hip.h declares;
* the Python binding covers the same set;
* the seeded Weibel input generator (host code) produces the configuration it
  states: 2*epc particles per cell, both flavours at each position, drift
  +-0.5c added with AddDriftVel, thermal spread uth, opposite drifts per call.
"""
import os
import re

import numpy as np

import pictor_amd
from pictor_amd import setup_weibel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pictor_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pictor_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = pictor_amd.lib()
    decl = header_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(pictor_amd.EXPORTS) == decl


def test_last_error_callable_without_gpu():
    L = pictor_amd.lib()
    assert isinstance(L.pictor_last_error(), bytes)


def test_setup_weibel_layout_and_statistics():
    nx, ny, epc = 16, 8, 8
    P = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=42)
    n = len(P["x"])
    assert n == 2 * epc * nx * ny
    # both flavours at every position, flavour 1 first
    assert np.array_equal(P["flv"][0::2], np.ones(n // 2)) and np.array_equal(P["flv"][1::2], np.full(n // 2, 2))
    assert np.array_equal(P["x"][0::2], P["x"][1::2]) and np.array_equal(P["y"][0::2], P["y"][1::2])
    assert np.array_equal(P["q"], np.where(P["flv"] == 1, 1.0, -1.0))
    # epc particles of each flavour in every cell, positions inside [3, 3+n)
    i = np.floor(P["x"]).astype(int) - 3
    j = np.floor(P["y"]).astype(int) - 3
    assert i.min() == 0 and i.max() == nx - 1 and j.min() == 0 and j.max() == ny - 1
    counts = np.bincount((i + nx * j)[P["flv"] == 1], minlength=nx * ny)
    assert np.all(counts == epc)
    assert np.all((P["z"] >= 1.0) & (P["z"] < 2.0))
    # drift: gamma*beta of 0.5c is 0.5/sqrt(0.75); half of each flavour each way
    ud = 0.5 / np.sqrt(0.75)
    for f in (1, 2):
        u = P["u"][P["flv"] == f]
        assert abs(np.mean(np.abs(u)) - ud) < 2e-3
        assert abs(np.mean(np.sign(u))) < 1e-12
    # the pair at one position drifts in opposite directions
    assert np.all(np.sign(P["u"][0::2]) == -np.sign(P["u"][1::2]))
    # thermal spread of the untouched components
    assert abs(np.std(P["v"]) - 0.01) < 1e-3 and abs(np.std(P["w"]) - 0.01) < 1e-3
    # determinism of the counter-based generator
    Q = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=42)
    assert all(np.array_equal(P[k], Q[k]) for k in P)
    R = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.01, seed=43)
    assert not np.array_equal(P["x"], R["x"])


def test_setup_weibel_rejects_bad_arguments():
    from pictor_amd._lib import PictorError
    try:
        setup_weibel(8, 8, epc=3)
    except PictorError as e:
        assert e.code == 1
    else:
        raise AssertionError("odd epc accepted")
