"""The committed golden fixtures (tests/golden/*.npz) against the oracle, on the CPU.

Each fixture is a 10-step run of the oracle's MAIN loop at 16 x 16 (2-D, Yee
and Lehe, and with reflecting / removing particle faces and conducting field
faces) or 8^3 (3-D), written by tests/golden/make_golden.py.  The oracle must
reproduce every stored array bit for bit: a change anywhere on the step path
(gather, push, deposit, halos, field update, faces, diagnostics) shows here.
"""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mg)


def load(name):
    with np.load(os.path.join(HERE, name + ".npz")) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", list(mg.CASES))
def test_oracle_reproduces_fixture(name):
    ref = load(name)
    res = mg.run(name)
    assert sorted(res) == sorted(ref)
    for k in ref:
        assert res[k].dtype == ref[k].dtype, k
        assert np.array_equal(res[k], ref[k]), k


@pytest.mark.parametrize("name", list(mg.CASES))
def test_fixture_invariants(name):
    # what the physics fixes about the stored runs: the particle charge is
    # conserved without removing faces; every particle stays inside the box
    # (and inside its faces); the inputs are what the generator makes
    ref = load(name)
    ndim, nx, ny, nz, epc, fdtd, nsteps, seed, pfaces, ffaces = mg.CASES[name]
    ncell = nx * ny * (nz if ndim == 3 else 1)
    n_in = len(ref["in_x"])
    if not pfaces:
        assert n_in == 2 * epc * ncell
        assert len(ref["out_x"]) == n_in
        assert ref["out_q"].sum() == ref["in_q"].sum() == 0.0
    else:
        assert len(ref["out_x"]) < n_in          # the removing face took some
    for axis, n in (("x", nx), ("y", ny)) + ((("z", nz),) if ndim == 3 else ()):
        c = ref["out_" + axis] - 3.0
        assert c.min() >= 0.0 and c.max() < n, axis
    for axis, side, kind, pos in pfaces:
        c = ref["out_" + axis] - 3.0
        assert (c >= pos).all() if side == "lower" else (c <= pos).all()
    assert np.isin(ref["out_flv"], (1, 2)).all()
    assert np.all(ref["out_energy"] > 0)
