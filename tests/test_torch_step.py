"""The torch restatement of the 2-D step used as the adjoint's reference
gradient agrees with the C oracle's forward (no GPU)."""
import numpy as np

from oracle.pyoracle import Oracle
from pictor_amd import setup_weibel
from tests._torch_step import Step2D


def test_torch_step_matches_oracle():
    nx, ny, epc, nsteps = 16, 16, 4, 12
    P = setup_weibel(nx, ny, epc=epc, drift=0.5, uth=0.1, seed=3)
    o = Oracle(2, nx, ny, epc=epc, cap=2 * len(P["x"]), reorder_period=0)
    o.load_particles(**P)
    for t in range(1, nsteps + 1):
        o.step(t)
    ts = Step2D(nx, ny, epc=epc)
    L, G = ts.gradient(P, nsteps)
    Lo = o.energy()[0]
    assert abs(L - Lo) <= 1e-12 * Lo, (L, Lo)
    assert all(np.all(np.isfinite(g)) for g in G)
