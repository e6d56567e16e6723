"""Host side of the adjoint (no GPU): the torch AddDriftVel boost and the
drift signs reproduce the library's Weibel generator, and the C ABI exports
the adjoint entry points."""
import numpy as np
import torch

from pictor_amd import setup_weibel
from pictor_amd.adjoint import drift_boost, drift_signs


def test_drift_boost_reproduces_generator():
    P0 = setup_weibel(16, 8, epc=6, drift=0.0, uth=0.02, seed=3)
    for drift in (0.2, 0.5, -0.3):
        P1 = setup_weibel(16, 8, epc=6, drift=drift, uth=0.02, seed=3)
        s = torch.from_numpy(drift_signs(len(P0["x"]), 6))
        u, v, w = drift_boost(*[torch.from_numpy(P0[k]) for k in "uvw"], s, torch.tensor(drift, dtype=torch.float64))
        assert np.max(np.abs(u.numpy() - P1["u"])) <= 4e-16 * np.max(np.abs(P1["u"])) + 1e-300
        assert np.array_equal(v.numpy(), P1["v"]) and np.array_equal(w.numpy(), P1["w"])
        for k in ("x", "y", "z", "q", "flv"):
            assert np.array_equal(P0[k], P1[k])


def test_drift_boost_derivative_closed_form():
    # du/d(beta_d) = gamma_d^3 (beta_d u_th + s gamma_th)
    rng = np.random.default_rng(1)
    u_th = torch.from_numpy(rng.normal(0, 0.05, 50))
    v_th = torch.from_numpy(rng.normal(0, 0.05, 50))
    w_th = torch.from_numpy(rng.normal(0, 0.05, 50))
    s = torch.from_numpy(np.where(rng.random(50) < 0.5, -1.0, 1.0))
    b = torch.tensor(0.4, dtype=torch.float64, requires_grad=True)
    u, _, _ = drift_boost(u_th, v_th, w_th, s, b)
    u.sum().backward()
    gth = torch.sqrt(1 + u_th ** 2 + v_th ** 2 + w_th ** 2)
    gd = 1 / np.sqrt(1 - 0.4 ** 2)
    want = float((gd ** 3 * (0.4 * u_th + s * gth)).sum())
    assert abs(float(b.grad) - want) <= 1e-13 * abs(want)
