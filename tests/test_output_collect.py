"""Pins of the CollectFld restatement the field-file tests compare against (no GPU).

On the Yee grid Ex sits at (i + 1/2, j, k), Ey at (i, j + 1/2, k), Ez at
(i, j, k + 1/2) (3-D; the node in 2-D), Bx at (i, j + 1/2, k + 1/2), By at
(i + 1/2, j, k + 1/2), Bz at (i + 1/2, j + 1/2, k) in the reference's index
convention (the averages of CollectFld, savedata_routines.F90:227-246, bring
each to node (i, j, k)).  A field linear in space, sampled at its own
staggered points, must therefore come back as the linear function at the
nodes; a uniform field must come back unchanged; the node set is 3, 3 + r,
..., 3 + r int(n / r).  A wrong stagger, index or weight fails one of these.
"""
import numpy as np
import pytest

from tests._collect import collect

STAG = {  # offsets of each component from the node, in cells (2-D / 3-D)
    2: {"Ex": (-0.5, 0, 0), "Ey": (0, -0.5, 0), "Ez": (0, 0, 0), "Bx": (0, -0.5, 0), "By": (-0.5, 0, 0),
        "Bz": (-0.5, -0.5, 0)},
    3: {"Ex": (-0.5, 0, 0), "Ey": (0, -0.5, 0), "Ez": (0, 0, -0.5), "Bx": (0, -0.5, -0.5), "By": (-0.5, 0, -0.5),
        "Bz": (-0.5, -0.5, 0)},
}


def _grid(ndim, n):
    m = n + 5
    return (m, m, m if ndim == 3 else 1)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("res", [1, 2, 3])
def test_linear_fields_come_back_at_the_nodes(ndim, res):
    n = 10
    shape = _grid(ndim, n)
    idx = np.meshgrid(*[np.arange(s, dtype=float) + 1 for s in shape], indexing="ij")   # 1-based indices
    coef = {"Ex": (0.25, -0.5, 0.125), "Ey": (-0.375, 0.75, 0.25), "Ez": (0.5, 0.25, -0.75),
            "Bx": (0.125, 0.5, 0.375), "By": (-0.25, -0.125, 0.5), "Bz": (0.75, -0.25, 0.125)}
    F = {}
    for name, (a, b, c) in coef.items():
        # value at the component's own position: node index (i, j, k) stores the point (i, j, k) - offset... the
        # component stored at index i lives at i + 1/2 (x-staggered), so its position is index - off
        off = STAG[ndim][name]
        pos = [idx[d] - off[d] for d in range(3)]
        zpos = pos[2] if ndim == 3 else 0.0
        F[name] = np.asfortranarray(1.0 + a * pos[0] + b * pos[1] + c * zpos)
    out = collect(F, ndim, res)
    nodes = np.arange(3, res * (n // res) + 3 + 1, res).astype(float)
    kn = nodes if ndim == 3 else np.array([1.0])
    I, J, K = np.meshgrid(nodes, nodes, kn, indexing="ij")
    for name, (a, b, c) in coef.items():
        want = 1.0 + a * I + b * J + (c * K if ndim == 3 else 0.0)
        assert out[name].shape == want.shape
        assert np.max(np.abs(out[name] - want.astype(np.float32))) <= 4e-6 * np.max(np.abs(want)), name


def test_uniform_fields_and_node_count():
    for ndim in (2, 3):
        shape = _grid(ndim, 12)
        F = {name: np.full(shape, v, order="F") for name, v in zip(("Ex", "Ey", "Ez", "Bx", "By", "Bz", "Jz"),
                                                                     (1.5, -2.0, 0.25, 3.0, -0.5, 4.0, 7.0))}
        for res, cnt in ((1, 13), (3, 5), (5, 3), (12, 2)):
            out = collect(F, ndim, res)
            assert out["Ex"].shape == ((cnt, cnt, cnt) if ndim == 3 else (cnt, cnt, 1))
            for name in F:
                assert np.all(out[name] == np.float32(F[name][0, 0, 0])), name
