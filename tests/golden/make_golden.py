"""Write the golden fixtures of the particle step: small seeded runs of the oracle.

Each fixture holds the inputs of a run (particles in the reference's SoA,
fields with their ghost layers) and what the oracle's MAIN loop
(orc_step: src/MAIN.F90:55-120 restated) leaves after ``nsteps`` steps: the
nine field arrays, the particles (sorted, holes dropped), the energies of
CalcFldEnergy / CalcPrtlEnergy, the Gauss residual and the momentum sums.
The inputs come from numpy's seeded generator, not from the product's
loader, so the fixtures depend on nothing but this script and the oracle.

    python tests/golden/make_golden.py          # rewrites tests/golden/*.npz

The CPU test test_golden.py checks that the oracle still reproduces every
fixture bit for bit; the GPU test test_gpu_golden.py runs the HIP path from
the same inputs against the stored outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.pyoracle import Oracle  # noqa: E402

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
PCOLS = ("q", "x", "y", "z", "u", "v", "w", "flv")

# name: (ndim, nx, ny, nz, epc, fdtd, nsteps, seed, particle faces, field faces)
CASES = {
    "weibel2d_16x16_yee": (2, 16, 16, 1, 2, "yee", 10, 101, [], []),
    "weibel2d_16x16_lehe": (2, 16, 16, 1, 2, "lehe", 10, 102, [], []),
    "weibel3d_8x8x8_yee": (3, 8, 8, 8, 1, "yee", 10, 103, [], []),
    "faces2d_16x16": (2, 16, 16, 1, 2, "yee", 10, 104,
                      [("x", "lower", "refl", 2.5), ("x", "upper", "remv", 13.5)],
                      [("x", "lower", "cond", 2.0), ("x", "upper", "cond", 14.0)]),
}


def plasma(ndim, nx, ny, nz, epc, seed, drift=0.5, uth=0.05):
    """Counter-streaming pair plasma: per cell epc pairs of (q = +1, flv 1) and
    (q = -1, flv 2) at one position, streaming along x with +-drift."""
    rng = np.random.default_rng(seed)
    ncell = nx * ny * (nz if ndim == 3 else 1)
    idx = np.repeat(np.arange(ncell), epc)
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    n = len(idx)
    x = 3.0 + i + rng.random(n)
    y = 3.0 + j + rng.random(n)
    z = (3.0 + k + rng.random(n)) if ndim == 3 else 1.0 + rng.random(n)
    sgn = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    out = {c: np.empty(2 * n) for c in PCOLS[:-1]}
    out["flv"] = np.empty(2 * n, np.int32)
    for s, (q, flv) in enumerate(((1.0, 1), (-1.0, 2))):
        sl = slice(s, 2 * n, 2)
        out["q"][sl], out["flv"][sl] = q, flv
        out["x"][sl], out["y"][sl], out["z"][sl] = x, y, z
        out["u"][sl] = (1 - 2 * s) * sgn * drift + uth * rng.standard_normal(n)
        out["v"][sl] = uth * rng.standard_normal(n)
        out["w"][sl] = uth * rng.standard_normal(n)
    return out


def inside(P, faces):
    keep = np.ones(len(P["x"]), bool)
    for axis, side, kind, pos in faces:
        c = P[axis] - 3.0
        keep &= (c > pos + 0.02) if side == "lower" else (c < pos - 0.02)
    return {k: v[keep] for k, v in P.items()}


def sort_particles(P):
    order = np.lexsort([P[c] for c in reversed(PCOLS)])
    return {c: np.asarray(P[c])[order] for c in PCOLS}


def build(name):
    """The oracle and inputs of one case (faces set, fields and particles loaded)."""
    ndim, nx, ny, nz, epc, fdtd, nsteps, seed, pfaces, ffaces = CASES[name]
    P = inside(plasma(ndim, nx, ny, nz, epc, seed), pfaces)
    o = Oracle(ndim, nx, ny, nz, epc=epc, fdtd=fdtd, cap=2 * len(P["x"]) + 1024)
    for axis, side, kind, pos in ffaces:
        o.set_bc_fld(axis, pos, kind, side)
    for axis, side, kind, pos in pfaces:
        o.set_bc_prtl(axis, pos, kind, side)
    rng = np.random.default_rng(seed + 1000)
    o.set_fields(**{f: np.asfortranarray(0.02 * rng.standard_normal(o.shape)) for f in FIELDS})
    o.field_halo("E")
    o.field_halo("B")
    F = {f: o.F[f].copy(order="F") for f in FIELDS}
    o.load_particles(**P)
    return o, P, F, nsteps


def run(name):
    o, P, F, nsteps = build(name)
    for t in range(1, nsteps + 1):
        o.step(t)
    res = {"in_" + k: v for k, v in P.items()}
    res.update({"in_" + f: F[f] for f in FIELDS})
    res.update({"out_" + f: o.F[f].copy(order="F") for f in o.FIELD_NAMES})
    res.update({"out_" + k: v for k, v in sort_particles(o.particles()).items()})
    res["out_energy"] = o.energy()
    res["out_gauss"] = o.gauss()
    res["out_momentum"] = o.momentum()
    return res


def main():
    for name in CASES:
        res = run(name)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **res)
        print(f"{name}: {len(res['in_x'])} -> {len(res['out_x'])} particles, "
              f"energies {res['out_energy']}, {os.path.getsize(path)} B")


if __name__ == "__main__":
    main()
