"""Shared helpers for the parity tests (oracle-side inputs and comparisons)."""
import sys

import numpy as np

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
PCOLS = ("q", "x", "y", "z", "u", "v", "w", "flv")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def smooth_fields(oracle, amp_e=0.01, amp_b=0.05, seed=0, modes=3):
    """Random smooth periodic fields on the physical cells; ghosts filled by the oracle's halo copy."""
    rng = np.random.default_rng(seed)
    p = oracle.p
    out = {}
    i = np.arange(p.mx)[:, None, None] - 3.0
    j = np.arange(p.my)[None, :, None] - 3.0
    k = np.arange(p.mz)[None, None, :] - 3.0
    for name in FIELDS:
        amp = amp_e if name[0] == "E" else amp_b
        f = np.zeros((p.mx, p.my, p.mz))
        for _ in range(modes):
            kx = 2 * np.pi * rng.integers(0, 3) / p.nx
            ky = 2 * np.pi * rng.integers(0, 3) / p.ny
            kz = 2 * np.pi * rng.integers(0, 3) / p.nz if p.ndim == 3 else 0.0
            ph = rng.uniform(0, 2 * np.pi)
            f += amp * rng.uniform(-1, 1) * np.cos(kx * i + ky * j + kz * k + ph)
        f += amp * 0.3 * rng.uniform(-1, 1)
        out[name] = np.asfortranarray(f)
    oracle.set_fields(**out)
    oracle.field_halo("E")
    oracle.field_halo("B")
    return {n: oracle.F[n].copy(order="F") for n in FIELDS}


def sort_particles(P):
    keys = [P[c] for c in reversed(PCOLS)]
    order = np.lexsort(keys)
    return {c: np.asarray(P[c])[order] for c in PCOLS}


def assert_same_multiset(A, B):
    a, b = sort_particles(A), sort_particles(B)
    assert len(a["x"]) == len(b["x"]), (len(a["x"]), len(b["x"]))
    for c in PCOLS:
        if not np.array_equal(a[c], b[c]):
            bad = np.nonzero(a[c] != b[c])[0]
            raise AssertionError(f"column {c}: {len(bad)} mismatches, first at {bad[:5]}: "
                                 f"{a[c][bad[:5]]} vs {b[c][bad[:5]]}")


def rel_err(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def tiles_of(P, nx, ny, T):
    cx = np.clip(np.floor(P["x"]).astype(np.int64) - 3, 0, nx - 1)
    cy = np.clip(np.floor(P["y"]).astype(np.int64) - 3, 0, ny - 1)
    return (cy // T) * (nx // T) + cx // T


def tiles_of3(P, nx, ny, nz, T):
    cx = np.clip(np.floor(P["x"]).astype(np.int64) - 3, 0, nx - 1)
    cy = np.clip(np.floor(P["y"]).astype(np.int64) - 3, 0, ny - 1)
    cz = np.clip(np.floor(P["z"]).astype(np.int64) - 3, 0, nz - 1)
    return ((cz // T) * (ny // T) + cy // T) * (nx // T) + cx // T
